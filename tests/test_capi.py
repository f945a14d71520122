"""The C-ABI library loads and exports every entry point include/*.h declares
(no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mbirtem import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(REPO, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                   if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mbir_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding knows all of them
    assert set(syms) <= set(capi._SIG), set(syms) - set(capi._SIG)


def test_host_only_entry_points():
    assert capi.lib().mbir_version() >= 10000
    with pytest.raises(capi.MbirError) as ei:
        capi.geometry_init(0, 4, np.zeros(3), 8)
    assert ei.value.code == -1


def test_engine_cli_rejects_wrong_argument_counts():
    """The CLI checks its arguments before any device call (runs without a GPU)."""
    import subprocess
    cli = os.path.join(os.path.dirname(capi.LIB_PATH), "engine_cli")
    for mode in ("--mrc", "--multires", "--bf-multires"):
        r = subprocess.run([cli, mode, "x"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (mode, r.stderr)
        assert mode in r.stderr
