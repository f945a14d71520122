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
