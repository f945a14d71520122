"""Copy the reference's own input data files into small test fixtures.

Run in a container that has the reference checkout (REFERENCE_ROOT, default
/root/reference).  The reference holds no expected outputs for its hot path; its
data files are a Shepp-Logan tilt series (3 identical detector rows x 87 views x
128 bins, FP64, layout [t][theta][r]) with its tilt list, and a gains/offsets file
(87 gains = 1, 87 offsets = 0).  They are used as inputs of GPU-vs-oracle tests.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("REFERENCE_ROOT", "/root/reference")
SRC = os.path.join(REF, "Resources", "InputFiles")


def main():
    sino = np.fromfile(os.path.join(SRC, "ScheppLoganSinogram128NoiselessLimitedAngle.bin"), "<f8")
    sino = sino.reshape(3, 87, 128)
    txt = open(os.path.join(SRC, "ParamsWithROIAndViewSelectShepLogan.txt")).read().split()
    i = txt.index("SinoAngles") + 1 if "SinoAngles" in txt else None
    if i is None:  # the file glues tokens: "...SinoDelta_t1SinoAngles"
        joined = " ".join(txt)
        rest = joined.split("SinoAngles", 1)[1].split()
    else:
        rest = txt[i:]
    angles = []
    for a, flag in zip(rest[0::2], rest[1::2]):
        if not a.lstrip("-").isdigit():
            break
        if int(flag) == 1:
            angles.append(float(a))
    assert len(angles) == 87, len(angles)
    go = np.fromfile(os.path.join(SRC, "GainOffsetParams_ShepLoganPhantom.bin"), "<f8")
    np.save(os.path.join(HERE, "shepp_logan_sino_t_theta_r.npy"), sino)
    np.save(os.path.join(HERE, "shepp_logan_angles_deg.npy"), np.array(angles))
    np.save(os.path.join(HERE, "shepp_logan_gains_offsets.npy"), go)
    print("wrote fixtures:", sino.shape, len(angles), go.shape)


if __name__ == "__main__":
    sys.exit(main())
