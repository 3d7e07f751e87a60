"""Writes sample0_thickness_q4.npy: the reference's brush stroke sample
thickness map (data/sample_0/thickness_map.png, 16-bit gray, 171 x 800, the
sample TextureBrushTest.cxx:17-18 paints with), every 4th row and column, as
the raw uint16 values.  The tests scale them by 1/65535 and convert from sRGB
to linear as io::imRead(path, Mat<double>&, true) does
(io/src/ImageIO.cxx:78-112).

Run in a checkout that has the reference next to it:
    python tests/golden/texture/make_fixture.py /root/reference
"""
import os
import sys

import numpy as np
from PIL import Image


def main(ref_root: str) -> None:
    src = os.path.join(ref_root, "data", "sample_0", "thickness_map.png")
    a = np.asarray(Image.open(src))
    assert a.dtype == np.uint16 and a.shape == (171, 800), (a.dtype, a.shape)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sample0_thickness_q4.npy")
    np.save(out, np.ascontiguousarray(a[::4, ::4]))
    print("wrote", out, a[::4, ::4].shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
