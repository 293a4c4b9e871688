"""The correction cap's square-root-free test in the kernels.

The kernels decide `sqrt_rn(s) > cap` as `s > CAP_SQ_LIMIT` (the root is
monotone), with CAP_SQ_LIMIT the largest float whose correctly rounded square
root is at most the cap.  This pins that constant against the cap written in
the same source, with numpy's correctly rounded float32 square root.
"""
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(__file__), "..", "bevy_xpbd_amd", "csrc", "kernels.hip")


def _const(text, name):
    m = re.search(rf"constexpr float {name} = ([^;]+?)f;", text)
    assert m, name
    return np.float32(float.fromhex(m.group(1)) if "0x" in m.group(1) else float(m.group(1)))


def test_cap_sq_limit_is_the_last_float_whose_root_is_within_the_cap():
    text = open(SRC).read()
    cap = _const(text, "JACOBI_CORRECTION_CAP")
    limit = _const(text, "CAP_SQ_LIMIT")
    up = np.nextafter(limit, np.float32(np.inf))
    assert np.sqrt(limit, dtype=np.float32) <= cap
    assert np.sqrt(up, dtype=np.float32) > cap
    # and the equivalence on a dense sweep around the boundary
    bits = limit.view(np.uint32) + np.arange(-4096, 4097, dtype=np.int64)
    s = bits.astype(np.uint32).view(np.float32)
    assert np.array_equal(np.sqrt(s, dtype=np.float32) > cap, s > limit)
