"""xpbd_residual_noalloc: the relative residual over caller-owned device
arrays, with its block partials in the library's fixed device array instead
of allocated memory.  Needs an MI355X.

Checked against xpbd_residual of a handle holding the same state, on two
calls in a row, the second on a smaller mesh (the fixed array keeps the
first call's partials; the second must not see them).  Tolerance: both are
f64 sums of the same terms in different orders (the handle sums rest^2 on
the host in row order and C^2 over its own constraint order), so they agree
to a few units of E * 2^-53 relative; 1e-10 is stated for E <= 2e4.
"""
import ctypes as C
import math

import numpy as np
import pytest

from _helpers import DeviceBuffer

pytestmark = pytest.mark.gpu


def _state(X, cols, rows, seed):
    mesh = X.ClothMeshData.sheet(cols=cols, rows=rows, spacing=0.02, pins=2, jitter=1e-3, seed=seed)
    f = mesh.fields()
    cfg = X.ClothSimConfig.default(solve_substeps=4, solve_inner_iterations=3, autotune=0)
    sim = X.ClothComputePlugin(mesh, cfg, X.ClothSimUniforms.default(0, 4))
    sim.step_frames(3)
    r = sim.residual()
    pos, _ = sim.read_state()
    sim.close()
    return f, pos, r


def _noalloc(X, pos, f, e=None, pad_index=None):
    n = len(pos)
    # one vec4 of padding past the positions: the index test below names
    # particle n, which the kernel must not read; were it read, it stays
    # inside this allocation
    bufs = {"pos": DeviceBuffer(16 * (n + 1)), "ci": DeviceBuffer(4 * len(f["constraint_i"])),
            "cj": DeviceBuffer(4 * len(f["constraint_j"])), "rest": DeviceBuffer(4 * len(f["constraint_rest"])),
            "out": DeviceBuffer(8)}
    p = np.zeros((n + 1, 4), np.float32)
    p[:n, :3] = pos[:, :3]
    bufs["pos"].upload(p)
    ci = np.ascontiguousarray(f["constraint_i"], np.uint32).copy()
    if pad_index is not None:
        ci[pad_index] = n
    bufs["ci"].upload(ci)
    bufs["cj"].upload(np.ascontiguousarray(f["constraint_j"], np.uint32))
    bufs["rest"].upload(np.ascontiguousarray(f["constraint_rest"], np.float32))
    bufs["out"].upload(np.array([-1.0], np.float64))
    code = X.lib().xpbd_residual_noalloc(bufs["pos"].ptr, n, bufs["ci"].ptr, bufs["cj"].ptr, bufs["rest"].ptr,
                                         len(ci) if e is None else e, None, bufs["out"].ptr)
    out = bufs["out"].download(np.float64, (1,))[0] if code == 0 else None
    for b in bufs.values():
        b.free()
    return code, out


def test_residual_noalloc_matches_xpbd_residual_twice(X):
    for cols, rows, seed in ((96, 80, 3), (40, 24, 4)):
        f, pos, r_ref = _state(X, cols, rows, seed)
        code, r = _noalloc(X, pos, f)
        assert code == 0, X.lib().xpbd_thread_last_error()
        print(f"noalloc {cols}x{rows}: {r!r} vs xpbd_residual {r_ref!r}", flush=True)
        assert r_ref > 1e-6
        assert abs(r - r_ref) <= 1e-10 * r_ref


def test_residual_noalloc_refuses_oversize_and_flags_bad_indices(X):
    f, pos, r_ref = _state(X, 24, 16, 5)
    code, _ = _noalloc(X, pos, f, e=0x80000001)
    assert code == X.XPBD_E_INVALID
    assert b"XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS" in X.lib().xpbd_thread_last_error()
    code, r = _noalloc(X, pos, f, pad_index=7)
    assert code == 0 and math.isnan(r)
    code, r = _noalloc(X, pos, f)           # and the next good call is unaffected
    assert code == 0 and abs(r - r_ref) <= 1e-10 * r_ref
