"""The C ABI surface (include/xpbd.h) without a GPU: the library loads,
exports every declared symbol, and rejects bad arguments with status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "xpbd.h")).read()
    return sorted(set(re.findall(r"XPBD_API\s+[\w\s\*]+?\b(xpbd_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(X):
    names = declared_symbols()
    assert len(names) >= 25
    L = X.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version(X):
    assert X.lib().xpbd_abi_version() == 1


def test_invalid_arguments_are_rejected(X):
    L = X.lib()
    h = C.c_void_p()
    assert L.xpbd_mesh_grid_cloth_hanging(0, 4, 0.1, C.byref(h)) == 1
    assert L.xpbd_mesh_grid_cloth_hanging(4, 4, -1.0, C.byref(h)) == 1
    d = X.SheetDesc.make(cols=1)
    assert L.xpbd_mesh_sheet(C.byref(d), C.byref(h)) == 1
    d = X.SheetDesc.make(pins=9)
    assert L.xpbd_mesh_sheet(C.byref(d), C.byref(h)) == 1
    assert L.xpbd_create(None, None, None, None, C.byref(h)) == 1
    with pytest.raises(X.XpbdError):
        X.ClothMeshData.grid_cloth_hanging(0, 0, 0.1)
    # config ranges are checked before any device call
    mesh = X.ClothMeshData.grid_cloth_hanging(4, 4, 0.1)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 8)
    for bad in (dict(solver=2), dict(gs_launch=4), dict(persistent_jacobi=3), dict(jacobi_block=96)):
        cfg = X.ClothSimConfig.default(**bad)
        sim = C.c_void_p()
        assert L.xpbd_create(mesh._h, C.byref(cfg), C.byref(u), None, C.byref(sim)) == X.XPBD_E_INVALID, bad


def test_residual_noalloc_refuses_oversize_before_any_device_call(X):
    """Argument checks come first: no HIP call is made for these."""
    L = X.lib()
    dummy = C.c_void_p(16)
    assert L.xpbd_residual_noalloc(dummy, 4, dummy, dummy, dummy, 0x80000001, None, dummy) == 1
    assert b"XPBD_RESIDUAL_NOALLOC_MAX_CONSTRAINTS" in L.xpbd_thread_last_error()
    assert L.xpbd_residual_noalloc(dummy, 4, None, dummy, dummy, 3, None, dummy) == 1
    assert L.xpbd_residual_noalloc(dummy, 4, dummy, dummy, dummy, 3, None, None) == 1


def test_mesh_read_rejects_short_buffer(X):
    m = X.ClothMeshData.grid_cloth_hanging(3, 3, 0.1)
    buf = np.zeros(2, np.float32)
    assert X.lib().xpbd_mesh_read(m._h, 0, buf.ctypes.data_as(C.c_void_p), buf.nbytes) == 1


def test_defaults_match_reference_constants(X):
    """ClothSimUniforms::default (cloth_compute.rs:204-243), dt = (1/60)/S."""
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 36)
    assert np.float32(u.dt) == np.float32(1.0) / np.float32(60.0) / np.float32(36)
    assert np.float32(u.jacobi_omega) == np.float32(0.32)
    assert np.float32(u.thickness) == np.float32(0.04) and np.float32(u.coll_scale) == np.float32(0.38)
    assert np.float32(u.linear_drag_per_sec) == np.float32(1.25) and u.grab_idx == -1
    assert np.float32(u.grab_stiffness) == np.float32(0.28)
    g = X.ClothSimUniforms.default(X.SOLVER_GAUSS_SEIDEL, 36)
    assert g.jacobi_omega == 1.0 and np.float32(g.grab_stiffness) == np.float32(0.45)
    c = X.ClothSimConfig.default()
    assert (c.solve_substeps, c.solve_inner_iterations, c.collision_every_n_substeps) == (36, 22, 1)


def _gpu_present():
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_create_without_device_reports_no_device(X):
    if _gpu_present():
        pytest.skip("a GPU is present")
    m = X.ClothMeshData.grid_cloth_hanging(3, 3, 0.1)
    with pytest.raises(X.XpbdError) as e:
        X.ClothComputePlugin(m)
    assert e.value.code == 6
