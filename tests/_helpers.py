"""Shared helpers for the tests: parameter conversion between the product's
ClothSimUniforms/ClothSimConfig and the oracle's params, and error metrics."""
import numpy as np


def oracle_params(O, X, u, cfg, collision=False, coll_mode=None):
    p = O.Params()
    p.dt = u.dt
    p.inv_dt = u.inv_dt
    p.jacobi_omega = u.jacobi_omega
    p.inner_iterations = cfg.solve_inner_iterations
    p.gravity[:] = list(u.gravity[:3])
    p.grab_idx = u.grab_idx
    p.grab_active = u.grab_active
    p.grab_target[:] = list(u.grab_target[:3])
    p.grab_stiffness = u.grab_stiffness
    p.linear_drag_per_sec = u.linear_drag_per_sec
    p.thickness = u.thickness
    p.coll_scale = u.coll_scale
    p.collision = 1 if collision else 0
    p.solver = 1 if cfg.solver == X.SOLVER_GAUSS_SEIDEL else 0
    p.coll_mode = O.COLL_GRID if coll_mode is None else coll_mode
    return p


def tiny_mesh(pos, inv_mass, cons=(), rest_pos=None, ground=None):
    """A tiny mesh dict with a CSR built by cursor scatter (mesh_prep.rs:588-634)."""
    pos = np.asarray(pos, np.float32)
    n = len(pos)
    ci = np.array([c[0] for c in cons], np.uint32)
    cj = np.array([c[1] for c in cons], np.uint32)
    crest = np.array([c[2] for c in cons], np.float32)
    ccomp = np.array([c[3] for c in cons], np.float32)
    counts = np.zeros(n, np.int64)
    for a, b in zip(ci, cj):
        counts[a] += 1
        counts[b] += 1
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    other = np.zeros(offs[-1], np.uint32)
    eid = np.zeros(offs[-1], np.uint32)
    cur = offs[:-1].astype(np.int64).copy()
    for e, (a, b) in enumerate(zip(ci, cj)):
        other[cur[a]], eid[cur[a]] = b, e
        cur[a] += 1
        other[cur[b]], eid[cur[b]] = a, e
        cur[b] += 1
    return {"num_particles": n, "inv_mass": np.asarray(inv_mass, np.float32),
            "rest_positions": pos if rest_pos is None else np.asarray(rest_pos, np.float32),
            "positions": pos, "ground": np.zeros(n, np.uint32) if ground is None else np.asarray(ground, np.uint32), "constraint_i": ci, "constraint_j": cj,
            "constraint_rest": crest, "constraint_comp": ccomp, "neighbor_offsets": offs,
            "neighbor_other": other, "neighbor_constraint_id": eid}


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64)[:, :3] - np.asarray(b, np.float64)[:, :3])))


def bits_equal(a, b):
    a = np.ascontiguousarray(np.asarray(a)[:, :3], dtype=np.float32)
    b = np.ascontiguousarray(np.asarray(b)[:, :3], dtype=np.float32)
    return a.tobytes() == b.tobytes()


def first_diff(a, b):
    a = np.asarray(a)[:, :3]
    b = np.asarray(b)[:, :3]
    bad = np.nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1))[0] if a.dtype == np.float32 else []
    return bad[:5]


class DeviceBuffer:
    """Raw device memory through the HIP runtime (for the render-buffer
    boundary, which takes device pointers)."""
    _hip = None

    @classmethod
    def hip(cls):
        import ctypes as C
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        import ctypes as C
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), nbytes) == 0

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
