"""TEST INFRASTRUCTURE ONLY — ctypes loader for the CPU oracle (liboracle.so).

The oracle is a C++ restatement of the reference's CPU solver
(src/xpbd_cpu.rs) and topology code (src/mesh_prep.rs).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker; the product package (bevy_xpbd_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# Mesh field ids (xpbd_oracle.h OM_*), same numbering as include/xpbd.h.
FIELDS = {
    "positions": (0, np.float32, 3),
    "rest_positions": (1, np.float32, 3),
    "normals": (2, np.float32, 3),
    "uvs": (3, np.float32, 2),
    "indices": (4, np.uint32, 1),
    "tri_edge_neighbors": (5, np.uint32, 1),
    "inv_mass": (6, np.float32, 1),
    "neighbor_other": (7, np.uint32, 1),
    "neighbor_rest_len": (8, np.float32, 1),
    "neighbor_compliance": (9, np.float32, 1),
    "neighbor_constraint_id": (10, np.uint32, 1),
    "neighbor_offsets": (11, np.uint32, 1),
    "constraint_i": (12, np.uint32, 1),
    "constraint_j": (13, np.uint32, 1),
    "constraint_rest": (14, np.float32, 1),
    "constraint_comp": (15, np.float32, 1),
    "batch_offsets": (16, np.uint32, 1),
    "ground": (17, np.uint32, 1),
    "cloth_offsets": (18, np.uint32, 1),
}

LAYOUT_HANGING_XY, LAYOUT_HORIZONTAL_XZ = 0, 1
PIN_NONE, PIN_TWO_CORNERS, PIN_TOP_ROW, PIN_CENTER = 0, 1, 2, 3
SOLVER_JACOBI, SOLVER_GS = 0, 1
COLL_BRUTE, COLL_GRID = 0, 1


class SheetDesc(C.Structure):
    _fields_ = [
        ("cols", C.c_uint32), ("rows", C.c_uint32), ("spacing", C.c_float),
        ("layout", C.c_uint32), ("height", C.c_float), ("angle", C.c_float),
        ("pins", C.c_uint32), ("ground", C.c_uint32), ("inv_mass", C.c_float),
        ("stretch_compliance", C.c_float), ("shear_compliance", C.c_float),
        ("bend_compliance", C.c_float), ("jitter", C.c_float), ("seed", C.c_uint64),
    ]


class SolverIn(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("inv_mass", C.c_void_p), ("rest_pos", C.c_void_p),
        ("ground", C.c_void_p), ("e", C.c_uint32), ("ci", C.c_void_p), ("cj", C.c_void_p),
        ("crest", C.c_void_p), ("ccomp", C.c_void_p), ("noffs", C.c_void_p),
        ("nother", C.c_void_p), ("neid", C.c_void_p), ("num_cloths", C.c_uint32),
        ("cloth_offsets", C.c_void_p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("inv_dt", C.c_float), ("jacobi_omega", C.c_float),
        ("inner_iterations", C.c_uint32), ("gravity", C.c_float * 3),
        ("grab_idx", C.c_int32), ("grab_active", C.c_uint32), ("grab_target", C.c_float * 3),
        ("grab_stiffness", C.c_float), ("linear_drag_per_sec", C.c_float),
        ("thickness", C.c_float), ("coll_scale", C.c_float), ("collision", C.c_uint32),
        ("solver", C.c_uint32), ("coll_mode", C.c_uint32),
    ]


class Grid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("hi", C.c_float * 3), ("cell", C.c_float),
                ("inv_cell", C.c_float), ("dims", C.c_uint32 * 3)]


_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        L.om_grid_cloth_hanging.restype = vp
        L.om_grid_cloth_hanging.argtypes = [u32, u32, f32]
        L.om_parse_welded_obj.restype = vp
        L.om_parse_welded_obj.argtypes = [C.c_char_p]
        L.om_assemble.restype = vp
        L.om_assemble.argtypes = [u32, vp, vp, u32, vp, vp, vp, vp]
        L.om_sheet.restype = vp
        L.om_sheet.argtypes = [C.POINTER(SheetDesc)]
        L.om_batch.restype = vp
        L.om_batch.argtypes = [C.POINTER(SheetDesc), u32]
        L.om_count.restype = C.c_uint64
        L.om_count.argtypes = [vp, C.c_int]
        L.om_copy.argtypes = [vp, C.c_int, vp]
        L.om_free.argtypes = [vp]
        L.om_derive_collision_grid.argtypes = [vp, u32, f32, vp, vp, vp, vp, vp]
        for suf in ("f32", "f64"):
            getattr(L, f"oracle_substep_{suf}").argtypes = [C.POINTER(SolverIn), vp, vp, C.POINTER(Params)]
            getattr(L, f"oracle_frames_{suf}").argtypes = [C.POINTER(SolverIn), vp, vp, C.POINTER(Params), u32, u32, u32]
            getattr(L, f"oracle_residual_{suf}").argtypes = [C.POINTER(SolverIn), vp]
            getattr(L, f"oracle_residual_{suf}").restype = C.c_double
        L.oracle_self_collision_f32.argtypes = [C.POINTER(SolverIn), vp, f32, f32, u32]
        L.oracle_self_collision_f64.argtypes = [C.POINTER(SolverIn), vp, C.c_double, C.c_double, u32]
        L.oracle_cell_grid.argtypes = [vp, u32, f32, C.POINTER(Grid)]
        L.oracle_cell_keys.argtypes = [vp, u32, C.POINTER(Grid), vp]
        L.oracle_sort_pairs.argtypes = [vp, u32, vp, vp]
        L.oracle_cell_keys_batch.argtypes = [vp, vp, u32, f32, vp, vp]
        L.oracle_vertex_normals.argtypes = [vp, u32, vp, u32, vp]
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _read_mesh(handle) -> dict:
    if not handle:
        raise ValueError("oracle mesh builder rejected its arguments")
    L = lib()
    out = {}
    try:
        for name, (fid, dt, width) in FIELDS.items():
            cnt = int(L.om_count(handle, fid))
            a = np.empty(cnt, dtype=dt)
            if cnt:
                L.om_copy(handle, fid, _ptr(a))
            out[name] = a.reshape(-1, width) if width > 1 else a
    finally:
        L.om_free(handle)
    out["num_particles"] = len(out["inv_mass"])
    out["num_constraints"] = len(out["constraint_i"])
    return out


def grid_cloth_hanging(quad_cols: int, quad_rows: int, cell_size: float) -> dict:
    return _read_mesh(lib().om_grid_cloth_hanging(quad_cols, quad_rows, cell_size))


def parse_welded_obj(text: str) -> dict:
    return _read_mesh(lib().om_parse_welded_obj(text.encode()))


def assemble(positions, inv_mass, constraints) -> dict:
    """finalize_cloth_mesh's colouring, colour sort and CSR scatter applied to
    the rows `constraints` = [(i, j, rest, compliance), ...] (mesh_prep.rs:153-221,
    580-635)."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    w = np.ascontiguousarray(inv_mass, np.float32)
    cons = list(constraints)
    ci = np.array([c[0] for c in cons], np.uint32)
    cj = np.array([c[1] for c in cons], np.uint32)
    rest = np.array([c[2] for c in cons], np.float32)
    comp = np.array([c[3] for c in cons], np.float32)
    return _read_mesh(lib().om_assemble(len(pos), _ptr(pos), _ptr(w), len(cons), _ptr(ci), _ptr(cj),
                                        _ptr(rest), _ptr(comp)))


def sheet_desc(**kw) -> SheetDesc:
    d = SheetDesc()
    defaults = dict(cols=32, rows=32, spacing=1.0 / 31.0, layout=LAYOUT_HANGING_XY, height=0.0,
                    angle=0.0, pins=PIN_TWO_CORNERS, ground=0, inv_mass=1000.0,
                    stretch_compliance=0.0, shear_compliance=1e-6, bend_compliance=1e-4,
                    jitter=0.0, seed=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(d, k, v)
    return d


def sheet(**kw) -> dict:
    d = sheet_desc(**kw)
    return _read_mesh(lib().om_sheet(C.byref(d)))


def batch(descs) -> dict:
    arr = (SheetDesc * len(descs))(*descs)
    return _read_mesh(lib().om_batch(arr, len(descs)))


def derive_collision_grid(rest_vec4: np.ndarray, thickness: float):
    rest_vec4 = np.ascontiguousarray(rest_vec4, dtype=np.float32)
    origin = np.zeros(3, np.float32)
    inv = np.zeros(1, np.float32)
    dims = np.zeros(3, np.uint32)
    cells = np.zeros(1, np.uint32)
    digits = np.zeros(1, np.uint32)
    lib().om_derive_collision_grid(_ptr(rest_vec4), len(rest_vec4), thickness, _ptr(origin),
                                   _ptr(inv), _ptr(dims), _ptr(cells), _ptr(digits))
    return origin, float(inv[0]), dims, int(cells[0]), int(digits[0])


class Solver:
    """Holds the mesh arrays in the layout the oracle's substep functions take."""

    def __init__(self, mesh: dict):
        self.mesh = mesh
        n = mesh["num_particles"]
        self.arrays = {
            "inv_mass": np.ascontiguousarray(mesh["inv_mass"], np.float32),
            "rest": np.ascontiguousarray(mesh["rest_positions"], np.float32).reshape(-1),
            "ground": np.ascontiguousarray(mesh["ground"], np.uint32) if len(mesh["ground"]) == n else None,
            "ci": np.ascontiguousarray(mesh["constraint_i"], np.uint32),
            "cj": np.ascontiguousarray(mesh["constraint_j"], np.uint32),
            "crest": np.ascontiguousarray(mesh["constraint_rest"], np.float32),
            "ccomp": np.ascontiguousarray(mesh["constraint_comp"], np.float32),
            "noffs": np.ascontiguousarray(mesh["neighbor_offsets"], np.uint32),
            "nother": np.ascontiguousarray(mesh["neighbor_other"], np.uint32),
            "neid": np.ascontiguousarray(mesh["neighbor_constraint_id"], np.uint32),
            "coffs": np.ascontiguousarray(mesh.get("cloth_offsets", np.array([0, n])), np.uint32),
        }
        a = self.arrays
        nc = max(len(a["coffs"]) - 1, 1)
        self.sin = SolverIn(n, _ptr(a["inv_mass"]), _ptr(a["rest"]), _ptr(a["ground"]),
                            len(a["ci"]), _ptr(a["ci"]), _ptr(a["cj"]), _ptr(a["crest"]),
                            _ptr(a["ccomp"]), _ptr(a["noffs"]), _ptr(a["nother"]), _ptr(a["neid"]),
                            nc, _ptr(a["coffs"]) if nc > 1 else None)

    @staticmethod
    def _state(pos, vel, dtype):
        pos = np.array(np.asarray(pos)[:, :3], dtype=dtype, order="C", copy=True)
        vel = np.array(np.asarray(vel)[:, :3], dtype=dtype, order="C", copy=True)
        return pos, vel

    def substep(self, pos, vel, params: Params, dtype=np.float32):
        pos, vel = self._state(pos, vel, dtype)
        fn = lib().oracle_substep_f32 if dtype == np.float32 else lib().oracle_substep_f64
        fn(C.byref(self.sin), _ptr(pos), _ptr(vel), C.byref(params))
        return pos, vel

    def frames(self, pos, vel, params: Params, substeps: int, stride: int, frames: int, dtype=np.float32):
        pos, vel = self._state(pos, vel, dtype)
        fn = lib().oracle_frames_f32 if dtype == np.float32 else lib().oracle_frames_f64
        fn(C.byref(self.sin), _ptr(pos), _ptr(vel), C.byref(params), substeps, stride, frames)
        return pos, vel

    def self_collision(self, pos, thickness, coll_scale, mode=COLL_BRUTE, dtype=np.float32):
        pos = np.array(np.asarray(pos)[:, :3], dtype=dtype, order="C", copy=True)
        if dtype == np.float32:
            lib().oracle_self_collision_f32(C.byref(self.sin), _ptr(pos), thickness, coll_scale, mode)
        else:
            lib().oracle_self_collision_f64(C.byref(self.sin), _ptr(pos), thickness, coll_scale, mode)
        return pos

    def residual(self, pos) -> float:
        pos = np.asarray(pos)[:, :3]
        if pos.dtype == np.float64:
            p = np.ascontiguousarray(pos)
            return lib().oracle_residual_f64(C.byref(self.sin), _ptr(p))
        p = np.ascontiguousarray(pos, dtype=np.float32)
        return lib().oracle_residual_f32(C.byref(self.sin), _ptr(p))


def f32(x):
    return np.float32(x)


REFERENCE_FRAME_DELTA_SECS = f32(1.0) / f32(60.0)   # cloth_compute.rs:65


def params(substeps=36, inner_iterations=22, omega=0.32, gravity=(0.0, -9.81, 0.0),
           drag=1.25, thickness=0.04, coll_scale=0.38, collision=False, solver=SOLVER_JACOBI,
           coll_mode=COLL_BRUTE, grab_idx=-1, grab_active=False, grab_target=(0.0, 0.0, 0.0),
           grab_stiffness=0.28) -> Params:
    """ClothSimUniforms::default (cloth_compute.rs:204-243) with dt = (1/60)/S
    (cloth_compute.rs:438-439)."""
    dt = REFERENCE_FRAME_DELTA_SECS / f32(substeps)
    p = Params()
    p.dt = dt
    p.inv_dt = f32(1.0) / dt
    p.jacobi_omega = omega
    p.inner_iterations = inner_iterations
    p.gravity[:] = [float(f32(g)) for g in gravity]
    p.grab_idx = grab_idx
    p.grab_active = 1 if grab_active else 0
    p.grab_target[:] = list(grab_target)
    p.grab_stiffness = grab_stiffness
    p.linear_drag_per_sec = drag
    p.thickness = thickness
    p.coll_scale = coll_scale
    p.collision = 1 if collision else 0
    p.solver = solver
    p.coll_mode = coll_mode
    return p


def cell_grid(pos4: np.ndarray, cell_size: float) -> Grid:
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    g = Grid()
    lib().oracle_cell_grid(_ptr(pos4), len(pos4), cell_size, C.byref(g))
    return g


def cell_keys(pos4: np.ndarray, g: Grid) -> np.ndarray:
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    keys = np.empty(len(pos4), np.uint32)
    lib().oracle_cell_keys(_ptr(pos4), len(pos4), C.byref(g), _ptr(keys))
    return keys


def cell_keys_batch(pos4: np.ndarray, cloth_offsets: np.ndarray, cell_size: float):
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    offs = np.ascontiguousarray(cloth_offsets, dtype=np.uint32)
    nc = len(offs) - 1
    grids = (Grid * nc)()
    keys = np.empty(len(pos4), np.uint32)
    lib().oracle_cell_keys_batch(_ptr(pos4), _ptr(offs), nc, cell_size, grids, _ptr(keys))
    return list(grids), keys


def sort_pairs(keys: np.ndarray):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    sk = np.empty_like(keys)
    si = np.empty_like(keys)
    lib().oracle_sort_pairs(_ptr(keys), len(keys), _ptr(sk), _ptr(si))
    return sk, si


def vertex_normals(pos3: np.ndarray, tris: np.ndarray) -> np.ndarray:
    pos3 = np.ascontiguousarray(np.asarray(pos3)[:, :3], dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    out = np.empty_like(pos3)
    lib().oracle_vertex_normals(_ptr(pos3), len(pos3), _ptr(tris), len(tris) // 3, _ptr(out))
    return out


def set_threads(t: int):
    lib().oracle_set_threads(int(t))
