"""ctypes binding of include/xpbd.h and the plugin-shaped Python API above it.

Names follow the reference crate: ClothMeshData (mesh_prep.rs:103-131),
ClothSimConfig (cloth_compute.rs:142-179), ClothSimUniforms
(cloth_compute.rs:182-243), ClothSimControl (cloth_compute.rs:101-118),
ClothComputePlugin (cloth_compute.rs:443-479), ClothMaterialPlugin
(cloth_material.rs:41-48).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxpbd_hip.so")

XPBD_OK, XPBD_E_INVALID, XPBD_E_HIP, XPBD_E_NOMEM, XPBD_E_UNSUPPORTED, XPBD_E_STATE, XPBD_E_NO_DEVICE = range(7)
SOLVER_JACOBI, SOLVER_GAUSS_SEIDEL = 0, 1
LAYOUT_HANGING_XY, LAYOUT_HORIZONTAL_XZ = 0, 1
PIN_NONE, PIN_TWO_CORNERS, PIN_TOP_ROW, PIN_CENTER = 0, 1, 2, 3
KERNEL_JACOBI_ITER, KERNEL_COLLISION_SWEEP, KERNEL_RADIX_PASS, KERNEL_GS_SUBSTEP = 0, 1, 2, 3
KERNEL_COPY, KERNEL_JACOBI_PERSISTENT = 4, 5

_STATUS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "unsupported",
           5: "bad state", 6: "no HIP device"}

# (field id, dtype, row width) — XPBD_MESH_* in include/xpbd.h
MESH_FIELDS = {
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


class XpbdError(RuntimeError):
    def __init__(self, code: int, what: str, detail: str = ""):
        self.code = code
        super().__init__(f"{what}: {_STATUS.get(code, code)}" + (f" ({detail})" if detail else ""))


class SheetDesc(C.Structure):
    _fields_ = [
        ("cols", C.c_uint32), ("rows", C.c_uint32), ("spacing", C.c_float),
        ("layout", C.c_uint32), ("height", C.c_float), ("angle", C.c_float),
        ("pins", C.c_uint32), ("ground", C.c_uint32), ("inv_mass", C.c_float),
        ("stretch_compliance", C.c_float), ("shear_compliance", C.c_float),
        ("bend_compliance", C.c_float), ("jitter", C.c_float), ("seed", C.c_uint64),
    ]

    @classmethod
    def make(cls, **kw) -> "SheetDesc":
        d = cls()
        vals = dict(cols=32, rows=32, spacing=1.0 / 31.0, layout=LAYOUT_HANGING_XY, height=0.0, angle=0.0,
                    pins=PIN_TWO_CORNERS, ground=0, inv_mass=1000.0, stretch_compliance=0.0,
                    shear_compliance=1e-6, bend_compliance=1e-4, jitter=0.0, seed=0)
        vals.update(kw)
        for k, v in vals.items():
            setattr(d, k, v)
        return d


class ClothSimConfig(C.Structure):
    """Scalar fields of ClothSimConfig (cloth_compute.rs:142-179)."""
    _fields_ = [
        ("solve_substeps", C.c_uint32), ("solve_inner_iterations", C.c_uint32),
        ("collision_every_n_substeps", C.c_uint32), ("solver", C.c_uint32),
        ("collision_cell", C.c_float), ("use_graph", C.c_uint32), ("autotune", C.c_uint32),
        ("jacobi_block", C.c_uint32), ("fused_predict", C.c_uint32), ("plain_adjacency", C.c_uint32),
        ("fast_division", C.c_uint32), ("xcd_remap", C.c_uint32), ("persistent_jacobi", C.c_uint32),
    ]

    @classmethod
    def default(cls, **kw) -> "ClothSimConfig":
        c = cls()
        lib().xpbd_sim_config_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class ClothSimUniforms(C.Structure):
    """ClothSimUniforms (cloth_compute.rs:182-202)."""
    _fields_ = [
        ("dt", C.c_float), ("inv_dt", C.c_float), ("num_particles", C.c_uint32), ("num_tris", C.c_uint32),
        ("jacobi_omega", C.c_float), ("inner_iterations", C.c_uint32), ("thickness", C.c_float),
        ("coll_scale", C.c_float), ("gravity", C.c_float * 4), ("grab_target", C.c_float * 4),
        ("grab_idx", C.c_int32), ("grab_active", C.c_uint32), ("grab_stiffness", C.c_float),
        ("linear_drag_per_sec", C.c_float), ("constraint_batch_idx", C.c_uint32),
    ]

    @classmethod
    def default(cls, solver: int = SOLVER_JACOBI, substeps: int = 36, **kw) -> "ClothSimUniforms":
        u = cls()
        lib().xpbd_sim_uniforms_default(solver, substeps, C.byref(u))
        for k, v in kw.items():
            if k in ("gravity", "grab_target"):
                getattr(u, k)[:] = list(v) + [0.0] * (4 - len(v))
            else:
                setattr(u, k, v)
        return u


class CollGrid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("inv_cell", C.c_float), ("dims", C.c_uint32 * 3),
                ("num_cells", C.c_uint32), ("radix_digits", C.c_uint32)]


_lib = None


def lib():
    """Load libxpbd_hip.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m bevy_xpbd_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, u32, i32, f32, u64 = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_uint64
        pp = C.POINTER(C.c_void_p)
        sig = {
            "xpbd_abi_version": (i32, []),
            "xpbd_mesh_grid_cloth_hanging": (i32, [u32, u32, f32, pp]),
            "xpbd_mesh_parse_welded_obj": (i32, [C.c_char_p, C.c_size_t, pp]),
            "xpbd_mesh_sheet": (i32, [C.POINTER(SheetDesc), pp]),
            "xpbd_mesh_batch": (i32, [C.POINTER(SheetDesc), u32, pp]),
            "xpbd_mesh_count": (u64, [vp, i32]),
            "xpbd_mesh_read": (i32, [vp, i32, vp, u64]),
            "xpbd_mesh_destroy": (None, [vp]),
            "xpbd_derive_collision_grid": (i32, [vp, u32, f32, C.POINTER(CollGrid)]),
            "xpbd_sim_uniforms_default": (None, [u32, u32, C.POINTER(ClothSimUniforms)]),
            "xpbd_sim_config_default": (None, [C.POINTER(ClothSimConfig)]),
            "xpbd_create": (i32, [vp, C.POINTER(ClothSimConfig), C.POINTER(ClothSimUniforms), vp, pp]),
            "xpbd_destroy": (None, [vp]),
            "xpbd_last_error": (C.c_char_p, [vp]),
            "xpbd_set_uniforms": (i32, [vp, C.POINTER(ClothSimUniforms)]),
            "xpbd_set_config": (i32, [vp, C.POINTER(ClothSimConfig)]),
            "xpbd_step_frames": (i32, [vp, u32]),
            "xpbd_step_substeps": (i32, [vp, u32, u32]),
            "xpbd_sync": (i32, [vp]),
            "xpbd_stream": (vp, [vp]),
            "xpbd_reset": (i32, [vp]),
            "xpbd_read_state": (i32, [vp, vp, vp]),
            "xpbd_write_state": (i32, [vp, vp, vp]),
            "xpbd_publish": (i32, [vp, vp, vp]),
            "xpbd_bind_render_buffers": (i32, [vp, vp, vp]),
            "xpbd_residual": (i32, [vp, C.POINTER(C.c_double)]),
            "xpbd_time_kernel": (i32, [vp, i32, u32, C.POINTER(C.c_float)]),
            "xpbd_profile_frames": (i32, [vp, u32, i32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
            "xpbd_debug_collision_keys": (i32, [vp, f32, vp, vp, vp]),
            "xpbd_selftest_division": (i32, [i32, i32, u64, u64, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(code: int, what: str, handle=None):
    if code != XPBD_OK:
        detail = ""
        if handle:
            msg = lib().xpbd_last_error(handle)
            detail = msg.decode() if msg else ""
        raise XpbdError(code, what, detail)


class ClothMeshData:
    """ClothMeshData (mesh_prep.rs:103-131), built by the library's host code."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.xpbd_mesh_destroy(h)
            self._h = None

    @classmethod
    def _build(cls, fn, *args) -> "ClothMeshData":
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)), fn.__name__)
        return cls(h)

    @classmethod
    def grid_cloth_hanging(cls, quad_cols: int, quad_rows: int, cell_size: float) -> "ClothMeshData":
        return cls._build(lib().xpbd_mesh_grid_cloth_hanging, quad_cols, quad_rows, cell_size)

    @classmethod
    def parse_welded_obj(cls, text: str) -> "ClothMeshData":
        b = text.encode()
        return cls._build(lib().xpbd_mesh_parse_welded_obj, b, len(b))

    @classmethod
    def sheet(cls, **kw) -> "ClothMeshData":
        d = SheetDesc.make(**kw)
        return cls._build(lib().xpbd_mesh_sheet, C.byref(d))

    @classmethod
    def batch(cls, descs) -> "ClothMeshData":
        arr = (SheetDesc * len(descs))(*descs)
        return cls._build(lib().xpbd_mesh_batch, arr, len(descs))

    def field(self, name: str) -> np.ndarray:
        fid, dt, width = MESH_FIELDS[name]
        cnt = int(lib().xpbd_mesh_count(self._h, fid))
        a = np.empty(cnt, dtype=dt)
        if cnt:
            _check(lib().xpbd_mesh_read(self._h, fid, _ptr(a), a.nbytes), "xpbd_mesh_read")
        return a.reshape(-1, width) if width > 1 else a

    def fields(self) -> dict:
        out = {k: self.field(k) for k in MESH_FIELDS}
        out["num_particles"] = len(out["inv_mass"])
        out["num_constraints"] = len(out["constraint_i"])
        return out

    @property
    def num_particles(self) -> int:
        return int(lib().xpbd_mesh_count(self._h, MESH_FIELDS["inv_mass"][0]))

    @property
    def num_constraints(self) -> int:
        return int(lib().xpbd_mesh_count(self._h, MESH_FIELDS["constraint_i"][0]))

    def to_sim_config(self, config: ClothSimConfig | None = None,
                      uniforms: ClothSimUniforms | None = None) -> "ClothComputePlugin":
        """ClothMeshData::to_sim_config + init_cloth_sim (cloth_compute.rs:1599, 571)."""
        return ClothComputePlugin(self, config, uniforms)


def derive_collision_grid(rest_vec4: np.ndarray, thickness: float):
    """derive_collision_grid (mesh_prep.rs:16-85): origin, inv_cell, dims, num_cells, radix_digits."""
    rest_vec4 = np.ascontiguousarray(rest_vec4, dtype=np.float32)
    g = CollGrid()
    _check(lib().xpbd_derive_collision_grid(_ptr(rest_vec4), len(rest_vec4), thickness, C.byref(g)),
           "xpbd_derive_collision_grid")
    return (np.array(g.origin[:], np.float32), np.float32(g.inv_cell), np.array(g.dims[:], np.uint32),
            int(g.num_cells), int(g.radix_digits))


class ClothSimControl:
    """ClothSimControl (cloth_compute.rs:101-118): pause / single step / reinit serials."""

    def __init__(self):
        self.sim_paused = False
        self.step_serial = 0
        self.reinit_serial = 0


class ClothComputePlugin:
    """One simulated cloth (or batch) on the GPU: ClothComputePlugin's render-
    world systems (cloth_compute.rs:443-479) as explicit calls."""

    def __init__(self, mesh: ClothMeshData, config: ClothSimConfig | None = None,
                 uniforms: ClothSimUniforms | None = None, stream: int | None = None):
        self.config = config if config is not None else ClothSimConfig.default()
        self.uniforms = uniforms if uniforms is not None else ClothSimUniforms.default(
            self.config.solver, self.config.solve_substeps)
        self.n = mesh.num_particles
        self.control = ClothSimControl()
        self._acked_step = 0
        self._acked_reinit = 0
        h = C.c_void_p()
        code = lib().xpbd_create(mesh._h, C.byref(self.config), C.byref(self.uniforms),
                                 C.c_void_p(stream) if stream else None, C.byref(h))
        _check(code, "xpbd_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().xpbd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- extraction of changed resources
    def set_uniforms(self, u: ClothSimUniforms):
        self.uniforms = u
        _check(lib().xpbd_set_uniforms(self._h, C.byref(u)), "xpbd_set_uniforms", self._h)

    def set_config(self, c: ClothSimConfig):
        self.config = c
        _check(lib().xpbd_set_config(self._h, C.byref(c)), "xpbd_set_config", self._h)

    # -- run_cloth_sim
    def step_frames(self, frames: int = 1):
        _check(lib().xpbd_step_frames(self._h, frames), "xpbd_step_frames", self._h)

    def step_substeps(self, count: int = 1, collision: bool = False):
        _check(lib().xpbd_step_substeps(self._h, count, 1 if collision else 0), "xpbd_step_substeps", self._h)

    def run_cloth_sim(self):
        """One render frame honouring ClothSimControl (cloth_compute.rs:1197-1221, 509-569)."""
        c = self.control
        if c.reinit_serial != self._acked_reinit:
            self._acked_reinit = c.reinit_serial
            self.reset()
        if c.sim_paused and c.step_serial == self._acked_step:
            return False
        self.step_frames(1)
        self._acked_step = c.step_serial
        return True

    def sync(self):
        _check(lib().xpbd_sync(self._h), "xpbd_sync", self._h)

    @property
    def stream(self) -> int:
        return lib().xpbd_stream(self._h) or 0

    def reset(self):
        _check(lib().xpbd_reset(self._h), "xpbd_reset", self._h)

    def read_state(self):
        pos = np.zeros((self.n, 4), np.float32)
        vel = np.zeros((self.n, 4), np.float32)
        _check(lib().xpbd_read_state(self._h, _ptr(pos), _ptr(vel)), "xpbd_read_state", self._h)
        return pos, vel

    def write_state(self, pos, vel=None):
        pos4 = np.zeros((self.n, 4), np.float32)
        pos4[:, :3] = np.asarray(pos)[:, :3]
        vel4 = None
        if vel is not None:
            vel4 = np.zeros((self.n, 4), np.float32)
            vel4[:, :3] = np.asarray(vel)[:, :3]
        _check(lib().xpbd_write_state(self._h, _ptr(pos4), _ptr(vel4)), "xpbd_write_state", self._h)

    def publish(self, dst_pos_ptr: int | None, dst_nrm_ptr: int | None):
        _check(lib().xpbd_publish(self._h, dst_pos_ptr, dst_nrm_ptr), "xpbd_publish", self._h)

    def bind_render_buffers(self, dst_pos_ptr: int | None, dst_nrm_ptr: int | None):
        _check(lib().xpbd_bind_render_buffers(self._h, dst_pos_ptr, dst_nrm_ptr), "xpbd_bind_render_buffers",
               self._h)

    def residual(self) -> float:
        r = C.c_double()
        _check(lib().xpbd_residual(self._h, C.byref(r)), "xpbd_residual", self._h)
        return r.value

    def time_kernel(self, which: int, reps: int = 20) -> float:
        ms = C.c_float()
        _check(lib().xpbd_time_kernel(self._h, which, reps, C.byref(ms)), "xpbd_time_kernel", self._h)
        return ms.value

    def profile_frames(self, frames: int, which: int):
        """Advance `frames` frames eagerly with HIP events around every launch
        of kernel kind `which`; returns (mean duration in ms, launches)."""
        ms = C.c_float()
        cnt = C.c_uint32()
        _check(lib().xpbd_profile_frames(self._h, frames, which, C.byref(ms), C.byref(cnt)), "xpbd_profile_frames",
               self._h)
        return ms.value, cnt.value

    def debug_collision_keys(self, cell_size: float, num_cloths: int = 1):
        grid = np.zeros((num_cloths, 12), np.uint32)
        keys = np.zeros(self.n, np.uint32)
        ids = np.zeros(self.n, np.uint32)
        _check(lib().xpbd_debug_collision_keys(self._h, cell_size, _ptr(grid), _ptr(keys), _ptr(ids)),
               "xpbd_debug_collision_keys", self._h)
        gs = [{"lo": g[0:3].view(np.float32).copy(), "hi": g[3:6].view(np.float32).copy(),
               "cell": g[6:7].view(np.float32)[0], "inv_cell": g[7:8].view(np.float32)[0],
               "dims": g[8:11].copy(), "key_base": int(g[11])} for g in grid]
        return (gs[0] if num_cloths == 1 else gs), keys, ids


def selftest_division(exp_lo: int = -60, exp_hi: int = 60, pairs: int = 1 << 28, seed: int = 1) -> dict:
    """Run the device self-test of the Jacobi gather's division (needs a GPU)."""
    out = np.zeros(5, np.uint32)
    _check(lib().xpbd_selftest_division(exp_lo, exp_hi, pairs, seed, _ptr(out)), "xpbd_selftest_division")
    return {"rcp_mismatches": int(out[0]), "first_bad_b": int(out[1]), "div_mismatches": int(out[2]),
            "first_bad_pair": (int(out[3]), int(out[4]))}


class ClothMaterialPlugin:
    """ClothMaterialPlugin / ClothMatExt (cloth_material.rs:14-48): the render
    side reads sim_positions (binding 101) and sim_normals (binding 102) as
    array<vec4<f32>>, w = 0, indexed by particle id.  This class only states
    that layout; the compute plugin writes it with publish()."""
    POSITION_BINDING = 101
    NORMAL_BINDING = 102
    ELEMENT_BYTES = 16

    @staticmethod
    def buffer_bytes(num_particles: int) -> int:
        return 16 * num_particles
