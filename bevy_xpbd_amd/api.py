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
    "cloth_cols": (19, np.uint32, 1),
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


class MeshDesc(C.Structure):
    """xpbd_mesh_desc: ClothSimConfig's topology arrays (cloth_compute.rs:142-179)."""
    _fields_ = [
        ("num_particles", C.c_uint32), ("num_tris", C.c_uint32), ("num_distance_constraints", C.c_uint32),
        ("constraint_batch_count", C.c_uint32), ("initial_pos", C.c_void_p), ("rest_pos", C.c_void_p),
        ("initial_normals", C.c_void_p), ("inv_mass", C.c_void_p), ("constraint_i", C.c_void_p),
        ("constraint_j", C.c_void_p), ("constraint_rest_len", C.c_void_p), ("constraint_compliance", C.c_void_p),
        ("constraint_batch_offsets", C.c_void_p), ("tri_indices", C.c_void_p), ("neighbor_offsets", C.c_void_p),
        ("neighbor_packed", C.c_void_p), ("ground", C.c_void_p), ("device_prepare", C.c_uint32),
        ("num_cloths", C.c_uint32),
        ("cloth_offsets", C.c_void_p),
    ]


class ClothSimConfig(C.Structure):
    """Scalar fields of ClothSimConfig (cloth_compute.rs:142-179)."""
    _fields_ = [
        ("solve_substeps", C.c_uint32), ("solve_inner_iterations", C.c_uint32),
        ("collision_every_n_substeps", C.c_uint32), ("solver", C.c_uint32),
        ("collision_cell", C.c_float), ("use_graph", C.c_uint32), ("autotune", C.c_uint32),
        ("jacobi_block", C.c_uint32), ("fused_predict", C.c_uint32), ("plain_adjacency", C.c_uint32),
        ("fast_division", C.c_uint32), ("xcd_remap", C.c_uint32), ("persistent_jacobi", C.c_uint32),
        ("grid_tiles", C.c_uint32),
        ("gs_launch", C.c_uint32),
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
            "xpbd_mesh_from_arrays": (i32, [C.POINTER(MeshDesc), pp]),
            "xpbd_thread_last_error": (C.c_char_p, []),
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
            "xpbd_import_render_buffers_fd": (i32, [vp, i32, u64, u64, i32, u64, u64]),
            "xpbd_residual": (i32, [vp, C.POINTER(C.c_double)]),
            "xpbd_residual_noalloc": (i32, [vp, u32, vp, vp, vp, u32, vp, vp]),
            "xpbd_time_kernel": (i32, [vp, i32, u32, C.POINTER(C.c_float)]),
            "xpbd_profile_frames": (i32, [vp, u32, i32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
            "xpbd_debug_collision_keys": (i32, [vp, f32, vp, vp, vp]),
            "xpbd_debug_raise_device_error": (i32, [vp]),
            "xpbd_debug_gs_tiles": (i32, [vp, u32, u32, vp]),
            "xpbd_debug_gs_grid": (i32, [vp, u32, vp]),
            "xpbd_debug_jacobi_resident": (i32, [vp, u32, vp]),
            "xpbd_get_layout": (i32, [vp, vp]),
            "xpbd_selftest_division": (i32, [i32, i32, u64, u64, vp]),
            "xpbd_selftest_sqrt": (i32, [i32, i32, vp]),
            "xpbd_mlp_create": (i32, [u64, pp]),
            "xpbd_mlp_destroy": (None, [vp]),
            "xpbd_mlp_last_error": (C.c_char_p, [vp]),
            "xpbd_mlp_get_params": (i32, [vp, vp, C.POINTER(C.c_float)]),
            "xpbd_mlp_set_params": (i32, [vp, vp, f32]),
            "xpbd_mlp_forward": (i32, [vp, vp, u32, vp]),
            "xpbd_mlp_gradient": (i32, [vp, vp, vp, u32, vp, C.POINTER(C.c_float)]),
            "xpbd_mlp_train": (i32, [vp, vp, vp, u32, u32, u32, f32, vp]),
            "xpbd_warm_sample": (i32, [vp, u32, vp, vp]),
            "xpbd_warm_features": (i32, [vp, vp]),
            "xpbd_set_warm_start": (i32, [vp, vp, u32, f32]),
            "xpbd_warm_stats": (i32, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), u32]),
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

    @classmethod
    def from_arrays(cls, initial_pos, inv_mass, constraint_i=(), constraint_j=(), constraint_rest_len=(),
                    constraint_compliance=(), rest_pos=None, initial_normals=None, tri_indices=None,
                    constraint_batch_offsets=None, neighbor_offsets=None, neighbor_packed=None,
                    ground=None, cloth_offsets=None, device_prepare=False) -> "ClothMeshData":
        """A caller's topology in ClothSimConfig's arrays (cloth_compute.rs:142-179):
        positions as (N, 3) or (N, 4) rows, constraint rows, optional colour
        batch offsets (None: coloured here as finalize_cloth_mesh does) and
        optional Jacobi adjacency (neighbor_offsets + neighbor_packed, the
        vec4 entries of pack_neighbor_gpu; None: built by cursor scatter);
        cloth_offsets splits the particles into cloths that collide only
        within themselves (None: one cloth); device_prepare runs the
        colouring, colour sort and adjacency scatter on the GPU (same
        results)."""
        def vec4(a):
            if a is None:
                return None
            a = np.asarray(a, np.float32)
            out = np.zeros((len(a), 4), np.float32)
            out[:, :a.shape[1]] = a[:, :4] if a.shape[1] >= 4 else a
            if a.shape[1] >= 4:
                out[:, 3] = 0.0
            return out

        def arr(a, dt):
            return None if a is None else np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=dt)

        keep = {
            "initial_pos": vec4(initial_pos), "rest_pos": vec4(rest_pos), "initial_normals": vec4(initial_normals),
            "inv_mass": arr(inv_mass, np.float32), "constraint_i": arr(constraint_i, np.uint32),
            "constraint_j": arr(constraint_j, np.uint32), "constraint_rest_len": arr(constraint_rest_len, np.float32),
            "constraint_compliance": arr(constraint_compliance, np.float32),
            "constraint_batch_offsets": arr(constraint_batch_offsets, np.uint32),
            "tri_indices": arr(tri_indices, np.uint32), "neighbor_offsets": arr(neighbor_offsets, np.uint32),
            "neighbor_packed": arr(neighbor_packed, np.float32), "ground": arr(ground, np.uint32),
            "cloth_offsets": arr(cloth_offsets, np.uint32),
        }
        d = MeshDesc()
        d.num_particles = len(keep["initial_pos"])
        d.num_tris = 0 if keep["tri_indices"] is None else len(keep["tri_indices"]) // 3
        d.num_distance_constraints = len(keep["constraint_i"])
        bo = keep["constraint_batch_offsets"]
        d.constraint_batch_count = 0 if bo is None else max(len(bo) - 1, 0)
        co = keep["cloth_offsets"]
        d.num_cloths = 0 if co is None else max(len(co) - 1, 0)
        d.device_prepare = 1 if device_prepare else 0
        for k, v in keep.items():
            setattr(d, k, None if v is None or v.size == 0 else v.ctypes.data)
        h = C.c_void_p()
        code = lib().xpbd_mesh_from_arrays(C.byref(d), C.byref(h))
        if code != XPBD_OK:
            raise XpbdError(code, "xpbd_mesh_from_arrays", lib().xpbd_thread_last_error().decode())
        return cls(h)

    def sim_config_arrays(self) -> dict:
        """The topology arrays of this mesh as ClothSimConfig carries them
        (cloth_compute.rs:142-179, 1599-1700): vec4 positions and the packed
        Jacobi adjacency of pack_neighbor_gpu (cloth_jacobi.rs:95-111)."""
        f = self.fields()
        n = f["num_particles"]

        def vec4(a):
            out = np.zeros((n, 4), np.float32)
            out[:, :3] = a
            return out

        packed = np.stack([f["neighbor_other"].astype(np.float32), f["neighbor_rest_len"],
                           f["neighbor_compliance"], f["neighbor_constraint_id"].astype(np.float32)], axis=1)
        return {"initial_pos": vec4(f["positions"]), "rest_pos": vec4(f["rest_positions"]),
                "initial_normals": vec4(f["normals"]) if len(f["normals"]) == n else None,
                "inv_mass": f["inv_mass"], "constraint_i": f["constraint_i"], "constraint_j": f["constraint_j"],
                "constraint_rest_len": f["constraint_rest"], "constraint_compliance": f["constraint_comp"],
                "constraint_batch_offsets": f["batch_offsets"], "tri_indices": f["indices"],
                "neighbor_offsets": f["neighbor_offsets"], "neighbor_packed": packed, "ground": f["ground"],
                "cloth_offsets": f["cloth_offsets"]}

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

    def import_render_buffers_fd(self, pos_fd: int, pos_size: int, pos_offset: int = 0, nrm_fd: int = -1,
                                 nrm_size: int = 0, nrm_offset: int = 0):
        """Zero-copy render buffers from exported device memory (POSIX fds,
        e.g. Vulkan opaque fds of the renderer's position / normal buffers);
        the library takes the descriptors over on success."""
        _check(lib().xpbd_import_render_buffers_fd(self._h, int(pos_fd), int(pos_size), int(pos_offset), int(nrm_fd),
                                                   int(nrm_size), int(nrm_offset)),
               "xpbd_import_render_buffers_fd", self._h)

    def raise_device_error(self):
        """Test hook: set the device error word, as an expired barrier wait does."""
        _check(lib().xpbd_debug_raise_device_error(self._h), "xpbd_debug_raise_device_error", self._h)

    def gs_tiles(self, tile: int = 0, blocks: int = 0):
        """Test/tuning hook: the per-tile-flag GS plan (gs_launch = 3), re-planned
        with `tile` particles per tile on `blocks` blocks when tile > 0.  Returns
        (particles per tile, tiles, blocks, most dependencies of one tile); the
        flag form then runs these index-range tiles."""
        out = (C.c_uint32 * 4)()
        _check(lib().xpbd_debug_gs_tiles(self._h, int(tile), int(blocks), out), "xpbd_debug_gs_tiles", self._h)
        return tuple(int(v) for v in out)

    def gs_grid(self, rows: int = 0):
        """Test/tuning hook: the flag form's LDS grid tiles (one grid cloth,
        64 columns x rows), re-planned at `rows` rows and switched on when
        rows > 0.  Returns (rows per tile, tiles, threads per block, in use)."""
        out = (C.c_uint32 * 4)()
        _check(lib().xpbd_debug_gs_grid(self._h, int(rows), out), "xpbd_debug_gs_grid", self._h)
        return tuple(int(v) for v in out)

    def jacobi_resident(self, rows: int = 0):
        """Test/tuning hook: the LDS-resident Jacobi tiles (persistent_jacobi = 2),
        re-planned at `rows` rows per tile when rows > 0.  Returns (rows per
        tile, tiles, pairs staged in LDS, in use)."""
        out = (C.c_uint32 * 4)()
        _check(lib().xpbd_debug_jacobi_resident(self._h, int(rows), out), "xpbd_debug_jacobi_resident", self._h)
        return tuple(int(v) for v in out)

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

    def warm_sample(self, iterations: int, features: DeviceArray, targets: DeviceArray):
        """Training pairs of the current state: features (N x 32 bf16) and
        (iterated - predicted) / spacing (N x 3 f32), both on the device."""
        _check(lib().xpbd_warm_sample(self._h, iterations, features.ptr, targets.ptr), "xpbd_warm_sample", self._h)

    def warm_features(self, features: DeviceArray):
        _check(lib().xpbd_warm_features(self._h, features.ptr), "xpbd_warm_features", self._h)

    def set_warm_start(self, mlp, warm_iterations: int = 2, residual_threshold: float = 1e-4):
        """Attach (mlp) or detach (None) the learned warm start."""
        self._warm = mlp   # keep the network alive while attached
        _check(lib().xpbd_set_warm_start(self._h, mlp._h if mlp is not None else None, warm_iterations,
                                         residual_threshold), "xpbd_set_warm_start", self._h)

    def warm_stats(self, reset: bool = False):
        sub, fb = C.c_uint32(), C.c_uint32()
        _check(lib().xpbd_warm_stats(self._h, C.byref(sub), C.byref(fb), 1 if reset else 0), "xpbd_warm_stats",
               self._h)
        return sub.value, fb.value

    def layout(self) -> dict:
        """Adjacency layout in use: "plain", "packed", "grid_tiles" or
        "shared_tiles" (grid tiles with each constraint evaluated once per
        tile), with the tile grid's columns and rows, and whether the fast
        correctly rounded division is active."""
        out = (C.c_uint32 * 4)()
        _check(lib().xpbd_get_layout(self._h, out), "xpbd_get_layout", self._h)
        return {"adjacency": ("plain", "packed", "grid_tiles", "shared_tiles")[out[0]], "tile_cols": int(out[1]),
                "tile_rows": int(out[2]), "fast_division": bool(out[3])}

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


class DeviceArray:
    """Device memory from the HIP runtime (hipMalloc via ctypes) with numpy
    upload / download; the boundary takes device buffers as plain pointers."""
    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMalloc.restype = C.c_int
            h.hipFree.argtypes = [C.c_void_p]
            h.hipFree.restype = C.c_int
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemcpy.restype = C.c_int
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            h.hipMemset.restype = C.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = self._rt().hipMalloc(C.byref(p), max(self.nbytes, 256))
        if rc != 0:
            raise XpbdError(XPBD_E_NOMEM, "hipMalloc", f"hip error {rc}")
        self.ptr = p.value
        self._rt().hipMemset(C.c_void_p(self.ptr), 0, max(self.nbytes, 256))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a, offset_bytes: int = 0):
        a = np.ascontiguousarray(a)
        if offset_bytes + a.nbytes > self.nbytes:
            raise ValueError("upload beyond the device array")
        rc = self._rt().hipMemcpy(C.c_void_p(self.ptr + offset_bytes), a.ctypes.data_as(C.c_void_p), a.nbytes, 1)
        if rc != 0:
            raise XpbdError(XPBD_E_HIP, "hipMemcpy", f"hip error {rc}")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        rc = self._rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2)
        if rc != 0:
            raise XpbdError(XPBD_E_HIP, "hipMemcpy", f"hip error {rc}")
        return out

    def at(self, offset_bytes: int) -> int:
        return self.ptr + offset_bytes

    def free(self):
        if getattr(self, "ptr", None):
            self._rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


MLP_PARAMS = 7312
MLP_IN = 32


class WarmStartMLP:
    """The learned warm start's network (30 -> 64 -> 64 -> 3, GELU; BF16
    MFMA forward / backward with FP32 accumulate, Adam) on the GPU."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(lib().xpbd_mlp_create(seed, C.byref(h)), "xpbd_mlp_create")
        self._h = h

    def _chk(self, code, what):
        if code != XPBD_OK:
            msg = lib().xpbd_mlp_last_error(self._h)
            raise XpbdError(code, what, msg.decode() if msg else "")

    def params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        sc = C.c_float()
        self._chk(lib().xpbd_mlp_get_params(self._h, _ptr(p), C.byref(sc)), "xpbd_mlp_get_params")
        return p, sc.value

    def set_params(self, params, out_scale: float):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        self._chk(lib().xpbd_mlp_set_params(self._h, _ptr(p), float(out_scale)), "xpbd_mlp_set_params")

    def forward(self, features: DeviceArray, rows: int) -> np.ndarray:
        out = DeviceArray((rows, 3), np.float32)
        self._chk(lib().xpbd_mlp_forward(self._h, features.ptr, rows, out.ptr), "xpbd_mlp_forward")
        return out.numpy()

    def gradient(self, features: DeviceArray, targets: DeviceArray, rows: int):
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._chk(lib().xpbd_mlp_gradient(self._h, features.ptr, targets.ptr, rows, _ptr(g), C.byref(loss)),
                  "xpbd_mlp_gradient")
        return g, loss.value

    def train(self, features: DeviceArray, targets: DeviceArray, rows: int, steps: int, batch: int,
              lr: float) -> np.ndarray:
        curve = np.zeros(steps, np.float32)
        self._chk(lib().xpbd_mlp_train(self._h, features.ptr, targets.ptr, rows, steps, batch, lr, _ptr(curve)),
                  "xpbd_mlp_train")
        return curve

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            lib().xpbd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest_sqrt(exp_lo: int = -100, exp_hi: int = 127) -> dict:
    """Device check of the fast square root against sqrtf, exhaustive over
    the positive normals with exponents in [exp_lo, exp_hi] and 0."""
    out = np.zeros(2, np.uint32)
    _check(lib().xpbd_selftest_sqrt(exp_lo, exp_hi, _ptr(out)), "xpbd_selftest_sqrt")
    return {"sqrt_mismatches": int(out[0]), "first_bad": int(out[1])}


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
