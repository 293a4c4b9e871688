"""Zero-copy render buffers through xpbd_import_render_buffers_fd.

The renderer's position / normal buffers reach the library as POSIX file
descriptors of exported device memory.  Here HIP's virtual-memory API stands
in for the renderer's Vulkan export: it creates an exportable allocation,
maps it for the test's own fill and readback, and exports the descriptor.
After frames with the import bound, the buffers hold exactly what the
publish writes into plain device buffers (test_publish_layout_and_normals):
positions bit-equal to the state, w = 0, normals bit-exact against the
oracle.  A box whose runtime cannot export or import such memory skips,
with the runtime's error as the reason.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from _helpers import bits_equal

pytestmark = pytest.mark.gpu


class _Loc(C.Structure):
    _fields_ = [("type", C.c_int), ("id", C.c_int)]


class _Flags(C.Structure):
    _fields_ = [("compressionType", C.c_ubyte), ("gpuDirectRDMACapable", C.c_ubyte), ("usage", C.c_ushort)]


class _Prop(C.Structure):   # hipMemAllocationProp
    _fields_ = [("type", C.c_int), ("requestedHandleType", C.c_int), ("location", _Loc),
                ("win32HandleMetaData", C.c_void_p), ("allocFlags", _Flags)]


class _Access(C.Structure):   # hipMemAccessDesc
    _fields_ = [("location", _Loc), ("flags", C.c_int)]


PINNED, POSIX_FD, LOC_DEVICE, PROT_RW, D2H, H2D = 1, 1, 1, 3, 2, 1
VMM_SUPPORTED_ATTR = 'hipDeviceAttributeVirtualMemoryManagementSupported'


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    h = C.CDLL("libamdhip64.so")
    vp, sz, ull = C.c_void_p, C.c_size_t, C.c_ulonglong
    sigs = {
        "hipMemGetAllocationGranularity": [C.POINTER(sz), C.POINTER(_Prop), C.c_int],
        "hipMemCreate": [C.POINTER(vp), sz, C.POINTER(_Prop), ull],
        "hipMemExportToShareableHandle": [vp, vp, C.c_int, ull],
        "hipMemAddressReserve": [C.POINTER(vp), sz, sz, vp, ull],
        "hipMemMap": [vp, sz, sz, vp, ull],
        "hipMemSetAccess": [vp, sz, C.POINTER(_Access), sz],
        "hipMemUnmap": [vp, sz],
        "hipMemRelease": [vp],
        "hipMemAddressFree": [vp, sz],
        "hipMemcpy": [vp, vp, sz, C.c_int],
        "hipDeviceSynchronize": [],
        "hipGetErrorString": [C.c_int],
    }
    for name, args in sigs.items():
        fn = getattr(h, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "hipGetErrorString" else C.c_int
    return h


class Exported:
    """An exportable device allocation of at least `nbytes`, mapped at `va`."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.handle = C.c_void_p()
        self.va = C.c_void_p()
        self.size = 0
        self.mapped = False
        prop = _Prop(type=PINNED, requestedHandleType=POSIX_FD, location=_Loc(LOC_DEVICE, 0))
        gran = C.c_size_t()
        self._ok(hip.hipMemGetAllocationGranularity(C.byref(gran), C.byref(prop), 0), "granularity")
        self.size = -(-nbytes // gran.value) * gran.value
        self._ok(hip.hipMemCreate(C.byref(self.handle), self.size, C.byref(prop), 0), "hipMemCreate")
        self._ok(hip.hipMemAddressReserve(C.byref(self.va), self.size, 0, None, 0), "hipMemAddressReserve")
        self._ok(hip.hipMemMap(self.va, self.size, 0, self.handle, 0), "hipMemMap")
        self.mapped = True
        acc = _Access(location=_Loc(LOC_DEVICE, 0), flags=PROT_RW)
        self._ok(hip.hipMemSetAccess(self.va, self.size, C.byref(acc), 1), "hipMemSetAccess")

    def _ok(self, rc, what):
        if rc != 0:
            reason = self.hip.hipGetErrorString(rc).decode()
            self.close()
            pytest.skip(f"{what}: {reason} (exportable device memory unavailable here)")

    def export_fd(self):
        fd = C.c_int(-1)
        self._ok(self.hip.hipMemExportToShareableHandle(C.byref(fd), self.handle, POSIX_FD, 0),
                 "hipMemExportToShareableHandle")
        return fd.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.va, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, shape):
        out = np.empty(shape, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.va, out.nbytes, D2H) == 0
        return out

    def close(self):
        if self.mapped:
            self.hip.hipMemUnmap(self.va, self.size)
            self.mapped = False
        if self.va.value:
            self.hip.hipMemAddressFree(self.va, self.size)
            self.va = C.c_void_p()
        if self.handle.value:
            self.hip.hipMemRelease(self.handle)
            self.handle = C.c_void_p()


def test_imported_render_buffers_receive_the_publish(O, X):
    note("case render import")
    hip = _hip()
    mesh = X.ClothMeshData.sheet(cols=33, rows=21, spacing=0.03, pins=2, jitter=1e-3, seed=3)
    f = mesh.fields()
    cfg = X.ClothSimConfig.default(solve_substeps=4, solve_inner_iterations=6, use_graph=1)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 4)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    n = sim.n
    bufs = [Exported(hip, 16 * n), Exported(hip, 16 * n)]
    try:
        for b in bufs:
            b.upload(np.full((n, 4), 7.0, np.float32))
        fds = [b.export_fd() for b in bufs]
        try:
            sim.import_render_buffers_fd(fds[0], bufs[0].size, 0, fds[1], bufs[1].size, 0)
        except RuntimeError as exc:
            for fd in fds:
                try:
                    os.close(fd)
                except OSError:
                    pass
            pytest.skip(f"import of an exported allocation refused here: {exc}")
        sim.step_frames(2)
        sim.sync()
        gp, _ = sim.read_state()
        rp = bufs[0].download((n, 4))
        rn = bufs[1].download((n, 4))
        assert (rp[:, 3] == 0).all() and (rn[:, 3] == 0).all()
        assert bits_equal(rp, gp)
        assert bits_equal(rn, O.vertex_normals(gp[:, :3], f["indices"]))
        # rebinding plain pointers releases the imports; the buffers keep the last frame
        sim.bind_render_buffers(None, None)
        sim.step_frames(1)
        sim.sync()
        assert bits_equal(bufs[0].download((n, 4)), gp)
    finally:
        sim.close()
        assert hip.hipDeviceSynchronize() == 0
        for b in bufs:
            b.close()


def test_import_rejects_short_or_misaligned_buffers(X):
    note("case render import bounds")
    mesh = X.ClothMeshData.sheet(cols=8, rows=8, spacing=0.1, pins=1)
    sim = X.ClothComputePlugin(mesh, X.ClothSimConfig.default(), X.ClothSimUniforms.default(X.SOLVER_JACOBI, 8))
    need = 16 * sim.n
    for size, off in ((need - 16, 0), (need, 16), (need + 64, 8)):
        with pytest.raises(RuntimeError):
            sim.import_render_buffers_fd(12345, size, off)   # rejected before the descriptor is touched
    with pytest.raises(RuntimeError):
        sim.import_render_buffers_fd(-1, 0, 0, -1, 0, 0)
    sim.close()
