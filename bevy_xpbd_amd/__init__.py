"""bevy_xpbd_amd — MI355X-native XPBD cloth compute path.

Python mirror of the reference crate's plugin API for the cloth compute step
(mate-h/bevy_xpbd, crate bevy_softbody): ClothMeshData and its builders
(src/mesh_prep.rs), ClothSimConfig / ClothSimUniforms / ClothSimControl and
ClothComputePlugin (src/cloth_compute.rs), ClothMaterialPlugin's buffer layout
(src/cloth_material.rs).  Every call goes through the C ABI of
libxpbd_hip.so (include/xpbd.h); there is no CPU fallback: without the built
library, or without a GPU for the compute entry points, calls raise.
"""
from __future__ import annotations

from .api import (  # noqa: F401
    LIB_PATH,
    ClothComputePlugin,
    ClothMaterialPlugin,
    ClothMeshData,
    ClothSimConfig,
    ClothSimControl,
    ClothSimUniforms,
    DeviceArray,
    MeshDesc,
    SheetDesc,
    WarmStartMLP,
    MLP_PARAMS,
    XpbdError,
    derive_collision_grid,
    lib,
    selftest_division,
    selftest_sqrt,
    KERNEL_JACOBI_ITER,
    KERNEL_COLLISION_SWEEP,
    KERNEL_RADIX_PASS,
    KERNEL_GS_SUBSTEP,
    KERNEL_COPY,
    KERNEL_JACOBI_PERSISTENT,
    SOLVER_JACOBI,
    SOLVER_GAUSS_SEIDEL,
    LAYOUT_HANGING_XY,
    LAYOUT_HORIZONTAL_XZ,
    PIN_NONE,
    PIN_TWO_CORNERS,
    PIN_TOP_ROW,
    PIN_CENTER,
    XPBD_OK,
    XPBD_E_INVALID,
    XPBD_E_HIP,
    XPBD_E_NOMEM,
    XPBD_E_UNSUPPORTED,
    XPBD_E_STATE,
    XPBD_E_NO_DEVICE,
)
