"""Batched cloths on the GPU vs the oracle, through the C ABI (tests marked gpu
need an MI355X).

A batch is many cloths in one set of buffers and launches; each keeps its
own pins, ground flag and self-collision grid (keys carry a cloth-id prefix),
and no contact forms between two cloths.  Bar: bit-exact against the f32
oracle stepping the same batch.

Cross-cloth contacts need pairs closer now than in the rest pose
(xpbd_cpu.rs:95-101); cloths built in one place have coincident rest poses
and would hide a missing cloth boundary.  Cloth 1 is therefore built 1 m
away (horizontal, at height 1) and its current positions are then written
onto cloth 0's, so that a cross-cloth contact would fire; the oracle shows
that a merged single cloth does form such contacts.
"""
import sys

import numpy as np
import pytest

from _helpers import bits_equal, first_diff, max_abs, oracle_params


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def descs():
    return [
        dict(cols=18, rows=14, spacing=0.02, pins=1, jitter=1e-3, seed=5),
        dict(cols=18, rows=14, spacing=0.02, layout=1, height=1.0, angle=0.2, pins=0, jitter=1e-3, seed=6),
        dict(cols=11, rows=23, spacing=0.015, layout=1, height=0.05, angle=0.4, pins=0, ground=1, jitter=2e-3,
             seed=7, stretch_compliance=1e-7),
        dict(cols=9, rows=9, spacing=0.03, pins=3, jitter=0.0, seed=8, bend_compliance=1e-2),
        dict(cols=40, rows=7, spacing=0.01, pins=2, jitter=5e-4, seed=9, shear_compliance=1e-4),
    ]


def overlapped_state(f):
    """Current positions: cloth 1 laid over cloth 0 (same shape), others as built."""
    offs = f["cloth_offsets"]
    pos = np.concatenate([f["positions"], f["inv_mass"][:, None]], axis=1).astype(np.float32)
    rng = np.random.default_rng(3)
    a0, a1, b0, b1 = offs[0], offs[1], offs[1], offs[2]
    assert a1 - a0 == b1 - b0
    pos[b0:b1, :3] = pos[a0:a1, :3] + rng.normal(0, 2e-3, (a1 - a0, 3)).astype(np.float32)
    return pos


def make(X, mesh, solver=0, S=8, I=6, persistent=0, gs_launch=0, **uk):
    cfg = X.ClothSimConfig.default(solve_substeps=S, solve_inner_iterations=I, solver=solver, use_graph=0,
                                   persistent_jacobi=persistent, gs_launch=gs_launch)
    u = X.ClothSimUniforms.default(solver, S, **uk)
    return X.ClothComputePlugin(mesh, cfg, u), cfg, u


@pytest.mark.gpu
def test_batch_collision_keys_and_sort_bit_exact(O, X):
    note("case batch keys")
    mesh = X.ClothMeshData.batch([X.SheetDesc.make(**d) for d in descs()])
    f = mesh.fields()
    offs = f["cloth_offsets"]
    nc = len(offs) - 1
    sim, _, _ = make(X, mesh)
    pos = overlapped_state(f)
    pos[:, :3] += np.random.default_rng(1).normal(0, 0.01, (len(pos), 3)).astype(np.float32)
    sim.write_state(pos)
    pos, _ = sim.read_state()
    for cell in (0.006, 0.012, 0.05):
        grids, keys, ids = sim.debug_collision_keys(cell, num_cloths=nc)
        ogrids, okeys = O.cell_keys_batch(pos, offs, cell)
        cb = int(np.ceil(np.log2(nc)))
        for k in range(nc):
            assert np.array(ogrids[k].lo[:], np.float32).tobytes() == grids[k]["lo"].tobytes(), (cell, k)
            assert list(ogrids[k].dims) == list(grids[k]["dims"]), (cell, k)
            assert np.float32(ogrids[k].inv_cell) == grids[k]["inv_cell"], (cell, k)
            assert grids[k]["key_base"] == k << (32 - cb)
        sk, si = O.sort_pairs(okeys)
        assert (keys == sk).all() and (ids == si).all(), cell
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["jacobi", "jacobi_persistent", "gs", "gs_colours", "gs_tiles"])
def test_batch_substeps_with_collision_bit_exact(O, X, mode):
    note(f"case batch substeps mode={mode}")
    solver = 1 if mode.startswith("gs") else 0
    mesh = X.ClothMeshData.batch([X.SheetDesc.make(**d) for d in descs()])
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, solver=solver, persistent=1 if mode == "jacobi_persistent" else 0,
                       gs_launch={"gs": 1, "gs_colours": 2, "gs_tiles": 3}.get(mode, 0),
                       thickness=0.012, coll_scale=1.0, jacobi_omega=0.32 if solver == 0 else 1.0)
    pos = overlapped_state(f)
    sim.write_state(pos)
    pos, _ = sim.read_state()
    vel = np.zeros_like(pos)
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    for k in range(6):
        note(f"  substep {k}")
        sim.step_substeps(1, True)
        gp, gv = sim.read_state()
        pos, vel = s.substep(pos, vel, p)
        assert bits_equal(gp, pos), f"substep {k}: max |d| {max_abs(gp, pos):.3e} at {first_diff(gp, pos)}"
        assert bits_equal(gv, vel), f"substep {k}: velocities differ"
    sim.close()


def test_batch_boundary_is_not_vacuous(O, X):
    """(CPU, oracle only.)  With the cloth boundary removed (one cloth of all particles) the same
    start does form contacts between cloths 0 and 1 and ends elsewhere."""
    mesh = X.ClothMeshData.batch([X.SheetDesc.make(**d) for d in descs()])
    f = mesh.fields()
    pos = overlapped_state(f)
    p = O.params(substeps=8, inner_iterations=6, thickness=0.012, coll_scale=1.0, collision=True,
                 coll_mode=O.COLL_GRID)
    merged = dict(f)
    merged["cloth_offsets"] = np.array([0, f["num_particles"]], np.uint32)
    a, _ = O.Solver(f).substep(pos[:, :3].copy(), np.zeros((len(pos), 3), np.float32), p)
    b, _ = O.Solver(merged).substep(pos[:, :3].copy(), np.zeros((len(pos), 3), np.float32), p)
    assert a.tobytes() != b.tobytes()
