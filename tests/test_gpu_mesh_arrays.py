"""Meshes that enter through ClothSimConfig's arrays (xpbd_mesh_from_arrays)
run bit-exact against the oracle on the same arrays.  Needs an MI355X.

* an oracle-built sheet with the caller's own inverse masses: extra pins,
  a heavy band, a light corner (ClothMeshData fields a user edits before
  to_sim_config, mesh_prep.rs:103-131, cloth_compute.rs:1599);
* the reference's own asset cloth.obj (651 welded vertices), built by the
  oracle's parse_welded_obj and handed over as arrays;
* the rows without batch offsets (coloured by the library), and a foreign
  adjacency order (slices reversed) that the Jacobi gather must follow.
Each runs 3 substeps with self-collision, Jacobi and Gauss-Seidel, compared
bit for bit after every substep.
"""
import os
import sys

import numpy as np
import pytest

from _helpers import bits_equal, first_diff, max_abs, oracle_params

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sim_config_arrays(f):
    """ClothSimConfig's arrays of an oracle mesh dict (to_sim_config,
    cloth_compute.rs:1599-1700; pack_neighbor_gpu, cloth_jacobi.rs:95-111)."""
    n = f["num_particles"]
    v4 = lambda a: np.concatenate([np.asarray(a, np.float32).reshape(n, 3), np.zeros((n, 1), np.float32)], 1)
    packed = np.stack([f["neighbor_other"].astype(np.float32), f["neighbor_rest_len"], f["neighbor_compliance"],
                       f["neighbor_constraint_id"].astype(np.float32)], axis=1)
    return dict(initial_pos=v4(f["positions"]), rest_pos=v4(f["rest_positions"]), inv_mass=f["inv_mass"],
                constraint_i=f["constraint_i"], constraint_j=f["constraint_j"],
                constraint_rest_len=f["constraint_rest"], constraint_compliance=f["constraint_comp"],
                constraint_batch_offsets=f["batch_offsets"], tri_indices=f["indices"],
                neighbor_offsets=f["neighbor_offsets"], neighbor_packed=packed, ground=f["ground"])


def run_arrays(O, X, f, arrays, solver, S=8, I=6, substeps=3, thickness=None, gs_launch=0):
    mesh = X.ClothMeshData.from_arrays(**arrays)
    cfg = X.ClothSimConfig.default(solve_substeps=S, solve_inner_iterations=I, solver=solver, autotune=0,
                                   gs_launch=gs_launch)
    uk = {} if thickness is None else dict(thickness=thickness)
    u = X.ClothSimUniforms.default(solver, S, **uk)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    pos, vel = f["positions"].copy(), np.zeros_like(f["positions"])
    contacts = 0
    for k in range(substeps):
        sim.step_substeps(1, True)
        gp, gv = sim.read_state()
        pre, _ = s.substep(pos, vel, oracle_params(O, X, u, cfg, collision=False))
        pos, vel = s.substep(pos, vel, p)
        contacts += int(np.any(pre != pos, axis=1).sum())
        assert bits_equal(gp, pos), f"substep {k}: max |d| {max_abs(gp, pos):.3e} at {first_diff(gp, pos)}"
        assert bits_equal(gv, vel), f"substep {k}: velocities differ"
    sim.close()
    return contacts


@pytest.mark.parametrize("solver", [0, 1])
def test_custom_inverse_mass_sheet(O, X, solver):
    note(f"case arrays: custom pins solver={solver}")
    f = O.sheet(cols=28, rows=20, spacing=0.02, layout=1, height=0.5, angle=0.7, pins=0, jitter=2e-3, seed=6)
    w = f["inv_mass"].copy()
    w[[0, 27, 13 * 28 + 5, 19 * 28 + 27]] = 0.0      # own pins
    w[5 * 28:7 * 28] *= np.float32(0.05)             # a heavy band
    w[-40:] *= np.float32(3.0)                       # a light corner
    f["inv_mass"] = w
    # squeeze the sheet so contacts start at once
    c = f["positions"].mean(axis=0)
    f["positions"] = ((f["positions"] - c) * np.float32(0.55) + c).astype(np.float32)
    contacts = run_arrays(O, X, f, sim_config_arrays(f), solver, thickness=0.016)
    assert contacts > 0


@pytest.mark.parametrize("solver,gs_launch", [(0, 0), (1, 0), (1, 3)], ids=["jacobi", "gs", "gs_tiles"])
def test_cloth_obj_through_arrays(O, X, solver, gs_launch):
    """The welded OBJ (its own vertex order: the per-tile-flag GS plans tiles of
    index ranges whose dependency lists follow that order)."""
    note(f"case arrays: cloth.obj solver={solver} gs_launch={gs_launch}")
    with open(os.path.join(GOLDEN, "cloth.obj")) as fh:
        f = O.parse_welded_obj(fh.read())
    assert f["num_particles"] == 651
    c = f["positions"].mean(axis=0)
    f["positions"] = ((f["positions"] - c) * np.float32(0.6) + c).astype(np.float32)
    contacts = run_arrays(O, X, f, sim_config_arrays(f), solver, S=36, I=22, thickness=0.04, gs_launch=gs_launch)
    assert contacts > 0


def test_library_colouring_and_foreign_adjacency_order(O, X):
    """Rows without batch offsets are coloured by the library exactly as the
    oracle's finalize_cloth_mesh restatement colours them; a given adjacency
    with every slice reversed is honoured (the gather sums in slice order)."""
    note("case arrays: library colouring, reversed slices")
    f = O.sheet(cols=20, rows=14, spacing=0.03, pins=2, jitter=1e-3, seed=8)
    a = sim_config_arrays(f)
    a0 = dict(a, constraint_batch_offsets=None, neighbor_offsets=None, neighbor_packed=None)
    for solver in (0, 1):
        run_arrays(O, X, f, a0, solver, substeps=2, thickness=0.024)
    offs = f["neighbor_offsets"].astype(np.int64)
    order = np.concatenate([np.arange(offs[i + 1] - 1, offs[i] - 1, -1) for i in range(len(offs) - 1)])
    assert len(order) == len(f["neighbor_other"])
    g = dict(f)
    for k in ("neighbor_other", "neighbor_rest_len", "neighbor_compliance", "neighbor_constraint_id"):
        g[k] = np.ascontiguousarray(f[k][order])
    run_arrays(O, X, g, sim_config_arrays(g), 0, substeps=2, thickness=0.024)
    # non-vacuous: the two summation orders give other f32 results
    p = O.params(substeps=8, inner_iterations=6, thickness=0.024, collision=True, coll_mode=O.COLL_GRID)
    a, _ = O.Solver(f).frames(f["positions"], np.zeros_like(f["positions"]), p, 2, 1, 1)
    b, _ = O.Solver(g).frames(f["positions"], np.zeros_like(f["positions"]), p, 2, 1, 1)
    assert not bits_equal(a, b)
