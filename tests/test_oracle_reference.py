"""The reference crate's own CPU tests, restated against the oracle.

Each test cites the reference test it restates (src/xpbd_cpu.rs,
src/cloth_compute.rs).  The reference runs one frame as SUBSTEPS = 36
substeps of INNER_ITERS = 22 iterations (cloth_compute.rs:59-61); both solver
features are covered (xpbd_cpu.rs is compiled once per feature).
"""
import numpy as np
import pytest

SOLVERS = [0, 1]   # oracle.SOLVER_JACOBI, oracle.SOLVER_GS


def _params(O, solver, **kw):
    base = dict(substeps=36, inner_iterations=22, solver=solver,
                omega=0.32 if solver == O.SOLVER_JACOBI else 1.0,
                grab_stiffness=0.28 if solver == O.SOLVER_JACOBI else 0.45)
    base.update(kw)
    return O.params(**base)


def _frame(O, mesh, p, frames=1):
    s = O.Solver(mesh)
    pos = mesh["positions"].copy()
    vel = np.zeros_like(pos)
    return s.frames(pos, vel, p, 36, 1, frames)


def test_cloth_mesh_inv_mass_and_neighbors_sane(O):
    # xpbd_cpu.rs:751-757
    m = O.grid_cloth_hanging(16, 16, 0.045)
    assert np.isfinite(m["inv_mass"]).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_one_frame_integrator_only_stays_finite(O, solver):
    # xpbd_cpu.rs:759-807 (jacobi_omega = 0)
    m = O.grid_cloth_hanging(16, 16, 0.045)
    pos, _ = _frame(O, m, _params(O, solver, omega=0.0))
    assert np.isfinite(pos).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_one_frame_full_solver_stays_finite(O, solver):
    # xpbd_cpu.rs:809-855
    m = O.grid_cloth_hanging(16, 16, 0.045)
    pos, _ = _frame(O, m, _params(O, solver))
    assert np.isfinite(pos).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_pins_hold_rest_positions(O, solver):
    # xpbd_cpu.rs:857-913
    m = O.grid_cloth_hanging(16, 16, 0.045)
    p = _params(O, solver, grab_idx=-1, grab_active=False, grab_stiffness=0.0)
    pos, _ = _frame(O, m, p)
    pinned = m["inv_mass"] <= 0
    assert pinned.sum() == 17
    assert (np.linalg.norm(pos[pinned] - m["positions"][pinned], axis=1) < 1e-4).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_single_triangle_one_frame_finite(O, solver):
    # xpbd_cpu.rs:915-974
    obj = "\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n"
    m = O.parse_welded_obj(obj)
    assert m["num_particles"] == 3
    pos, _ = _frame(O, m, _params(O, solver, gravity=(-1.0, -2.0, 0.5)))
    assert np.isfinite(pos).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_cloth_15_frames_extent_stays_reasonable(O, solver):
    # xpbd_cpu.rs:976-1034
    m = O.grid_cloth_hanging(16, 16, 0.045)
    pos, _ = _frame(O, m, _params(O, solver), frames=15)
    free = m["inv_mass"] > 0
    ext = pos[free].max(0) - pos[free].min(0)
    assert ext[0] > 0.18 and ext[1] > 0.04, ext


def _pair(O, inv_mass, sim):
    mesh = {"num_particles": 2, "inv_mass": np.array(inv_mass, np.float32),
            "rest_positions": np.array([[0, 0, 0], [0.1, 0, 0]], np.float32),
            "ground": np.zeros(2, np.uint32), "constraint_i": np.zeros(0, np.uint32),
            "constraint_j": np.zeros(0, np.uint32), "constraint_rest": np.zeros(0, np.float32),
            "constraint_comp": np.zeros(0, np.float32), "neighbor_offsets": np.zeros(3, np.uint32),
            "neighbor_other": np.zeros(0, np.uint32), "neighbor_constraint_id": np.zeros(0, np.uint32)}
    s = O.Solver(mesh)
    return s.self_collision(np.array(sim, np.float32), 0.04, 0.38, O.COLL_BRUTE)


def test_self_collision_pair_separates_when_penetrating(O):
    # xpbd_cpu.rs:1036-1047
    sim = [[0, 0, 0], [0.015, 0, 0]]
    out = _pair(O, [1.0, 1.0], sim)
    assert np.linalg.norm(out[1] - out[0]) > 0.015 + 1e-4


def test_self_collision_skips_when_world_separation_exceeds_rest(O):
    # xpbd_cpu.rs:1049-1060 (bit-equal)
    sim = np.array([[0, 0, 0], [0.15, 0, 0]], np.float32)
    out = _pair(O, [1.0, 1.0], sim)
    assert out.tobytes() == sim.tobytes()


def test_self_collision_pinned_neighbor_takes_no_delta(O):
    # xpbd_cpu.rs:1062-1073
    sim = np.array([[0, 0, 0], [0.015, 0, 0]], np.float32)
    out = _pair(O, [0.0, 1.0], sim)
    assert out[0].tobytes() == sim[0].tobytes()
    assert out[1, 0] > 0.015


@pytest.mark.parametrize("solver", SOLVERS)
def test_one_frame_with_self_collision_stays_finite(O, solver):
    # xpbd_cpu.rs:1075-1128 (thickness 0.04, coll_scale 0.38)
    m = O.grid_cloth_hanging(16, 16, 0.045)
    pos, _ = _frame(O, m, _params(O, solver, collision=True, coll_mode=O.COLL_BRUTE))
    assert np.isfinite(pos).all()


def test_cloth_neighbor_and_tri_indices_in_range(O):
    # cloth_compute.rs:1528-1548
    m = O.grid_cloth_hanging(12, 10, 0.05)
    n = m["num_particles"]
    assert (m["neighbor_other"] < n).all()
    assert (m["indices"] < n).all()


def test_cloth_neighbor_slices_match_constraints(O):
    # cloth_compute.rs:1551-1577
    m = O.grid_cloth_hanging(10, 8, 0.05)
    edges = {(min(a, b), max(a, b)) for a, b in zip(m["constraint_i"], m["constraint_j"])}
    offs = m["neighbor_offsets"]
    for i in range(m["num_particles"]):
        for k in range(offs[i], offs[i + 1]):
            j = int(m["neighbor_other"][k])
            assert (min(i, j), max(i, j)) in edges
