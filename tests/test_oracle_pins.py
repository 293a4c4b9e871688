"""Pins for the oracle functions that the solver tests lean on but that no
closed form covered before: the relative residual, the ground plane, the
exact greedy colouring order, and OBJ welding on the reference's own asset.

Each expectation below is derived by hand (or is a closed form), not by
running either implementation:

* residual: sqrt(sum C^2) / sqrt(sum rest^2) on meshes whose constraint
  lengths are exact binary fractions, so the expected value is exact;
* ground plane: a flagged free particle predicted below y = 0 ends at y = 0
  and leaves with v_y = (0 - y_prev) / dt (drag 0); pinned, unflagged and
  above-ground particles are untouched;
* colouring: grid_cloth_hanging(2, 1, 0.5) traced through
  partition_constraints_for_gs_batches (mesh_prep.rs:153-221) by hand;
* cloth.obj (the reference's asset, a 21 x 31 vertex grid of 0.05 m quads):
  counts that follow from the grid shape (mesh_prep.rs:223-357, 444-501,
  532-568).
"""
import os

import numpy as np
import pytest

from _helpers import tiny_mesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STRETCH, BEND, SKIP2, CROSS = np.float32(2e-8), np.float32(3.8), np.float32(2.5e-8), np.float32(6e-9)


# ----------------------------------------------------------------- residual
def _chain(n, spacing):
    """Particles on the x axis, axial constraints of rest `spacing`."""
    pos = [[spacing * k, 0.0, 0.0] for k in range(n)]
    cons = [(k, k + 1, spacing, 0.0) for k in range(n - 1)]
    return tiny_mesh(pos, [1.0] * n, cons)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_residual_of_the_rest_pose_is_zero(O, dtype):
    m = _chain(9, 0.25)
    assert O.Solver(m).residual(m["positions"].astype(dtype)) == 0.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_residual_of_a_uniform_stretch_is_the_stretch(O, dtype):
    """Every length scaled by 1 + s: C_e = s rest_e, so ||C|| / ||rest|| = s.
    With s = 1/16 and rest = 1/4 every quantity is a binary fraction and
    the sums of squares differ by the factor 2^-8, so the ratio is exact."""
    m = _chain(9, 0.25)
    s = 1.0 / 16.0
    pos = (m["positions"].astype(np.float64) * (1.0 + s)).astype(dtype)
    assert O.Solver(m).residual(pos) == s


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_residual_two_constraint_closed_form(O, dtype):
    """C = (0.5, -0.75) against rest (1, 2): sqrt(0.8125) / sqrt(5)."""
    pos = [[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [1.5, 1.25, 0.0]]
    m = tiny_mesh(pos, [1.0, 1.0, 1.0], [(0, 1, 1.0, 0.0), (1, 2, 2.0, 0.0)])
    got = O.Solver(m).residual(np.asarray(pos, dtype))
    assert got == pytest.approx(np.sqrt(0.8125) / np.sqrt(5.0), rel=1e-15)
    # a dropped square root, an L1 norm or a missing normalisation each miss it
    assert abs(got - 0.8125 / 5.0) > 0.1 and abs(got - 1.25 / 3.0) > 0.01 and abs(got - np.sqrt(0.8125)) > 0.1


# ------------------------------------------------------------- ground plane
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ground_plane_closed_form(O, dtype):
    """One substep, no constraints, drag 0.  Particle 0 (free, flagged)
    falls through y = 0: it ends at y = 0 exactly with v_y = -y_prev / dt,
    and x, z move as in free flight.  1 is pinned and flagged below ground:
    untouched.  2 is free and unflagged: it ends below ground in free
    flight.  3 is flagged but stays above ground: identical to a run
    without the flag."""
    y0 = 1e-3
    pos = [[0.25, y0, -0.5], [0.0, -0.5, 0.0], [1.0, y0, 0.0], [2.0, 1.0, 0.0]]
    vel = np.array([[0.5, -5.0, 0.25], [0.0, 0.0, 0.0], [0.0, -5.0, 0.0], [0.0, -1.0, 0.0]], dtype)
    w = [1.0, 0.0, 1.0, 1.0]
    p = O.params(substeps=8, drag=0.0)
    flagged = O.Solver(tiny_mesh(pos, w, ground=[1, 1, 0, 1]))
    free = O.Solver(tiny_mesh(pos, w))
    x_g, v_g = flagged.substep(np.asarray(pos, dtype), vel, p, dtype=dtype)
    x_f, v_f = free.substep(np.asarray(pos, dtype), vel, p, dtype=dtype)
    dt = dtype(p.dt)
    inv_dt = dtype(p.inv_dt)
    # free flight below ground without the flag (sanity of the set-up)
    assert x_f[0, 1] < 0 and x_f[2, 1] < 0
    # particle 0: lifted to exactly 0, velocity from the lifted position
    assert x_g[0, 1] == 0.0
    assert v_g[0, 1] == dtype(dtype(0.0) - dtype(y0)) * inv_dt
    assert x_g[0, 0] == x_f[0, 0] and x_g[0, 2] == x_f[0, 2]
    assert v_g[0, 0] == v_f[0, 0] and v_g[0, 2] == v_f[0, 2]
    assert v_g[0, 1] != v_f[0, 1]
    # particle 1: pinned, untouched below ground
    assert x_g[1].tolist() == [0.0, -0.5, 0.0] and v_g[1].tolist() == [0.0, 0.0, 0.0]
    # particles 2 and 3: bit-identical to the run without ground flags
    assert x_g[2:].tobytes() == x_f[2:].tobytes() and v_g[2:].tobytes() == v_f[2:].tobytes()
    # free flight of 3: v = v0 + g dt, x = x0 + v dt; the velocity comes back
    # as (x - x0) / dt, which in f32 carries the rounding of x (|x| = 1 against
    # |x - x0| = 2e-3: about 3e-5 relative; in f64 the f32 parameters dt and
    # 1/dt differ from exact reciprocals by about 1e-7)
    tol = 1e-4 if dtype == np.float32 else 1e-7
    v3 = -1.0 + float(p.gravity[1]) * float(dt)
    assert v_g[3, 1] == pytest.approx(v3, rel=tol)
    assert x_g[3, 1] == pytest.approx(1.0 + v3 * float(dt), rel=1e-7)


# ---------------------------------------------------------------- colouring
# grid_cloth_hanging(2, 1, 0.5): vertices 0 1 2 (y = 0) and 3 4 5 (y = 0.5,
# pinned).  Constraints before colouring (finalize_cloth_mesh order: triangle
# edges, hinge pairs, skip-2 braces, cross diagonals):
HAND_RAW = [
    (0, 1, 0.5, STRETCH), (1, 3, 0.70710677, STRETCH), (0, 3, 0.5, STRETCH), (1, 4, 0.5, STRETCH),
    (3, 4, 0.5, STRETCH), (1, 2, 0.5, STRETCH), (2, 4, 0.70710677, STRETCH), (2, 5, 0.5, STRETCH),
    (4, 5, 0.5, STRETCH), (0, 4, 0.70710677, BEND), (2, 3, 1.118034, BEND), (1, 5, 0.70710677, BEND),
    (0, 2, 1.0, SKIP2), (3, 5, 1.0, SKIP2), (0, 4, 0.70710677, CROSS), (1, 5, 0.70710677, CROSS),
]
# Degrees: 1 and 4 have 6, the others 5.  Processing order by (max endpoint
# degree descending, min, max, input index):
#   e0 e9 e14 e5 e1 e3 e11 e15 e6 e4 e8 | e12 e2 e10 e7 e13
# and the smallest colour free at both ends gives
#   e0:0 e9:1 e14:2 e5:1 e1:2 e3:3 e11:4 e15:5 e6:0 e4:4 e8:6 e12:3 e2:5 e10:6 e7:2 e13:0.
# Sorted by (colour, input index):
HAND_ORDER = [0, 6, 13, 5, 9, 1, 7, 14, 3, 12, 4, 11, 2, 15, 8, 10]
HAND_OFFSETS = [0, 3, 5, 8, 10, 12, 14, 16]


def _check_hand_colouring(m):
    want = [HAND_RAW[k] for k in HAND_ORDER]
    assert m["constraint_i"].tolist() == [c[0] for c in want]
    assert m["constraint_j"].tolist() == [c[1] for c in want]
    assert m["constraint_comp"].tolist() == [float(c[3]) for c in want]
    assert np.allclose(m["constraint_rest"], [c[2] for c in want], rtol=1e-7)
    assert m["batch_offsets"].tolist() == HAND_OFFSETS
    # CSR by cursor scatter in that order: vertex 0's slice lists its
    # constraints by sorted row, 1 before 4 (bend) before 4 (cross) ...
    offs, other, eid = m["neighbor_offsets"], m["neighbor_other"], m["neighbor_constraint_id"]
    assert other[offs[0]:offs[1]].tolist() == [1, 4, 4, 2, 3]
    assert eid[offs[0]:offs[1]].tolist() == [0, 4, 7, 9, 12]


def test_colouring_hand_traced_oracle_grid(O):
    _check_hand_colouring(O.grid_cloth_hanging(2, 1, 0.5))


def test_colouring_hand_traced_oracle_assemble(O):
    pos = [[-0.5, 0, 0], [0, 0, 0], [0.5, 0, 0], [-0.5, 0.5, 0], [0, 0.5, 0], [0.5, 0.5, 0]]
    _check_hand_colouring(O.assemble(pos, [1.0] * 6, HAND_RAW))


def test_colouring_hand_traced_product(X):
    _check_hand_colouring(X.ClothMeshData.grid_cloth_hanging(2, 1, 0.5).fields())


def test_inverse_mass_hand_traced(O, X):
    """Triangle areas 1/8: each incident triangle adds f32(1 / (1/8) / 3);
    vertex 0 has 1 triangle, 1 has 3, 2 has 2; the top row is pinned."""
    third = np.float32(np.float32(1.0) / np.float32(0.125)) / np.float32(3.0)
    want = [third, np.float32(np.float32(third + third) + third), np.float32(third + third), 0.0, 0.0, 0.0]
    for m in (O.grid_cloth_hanging(2, 1, 0.5), X.ClothMeshData.grid_cloth_hanging(2, 1, 0.5).fields()):
        assert m["inv_mass"].tolist() == [float(v) for v in np.asarray(want, np.float32)]


def test_colouring_small_closed_forms(O):
    """A path 0-1-2-3 takes colours 0 1 0 (greedy in (min, max) order); a
    star of k edges takes colours 0..k-1 and k batches."""
    z = [[0, 0, 0]] * 4
    m = O.assemble(z, [1.0] * 4, [(2, 3, 1.0, 0.0), (0, 1, 1.0, 0.0), (1, 2, 1.0, 0.0)])
    assert list(zip(m["constraint_i"].tolist(), m["constraint_j"].tolist())) == [(2, 3), (0, 1), (1, 2)]
    assert m["batch_offsets"].tolist() == [0, 2, 3]
    star = O.assemble([[0, 0, 0]] * 6, [1.0] * 6, [(0, k, 1.0, 0.0) for k in (5, 3, 1, 4, 2)])
    assert star["constraint_j"].tolist() == [1, 2, 3, 4, 5]
    assert star["batch_offsets"].tolist() == [0, 1, 2, 3, 4, 5]


# ---------------------------------------------------------------- cloth.obj
def _cloth_obj():
    with open(os.path.join(GOLDEN, "cloth.obj")) as f:
        return f.read()


def _check_cloth_obj(m):
    cols, rows = 21, 31                        # x: -0.5..0.5, y: -0.75..0.75, step 0.05
    quads = (cols - 1) * (rows - 1)             # 600
    tri_edges = (cols - 1) * rows + (rows - 1) * cols + quads    # 620 + 630 + 600 = 1850
    boundary = 2 * (cols - 1) + 2 * (rows - 1)  # 100
    skip2 = rows * (cols - 2) + (rows - 2) * cols               # 589 + 609 = 1198
    assert m["num_particles"] == cols * rows == 651
    assert len(m["indices"]) == 3 * 2 * quads == 3600
    comp = m["constraint_comp"]
    assert int((comp == STRETCH).sum()) == tri_edges == 1850
    assert int((comp == BEND).sum()) == tri_edges - boundary == 1750
    assert int((comp == SKIP2).sum()) == skip2 == 1198
    assert int((comp == CROSS).sum()) == quads == 600
    assert m["num_constraints"] == 1850 + 1750 + 1198 + 600
    pos = m["positions"]
    pinned = m["inv_mass"] == 0
    assert int(pinned.sum()) == cols and np.all(pos[pinned, 1] == np.float32(0.75))
    # rest lengths by class: skip-2 braces span two cells, cross diagonals one
    rest = m["constraint_rest"]
    assert np.allclose(rest[comp == SKIP2], 0.1, atol=1e-6)
    assert np.allclose(rest[comp == CROSS], 0.05 * np.sqrt(2.0), atol=1e-6)
    assert set(np.round(rest[comp == STRETCH].astype(np.float64), 5).tolist()) == {0.05, 0.07071}
    # every batch is vertex-disjoint
    ci, cj, bo = m["constraint_i"], m["constraint_j"], m["batch_offsets"]
    for b in range(len(bo) - 1):
        v = np.concatenate([ci[bo[b]:bo[b + 1]], cj[bo[b]:bo[b + 1]]])
        assert len(np.unique(v)) == len(v)


def test_cloth_obj_counts_oracle(O):
    _check_cloth_obj(O.parse_welded_obj(_cloth_obj()))


def test_cloth_obj_counts_product(X):
    _check_cloth_obj(X.ClothMeshData.parse_welded_obj(_cloth_obj()).fields())
