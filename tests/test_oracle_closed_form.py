"""Pin the oracle against what the mathematics fixes.

Closed forms (free fall with drag as a geometric series, the speed cap, a
single distance constraint projected by Jacobi and by Gauss-Seidel, the
correction clamp), invariants (pinned particles are bit-invariant) and brute
force (hash-grid collision == the reference's O(N^2) loop, bit for bit).  A
dropped term, a wrong sign or a swapped operand in the oracle fails one of
these.
"""
import numpy as np
import pytest

from _helpers import tiny_mesh as _mesh


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_free_fall_with_drag_is_a_geometric_series(O, dtype, tol):
    S, K = 8, 40
    p = O.params(substeps=S, inner_iterations=3, drag=1.25)
    m = _mesh([[0.0, 1.0, 0.0]], [1.0])
    s = O.Solver(m)
    pos = m["positions"].astype(dtype)
    vel = np.array([[0.3, 0.0, -0.2]], dtype)
    v0 = vel[0].astype(np.float64).copy()
    x0 = pos[0].astype(np.float64).copy()
    for _ in range(K):
        pos, vel = s.substep(pos, vel, p, dtype=dtype)
    dt, inv_dt = float(p.dt), float(p.inv_dt)
    d = np.exp(-1.25 * dt) if dtype == np.float64 else float(np.exp(np.float32(-1.25) * np.float32(dt)))
    g = np.array(p.gravity[:], np.float64)
    a = dt * inv_dt * d            # v_k = a v_{k-1} + a g dt
    v = v0.copy()
    xs = x0.copy()
    vs = []
    for _ in range(K):
        vs.append(v.copy())
        v = a * v + a * g * dt
    # closed form of the recurrence
    ak = a ** K
    v_closed = ak * v0 + a * g * dt * (1 - ak) / (1 - a)
    x_closed = x0 + dt * (np.sum(vs, axis=0) + K * g * dt)
    assert np.allclose(vel[0], v_closed, rtol=tol, atol=tol)
    assert np.allclose(pos[0], x_closed, rtol=tol, atol=tol)
    assert np.allclose(v, v_closed, rtol=1e-12, atol=1e-12)


def test_speed_cap_clamps_to_12(O):
    p = O.params(substeps=8, gravity=(0.0, 0.0, 0.0), drag=0.0)
    m = _mesh([[0.0, 0.0, 0.0]], [1.0])
    pos, vel = O.Solver(m).substep(m["positions"], np.array([[0.0, -20.0, 0.0]], np.float32), p,
                                   dtype=np.float64)
    dt, inv_dt = float(p.dt), float(p.inv_dt)
    assert np.isclose(pos[0, 1], -12.0 * dt, rtol=1e-12)
    assert np.isclose(vel[0, 1], -12.0 * dt * inv_dt, rtol=1e-12)


def test_pinned_particle_is_bit_invariant(O):
    p = O.params(substeps=8)
    m = _mesh([[0.1, 0.2, 0.3], [0.1, 0.0, 0.3]], [0.0, 1.0], [(0, 1, 0.2, 0.0)])
    s = O.Solver(m)
    pos = m["positions"].copy()
    vel = np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0]], np.float32)
    for _ in range(20):
        pos, vel = s.substep(pos, vel, p)
    assert pos[0].tobytes() == m["positions"][0].tobytes()
    assert vel[0].tobytes() == np.array([1.0, 2.0, 3.0], np.float32).tobytes()


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("comp,omega", [(0.0, 1.0), (1e-4, 1.0), (0.0, 0.32), (1e-6, 0.5)])
def test_single_constraint_projection(O, solver, comp, omega):
    """One constraint, equal masses, one iteration: dlambda = -C/(2w + alpha/dt^2)
    and each endpoint moves omega*w*dlambda along the constraint direction."""
    L, rest, w = 1.1, 1.0, 1.0
    p = O.params(substeps=8, inner_iterations=1, omega=omega, gravity=(0, 0, 0), drag=0.0, solver=solver)
    m = _mesh([[0.0, 0.0, 0.0], [L, 0.0, 0.0]], [w, w], [(0, 1, rest, comp)])
    pos, _ = O.Solver(m).substep(m["positions"], np.zeros((2, 3), np.float32), p, dtype=np.float64)
    alpha_t = comp * float(p.inv_dt) ** 2
    C = L - rest
    dlam = -C / (2 * w + alpha_t)
    move = omega * w * abs(dlam)
    assert np.isclose(pos[0, 0], move, rtol=1e-6)
    if solver == O.SOLVER_JACOBI:
        assert np.isclose(pos[1, 0], L - move, rtol=1e-6)
    else:
        # Gauss-Seidel updates both endpoints from the same dlambda too
        assert np.isclose(pos[1, 0], L - move, rtol=1e-6)
    if comp == 0.0 and omega == 1.0:
        assert np.isclose(pos[1, 0] - pos[0, 0], rest, rtol=1e-12)


def test_gauss_seidel_chain_sees_updated_positions(O):
    """Chain 0-1-2 at spacing 1.2, rest 1: GS projects (0,1) first (colour 0),
    then (1,2) with the moved particle 1 — Jacobi would use the old one."""
    p = O.params(substeps=8, inner_iterations=1, omega=1.0, gravity=(0, 0, 0), drag=0.0, solver=O.SOLVER_GS)
    m = _mesh([[0, 0, 0], [1.2, 0, 0], [2.4, 0, 0]], [1.0, 1.0, 1.0], [(0, 1, 1.0, 0.0), (1, 2, 1.0, 0.0)])
    s = O.Solver(m)
    pos, _ = s.substep(m["positions"], np.zeros((3, 3), np.float32), p, dtype=np.float64)
    x1 = 1.2 - 0.1            # after (0,1): 0 -> 0.1, 1 -> 1.1
    c2 = (2.4 - x1) - 1.0     # then (1,2) from 1.1
    assert np.allclose(pos[:, 0], [0.1, x1 + c2 / 2, 2.4 - c2 / 2], rtol=1e-6)   # inputs are f32
    pj, _ = s.substep(m["positions"], np.zeros((3, 3), np.float32),
                      O.params(substeps=8, inner_iterations=1, omega=1.0, gravity=(0, 0, 0), drag=0.0),
                      dtype=np.float64)
    assert np.allclose(pj[:, 0], [0.1, 1.2, 2.3], rtol=1e-6)   # Jacobi: both constraints see old x1


def test_correction_clamp_is_0_28(O):
    p = O.params(substeps=8, inner_iterations=1, omega=1.0, gravity=(0, 0, 0), drag=0.0)
    m = _mesh([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]], [1.0, 1.0], [(0, 1, 1.0, 0.0)])
    pos, _ = O.Solver(m).substep(m["positions"], np.zeros((2, 3), np.float32), p, dtype=np.float64)
    assert np.isclose(pos[0, 0], 0.28, rtol=1e-7)
    assert np.isclose(pos[1, 0], 2.0 - 0.28, rtol=1e-7)


def test_collision_pair_closed_form(O):
    """Two free particles 0.015 apart, thickness 0.04, rest 0.1: each moves
    (0.04 - 0.015) * 0.5 * coll_scale * w/(w+w) apart."""
    m = _mesh([[0, 0, 0], [0.015, 0, 0]], [1.0, 1.0], rest_pos=[[0, 0, 0], [0.1, 0, 0]])
    out = O.Solver(m).self_collision(m["positions"], 0.04, 0.38, O.COLL_BRUTE, dtype=np.float64)
    corr = (0.04 - 0.015) * 0.5 * 0.38
    assert np.allclose(out[:, 0], [-corr * 0.5, 0.015 + corr * 0.5], rtol=1e-6)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_collision_grid_equals_brute_force_bitwise(O, seed):
    rng = np.random.default_rng(seed)
    n = 1500
    rest = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    pos = (rng.uniform(0, 0.25, (n, 3))).astype(np.float32)     # dense: many contacts
    w = rng.choice([0.0, 1.0, 2.0], n, p=[0.1, 0.6, 0.3]).astype(np.float32)
    m = _mesh(pos, w, rest_pos=rest)
    s = O.Solver(m)
    for dtype in (np.float32, np.float64):
        a = s.self_collision(pos, 0.02, 0.38, O.COLL_BRUTE, dtype=dtype)
        b = s.self_collision(pos, 0.02, 0.38, O.COLL_GRID, dtype=dtype)
        assert a.tobytes() == b.tobytes()
        assert not np.array_equal(a, pos.astype(dtype))   # contacts happened


def test_sort_pairs_is_stable_by_key(O):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 50, 10000).astype(np.uint32)
    sk, si = O.sort_pairs(keys)
    order = np.argsort(keys, kind="stable")
    assert (si == order).all() and (sk == keys[order]).all()


def test_cell_keys_are_dense_and_neighbourhood_complete(O):
    """Every pair within cell_size lies in adjacent cells (|d cell| <= 1 per axis)."""
    rng = np.random.default_rng(9)
    pos = np.zeros((4000, 4), np.float32)
    pos[:, :3] = rng.uniform(-3, 5, (4000, 3))
    cell = 0.37
    g = O.cell_grid(pos, cell)
    keys = O.cell_keys(pos, g)
    gx, gy, gz = g.dims
    assert gx * gy * gz <= 2**27 and max(g.dims) < 16383
    cx, cy, cz = keys % gx, (keys // gx) % gy, keys // (gx * gy)
    d = np.linalg.norm(pos[:, None, :3] - pos[None, :, :3], axis=2)
    ii, jj = np.nonzero(d <= cell)
    assert (np.abs(cx[ii].astype(int) - cx[jj]) <= 1).all()
    assert (np.abs(cy[ii].astype(int) - cy[jj]) <= 1).all()
    assert (np.abs(cz[ii].astype(int) - cz[jj]) <= 1).all()
    # dense key: decodes back to the cell of the position
    q = np.floor((pos[:, :3] - np.array(g.lo[:], np.float32)) * np.float32(g.inv_cell))
    assert (cx == np.clip(q[:, 0], 0, gx - 1)).all()


def test_cell_grid_doubles_until_the_one_cloth_key_fits_27_bits(O):
    """Bounds 1000 x 1000 x 200 with cell_size 1: the first cell, 1 * 65/64,
    gives 985 x 985 x 197 = 191,136,325 cells > 2^27 (every axis is under
    16383, so only the cell count forces a doubling); at 2 * 65/64 the grid is
    493 x 493 x 99 = 24,062,331 cells, which fits."""
    pos = np.zeros((2, 4), np.float32)
    pos[1, :3] = [1000.0, 1000.0, 200.0]
    g = O.cell_grid(pos, 1.0)
    assert g.cell == np.float32(2.03125)
    assert list(g.dims) == [493, 493, 99]
    assert 985 * 985 * 197 > 2**27 >= 493 * 493 * 99


def test_vertex_normals_flat_sheet_is_plus_z(O):
    m = O.grid_cloth_hanging(4, 3, 0.1)
    nrm = O.vertex_normals(m["positions"], m["indices"])
    assert np.allclose(nrm, [0, 0, 1], atol=1e-6)
