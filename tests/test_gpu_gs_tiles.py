"""Coloured Gauss-Seidel ordered by per-tile flags (gs_launch = 3) against the
CPU oracle, through the C ABI.  Needs an MI355X.

The particles are cut into tiles of consecutive indices; a tile's step (its
predict, each colour's constraints of the tile, its final write) waits only
for the tiles whose particles its constraints can share.  Every particle still
sees the colours in order, so the bar is that of the other GS forms: bit-exact
f32 positions and velocities, substep by substep, whatever the tile size and
however many tiles one block runs in turn.

On one grid cloth whose column count is a multiple of 64 the flag form runs
LDS-resident grid tiles (64 columns x 8..64 rows, one block each): a tile's
interior stays in LDS for the whole substep and only a 2-particle ring goes
through memory.  Same bar.
"""
import numpy as np
import pytest

from _helpers import bits_equal, first_diff, max_abs, oracle_params
from test_gpu_parity import make, note

pytestmark = pytest.mark.gpu


def _sheet(X, cols=48, rows=40, seed=13):
    return X.ClothMeshData.sheet(cols=cols, rows=rows, spacing=0.02, pins=2, jitter=1e-3, seed=seed)


def _check_substeps(O, X, sim, cfg, u, f, substeps, collision=True):
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=collision)
    pos = f["positions"].copy()
    vel = np.zeros_like(pos)
    for k in range(substeps):
        note(f"  substep {k}")
        sim.step_substeps(1, collision)
        gp, gv = sim.read_state()
        pos, vel = s.substep(pos, vel, p)
        assert bits_equal(gp, pos), f"substep {k}: positions differ (max |d| {max_abs(gp, pos):.3e}) at " \
                                    f"{first_diff(gp, pos)}"
        assert bits_equal(gv, vel), f"substep {k}: velocities differ"


def test_gs_tile_plan(X):
    """The plan covers every particle, its dependency lists fit one
    wavefront, and a re-plan below the minimum tile is refused and keeps the
    plan in use."""
    note("case gs tile plan")
    mesh = _sheet(X)
    n = mesh.fields()["positions"].shape[0]
    sim, _, _ = make(X, mesh, solver=1, S=4, I=6, gs_launch=3, thickness=0.016)
    tile, nt, blocks, deps = sim.gs_tiles()
    assert tile >= 64 and nt == -(-n // tile) and 1 <= blocks <= nt and deps <= 64
    assert sim.gs_tiles(64, 3) == (64, -(-n // 64), 3, sim.gs_tiles()[3])
    # a 48-wide sheet reaches two rows (96 particles) ahead: 2 tiles of 64 each way at least
    assert 4 <= sim.gs_tiles()[3] <= 64
    with pytest.raises(X.XpbdError):
        sim.gs_tiles(32)
    assert sim.gs_tiles()[:3] == (64, -(-n // 64), 3)
    sim.close()


@pytest.mark.parametrize("tile,blocks", [(64, 1), (64, 3), (64, 0), (128, 5), (512, 0), (4096, 0)],
                         ids=["64x1block", "64x3blocks", "64", "128x5blocks", "512", "one_tile"])
def test_gs_tiles_bit_exact(O, X, tile, blocks):
    """One block running every tile in turn, a few blocks with several tiles
    each, one tile per block, and a single tile, with collision."""
    note(f"case gs tiles tile={tile} blocks={blocks}")
    mesh = _sheet(X)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, solver=1, S=8, I=6, gs_launch=3, thickness=0.016, jacobi_omega=1.0)
    plan = sim.gs_tiles(tile, blocks)
    assert plan[0] == tile and (blocks == 0 or plan[2] == min(blocks, plan[1])), plan
    _check_substeps(O, X, sim, cfg, u, f, substeps=4)
    sim.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_gs_tiles_frames_and_graph(O, X, graph):
    """Whole frames (the substep loop, replayed from a hipGraph or not) on the
    256-wide folding sheet, with its contacts."""
    note(f"case gs tiles 256 fold graph={graph}")
    h = 1 / 255
    kw = dict(cols=256, rows=96, spacing=h, layout=1, height=1.0, angle=1.643758, pins=3)
    mesh = X.ClothMeshData.sheet(**kw)
    f = mesh.fields()
    S = 4
    sim, cfg, u = make(X, mesh, solver=1, S=S, I=8, gs_launch=3, use_graph=graph, thickness=2 * (0.4 * h),
                       jacobi_omega=1.0)
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    pos, vel = f["positions"].copy(), np.zeros_like(f["positions"])
    for frame in range(3):
        sim.step_frames(1)
        pos, vel = s.frames(pos, vel, p, S, 1, 1)
        gp, gv = sim.read_state()
        assert bits_equal(gp, pos), f"frame {frame}: max |d| {max_abs(gp, pos):.3e}"
        assert bits_equal(gv, vel), f"frame {frame}: velocities differ"
    sim.close()


def test_gs_tiles_equal_colours_on_1024_sheet(X):
    """The bench sheet (1024x1024): per-tile flags (LDS grid tiles of 64 and
    32 rows, index-range tiles of 256 particles) and one launch per colour
    (itself pinned to the oracle at smaller sizes) give the same bits after a
    frame of 8 substeps with collision."""
    note("case gs tiles 1024")
    h = 1 / 1023
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    out = []
    for form, tile, rows in ((2, 0, 0), (3, 0, 0), (3, 256, 0), (3, 0, 32)):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), solver=1, S=8, I=10, gs_launch=form, use_graph=1,
                         thickness=2 * (0.4 * h), jacobi_omega=1.0)
        if tile:
            sim.gs_tiles(tile)
        if rows:
            sim.gs_grid(rows)
        if form == 3:
            assert sim.gs_grid()[3] == (0 if tile else 1)
        sim.step_frames(1)
        out.append(sim.read_state())
        sim.close()
    for k in (1, 2, 3):
        assert bits_equal(out[k][0], out[0][0]), f"form {k}: max |d| {max_abs(out[k][0], out[0][0]):.3e}"
        assert bits_equal(out[k][1], out[0][1])


def test_gs_grid_plan(X):
    """Grid tiles cover the cloth (partial last row of tiles included), the
    height is re-planned, refused below 8 rows (plan kept), and index-range
    tiles switch them off; a cloth whose columns are not a multiple of 64 has
    none."""
    note("case gs grid plan")
    sim, _, _ = make(X, _sheet(X, cols=128, rows=72), solver=1, S=4, I=6, gs_launch=3, thickness=0.016)
    th, nt, threads, on = sim.gs_grid()
    assert on == 1 and nt == 2 * -(-72 // th) and threads >= 256, (th, nt, threads, on)
    assert sim.gs_grid(16) == (16, 2 * 5, 256, 1)
    assert sim.gs_grid(8) == (8, 2 * 9, 256, 1)
    with pytest.raises(X.XpbdError):
        sim.gs_grid(4)
    assert sim.gs_grid() == (8, 18, 256, 1)
    sim.gs_tiles(128)
    assert sim.gs_grid()[3] == 0
    sim.close()
    sim, _, _ = make(X, _sheet(X), solver=1, S=4, I=6, gs_launch=3, thickness=0.016)
    assert sim.gs_grid() == (0, 0, 0, 0)
    sim.close()


@pytest.mark.parametrize("cols,rows,th", [(128, 72, 8), (128, 72, 16), (128, 72, 32), (128, 72, 64), (64, 100, 16),
                                          (192, 40, 0)])
def test_gs_grid_bit_exact(O, X, cols, rows, th):
    """LDS grid tiles of each height, a partial last row of tiles, a single
    column of tiles, and the create-time height, with collision."""
    note(f"case gs grid {cols}x{rows} th={th}")
    mesh = _sheet(X, cols=cols, rows=rows, seed=cols + rows)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, solver=1, S=8, I=6, gs_launch=3, thickness=0.016, jacobi_omega=1.0)
    plan = sim.gs_grid(th) if th else sim.gs_grid()
    assert plan[3] == 1 and (th == 0 or plan[0] == th), plan
    _check_substeps(O, X, sim, cfg, u, f, substeps=4)
    sim.close()


@pytest.mark.parametrize("iters", [1, 2])
def test_gs_grid_iteration_counts(O, X, iters):
    """One and two iterations: the first colour step follows the predict, the
    final write follows the last colour step directly."""
    note(f"case gs grid I={iters}")
    mesh = _sheet(X, cols=128, rows=48, seed=7)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, solver=1, S=4, I=iters, gs_launch=3, thickness=0.016, jacobi_omega=1.0)
    assert sim.gs_grid(16)[3] == 1
    _check_substeps(O, X, sim, cfg, u, f, substeps=3)
    sim.close()


def test_gs_grid_long_run_equals_per_colour(X):
    """Many colour steps of the flag protocol: 8 frames (64 substeps, 7 680
    colour steps per tile) of the 1024x1024 sheet on LDS grid tiles stay the
    bits of one launch per colour, frame by frame."""
    note("case gs grid 8 frames")
    h = 1 / 1023
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    sims = []
    for form in (2, 3):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), solver=1, S=8, I=10, gs_launch=form, use_graph=1,
                         thickness=2 * (0.4 * h), jacobi_omega=1.0)
        sims.append(sim)
    assert sims[1].gs_grid()[3] == 1
    for frame in range(8):
        for sim in sims:
            sim.step_frames(1)
        a, b = sims[0].read_state(), sims[1].read_state()
        assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1]), f"frame {frame}"
    for sim in sims:
        sim.close()


@pytest.mark.parametrize("cols,rows", [(64, 17), (128, 200), (320, 96), (640, 333)])
def test_gs_grid_shapes_equal_per_colour(X, cols, rows):
    """Cloth shapes with partial last rows of grid tiles: 3 frames with
    collision on LDS grid tiles equal one launch per colour bit for bit."""
    note(f"case gs grid shape {cols}x{rows}")
    h = 1 / (cols - 1)
    kw = dict(cols=cols, rows=rows, spacing=h, layout=1, height=0.6, angle=0.4, pins=3)
    out = []
    for form in (2, 3):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), solver=1, S=4, I=6, gs_launch=form, use_graph=1,
                         thickness=2 * (0.4 * h), jacobi_omega=1.0)
        if form == 3:
            assert sim.gs_grid()[3] == 1
        sim.step_frames(3)
        out.append(sim.read_state())
        sim.close()
    assert bits_equal(out[0][0], out[1][0]) and bits_equal(out[0][1], out[1][1])
