"""Jacobi substeps on LDS-resident grid tiles (persistent_jacobi = 2) against
the CPU oracle, through the C ABI.  Needs an MI355X.

One launch runs the predict and every iteration of a substep; each block
keeps its tile (64 columns x 16..64 rows) and a 2-particle ring in LDS, and
only the ring crosses tiles between iterations, ordered by neighbour flags.
The per-particle gather is the grid-tile kernel's, so the bar is the one of
every other layout: bit-exact f32 positions and velocities, substep by
substep.
"""
import time

import numpy as np
import pytest

from _helpers import bits_equal, first_diff, max_abs, oracle_params
from test_gpu_parity import make, note

pytestmark = pytest.mark.gpu


def _sheet(X, cols, rows, seed=0):
    return X.ClothMeshData.sheet(cols=cols, rows=rows, spacing=0.02, pins=2, jitter=1e-3, seed=seed or cols + rows)


def _check_substeps(O, X, sim, cfg, u, f, substeps, collision):
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


def test_resident_plan(X):
    """Tiles cover the cloth (a partial last row of tiles included), the
    height is re-planned and refused outside 16..64 (plan kept); a cloth
    whose columns are not a multiple of 64 has none and refuses the form."""
    note("case jacobi resident plan")
    sim, cfg, _ = make(X, _sheet(X, 128, 72), S=4, I=6, persistent_jacobi=2, thickness=0.016)
    th, nt, npairs, on = sim.jacobi_resident()
    assert on == 1 and nt == 2 * -(-72 // th) and npairs > 0, (th, nt, npairs, on)
    assert sim.jacobi_resident(16)[:2] == (16, 2 * 5)
    assert sim.jacobi_resident(32)[:2] == (32, 2 * 3)
    with pytest.raises(X.XpbdError):
        sim.jacobi_resident(8)
    assert sim.jacobi_resident()[:2] == (32, 6)
    cfg.persistent_jacobi = 0
    sim.set_config(cfg)
    assert sim.jacobi_resident()[3] == 0
    sim.close()
    sim, cfg, _ = make(X, _sheet(X, 48, 40), S=4, I=6, thickness=0.016)
    assert sim.jacobi_resident() == (0, 0, 0, 0)
    cfg.persistent_jacobi = 2
    with pytest.raises(X.XpbdError):
        sim.set_config(cfg)
    sim.close()


@pytest.mark.parametrize("cols,rows,th,fast,coll", [
    (128, 72, 16, 1, True), (128, 72, 32, 1, True), (128, 72, 64, 1, True), (128, 72, 64, 0, True),
    (64, 100, 16, 1, False), (192, 40, 32, 0, False), (256, 33, 16, 1, True)])
def test_resident_bit_exact(O, X, cols, rows, th, fast, coll):
    """Each tile height, a partial last row of tiles, a single column of
    tiles, both division forms, with and without collision (the bounds of
    the last iterate per tile, or the state written in place)."""
    note(f"case jacobi resident {cols}x{rows} th={th} fast={fast} coll={coll}")
    mesh = _sheet(X, cols, rows)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, S=8, I=6, persistent_jacobi=2, fast_division=fast, thickness=0.016)
    assert sim.jacobi_resident(th)[::3] == (th, 1)
    _check_substeps(O, X, sim, cfg, u, f, substeps=3, collision=coll)
    sim.close()


@pytest.mark.parametrize("iters", [1, 2, 3, 10])
def test_resident_iteration_counts(O, X, iters):
    """One iteration (the predict's ring read by the only iteration), two,
    odd and even counts (the last iterate in B or A)."""
    note(f"case jacobi resident I={iters}")
    mesh = _sheet(X, 128, 48, seed=7)
    f = mesh.fields()
    for coll in (True, False):
        sim, cfg, u = make(X, mesh, S=4, I=iters, persistent_jacobi=2, fast_division=1, thickness=0.016)
        assert sim.jacobi_resident(16)[3] == 1
        _check_substeps(O, X, sim, cfg, u, f, substeps=2, collision=coll)
        sim.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_resident_frames_fold(O, X, graph):
    """Frames of the 256x256 folding sheet (contacts, the fast division on
    and off as the window's range changes), replayed from a hipGraph or not."""
    note(f"case jacobi resident fold graph={graph}")
    h = 1 / 255
    kw = dict(cols=256, rows=256, spacing=h, layout=1, height=1.0, angle=1.643758, pins=3)
    mesh = X.ClothMeshData.sheet(**kw)
    f = mesh.fields()
    S = 4
    sim, cfg, u = make(X, mesh, S=S, I=8, persistent_jacobi=2, fast_division=1, use_graph=graph,
                       thickness=2 * (0.4 * h), jacobi_omega=0.32)
    assert sim.jacobi_resident()[3] == 1
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


def test_resident_equals_per_iteration_on_1024_sheet(X):
    """The bench sheet (1024x1024, 8 x 10, collision every substep): the
    resident tiles at 64 and 32 rows and the per-iteration launches (pinned
    to the oracle through the bench's whole run elsewhere) give the same bits
    after two frames."""
    note("case jacobi resident 1024")
    h = 1 / 1023
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    out = []
    for form, th in ((0, 0), (2, 64), (2, 32)):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=8, I=10, persistent_jacobi=form, fast_division=1,
                         use_graph=1, thickness=2 * (0.4 * h))
        if th:
            assert sim.jacobi_resident(th)[::3] == (th, 1)
        sim.step_frames(2)
        out.append(sim.read_state())
        sim.close()
    for k in (1, 2):
        assert bits_equal(out[k][0], out[0][0]), f"form {k}: max |d| {max_abs(out[k][0], out[0][0]):.3e}"
        assert bits_equal(out[k][1], out[0][1])


def test_resident_error_word_ends_launch_promptly(O, X):
    """An expired wait (the error word, set from the host) ends the resident
    launch at once: XPBD_E_STATE, then exact again after a reset."""
    note("case jacobi resident error word")
    mesh = _sheet(X, 128, 40, seed=13)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, S=4, I=8, persistent_jacobi=2, fast_division=1, thickness=0.016)
    sim.raise_device_error()
    t0 = time.perf_counter()
    sim.step_substeps(4, True)
    with pytest.raises(X.XpbdError) as e:
        sim.sync()
    assert e.value.code == X.XPBD_E_STATE
    assert time.perf_counter() - t0 < 2.0
    sim.reset()
    _check_substeps(O, X, sim, cfg, u, f, substeps=2, collision=True)
    sim.close()


def test_resident_long_run_equals_per_iteration(X):
    """Many substeps of the flag protocol: 24 frames (192 substeps, 1 920
    ring exchanges per tile) of the bench sheet on resident tiles stay the
    bits of the per-iteration launches, frame by frame."""
    note("case jacobi resident 24 frames")
    h = 1 / 1023
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    sims = []
    for form in (0, 2):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=8, I=10, persistent_jacobi=form, fast_division=1,
                         use_graph=1, thickness=2 * (0.4 * h))
        assert sim.jacobi_resident()[3] == (1 if form == 2 else 0)
        sims.append(sim)
    for frame in range(24):
        for sim in sims:
            sim.step_frames(1)
        a, b = sims[0].read_state(), sims[1].read_state()
        assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1]), f"frame {frame}"
    for sim in sims:
        sim.close()


@pytest.mark.parametrize("cols,rows", [(64, 17), (128, 200), (320, 96), (640, 333)])
def test_resident_shapes_equal_per_iteration(X, cols, rows):
    """Cloth shapes with partial last rows of tiles at each height the plan
    picks: 4 frames with collision on resident tiles equal the per-iteration
    launches bit for bit."""
    note(f"case jacobi resident shape {cols}x{rows}")
    h = 1 / (cols - 1)
    kw = dict(cols=cols, rows=rows, spacing=h, layout=1, height=0.6, angle=0.4, pins=3)
    out = []
    for form in (0, 2):
        sim, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=8, persistent_jacobi=form, fast_division=1,
                         use_graph=1, thickness=2 * (0.4 * h))
        assert sim.jacobi_resident()[3] == (form == 2)
        sim.step_frames(4)
        out.append(sim.read_state())
        sim.close()
    assert bits_equal(out[0][0], out[1][0]) and bits_equal(out[0][1], out[1][1])


def test_resident_form_switched_at_run_time(X):
    """set_config between frames moves Jacobi between per-iteration launches,
    resident tiles and the grid-barrier launch (the frame graph rebuilt each
    time); the trajectory equals a run that never switches."""
    note("case jacobi resident switch")
    kw = dict(cols=128, rows=96, spacing=1 / 127, layout=1, height=1.0, angle=0.7, pins=3)
    th = 2 * 0.4 / 127
    a, cfg, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=6, use_graph=1, fast_division=1, thickness=th)
    b, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=6, use_graph=1, fast_division=1, thickness=th)
    for k, form in enumerate((2, 0, 1, 2, 0)):
        cfg.persistent_jacobi = form
        a.set_config(cfg)
        assert a.jacobi_resident()[3] == (form == 2)
        a.step_frames(2)
        a.sync()
        b.step_frames(2)
        b.sync()
        pa, va = a.read_state()
        pb, vb = b.read_state()
        assert bits_equal(pa, pb) and bits_equal(va, vb), f"after switch {k} to persistent_jacobi={form}"
    a.close()
    b.close()
