"""HIP path vs the CPU oracle, through the C ABI.  Needs an MI355X.

Bar: bit-exact f32 positions and velocities against the f32 oracle, substep
by substep (the kernels restate the CPU arithmetic operation by operation);
the reference's own looser GPU/CPU tolerances (2e-3, 3e-2,
gpu_cpu_parity_jacobi.rs:912,1015; 2.5e-2, gpu_cpu_parity.rs:977) are
asserted too, as the reference tests do.
"""
import sys

import numpy as np
import pytest

from _helpers import DeviceBuffer, bits_equal, first_diff, max_abs, oracle_params

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# (solver, gs_launch): Jacobi; coloured GS as one persistent launch per
# substep with grid barriers; coloured GS as one launch per colour and
# iteration; coloured GS as one persistent launch ordered by per-tile flags
SOLVER_FORMS = [(0, 0), (1, 1), (1, 2), (1, 3)]
SOLVER_FORM_IDS = ["jacobi", "gs_persistent", "gs_colours", "gs_tiles"]


def make(X, mesh, solver=0, S=8, I=10, coll_stride=1, use_graph=0, fused_predict=0, jacobi_block=0,
         plain_adjacency=0, fast_division=0, persistent_jacobi=0, grid_tiles=1, autotune=0, gs_launch=0, **uk):
    cfg = X.ClothSimConfig.default(solve_substeps=S, solve_inner_iterations=I, solver=solver,
                                   collision_every_n_substeps=coll_stride, use_graph=use_graph,
                                   fused_predict=fused_predict, jacobi_block=jacobi_block,
                                   plain_adjacency=plain_adjacency, fast_division=fast_division,
                                   persistent_jacobi=persistent_jacobi, grid_tiles=grid_tiles, autotune=autotune,
                                   gs_launch=gs_launch)
    u = X.ClothSimUniforms.default(solver, S, **uk)
    return X.ClothComputePlugin(mesh, cfg, u), cfg, u


def run_pair(O, X, mesh_kw=None, mesh=None, solver=0, S=8, I=10, substeps=8, collision=False, fused_predict=0,
             jacobi_block=0, plain_adjacency=0, fast_division=0, persistent_jacobi=0, grid_tiles=1, layout=None,
             autotune=0, gs_launch=0, **uk):
    mesh = mesh or X.ClothMeshData.sheet(**mesh_kw)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, solver, S, I, fused_predict=fused_predict, jacobi_block=jacobi_block,
                       plain_adjacency=plain_adjacency, fast_division=fast_division,
                       persistent_jacobi=persistent_jacobi, grid_tiles=grid_tiles, autotune=autotune,
                       gs_launch=gs_launch, **uk)
    if layout is not None:
        assert sim.layout()["adjacency"] == layout, sim.layout()
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=collision)
    pos = f["positions"].copy()
    vel = np.zeros_like(pos)
    for k in range(substeps):
        note(f"  substep {k}")
        sim.step_substeps(1, collision)
        gp, gv = sim.read_state()
        pos, vel = s.substep(pos, vel, p)
        err = max_abs(gp, pos)
        assert bits_equal(gp, pos), f"substep {k}: positions differ (max |d| {err:.3e}) at {first_diff(gp, pos)}"
        assert bits_equal(gv, vel), f"substep {k}: velocities differ"
    sim.close()
    return pos, vel


@pytest.mark.parametrize("solver,gs_launch", SOLVER_FORMS, ids=SOLVER_FORM_IDS)
def test_config1_sheet_bit_exact_per_substep(O, X, solver, gs_launch):
    note(f"case config1 32x32 solver={solver} gs_launch={gs_launch}")
    run_pair(O, X, dict(cols=32, rows=32, spacing=1 / 31, pins=1, jitter=1e-3, seed=1), solver=solver,
             S=8, I=10, substeps=16, gs_launch=gs_launch)


@pytest.mark.parametrize("plain", [0, 1])
@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("iters", [1, 2, 3])
def test_jacobi_iteration_count_edge_cases(O, X, iters, fused, plain):
    note(f"case iters={iters} fused_predict={fused} plain_adjacency={plain}")
    for coll in (False, True):
        run_pair(O, X, dict(cols=20, rows=9, spacing=0.05, pins=2, jitter=1e-3, seed=2), S=8, I=iters,
                 substeps=4, collision=coll, fused_predict=fused, plain_adjacency=plain)


@pytest.mark.parametrize("fast", [0, 1])
@pytest.mark.parametrize("plain", [0, 1])
@pytest.mark.parametrize("iters", [1, 2, 3, 10])
def test_persistent_jacobi_bit_exact(O, X, iters, plain, fast):
    """Predict and all iterations of a substep in one launch (grid barriers
    between iterations, the slice tame bytes carried across them for the fast
    division): the same results as one launch per iteration."""
    note(f"case persistent iters={iters} plain_adjacency={plain} fast={fast}")
    for coll in (False, True):
        run_pair(O, X, dict(cols=70, rows=33, spacing=0.03, pins=2, jitter=1e-3, seed=9), S=8, I=iters,
                 substeps=3, collision=coll, plain_adjacency=plain, persistent_jacobi=1, fast_division=fast)


def test_reference_grid_plain_adjacency(O, X):
    """The 12-byte adjacency path on the reference's own grid builder."""
    note("case grid18 plain")
    mesh = X.ClothMeshData.grid_cloth_hanging(18, 18, 0.042)
    run_pair(O, X, mesh=mesh, S=36, I=22, substeps=2, collision=True, plain_adjacency=1)


@pytest.mark.parametrize("fast_division", [0, 1])
@pytest.mark.parametrize("block", [64, 128, 512])
def test_jacobi_block_sizes(O, X, block, fast_division):
    note(f"case block={block} fast_division={fast_division}")
    run_pair(O, X, dict(cols=70, rows=33, spacing=0.03, pins=2, jitter=1e-3, seed=8), S=8, I=5, substeps=3,
             collision=True, jacobi_block=block, fast_division=fast_division)


@pytest.mark.parametrize("cols,rows,fast,fused", [(64, 37, 0, 0), (64, 37, 1, 0), (128, 16, 1, 1), (128, 16, 0, 1),
                                                (64, 37, 1, 2), (128, 16, 0, 2)])
def test_grid_tiles_bit_exact(O, X, cols, rows, fast, fused):
    """Grid-tiled Jacobi (64 x 8 LDS tiles with a 2-particle halo): columns a
    multiple of 64, a partial last band of tiles (37 rows), predict fused
    into the first iteration (0: automatic = on for tiles, 1: on) or separate
    (2), both division forms, with collision."""
    note(f"case grid tiles {cols}x{rows} fast={fast} fused={fused}")
    kw = dict(cols=cols, rows=rows, spacing=0.02, pins=2, jitter=1e-3, seed=cols + rows)
    run_pair(O, X, kw, S=8, I=5, substeps=3, collision=True, fast_division=fast, fused_predict=fused,
             thickness=0.016, layout="grid_tiles")
    run_pair(O, X, kw, S=8, I=5, substeps=1, collision=True, fast_division=fast, fused_predict=fused,
             thickness=0.016, grid_tiles=0, layout="packed")


def test_grid_tiles_only_where_they_fit(X):
    """Tiles need one grid cloth whose column count is a multiple of 64."""
    note("case grid tiles eligibility")
    cases = [(dict(cols=64, rows=9), "grid_tiles"), (dict(cols=70, rows=9), "packed"),
             (dict(cols=32, rows=32), "packed")]
    for kw, want in cases:
        sim, _, _ = make(X, X.ClothMeshData.sheet(spacing=0.02, pins=2, **kw))
        assert sim.layout()["adjacency"] == want, (kw, sim.layout())
        sim.close()
    sim, _, _ = make(X, X.ClothMeshData.sheet(cols=64, rows=9, spacing=0.02, pins=2), plain_adjacency=1)
    assert sim.layout()["adjacency"] == "plain"
    sim.close()


def test_collision_dense_cluster_overflow_path(O, X):
    """A sheet squeezed to a fiftieth of its size: every particle touches many
    partners, exercising the sweep's overflow kernel: its 32-entry LDS list
    (9..32 candidates) and its selection fallback (more than 32)."""
    note("case dense cluster")
    mesh = X.ClothMeshData.sheet(cols=24, rows=24, spacing=0.01, pins=0, jitter=1e-3, seed=11)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, 0, S=8, I=2, thickness=0.004)
    pos = f["positions"].copy()
    pos *= np.float32(0.02)
    sim.write_state(pos)
    sim.step_substeps(1, True)
    gp, gv = sim.read_state()
    p = oracle_params(O, X, u, cfg, collision=True)
    s = O.Solver(f)
    op, ov = s.substep(pos, np.zeros_like(pos), p)
    assert bits_equal(gp, op) and bits_equal(gv, ov), max_abs(gp, op)
    # candidates per particle in the pre-collision iterate: beyond the sweep's
    # 8-entry list, both inside and beyond the overflow kernel's 32-entry list
    p0 = O.params(substeps=8, inner_iterations=2, omega=float(u.jacobi_omega), thickness=0.004)
    q, _ = s.substep(pos, np.zeros_like(pos), p0)
    d = np.linalg.norm(q[:, None, :] - q[None, :, :], axis=2)
    cand = np.sum(d <= 0.004, axis=1) - 1
    assert ((cand > 8) & (cand <= 32)).any() and (cand > 32).any(), (cand.min(), cand.max())


@pytest.mark.parametrize("solver,gs_launch", SOLVER_FORMS, ids=SOLVER_FORM_IDS)
def test_reference_triangle_one_substep(O, X, solver, gs_launch):
    """gpu_cpu_one_substep_single_triangle_positions_close
    (gpu_cpu_parity_jacobi.rs:817-924; gpu_cpu_parity.rs:821-907): one substep
    with collision on, gravity (-1,-2,0.5), SUBSTEPS=36, INNER_ITERS=22."""
    note(f"case triangle solver={solver}")
    mesh = X.ClothMeshData.parse_welded_obj("\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n")
    pos, _ = run_pair(O, X, mesh=mesh, solver=solver, S=36, I=22, substeps=1, collision=True,
                      gravity=(-1.0, -2.0, 0.5), grab_stiffness=0.0, gs_launch=gs_launch)
    assert np.isfinite(pos).all()


@pytest.mark.parametrize("solver,gs_launch", SOLVER_FORMS, ids=SOLVER_FORM_IDS)
def test_reference_grid_18x18_one_substep(O, X, solver, gs_launch):
    """gpu_cpu_grid_cloth_substep_positions_close (gpu_cpu_parity_jacobi.rs:926-1027,
    tolerance 3e-2; gpu_cpu_parity.rs:909-989, 2.5e-2) — here bit-exact."""
    note(f"case grid18 solver={solver}")
    mesh = X.ClothMeshData.grid_cloth_hanging(18, 18, 0.042)
    run_pair(O, X, mesh=mesh, solver=solver, S=36, I=22, substeps=3, collision=True, gs_launch=gs_launch)


def test_reference_grab_and_frames(O, X):
    """A whole frame schedule with the grab term active (predict, xpbd_cpu.rs:56-63)."""
    note("case grab frames")
    mesh = X.ClothMeshData.grid_cloth_hanging(16, 16, 0.045)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh, 0, S=12, I=8, grab_idx=5, grab_active=1, grab_target=(0.3, 0.2, 0.4))
    sim.step_frames(3)
    gp, gv = sim.read_state()
    p = oracle_params(O, X, u, cfg, collision=True)
    pos, vel = O.Solver(f).frames(f["positions"], np.zeros_like(f["positions"]), p, 12, 1, 3)
    assert bits_equal(gp, pos) and bits_equal(gv, vel), max_abs(gp, pos)
    sim.close()


def test_folding_sheet_with_collision_bit_exact(O, X):
    """Config-2 style: horizontal sheet, centre pin, falls and folds onto
    itself; contacts must be resolved identically to the brute-force order."""
    note("case fold 40x40")
    kw = dict(cols=40, rows=40, spacing=1 / 39, layout=1, height=1.0, angle=0.9, pins=3)
    mesh = X.ClothMeshData.sheet(**kw)
    f = mesh.fields()
    h = 1 / 39
    th = 2 * (0.4 * h)
    sim, cfg, u = make(X, mesh, 0, S=4, I=8, thickness=th)
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    pos = f["positions"].copy()
    vel = np.zeros_like(pos)
    contacts = 0
    for frame in range(40):
        sim.step_frames(1)
        pos, vel = s.frames(pos, vel, p, 4, 1, 1)
        if frame % 5 == 4:
            note(f"  frame {frame}")
            gp, gv = sim.read_state()
            assert bits_equal(gp, pos), f"frame {frame}: max |d| {max_abs(gp, pos):.3e}"
            assert bits_equal(gv, vel)
    # the brute-force loop agrees with the grid oracle on the final state
    pb = O.params()
    q1 = s.self_collision(pos, th, 0.38, O.COLL_BRUTE)
    q2 = s.self_collision(pos, th, 0.38, O.COLL_GRID)
    assert q1.tobytes() == q2.tobytes()
    contacts = int(np.any(q1 != pos, axis=1).sum())
    note(f"  particles in contact at the end: {contacts}")
    assert contacts > 0
    sim.close()


def test_collision_keys_and_sort_bit_exact(O, X):
    note("case keys")
    mesh = X.ClothMeshData.sheet(cols=64, rows=48, spacing=0.01, layout=1, height=1.0, angle=0.3, pins=3,
                                 jitter=2e-3, seed=4)
    sim, _, _ = make(X, mesh)
    pos, _ = sim.read_state()
    pos[:, :3] += np.random.default_rng(1).normal(0, 0.05, (len(pos), 3)).astype(np.float32)
    sim.write_state(pos)
    for cell in (0.008, 0.02, 0.1):
        g, keys, ids = sim.debug_collision_keys(cell)
        og = O.cell_grid(pos, cell)
        assert np.array(og.lo[:], np.float32).tobytes() == g["lo"].tobytes()
        assert list(og.dims) == list(g["dims"])
        assert np.float32(og.inv_cell) == g["inv_cell"]
        ok = O.cell_keys(pos, og)
        sk, si = O.sort_pairs(ok)
        assert (keys == sk).all() and (ids == si).all()
    sim.close()


def test_collision_sort_wide_keys_many_tiles(O, X):
    """One cloth's keys sort in 3 passes of 9-bit digits: 90,000 particles
    (44 sort tiles) scattered through a cube whose grid nearly fills the
    2^27-cell key space, so every digit of every pass varies; and the same
    particles packed into a few cells (long equal-key runs)."""
    note("case wide keys")
    mesh = X.ClothMeshData.sheet(cols=300, rows=300, spacing=0.01, pins=0, jitter=0.0, seed=1)
    sim, _, _ = make(X, mesh)
    pos, _ = sim.read_state()
    rng = np.random.default_rng(12)
    for cell, spread in ((0.00215, 1.0), (0.1, 0.15)):
        pos[:, :3] = rng.uniform(0.0, spread, (len(pos), 3)).astype(np.float32)
        sim.write_state(pos)
        g, keys, ids = sim.debug_collision_keys(cell)
        og = O.cell_grid(pos, cell)
        assert list(og.dims) == list(g["dims"])
        cells = int(np.prod(np.asarray(g["dims"], np.int64)))
        assert cells <= 2**27
        ok = O.cell_keys(pos, og)
        if spread == 1.0:
            assert cells > 2**26 and int(ok.max()) >= 2**26   # the top digit is in use
        sk, si = O.sort_pairs(ok)
        assert (keys == sk).all() and (ids == si).all()
    sim.close()


@pytest.mark.parametrize("bound", [False, True])
def test_publish_layout_and_normals(O, X, bound):
    """finalize_normals / render_positions layout (cloth_sim_jacobi.wgsl:527-547):
    vec4<f32>, w = 0, indexed by particle id; normals = normalised sum of
    incident face normals, checked bit-exact against the oracle."""
    note(f"case publish bound={bound}")
    mesh = X.ClothMeshData.sheet(cols=33, rows=21, spacing=0.03, pins=2, jitter=1e-3, seed=3)
    f = mesh.fields()
    sim, _, _ = make(X, mesh, S=4, I=6, use_graph=1)
    dpos = DeviceBuffer(16 * sim.n)
    dnrm = DeviceBuffer(16 * sim.n)
    dpos.upload(np.full((sim.n, 4), 7.0, np.float32))
    dnrm.upload(np.full((sim.n, 4), 7.0, np.float32))
    if bound:
        sim.bind_render_buffers(dpos.ptr.value, dnrm.ptr.value)
        sim.step_frames(2)
    else:
        sim.step_frames(2)
        sim.publish(dpos.ptr.value, dnrm.ptr.value)
    sim.sync()
    gp, _ = sim.read_state()
    rp = dpos.download(np.float32, (sim.n, 4))
    rn = dnrm.download(np.float32, (sim.n, 4))
    assert (rp[:, 3] == 0).all() and (rn[:, 3] == 0).all()
    assert bits_equal(rp, gp)
    on = O.vertex_normals(gp[:, :3], f["indices"])
    assert bits_equal(rn, on)
    sim.close()
    dpos.free()
    dnrm.free()


@pytest.mark.parametrize("persistent", [0, 1])
def test_graph_replay_equals_eager(X, persistent):
    note(f"case graph persistent={persistent}")
    kw = dict(cols=48, rows=48, spacing=1 / 47, layout=1, height=1.0, angle=0.4, pins=3)
    th = 2 * 0.4 / 47
    a, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=8, use_graph=1, thickness=th,
                   persistent_jacobi=persistent)
    b, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=8, use_graph=0, thickness=th)
    for k in range(4):
        a.step_frames(5)
        a.sync()   # a persistent launch runs alone on the GPU (xpbd.h threading)
        b.step_frames(5)
        pa, va = a.read_state()
        pb, vb = b.read_state()
        assert bits_equal(pa, pb) and bits_equal(va, vb), f"replay {k}"
    a.close()
    b.close()


def test_gs_launch_form_switched_at_run_time(X):
    """set_config between frames moves coloured GS between its three forms
    (barriers, one launch per colour, per-tile flags; the frame graph is
    rebuilt each time); the trajectory equals a run that never switches, and
    the GS substep is timed in each form."""
    note("case gs launch switch")
    kw = dict(cols=48, rows=40, spacing=1 / 47, layout=1, height=1.0, angle=0.7, pins=3)
    th = 2 * 0.4 / 47
    a, cfg, _ = make(X, X.ClothMeshData.sheet(**kw), solver=1, S=4, I=6, use_graph=1, thickness=th, gs_launch=1,
                     jacobi_omega=1.0)
    b, _, _ = make(X, X.ClothMeshData.sheet(**kw), solver=1, S=4, I=6, use_graph=1, thickness=th, gs_launch=1,
                   jacobi_omega=1.0)
    for k, form in enumerate((2, 3, 1, 3, 2)):
        cfg.gs_launch = form
        a.set_config(cfg)
        a.step_frames(3)
        a.sync()
        b.step_frames(3)
        b.sync()
        pa, va = a.read_state()
        pb, vb = b.read_state()
        assert bits_equal(pa, pb) and bits_equal(va, vb), f"after switch {k} to gs_launch={form}"
        assert a.time_kernel(X.KERNEL_GS_SUBSTEP, 3) > 0.0
    a.close()
    b.close()


def test_profile_frames_counts_launches_and_steps_like_frames(X):
    """The instrumented frames advance the state exactly as step_frames does
    and time one event pair per launch of the chosen kind."""
    note("case profile_frames")
    kw = dict(cols=40, rows=30, spacing=0.02, pins=2, jitter=1e-3, seed=12)
    a, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=6, use_graph=1, thickness=0.016)
    b, _, _ = make(X, X.ClothMeshData.sheet(**kw), S=4, I=6, use_graph=1, thickness=0.016)
    ms, cnt = a.profile_frames(3, X.KERNEL_JACOBI_ITER)
    assert cnt == 3 * 4 * 5 and ms > 0
    ms, cnt = a.profile_frames(2, X.KERNEL_COLLISION_SWEEP)
    assert cnt == 2 * 4 and ms > 0
    b.step_frames(5)
    pa, va = a.read_state()
    pb, vb = b.read_state()
    assert bits_equal(pa, pb) and bits_equal(va, vb)
    a.close()
    b.close()


def test_reset_and_state_roundtrip(X):
    note("case reset")
    mesh = X.ClothMeshData.sheet(cols=16, rows=16, spacing=0.05, pins=1, jitter=1e-3, seed=5)
    f = mesh.fields()
    sim, _, _ = make(X, mesh, S=4, I=4)
    p0, v0 = sim.read_state()
    assert bits_equal(p0, f["positions"]) and (v0 == 0).all() and (p0[:, 3] == 0).all()
    sim.step_frames(3)
    p1, _ = sim.read_state()
    assert not bits_equal(p1, p0)
    sim.control.reinit_serial += 1
    sim.control.sim_paused = True
    assert sim.run_cloth_sim() is False
    p2, v2 = sim.read_state()
    assert bits_equal(p2, f["positions"]) and (v2 == 0).all()
    sim.control.step_serial += 1
    assert sim.run_cloth_sim() is True
    sim.close()


def test_residual_matches_oracle(O, X):
    note("case residual")
    mesh = X.ClothMeshData.sheet(cols=40, rows=30, spacing=0.02, pins=2, jitter=1e-3, seed=6)
    f = mesh.fields()
    sim, _, _ = make(X, mesh, S=8, I=10)
    sim.step_frames(2)
    gp, _ = sim.read_state()
    r_gpu = sim.residual()
    r_cpu = O.Solver(f).residual(gp)
    assert r_gpu > 0 and abs(r_gpu - r_cpu) <= 1e-9 * max(r_cpu, 1e-30) + 1e-15
    sim.close()


@pytest.mark.parametrize("solver,gs_launch", SOLVER_FORMS, ids=SOLVER_FORM_IDS)
def test_config2_256_collision_substeps_bit_exact(O, X, solver, gs_launch):
    """256x256 horizontal sheet, centre pin, seeded rotation, self-collision."""
    note(f"case 256 solver={solver} gs_launch={gs_launch}")
    h = 1 / 255
    rng = np.random.default_rng(2)
    kw = dict(cols=256, rows=256, spacing=h, layout=1, height=1.0, angle=float(rng.uniform(0, 2 * np.pi)), pins=3)
    run_pair(O, X, kw, solver=solver, S=4, I=8, substeps=3, collision=True, thickness=2 * (0.4 * h),
             jacobi_omega=0.32 if solver == 0 else 1.0, gs_launch=gs_launch)


@pytest.mark.parametrize("solver", [0, 1])
def test_autotuned_launch_shapes_bit_exact(O, X, solver):
    """Autotuned launch shapes (Jacobi block and block order; the coloured
    GS grid size) give the same results as the defaults."""
    note(f"case autotune solver={solver}")
    kw = dict(cols=128, rows=96, spacing=0.01, pins=2, jitter=1e-3, seed=21)
    run_pair(O, X, kw, solver=solver, S=8, I=6, substeps=3, collision=True, thickness=0.008, autotune=1,
             jacobi_omega=0.32 if solver == 0 else 1.0)


def test_config3_1024_two_substeps_bit_exact(O, X):
    note("case 1024")
    h = 1 / 1023
    O.set_threads(16)
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    run_pair(O, X, kw, S=8, I=10, substeps=2, collision=True, thickness=2 * (0.4 * h))


def test_residual_is_reproducible(X):
    """The reported residual sums fixed per-block partials in a fixed order
    (no atomics): the same state gives the same bits on every call."""
    note("case residual reproducible")
    mesh = X.ClothMeshData.sheet(cols=96, rows=80, spacing=0.01, pins=2, jitter=1e-3, seed=8)
    sim, _, _ = make(X, mesh, S=8, I=10)
    sim.step_frames(3)
    rs = {sim.residual() for _ in range(5)}
    assert len(rs) == 1 and next(iter(rs)) > 0, rs
    sim.close()


@pytest.mark.parametrize("case", ["heavy", "negative_compliance", "plain"])
def test_fast_division_preconditions(O, X, case):
    """The short operations (and the tests they let the Jacobi entry drop)
    apply only where their preconditions hold: inverse masses 0 or >= 2^-26,
    compliances >= 0.  A mesh outside them runs the general path, still bit
    for bit; a mesh inside runs the fast path, bit for bit."""
    note(f"case fast preconditions {case}")
    kw = dict(cols=64, rows=24, spacing=0.02, pins=2, jitter=1e-3, seed=5)
    if case == "heavy":
        kw["inv_mass"] = 1e-9
    elif case == "negative_compliance":
        kw["bend_compliance"] = -1e-7
    descs = [X.SheetDesc.make(**kw)]
    mesh = X.ClothMeshData.batch(descs)
    sim, _, _ = make(X, mesh, fast_division=1)
    assert sim.layout()["fast_division"] == (case == "plain"), sim.layout()
    sim.close()
    run_pair(O, X, mesh=mesh, I=6, substeps=4, fast_division=1)
