"""The learned warm start on the GPU, through the C ABI, against the CPU
oracle of the network (oracle/mlp_oracle.py) and the plain solver.

- features of a predicted state: bit-exact (bf16) against the oracle;
- BF16-MFMA forward and backward: within the tolerance the bf16 rounding
  points leave (the oracle rounds at the same points, sums in float64);
- the warm-start substep: with a zero network and I warm iterations it is
  the plain substep bit for bit; with threshold 0 every substep falls back
  to the plain substep, bit for bit.
"""
import sys

import numpy as np
import pytest

from _helpers import bits_equal, oracle_params
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_batch(X):
    descs = [X.SheetDesc.make(cols=32, rows=32, spacing=1 / 31, pins=1, jitter=1e-3, seed=s,
                              stretch_compliance=c) for s, c in ((1, 0.0), (2, 1e-7))]
    descs.append(X.SheetDesc.make(cols=40, rows=24, spacing=0.02, layout=1, height=1.0, angle=0.7, pins=3,
                                  jitter=1e-3, seed=3, shear_compliance=1e-4, bend_compliance=1e-2))
    return X.ClothMeshData.batch(descs)


def make(X, mesh, I=10, use_graph=0, coll=False):
    cfg = X.ClothSimConfig.default(solve_substeps=8, solve_inner_iterations=I, use_graph=use_graph)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 8, coll_scale=0.38 if coll else 0.0)
    return X.ClothComputePlugin(mesh, cfg, u), cfg, u


def predicted(O, X, f, pos, vel, u, cfg):
    """Predicted positions of the state (pos, vel): the oracle's predict, i.e.
    one substep with zero iterations and no constraints would be wrong; use
    the product's predict through the features' own definition instead:
    x~ = x + dt * v', v' = v + g dt (speed-capped), as xpbd_cpu.rs:38-65."""
    dt = np.float32(u.dt)
    g = np.asarray(u.gravity[:3], np.float32)
    v = (vel[:, :3] + g * dt).astype(np.float32)
    sp = np.sqrt(((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(np.float32))
    cap = sp > np.float32(12.0)
    v[cap] = (v[cap] * (np.float32(12.0) / sp[cap])[:, None]).astype(np.float32)
    x = (pos[:, :3] + v * dt).astype(np.float32)
    pinned = f["inv_mass"] <= 0
    x[pinned] = pos[pinned, :3]
    return x


def test_features_bit_exact(O, X):
    note("case warm features")
    mesh = small_batch(X)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh)
    sim.step_frames(3)
    pos, vel = sim.read_state()
    n = len(pos)
    d = X.DeviceArray((n, 32), np.uint16)
    sim.warm_features(d)
    got = d.numpy().astype(np.uint32) << 16
    got = got.view(np.float32)
    nbr, aux = M.grid_topology(f)
    xp = predicted(O, X, f, pos, vel, u, cfg)
    want = M.features(xp, vel, f["rest_positions"], nbr, aux, u.dt)
    assert got.tobytes() == want.tobytes(), np.argwhere(got != want)[:5]
    sim.close()


def test_warm_substep_start_is_xp_plus_scaled_forward(O, X):
    """With 0 warm iterations and no gate the substep's result is x0 itself:
    it must equal the predicted state plus spacing * forward / out_scale,
    where forward is the network on the features of that state (both taken
    through their own entry points), bit for bit; pinned vertices keep the
    prediction and grounded ones are clamped to y >= 0."""
    note("case warm x0")
    mesh = small_batch(X)
    f = mesh.fields()
    sim, cfg, u = make(X, mesh)
    sim.step_frames(2)
    pos, vel = sim.read_state()
    n = len(pos)
    d = X.DeviceArray((n, 32), np.uint16)
    sim.warm_features(d)
    out_scale = 2.5
    mlp = X.WarmStartMLP(seed=1)
    mlp.set_params(M.init_params(seed=3), out_scale=out_scale)
    out = mlp.forward(d, n)
    assert np.abs(out).max() > 1e-3
    sim.set_warm_start(mlp, warm_iterations=0, residual_threshold=float("inf"))
    sim.step_substeps(1, False)
    got, _ = sim.read_state()
    _, aux = M.grid_topology(f)
    ih = aux[:, 3].astype(np.float32)
    xp = predicted(O, X, f, pos, vel, u, cfg)
    live = (f["inv_mass"] > 0) & (ih > 0)
    sc = np.float32(np.float32(1.0) / np.float32(out_scale)) / np.where(live, ih, np.float32(1))
    want = xp.copy()
    want[live] = (xp[live] + (out[live] * sc[live, None]).astype(np.float32)).astype(np.float32)
    low = (f["ground"] != 0) & (want[:, 1] < 0)
    want[low, 1] = 0.0
    assert bits_equal(got[:, :3], want), np.argwhere(got[:, :3] != want)[:5]
    sim.close()
    mlp.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradient_against_oracle(X):
    note("case mlp forward/gradient")
    rng = np.random.default_rng(5)
    rows = 1000
    P = M.init_params(seed=7)
    feat = M.bf16(rng.normal(0, 0.5, (rows, 32)).astype(np.float32))
    feat[:, 30:] = 0
    T = rng.normal(0, 1, (rows, 3)).astype(np.float32)
    mlp = X.WarmStartMLP(seed=1)
    mlp.set_params(P, out_scale=1.7)
    dfeat = X.DeviceArray.from_numpy((feat.view(np.uint32) >> 16).astype(np.uint16))
    dT = X.DeviceArray.from_numpy(T)
    out = mlp.forward(dfeat, rows)
    ref, _ = M.forward(P, feat, bf16_points=True)
    note(f"  forward rel err {rel(out, ref):.2e}")
    assert rel(out, ref) < 2e-3
    assert np.max(np.abs(out - ref)) < 2e-2 * np.max(np.abs(ref))
    g, loss = mlp.gradient(dfeat, dT, rows)
    rloss, rg = M.loss_and_grad(P, feat, T, 1.7, bf16_points=True)
    note(f"  loss {loss:.6f} vs {rloss:.6f}")
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    for name, a, b in (("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
                       ("b2", M.OFF_B2, M.OFF_W3), ("W3", M.OFF_W3, M.OFF_B3), ("b3", M.OFF_B3, M.PARAMS)):
        e = rel(g[a:b], rg[a:b])
        note(f"  grad {name}: rel err {e:.2e}")
        assert e < 1e-2, name
    assert np.all(g[M.fixed_zero_mask()] == 0)
    mlp.close()


def test_training_reduces_loss(X):
    note("case mlp training")
    rng = np.random.default_rng(9)
    rows = 8192
    P = M.init_params(seed=3)
    feat = M.bf16(rng.normal(0, 0.5, (rows, 32)).astype(np.float32))
    feat[:, 30:] = 0
    teacher = M.init_params(seed=11)
    T, _ = M.forward(teacher, feat, bf16_points=False)
    T = T.astype(np.float32)
    mlp = X.WarmStartMLP(seed=1)
    mlp.set_params(P, out_scale=1.0)
    dfeat = X.DeviceArray.from_numpy((feat.view(np.uint32) >> 16).astype(np.uint16))
    dT = X.DeviceArray.from_numpy(T)
    curve = mlp.train(dfeat, dT, rows, steps=400, batch=1024, lr=2e-3)
    note(f"  loss {curve[0]:.4f} -> {curve[-1]:.4f}")
    assert curve[-1] < 0.2 * curve[0]
    mlp.close()


@pytest.mark.parametrize("threshold,use_graph,coll", [(float("inf"), 0, False), (1e30, 0, False), (1e30, 1, False),
                                                     (float("inf"), 0, True), (1e30, 1, True)])
def test_zero_network_warm_substep_equals_plain(X, threshold, use_graph, coll):
    """x0 = xp exactly for a zero network, so I warm iterations reproduce the
    plain substep bit for bit (features, apply, iteration chain, the gate
    accepting every substep, selection, collision and velocity update)."""
    note(f"case warm zero network threshold={threshold} graph={use_graph} coll={coll}")
    mesh = small_batch(X)
    a, _, _ = make(X, mesh, I=6, use_graph=use_graph, coll=coll)
    b, _, _ = make(X, mesh, I=6, use_graph=use_graph, coll=coll)
    mlp = X.WarmStartMLP(seed=1)
    mlp.set_params(np.zeros(X.MLP_PARAMS, np.float32), out_scale=1.0)
    a.set_warm_start(mlp, warm_iterations=6, residual_threshold=threshold)
    for k in range(3):
        a.step_substeps(4, False)
        b.step_substeps(4, False)
        pa, va = a.read_state()
        pb, vb = b.read_state()
        assert bits_equal(pa, pb) and bits_equal(va, vb), k
    assert a.warm_stats() == (12, 0)
    a.close()
    b.close()
    mlp.close()


@pytest.mark.parametrize("use_graph,coll", [(0, False), (1, False), (1, True)])
def test_threshold_zero_always_falls_back_to_plain(X, use_graph, coll):
    note(f"case warm fallback graph={use_graph} coll={coll}")
    mesh = small_batch(X)
    a, _, _ = make(X, mesh, I=10, use_graph=use_graph, coll=coll)
    b, _, _ = make(X, mesh, I=10, use_graph=use_graph, coll=coll)
    mlp = X.WarmStartMLP(seed=2)
    P, _ = mlp.params()
    mlp.set_params(P, out_scale=10.0)
    a.set_warm_start(mlp, warm_iterations=2, residual_threshold=0.0)
    # the fallback is a persistent launch: it runs alone (xpbd.h threading)
    a.step_frames(2)
    a.sync()
    b.step_frames(2)
    pa, va = a.read_state()
    pb, vb = b.read_state()
    assert bits_equal(pa, pb) and bits_equal(va, vb)
    assert a.warm_stats() == (16, 16)
    a.close()
    b.close()
    mlp.close()


def test_gate_residual_matches_reported_residual(X):
    """The gate's f32 residual agrees with the reported (f64) one: a threshold
    just above the reported relative residual of the warm result accepts it,
    one just below falls back."""
    note("case warm gate threshold")
    mesh = small_batch(X)
    mlp = X.WarmStartMLP(seed=1)
    mlp.set_params(np.zeros(X.MLP_PARAMS, np.float32), out_scale=1.0)
    ref, _, _ = make(X, mesh, I=6)
    ref.step_substeps(1, False)
    r = ref.residual()
    ref.close()
    assert r > 0
    for scale, want in ((1.0 + 1e-3, 0), (1.0 - 1e-3, 1)):
        a, _, _ = make(X, mesh, I=6)
        a.set_warm_start(mlp, warm_iterations=6, residual_threshold=r * scale)
        a.step_substeps(1, False)
        assert a.warm_stats() == (1, want), scale
        a.close()
    mlp.close()


def test_gated_warm_start_is_reproducible(X):
    """The gate sums its block partials in block order, so two runs of the
    same gated warm start give the same states and the same fallback counts,
    bit for bit (the threshold sits inside the range of the warm residuals,
    where an arrival-order sum could flip a decision)."""
    note("case warm gate reproducible")
    mesh = small_batch(X)
    mlp = X.WarmStartMLP(seed=2)
    P, _ = mlp.params()
    mlp.set_params(P, out_scale=40.0)
    probe, _, _ = make(X, mesh, I=10)
    probe.set_warm_start(mlp, warm_iterations=2, residual_threshold=float("inf"))
    rs = []
    for _ in range(8):
        probe.step_substeps(1, False)
        rs.append(probe.residual())
    probe.close()
    threshold = float(np.median(rs))
    runs = []
    for _ in range(2):
        a, _, _ = make(X, mesh, I=10, use_graph=1)
        a.set_warm_start(mlp, warm_iterations=2, residual_threshold=threshold)
        a.step_frames(2)
        p, v = a.read_state()
        runs.append((p, v, a.warm_stats()))
        a.close()
    note(f"  fallbacks {runs[0][2]} and {runs[1][2]} at threshold {threshold:.3e}")
    assert bits_equal(runs[0][0], runs[1][0]) and bits_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]
    mlp.close()
