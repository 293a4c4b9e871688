"""Training and evaluation of the learned warm start, on the GPU.

- dataset: seeded cloths of the two training shapes (32x32 hanging sheets,
  two corners pinned, and a 256x256 horizontal sheet, centre pinned, at a
  seeded rotation), log-uniform compliances, simulated by the plain Jacobi
  solver; at chosen frames each vertex gives one pair (features of the
  predicted state, (state after `target_iterations` plain iterations -
  predicted) / spacing), written straight into device buffers;
- training: Adam on the BF16 MFMA kernels (WarmStartMLP.train);
- evaluation: from one state, the relative residual after one substep for
  the plain solver at several iteration counts and for warm start + k
  iterations.

No reference counterpart (the reference has no learned component).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import api as A
from .workloads import SplitMix64


def _log_uniform(rng: SplitMix64, lo: float, hi: float) -> float:
    a, b = math.log10(lo), math.log10(hi)
    return 10.0 ** (a + (b - a) * rng.unit())


def training_descs(seed: int = 1, small: int = 32, large: int = 1):
    """`small` 32x32 hanging sheets and `large` 256x256 horizontal sheets."""
    rng = SplitMix64(seed)
    descs = []
    for _ in range(small):
        d = A.SheetDesc.make(cols=32, rows=32, spacing=1.0 / 31.0, layout=A.LAYOUT_HANGING_XY, height=0.0,
                             pins=A.PIN_TWO_CORNERS if rng.below(2) == 0 else A.PIN_TOP_ROW)
        d.jitter, d.seed = 1e-3, rng.next()
        d.stretch_compliance = 0.0 if rng.below(4) == 0 else _log_uniform(rng, 1e-8, 1e-3)
        d.shear_compliance = _log_uniform(rng, 1e-8, 1e-3)
        d.bend_compliance = _log_uniform(rng, 1e-8, 1e-3)
        descs.append(d)
    for _ in range(large):
        d = A.SheetDesc.make(cols=256, rows=256, spacing=1.0 / 255.0, layout=A.LAYOUT_HORIZONTAL_XZ, height=1.0,
                             angle=2.0 * math.pi * rng.unit(), pins=A.PIN_CENTER)
        d.jitter, d.seed = 1e-3, rng.next()
        d.stretch_compliance = _log_uniform(rng, 1e-8, 1e-3)
        d.shear_compliance = _log_uniform(rng, 1e-8, 1e-3)
        d.bend_compliance = _log_uniform(rng, 1e-8, 1e-3)
        descs.append(d)
    return descs


def eval_descs(seed: int = 101, count: int = 4):
    """The evaluation layout: `count` 256x256 sheets in one batch, hanging
    (two corners or top row) or horizontal (centre pin), seeded compliances."""
    rng = SplitMix64(seed)
    descs = []
    for k in range(count):
        if k % 2 == 0:
            d = A.SheetDesc.make(cols=256, rows=256, spacing=1.0 / 255.0, layout=A.LAYOUT_HANGING_XY, height=0.0,
                                 pins=A.PIN_TWO_CORNERS if rng.below(2) == 0 else A.PIN_TOP_ROW)
        else:
            d = A.SheetDesc.make(cols=256, rows=256, spacing=1.0 / 255.0, layout=A.LAYOUT_HORIZONTAL_XZ,
                                 height=1.0, angle=2.0 * math.pi * rng.unit(), pins=A.PIN_CENTER)
        d.jitter, d.seed = 1e-3, rng.next()
        d.stretch_compliance = _log_uniform(rng, 1e-8, 1e-3)
        d.shear_compliance = _log_uniform(rng, 1e-8, 1e-3)
        d.bend_compliance = _log_uniform(rng, 1e-8, 1e-3)
        descs.append(d)
    return descs


def make_sim(mesh, substeps: int = 8, iterations: int = 10, use_graph: int = 0):
    cfg = A.ClothSimConfig.default(solve_substeps=substeps, solve_inner_iterations=iterations, use_graph=use_graph)
    u = A.ClothSimUniforms.default(A.SOLVER_JACOBI, substeps, coll_scale=0.0)
    return A.ClothComputePlugin(mesh, cfg, u)


def generate_dataset(seed: int = 1, frames=(3, 12), target_iterations: int = 200, small: int = 32,
                     large: int = 1, log=None):
    """Pairs from the plain solver: returns (features DeviceArray rows x 32
    bf16, targets DeviceArray rows x 3 f32, rows).  Rows are shuffled."""
    mesh = A.ClothMeshData.batch(training_descs(seed, small, large))
    sim = make_sim(mesh)
    n = sim.n
    rows = n * len(frames)
    feat = A.DeviceArray((rows, A.MLP_IN), np.uint16)
    tgt = A.DeviceArray((rows, 3), np.float32)
    done = 0
    for k, f in enumerate(frames):
        sim.step_frames(f - done)
        done = f
        fk = A.DeviceArray((n, A.MLP_IN), np.uint16)
        tk = A.DeviceArray((n, 3), np.float32)
        sim.warm_sample(target_iterations, fk, tk)
        feat.upload(fk.numpy(), k * n * A.MLP_IN * 2)
        tgt.upload(tk.numpy(), k * n * 3 * 4)
        if log:
            log(f"warmstart: sampled frame {f}: {n} pairs")
    sim.close()
    # shuffle the rows once (the trainer walks consecutive minibatches)
    perm = np.random.default_rng(seed).permutation(rows)
    fh, th = feat.numpy()[perm], tgt.numpy()[perm]
    feat.upload(fh)
    tgt.upload(th)
    return feat, tgt, rows, th


def train(feat, tgt, rows: int, targets_host=None, seed: int = 1, steps: int = 3000, batch: int = 4096,
          lr: float = 2e-3, log=None):
    """A trained WarmStartMLP; the output scale is 1 / std of the targets."""
    th = targets_host if targets_host is not None else tgt.numpy()
    std = float(np.std(th))
    scale = 1.0 / std if std > 0 else 1.0
    mlp = A.WarmStartMLP(seed=seed)
    p, _ = mlp.params()
    mlp.set_params(p, out_scale=scale)
    t0 = time.perf_counter()
    curve = mlp.train(feat, tgt, rows, steps=steps, batch=batch, lr=lr)
    if log:
        log(f"warmstart: trained {steps} steps x {batch} rows in {time.perf_counter() - t0:.2f}s, "
            f"loss {curve[0]:.4f} -> {float(np.mean(curve[-50:])):.4f} (scaled units; target std {std:.3e})")
    return mlp, curve


def substep_residuals(mesh, state, iterations_list, mlp=None, warm_iterations=(2,), threshold=float("inf")):
    """Relative residual after one substep from `state` (pos, vel) for the
    plain solver at each iteration count and, with `mlp`, for warm start + k
    iterations.  Returns {"plain": {I: r}, "warm": {k: r}}."""
    pos, vel = state
    out = {"plain": {}, "warm": {}}
    for it in iterations_list:
        sim = make_sim(mesh, iterations=it)
        sim.write_state(pos, vel)
        sim.step_substeps(1, False)
        out["plain"][it] = sim.residual()
        sim.close()
    if mlp is not None:
        for k in warm_iterations:
            sim = make_sim(mesh, iterations=10)
            sim.set_warm_start(mlp, warm_iterations=k, residual_threshold=threshold)
            sim.write_state(pos, vel)
            sim.step_substeps(1, False)
            out["warm"][k] = sim.residual()
            sim.close()
    return out


def stiff_eval_descs(seed: int = 77, count: int = 16):
    """Held-out sheets of the first training shape: 32x32 hanging, inextensible
    stretch (compliance 0), shear 1e-6, bend 1e-4, two corners or top row."""
    rng = SplitMix64(seed)
    descs = []
    for _ in range(count):
        d = A.SheetDesc.make(cols=32, rows=32, spacing=1.0 / 31.0, layout=A.LAYOUT_HANGING_XY, height=0.0,
                             pins=A.PIN_TWO_CORNERS if rng.below(2) == 0 else A.PIN_TOP_ROW)
        d.jitter, d.seed = 1e-3, rng.next()
        d.stretch_compliance, d.shear_compliance, d.bend_compliance = 0.0, 1e-6, 1e-4
        descs.append(d)
    return descs


def equivalent_iterations(plain: dict, r: float) -> float:
    """Plain iteration count reaching residual r, by log-linear interpolation
    of the measured (iterations, residual) points (clamped to their range)."""
    pts = sorted(plain.items())
    if r >= pts[0][1]:
        return float(pts[0][0])
    for (i0, r0), (i1, r1) in zip(pts, pts[1:]):
        if r1 <= r <= r0:
            if r0 == r1:
                return float(i0)
            t = (math.log(r0) - math.log(r)) / (math.log(r0) - math.log(r1))
            return float(i0 + t * (i1 - i0))
    return float(pts[-1][0])


def substep_time_ms(mesh, iterations: int, mlp=None, warm_iterations: int = 2, threshold: float = float("inf"),
                    frames: int = 10) -> float:
    """Mean wall time of one substep (hipGraph replay of whole frames)."""
    sim = make_sim(mesh, iterations=iterations, use_graph=1)
    if mlp is not None:
        sim.set_warm_start(mlp, warm_iterations=warm_iterations, residual_threshold=threshold)
    sim.step_frames(2)
    sim.sync()
    t0 = time.perf_counter()
    sim.step_frames(frames)
    sim.sync()
    el = time.perf_counter() - t0
    sim.close()
    return 1e3 * el / (frames * 8)


def evaluate(mlp, descs=None, start_frames: int = 2, plain_iterations=(2, 4, 10, 20, 50, 100),
             warm_iterations: int = 2, threshold: float = 1e-4):
    """The gain of the warm start on held-out cloths.

    Single substep from a plain-solver state: residual of plain at each
    iteration count and of warm start + `warm_iterations`; the plain
    iteration count of equal residual; the time per substep of both at that
    equality.  Multi-step: over 10 frames with the residual gate at
    `threshold`, how many substeps fell back, and the resulting rate."""
    mesh = A.ClothMeshData.batch(descs if descs is not None else stiff_eval_descs())
    sim = make_sim(mesh)
    sim.step_frames(start_frames)
    state = sim.read_state()
    sim.close()
    r = substep_residuals(mesh, state, plain_iterations, mlp, warm_iterations=(warm_iterations,))
    rw = r["warm"][warm_iterations]
    eq = equivalent_iterations(r["plain"], rw)
    i_eq = max(1, int(round(eq)))
    t_plain = substep_time_ms(mesh, i_eq)
    t_warm = substep_time_ms(mesh, 10, mlp, warm_iterations, float("inf"))
    gated = make_sim(mesh, use_graph=1)
    gated.set_warm_start(mlp, warm_iterations=warm_iterations, residual_threshold=threshold)
    gated.step_frames(2)
    gated.sync()
    gated.warm_stats(reset=True)
    t0 = time.perf_counter()
    gated.step_frames(10)
    gated.sync()
    tg = time.perf_counter() - t0
    sub, fb = gated.warm_stats()
    gated.close()
    return {
        "particles": mesh.num_particles,
        "residual_plain": {str(k): v for k, v in r["plain"].items()},
        f"residual_warm_plus_{warm_iterations}": rw,
        "equivalent_plain_iterations": round(eq, 1),
        "substep_ms_plain_at_equal_residual": round(t_plain, 4),
        "substep_ms_warm": round(t_warm, 4),
        "single_substep_speedup_at_equal_residual": round(t_plain / t_warm, 2),
        "gated_threshold": threshold,
        "gated_substeps": sub,
        "gated_fallbacks": fb,
        "gated_substeps_per_s": round(80 / tg, 1),
    }


def evaluate_batch(mlp, count: int = 512, iterations: int = 10, warm_iterations=(3, 4), frames: int = 20,
                   start_frames: int = 2):
    """The warm start at scale behind its gate: `count` held-out stiff sheets
    in one batch.  The plain solver at `iterations` sets the gate: the
    largest relative residual it leaves over 10 frames.  Then, for each warm
    iteration count, `frames` frames with the gate at that threshold: the
    substeps that fell back to the plain substep, the substep rate against
    the plain one and the residual at the end."""
    mesh = A.ClothMeshData.batch(stiff_eval_descs(count=count))
    sim = make_sim(mesh, iterations=iterations, use_graph=1)
    sim.step_frames(start_frames)
    res = []
    for _ in range(10):
        sim.step_frames(1)
        res.append(sim.residual())
    sim.sync()
    t0 = time.perf_counter()
    sim.step_frames(frames)
    sim.sync()
    plain_rate = frames * 8 / (time.perf_counter() - t0)
    sim.close()
    threshold = max(res)
    out = {"sheets": count, "particles": mesh.num_particles, "plain_iterations": iterations,
           "plain_residual_range": [min(res), threshold], "plain_substeps_per_s": round(plain_rate, 1),
           "gate_threshold": threshold, "warm": {}}
    for k in warm_iterations:
        g = make_sim(mesh, use_graph=1)
        g.set_warm_start(mlp, warm_iterations=k, residual_threshold=threshold)
        g.step_frames(start_frames)
        g.sync()
        g.warm_stats(reset=True)
        t0 = time.perf_counter()
        g.step_frames(frames)
        g.sync()
        rate = frames * 8 / (time.perf_counter() - t0)
        sub, fb = g.warm_stats()
        out["warm"][str(k)] = {"fallbacks": fb, "substeps": sub, "substeps_per_s": round(rate, 1),
                               "speedup_vs_plain": round(rate / plain_rate, 3), "final_residual": g.residual()}
        g.close()
    return out


def _rollout(mesh, state, frames, iterations=10, mlp=None, warm_iterations=2, threshold=float("inf")):
    """`frames` frames (hipGraph) from `state`: substeps/s (frames timed one by
    one, residual read outside the timing), the relative residual after each
    frame, and the warm start's (substeps, fallbacks)."""
    sim = make_sim(mesh, iterations=iterations, use_graph=1)
    if mlp is not None:
        sim.set_warm_start(mlp, warm_iterations=warm_iterations, residual_threshold=threshold)
    sim.write_state(*state)
    sim.step_frames(1)   # capture the graph, then start again from the state
    sim.sync()
    sim.write_state(*state)
    if mlp is not None:
        sim.warm_stats(reset=True)
    res, t = [], 0.0
    for _ in range(frames):
        sim.sync()
        t0 = time.perf_counter()
        sim.step_frames(1)
        sim.sync()
        t += time.perf_counter() - t0
        res.append(sim.residual())
    stats = sim.warm_stats() if mlp is not None else (0, 0)
    sim.close()
    return {"substeps_per_s": round(frames * 8 / t, 1), "final_residual": res[-1], "max_residual": max(res),
            "residuals": res, "warm_substeps": stats[0], "fallbacks": stats[1]}


def residual_curve(mlp, count: int = 8, seed: int = 101, frames: int = 10, start_frames: int = 2,
                   plain_iterations=(2, 4, 6, 8, 10, 15, 20, 30, 40), warm_iterations=(2, 3, 4), log=None):
    """Time to residual on the learned warm start's stated workload: `count`
    256x256 sheets in one batch (eval_descs: hanging and horizontal, log-uniform
    compliances).  From one plain-solver state, `frames` frames of plain Jacobi
    at each iteration count give the curve (iterations -> residual after the
    last frame, substeps/s); warm + k (ungated) ends at some residual r_k, and
    the plain iteration count I* ending at r_k (log-linear interpolation of the
    curve) gives the speedup at equal residual, rate(warm + k) / rate(plain
    at I*).  Warm + k gated at the largest per-frame residual of plain at the
    iteration count just above I* shows what the gate keeps."""
    mesh = A.ClothMeshData.batch(eval_descs(seed=seed, count=count))
    sim = make_sim(mesh, iterations=10)
    sim.step_frames(start_frames)
    state = sim.read_state()
    sim.close()
    plain = {}
    for it in plain_iterations:
        plain[it] = _rollout(mesh, state, frames, iterations=it)
        if log:
            log(f"warmstart: 256x256 x {count}: plain I={it}: {plain[it]['substeps_per_s']} substeps/s, "
                f"residual {plain[it]['final_residual']:.4e}")
    pts = sorted((it, r["final_residual"]) for it, r in plain.items())

    def rate_at(i_star):
        lo = max([p for p in pts if p[0] <= i_star], default=pts[0])
        hi = min([p for p in pts if p[0] >= i_star], default=pts[-1])
        if lo[0] == hi[0]:
            return plain[lo[0]]["substeps_per_s"]
        # substep time is affine in the iteration count: interpolate it
        t_lo, t_hi = 1.0 / plain[lo[0]]["substeps_per_s"], 1.0 / plain[hi[0]]["substeps_per_s"]
        return 1.0 / (t_lo + (i_star - lo[0]) / (hi[0] - lo[0]) * (t_hi - t_lo))

    warm = {}
    for k in warm_iterations:
        w = _rollout(mesh, state, frames, mlp=mlp, warm_iterations=k)
        i_star = equivalent_iterations(dict(pts), w["final_residual"])
        above = min([it for it, _ in pts if it >= i_star], default=pts[-1][0])
        gate = plain[above]["max_residual"]
        g = _rollout(mesh, state, frames, mlp=mlp, warm_iterations=k, threshold=gate)
        rec = {"ungated": {k2: v for k2, v in w.items() if k2 != "residuals"},
               "equivalent_plain_iterations": round(i_star, 1),
               "plain_substeps_per_s_at_equal_residual": round(rate_at(i_star), 1),
               "speedup_at_equal_residual": round(w["substeps_per_s"] / rate_at(i_star), 3),
               "gated": {"threshold": gate, "plain_iterations_of_threshold": above,
                         **{k2: v for k2, v in g.items() if k2 != "residuals"},
                         "speedup_vs_that_plain": round(g["substeps_per_s"] / plain[above]["substeps_per_s"], 3)}}
        warm[str(k)] = rec
        if log:
            log(f"warmstart: 256x256 x {count}: warm+{k}: residual {w['final_residual']:.4e} = plain I~{i_star:.1f}, "
                f"{rec['speedup_at_equal_residual']}x at equal residual; gated at plain I={above}'s max "
                f"({gate:.3e}): {g['fallbacks']}/{g['warm_substeps']} fallbacks, "
                f"{rec['gated']['speedup_vs_that_plain']}x")
    return {"workload": f"{count} x 256x256 sheets in one batch (eval_descs seed {seed}), {frames} frames of "
                        f"8 substeps from a shared plain-solver state", "particles": mesh.num_particles,
            "plain": {str(it): {k2: v for k2, v in r.items() if k2 not in ("residuals", "warm_substeps", "fallbacks")}
                      for it, r in plain.items()},
            "warm": warm}


def validation_fallbacks(mlp, count: int = 64, seed: int = 91, iterations: int = 10, warm_iterations: int = 2,
                         frames: int = 6, start_frames: int = 2):
    """Fallbacks of the gated warm start over a short rollout of `count`
    validation sheets (the evaluation family, another seed), the gate set at
    the plain solver's largest residual there.  Returns (fallbacks, substeps)."""
    mesh = A.ClothMeshData.batch(stiff_eval_descs(seed=seed, count=count))
    sim = make_sim(mesh, iterations=iterations, use_graph=1)
    sim.step_frames(start_frames)
    res = []
    for _ in range(frames):
        sim.step_frames(1)
        res.append(sim.residual())
    sim.close()
    g = make_sim(mesh, use_graph=1)
    g.set_warm_start(mlp, warm_iterations=warm_iterations, residual_threshold=max(res))
    g.step_frames(start_frames)
    g.warm_stats(reset=True)
    g.step_frames(frames)
    fb = g.warm_stats()
    g.close()
    return fb[1], fb[0]


def train_selected(target_iterations=(20, 30, 40), steps: int = 10000, batch: int = 4096, lr: float = 2e-3,
                   seed: int = 1, log=None):
    """One network per target-iteration count, kept by rollout validation: the
    one whose gated warm + 2 falls back least on the validation sheets (ties:
    the lower final training loss).  The rollout, not the regression loss,
    decides: networks of equal loss differ widely in how their small errors
    feed back through the velocity update.  Returns (mlp, report)."""
    best, best_key, report = None, None, []
    for ti in target_iterations:
        feat, tgt, rows, th = generate_dataset(seed=seed, frames=(3, 12), target_iterations=ti)
        mlp, curve = train(feat, tgt, rows, th, seed=seed, steps=steps, batch=batch, lr=lr)
        feat.free()
        tgt.free()
        fb, sub = validation_fallbacks(mlp)
        loss = float(np.mean(curve[-50:]))
        report.append({"target_iterations": ti, "train_loss": round(loss, 4), "validation_fallbacks": fb,
                       "validation_substeps": sub})
        if log:
            log(f"warmstart: targets I={ti}: loss {loss:.4f}, validation fallbacks {fb}/{sub}")
        key = (fb, loss)
        if best is None or key < best_key:
            if best is not None:
                best.close()
            best, best_key = mlp, key
        else:
            mlp.close()
    return best, report


def main(argv=None):
    """Train a rollout-selected network and evaluate it behind the gate at
    scale; prints one JSON line (warmstart.evaluate_batch's record)."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description="learned warm start: train, select, evaluate at scale")
    ap.add_argument("--sheets", type=int, default=512, help="held-out stiff 32x32 sheets in the batch")
    ap.add_argument("--frames", type=int, default=20, help="timed frames per warm iteration count")
    ap.add_argument("--quick", action="store_true", help="one candidate (40-iteration targets)")
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)   # noqa: E731
    mlp, selection = train_selected(target_iterations=(40,) if args.quick else (20, 30, 40), log=log)
    r = evaluate_batch(mlp, count=args.sheets, warm_iterations=(2, 3, 4), frames=args.frames)
    r["selection"] = selection
    mlp.close()
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
