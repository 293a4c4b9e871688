"""Benchmark: cloth substeps/s of the XPBD hot path on one MI355X.

Workload (default): the 1024x1024 self-colliding sheet — 1,048,576 particles,
6,281,218 distance constraints (stretch/shear/bend), top row pinned, seeded
+-1e-3 m jitter, particle radius 0.4 x spacing (contact distance = cell =
2r), Jacobi, 8 substeps x 10 iterations per frame.  A "step" is one frame
(8 substeps) replayed from the captured hipGraph; inputs are resident in HBM.

`--workload batch512`: 512 cloths of uneven sizes (sides U{16..192} per axis,
log-uniform compliances, pins from {two corners, top row, none + ground}),
self-collision within each cloth, same solver settings.  `--persistent` runs
predict + all iterations of a substep as one persistent launch.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 exits: the library runs
on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CALIBRATION_BYTES = 1 << 30
PROFILE_FRAMES = 10


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def sheet_1024():
    return dict(cols=1024, rows=1024, spacing=1.0 / 1023.0, pins=2, jitter=1e-3, seed=3)


def algorithmic_bytes_per_iteration(n: int, e: int) -> int:
    """Per Jacobi iteration: read x (12 B), write x (12 B), inverse mass (4 B),
    CSR offsets (4 B) per particle; 12 B (neighbour id, rest, compliance) per
    directed constraint entry (2 per constraint)."""
    return 32 * n + 24 * e


def algorithmic_bytes_per_substep(n: int, e: int, iters: int, collision: bool, table_slots: int) -> int:
    b = 80 * n + iters * algorithmic_bytes_per_iteration(n, e)
    if collision:
        b += 152 * n + 8 * table_slots
    return b


class TraceRange:
    """A named ROCTX range around a region (rocprofv3 --marker-trace records
    it; outside a profiler the calls are no-ops).  bench.py brackets its
    timed frames with one, so --trace-summary can keep exactly their kernels."""

    def __init__(self, name: str):
        self.name = name.encode()
        self.lib = None
        try:
            import ctypes
            self.lib = ctypes.CDLL("librocprofiler-sdk-roctx.so")
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        except OSError:
            self.lib = None

    def __enter__(self):
        if self.lib is not None:
            self.lib.roctxRangePushA(self.name)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxRangePop()


def trace_summary(trace_dir: str, range_name: str = "timed frames", frames: int | None = None,
                  by_grid: bool = False) -> str:
    """Per-kernel durations from a rocprofv3 CSV trace (--kernel-trace -f csv,
    with --marker-trace for a bench.py run), restricted to the dispatches that
    ran inside the ROCTX range `range_name` when the trace holds it (else all
    dispatches); with by_grid, one row per (kernel, grid size): a markdown
    table."""
    import csv
    import glob
    kt = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    mt = glob.glob(os.path.join(trace_dir, "**", "*marker_api_trace.csv"), recursive=True)
    if not kt:
        raise SystemExit(f"bench.py: no kernel trace CSV under {trace_dir}")
    lo = hi = None
    if mt:
        with open(mt[0]) as fh:
            for row in csv.DictReader(fh):
                text = " ".join(str(v) for v in row.values())
                if range_name in text:
                    lo, hi = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
    stats = {}
    with open(kt[0]) as fh:
        for row in csv.DictReader(fh):
            a, b = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
            if lo is None or (a >= lo and b <= hi):
                k = row["Kernel_Name"]
                if by_grid:
                    k = (k, int(row.get("Grid_Size") or row.get("Grid_Size_X") or 0))
                c, t = stats.get(k, (0, 0))
                stats[k] = (c + 1, t + (b - a))
    total = sum(t for _, t in stats.values())
    head = (f"Window: ROCTX range '{range_name}', {(hi - lo) / 1e3:.1f} us wall" if lo is not None
            else "Window: the whole trace") + (f", {frames} frames" if frames and lo is not None else "")
    lines = [head + f"; kernel time {total / 1e3:.1f} us.", ""]
    per_frame = bool(frames) and lo is not None
    cols = "| kernel |" + (" grid threads |" if by_grid else "") + " calls | total us | average us | % |"
    lines += [cols + (" per frame us |" if per_frame else ""),
              "|---|" + ("---:|" if by_grid else "") + "---:|---:|---:|---:|" + ("---:|" if per_frame else "")]
    for k, (c, t) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        name, grid = (k if by_grid else (k, None))
        name = name if len(name) <= 90 else name[:90]
        row = f"| `{name}` |" + (f" {grid} |" if by_grid else "")
        row += f" {c} | {t / 1e3:.1f} | {t / c / 1e3:.2f} | {100.0 * t / total:.2f} |"
        if per_frame:
            row += f" {t / frames / 1e3:.1f} |"
        lines.append(row)
    return "\n".join(lines)


def counter_summary(fetch_dir: str, write_dir: str, kernels=("k_jacobi_tiled<0u", "k_jacobi<0u")):
    """Bytes per dispatch from two rocprofv3 counter passes (--pmc FETCH_SIZE,
    then --pmc WRITE_SIZE, -f csv) of `bench.py --kernels-only`.  Both counters
    are KiB per dispatch; each is scaled by what it reads for the 1 GiB
    device-to-device calibration copy (the largest runtime copy dispatch),
    which reads and writes CALIBRATION_BYTES.  Returns (markdown table, the
    JSON record for the plain Jacobi iteration: the first kernel whose name
    contains one of `kernels`)."""
    import csv
    import glob

    def load(d, counter):
        f = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not f:
            raise SystemExit(f"bench.py: no counter CSV under {d}")
        per = {}
        with open(f[0]) as fh:
            for row in csv.DictReader(fh):
                if row["Counter_Name"] == counter:
                    per.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return per

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise SystemExit("bench.py: no calibration copy in the counter CSV")
    kib = CALIBRATION_BYTES / 1024.0
    wmax = max(max(write.get(k, [0.0])) for k in copy)
    if wmax < 0.5 * kib:   # only the small state uploads: the 1 GiB copy did not run
        raise SystemExit(f"bench.py: no {CALIBRATION_BYTES}-byte calibration copy in the counter CSV")
    fs = kib / max(max(fetch[k]) for k in copy)
    ws = kib / wmax
    rows = []
    for k in fetch:
        if k not in write or k.startswith("__amd_rocclr"):
            continue
        fk, wk = sum(fetch[k]) / len(fetch[k]), sum(write[k]) / len(write[k])
        rows.append((k, len(fetch[k]), fk, wk, fk * fs * 1024, wk * ws * 1024))
    rows.sort(key=lambda r: -(r[4] + r[5]))
    lines = [f"Calibration: one device-to-device copy of {CALIBRATION_BYTES} B (larger than the 256 MB "
             "Infinity Cache), read and written once.", "",
             f"- FETCH_SIZE reads {kib / fs:.1f} KiB for it -> scale {fs:.5f}",
             f"- WRITE_SIZE reads {kib / ws:.1f} KiB for it -> scale {ws:.5f}", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled read MB | scaled write MB "
             "| traffic MB |", "|---|---:|---:|---:|---:|---:|---:|"]
    for k, c, fk, wk, rb, wb in rows:
        name = k if len(k) <= 60 else k[:60]
        lines.append(f"| `{name}` | {c} | {fk:.1f} | {wk:.1f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} "
                     f"| {(rb + wb) / 1e6:.1f} |")
    rec = None
    for k, c, fk, wk, rb, wb in rows:
        if any(x in k for x in kernels):
            sweep = [r for r in rows if "k_coll_sweep(" in r[0]]
            rec = {"kernel": k, "fetch_kib_raw": fk, "write_kib_raw": wk, "fetch_scale": fs,
                   "write_scale": ws, "calibration_bytes": CALIBRATION_BYTES,
                   "bytes_per_launch": int(round(rb + wb)), "read_bytes": int(round(rb)),
                   "write_bytes": int(round(wb))}
            if sweep:
                rec["collision_sweep_bytes_per_launch"] = int(round(sweep[0][4] + sweep[0][5]))
            break
    return "\n".join(lines), rec


def cpu_baseline(mesh_fields, u, cfg, budget_s: float, threads: int):
    """Oracle (C++ restatement of the reference CPU solver) on this host, a
    bounded number of substeps of the same workload."""
    import numpy as np
    from oracle import oracle as O
    O.set_threads(threads)
    p = O.params(substeps=cfg.solve_substeps, inner_iterations=cfg.solve_inner_iterations,
                 omega=float(u.jacobi_omega), thickness=float(u.thickness), coll_scale=float(u.coll_scale),
                 collision=True, coll_mode=O.COLL_GRID)
    s = O.Solver(mesh_fields)
    pos = mesh_fields["positions"].copy()
    vel = np.zeros_like(pos)
    done = 0
    t0 = time.perf_counter()
    while True:
        pos, vel = s.substep(pos, vel, p)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 64:
            break
    return done / el, done, el


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def learned_warm_start():
    """Train the warm-start network on the GPU (solver-generated pairs of the
    32x32 and 256x256 training sheets) and measure what it gains on held-out
    32x32 stiff sheets: one substep of warm start + 2 iterations against
    plain Jacobi at several iteration counts (equal-residual speedup), and
    10 frames behind the residual gate (threshold 1e-4).  Then at scale: a
    network chosen by a gated rollout on validation sheets
    (warmstart.train_selected), on 512 held-out stiff sheets in one batch,
    gated at the plain 10-iteration residual (warmstart.evaluate_batch).
    Last, on the stated workload (8 x 256x256 sheets in one batch): the time to
    residual against plain Jacobi at 2..40 iterations (warmstart.residual_curve)."""
    try:
        from bevy_xpbd_amd import warmstart as W
        t0 = time.perf_counter()
        feat, tgt, rows, th = W.generate_dataset(seed=1, frames=(3, 12), target_iterations=200)
        mlp, curve = W.train(feat, tgt, rows, th, steps=3000, batch=4096, lr=2e-3)
        res = W.evaluate(mlp, threshold=1e-4)
        res.update({"training_pairs": rows, "train_steps": 3000, "train_batch": 4096,
                    "train_loss_first_last": [round(float(curve[0]), 4), round(float(curve[-50:].mean()), 4)],
                    "seconds": round(time.perf_counter() - t0, 1)})
        mlp.close()
        log(f"bench: warm start: warm+2 ~ plain I={res['equivalent_plain_iterations']}, single-substep speedup "
            f"{res['single_substep_speedup_at_equal_residual']}x; gated fallbacks "
            f"{res['gated_fallbacks']}/{res['gated_substeps']}")
        # at scale: networks on 20/30/40-iteration targets, one kept by a gated
        # rollout on validation sheets, then gated at the plain I=10 residual
        # on the held-out 512 sheets
        mlp, selection = W.train_selected(log=log)
        batch = W.evaluate_batch(mlp, count=512, warm_iterations=(2, 3, 4))
        batch.update({"train_steps": 10000, "selection": selection})
        # the stated workload: 256x256 sheets in the batch layout, time to
        # residual against plain Jacobi at 2..40 iterations
        curve = W.residual_curve(mlp, log=log)
        mlp.close()
        res["batch_512_gated"] = batch
        res["sheets256_time_to_residual"] = curve
        best = max(curve["warm"].items(), key=lambda kv: kv[1]["speedup_at_equal_residual"])
        res["headline"] = {"workload": curve["workload"], "warm_iterations": int(best[0]),
                           "speedup_at_equal_residual": best[1]["speedup_at_equal_residual"],
                           "equivalent_plain_iterations": best[1]["equivalent_plain_iterations"],
                           "gated_speedup_vs_plain_at_gate": best[1]["gated"]["speedup_vs_that_plain"],
                           "gated_fallbacks": [best[1]["gated"]["fallbacks"], best[1]["gated"]["warm_substeps"]],
                           "gate_threshold": best[1]["gated"]["threshold"],
                           "gate_plain_iterations": best[1]["gated"]["plain_iterations_of_threshold"]}
        for k, v in batch["warm"].items():
            log(f"bench: warm start, 512 sheets: warm+{k} gated at the plain I=10 residual: {v['fallbacks']}/"
                f"{v['substeps']} fallbacks, {v['speedup_vs_plain']}x plain I=10")
        res["seconds"] = round(time.perf_counter() - t0, 1)
        return res
    except Exception as exc:   # reported, never fatal to the throughput line
        log(f"bench: warm start evaluation failed: {exc!r}")
        return {"error": repr(exc)}


def config2_rates(X, frames: int = 40, warmup: int = 5):
    """The 256x256 sheet of the second configuration (horizontal at 1 m,
    centre pinned, seeded rotation, self-collision every substep, contact
    distance 0.8 x spacing), 4 substeps x 8 iterations, both solvers (Jacobi
    omega 0.32, coloured Gauss-Seidel omega 1.0; GS in its autotuned launch
    form and in each form: one persistent launch per substep with grid
    barriers, one launch per colour, one persistent launch ordered by per-tile
    flags), frames replayed from the hipGraph: substeps/s, the GS substep's
    time and its colours."""
    import numpy as np
    h = 1.0 / 255.0
    angle = float(np.random.default_rng(2).uniform(0, 2 * np.pi))
    out = {"workload": "256x256 horizontal sheet, centre pin, seeded rotation, self-collision every substep, "
                       "4 substeps x 8 iterations (step = 1 frame)", "angle_rad": round(angle, 6)}
    for name, solver, form in (("jacobi", X.SOLVER_JACOBI, 0), ("gauss_seidel", X.SOLVER_GAUSS_SEIDEL, 0),
                               ("gauss_seidel_persistent", X.SOLVER_GAUSS_SEIDEL, 1),
                               ("gauss_seidel_per_colour", X.SOLVER_GAUSS_SEIDEL, 2),
                               ("gauss_seidel_tile_flags", X.SOLVER_GAUSS_SEIDEL, 3)):
        mesh = X.ClothMeshData.sheet(cols=256, rows=256, spacing=h, layout=X.LAYOUT_HORIZONTAL_XZ, height=1.0,
                                     angle=angle, pins=X.PIN_CENTER)
        cfg = X.ClothSimConfig.default(solve_substeps=4, solve_inner_iterations=8, solver=solver, use_graph=1,
                                       gs_launch=form)
        u = X.ClothSimUniforms.default(solver, 4, thickness=2.0 * (0.4 * h))
        sim = X.ClothComputePlugin(mesh, cfg, u)
        sim.step_frames(warmup)
        sim.sync()
        t0 = time.perf_counter()
        sim.step_frames(frames)
        sim.sync()
        rate = 4 * frames / (time.perf_counter() - t0)
        rec = {"substeps_per_s": round(rate, 1)}
        if solver == X.SOLVER_GAUSS_SEIDEL:
            ms, cnt = sim.profile_frames(5, X.KERNEL_GS_SUBSTEP)
            rec.update({"gs_substep_launch_us": round(ms * 1e3, 2), "colours": len(mesh.field("batch_offsets")) - 1})
            if form in (0, 3):
                rec["tile_plan"] = sim.gs_tiles()
        sim.close()
        out[name] = rec
        log(f"bench: 256x256 sheet, {name}: {rate:.1f} substeps/s")
    return out


def gs_forms_1024(X, frames: int = 5, warmup: int = 2, only=None):
    """Coloured Gauss-Seidel on the 1024x1024 bench sheet (8 substeps x 10
    iterations, omega 1.0, collision every substep), frames replayed from the
    hipGraph, in its autotuned launch form and in each form: substeps/s
    (and the tile plan: particles per tile, tiles, blocks, most dependencies)."""
    out = {"workload": "1024x1024 bench sheet, coloured Gauss-Seidel 8 substeps x 10 iterations, "
                       "self-collision every substep"}
    for name, form in (("autotuned", 0), ("persistent", 1), ("per_colour", 2), ("tile_flags", 3)):
        if only and name not in only:
            continue
        mesh, thickness, _, _ = build_workload(X, "sheet1024")
        cfg = X.ClothSimConfig.default(solve_substeps=8, solve_inner_iterations=10, solver=X.SOLVER_GAUSS_SEIDEL,
                                       use_graph=1, gs_launch=form)
        u = X.ClothSimUniforms.default(X.SOLVER_GAUSS_SEIDEL, 8, thickness=float(thickness))
        sim = X.ClothComputePlugin(mesh, cfg, u)
        sim.step_frames(warmup)
        sim.sync()
        with TraceRange(f"gauss-seidel {name} frames"):
            t0 = time.perf_counter()
            sim.step_frames(frames)
            sim.sync()
            rate = 8 * frames / (time.perf_counter() - t0)
        out[name] = {"substeps_per_s": round(rate, 1)}
        if form in (0, 3):
            out[name]["tile_plan"] = sim.gs_tiles()
        sim.close()
        log(f"bench: 1024x1024 sheet, Gauss-Seidel {name}: {rate:.1f} substeps/s")
    return out


def build_workload(X, name: str):
    """Returns (mesh, thickness, description, config extras)."""
    if name == "sheet1024":
        kw = sheet_1024()
        h = kw["spacing"]
        mesh = X.ClothMeshData.sheet(**kw)
        return mesh, 2.0 * (0.4 * h), ("1024x1024 self-colliding sheet, plain Jacobi, 8 substeps x 10 iterations "
                                       "per frame (step = 1 frame)"), {}
    if name == "batch512":
        from bevy_xpbd_amd.workloads import batch_descs
        spacing = 0.01
        descs, modes = batch_descs(count=512, seed=4, spacing=spacing)
        mesh = X.ClothMeshData.batch(descs)
        extra = {"cloths": len(descs), "sides": "U{16..192} per axis", "compliance": "log-uniform [1e-8, 1e-3] per class",
                 "pins": {m: modes.count(m) for m in sorted(set(modes))}, "seed": 4, "spacing_m": spacing}
        return mesh, 2.0 * (0.4 * spacing), ("512-cloth uneven batch, self-collision within each cloth, plain "
                                             "Jacobi, 8 substeps x 10 iterations per frame (step = 1 frame)"), extra
    raise SystemExit(f"bench.py: unknown workload {name!r}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sheet1024", choices=["sheet1024", "batch512"])
    ap.add_argument("--persistent", action="store_true", help="one persistent Jacobi launch per substep")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (capped at 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="skip the 256x256 Jacobi / Gauss-Seidel side record")
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm start's train + evaluation")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the per-kernel timing launches (for rocprofv3 counter passes)")
    ap.add_argument("--trace-summary", metavar="DIR",
                    help="print the per-kernel table of the timed frames from a rocprofv3 CSV trace "
                         "(--kernel-trace --marker-trace -f csv) of an earlier bench.py run, and exit")
    ap.add_argument("--by-grid", action="store_true", help="--trace-summary: one row per kernel and grid size")
    ap.add_argument("--trace-range", default="timed frames",
                    help="--trace-summary: the ROCTX range to keep (default: the timed Jacobi frames; "
                         "'gauss-seidel per_colour frames' etc. for the 1024x1024 Gauss-Seidel records)")
    ap.add_argument("--gs-forms", metavar="NAMES",
                    help="only the 1024x1024 Gauss-Seidel records of these forms (comma-separated: autotuned, "
                         "persistent, per_colour, tile_flags; --steps frames each), printed as JSON, then exit")
    ap.add_argument("--counter-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the bytes-per-dispatch table from FETCH_SIZE and WRITE_SIZE counter passes "
                         "(-f csv) of `bench.py --kernels-only`, then the Jacobi iteration's JSON record, and exit")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary, range_name=args.trace_range, frames=args.steps, by_grid=args.by_grid))
        return
    if args.gs_forms:
        import bevy_xpbd_amd as X
        print(json.dumps(gs_forms_1024(X, frames=args.steps, only=args.gs_forms.split(","))), flush=True)
        return
    if args.counter_summary:
        table, rec = counter_summary(*args.counter_summary)
        print(table)
        print(json.dumps(dict(rec or {}, workload=args.workload), indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    import bevy_xpbd_amd as X

    S, I = 8, 10
    log(f"bench: building {args.workload} topology")
    t0 = time.perf_counter()
    mesh, thickness, workload, extra = build_workload(X, args.workload)
    n, e = mesh.num_particles, mesh.num_constraints
    log(f"bench: topology {n} particles, {e} constraints in {time.perf_counter() - t0:.1f}s")
    cfg = X.ClothSimConfig.default(solve_substeps=S, solve_inner_iterations=I, use_graph=1,
                                   autotune=0 if args.no_autotune else 1,
                                   persistent_jacobi=1 if args.persistent else 0)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, S, thickness=float(thickness))
    sim = X.ClothComputePlugin(mesh, cfg, u)
    if args.kernels_only:
        jac_ms = sim.time_kernel(X.KERNEL_JACOBI_ITER, 20)
        sweep_ms = sim.time_kernel(X.KERNEL_COLLISION_SWEEP, 5)
        copy_ms = sim.time_kernel(X.KERNEL_COPY, 10)
        log(f"bench: kernels only: jacobi {jac_ms * 1e3:.1f} us, sweep {sweep_ms * 1e3:.1f} us, "
            f"copy of {32 * n} B {copy_ms * 1e3:.1f} us")
        sim.close()
        # counter calibration beyond the 256 MB Infinity Cache: one 1 GiB
        # device-to-device copy (the runtime's copy kernel reads and writes
        # CALIBRATION_BYTES each)
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(src), CALIBRATION_BYTES) == 0 and \
                hip.hipMalloc(ctypes.byref(dst), CALIBRATION_BYTES) == 0:
            hip.hipMemcpy(dst, src, CALIBRATION_BYTES, 3)   # hipMemcpyDeviceToDevice
            hip.hipDeviceSynchronize()
            log(f"bench: calibration copy of {CALIBRATION_BYTES} B done")
        hip.hipFree(src)
        hip.hipFree(dst)
        return
    log(f"bench: device state uploaded; warmup {args.warmup} frames")
    sim.step_frames(args.warmup)
    sim.sync()
    log(f"bench: timing {args.steps} frames")
    with TraceRange("timed frames"):
        t0 = time.perf_counter()
        sim.step_frames(args.steps)
        sim.sync()
        el = time.perf_counter() - t0
    substeps = args.steps * S
    value = substeps / el
    ms_per_step = 1000.0 * el / args.steps
    log(f"bench: {value:.1f} substeps/s ({ms_per_step:.3f} ms/frame)")

    # dominant kernel: one Jacobi iteration.  Its mean launch duration comes
    # from HIP events on the handle's stream around every iteration launch of
    # PROFILE_FRAMES further frames of the same run (eager, live state).
    jac_ms, jac_launches = sim.profile_frames(PROFILE_FRAMES, X.KERNEL_JACOBI_ITER)
    sweep_ms, _ = sim.profile_frames(PROFILE_FRAMES, X.KERNEL_COLLISION_SWEEP)
    pers_live_ms = None
    if args.persistent:
        pers_live_ms, pers_launches = sim.profile_frames(PROFILE_FRAMES, X.KERNEL_JACOBI_PERSISTENT)
    jac_rep_ms = sim.time_kernel(X.KERNEL_JACOBI_ITER, 50)
    radix_ms = sim.time_kernel(X.KERNEL_RADIX_PASS, 20)
    copy_ms = sim.time_kernel(X.KERNEL_COPY, 20)
    pers_ms = sim.time_kernel(X.KERNEL_JACOBI_PERSISTENT, 5)
    b_iter = algorithmic_bytes_per_iteration(n, e)
    table_slots = 1 << max(6, (2 * n - 1).bit_length())
    b_sub = algorithmic_bytes_per_substep(n, e, I, True, table_slots)
    substep_gbs = b_sub * value / 1e9
    traffic = None
    tname = {"sheet1024": "jacobi_iter_1024_traffic.json", "batch512": "jacobi_iter_batch512_traffic.json"}
    tpath = os.path.join(ROOT, "profiles", tname[args.workload])
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    residual = sim.residual()
    sim.close()
    if jac_launches:
        achieved = b_iter / (jac_ms * 1e-3) / 1e9
        rkernel, r_us, r_bytes = "k_jacobi (one iteration)", jac_ms * 1e3, b_iter
    else:   # persistent: the dominant launch is the whole predict + I iterations
        achieved = (I * b_iter) / (pers_live_ms * 1e-3) / 1e9
        rkernel, r_us, r_bytes = "k_jacobi_persistent (predict + I iterations)", pers_live_ms * 1e3, I * b_iter
        jac_launches = pers_launches
        traffic = None   # the counter summaries measure the per-iteration kernel, not this launch
    log(f"bench: dominant kernel {rkernel} {r_us:.1f} us = {achieved:.0f} GB/s algorithmic; "
        f"collision sweep {sweep_ms * 1e3:.1f} us; radix pass {radix_ms * 1e3:.1f} us")

    side = gs1024 = None
    if args.workload == "sheet1024" and not args.no_side:
        side = config2_rates(X)
        gs1024 = gs_forms_1024(X)

    warm = None
    if not args.no_warm:
        warm = learned_warm_start()

    cpu = None
    if not args.no_cpu:
        host_cpus = os.cpu_count() or 1
        threads = args.cpu_threads or min(host_cpus, 16)
        fields = mesh.fields()
        log(f"bench: CPU baseline (oracle, {threads} threads, ~{args.cpu_budget:.0f}s; then 1 thread)")
        rate, done, cel = cpu_baseline(fields, u, cfg, args.cpu_budget, threads)
        rate1, done1, cel1 = cpu_baseline(fields, u, cfg, args.cpu_budget, 1)
        cpu = {"value": round(rate, 4), "unit": "substeps/s", "cores": threads, "kind": "port",
               "sample": f"{done} substeps of the same workload ({args.workload}: 8x10 Jacobi + hash-grid "
                         f"self-collision) from the initial state, {cel:.1f}s",
               "host_cpus": host_cpus, "host_cpu_model": cpu_model(),
               "single_thread": {"value": round(rate1, 4), "unit": "substeps/s", "cores": 1,
                                 "sample": f"{done1} substeps, same start, {cel1:.1f}s"}}

    config = {"workload": workload, "particles": n, "constraints": e, "substeps_per_step": S, "iterations": I,
              "collision": "every substep, contact distance 0.8 x spacing",
              "solve_launch": "persistent (one launch per substep)" if args.persistent else
                              "one launch per iteration (hipGraph)",
              "residual_after_timed_run": residual,
              "substep_algorithmic_GBps": round(substep_gbs, 1),
              "substep_hbm_frac": round(substep_gbs / HBM_PEAK_GBS, 4),
              "collision_sweep_us": round(sweep_ms * 1e3, 2), "radix_pass_us": round(radix_ms * 1e3, 2),
              "copy_kernel_GBps": round(32 * n / (copy_ms * 1e-3) / 1e9, 1),
              "persistent_predict_plus_iterations_us": round(pers_ms * 1e3, 2)}
    config.update(extra)
    if side is not None:
        config["sheet256_both_solvers"] = side
    if gs1024 is not None:
        config["sheet1024_gauss_seidel"] = gs1024
    if warm is not None:
        config["learned_warm_start"] = warm
    out = {
        "metric": "cloth substeps/s at fixed constraint residual; learned vs plain Jacobi; HBM %",
        "value": round(value, 2),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "kernel": rkernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_launch_us": round(r_us, 2),
                     "launches_timed": jac_launches, "algorithmic_bytes_per_launch": r_bytes,
                     "back_to_back_iteration_us": round(jac_rep_ms * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
