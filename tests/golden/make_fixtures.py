"""Regenerates tests/golden/oracle_config1.json (and, with --sheet1024,
tests/golden/oracle_sheet1024.json).

The reference ships no stored vectors or trajectories, so these fixtures are
not reference output: they freeze this repository's f32 and f64 oracles on
the 32x32 hanging sheet (spacing 1/31 m, two top corners pinned, +-1e-3 m
jitter from seed 1, Jacobi, 8 substeps x 10 iterations, no collision), so
that any change of the oracle's arithmetic shows up as a changed digest.

oracle_sheet1024.json freezes the f32 oracle on the 1024x1024 self-colliding
sheet (spacing 1/1023 m, two top corners pinned, +-1e-3 m jitter from seed 3,
contact distance 2 x 0.4 x spacing, Jacobi, 8 substeps x 10 iterations,
collision every substep) at frames 5, 25 and 65, the frames that bracket a
60-frame timed run after 5 warmup frames; it takes a few minutes on 8 cores.

Run from the repository root:  python tests/golden/make_fixtures.py [--sheet1024]
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

FRAMES = (1, 10, 120)
SHEET = dict(cols=32, rows=32, spacing=1.0 / 31.0, pins=1, jitter=1e-3, seed=1)
SAMPLE = (0, 31, 500, 1023)


def trajectory(O):
    m = O.sheet(**SHEET)
    p = O.params(substeps=8, inner_iterations=10, omega=0.32, collision=False)
    s = O.Solver(m)
    x32, v32 = m["positions"].copy(), np.zeros_like(m["positions"])
    x64 = m["positions"].astype(np.float64)
    v64 = np.zeros_like(x64)
    snaps = {}
    for f in range(1, max(FRAMES) + 1):
        x32, v32 = s.frames(x32, v32, p, 8, 1, 1)
        x64, v64 = s.frames(x64, v64, p, 8, 1, 1, dtype=np.float64)
        if f in FRAMES:
            snaps[f] = (x32.copy(), x64.copy())
    return snaps


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def fixtures(O):
    out = {"sheet": SHEET, "substeps": 8, "iterations": 10, "frames": {}}
    for f, (x32, x64) in trajectory(O).items():
        out["frames"][str(f)] = {
            "f32_sha256": digest(x32),
            "f32_sample": {str(i): [float(v) for v in x32[i]] for i in SAMPLE},
            "max_abs_f32_vs_f64": float(np.abs(x32.astype(np.float64) - x64).max()),
        }
    return out


SHEET1024 = dict(cols=1024, rows=1024, spacing=1.0 / 1023.0, pins=2, jitter=1e-3, seed=3)
SHEET1024_FRAMES = (5, 25, 65)
SHEET1024_SAMPLE = (0, 1023, 524800, 1048575)


def sheet1024_fixtures(O, threads, frames=SHEET1024_FRAMES):
    h = SHEET1024["spacing"]
    m = O.sheet(**SHEET1024)
    p = O.params(substeps=8, inner_iterations=10, thickness=2.0 * (0.4 * h), collision=True,
                 coll_mode=O.COLL_GRID)
    s = O.Solver(m)
    O.set_threads(threads)
    x, v = m["positions"].copy(), np.zeros_like(m["positions"])
    out = {"sheet": SHEET1024, "substeps": 8, "iterations": 10, "thickness": float(np.float32(p.thickness)),
           "frames": {}}
    done = 0
    for f in frames:
        x, v = s.frames(x, v, p, 8, 1, f - done)
        done = f
        out["frames"][str(f)] = {
            "f32_sha256": digest(x),
            "f32_vel_sha256": digest(v),
            "f32_sample": {str(i): [float(c) for c in x[i]] for i in SHEET1024_SAMPLE},
            "residual": s.residual(x),
        }
        print(f"frame {f} done", flush=True)
    return out


if __name__ == "__main__":
    from oracle import oracle as O
    if "--sheet1024" in sys.argv:
        path = os.path.join(ROOT, "tests", "golden", "oracle_sheet1024.json")
        data = sheet1024_fixtures(O, os.cpu_count() or 1)
    else:
        path = os.path.join(ROOT, "tests", "golden", "oracle_config1.json")
        data = fixtures(O)
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", path)
