"""Regenerates tests/golden/oracle_config1.json.

The reference ships no stored vectors or trajectories, so these fixtures are
not reference output: they freeze this repository's f32 and f64 oracles on
the 32x32 hanging sheet (spacing 1/31 m, two top corners pinned, +-1e-3 m
jitter from seed 1, Jacobi, 8 substeps x 10 iterations, no collision), so
that any change of the oracle's arithmetic shows up as a changed digest.

Run from the repository root:  python tests/golden/make_fixtures.py
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


if __name__ == "__main__":
    from oracle import oracle as O
    path = os.path.join(ROOT, "tests", "golden", "oracle_config1.json")
    with open(path, "w") as fh:
        json.dump(fixtures(O), fh, indent=1)
    print("wrote", path)
