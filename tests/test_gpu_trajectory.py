"""Long trajectories on the GPU (through the C ABI), against the oracle.

- 32x32 hanging sheet, Jacobi 8x10, 120 frames replayed from the hipGraph:
  bit-exact against the f32 oracle's frozen trajectory (tests/golden), and
  within the stated trajectory tolerance of the f64 oracle (1e-4 m; the f32
  solver itself is 1.8e-5 m away after 120 frames).
- The error against the f64 oracle is printed for each config run here.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from _helpers import bits_equal, max_abs, oracle_params

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config1_120_frames_graph_matches_oracle(O, X):
    want = json.load(open(os.path.join(HERE, "golden", "oracle_config1.json")))
    sheet = want["sheet"]
    note("case config1 120 frames")
    mesh = X.ClothMeshData.sheet(**sheet)
    cfg = X.ClothSimConfig.default(solve_substeps=8, solve_inner_iterations=10, use_graph=1)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 8, coll_scale=0.0)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    f = mesh.fields()
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=False)
    x64 = f["positions"].astype(np.float64)
    v64 = np.zeros_like(x64)
    done = 0
    for frames in (1, 10, 120):
        sim.step_frames(frames - done)
        x64, v64 = s.frames(x64, v64, p, 8, 1, frames - done, dtype=np.float64)
        done = frames
        gp, _ = sim.read_state()
        w = want["frames"][str(frames)]
        digest = hashlib.sha256(np.ascontiguousarray(gp[:, :3], np.float32).tobytes()).hexdigest()
        assert digest == w["f32_sha256"], f"frame {frames}: GPU trajectory differs from the f32 oracle"
        err = max_abs(gp, x64)
        note(f"  config1 frame {frames}: max |GPU f32 - oracle f64| = {err:.3e} m")
        assert err < 1e-4
    sim.close()


def test_config3_one_frame_error_against_f64(O, X):
    """1024x1024 sheet with self-collision, one frame (8 substeps x 10):
    bit-exact against the f32 oracle; the f64 error is reported."""
    note("case config3 one frame")
    h = 1 / 1023
    mesh = X.ClothMeshData.sheet(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    cfg = X.ClothSimConfig.default(solve_substeps=8, solve_inner_iterations=10, use_graph=1)
    u = X.ClothSimUniforms.default(X.SOLVER_JACOBI, 8, thickness=2 * 0.4 * h)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    f = mesh.fields()
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    sim.step_frames(1)
    gp, _ = sim.read_state()
    x32, _ = s.frames(f["positions"].copy(), np.zeros_like(f["positions"]), p, 8, 1, 1)
    assert bits_equal(gp, x32)
    x64, _ = s.frames(f["positions"].astype(np.float64), np.zeros(f["positions"].shape), p, 8, 1, 1,
                      dtype=np.float64)
    note(f"  config3 frame 1: max |GPU f32 - oracle f64| = {max_abs(gp, x64):.3e} m")
    sim.close()
