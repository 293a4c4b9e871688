"""Collision every N substeps, and the device guard of persistent launches.
Needs an MI355X.

* collision_every_n_substeps = 4 and 8 (the reference demo's budget uses 8,
  examples/cloth.rs:106; rule si % N == N - 1, cloth_jacobi.rs:700-704):
  whole frames, eager and replayed from the hipGraph, Jacobi and coloured
  Gauss-Seidel, bit-exact against the oracle's frame driver after every
  frame, on a squeezed sheet that is in contact from the first substep;
* the error word: once set (here from the host, as an expired barrier wait
  sets it), every barrier of a persistent launch returns at once, the next
  sync reports XPBD_E_STATE and clears the word, and after a reset the
  handle steps bit-exact again.
"""
import sys
import time

import numpy as np
import pytest

from _helpers import bits_equal, max_abs, oracle_params

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _squeezed(X):
    mesh = X.ClothMeshData.sheet(cols=36, rows=30, spacing=0.02, layout=1, height=0.4, angle=1.1, pins=3,
                                 jitter=2e-3, seed=21)
    f = mesh.fields()
    c = f["positions"].mean(axis=0)
    f["positions"] = ((f["positions"] - c) * np.float32(0.6) + c).astype(np.float32)
    return mesh, f


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("solver,gs_launch", [(0, 0), (1, 1), (1, 2), (1, 3)],
                         ids=["jacobi", "gs_persistent", "gs_colours", "gs_tiles"])
@pytest.mark.parametrize("stride", [4, 8])
def test_collision_stride_frames_bit_exact(O, X, stride, solver, gs_launch, graph):
    note(f"case stride={stride} solver={solver} gs_launch={gs_launch} graph={graph}")
    mesh, f = _squeezed(X)
    S = 8
    cfg = X.ClothSimConfig.default(solve_substeps=S, solve_inner_iterations=6, solver=solver,
                                   collision_every_n_substeps=stride, use_graph=graph, autotune=0,
                                   gs_launch=gs_launch)
    u = X.ClothSimUniforms.default(solver, S, thickness=0.016)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    sim.write_state(f["positions"])
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    pos, vel = f["positions"].copy(), np.zeros_like(f["positions"])
    for frame in range(4):
        sim.step_frames(1)
        pos, vel = s.frames(pos, vel, p, S, stride, 1)
        gp, gv = sim.read_state()
        assert bits_equal(gp, pos), f"frame {frame}: max |d| {max_abs(gp, pos):.3e}"
        assert bits_equal(gv, vel), f"frame {frame}: velocities differ"
    # the stride matters here: every-substep collision gives other results
    q, _ = s.frames(f["positions"], np.zeros_like(pos), p, S, 1, 4)
    assert not bits_equal(q, pos)
    sim.close()


@pytest.mark.parametrize("mode", ["gauss_seidel", "gauss_seidel_tiles", "persistent_jacobi"])
def test_error_word_ends_persistent_launches_promptly(O, X, mode):
    note(f"case error word {mode}")
    solver = 1 if mode.startswith("gauss_seidel") else 0
    mesh = X.ClothMeshData.sheet(cols=48, rows=40, spacing=0.02, pins=2, jitter=1e-3, seed=13)
    f = mesh.fields()
    cfg = X.ClothSimConfig.default(solve_substeps=4, solve_inner_iterations=8, solver=solver, autotune=0,
                                   persistent_jacobi=1 if solver == 0 else 0,
                                   gs_launch=3 if mode == "gauss_seidel_tiles" else 0)
    u = X.ClothSimUniforms.default(solver, 4, thickness=0.016)
    sim = X.ClothComputePlugin(mesh, cfg, u)
    sim.raise_device_error()
    t0 = time.perf_counter()
    sim.step_substeps(4, True)
    with pytest.raises(X.XpbdError) as e:
        sim.sync()
    elapsed = time.perf_counter() - t0
    assert e.value.code == X.XPBD_E_STATE
    assert elapsed < 2.0, elapsed
    # the report cleared the word: reset, and the handle is exact again
    sim.reset()
    s = O.Solver(f)
    p = oracle_params(O, X, u, cfg, collision=True)
    pos, vel = f["positions"].copy(), np.zeros_like(f["positions"])
    for k in range(2):
        sim.step_substeps(1, True)
        pos, vel = s.substep(pos, vel, p)
    sim.sync()
    gp, gv = sim.read_state()
    assert bits_equal(gp, pos) and bits_equal(gv, vel), max_abs(gp, pos)
    sim.close()
