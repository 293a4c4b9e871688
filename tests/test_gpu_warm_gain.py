"""What the trained warm start gains over the method without it (GPU).

The network is trained on the GPU (solver-generated pairs, BF16 MFMA
kernels), then, from a held-out plain-solver state of 32x32 stiff sheets,
one substep of warm start + 2 Jacobi iterations must end with a lower
relative constraint residual than 2 plain iterations (the same iteration
budget without the network), and a lower one than 10 plain iterations.  The
plain-iteration equivalent is printed.  Multi-step use is guarded by the
residual gate, tested in test_gpu_warmstart.py.
"""
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_trained_warm_start_beats_the_same_budget_without_it(X):
    from bevy_xpbd_amd import warmstart as W
    print("case warm gain", flush=True)
    feat, tgt, rows, th = W.generate_dataset(seed=1, frames=(3, 12), target_iterations=200)
    mlp, curve = W.train(feat, tgt, rows, th, steps=2000, batch=4096, lr=2e-3)
    assert curve[-50:].mean() < 0.25 * curve[:10].mean()
    mesh = X.ClothMeshData.batch(W.stiff_eval_descs())
    sim = W.make_sim(mesh)
    sim.step_frames(2)
    state = sim.read_state()
    sim.close()
    r = W.substep_residuals(mesh, state, (2, 10, 20, 50, 100), mlp, warm_iterations=(2,))
    rw = r["warm"][2]
    eq = W.equivalent_iterations(r["plain"], rw)
    print(f"  plain {r['plain']}  warm+2 {rw:.4e}  ~ plain I = {eq:.1f}", flush=True)
    sys.stdout.flush()
    assert rw < r["plain"][2]
    assert rw < r["plain"][10]
    mlp.close()
