"""The warm-start network's CPU oracle (oracle/mlp_oracle.py), pinned by
finite differences, closed forms and invariants (CPU only)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_gelu_derivative_matches_finite_differences():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(x))) < 1e-8


def test_gelu_known_values():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([10.0]))[0] - 10.0) < 1e-12       # identity far right
    assert abs(M.gelu(np.array([-10.0]))[0]) < 1e-12             # zero far left


@pytest.mark.parametrize("bf16_points", [False, True])
def test_zero_weights_output_last_bias(bf16_points):
    P = np.zeros(M.PARAMS, np.float32)
    P[M.OFF_B3:M.OFF_B3 + 3] = [0.25, -1.5, 3.0]
    P[M.OFF_B1:M.OFF_W2] = 0.7   # hidden biases do not reach the output
    X = np.random.default_rng(0).normal(size=(9, 32)).astype(np.float32)
    out, _ = M.forward(P, X, bf16_points)
    assert np.all(out == np.array([0.25, -1.5, 3.0]))


def test_backward_matches_central_differences():
    rng = np.random.default_rng(3)
    P = M.init_params(seed=4).astype(np.float64)
    X = rng.normal(size=(24, 32))
    X[:, 30:] = 0
    T = rng.normal(size=(24, 3))
    _, g = M.loss_and_grad(P.astype(np.float32), X, T, 1.0, bf16_points=False)
    # float64 loss as a function of float64 parameters (no float32 rounding of P)
    def loss64(Pd):
        W1 = Pd[M.OFF_W1:M.OFF_B1].reshape(32, 64)
        W2 = Pd[M.OFF_W2:M.OFF_B2].reshape(64, 64)
        W3 = Pd[M.OFF_W3:M.OFF_B3].reshape(64, 16)
        h1 = M.gelu(X @ W1 + Pd[M.OFF_B1:M.OFF_W2])
        h2 = M.gelu(h1 @ W2 + Pd[M.OFF_B2:M.OFF_W3])
        o = (h2 @ W3 + Pd[M.OFF_B3:])[:, :3]
        return np.sum((o - T) ** 2) / (3 * len(X))
    P64 = P.astype(np.float32).astype(np.float64)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_W1 + 30 * 64), 60, replace=False),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 10, replace=False),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 60, replace=False),
                          rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 10, replace=False),
                          [M.OFF_W3 + 16 * k + o for k in (0, 5, 63) for o in range(3)],
                          [M.OFF_B3, M.OFF_B3 + 1, M.OFF_B3 + 2]])
    h = 1e-6
    for p in idx:
        a, b = P64.copy(), P64.copy()
        a[p] += h
        b[p] -= h
        fd = (loss64(a) - loss64(b)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(fd)), (p, fd, g[p])
    # padded parameters carry no gradient
    assert np.all(g[M.fixed_zero_mask()] == 0)


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (mantissa 0)
    assert r[2] == np.float32(1.015625)     # tie -> even (up)
    assert r[3] == 1.0                      # below half ulp
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_grid_topology_and_features_closed_forms(X):
    """3-column sheet: neighbour slots and compliance classes; a uniform
    stretch along x gives the exact strain in the +x / -x features; a rigid
    translation gives zero displacement features."""
    mesh = X.ClothMeshData.sheet(cols=5, rows=4, spacing=0.25, pins=0, stretch_compliance=1e-6,
                                 shear_compliance=1e-5, bend_compliance=1e-3)
    f = mesh.fields()
    nbr, aux = M.grid_topology(f)
    i = 1 * 5 + 2   # col 2, row 1
    assert list(nbr[i]) == [i + 1, i - 1, i + 5, i - 5, i + 6, i - 6, i - 4, i + 4]
    assert list(nbr[0]) == [1, -1, 5, -1, 6, -1, -1, -1]
    assert np.allclose(aux[i, :3], [(np.log10(c) + 5.5) / 2.5 for c in (1e-6, 1e-5, 1e-3)], atol=1e-6)
    assert aux[i, 3] == np.float32(4.0)
    rest = f["rest_positions"]
    xp = rest.copy()
    xp[:, 0] *= np.float32(1.5)          # stretch 50% along x
    feats = M.features(xp, np.zeros_like(xp), rest, nbr, aux, 1.0 / 480.0)
    assert feats[i, 0] == np.float32(0.5) and feats[i, 3] == np.float32(-0.5)   # +x: +0.5, -x: -0.5 (x comp)
    assert np.all(feats[i, [1, 2, 4, 5]] == 0)
    moved = rest + np.float32([0.5, 0.25, -0.125])
    feats = M.features(moved, np.zeros_like(rest), rest, nbr, aux, 1.0 / 480.0)
    assert np.all(feats[:, :24] == 0)


def test_grid_topology_array_form_equals_loop():
    """The array form used at 256x256 scale gives the loop's neighbours and
    feature constants, bit for bit, on batches with uneven cloths."""
    import bevy_xpbd_amd as X
    descs = [X.SheetDesc.make(cols=32, rows=32, spacing=1 / 31, pins=1, jitter=1e-3, seed=1, stretch_compliance=0.0),
             X.SheetDesc.make(cols=40, rows=24, spacing=0.02, layout=1, height=1.0, angle=0.7, pins=3, jitter=1e-3,
                              seed=3, shear_compliance=1e-4, bend_compliance=1e-2),
             X.SheetDesc.make(cols=3, rows=7, spacing=0.02, stretch_compliance=1e-7),
             X.SheetDesc.make(cols=2, rows=2, spacing=0.05)]
    f = X.ClothMeshData.batch(descs).fields()
    a, b = M.grid_topology(f), M.grid_topology_loop(f)
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
