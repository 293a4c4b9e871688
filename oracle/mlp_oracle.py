"""CPU oracle of the learned warm start (test infrastructure only).

The warm-start network has no counterpart in the reference crate; this file
restates the build's own definitions in float64 numpy so the GPU kernels can
be checked against it:

- grid neighbours and per-vertex feature constants (neighbour slots +x, -x,
  +y, -y, (+x,+y), (-x,-y), (+x,-y), (-x,+y); scaled log10 compliance of the
  stretch / shear / bend partners; 1 / spacing);
- the 30 features of a predicted state, rounded to bf16 as the GPU stores them;
- the MLP 30 -> 64 -> 64 -> 3 with tanh-form GELU, forward and backward of
  the mean squared error.  With `bf16=True` the values the GPU feeds to its
  BF16 matrix cores (inputs, weights, hidden activations, output and hidden
  gradients) are rounded to bf16 at the same points; all sums are float64.

Pinned by: central finite differences of the float64 loss (backward), the
derivative of GELU against finite differences, and closed forms (a network
with zero weights outputs its last bias).  Only tests/ use this module.
"""
from __future__ import annotations

import numpy as np

FEAT, IN, HID, OUT, OUTP = 30, 32, 64, 3, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * OUTP
PARAMS = OFF_B3 + OUTP
GELU_C = 0.7978845608028654
GELU_A = 0.044715

DX = (1, -1, 0, 0, 1, -1, 1, -1)
DY = (0, 0, 1, -1, 1, -1, -1, 1)


def bf16(x):
    """Round float32 values to bf16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def scaled_log_compliance(c):
    c = np.float32(max(np.float32(c), np.float32(1e-10)))
    lg = np.float32(np.log10(np.float64(c)))
    return np.float32((lg + np.float32(5.5)) / np.float32(2.5))


def grid_topology(f):
    """nbr8 (N x 8 int32, -1 absent) and vaux (N x 4 float32) from mesh
    fields; array form of grid_topology_loop (same results, checked against
    it on small meshes), for meshes of a million vertices."""
    offs = np.asarray(f["cloth_offsets"], np.int64)
    colsv = np.asarray(f["cloth_cols"], np.int64)
    noff = np.asarray(f["neighbor_offsets"], np.int64)
    no = np.asarray(f["neighbor_other"], np.int64)
    nr, nc = np.asarray(f["neighbor_rest_len"], np.float32), np.asarray(f["neighbor_compliance"], np.float32)
    n = len(f["inv_mass"])
    nbr = np.full((n, 8), -1, np.int32)
    aux = np.zeros((n, 4), np.float32)
    cloth = np.repeat(np.arange(len(colsv)), np.diff(offs)[:len(colsv)])
    idx = np.arange(len(cloth))
    off, cols = offs[cloth], colsv[cloth]
    rows = (offs[cloth + 1] - off) // cols
    col, row = (idx - off) % cols, (idx - off) // cols
    for k in range(8):
        x, y = col + DX[k], row + DY[k]
        ok = (x >= 0) & (x < cols) & (y >= 0) & (y < rows)
        nbr[idx[ok], k] = (off + y * cols + x)[ok]
    # CSR entries of the grid vertices: owner, partner offsets in its cloth
    owner = np.repeat(np.arange(n), np.diff(noff))
    keep = owner < len(cloth)
    q = np.nonzero(keep)[0]
    i, j = owner[q], no[q]
    same = (j >= off[i]) & (j < offs[cloth[i] + 1])
    q, i, j = q[same], i[same], j[same]
    dc = np.abs((j - off[i]) % cols[i] - col[i])
    dr = np.abs((j - off[i]) // cols[i] - row[i])
    cls = {"stretch": dc + dr == 1, "shear": (dc == 1) & (dr == 1),
           "bend": ((dc == 2) & (dr == 0)) | ((dc == 0) & (dr == 2))}
    comp = np.full((len(cloth), 3), np.float32(0))
    hs = np.zeros(len(cloth), np.float32)
    for a, name in enumerate(("stretch", "shear", "bend")):
        m = cls[name]
        first_owner, first = np.unique(i[m], return_index=True)   # entries are in CSR order
        comp[first_owner, a] = nc[q[m][first]]
        if name == "stretch":
            hs[first_owner] = nr[q[m][first]]
    lc = np.log10(np.maximum(comp, np.float32(1e-10)).astype(np.float64)).astype(np.float32)
    aux[:len(cloth), :3] = ((lc + np.float32(5.5)) / np.float32(2.5)).astype(np.float32)
    inv = np.zeros_like(hs)
    inv[hs > 0] = (np.float32(1.0) / hs[hs > 0]).astype(np.float32)
    aux[:len(cloth), 3] = inv
    return nbr, aux


def grid_topology_loop(f):
    """nbr8 (N x 8 int32, -1 absent) and vaux (N x 4 float32) from mesh fields."""
    offs, colsv = f["cloth_offsets"], f["cloth_cols"]
    no, nr, nc, noff = f["neighbor_other"], f["neighbor_rest_len"], f["neighbor_compliance"], f["neighbor_offsets"]
    n = len(f["inv_mass"])
    nbr = np.full((n, 8), -1, np.int32)
    aux = np.zeros((n, 4), np.float32)
    for c in range(len(colsv)):
        off, end, cols = int(offs[c]), int(offs[c + 1]), int(colsv[c])
        rows = (end - off) // cols
        for i in range(off, end):
            col, row = (i - off) % cols, (i - off) // cols
            for k in range(8):
                x, y = col + DX[k], row + DY[k]
                if 0 <= x < cols and 0 <= y < rows:
                    nbr[i, k] = off + y * cols + x
            cs = csh = cb = hs = np.float32(0)
            fs = fsh = fb = False
            for q in range(int(noff[i]), int(noff[i + 1])):
                j = int(no[q])
                if j < off or j >= end:
                    continue
                dc, dr = abs((j - off) % cols - col), abs((j - off) // cols - row)
                if not fs and dc + dr == 1:
                    fs, cs, hs = True, nc[q], nr[q]
                elif not fsh and dc == 1 and dr == 1:
                    fsh, csh = True, nc[q]
                elif not fb and ((dc == 2 and dr == 0) or (dc == 0 and dr == 2)):
                    fb, cb = True, nc[q]
            aux[i] = (scaled_log_compliance(cs), scaled_log_compliance(csh), scaled_log_compliance(cb),
                      np.float32(1.0) / np.float32(hs) if hs > 0 else np.float32(0))
    return nbr, aux


def features(xp, vel, rest, nbr, aux, dt):
    """30 features (+2 zero pads) of the predicted positions xp (N x >=3 f32),
    float32 arithmetic in the GPU's order, then rounded to bf16."""
    xp = np.asarray(xp, np.float32)[:, :3]
    rest = np.asarray(rest, np.float32)[:, :3]
    vel = np.asarray(vel, np.float32)[:, :3]
    n = len(xp)
    out = np.zeros((n, IN), np.float32)
    ih = aux[:, 3]
    for k in range(8):
        j = nbr[:, k]
        ok = j >= 0
        jj = np.where(ok, j, 0)
        d = (xp[jj] - xp) - (rest[jj] - rest)
        d = (d * ih[:, None]).astype(np.float32)
        out[:, 3 * k:3 * k + 3] = np.where(ok[:, None], d, np.float32(0))
    sv = (np.float32(dt) * ih).astype(np.float32)
    out[:, 24:27] = vel * sv[:, None]
    out[:, 27:30] = aux[:, :3]
    return bf16(out)


def unpack(P, round_weights):
    P = np.asarray(P, np.float32)
    W1 = P[OFF_W1:OFF_B1].reshape(IN, HID)
    W2 = P[OFF_W2:OFF_B2].reshape(HID, HID)
    W3 = P[OFF_W3:OFF_B3].reshape(HID, OUTP)
    if round_weights:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    f = np.float64
    return (W1.astype(f), P[OFF_B1:OFF_W2].astype(f), W2.astype(f), P[OFF_B2:OFF_W3].astype(f), W3.astype(f),
            P[OFF_B3:PARAMS].astype(f))


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def _r(x, on):
    return bf16(x).astype(np.float64) if on else x


def forward(P, X, bf16_points=True):
    """X: rows x 32 (float); returns (out rows x 3, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(P, bf16_points)
    X = _r(np.asarray(X, np.float32), bf16_points).astype(np.float64)
    z1 = X @ W1 + b1
    h1 = _r(gelu(z1), bf16_points)
    z2 = h1 @ W2 + b2
    h2 = _r(gelu(z2), bf16_points)
    out = h2 @ W3 + b3
    return out[:, :OUT], (X, z1, h1, z2, h2, out, W1, W2, W3)


def loss_and_grad(P, X, T, target_scale=1.0, bf16_points=True):
    """Mean squared error over rows x 3 against T * target_scale and its
    gradient (PARAMS float64), with the GPU's rounding points when asked."""
    out, (X, z1, h1, z2, h2, o, W1, W2, W3) = forward(P, X, bf16_points)
    rows = len(X)
    d = out - np.asarray(T, np.float64) * np.float64(np.float32(target_scale))
    loss = float(np.sum(d * d) / (3.0 * rows))
    dO = np.zeros((rows, OUTP))
    dO[:, :OUT] = 2.0 * d / (3.0 * rows)
    dOr = _r(dO, bf16_points)
    dz2 = (dOr @ W3.T) * gelu_grad(z2)
    dz2r = _r(dz2, bf16_points)
    dz1 = (dz2r @ W2.T) * gelu_grad(z1)
    dz1r = _r(dz1, bf16_points)
    g = np.zeros(PARAMS)
    g[OFF_W1:OFF_B1] = (X.T @ dz1r).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2r).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = (h2.T @ dOr).ravel()
    g[OFF_B3:PARAMS] = dO.sum(0)
    fixed = fixed_zero_mask()
    g[fixed] = 0.0
    return loss, g


def fixed_zero_mask():
    m = np.zeros(PARAMS, bool)
    m[OFF_W1 + FEAT * HID:OFF_B1] = True
    w3 = np.arange(OFF_W3, OFF_B3)
    m[w3[(w3 - OFF_W3) % OUTP >= OUT]] = True
    m[OFF_B3 + OUT:PARAMS] = True
    return m


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(PARAMS, np.float32)
    P[OFF_W1:OFF_B1] = (rng.normal(0, np.sqrt(2 / FEAT), (IN, HID)) * scale).ravel()
    P[OFF_B1:OFF_W2] = rng.normal(0, 0.1, HID)
    P[OFF_W2:OFF_B2] = (rng.normal(0, np.sqrt(2 / HID), (HID, HID)) * scale).ravel()
    P[OFF_B2:OFF_W3] = rng.normal(0, 0.1, HID)
    P[OFF_W3:OFF_B3] = (rng.normal(0, np.sqrt(1 / HID), (HID, OUTP)) * scale).ravel()
    P[OFF_B3:PARAMS] = rng.normal(0, 0.1, OUTP)
    P[fixed_zero_mask()] = 0
    return P
