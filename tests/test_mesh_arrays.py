"""The topology boundary: xpbd_mesh_from_arrays takes ClothSimConfig's arrays
(cloth_compute.rs:142-179) -- the caller's own mesh, pins and inverse masses
-- and checks every one of them.  Host code only (no GPU).

* A mesh handed back through its own ClothSimConfig arrays is unchanged,
  field for field (rows, batch offsets, CSR).
* Rows without batch offsets are coloured as finalize_cloth_mesh does: the
  product equals the oracle's restatement (mesh_prep.rs:153-221, 580-635),
  and the hand-traced 2 x 1 grid case.
* Inconsistent arrays are refused with a reason.
"""
import os

import numpy as np
import pytest

from test_oracle_pins import HAND_RAW, _check_hand_colouring

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SKIP = {"uvs", "tri_edge_neighbors", "cloth_cols"}   # not part of ClothSimConfig


def _same(a, b):
    for k in a:
        if k in SKIP:
            continue
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape and x.tobytes() == y.tobytes(), k


def _meshes(X):
    with open(os.path.join(GOLDEN, "cloth.obj")) as f:
        obj = f.read()
    return [X.ClothMeshData.sheet(cols=12, rows=9, spacing=0.1, pins=2, jitter=1e-3, seed=3),
            X.ClothMeshData.grid_cloth_hanging(7, 5, 0.05), X.ClothMeshData.parse_welded_obj(obj),
            X.ClothMeshData.batch([X.SheetDesc.make(cols=6, rows=5, spacing=0.1, ground=1),
                                   X.SheetDesc.make(cols=9, rows=4, spacing=0.1, pins=0)])]


def test_round_trip_through_sim_config_arrays(X):
    for m in _meshes(X):
        _same(m.fields(), X.ClothMeshData.from_arrays(**m.sim_config_arrays()).fields())


def test_batch_offsets_without_csr_rebuilds_the_cursor_scatter(X):
    for m in _meshes(X):
        a = m.sim_config_arrays()
        a["neighbor_offsets"] = a["neighbor_packed"] = None
        _same(m.fields(), X.ClothMeshData.from_arrays(**a).fields())


def test_rows_without_batch_offsets_are_coloured_like_the_oracle(O, X):
    rng = np.random.default_rng(5)
    for n, e in [(6, 16), (40, 150), (300, 2000)]:
        pos = rng.normal(size=(n, 3)).astype(np.float32)
        w = rng.uniform(0, 2, n).astype(np.float32)
        i = rng.integers(0, n, e)
        j = (i + 1 + rng.integers(0, n - 1, e)) % n
        rows = [(int(a), int(b), float(r), float(c)) for a, b, r, c in
                zip(i, j, rng.uniform(0, 1, e).astype(np.float32), rng.uniform(0, 1e-3, e).astype(np.float32))]
        want = O.assemble(pos, w, rows)
        got = X.ClothMeshData.from_arrays(pos, w, [r[0] for r in rows], [r[1] for r in rows],
                                          [r[2] for r in rows], [r[3] for r in rows]).fields()
        for k in ("constraint_i", "constraint_j", "constraint_rest", "constraint_comp", "batch_offsets",
                  "neighbor_offsets", "neighbor_other", "neighbor_rest_len", "neighbor_compliance",
                  "neighbor_constraint_id", "inv_mass", "positions", "rest_positions"):
            assert np.asarray(got[k]).tobytes() == np.asarray(want[k]).tobytes(), (n, e, k)


def test_hand_traced_colouring_through_arrays(X):
    pos = [[-0.5, 0, 0], [0, 0, 0], [0.5, 0, 0], [-0.5, 0.5, 0], [0, 0.5, 0], [0.5, 0.5, 0]]
    m = X.ClothMeshData.from_arrays(pos, [1.0] * 6, [c[0] for c in HAND_RAW], [c[1] for c in HAND_RAW],
                                    [c[2] for c in HAND_RAW], [c[3] for c in HAND_RAW])
    _check_hand_colouring(m.fields())


def test_custom_inverse_mass_and_ground_are_kept(X):
    m = X.ClothMeshData.sheet(cols=10, rows=10, spacing=0.1, pins=0)
    a = m.sim_config_arrays()
    w = a["inv_mass"].copy()
    w[[3, 57, 99]] = 0.0
    w[10:20] = 250.0
    g = np.zeros(len(w), np.uint32)
    g[::7] = 1
    a["inv_mass"], a["ground"] = w, g
    f = X.ClothMeshData.from_arrays(**a).fields()
    assert f["inv_mass"].tobytes() == w.tobytes() and f["ground"].tolist() == g.tolist()


def _expect_invalid(X, what, **a):
    with pytest.raises(X.XpbdError) as e:
        X.ClothMeshData.from_arrays(**a)
    assert e.value.code == X.XPBD_E_INVALID
    assert what in str(e.value), str(e.value)


def test_inconsistent_arrays_are_refused(X):
    base = X.ClothMeshData.sheet(cols=6, rows=5, spacing=0.1, pins=2).sim_config_arrays()

    def edit(**kw):
        a = {k: (None if v is None else np.array(v, copy=True)) for k, v in base.items()}
        a.update(kw)
        return a

    ci = base["constraint_i"].copy()
    ci[4] = 30
    _expect_invalid(X, "out of range", **edit(constraint_i=ci))
    cj = base["constraint_j"].copy()
    cj[2] = base["constraint_i"][2]
    _expect_invalid(X, "itself", **edit(constraint_j=cj))
    w = base["inv_mass"].copy()
    w[1] = -1.0
    _expect_invalid(X, "inv_mass", **edit(inv_mass=w))
    p = base["initial_pos"].copy()
    p[2, 1] = np.nan
    _expect_invalid(X, "not finite", **edit(initial_pos=p))
    r = base["constraint_rest_len"].copy()
    r[0] = -0.1
    _expect_invalid(X, "rest length", **edit(constraint_rest_len=r))
    _expect_invalid(X, "vertex-disjoint", **edit(constraint_batch_offsets=[0, len(ci)]))
    _expect_invalid(X, "end at E", **edit(constraint_batch_offsets=base["constraint_batch_offsets"][:-1]))
    pk = base["neighbor_packed"].copy()
    pk[5, 1] += np.float32(1e-3)
    _expect_invalid(X, "rest or compliance", **edit(neighbor_packed=pk))
    pk = base["neighbor_packed"].copy()
    pk[7, 3] = (pk[7, 3] + 1) % len(ci)
    _expect_invalid(X, "does not join", **edit(neighbor_packed=pk))
    pk = base["neighbor_packed"].copy()
    pk[3, 0] += np.float32(0.5)
    _expect_invalid(X, "bad particle or row index", **edit(neighbor_packed=pk))
    _expect_invalid(X, "come together", **edit(neighbor_packed=None))
    _expect_invalid(X, "needs the constraint_batch_offsets", **edit(constraint_batch_offsets=None))
    t = base["tri_indices"].copy()
    t[0] = 99
    _expect_invalid(X, "tri_indices", **edit(tri_indices=t))


def test_device_prepare_without_a_gpu_reports_no_device(X):
    import ctypes as C
    hip = None
    try:
        hip = C.CDLL("libamdhip64.so")
        cnt = C.c_int(0)
        present = hip.hipGetDeviceCount(C.byref(cnt)) == 0 and cnt.value > 0
    except OSError:
        present = False
    if present:
        pytest.skip("a GPU is present")
    with pytest.raises(X.XpbdError) as e:
        X.ClothMeshData.from_arrays([[0, 0, 0], [1, 0, 0]], [1.0, 1.0], [0], [1], [1.0], [0.0], device_prepare=True)
    assert e.value.code == X.XPBD_E_NO_DEVICE


def test_device_prepare_refuses_given_colouring(X):
    base = X.ClothMeshData.sheet(cols=6, rows=5, spacing=0.1, pins=2).sim_config_arrays()
    _expect_invalid(X, "give neither", **dict(base, device_prepare=True))
