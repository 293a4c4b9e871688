"""Topology preparation on the GPU (xpbd_mesh_from_arrays with device_prepare):
the greedy colouring (mesh_prep.rs:153-221), the colour sort and the CSR
cursor scatter (mesh_prep.rs:580-635) computed on the device give the host
path's mesh, field for field, bit for bit.  Needs an MI355X.

Cases: the hand-traced 2 x 1 grid; random rows in uneven cloths; the
reference's cloth.obj; at scale, the 1024 x 1024 sheet (one cloth: one
wavefront walks all 6.3 M rows) and the 512-cloth batch (35 M rows, the
cloths walked in parallel), with the host and device times printed.
"""
import os
import sys
import time

import numpy as np
import pytest

from test_oracle_pins import HAND_RAW, _check_hand_colouring

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("constraint_i", "constraint_j", "constraint_rest", "constraint_comp", "batch_offsets", "neighbor_offsets",
          "neighbor_other", "neighbor_rest_len", "neighbor_compliance", "neighbor_constraint_id", "inv_mass",
          "positions", "cloth_offsets", "ground")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _raw(f):
    """A mesh's rows as a caller would hand them in (no colouring, no CSR)."""
    n = f["num_particles"]
    return dict(initial_pos=f["positions"], inv_mass=f["inv_mass"], constraint_i=f["constraint_i"],
                constraint_j=f["constraint_j"], constraint_rest_len=f["constraint_rest"],
                constraint_compliance=f["constraint_comp"], tri_indices=f["indices"],
                cloth_offsets=f["cloth_offsets"] if len(f["cloth_offsets"]) > 2 else None) if n else None


def _both(X, arrays, label):
    t0 = time.perf_counter()
    hm = X.ClothMeshData.from_arrays(**arrays)
    t1 = time.perf_counter()
    dm = X.ClothMeshData.from_arrays(**arrays, device_prepare=True)
    t2 = time.perf_counter()
    host, dev = hm.fields(), dm.fields()
    note(f"  {label}: xpbd_mesh_from_arrays host {t1 - t0:.3f} s, device {t2 - t1:.3f} s, "
         f"{len(host['constraint_i'])} rows, {len(host['batch_offsets']) - 1} colours")
    for k in FIELDS:
        a, b = np.asarray(host[k]), np.asarray(dev[k])
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), (label, k)
    return host, dev


def test_hand_traced_colouring_on_device(X):
    note("case topology: hand-traced grid")
    pos = [[-0.5, 0, 0], [0, 0, 0], [0.5, 0, 0], [-0.5, 0.5, 0], [0, 0.5, 0], [0.5, 0.5, 0]]
    m = X.ClothMeshData.from_arrays(pos, [1.0] * 6, [c[0] for c in HAND_RAW], [c[1] for c in HAND_RAW],
                                    [c[2] for c in HAND_RAW], [c[3] for c in HAND_RAW], device_prepare=True)
    _check_hand_colouring(m.fields())


def test_random_rows_in_uneven_cloths(X):
    note("case topology: random rows")
    rng = np.random.default_rng(17)
    for sizes in ((7,), (300,), (50, 3, 900, 41), (2000,) * 3):
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
        ci, cj = [], []
        for c, sz in enumerate(sizes):
            e = 3 * sz
            i = rng.integers(0, sz, e)
            j = (i + 1 + rng.integers(0, max(sz - 1, 1), e)) % sz
            keep = i != j
            ci.append(i[keep] + offs[c])
            cj.append(j[keep] + offs[c])
        ci, cj = np.concatenate(ci), np.concatenate(cj)
        # a few exact duplicates: the row index breaks their tie
        dup = rng.integers(0, len(ci), max(1, len(ci) // 20))
        ci, cj = np.concatenate([ci, cj[dup]]), np.concatenate([cj, ci[dup]])
        n = int(offs[-1])
        arrays = dict(initial_pos=rng.normal(size=(n, 3)).astype(np.float32),
                      inv_mass=rng.uniform(0, 2, n).astype(np.float32), constraint_i=ci, constraint_j=cj,
                      constraint_rest_len=rng.uniform(0, 1, len(ci)).astype(np.float32),
                      constraint_compliance=rng.uniform(0, 1e-3, len(ci)).astype(np.float32),
                      cloth_offsets=offs if len(sizes) > 1 else None)
        try:
            _both(X, arrays, f"random {sizes}")
        except X.XpbdError as exc:
            # random graphs may allow more than 64 colours: the device path refuses them
            assert exc.code == X.XPBD_E_UNSUPPORTED, exc


def test_cloth_obj_on_device(X):
    note("case topology: cloth.obj")
    with open(os.path.join(GOLDEN, "cloth.obj")) as fh:
        f = X.ClothMeshData.parse_welded_obj(fh.read()).fields()
    _both(X, _raw(f), "cloth.obj")


def test_too_many_colours_is_refused_before_launch(X):
    note("case topology: star of 40")
    n = 41
    with pytest.raises(X.XpbdError) as e:
        X.ClothMeshData.from_arrays(np.zeros((n, 3), np.float32), np.ones(n, np.float32), [0] * 40,
                                    list(range(1, 41)), [1.0] * 40, [0.0] * 40, device_prepare=True)
    assert e.value.code == X.XPBD_E_UNSUPPORTED


@pytest.mark.parametrize("which", ["sheet1024", "batch512"])
def test_at_scale_equals_host(X, which):
    note(f"case topology: {which}")
    if which == "sheet1024":
        f = X.ClothMeshData.sheet(cols=1024, rows=1024, spacing=1 / 1023, pins=2, jitter=1e-3, seed=3).fields()
    else:
        from bevy_xpbd_amd.workloads import batch_descs
        descs, _ = batch_descs(count=512, seed=4, spacing=0.01)
        f = X.ClothMeshData.batch(descs).fields()
    _both(X, _raw(f), which)
