"""Batched cloths in the oracle (CPU).

Cloths in one batch share the launch but not their physics: each keeps its
own constraints, pins, ground flag and self-collision grid, and no contact is
ever formed between two cloths even where they occupy the same space.  The
check is therefore exact and independent of the solver's arithmetic: a batch
stepped together equals each cloth stepped alone, bit for bit.

A contact needs the pair closer now than in the rest pose
(xpbd_cpu.rs:95-101); cloths built in the same place have coincident rest
poses, which would hide any cross-cloth contact.  The tests therefore move
cloth k's rest pose by 3k m along x (current positions stay overlapping), so
a missing cloth boundary would show.
"""
import numpy as np
import pytest


def descs_overlapping(O):
    # three cloths in the same region of space (a contact test across cloths
    # would fire on every particle), mixed layouts, pins and ground flags
    return [
        dict(cols=18, rows=14, spacing=0.02, pins=1, jitter=1e-3, seed=5),
        dict(cols=18, rows=14, spacing=0.02, pins=1, jitter=1e-3, seed=6),
        dict(cols=11, rows=23, spacing=0.015, layout=O.LAYOUT_HORIZONTAL_XZ, height=0.05, angle=0.4,
             pins=0, ground=1, jitter=2e-3, seed=7, stretch_compliance=1e-7),
        dict(cols=9, rows=9, spacing=0.03, pins=3, jitter=0.0, seed=8, bend_compliance=1e-2),
    ]


def shift_rest(m, offs=None):
    m = dict(m)
    r = m["rest_positions"].copy()
    offs = m["cloth_offsets"] if offs is None else offs
    for k in range(len(offs) - 1):
        r[offs[k]:offs[k + 1], 0] += np.float32(3.0 * k)
    m["rest_positions"] = r
    return m


def params(O, solver, collision=True):
    return O.params(substeps=8, inner_iterations=6, omega=0.32 if solver == 0 else 1.0, thickness=0.012,
                    coll_scale=1.0, collision=collision, coll_mode=O.COLL_GRID, solver=solver)


def test_batch_positions_are_the_sheets(O):
    ds = descs_overlapping(O)
    b = O.batch([O.sheet_desc(**d) for d in ds])
    offs = b["cloth_offsets"]
    assert len(offs) == len(ds) + 1
    for k, d in enumerate(ds):
        s = O.sheet(**d)
        assert b["positions"][offs[k]:offs[k + 1]].tobytes() == s["positions"].tobytes()
        assert b["inv_mass"][offs[k]:offs[k + 1]].tobytes() == s["inv_mass"].tobytes()


@pytest.mark.parametrize("coll_mode", ["grid", "brute"])
@pytest.mark.parametrize("solver", [0, 1])
def test_batch_equals_cloths_stepped_alone(O, solver, coll_mode):
    ds = descs_overlapping(O)
    b = shift_rest(O.batch([O.sheet_desc(**d) for d in ds]))
    offs = b["cloth_offsets"]
    p = params(O, solver)
    p.coll_mode = O.COLL_GRID if coll_mode == "grid" else O.COLL_BRUTE
    sb = O.Solver(b)
    pos, vel = b["positions"].copy(), np.zeros_like(b["positions"])
    alone = []
    for k, d in enumerate(ds):
        m = dict(O.sheet(**d))
        m["rest_positions"] = m["rest_positions"].copy()
        m["rest_positions"][:, 0] += np.float32(3.0 * k)
        alone.append((O.Solver(m), m["positions"].copy(), np.zeros_like(m["positions"])))
    for _ in range(12):
        pos, vel = sb.substep(pos, vel, p)
        alone = [(s, *s.substep(x, v, p)) for s, x, v in alone]
    for k, (_, x, v) in enumerate(alone):
        assert pos[offs[k]:offs[k + 1]].tobytes() == x.tobytes(), f"cloth {k} positions"
        assert vel[offs[k]:offs[k + 1]].tobytes() == v.tobytes(), f"cloth {k} velocities"


def test_batch_overlap_would_collide_across_cloths(O):
    """The invariant above is not vacuous: cloths 0 and 1 overlap, and a
    single cloth made of both (no cloth boundary) does form contacts between
    them, so its result differs."""
    ds = descs_overlapping(O)[:2]
    b = shift_rest(O.batch([O.sheet_desc(**d) for d in ds]))
    merged = dict(b)
    merged["cloth_offsets"] = np.array([0, b["num_particles"]], np.uint32)
    p = params(O, 0)
    x0 = b["positions"].copy()
    a, _ = O.Solver(b).substep(x0, np.zeros_like(x0), p)
    m, _ = O.Solver(merged).substep(x0, np.zeros_like(x0), p)
    assert a.tobytes() != m.tobytes()


def test_batch_cell_keys_carry_cloth_prefix(O):
    ds = descs_overlapping(O)
    b = O.batch([O.sheet_desc(**d) for d in ds])
    offs = b["cloth_offsets"]
    pos = np.concatenate([b["positions"], b["inv_mass"][:, None]], axis=1)   # vec4, w = inverse mass
    grids, keys = O.cell_keys_batch(pos, offs, 0.012)
    nc = len(ds)
    cb = int(np.ceil(np.log2(nc)))
    for k in range(nc):
        kk = keys[offs[k]:offs[k + 1]]
        assert (kk >> np.uint32(32 - cb) == k).all()
        # local part: the single-cloth key of the cloth's own grid
        g = O.cell_grid(pos[offs[k]:offs[k + 1]], 0.012)
        assert list(g.dims) == list(grids[k].dims)
        local = O.cell_keys(pos[offs[k]:offs[k + 1]], g)
        assert ((kk - (np.uint32(k) << np.uint32(32 - cb))) == local).all()
