"""Host topology of the product (libxpbd_hip.so, no GPU needed) against the
oracle's restatement of src/mesh_prep.rs: bit-exact, field by field —
colouring order, CSR cursor order, rest lengths, inverse masses, pins.
Plus the colouring invariants the reference relies on (mesh_prep.rs:153-221)."""
import numpy as np
import pytest

FIELDS = ["positions", "rest_positions", "normals", "uvs", "indices", "tri_edge_neighbors", "inv_mass",
          "neighbor_other", "neighbor_rest_len", "neighbor_compliance", "neighbor_constraint_id",
          "neighbor_offsets", "constraint_i", "constraint_j", "constraint_rest", "constraint_comp",
          "batch_offsets", "ground", "cloth_offsets"]


def assert_same_mesh(a: dict, b: dict):
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert x.shape == y.shape, (f, x.shape, y.shape)
        assert x.tobytes() == y.tobytes(), f"field {f} differs"


def check_colouring(m: dict):
    ci, cj, offs = m["constraint_i"], m["constraint_j"], m["batch_offsets"]
    assert offs[0] == 0 and offs[-1] == len(ci) and (np.diff(offs) > 0).all()
    for b in range(len(offs) - 1):
        a, c = ci[offs[b]:offs[b + 1]], cj[offs[b]:offs[b + 1]]
        keep = a != c   # a degenerate face yields an (i, i) row; it shares only with itself
        v = np.concatenate([a, c[keep]])
        assert len(np.unique(v)) == len(v), f"colour {b} shares a particle"


def grid_obj(cols, rows, h):
    """A welded quad-grid OBJ like a Blender export (v / vt / f v/vt)."""
    lines = ["# generated grid", "o Grid"]
    for r in range(rows):
        for c in range(cols):
            lines.append(f"v {c * h - 0.5:.6f} {r * h - 0.75:.6f} -0.000000")
    for r in range(rows):
        for c in range(cols):
            lines.append(f"vt {c / (cols - 1):.6f} {r / (rows - 1):.6f}")
    for r in range(rows - 1):
        for c in range(cols - 1):
            a = r * cols + c + 1
            lines.append(f"f {a}/{a} {a + 1}/{a + 1} {a + cols + 1}/{a + cols + 1} {a + cols}/{a + cols}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("qc,qr,cell", [(1, 1, 0.1), (16, 16, 0.045), (18, 18, 0.042), (12, 10, 0.05),
                                        (31, 7, 0.033)])
def test_grid_cloth_hanging_bit_exact(O, X, qc, qr, cell):
    a = O.grid_cloth_hanging(qc, qr, cell)
    b = X.ClothMeshData.grid_cloth_hanging(qc, qr, cell).fields()
    assert_same_mesh(a, b)
    check_colouring(b)


@pytest.mark.parametrize("text", [
    "\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n",
    "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0.2\nf 1 2 3 4\nf 4 3 5\n",
    "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0.5 1.6 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1 5//1\n# c\nf 9 9 9\n",
    grid_obj(21, 31, 0.05),
])
def test_parse_welded_obj_bit_exact(O, X, text):
    a = O.parse_welded_obj(text)
    b = X.ClothMeshData.parse_welded_obj(text).fields()
    assert_same_mesh(a, b)
    check_colouring(b)


SHEETS = [
    dict(cols=32, rows=32, spacing=1 / 31, pins=1, jitter=1e-3, seed=7),
    dict(cols=64, rows=40, spacing=0.01, layout=1, height=1.0, angle=2.1, pins=3),
    dict(cols=17, rows=5, spacing=0.05, pins=2, ground=1, stretch_compliance=1e-8, shear_compliance=3e-4,
         bend_compliance=1e-3),
    dict(cols=2, rows=2, spacing=0.1, pins=0),
]


@pytest.mark.parametrize("kw", SHEETS)
def test_sheet_bit_exact(O, X, kw):
    a = O.sheet(**kw)
    b = X.ClothMeshData.sheet(**kw).fields()
    assert_same_mesh(a, b)
    check_colouring(b)


def test_config_sheet_counts(X):
    """32x32 config: 1984 stretch + 1922 shear + 1920 bend = 5826, 12 colours."""
    m = X.ClothMeshData.sheet(cols=32, rows=32, spacing=1 / 31, pins=1).fields()
    assert m["num_constraints"] == 5826
    assert len(m["batch_offsets"]) - 1 == 12
    comp = m["constraint_comp"]
    assert (comp == 0).sum() == 1984 and (comp == np.float32(1e-6)).sum() == 1922
    assert (comp == np.float32(1e-4)).sum() == 1920
    w = m["inv_mass"]
    assert (w == 0).sum() == 2 and w[31 * 32] == 0 and w[31 * 32 + 31] == 0


def test_batch_bit_exact(O, X):
    rng = np.random.default_rng(3)
    descs = []
    for c in range(6):
        kw = dict(cols=int(rng.integers(4, 20)), rows=int(rng.integers(4, 20)), spacing=0.02,
                  pins=int(rng.integers(0, 3)), ground=int(rng.integers(0, 2)),
                  stretch_compliance=float(10 ** rng.uniform(-8, -3)),
                  shear_compliance=float(10 ** rng.uniform(-8, -3)),
                  bend_compliance=float(10 ** rng.uniform(-8, -3)), jitter=1e-3, seed=c, height=0.5)
        descs.append(kw)
    a = O.batch([O.sheet_desc(**d) for d in descs])
    b = X.ClothMeshData.batch([X.SheetDesc.make(**d) for d in descs]).fields()
    assert_same_mesh(a, b)
    check_colouring(b)
    assert len(b["cloth_offsets"]) == 7


@pytest.mark.slow
def test_sheet_256_bit_exact(O, X):
    kw = dict(cols=256, rows=256, spacing=1 / 255, layout=1, height=1.0, angle=0.7, pins=3)
    a = O.sheet(**kw)
    b = X.ClothMeshData.sheet(**kw).fields()
    assert_same_mesh(a, b)
    assert b["num_constraints"] == 390658 and len(b["batch_offsets"]) - 1 == 12


@pytest.mark.parametrize("th", [0.04, 0.01, 0.2])
def test_derive_collision_grid_bit_exact(O, X, th):
    m = O.grid_cloth_hanging(24, 18, 0.045)
    rest = np.zeros((m["num_particles"], 4), np.float32)
    rest[:, :3] = m["rest_positions"]
    a = O.derive_collision_grid(rest, th)
    b = X.derive_collision_grid(rest, th)
    assert a[0].tobytes() == b[0].tobytes()
    assert np.float32(a[1]).tobytes() == np.float32(b[1]).tobytes()
    assert (a[2] == b[2]).all() and a[3] == b[3] and a[4] == b[4]
    assert a[3] <= 32768 and (a[2] <= 96).all()


def test_derive_collision_grid_hand_case(X):
    """Unit cube of rest points, thickness 0.04: margin = 0.04*6 + 0.02 = 0.26,
    extent 1.52 per axis, target cells 256, crude side (1.52^3/256)^(1/3),
    which already gives ceil(1.52/side) = 7 per axis -> 343 cells, 2 digits."""
    pts = np.array([[x, y, z, 0] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
    origin, inv, dims, cells, digits = X.derive_collision_grid(pts, 0.04)
    assert np.allclose(origin, -0.26, atol=1e-6)
    side = (1.52 ** 3 / 256) ** (1 / 3)
    assert np.isclose(1 / inv, side, rtol=1e-5)
    assert list(dims) == [7, 7, 7] and cells == 343 and digits == 2
