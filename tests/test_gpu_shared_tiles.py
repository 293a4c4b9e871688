"""Grid tiles with each constraint evaluated once per tile (grid_tiles = 3)
against the CPU oracle, through the C ABI.  Needs an MI355X.

The tile's constraints are evaluated once, as their first endpoint sees them,
and each particle adds the stored normal and dlam with its own inverse mass,
negated for the second endpoint; the bar is the same as for every other
layout: bit-exact f32 positions and velocities, substep by substep.
"""
import numpy as np
import pytest

from test_gpu_parity import make, note, run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cols,rows,fast,fused", [(64, 37, 0, 0), (64, 37, 1, 0), (128, 16, 1, 1), (128, 16, 0, 2),
                                                (192, 9, 1, 0)])
def test_shared_tiles_bit_exact(O, X, cols, rows, fast, fused):
    """Columns a multiple of 64, a partial last band of tiles, predict fused
    or separate, both division forms, with collision (its bounds iteration
    runs on the shared tiles too)."""
    note(f"case shared tiles {cols}x{rows} fast={fast} fused={fused}")
    kw = dict(cols=cols, rows=rows, spacing=0.02, pins=2, jitter=1e-3, seed=cols + rows)
    run_pair(O, X, kw, S=8, I=5, substeps=3, collision=True, fast_division=fast, fused_predict=fused,
             thickness=0.016, grid_tiles=3, layout="shared_tiles")


@pytest.mark.parametrize("I", [1, 2, 10])
def test_shared_tiles_iteration_counts(O, X, I):
    """First, last and middle iterations (predict, bounds, final forms)."""
    note(f"case shared tiles I={I}")
    kw = dict(cols=64, rows=24, spacing=0.02, pins=2, jitter=1e-3, seed=5)
    run_pair(O, X, kw, S=4, I=I, substeps=2, collision=I != 2, fast_division=1, thickness=0.016, grid_tiles=3,
             layout="shared_tiles")


def test_shared_tiles_folding_sheet(O, X):
    """A 256x256 horizontal sheet folding onto itself (config-2 shape): the
    per-constraint zeros of skipped entries and the both-endpoint sign flips
    across tile edges, with contacts."""
    note("case shared tiles 256 fold")
    h = 1 / 255
    rng = np.random.default_rng(2)
    kw = dict(cols=256, rows=256, spacing=h, layout=1, height=1.0, angle=float(rng.uniform(0, 2 * np.pi)), pins=3)
    run_pair(O, X, kw, S=4, I=8, substeps=3, collision=True, thickness=2 * (0.4 * h), jacobi_omega=0.32,
             fast_division=1, grid_tiles=3, layout="shared_tiles")


def test_shared_tiles_1024(O, X):
    """The bench sheet (1024x1024, top row pinned), two substeps."""
    note("case shared tiles 1024")
    h = 1 / 1023
    O.set_threads(16)
    kw = dict(cols=1024, rows=1024, spacing=h, pins=2, jitter=1e-3, seed=3)
    run_pair(O, X, kw, S=8, I=10, substeps=2, collision=True, thickness=2 * (0.4 * h), fast_division=1,
             grid_tiles=3, layout="shared_tiles")


def test_shared_tiles_layout_selection(X):
    """grid_tiles = 3 takes the shared form where tiles apply and falls back
    to the packed layout elsewhere; 2 keeps the per-entry tiles; 1 without
    autotuning keeps the per-entry tiles."""
    note("case shared tiles layout")
    cases = [(dict(cols=64, rows=9), 3, "shared_tiles"), (dict(cols=70, rows=9), 3, "packed"),
             (dict(cols=64, rows=9), 2, "grid_tiles"), (dict(cols=64, rows=9), 1, "grid_tiles")]
    for kw, gt, want in cases:
        sim, _, _ = make(X, X.ClothMeshData.sheet(spacing=0.02, pins=2, **kw), grid_tiles=gt)
        assert sim.layout()["adjacency"] == want, (kw, gt, sim.layout())
        sim.close()
