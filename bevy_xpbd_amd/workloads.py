"""Seeded synthetic workloads of the measured configurations.

The reference has no workload generator (its example builds one hanging grid,
examples/cloth.rs); these are the benchmark configurations of this project.
All randomness comes from one SplitMix64 stream per workload, with floats
drawn as (u32 >> 8) * 2^-24, the same rule the sheet builder uses for its
jitter (xpbd_sheet_desc.seed).
"""
from __future__ import annotations

import math

from .api import (LAYOUT_HANGING_XY, LAYOUT_HORIZONTAL_XZ, PIN_NONE, PIN_TOP_ROW, PIN_TWO_CORNERS,
                  SheetDesc)

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def unit(self) -> float:
        return ((self.next() >> 32) >> 8) * (1.0 / (1 << 24))

    def below(self, n: int) -> int:
        """Uniform integer in [0, n) (n small: the modulo bias is below 2^-32)."""
        return self.next() % n


PIN_MODES = ("two_corners", "top_row", "none_ground")


def batch_descs(count: int = 512, seed: int = 4, spacing: float = 0.01, side_lo: int = 16, side_hi: int = 192,
                comp_lo: float = 1e-8, comp_hi: float = 1e-3, jitter: float = 1e-3):
    """The uneven-cloth batch: `count` cloths, each side drawn uniformly from
    {side_lo..side_hi} vertices (independently per axis), stretch / shear /
    bend compliance each log-uniform in [comp_lo, comp_hi], pinning drawn from
    {two top corners, full top row, none + ground plane y = 0}.  Hanging
    cloths start vertical (top row at y = rows*spacing + 0.5); ground cloths
    start horizontal at y = 0.3 m, rotated about +Y by a drawn angle.  All
    cloths share the rest spacing, so one contact distance (0.8 x spacing)
    fits every cloth.  Returns (descs, modes)."""
    rng = SplitMix64(seed)
    la, lb = math.log10(comp_lo), math.log10(comp_hi)
    descs, modes = [], []
    for _ in range(count):
        cols = side_lo + rng.below(side_hi - side_lo + 1)
        rows = side_lo + rng.below(side_hi - side_lo + 1)
        comps = [10.0 ** (la + (lb - la) * rng.unit()) for _ in range(3)]
        mode = rng.below(3)
        angle = 2.0 * math.pi * rng.unit()
        cloth_seed = rng.next()
        if mode == 2:
            d = SheetDesc.make(cols=cols, rows=rows, spacing=spacing, layout=LAYOUT_HORIZONTAL_XZ, height=0.3,
                               angle=angle, pins=PIN_NONE, ground=1)
        else:
            d = SheetDesc.make(cols=cols, rows=rows, spacing=spacing, layout=LAYOUT_HANGING_XY, height=0.5,
                               pins=PIN_TWO_CORNERS if mode == 0 else PIN_TOP_ROW, ground=0)
        d.stretch_compliance, d.shear_compliance, d.bend_compliance = comps
        d.jitter = jitter
        d.seed = cloth_seed
        descs.append(d)
        modes.append(PIN_MODES[mode])
    return descs, modes
