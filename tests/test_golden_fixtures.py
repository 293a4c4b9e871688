"""The oracle against its frozen fixtures (tests/golden/, made by
tests/golden/make_fixtures.py).  The reference holds no stored vectors; these
pin the oracle against unnoticed change of its own arithmetic."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures  # noqa: E402


def test_oracle_config1_trajectory_matches_fixture(O):
    want = json.load(open(os.path.join(HERE, "golden", "oracle_config1.json")))
    got = make_fixtures.fixtures(O)
    for f, w in want["frames"].items():
        g = got["frames"][f]
        assert g["f32_sha256"] == w["f32_sha256"], f"frame {f}"
        assert g["f32_sample"] == w["f32_sample"], f"frame {f}"
        assert g["max_abs_f32_vs_f64"] == w["max_abs_f32_vs_f64"], f"frame {f}"
    # the stated trajectory tolerance of the f32 solver against f64 on this sheet
    assert want["frames"]["120"]["max_abs_f32_vs_f64"] < 1e-4


def test_oracle_sheet1024_frame5_matches_fixture(O):
    """The first frame of the frozen 1024x1024 trajectory (frame 5 of 65;
    about 16 s on 8 cores): positions, velocities and residual."""
    want = json.load(open(os.path.join(HERE, "golden", "oracle_sheet1024.json")))
    got = make_fixtures.sheet1024_fixtures(O, os.cpu_count() or 1, frames=(5,))
    assert got["thickness"] == want["thickness"]
    g, w = got["frames"]["5"], want["frames"]["5"]
    assert g["f32_sha256"] == w["f32_sha256"] and g["f32_vel_sha256"] == w["f32_vel_sha256"]
    assert g["f32_sample"] == w["f32_sample"] and g["residual"] == w["residual"]
