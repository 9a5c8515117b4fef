"""The agreement metric on hand-built depth series with known answers."""
import numpy as np
import pytest

from his_amd.criteria import LIMITS, agreement


def _series():
    rng = np.random.default_rng(0)
    h = np.clip(rng.normal(0.3, 0.2, (5, 12, 10)), 0, None)
    land = np.ones((12, 10), bool)
    land[:, -2:] = False
    return h, land


def test_identical_runs_agree_exactly():
    h, land = _series()
    r = agreement(h, h.copy(), land, 20.0, 1e-3)
    assert r["passed"] and r["depth_rmse_wet_m"] == 0.0 and r["flood_volume_rel"] == 0.0


def test_uniform_offset_gives_closed_form_metrics():
    h, land = _series()
    hb = h + 0.004 * (h > 1e-3)            # +4 mm on every wet sample
    r = agreement(h, hb, land, 20.0, 1e-3)
    assert r["depth_rmse_wet_m"] == pytest.approx(0.004)
    assert r["depth_p99_abs_m"] == pytest.approx(0.004)
    assert not r["checks"]["depth_rmse_wet_m"]["passed"] and r["checks"]["depth_p99_abs_m"]["passed"]


def test_one_front_cell_only_moves_the_unjudged_maximum():
    h, land = _series()
    hb = h.copy()
    hb[2, 3, 4] = h[2, 3, 4] + 0.04          # one chaotic cell, 40 mm
    r = agreement(h, hb, land, 20.0, 1e-3)
    assert r["depth_max_abs_any_m"] == pytest.approx(0.04)
    assert r["checks"]["depth_rmse_wet_m"]["passed"]


def test_reach_discharge_and_empty_comparison():
    h, land = _series()
    q = np.abs(np.random.default_rng(1).normal(10, 3, (5, 6)))
    reach = np.array([0, 0, 1, 1, 2, 2])
    qb = q.copy()
    qb[:, 4:] *= 1.03
    r = agreement(h, h, land, 20.0, 1e-3, q, qb, reach)
    assert r["reach_peak_discharge"][2]["rel"] == pytest.approx(0.03)
    assert not r["passed"] and LIMITS["reach_peak_discharge_rel"] == 0.02
    with pytest.raises(ValueError):
        agreement(np.zeros((2, 3, 3)), np.zeros((2, 3, 3)), np.ones((3, 3), bool), 20.0, 1e-3)
