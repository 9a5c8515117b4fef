"""The dataset record format (hydrointel/data/generate.py:37-63, 104-110, 355-425) on the CPU:
block means, the rain resampling, file naming, and a record built from a small FloodResult
against brute-force recomputation of every derived array."""
import numpy as np
import pytest
import torch

from his_amd import records as R
from his_amd.api import FloodResult, SiteState
from his_amd.engine import Forcing
from his_amd.forcing import PiecewiseConstant, PiecewiseLinear, sine_tide
from his_amd.solver.coupling import RunLog

MMH = 1.0 / 3.6e6


@pytest.mark.parametrize("f", [1, 2, 3])
def test_coarsen_mean_is_the_block_mean(f):
    rng = np.random.default_rng(3)
    a = rng.random((2, 7, 9))
    c = R.coarsen_mean(a, f)
    cy, cx = 7 // f, 9 // f
    assert c.shape == (2, cy, cx)
    for k in range(2):
        for j in range(cy):
            for i in range(cx):
                want = a[k, j * f:(j + 1) * f, i * f:(i + 1) * f].mean()
                assert abs(c[k, j, i] - want) < 1e-14


def test_rain_resampled_as_block_means():
    rain = PiecewiseConstant(0.0, 600.0, [[10 * MMH, 20 * MMH, 0.0, 40 * MMH], [1 * MMH, 2 * MMH, 3 * MMH, 4 * MMH]])
    assert np.allclose(R.resampled_mmh(rain, 2400.0, 4, 0), [10, 20, 0, 40], atol=1e-9)
    assert np.allclose(R.resampled_mmh(rain, 2400.0, 2, 0), [15, 20], atol=1e-9)
    assert np.allclose(R.resampled_mmh(rain, 2400.0, 4, 1), [1, 2, 3, 4], atol=1e-9)
    # 800 s steps over 600 s blocks: step 0 holds 15 samples of 10 and 5 of 20, step 1 ten of 20
    # and ten of 0, step 2 five of 0 and fifteen of 40, step 3 lies after the storm
    assert np.allclose(R.resampled_mmh(rain, 3200.0, 4, 0), [12.5, 10.0, 30.0, 0.0], atol=1e-9)


def test_record_paths_and_round_trip(tmp_path):
    assert R.record_path(tmp_path, 7, False).name == "sim_00007.pt"
    assert R.record_path(tmp_path, 7, True).name == "sim_00007.pt.gz"
    assert R.record_path(tmp_path, 7).name == "sim_00007.pt"           # neither exists: plain
    rec = {"a": torch.arange(5, dtype=torch.float32), "n": {"x": 1.5, "k": 3, "s": "text", "ok": True},
           "series": [{"t": 0.0, "v": 2.0}, {"t": 1.0, "v": 3.0}]}
    R.save_record(rec, R.record_path(tmp_path, 7, True))
    assert R.record_path(tmp_path, 7).name == "sim_00007.pt.gz"        # only the compressed one exists
    back = R.load_record(R.record_path(tmp_path, 7))
    assert torch.equal(back["a"], rec["a"]) and back["n"] == rec["n"] and back["series"] == rec["series"]
    R.save_record(rec, R.record_path(tmp_path, 7, False))
    assert R.record_path(tmp_path, 7).name == "sim_00007.pt"           # plain wins when both exist
    assert torch.equal(R.load_record(R.record_path(tmp_path, 7))["a"], rec["a"])


def _result(nt=3, ny=5, nx=7, nc=4, batched=False):
    rng = np.random.default_rng(11)
    h = np.clip(rng.random((nt, ny, nx)) - 0.3, 0.0, None).astype(np.float32)
    u = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    v = rng.standard_normal((nt, ny, nx)).astype(np.float32)
    B = 2 if batched else 1
    fo = Forcing(PiecewiseConstant(0.0, 600.0, [[10 * MMH, 30 * MMH]] * B), PiecewiseLinear(0.0, 600.0, [[5.0, 9.0]] * B),
                 sine_tide(0.8, 44712.0, [0.3] * B), 1200.0)
    ledger = [{k: float(10 * t + j) for j, k in enumerate(R.LEDGER_KEYS)} for t in range(nt)]
    series = [{"t": 600.0 * t, "volume": 1.0 + t, "rel_error": 1e-7} for t in range(nt)]
    sources = {"rain": 3.0, "infiltration": 1.0}
    extras = {"forcing": fo, "run_log": RunLog(steps=12, wall_s=0.5, dt_min=0.2, dt_max=0.9), "ledger": ledger,
              "mass_series": series, "sources": sources}
    if batched:
        extras["batch_index"] = 1
    res = FloodResult(depth_max=torch.as_tensor(h.max(0)), depth_series=torch.as_tensor(h), u=torch.as_tensor(u),
                      v=torch.as_tensor(v), channel_eta=torch.zeros(nt, nc), channel_q=torch.ones(nt, nc),
                      times=torch.as_tensor(np.arange(nt) * 600.0), infiltrated_volume_m3=1.0, stored_volume_m3=2.0,
                      mass_balance_error=1e-8, provenance={}, extras=extras)
    site = SiteState(torch.full((ny, nx), 0.005), torch.ones(ny, nx), torch.full((ny, nx), 0.035), torch.ones(3))
    return res, site, h, u, v


@pytest.mark.parametrize("batched", [False, True])
def test_simulation_record_against_brute_force(tmp_path, batched):
    res, site, h, u, v = _result(batched=batched)
    f, dx = 2, 20.0
    rec = R.simulation_record(res, f, dx, site=site, meta={"rp": 10}, forcing_steps=8, wall_s=3.0,
                              forcing_factors={"rain_factor": 1.2})
    nt = h.shape[0]
    assert tuple(rec["h"].shape) == (nt, 2, 3) and rec["store_coarsen"] == 2 and rec["meta"] == {"rp": 10}
    assert torch.equal(rec["depth_max_full"], torch.as_tensor(h.max(0)))
    for t in range(nt):
        for j in range(2):
            for i in range(3):
                hb, ub = h[t, 2 * j:2 * j + 2, 2 * i:2 * i + 2], u[t, 2 * j:2 * j + 2, 2 * i:2 * i + 2]
                assert abs(float(rec["h"][t, j, i]) - hb.mean()) < 1e-6
                want_u = (hb * ub).mean() / hb.mean() if hb.mean() > 0 else 0.0
                assert abs(float(rec["u"][t, j, i]) - want_u) < 1e-5
    L = rec["ledger"]
    assert torch.allclose(L["v2d"], torch.tensor([0.0, 10.0, 20.0], dtype=torch.float64))
    want_v2d_coarse = h[:, :4, :6].astype(np.float64).sum(axis=(1, 2)) * dx * dx
    assert np.allclose(L["v2d_coarse"].numpy(), want_v2d_coarse)
    fz = rec["forcing"]
    assert fz["rain_mmh"].shape == (8,) and abs(float(fz["rain_mmh"][0]) - 10.0) < 1e-4
    assert abs(float(fz["rain_mmh"][-1]) - 30.0) < 1e-4 and fz["rain_factor"] == 1.2 and fz["t_end"] == 1200.0
    assert abs(float(fz["inflow_m3s"][-1]) - 9.0) < 1e-5 and fz["tide_m"].shape == (8,)
    assert rec["run_log"]["steps"] == 12 and "series" not in rec["run_log"]
    assert rec["volumes"]["series"][1]["volume"] == 2.0
    path = R.record_path(tmp_path, 0, True)
    R.save_record(rec, path)
    back = R.load_record(path)
    assert torch.equal(back["h"], rec["h"]) and back["ledger"].keys() == rec["ledger"].keys()
