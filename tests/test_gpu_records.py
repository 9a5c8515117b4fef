"""Records written from real runs: two storms in one batch through ``write_storms``, read back
and checked against the run's own results; the index; resuming; the compressed format."""
import json

import numpy as np
import pytest
import torch

from his_amd import api
from his_amd import records as R

pytestmark = pytest.mark.gpu


def _case():
    from his_amd import workloads as W
    from his_amd.engine import Forcing
    wl = W.plain_cpu()
    ter, net = wl.terrain, wl.network
    wl.members = [wl.members[0], W.Member(rain_peak_mmh=60.0, tide_phase=0.7, inflow_peak=150.0)]
    rain, tide, inflow = wl.forcing(batched=True)
    site = api.baseline_site(ter, net)
    return ter, net, Forcing(rain, inflow, tide, wl.storm.t_end), site, site.replace(storage_depth=site.storage_depth * 0.0)


def test_write_storms_and_read_back(gpu, tmp_path):
    ter, net, fo, site, wetter = _case()
    storms = [(0, site, {"name": "baseline"}), (1, wetter, {"name": "no retention"})]
    index = R.write_storms(tmp_path, storms, fo, ter, net, f=2, device=gpu, forcing_steps=24)
    assert set(index) == {"0", "1"}
    results = api.simulate_batch([site, wetter], fo, ter, net, device=gpu)
    nt = len(fo.times(900.0))
    for i, res in enumerate(results):
        rec = R.load_record(R.record_path(tmp_path, i))
        assert rec["meta"] == storms[i][2] and rec["store_coarsen"] == 2
        assert tuple(rec["h"].shape) == (nt, ter.ny // 2, ter.nx // 2)
        # the batch is deterministic: the record holds this run's own arrays
        assert torch.equal(rec["depth_max_full"], res.depth_max.float())
        assert torch.allclose(rec["h"], torch.as_tensor(R.coarsen_mean(res.depth_series.numpy(), 2), dtype=torch.float32))
        assert torch.equal(rec["eta1"], res.channel_eta) and torch.equal(rec["q1"], res.channel_q)
        L = rec["ledger"]
        assert all(L[k].shape == (nt,) for k in R.LEDGER_KEYS + ("v2d_coarse",))
        # the ledger's surface volume is the snapshot's: depth summed over the grid
        v2d = res.depth_series.double().sum(dim=(1, 2)).numpy() * ter.dx * ter.dx
        assert np.allclose(L["v2d"].numpy(), v2d, rtol=1e-5)
        assert np.all(L["v2d_coarse"].numpy() <= L["v2d"].numpy() * (1 + 1e-6))
        assert abs(rec["mass_balance_error"] - res.mass_balance_error) < 1e-12 and rec["mass_balance_error"] < 1e-3
        assert len(rec["volumes"]["series"]) > 0 and rec["forcing"]["rain_mmh"].shape == (24,)
        e = index[str(i)]
        assert e["record"] == f"sim_{i:05d}.pt" and e["batch_size"] == 2 and e["name"] == storms[i][2]["name"]
        land = ~ter.sea
        assert abs(e["peak_depth_land"] - float(res.depth_max.numpy()[land].max())) < 1e-6
    assert json.loads((tmp_path / "index.json").read_text()) == index
    # resuming runs nothing and leaves the records as they are
    stamp = R.record_path(tmp_path, 0).stat().st_mtime_ns
    assert R.write_storms(tmp_path, storms, fo, ter, net, f=2, device=gpu) == index
    assert R.record_path(tmp_path, 0).stat().st_mtime_ns == stamp
    # a storm added later runs alone and may be stored compressed
    from his_amd.forcing import select
    from his_amd.engine import Forcing
    fo1 = Forcing(select(fo.rain, [1]), select(fo.inflow, [1]), select(fo.tide, [1]), fo.t_end)
    index = R.write_storms(tmp_path, [(2, wetter, {"name": "again"})], fo1, ter, net, f=1, compress=True, device=gpu)
    rec2 = R.load_record(R.record_path(tmp_path, 2))
    assert R.record_path(tmp_path, 2).name == "sim_00002.pt.gz" and index["2"]["batch_size"] == 1
    assert tuple(rec2["h"].shape) == (nt, ter.ny, ter.nx)
