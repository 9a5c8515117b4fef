"""Records written from real runs: two storms in one batch through ``write_storms`` -- which calls
``api.simulate_batch(sites, scenarios)`` as generate.py:288 does -- read back and checked against the
run's own results; the index; resuming; the compressed format."""
import json

import numpy as np
import pytest
import torch

from his_amd import api
from his_amd import records as R
from his_amd import scenarios as S

pytestmark = pytest.mark.gpu


def _case(gpu):
    from his_amd import workloads as W
    from his_amd.domain.site import domain_from
    from his_amd.scenarios import Scenario
    ter = W.make_terrain(100, 81, 80.0)
    ctx = api.configure(domain_from(ter, W.chain_network(ter, 20)), device=gpu)
    sc = [Scenario(25, "RCP4.5", "SSP2", 2040), Scenario(100, "RCP8.5", "SSP3", 2070, True, -2.0)]
    site = api.baseline_site(sc[0])
    return ctx.domain, sc, site, site.replace(storage_depth=site.storage_depth * 0.0)


def test_write_storms_and_read_back(gpu, tmp_path):
    dom, sc, site, wetter = _case(gpu)
    try:
        storms = [(0, site, sc[0], {"name": "baseline"}), (1, wetter, sc[1], {"name": "no retention"})]
        index = R.write_storms(tmp_path, storms, f=2, forcing_steps=24)
        assert set(index) == {"0", "1"}
        results = api.simulate_batch([site, wetter], sc)
        nt = len(results[0].extras["forcing"].times(900.0))
        for i, res in enumerate(results):
            rec = R.load_record(R.record_path(tmp_path, i))
            assert rec["meta"]["name"] == storms[i][3]["name"] and rec["meta"]["rcp"] == sc[i].rcp
            assert rec["store_coarsen"] == 2
            assert tuple(rec["h"].shape) == (nt, dom.ny // 2, dom.nx // 2)
            # the batch is deterministic: the record holds this run's own arrays
            assert torch.equal(rec["depth_max_full"], res.depth_max.float())
            assert torch.allclose(rec["h"], torch.as_tensor(R.coarsen_mean(res.depth_series.numpy(), 2),
                                                            dtype=torch.float32))
            assert torch.equal(rec["eta1"], res.channel_eta) and torch.equal(rec["q1"], res.channel_q)
            L = rec["ledger"]
            assert all(L[k].shape == (nt,) for k in R.LEDGER_KEYS + ("v2d_coarse",))
            v2d = res.depth_series.double().sum(dim=(1, 2)).numpy() * dom.dx * dom.dx
            assert np.allclose(L["v2d"].numpy(), v2d, rtol=1e-5)
            assert np.all(L["v2d_coarse"].numpy() <= L["v2d"].numpy() * (1 + 1e-6))
            assert abs(rec["mass_balance_error"] - res.mass_balance_error) < 1e-12 and rec["mass_balance_error"] < 1e-3
            assert len(rec["volumes"]["series"]) > 0 and rec["forcing"]["rain_mmh"].shape == (24,)
            assert rec["fields"]["landuse_frac"].shape == (8, dom.ny // 2, dom.nx // 2)
            assert torch.allclose(rec["fields"]["landuse_frac"].sum(0), torch.ones(dom.ny // 2, dom.nx // 2))
            assert rec["forcing"]["horton_rate_mmh"].shape == (nt, dom.ny // 2, dom.nx // 2)
            # this package's design storm applies no areal reduction; hydrointel's would carry its own
            assert rec["forcing"]["arf"] == 1.0 and rec["forcing"]["storm_start"] == 0.0
            e = index[str(i)]
            assert e["record"] == f"sim_{i:05d}.pt" and e["batch_size"] == 2 and e["name"] == storms[i][3]["name"]
            land = ~dom.sea
            assert abs(e["peak_depth_land"] - float(res.depth_max.numpy()[land].max())) < 1e-6
        assert json.loads((tmp_path / "index.json").read_text()) == index
        # resuming runs nothing and leaves the records as they are
        stamp = R.record_path(tmp_path, 0).stat().st_mtime_ns
        assert R.write_storms(tmp_path, storms, f=2) == index
        assert R.record_path(tmp_path, 0).stat().st_mtime_ns == stamp
        # a storm added later runs alone and may be stored compressed
        index = R.write_storms(tmp_path, [(2, wetter, sc[1], {"name": "again"})], f=1, compress=True)
        rec2 = R.load_record(R.record_path(tmp_path, 2))
        assert R.record_path(tmp_path, 2).name == "sim_00002.pt.gz" and index["2"]["batch_size"] == 1
        assert tuple(rec2["h"].shape) == (nt, dom.ny, dom.nx)
        # a resolver whose forcing records its design factors, as hydrointel's does, has them written
        # (generate.py:383-384) in place of the pathway values
        design = S.DesignStorm()

        def with_factors(s, domain=None):
            fo = design(s, domain)
            fo.rain_factor, fo.slr_m, fo.arf_domain = 1.25, 0.5, 0.93
            return fo
        S.register_resolver(with_factors)
        R.write_storms(tmp_path, [(3, site, sc[0], {"name": "factors"})], f=2)
        got = R.load_record(R.record_path(tmp_path, 3))["forcing"]
        assert (got["rain_factor"], got["slr_m"], got["arf"]) == (1.25, 0.5, 0.93)
    finally:
        S.register_resolver(S.DesignStorm())
        api.configure()
