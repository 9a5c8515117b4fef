"""The verification suite's pass/fail driver and the batch study on the GPU (quick forms): the closed
design storm conserves mass in both precisions, the suite's cases pass and write a results file that
``generate``'s check accepts, and the batch study's three stages run and measure."""
import pytest

pytestmark = pytest.mark.gpu


def test_closed_storm_and_suite_results_file(gpu, tmp_path):
    from his_amd import api
    from his_amd import benchmarks as BM
    results = BM.suite(gpu, fast=True, only=("1", "7", "8"))
    by = {r["key"]: r for r in results}
    assert set(by) == {"1", "7", "8"}
    for r in results:
        assert r["passed"], (r["key"], r["metrics"], r["notes"])
    cs = by["8"]["metrics"]
    assert cs["fp64"]["mass_err"] < 1e-12 and cs["fp32bulk"]["mass_err"] < 1e-3
    assert cs["fp64"]["rain_m3"] > 0 and cs["fp64"]["infiltration_m3"] > 0 and cs["fp64"]["stored_m3"] > 0
    assert cs["fp64"]["boundary_out_m3"] == 0.0 and cs["fp64"]["boundary_in_m3"] == 0.0     # closed
    ctx = api.context()
    h = BM.config_hash(ctx.cfg, ctx.domain)
    BM.write_results(results, tmp_path, h)
    BM.require_benchmarks(tmp_path, h)


def test_batch_study_quick(gpu):
    from his_amd import api
    from his_amd import throughput as TP
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.domain.site import domain_from
    ter = W.make_terrain(100, 81, 80.0)
    ctx = api.configure(domain_from(ter, W.chain_network(ter, 20)), SolverConfig(), gpu)
    try:
        d = TP.run_study(ctx, quick=True)
    finally:
        api.configure()
    assert [r["batch"] for r in d["scaling"]] == [1, 2]
    assert all(r["max_mass_error"] < 1e-3 and r["s_per_storm"] > 0 for r in d["scaling"])
    assert d["chosen_batch"] in (1, 2)
    f = d["full_length"]
    assert f["max_mass_error"] < 1e-3 and len(f["sequential"]) == 1 and f["worst_peak_depth_diff_m"] >= 0.0
    assert all(r["engine_graph"] > 0 and r["total"] > 0 for r in d["profile"])
    assert "batch" in TP.report(d)
