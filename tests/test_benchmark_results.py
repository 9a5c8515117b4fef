"""The benchmark-results file that hydrointel's ``generate`` demands before it writes a dataset
(hydrointel/data/generate.py:112-124): written here, accepted when every case passed and the
configuration matches, refused otherwise (no GPU needed)."""
import json

import pytest

from his_amd import benchmarks as BM


def _results(ok=True):
    return [BM._result("1", "Lake at rest", True, "c", {"x": 1.0}),
            BM._result("8", "Closed storm", ok, "c", {"fp64": {"mass_err": 1e-12}})]


def test_results_file_is_accepted_only_when_it_holds(tmp_path):
    p = BM.write_results(_results(), tmp_path, "abcd1234abcd1234", {"device": "cpu"})
    d = json.loads(p.read_text())
    assert d["provenance"]["config_hash"] == "abcd1234abcd1234" and len(d["results"]) == 2
    assert (tmp_path / "benchmark_report.md").read_text().count("PASS") == 2
    BM.require_benchmarks(tmp_path, "abcd1234abcd1234")
    with pytest.raises(RuntimeError, match="configuration"):
        BM.require_benchmarks(tmp_path, "0000000000000000")
    BM.write_results(_results(ok=False), tmp_path, "abcd1234abcd1234")
    with pytest.raises(RuntimeError, match=r"\['8'\] failed"):
        BM.require_benchmarks(tmp_path, "abcd1234abcd1234")
    with pytest.raises(RuntimeError, match="not found"):
        BM.require_benchmarks(tmp_path / "elsewhere", "abcd1234abcd1234")


def test_config_hash_follows_solver_and_domain():
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.domain.site import domain_from
    ter = W.make_terrain(100, 81, 80.0)
    dom = domain_from(ter, W.chain_network(ter, 20))
    h = BM.config_hash(SolverConfig(), dom)
    assert len(h) == 16 and h == BM.config_hash(SolverConfig(), dom)
    assert h != BM.config_hash(SolverConfig(cfl=0.4), dom)
    ter2 = W.make_terrain(100, 81, 80.0, seed=5)
    assert h != BM.config_hash(SolverConfig(), domain_from(ter2, W.chain_network(ter2, 20)))
