"""Host-side ledger and channel-network rules, checked against closed forms (no GPU needed).

* MassBalance: a closed box that gains rain R and loses infiltration I balances exactly
  (error 0); a volume off by d gives error d / (R + I) (the gross sources), and with no sources
  d / (1e-9 |V0| + 1); clip volume is kept but enters neither sum; per-member assertion names the
  failing members only; the device engine's ledger rows fold in as rain, infiltration, boundary
  terms and the two clips together.
* Reach geometry: the cell lengths of a straight reach with uneven node spacing, and the chainage,
  against hand values; a network breaking each validation rule raises ValueError.
* apply_gamma: a rectangular reach's bankfull area scales by gamma; a junction at which two
  reaches end is rescaled by both factors.
"""
import math

import numpy as np
import pytest
import torch

from his_amd.domain.channels import ChannelNetwork, Junction, Reach, apply_gamma
from his_amd.solver.massbalance import TERMS, MassBalance


def test_ledger_balances_a_closed_box_and_measures_a_discrepancy():
    v0 = np.array([1000.0, 50.0])
    mb = MassBalance("cpu", v0)
    for _ in range(4):
        mb.add("rain", torch.tensor([25.0, 10.0], dtype=torch.float64))
        mb.add("infiltration", np.array([5.0, 2.0]))
        mb.add("clip", 3.0)
    now = v0 + np.array([80.0, 32.0])
    assert np.array_equal(mb.check(now, 1.0), [0.0, 0.0])
    err = mb.check(now + np.array([12.0, -6.0]), 2.0)
    assert err == pytest.approx([12.0 / 120.0, 6.0 / 48.0], rel=1e-15)
    src = mb.sources()
    assert src["clip"].tolist() == [12.0, 12.0] and src["rain"].tolist() == [100.0, 40.0]
    assert len(mb.series) == 2 and mb.series[1]["rel_error"].tolist() == pytest.approx([0.1, 0.125])
    with pytest.raises(AssertionError, match=r"member 1 error 1\.250e-01") as ei:
        mb.assert_ok(now + np.array([0.0, -6.0]), 3.0, 0.05)
    assert "member 0" not in str(ei.value)


def test_ledger_without_sources_uses_the_volume_floor():
    mb = MassBalance(None, 2.0e6)
    assert mb.batch is None
    err = mb.check(2.0e6 + 0.25, 0.0)
    assert isinstance(err, float) and err == pytest.approx(0.25 / (1e-9 * 2.0e6 + 1.0), rel=1e-15)
    with pytest.raises(KeyError):
        mb.add("evaporation", 1.0)
    with pytest.raises(ValueError):
        MassBalance(None, np.zeros(3)).add("rain", np.ones(2))


def test_engine_rows_fold_into_the_terms():
    mb = MassBalance("cpu", np.zeros(2))
    acc = torch.arange(20, dtype=torch.float64).reshape(10, 2)
    mb.fold_engine_rows(acc)
    src = mb.sources()
    for k, name in enumerate(TERMS[:6]):
        assert src[name].tolist() == acc[k].tolist()
    assert src["clip"].tolist() == (acc[6] + acc[7]).tolist()


def _reach(rid, xs, up=("inflow",), down=("stage",), width=10.0, bank=2.0):
    n = len(xs)
    return Reach(rid, np.asarray(xs, float), np.zeros(n), np.zeros(n), np.full(n, width), np.zeros(n),
                 np.full(n, 0.03), np.full(n, bank), up, down)


def test_reach_cell_lengths_and_chainage():
    r = _reach("a", [0.0, 10.0, 30.0, 60.0])
    assert r.spacing.tolist() == [10.0, 15.0, 25.0, 30.0]
    assert r.chainage.tolist() == [0.0, 10.0, 30.0, 60.0]
    assert _reach("b", [5.0]).spacing.tolist() == [1.0]


@pytest.mark.parametrize("breaker, needle", [
    (lambda r: setattr(r, "width", np.zeros(r.n)), "positive"),
    (lambda r: setattr(r, "manning", np.full(r.n, np.nan)), "finite"),
    (lambda r: setattr(r, "bed", np.zeros(r.n + 1)), "nodes"),
    (lambda r: setattr(r, "z_bank", np.full(r.n, 0.1)), "bank crest"),
    (lambda r: setattr(r, "up", ("weir",)), "end kind"),
    (lambda r: setattr(r, "down", ("junction", 3)), "junction 3"),
])
def test_validation_rules(breaker, needle):
    net = ChannelNetwork([_reach("a", [0.0, 10.0, 20.0])])
    net.validate()
    breaker(net.reaches[0])
    with pytest.raises(ValueError, match=needle):
        net.validate()


def test_gamma_scales_bankfull_area_and_rescales_shared_junctions():
    j = Junction(50.0, 0.0, 0.0, 10.0, 0.0, 0.03, 2.0, 20.0)
    net = ChannelNetwork([_reach("a", [0.0, 10.0, 20.0], down=("junction", 0)),
                          _reach("b", [50.0, 50.0, 50.0], down=("junction", 0)),
                          _reach("c", [60.0, 70.0], up=("junction", 0))], [j])
    g = np.array([1.44, 0.81, 2.0])
    out = apply_gamma(net, g)
    for r0, r1, gi in zip(net.reaches, out.reaches, g):
        area0 = r0.width * (r0.z_bank - r0.bed)
        area1 = r1.width * (r1.z_bank - r1.bed)
        assert area1 == pytest.approx(gi * area0, rel=1e-14)
        assert np.array_equal(r1.z_bank, r0.z_bank)
    jo = out.junctions[0]
    s = math.sqrt(1.44) * math.sqrt(0.81)
    assert jo.width == pytest.approx(10.0 * s, rel=1e-15)
    assert jo.z_bank - jo.bed == pytest.approx(2.0 * s, rel=1e-15)
    assert net.reaches[0].width[0] == 10.0 and net.junctions[0].width == 10.0      # input untouched
    with pytest.raises(ValueError):
        apply_gamma(net, [1.0, 1.0])
