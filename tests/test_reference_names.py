"""Names hydrointel's callers import from its solver and engine modules, here with the same meaning
(no GPU needed):

* ``Sections.i1`` is the first moment of the wetted area about the surface, checked against a
  numerical integral of the top width below the bank, at the bank and in the slot above it;
* ``Step1DVolumes`` lives in ``solver.swe1d`` (hydrointel/solver/swe1d.py:169-173) and is the type
  ``solver.coupling`` uses;
* ``engine.build_forcing`` with hydrointel's signature is the registered resolver's storm for the same
  scenario, ``engine.BatchForcing`` stacks members like ``Forcing.stack``, and
  ``engine.scenario_landuse`` takes an ``SSPState`` as hydrointel's does;
* ``scenarios.slr_m`` refuses a year beyond the projection table, as hydrointel's does;
* ``EdgeBC`` refuses unknown kinds and value-less stage or discharge edges, and gives per-member
  values as the callable returns them; ``all_edges`` and ``edge_values`` keep the edge order.
"""
import numpy as np
import pytest
import torch
from scipy.integrate import quad

from his_amd import engine as E
from his_amd import scenarios as S
from his_amd.solver import coupling, swe1d


def test_section_first_moment_is_the_integral_of_the_top_width():
    b, m, yb, ts = 4.0, 1.5, 2.0, 3.0
    sec = swe1d.Sections(*(torch.tensor([v], dtype=torch.float64) for v in (b, m, yb, ts)))
    top = lambda s: b + 2.0 * m * s if s <= yb else ts
    for y in (0.0, 0.3, 1.2, yb, 2.5, 4.75):
        want = quad(lambda s: (y - s) * top(s), 0.0, y, points=[yb] if y > yb else None, epsabs=0, epsrel=1e-13)[0]
        got = float(sec.i1(torch.tensor([y], dtype=torch.float64))[0])
        assert got == pytest.approx(want, rel=1e-12, abs=1e-14), y
    # the area is the derivative of the first moment
    y, e = torch.tensor([1.1, 3.3], dtype=torch.float64), 1e-6
    assert torch.allclose((sec.i1(y + e) - sec.i1(y - e)) / (2 * e), sec.area(y), rtol=1e-8, atol=0)


def test_step_volumes_type_is_shared():
    assert coupling.Step1DVolumes is swe1d.Step1DVolumes
    v = swe1d.Step1DVolumes(torch.zeros(2), torch.ones(2), torch.zeros(2))
    assert v.bnd_out.sum() == 2


def test_build_forcing_is_the_resolved_scenario():
    fo = E.build_forcing(None, 50, "RCP8.5", "SSP3", 2050, True, 1.5, 52.0)
    ref = S.resolve(S.Scenario(50, "RCP8.5", "SSP3", 2050, True, 1.5))
    t = np.linspace(0.0, ref.t_end, 97)
    assert fo.t_end == ref.t_end and fo.storm_start == ref.storm_start and fo.ssp == "SSP3"
    for name in ("rain", "inflow", "tide"):
        assert np.array_equal(np.array([getattr(fo, name)(x) for x in t]), np.array([getattr(ref, name)(x) for x in t]))
    with pytest.raises(ValueError):
        E.build_forcing(None, 20, "RCP8.5", "SSP3", 2050, True, 0.0, 52.0)
    two = E.BatchForcing([fo, E.build_forcing(None, 10, "RCP2.6", "SSP1", 2030, False, 0.0, 52.0)])
    assert two.rain(0.5 * fo.t_end).shape == (2,) and two.rain(0.5 * fo.t_end)[0] == fo.rain(0.5 * fo.t_end)[0]


def test_scenario_landuse_takes_a_pathway_state():
    from his_amd import workloads as W
    from his_amd.api import grown_landuse
    from his_amd.domain.site import domain_from
    ter = W.make_terrain(100, 81, 80.0)
    dom = domain_from(ter, W.chain_network(ter, 20))
    a = E.scenario_landuse(dom, S.ssp_state("SSP5"), 2070)
    assert np.array_equal(a, grown_landuse(dom, "SSP5", 2070)) and not np.array_equal(a, dom.landuse)
    # the growth starts at hydrointel's base year
    assert np.array_equal(E.scenario_landuse(dom, S.ssp_state("SSP5"), 2025), dom.landuse)


def test_sea_level_rise_refuses_years_past_the_table():
    assert S.slr_m("RCP4.5", 2100) == S.sea_level_rise("RCP4.5", 2100)
    with pytest.raises(ValueError):
        S.slr_m("RCP4.5", 2101)
    with pytest.raises(ValueError):
        S.slr_m("RCP7.0", 2050)


def test_edge_conditions():
    from his_amd.solver.boundaries import EDGES, EdgeBC, all_edges, edge_values
    with pytest.raises(ValueError):
        EdgeBC("wall")
    with pytest.raises(ValueError):
        EdgeBC("stage")
    tide = EdgeBC("stage", lambda t: 0.5 + t)
    assert tide.at(2.0) == 2.5 and isinstance(tide.at(2.0), float)
    per_member = EdgeBC("discharge", lambda t: np.array([1.0, 2.0]) * t)
    assert np.array_equal(per_member.at(3.0), [3.0, 6.0])
    bcs = all_edges("closed", east=EdgeBC("stage", 1.25), south="transmissive")
    assert [bcs[e].kind for e in EDGES] == ["closed", "stage", "transmissive", "closed"]
    assert edge_values(bcs, 0.0) == [0.0, 1.25, 0.0, 0.0]
