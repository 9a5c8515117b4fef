"""Land use, scenarios and the run-level API's host logic (no GPU needed).

* Horton and roughness fields of a mixed-class raster equal the class table's values, the capacities
  scaled by kappa and the decay not (hand-written expected values per class), also as
  ``Horton.from_landuse`` builds them, whose rate decays from its clock start;
* scenario validation raises ValueError like hydrointel's; the default design storm peaks at the
  return period's intensity times the RCP's rain factor at mid-storm, has high water at the requested
  offset, and holds the sea level of the RCP and horizon year when the tide is off;
* batched forcing must share a time base;
* land use grows with the SSP and the horizon year, deterministically, and only onto convertible land;
* ``validate_site`` rejects wrong shapes and out-of-range fields.
"""
import math

import numpy as np
import pytest
import torch

from his_amd import scenarios as S
from his_amd.domain import landuse as LU
from his_amd.engine import Forcing


def test_fields_of_a_mixed_class_raster():
    raster = np.array([[LU.WATER, LU.ROAD, LU.DENSE_URBAN, LU.RESIDENTIAL],
                       [LU.BARE, LU.GRASS, LU.CROPLAND, LU.MANGROVE]])
    kappa = np.array([[1.0, 2.0, 3.0, 1.5], [1.0, 1.2, 2.0, 4.0]])
    f0, fc, k = LU.horton_fields(raster, kappa)
    mmh = 1.0 / 3.6e6
    want_f0 = np.array([[0.0, 0.0, 2.0 * 3.0, 10.0 * 1.5], [40.0, 60.0 * 1.2, 30.0 * 2.0, 20.0 * 4.0]]) * mmh
    want_fc = np.array([[0.0, 0.0, 0.5 * 3.0, 2.0 * 1.5], [8.0, 12.0 * 1.2, 6.0 * 2.0, 5.0 * 4.0]]) * mmh
    want_k = np.array([[0.0, 0.0, 4.0, 4.0], [3.0, 3.0, 2.0, 2.0]]) / 3600.0
    assert np.allclose(f0, want_f0, rtol=1e-15, atol=0) and np.allclose(fc, want_fc, rtol=1e-15, atol=0)
    assert np.allclose(k, want_k, rtol=1e-15, atol=0)
    from his_amd.solver.infiltration import Horton
    hz = Horton.from_landuse(raster, torch.as_tensor(kappa), "cpu", torch.float64, t0=600.0)
    assert np.array_equal(hz.f0.numpy(), f0) and np.array_equal(hz.fc.numpy(), fc) and np.array_equal(hz.k.numpy(), k)
    assert hz.t0 == 600.0 and torch.equal(hz.rate(600.0), hz.f0) and torch.equal(hz.rate(0.0), hz.f0)
    late = hz.rate(600.0 + 3600.0).numpy()
    assert np.allclose(late, want_fc + (want_f0 - want_fc) * np.exp(-want_k * 3600.0), rtol=1e-14, atol=0)
    assert LU.lookup(raster, "manning").tolist() == [[0.030, 0.014, 0.050, 0.045], [0.025, 0.035, 0.040, 0.120]]
    assert LU.lookup(raster, "storage_m")[1, 2] == 0.05
    with pytest.raises(ValueError):
        LU.lookup(np.array([0, 8]), "manning")


@pytest.mark.parametrize("bad", [S.Scenario(20, "RCP4.5", "SSP2", 2050), S.Scenario(10, "RCP7.0", "SSP2", 2050),
                                 S.Scenario(10, "RCP4.5", "SSP9", 2050), S.Scenario(10, "RCP4.5", "SSP2", 2101)])
def test_scenario_validation(bad):
    with pytest.raises(ValueError):
        S.validate(bad)


def test_design_storm():
    st = S.DesignStorm()
    sc = S.Scenario(50, "RCP8.5", "SSP3", 2050, True, 1.5)
    fo = st(sc)
    mid = 0.5 * st.storm_s
    assert fo.rain(mid)[0] * 3.6e6 == pytest.approx(80.0 * 1.22, rel=1e-12)
    assert fo.rain(0.0)[0] == 0.0 and fo.rain(st.storm_s + 10)[0] == 0.0
    assert fo.inflow(mid)[0] == pytest.approx(300.0) and fo.inflow(0.0)[0] == pytest.approx(10.0)
    slr = 0.28
    t_high = mid + 1.5 * 3600.0
    assert fo.tide(t_high)[0] == pytest.approx(slr + 0.8, abs=1e-12)
    assert fo.tide(t_high + 0.5 * st.tide_period_s)[0] == pytest.approx(slr - 0.8, abs=1e-12)
    assert fo.t_end == st.storm_s + st.recession_s and fo.ssp == "SSP3"
    off = st(S.Scenario(50, "RCP8.5", "SSP3", 2100, False))
    assert off.tide(1234.0)[0] == pytest.approx(0.85) and off.tide(9999.0)[0] == pytest.approx(0.85)
    assert S.sea_level_rise("RCP2.6", 2035) == pytest.approx(0.09) and S.sea_level_rise("RCP2.6", 2020) == 0.0


def test_batched_forcing_shares_a_time_base():
    a = S.DesignStorm()(S.Scenario(10, "RCP4.5", "SSP2", 2050))
    b = S.DesignStorm()(S.Scenario(100, "RCP2.6", "SSP1", 2030, False))
    both = Forcing.stack([a, b])
    assert both.rain.n == 2 and both.tide.n == 2
    assert np.allclose(both.rain(5000.0), [a.rain(5000.0)[0], b.rain(5000.0)[0]], rtol=1e-15)
    assert np.allclose(both.tide(5000.0), [a.tide(5000.0)[0], b.tide(5000.0)[0]], rtol=1e-15)
    short = Forcing(a.rain, a.inflow, a.tide, a.t_end - 60.0)
    with pytest.raises(ValueError):
        Forcing.stack([a, short])


def _small_domain():
    from his_amd import workloads as W
    from his_amd.domain.site import domain_from
    ter = W.make_terrain(100, 81, 80.0)
    return domain_from(ter, W.chain_network(ter, 20))


def test_land_use_grows_with_pathway_and_year():
    from his_amd.api import grown_landuse
    dom = _small_domain()
    urban = lambda lu: int(np.isin(lu, (LU.DENSE_URBAN, LU.RESIDENTIAL)).sum())
    base = grown_landuse(dom, "SSP5", 2020)
    assert np.array_equal(base, dom.landuse)
    a, b = grown_landuse(dom, "SSP5", 2060), grown_landuse(dom, "SSP5", 2100)
    assert urban(base) < urban(a) < urban(b)
    assert urban(grown_landuse(dom, "SSP1", 2100)) < urban(b)
    assert np.array_equal(grown_landuse(dom, "SSP5", 2060), a)
    changed = a != base
    assert np.isin(base[changed], (LU.CROPLAND, LU.BARE, LU.GRASS)).all()
    assert not (changed & (dom.sea | dom.channel)).any()


def test_site_validation():
    from his_amd.api import SiteState, validate_site
    dom = _small_domain()
    good = SiteState(torch.zeros(dom.shape), torch.ones(dom.shape), torch.full(dom.shape, 0.03),
                     torch.ones(dom.network.n_reaches))
    validate_site(good, dom)
    for bad in (good.replace(storage_depth=torch.zeros(3, 3)), good.replace(manning=torch.zeros(dom.shape)),
                good.replace(infil_multiplier=torch.full(dom.shape, 0.5)),
                good.replace(channel_gamma=torch.ones(dom.network.n_reaches + 1)),
                good.replace(storage_depth=torch.full(dom.shape, math.nan))):
        with pytest.raises(ValueError):
            validate_site(bad, dom)
