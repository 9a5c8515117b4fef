"""The run-level interface: ``simulate`` and ``simulate_batch`` with hydrointel's signatures.

    simulate(site, scenario, output_interval_s=None, compare_to_baseline=False) -> FloodResult
    simulate_batch(sites, scenarios, output_interval_s=None) -> [FloodResult]

as hydrointel/api.py:211-324 declares them, so ``generate``, the benchmark stages and an optimiser
call them unchanged.  Like hydrointel, they work on a process-wide context (``configure``): the
solver configuration, the device and the ``Domain`` (terrain, channel network, land use).  Each call

1. validates the scenario and the site (ValueError for shapes, non-finite values and ranges);
2. turns each scenario into rain, tide and inflow through the registered resolver
   (``his_amd.scenarios``) and the scenario's land use into Horton fields scaled by the site's
   infiltration multiplier;
3. builds the coupled engine (``build_engine`` / ``build_engine_batch``), runs it to the end of the
   storm stopping at every output time, and asserts mass balance per member (AssertionError);
4. assembles ``FloodResult``: ``depth_max`` is the maximum over the output-time snapshots
   (hydrointel/api.py:239) and ``speed_max`` likewise; channel series cover the interior 1-D cells;
   volumes and the mass-balance error are per member; ``extras`` carries the forcing, the run log,
   the ledger and mass series, the boundary volumes and the exposure metrics.

A batch of one is routed to ``simulate`` (api.py:274-276).  ``compare_to_baseline`` fills
``max_depth_increase_m`` and ``area_worsened_ha`` against a run of ``baseline_site(scenario)``,
cached per scenario and output interval.  ``predict`` (the neural surrogate) is not part of this
package.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np
import torch

from .config import SolverConfig
from .domain import landuse as LU
from .engine import build_engine, build_engine_batch
from .scenarios import Scenario, label, population_factor, resolve, validate

WORSENED_THRESHOLD_M = 0.01


@dataclass(frozen=True)
class SiteState:
    """The physical state a design may change (hydrointel/api.py:42-52)."""
    storage_depth: torch.Tensor      # S(x, y) [m], (ny, nx)
    infil_multiplier: torch.Tensor   # kappa(x, y) [-], (ny, nx)
    manning: torch.Tensor            # n(x, y) [s m^-1/3], (ny, nx)
    channel_gamma: torch.Tensor      # gamma per reach [-], (n_reaches,)

    def replace(self, **kw) -> "SiteState":
        return dataclasses.replace(self, **kw)


@dataclass(frozen=True)
class FloodResult:
    """hydrointel/api.py:66-86; provenance is a plain dict here."""
    depth_max: torch.Tensor          # (ny, nx) [m]
    depth_series: torch.Tensor       # (nt, ny, nx) [m]
    u: torch.Tensor                  # (nt, ny, nx) [m/s]
    v: torch.Tensor                  # (nt, ny, nx) [m/s]
    channel_eta: torch.Tensor        # (nt, n_nodes) [m]
    channel_q: torch.Tensor          # (nt, n_nodes) [m^3/s]
    times: torch.Tensor              # (nt,) [s]
    infiltrated_volume_m3: float
    stored_volume_m3: float
    mass_balance_error: float
    provenance: dict
    in_distribution: bool = True
    distribution_warnings: list = field(default_factory=list)
    extras: dict = field(default_factory=dict)
    speed_max: torch.Tensor | None = None
    max_depth_increase_m: float | None = None
    area_worsened_ha: float | None = None


# ---------------------------------------------------------------------------- context
class Context:
    """What every call shares: solver configuration, device and domain; land use per SSP and year."""

    def __init__(self, domain=None, cfg: SolverConfig | None = None, device=None):
        self.cfg = cfg or SolverConfig()
        self._domain = domain
        self.device = device
        self._landuse: dict = {}

    @property
    def domain(self):
        if self._domain is None:
            self._domain = default_domain()
        return self._domain

    def landuse(self, ssp: str, year: int) -> np.ndarray:
        key = (ssp, int(year))
        if key not in self._landuse:
            self._landuse[key] = grown_landuse(self.domain, ssp, int(year))
        return self._landuse[key]

    def provenance(self, engine: str = "engine") -> dict:
        return dict(engine=engine, backend="his_amd (gfx950 HIP)", precision=self.cfg.precision,
                    device=str(self.device or "cuda"), domain=getattr(self.domain, "name", ""))


_CTX: list = [None]
_BASELINE_PEAK: dict = {}


def default_domain():
    """The 8 km x 6.5 km coastal plain at 20 m (400 x 325) with the 60-reach network and the seeded
    land-use map."""
    from . import workloads as W
    from .domain.site import domain_from
    ter = W.make_terrain(400, 325, 20.0)
    return domain_from(ter, W.network60(ter), name="coastal plain 400x325@20m")


def configure(domain=None, cfg: SolverConfig | None = None, device=None) -> Context:
    _CTX[0] = Context(domain, cfg, device)
    _BASELINE_PEAK.clear()
    return _CTX[0]


def context() -> Context:
    return _CTX[0] or configure()


def grown_landuse(domain, ssp, year: int, base_year: int = 2025) -> np.ndarray:
    """The domain's land use at a horizon year under an SSP (a name or an ``SSPState``): cropland, bare
    ground and grass turn residential (dense urban nearest the core) at the pathway's expansion rate
    per decade since ``base_year`` (hydrointel's origin, engine.py:73-75), nearest the existing town
    first.  Deterministic.  The growth rule is this package's; hydrointel's own expansion model sits
    upstream of the engine."""
    from .scenarios import ssp_state
    base = np.asarray(domain.landuse)
    rate = (ssp_state(ssp) if isinstance(ssp, str) else ssp).urban_expansion_rate
    grow = np.isin(base, (LU.CROPLAND, LU.BARE, LU.GRASS)) & ~domain.channel & ~domain.sea
    n_new = int(round(rate * max(year - base_year, 0) / 10.0 * grow.sum()))
    if n_new <= 0:
        return base.copy()
    from scipy.ndimage import distance_transform_edt
    town = np.isin(base, (LU.DENSE_URBAN, LU.RESIDENTIAL, LU.ROAD))
    dist = distance_transform_edt(~town) if town.any() else np.zeros(base.shape)
    order = np.argsort(np.where(grow, dist, np.inf), axis=None, kind="stable")[:n_new]
    out = base.copy().reshape(-1)
    near = dist.reshape(-1)[order] <= np.percentile(dist.reshape(-1)[order], 25)
    out[order] = np.where(near, LU.DENSE_URBAN, LU.RESIDENTIAL)
    return out.reshape(base.shape)


# ------------------------------------------------------------------------- sites
def baseline_site(scenario: Scenario) -> SiteState:
    """No interventions on the scenario's land surface: the class table's retention and roughness,
    kappa = 1 and gamma = 1 (hydrointel/api.py:175-184)."""
    ctx = context()
    lu = ctx.landuse(scenario.ssp, scenario.horizon_year)
    f32 = lambda a: torch.as_tensor(np.asarray(a, np.float32))
    return SiteState(f32(LU.lookup(lu, "storage_m")), f32(np.ones(ctx.domain.shape)), f32(LU.lookup(lu, "manning")),
                     f32(np.ones(ctx.domain.network.n_reaches)))


def _np(a) -> np.ndarray:
    if torch.is_tensor(a):
        return a.detach().to(torch.float64).cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def validate_site(site: SiteState, domain) -> None:
    """Shapes, finiteness and ranges of a site (hydrointel/api.py:456-473), as ValueError."""
    for name in ("storage_depth", "infil_multiplier", "manning"):
        a = _np(getattr(site, name))
        if a.shape != tuple(domain.shape):
            raise ValueError(f"SiteState.{name} is {a.shape} but the domain grid is {tuple(domain.shape)} (ny, nx)")
        if not np.isfinite(a).all():
            raise ValueError(f"SiteState.{name} holds non-finite values")
    n_reaches = domain.network.n_reaches
    if _np(site.channel_gamma).shape != (n_reaches,):
        raise ValueError(f"SiteState.channel_gamma needs one value per reach ({n_reaches},)")
    checks = (("storage_depth", lambda a: a.min() >= 0, "must be >= 0 m"),
              ("infil_multiplier", lambda a: a.min() >= 1.0 - 1e-6, "must be >= 1"),
              ("manning", lambda a: a.min() > 0, "must be > 0"),
              ("channel_gamma", lambda a: a.min() > 0, "must be > 0"))
    for name, ok, why in checks:
        if not ok(_np(getattr(site, name))):
            raise ValueError(f"SiteState.{name} {why}")


def check_site(site: SiteState, scenario: Scenario | None = None):
    """(in distribution, warnings).  The site is validated; a training envelope belongs to the
    surrogate, which is not part of this package, so no site is vouched for."""
    validate_site(site, context().domain)
    return False, ["no training envelope here (the surrogate is not part of this package): the engine result "
                   "is the ground truth"]


# -------------------------------------------------------------------------- runs
def _snapshots(eng, t_end: float, times, label_: str):
    snaps = []
    # the snapshots are only collected here, so their arrays may reach the host while the run steps on
    err = eng.run(t_end, out_times=times, on_snapshot=snaps.append, label=label_, async_outputs=True)
    return snaps, err


def _member(v, b: int, batched: bool):
    if batched and getattr(v, "ndim", 0):
        return float(np.asarray(v, dtype=np.float64).reshape(-1)[b])
    return float(v) if isinstance(v, (int, float, np.floating, np.integer)) else v


def _result(eng, snaps, err, b: int, batched: bool, prov: dict, ok_warn, scenario, extras: dict) -> FloodResult:
    interior = torch.as_tensor(eng.one.topo.kind <= 1)
    pick = (lambda a: a[b]) if batched else (lambda a: a)
    st = lambda k: torch.as_tensor(np.stack([pick(getattr(s, k)) for s in snaps]))
    h, u, v = st("h"), st("u"), st("v")
    src = eng.mb.sources()
    s_b = {k: float(np.asarray(x, dtype=np.float64).reshape(-1)[b]) for k, x in src.items()}
    stored = np.asarray(eng.two.storage_volume().detach().cpu().numpy()).reshape(-1)[b]
    depth_max = h.amax(0)
    log = eng.log
    if batched:
        log = dataclasses.replace(log, exchanged_in_m3=_member(log.exchanged_in_m3, b, True),
                                  exchanged_out_m3=_member(log.exchanged_out_m3, b, True), series=[])
    return FloodResult(
        depth_max=depth_max, depth_series=h, u=u, v=v,
        speed_max=torch.sqrt(u.double() ** 2 + v.double() ** 2).amax(0).float(),
        channel_eta=st("eta1")[:, interior], channel_q=st("q1")[:, interior],
        times=torch.as_tensor(np.array([x.t for x in snaps])),
        infiltrated_volume_m3=s_b["infiltration"], stored_volume_m3=float(stored),
        mass_balance_error=float(np.asarray(err, dtype=np.float64).reshape(-1)[b]), provenance=prov,
        in_distribution=ok_warn[0], distribution_warnings=list(ok_warn[1]),
        extras={"run_log": log, "sources": s_b, "channel_cells": np.nonzero(eng.one.topo.kind <= 1)[0],
                "topology": eng.one.topo, "boundary_in_m3": s_b["bnd2d_in"] + s_b["bnd1d_in"],
                "boundary_out_m3": s_b["bnd2d_out"] + s_b["bnd1d_out"],
                # volume ledger at each output time and the mass-balance checks, this member's
                # (hydrointel/api.py:245-247, 315-320)
                "ledger": [{k: _member(x, b, batched) for k, x in s.ledger.items()} for s in snaps],
                "mass_series": [{k: _member(x, b, batched) for k, x in r.items()} for r in eng.mb.series],
                **extras, **exposure_metrics(depth_max, scenario)})


def simulate(site: SiteState, scenario: Scenario, output_interval_s: float | None = None,
             compare_to_baseline: bool = False) -> FloodResult:
    """One storm through the coupled engine; raises AssertionError if mass balance fails."""
    ctx = context()
    validate(scenario)
    ok_warn = check_site(site, scenario)
    fo = resolve(scenario, ctx.domain)
    lu = ctx.landuse(scenario.ssp, scenario.horizon_year)
    eng = build_engine(ctx.cfg, ctx.domain, fo, _np(site.storage_depth), _np(site.infil_multiplier),
                       _np(site.manning), _np(site.channel_gamma), landuse=lu, device=ctx.device)
    times = fo.times(output_interval_s or ctx.cfg.output_interval_s)
    snaps, err = _snapshots(eng, fo.t_end, times, label(scenario))
    res = _result(eng, snaps, err, 0, False, ctx.provenance(), ok_warn, scenario, {"forcing": fo})
    if compare_to_baseline:
        ref = _baseline_peak(scenario, output_interval_s)
        res = dataclasses.replace(res, **disbenefit(res.depth_max, ref))
    return res


def simulate_batch(sites: list, scenarios: list, output_interval_s: float | None = None) -> list:
    """B storms on one set of kernels: shared terrain, time base and time step (the smallest any
    member needs); per-member site, return period, pathway, tide and mass balance."""
    ctx = context()
    if len(sites) != len(scenarios):
        raise ValueError(f"{len(sites)} sites but {len(scenarios)} scenarios")
    if not sites:
        return []
    if len(sites) == 1:
        return [simulate(sites[0], scenarios[0], output_interval_s)]
    checks = []
    for s, sc in zip(sites, scenarios):
        validate(sc)
        checks.append(check_site(s, sc))
    fos = [resolve(sc, ctx.domain) for sc in scenarios]
    lus = [ctx.landuse(sc.ssp, sc.horizon_year) for sc in scenarios]
    eng, fo = build_engine_batch(ctx.cfg, ctx.domain, fos, [_np(s.storage_depth) for s in sites],
                                 [_np(s.infil_multiplier) for s in sites], [_np(s.manning) for s in sites],
                                 [_np(s.channel_gamma) for s in sites], lus, device=ctx.device)
    times = fo.times(output_interval_s or ctx.cfg.output_interval_s)
    snaps, err = _snapshots(eng, fo.t_end, times,
                            f"batch of {len(sites)}: " + "; ".join(label(sc) for sc in scenarios[:4]))
    prov = ctx.provenance()
    return [_result(eng, snaps, err, b, True, prov, checks[b], sc,
                    {"forcing": fos[b], "batch_index": b, "batch_size": len(sites)})
            for b, sc in enumerate(scenarios)]


def _baseline_peak(scenario: Scenario, output_interval_s=None) -> torch.Tensor:
    key = (scenario, float(output_interval_s or context().cfg.output_interval_s))
    if key not in _BASELINE_PEAK:
        _BASELINE_PEAK[key] = simulate(baseline_site(scenario), scenario, output_interval_s).depth_max
    return _BASELINE_PEAK[key]


def baseline(scenario: Scenario, engine: str = "simulate") -> FloodResult:
    """The flood with no interventions.  Only the engine is part of this package."""
    if engine != "simulate":
        raise ValueError("engine must be 'simulate' (the surrogate 'predict' is not part of this package)")
    return simulate(baseline_site(scenario), scenario)


# ------------------------------------------------------------------------ metrics
def land_mask(domain=None) -> np.ndarray:
    """Land = neither sea nor channel footprint."""
    return (domain or context().domain).land


def disbenefit(depth_max, baseline_depth_max) -> dict:
    """Largest rise of peak depth on land (0 if nowhere rises), and the land area where it rose by more
    than 1 cm (hydrointel/api.py:400-418)."""
    dom = context().domain
    land = torch.as_tensor(land_mask(dom))
    d = depth_max.detach().float().cpu() - torch.as_tensor(baseline_depth_max).detach().float().cpu()
    rise = torch.clamp(d[land].max(), min=0.0)
    return {"max_depth_increase_m": float(rise),
            "area_worsened_ha": float((d[land] > WORSENED_THRESHOLD_M).sum()) * dom.dx ** 2 / 1e4}


def exposure_metrics(depth_max, scenario: Scenario, threshold_m: float = 0.15) -> dict:
    """The exposure raster (class exposure index times the SSP's population factor) and the
    exposure-weighted land area flooded deeper than ``threshold_m`` (hydrointel/api.py:430-442)."""
    ctx = context()
    dom = ctx.domain
    lu = ctx.landuse(scenario.ssp, scenario.horizon_year)
    expo = LU.lookup(lu, "exposure") * population_factor(scenario.ssp)
    wet = depth_max.detach().cpu().numpy() >= threshold_m
    return {"exposure": expo.astype(np.float32),
            f"exposure_weighted_area_gt_{threshold_m:g}m_km2": float(np.sum(expo[wet & dom.land]) * dom.dx ** 2 / 1e6)}
