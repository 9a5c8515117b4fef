"""``simulate`` / ``simulate_batch``: one storm, or B storms over one terrain, to FloodResults.

Mirrors hydrointel/api.py:211-324 for the part on the engine's path. Each call:
1. checks the site arrays;
2. builds the coupled engine (``build_engine`` / ``build_engine_batch``);
3. runs it to the end of the storm, stopping at every output time;
4. assembles ``FloodResult``:
   * ``depth_max`` is the maximum over the output-time snapshots (api.py:239), not over
     every step, and ``speed_max`` likewise from the snapshot velocities;
   * channel series cover the interior 1-D cells (kind <= 1);
   * infiltrated and stored volumes and the mass-balance error come per member.

hydrointel turns a ``Scenario`` (return period, RCP, SSP, horizon) into rainfall, tide and
inflow through its forcing package and land-use tables. Those sit upstream of the engine
and are not part of this package, so these functions take the storm's ``Forcing`` and the
terrain directly. Everything downstream of that matches. So does the B=1 routing: a
batch of one is the plain solver, as in api.py:274-276.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np
import torch

from .config import SolverConfig
from .engine import Forcing, build_engine_batch
from .workloads import MMH

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


def baseline_site(terrain, network, manning: float = 0.035, storage_m: float = 0.005) -> SiteState:
    """No interventions: uniform roughness and retention on land, kappa = 1, gamma = 1."""
    land = ~terrain.sea
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32))
    return SiteState(t(np.where(land, storage_m, 0.0)), t(np.ones(terrain.shape)),
                     t(np.full(terrain.shape, manning)), t(np.ones(network.n_reaches)))


def _np(a) -> np.ndarray:
    if torch.is_tensor(a):
        return a.detach().to(torch.float64).cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def validate_site(site: SiteState, terrain, network) -> None:
    """hydrointel/api.py:461-478: shapes, finiteness and ranges, as ValueError."""
    for name in ("storage_depth", "infil_multiplier", "manning"):
        a = getattr(site, name)
        if tuple(a.shape) != terrain.shape:
            raise ValueError(f"SiteState.{name} has shape {tuple(a.shape)}, domain grid is {terrain.shape} (ny, nx)")
        if not torch.isfinite(torch.as_tensor(a)).all():
            raise ValueError(f"SiteState.{name} contains non-finite values")
    if tuple(site.channel_gamma.shape) != (network.n_reaches,):
        raise ValueError(f"SiteState.channel_gamma must have shape ({network.n_reaches},)")
    if float(site.storage_depth.min()) < 0:
        raise ValueError("SiteState.storage_depth must be >= 0 m")
    if float(site.infil_multiplier.min()) < 1.0 - 1e-6:
        raise ValueError("SiteState.infil_multiplier must be >= 1")
    if float(site.manning.min()) <= 0:
        raise ValueError("SiteState.manning must be > 0")
    if float(site.channel_gamma.min()) <= 0:
        raise ValueError("SiteState.channel_gamma must be > 0")


def _horton_fields(terrain, kappa):
    """Horton f0 = 60, fc = 10 mm/h and k = 2 /h on land, infiltration capacity scaled by kappa."""
    land = ~terrain.sea
    return (np.where(land, 60.0 * MMH, 0.0) * kappa, np.where(land, 10.0 * MMH, 0.0) * kappa,
            np.where(land, 2.0 / 3600.0, 0.0))


def _run(sites, forcing: Forcing, terrain, network, output_interval_s, cfg, device, label):
    f0, fc, k = zip(*(_horton_fields(terrain, _np(s.infil_multiplier)) for s in sites))
    eng = build_engine_batch(cfg, terrain, network, forcing, [_np(s.manning) for s in sites],
                             [_np(s.storage_depth) for s in sites], list(f0), list(fc), list(k),
                             [_np(s.channel_gamma) for s in sites], device=device, batched=len(sites) > 1,
                             horton_t0=forcing.storm_start)
    snaps = []
    times = forcing.times(output_interval_s or cfg.output_interval_s)
    err = eng.run(forcing.t_end, out_times=times, on_snapshot=snaps.append, label=label)
    return eng, snaps, err


def _member(v, b: int, batched: bool):
    """Member b's value of a ledger or mass-series entry (arrays over the batch when batched)."""
    if batched and getattr(v, "ndim", 0):
        return float(np.asarray(v, dtype=np.float64).reshape(-1)[b])
    return float(v) if isinstance(v, (int, float, np.floating, np.integer)) else v


def _result(eng, snaps, err, b, batched: bool, prov: dict, extras: dict) -> FloodResult:
    interior = torch.as_tensor(eng.one.topo.kind <= 1)
    pick = (lambda a: a[b]) if batched else (lambda a: a)
    st = lambda k: torch.as_tensor(np.stack([pick(getattr(s, k)) for s in snaps]))
    h, u, v = st("h"), st("u"), st("v")
    src = eng.mb.sources()
    s_b = {k: float(np.asarray(v_, dtype=np.float64).reshape(-1)[b]) for k, v_ in src.items()}
    sv = eng.two.storage_volume().detach().cpu().numpy()
    return FloodResult(
        depth_max=h.amax(0), depth_series=h, u=u, v=v,
        speed_max=torch.sqrt(u.double() ** 2 + v.double() ** 2).amax(0).float(),
        channel_eta=st("eta1")[:, interior], channel_q=st("q1")[:, interior],
        times=torch.as_tensor(np.array([x.t for x in snaps])),
        infiltrated_volume_m3=s_b["infiltration"], stored_volume_m3=float(np.asarray(sv).reshape(-1)[b]),
        mass_balance_error=float(np.asarray(err, dtype=np.float64).reshape(-1)[b]), provenance=prov,
        extras={"run_log": eng.log, "sources": s_b, "channel_cells": np.nonzero(eng.one.topo.kind <= 1)[0],
                "topology": eng.one.topo, "boundary_in_m3": s_b["bnd2d_in"] + s_b["bnd1d_in"],
                "boundary_out_m3": s_b["bnd2d_out"] + s_b["bnd1d_out"],
                # volume ledger at each output time and the mass-balance checks, this member's
                # (hydrointel/api.py:245-247, 315-320)
                "ledger": [{k: _member(v, b, batched) for k, v in x.ledger.items()} for x in snaps],
                "mass_series": [{k: _member(v, b, batched) for k, v in rec.items()} for rec in eng.mb.series],
                **extras})


def simulate(site: SiteState, forcing: Forcing, terrain, network, output_interval_s: float | None = None,
             cfg: SolverConfig | None = None, device=None, baseline_depth_max=None) -> FloodResult:
    """One storm through the coupled engine; raises AssertionError if mass balance fails.
    ``baseline_depth_max`` fills max_depth_increase_m / area_worsened_ha (api.py:411-427)."""
    cfg = cfg or SolverConfig()
    validate_site(site, terrain, network)
    if forcing.rain.n != 1:
        raise ValueError(f"simulate takes one storm; the forcing holds {forcing.rain.n}")
    eng, snaps, err = _run([site], forcing, terrain, network, output_interval_s, cfg, device, "simulate")
    prov = dict(engine="his_amd", precision=cfg.precision, device=str(eng.two.device))
    res = _result(eng, snaps, err, 0, False, prov, {"forcing": forcing})
    if baseline_depth_max is not None:
        res = dataclasses.replace(res, **disbenefit(res.depth_max, baseline_depth_max, terrain, network))
    return res


def simulate_batch(sites: list, forcing: Forcing, terrain, network, output_interval_s: float | None = None,
                   cfg: SolverConfig | None = None, device=None) -> list:
    """B storms over one terrain on one set of kernels (shared dt, as hydrointel/api.py:272-324);
    per-member forcing rows, per-member mass balance; a batch of one is ``simulate``."""
    cfg = cfg or SolverConfig()
    if not sites:
        return []
    if forcing.rain.n != len(sites):
        raise ValueError(f"{len(sites)} sites but {forcing.rain.n} forcing members")
    if len(sites) == 1:
        return [simulate(sites[0], forcing, terrain, network, output_interval_s, cfg, device)]
    for s in sites:
        validate_site(s, terrain, network)
    eng, snaps, err = _run(sites, forcing, terrain, network, output_interval_s, cfg, device,
                           f"batch of {len(sites)}")
    prov = dict(engine="his_amd", precision=cfg.precision, device=str(eng.two.device))
    return [_result(eng, snaps, err, b, True, prov, {"forcing": forcing, "batch_index": b, "batch_size": len(sites)})
            for b in range(len(sites))]


def land_mask(terrain, network) -> np.ndarray:
    """Land = not sea and not under a channel reach (a 1-D cell's 2-D cell)."""
    from .solver.swe1d import build_topology
    tp = build_topology(network)
    ch = np.zeros(terrain.shape, bool)
    i = np.clip((tp.x / terrain.dx).astype(int), 0, terrain.nx - 1)
    j = np.clip((tp.y / terrain.dx).astype(int), 0, terrain.ny - 1)
    ch[j[tp.kind <= 1], i[tp.kind <= 1]] = True
    return ~(terrain.sea | ch)


def disbenefit(depth_max, baseline_depth_max, terrain, network) -> dict:
    """Largest rise of peak depth on land, and land area where it rose by more than 1 cm
    (hydrointel/api.py:411-427)."""
    land = torch.as_tensor(land_mask(terrain, network))
    d = depth_max.detach().float().cpu() - torch.as_tensor(baseline_depth_max).detach().float().cpu()
    rise = torch.clamp(d[land].max(), min=0.0)
    return {"max_depth_increase_m": float(rise),
            "area_worsened_ha": float((d[land] > WORSENED_THRESHOLD_M).sum()) * terrain.dx ** 2 / 1e4}
