"""Assemble the coupled engine for one storm or for B storms over shared terrain.

Two entry points keep hydrointel's signatures (hydrointel/engine.py:169-238), so its callers bind
unchanged:

* ``build_engine(cfg, domain, forcing, storage, kappa, manning, gamma, landuse=None, device=None,
  closed=False)`` -- one storm;
* ``build_engine_batch(cfg, domain, forcings, storages, kappas, mannings, gammas, landuses,
  device=None) -> (engine, forcing)`` -- B storms sharing the terrain and a time base.

Both take the site fields (retention depth, infiltration multiplier kappa, roughness, per-reach
channel conveyance gamma) and the land-use raster, from which the Horton fields follow (class table
times kappa); a forcing carrying an SSP scales gamma by its drainage-investment factor, clipped to
[0.5, 3], as hydrointel does.  The sea edge of the domain takes the tidal stage, the other edges are
transmissive, or every edge and 1-D end is closed with ``closed=True``; the sea starts at the tide
level and the channels at max(normal depth, tide).

``engine_from_fields`` is the same assembly from per-member fields directly (roughness, retention,
Horton f0/fc/k, gamma), which the seeded benchmark workloads and the serving pools use.
``normal_depth`` is Manning normal depth by 60-step bisection (hydrointel/engine.py:78-101).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .config import SolverConfig
from .domain.channels import apply_gamma
from .forcing import Series
from .solver.boundaries import EdgeBC
from .solver.coupling import CoupledEngine
from .solver.infiltration import Horton
from .solver.swe1d import SWE1D
from .solver.swe2d import SWE2D
from .workloads import Workload


@dataclass
class Forcing:
    rain: Series            # m/s per member
    inflow: Series          # m^3/s per member
    tide: Series            # m per member
    t_end: float
    storm_start: float = 0.0
    ssp: str | None = None  # socioeconomic pathway whose drainage investment scales channel conveyance

    def times(self, interval: float) -> np.ndarray:
        n = int(np.floor(self.t_end / interval + 1e-9))
        return np.arange(n + 1) * interval

    @staticmethod
    def stack(members: list) -> "Forcing":
        """One batched forcing from one-member forcings on a shared time base (same end, same start)."""
        from .forcing import stack
        f0 = members[0]
        if any(m.t_end != f0.t_end or m.storm_start != f0.storm_start for m in members):
            raise ValueError("batched storms must share their time base (t_end and storm start)")
        return Forcing(stack([m.rain for m in members]), stack([m.inflow for m in members]),
                       stack([m.tide for m in members]), f0.t_end, f0.storm_start, None)


def BatchForcing(forcings: list) -> Forcing:  # noqa: N802 -- hydrointel's class name
    """B members' forcing on one shared time base, as hydrointel's ``BatchForcing(forcings)``
    (engine.py:104-143) builds it: ``Forcing.stack``; ValueError when the members differ in it."""
    return Forcing.stack(list(forcings))


def build_forcing(cfg, return_period: int, rcp: str, ssp: str, horizon_year: int, tide_on: bool,
                  storm_tide_offset_h: float, domain_area_km2: float | None = None) -> Forcing:
    """A scenario's forcing with hydrointel's signature (engine.py:53-70), through the registered
    resolver (``his_amd.scenarios.register_resolver``; the default is this package's design storm).
    ``cfg`` and the domain area belong to hydrointel's own storm construction, which a resolver
    installed from it applies; the default storm does not use them."""
    from .scenarios import Scenario, resolve
    return resolve(Scenario(int(return_period), rcp, ssp, int(horizon_year), bool(tide_on),
                            float(storm_tide_offset_h)))


def scenario_landuse(domain, ssp, horizon_year: int, base_year: int = 2025) -> np.ndarray:
    """The domain's land use at the horizon year under an SSP (an ``SSPState`` or its name), with
    hydrointel's signature (engine.py:73-75); the growth rule is ``api.grown_landuse``'s."""
    from .api import grown_landuse
    return grown_landuse(domain, ssp, int(horizon_year), base_year)


def normal_depth(one: SWE1D, Q, slope_min: float = 1e-4) -> torch.Tensor:
    """Manning normal depth for discharge Q in every cell, by bisection (engine.py:78-101)."""
    sec = one.sec
    y = normal_depth_arrays(one.topo, one.net.n_reaches, one.z.detach().cpu().numpy(), sec.b.cpu().numpy(),
                            sec.m.cpu().numpy(), sec.yb.cpu().numpy(), sec.ts.cpu().numpy(), one.nman.cpu().numpy(),
                            Q, one.bshape, slope_min)
    return torch.as_tensor(y, dtype=torch.float64, device=one.device)


def normal_depth_arrays(tp, n_reaches: int, z, b, m, yb, ts, nman, Q, bshape=(), slope_min: float = 1e-4):
    """``normal_depth`` on host arrays: the bed slope along each reach's nodes (at least slope_min),
    then 60 bisection steps of Manning's equation between 0 and the bankfull depth."""
    s = np.full(z.shape, slope_min)
    for ri in range(n_reaches):
        idx = np.nonzero(tp.reach_of == ri)[0]
        idx = idx[tp.kind[idx] == 0]
        if len(idx) > 1:
            dz = -np.gradient(z[..., idx], tp.chain[idx], axis=-1)
            s[..., idx] = np.maximum(dz, slope_min)
    if np.ndim(Q):
        Q = np.asarray(Q, dtype=float).reshape(tuple(bshape) + (1,))
    lo, hi = np.zeros_like(s), np.broadcast_to(yb, s.shape).copy()
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        yt = np.minimum(mid, yb)
        A = yt * (b + m * yt) + ts * np.clip(mid - yb, 0.0, None)
        R = A / (b + 2.0 * yt * np.sqrt(1.0 + m * m))
        q = A * R ** (2 / 3) * np.sqrt(s) / nman
        lo, hi = np.where(q < Q, mid, lo), np.where(q < Q, hi, mid)
    return 0.5 * (lo + hi)


def _solver_cfg(cfg) -> SolverConfig:
    """A SolverConfig, or the solver part of a run configuration that carries one."""
    return cfg.solver if hasattr(cfg, "solver") else (cfg or SolverConfig())


def _effective_gamma(gamma, ssp: str | None) -> np.ndarray:
    g = np.asarray(gamma, dtype=np.float64)
    if ssp is None:
        return g
    from .scenarios import drainage_factor
    return np.clip(g * drainage_factor(ssp), 0.5, 3.0)


def _bcs(tide, closed: bool, sea_edge: str = "east"):
    if closed:
        return {e: EdgeBC("closed") for e in ("west", "east", "south", "north")}
    bcs = {e: EdgeBC("transmissive") for e in ("west", "east", "south", "north")}
    bcs[sea_edge] = EdgeBC("stage", tide)
    return bcs


def _close_network(net):
    import copy
    out = copy.deepcopy(net)
    for r in out.reaches:
        r.up = r.up if r.up[0] == "junction" else ("closed",)
        r.down = r.down if r.down[0] == "junction" else ("closed",)
    return out


def engine_from_fields(cfg: SolverConfig, terrain, network, forcing: Forcing, mannings, storages, f0s, fcs, ks,
                       gammas, device=None, closed: bool = False, horton_t0: float = 0.0,
                       batched: bool = True, sea_edge: str = "east") -> CoupledEngine:
    """One engine advancing B storms over one terrain; per-member (B, ny, nx) fields and gammas.
    ``batched=False`` with B=1 builds the unbatched (ny, nx) engine."""
    cfg = _solver_cfg(cfg)
    B = len(gammas)
    if not (len(mannings) == len(storages) == len(f0s) == len(fcs) == len(ks) == B):
        raise ValueError(f"batched inputs disagree in length: {B} gammas but "
                         f"{len(mannings)}/{len(storages)}/{len(f0s)}/{len(fcs)}/{len(ks)} fields")
    if not batched and B != 1:
        raise ValueError("an unbatched engine holds exactly one member")
    dtype = cfg.dtype
    net0 = _close_network(network) if closed else network
    nets = [apply_gamma(net0, g) for g in gammas]
    bcs = _bcs(forcing.tide, closed, sea_edge)
    st = (lambda xs: np.stack([np.asarray(x, dtype=np.float64) for x in xs])) if batched else \
        (lambda xs: np.asarray(xs[0], dtype=np.float64))
    dev = torch.device(device) if device is not None else torch.device("cuda")
    T = lambda a: torch.as_tensor(a, dtype=dtype, device=dev)
    horton = Horton(T(st(f0s)), T(st(fcs)), T(st(ks)), horton_t0)
    two = SWE2D(terrain.z, st(mannings), terrain.dx, cfg, bcs, dev, dtype, storage=st(storages), horton=horton,
                batch=B if batched else None)
    one = SWE1D(nets if batched else nets[0], cfg, dev, inflow=None if closed else forcing.inflow,
                stage=forcing.tide)
    eta0 = np.asarray(forcing.tide(0.0), dtype=np.float64).reshape(-1)
    sea = torch.as_tensor(terrain.sea, device=dev)
    eta0_2d = torch.as_tensor(eta0, dtype=dtype, device=dev).reshape((B, 1, 1) if batched else (1, 1))
    h0 = torch.where(sea, torch.clamp(eta0_2d - two.z, min=0.0), torch.zeros_like(two.z))
    two.set_state(h0.expand(two.bshape + (two.ny, two.nx)))
    yn = normal_depth(one, forcing.inflow(0.0))
    floor = torch.as_tensor(eta0 if not closed else np.full(B, -1e9), dtype=torch.float64, device=dev)
    floor = floor.reshape((B, 1) if batched else (1,)).expand_as(one.z)
    one.set_level(torch.maximum(one.z + yn, floor))
    return CoupledEngine(two, one, cfg, rain=forcing.rain)


def build_from_workload(wl: Workload, cfg: SolverConfig | None = None, device=None, batched: bool | None = None,
                        closed: bool = False):
    """(engine, forcing) for a seeded workload; precision and slot width from the workload."""
    cfg = (cfg or SolverConfig()).replace(precision=wl.precision, slot_width_frac=wl.slot_width_frac)
    batched = wl.batch > 1 if batched is None else batched
    rain, tide, inflow = wl.forcing(batched=batched)
    fo = Forcing(rain, inflow, tide, wl.storm.t_end)
    n, S, f0, fc, k = wl.fields()
    eng = engine_from_fields(cfg, wl.terrain, wl.network, fo, list(n), list(S), list(f0), list(fc), list(k),
                             [m.gamma for m in wl.members], device=device, closed=closed, batched=batched)
    return eng, fo


def _field(a, shape) -> np.ndarray:
    a = a.detach().to(torch.float64).cpu().numpy() if torch.is_tensor(a) else np.asarray(a, dtype=np.float64)
    return np.broadcast_to(a, shape).astype(np.float64)


def build_engine(cfg, domain, forcing: Forcing, storage, kappa, manning, gamma, landuse=None, device=None,
                 closed: bool = False) -> CoupledEngine:
    """One storm over ``domain`` (hydrointel/engine.py:206-238): the site's retention depth, infiltration
    multiplier and roughness fields, its per-reach conveyance gamma, and the Horton fields of the land use
    (the domain's unless ``landuse`` is given) scaled by kappa."""
    from .domain import landuse as LU
    shp = domain.shape
    lu = domain.landuse if landuse is None else np.asarray(landuse)
    f0, fc, k = LU.horton_fields(lu, _field(kappa, shp))
    return engine_from_fields(cfg, domain.terrain, domain.network, forcing, [_field(manning, shp)],
                              [_field(storage, shp)], [f0], [fc], [k], [_effective_gamma(gamma, forcing.ssp)],
                              device=device, closed=closed, horton_t0=forcing.storm_start, batched=False,
                              sea_edge=getattr(domain, "sea_edge", "east"))


def build_engine_batch(cfg, domain, forcings: list, storages, kappas, mannings, gammas, landuses, device=None):
    """B storms over ``domain`` advanced together (hydrointel/engine.py:169-203): every member keeps its
    own roughness, retention, infiltration, conveyance, rain, tide, inflow and mass balance; the storms
    share the terrain, a time base and the time step.  Returns (engine, batched forcing)."""
    from .domain import landuse as LU
    b = len(forcings)
    if not (len(storages) == len(kappas) == len(mannings) == len(gammas) == len(landuses) == b):
        raise ValueError(f"batched inputs disagree in length: {b} forcings but {len(storages)}/{len(kappas)}/"
                         f"{len(mannings)}/{len(gammas)}/{len(landuses)} fields")
    shp = domain.shape
    fo = Forcing.stack(forcings)
    hz = [LU.horton_fields(np.asarray(lu), _field(kap, shp)) for lu, kap in zip(landuses, kappas)]
    gam = [_effective_gamma(g, f.ssp) for g, f in zip(gammas, forcings)]
    eng = engine_from_fields(cfg, domain.terrain, domain.network, fo, [_field(m, shp) for m in mannings],
                             [_field(x, shp) for x in storages], [h[0] for h in hz], [h[1] for h in hz],
                             [h[2] for h in hz], gam, device=device, horton_t0=fo.storm_start, batched=True,
                             sea_edge=getattr(domain, "sea_edge", "east"))
    return eng, fo
