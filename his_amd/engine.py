"""Assemble the coupled engine for one storm or for B storms over shared terrain.

Mirrors hydrointel/engine.py:78-238: ``normal_depth`` (Manning normal depth by
60-step bisection, slopes from the reach chainage), ``build_engine`` and
``build_engine_batch`` (edge conditions: the sea edge is a tidal stage, the others
transmissive, or every edge and 1-D end closed with ``closed=True``; initial state:
sea at the tide level, channels at max(normal depth, tide)).

The domain and forcing types are this repository's (``his_amd.workloads``): a
``Terrain`` with its sea mask and channel network, and per-member ``Series`` for rain,
tide and upstream inflow, which the device evaluates inside the coupled step.
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

    def times(self, interval: float) -> np.ndarray:
        n = int(np.floor(self.t_end / interval + 1e-9))
        return np.arange(n + 1) * interval


def normal_depth(one: SWE1D, Q, slope_min: float = 1e-4) -> torch.Tensor:
    """Manning normal depth for discharge Q in every cell, by bisection (engine.py:78-101)."""
    tp = one.topo
    z = one.z.detach().cpu().numpy()
    s = np.full(z.shape, slope_min)
    for ri in range(one.net.n_reaches):
        idx = np.nonzero(tp.reach_of == ri)[0]
        idx = idx[tp.kind[idx] == 0]
        if len(idx) > 1:
            dz = -np.gradient(z[..., idx], tp.chain[idx], axis=-1)
            s[..., idx] = np.maximum(dz, slope_min)
    b, m = one.sec.b.cpu().numpy(), one.sec.m.cpu().numpy()
    yb, ts = one.sec.yb.cpu().numpy(), one.sec.ts.cpu().numpy()
    nman = one.nman.cpu().numpy()
    if np.ndim(Q):
        Q = np.asarray(Q, dtype=float).reshape(one.bshape + (1,))
    lo, hi = np.zeros_like(s), np.broadcast_to(yb, s.shape).copy()
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        yt = np.minimum(mid, yb)
        A = yt * (b + m * yt) + ts * np.clip(mid - yb, 0.0, None)
        R = A / (b + 2.0 * yt * np.sqrt(1.0 + m * m))
        q = A * R ** (2 / 3) * np.sqrt(s) / nman
        lo, hi = np.where(q < Q, mid, lo), np.where(q < Q, hi, mid)
    return torch.as_tensor(0.5 * (lo + hi), dtype=torch.float64, device=one.device)


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


def build_engine_batch(cfg: SolverConfig, terrain, network, forcing: Forcing, mannings, storages, f0s, fcs, ks,
                       gammas, device=None, closed: bool = False, horton_t0: float = 0.0,
                       batched: bool = True) -> CoupledEngine:
    """One engine advancing B storms over one terrain; per-member (B, ny, nx) fields and gammas.
    ``batched=False`` with B=1 builds the unbatched (ny, nx) engine."""
    B = len(gammas)
    if not (len(mannings) == len(storages) == len(f0s) == len(fcs) == len(ks) == B):
        raise ValueError(f"batched inputs disagree in length: {B} gammas but "
                         f"{len(mannings)}/{len(storages)}/{len(f0s)}/{len(fcs)}/{len(ks)} fields")
    if not batched and B != 1:
        raise ValueError("an unbatched engine holds exactly one member")
    dtype = cfg.dtype
    net0 = _close_network(network) if closed else network
    nets = [apply_gamma(net0, g) for g in gammas]
    bcs = _bcs(forcing.tide, closed)
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
    eng = build_engine_batch(cfg, wl.terrain, wl.network, fo, list(n), list(S), list(f0), list(fc), list(k),
                             [m.gamma for m in wl.members], device=device, closed=closed, batched=batched)
    return eng, fo


def build_engine(cfg: SolverConfig, terrain, network, forcing: Forcing, storage, kappa, manning, gamma,
                 f0=None, fc=None, k=None, device=None, closed: bool = False) -> CoupledEngine:
    """Unbatched engine for one storm (the reference's build_engine, engine.py:206-238).
    Horton fields default to 60 / 10 mm/h and 2 /h on land, scaled by ``kappa``."""
    land = ~terrain.sea
    kap = np.asarray(kappa, dtype=np.float64) * np.ones(terrain.shape)
    f0 = np.where(land, 60.0 / 3.6e6, 0.0) if f0 is None else np.asarray(f0, float)
    fc = np.where(land, 10.0 / 3.6e6, 0.0) if fc is None else np.asarray(fc, float)
    k = np.where(land, 2.0 / 3600.0, 0.0) if k is None else np.asarray(k, float)
    return build_engine_batch(cfg, terrain, network, forcing, [manning * np.ones(terrain.shape)],
                              [storage * np.ones(terrain.shape)], [f0 * kap], [fc * kap], [k],
                              [gamma], device=device, closed=closed, batched=False)
