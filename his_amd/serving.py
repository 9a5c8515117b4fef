"""Continuous batching of storm requests on one GPU.

hydrointel solves one storm (``api.simulate``, hydrointel/api.py:211-270) or one batch
that starts and ends together (``simulate_batch``, api.py:272-324). A serving loop
instead has requests arriving at any time, over several grid sizes, with storms of any
length. Here:

* One ``his_engine`` holds a pool of slots per grid size, one grid group each (at most
  four). Every slot has its own dt (``dt_group_size = 1``), its own clock from t = 0,
  its own output times and its own end. One hipGraph replay advances every busy slot
  one coupled step; idle slots return at the top of each kernel.
* Admission: an arriving request waits in its size's queue until a slot is free. It is
  admitted only if the HBM footprint of all busy slots plus its own stays within the
  budget. Its state, fields, channel sections and forcing rows are written into the slot
  straight from its member draw, with the operations a standalone build of the same storm
  performs (``Pool.load_request``); no engine is built per request. Only that slot's CFL
  rates are recomputed (``his_engine_refresh_members``).
* A request's result is therefore the result of the same storm run alone, bit for bit:
  every kernel computes a member from that member's data only, and partial sums are
  reduced per member in a fixed order. ``tests/test_gpu_parity.py`` checks this.
* Depth maxima are taken over the request's output times, t = 0 included, as
  ``FloodResult.depth_max`` is (hydrointel/api.py:239). Mass balance is judged per
  request (hydrointel/solver/massbalance.py:58-76). A request whose state turns
  implausible (hydrointel/solver/coupling.py:169-192) or unstable fails alone.

    python -m his_amd.serving --requests 2000 --out results/serving.json
"""
from __future__ import annotations

import argparse
import json
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import workloads as W
from .config import SolverConfig
from .engine import Forcing, engine_from_fields
from .forcing import PiecewiseLinear, sine_tide
from .solver.coupling import (ERROR_CODES, MAX_PLAUSIBLE_DEPTH_2D_M, MAX_PLAUSIBLE_RISE_1D_M, CoupledEngine)

STEP_S = 60.0                     # knot spacing of the per-slot forcing tables
MAX_STORM_S = 6 * 3600.0
K_KNOTS = int(round(MAX_STORM_S / STEP_S)) + 1
OUTPUT_S = 900.0
INFLOW_BASE = 10.0
TIDE_AMP = 0.8


def storm_length(raw_s: float) -> float:
    """Storm length on the knot grid: a multiple of two knots, so the hyetograph peak is a knot."""
    q = 2 * STEP_S
    return float(min(max(round(raw_s / q), 1) * q, MAX_STORM_S))


def triangle_knots(duration: float, peak: float, base: float) -> np.ndarray:
    t = np.arange(K_KNOTS) * STEP_S
    half = 0.5 * duration
    return base + (peak - base) * np.clip(1.0 - np.abs(t - half) / half, 0.0, None)


def request_forcing(req: dict) -> Forcing:
    """Rain and upstream inflow rising linearly to their peaks at mid-storm and back, tide in phase
    ``tide_phase``; one member, on the pool's knot grid."""
    m = req["member"]
    dur = storm_length(req["storm_s"])
    rain = PiecewiseLinear(0.0, STEP_S, triangle_knots(dur, m.rain_peak_mmh * W.MMH, 0.0)[None])
    inflow = PiecewiseLinear(0.0, STEP_S, triangle_knots(dur, m.inflow_peak, INFLOW_BASE)[None])
    tide = sine_tide(TIDE_AMP, W.TIDE_PERIOD_S, [m.tide_phase])
    return Forcing(rain, inflow, tide, dur)


def output_times(t_end: float) -> list:
    n = int(math.floor(t_end / OUTPUT_S + 1e-9))
    out = [k * OUTPUT_S for k in range(n + 1)]
    if t_end - out[-1] > 1e-9:
        out.append(t_end)
    return out


def member_fields(ter, m):
    land = ~ter.sea
    n = np.full(ter.shape, m.manning)
    S = np.where(land, m.retention_m, 0.0)
    f0 = np.where(land, m.f0_mmh * W.MMH, 0.0)
    fc = np.where(land, m.fc_mmh * W.MMH, 0.0)
    k = np.where(land, m.k_per_h / 3600.0, 0.0)
    return n, S, f0, fc, k


@dataclass
class GridSpec:
    nx: int
    ny: int
    dx: float
    network: str          # "chain" | "net60"
    seed: int = 20260917

    def build(self):
        ter = W.make_terrain(self.nx, self.ny, self.dx, self.seed)
        net = W.chain_network(ter, 20) if self.network == "chain" else W.network60(ter)
        return ter, net

    @property
    def slot_width_frac(self) -> float:
        return 1.0 if self.network == "chain" else 0.25


SERVING_GRIDS = (GridSpec(100, 81, 80.0, "chain"), GridSpec(200, 163, 40.0, "net60"),
                 GridSpec(400, 325, 20.0, "net60"))


def build_engine_for(spec: GridSpec, ter, net, members: list, forcing: Forcing, cfg: SolverConfig, device):
    cfg = cfg.replace(slot_width_frac=spec.slot_width_frac)
    flds = [member_fields(ter, m) for m in members]
    return engine_from_fields(cfg, ter, net, forcing, [f[0] for f in flds], [f[1] for f in flds],
                              [f[2] for f in flds], [f[3] for f in flds], [f[4] for f in flds],
                              [m.gamma for m in members], device=device, batched=True)


def request_engine(spec: GridSpec, ter, net, req: dict, cfg: SolverConfig, device) -> CoupledEngine:
    """The one-member engine of a request: what a standalone run of it builds."""
    return build_engine_for(spec, ter, net, [req["member"]], request_forcing(req), cfg, device)


def outlets(net, tp) -> tuple:
    """(cells, signs, main): the interior cells that drain through a stage end into the sea, the sign
    that makes their discharge positive seaward, and the index among them of the river mouth (the
    stage outlet reached downstream from the inflow reach)."""
    stage = np.flatnonzero(tp.kind == 3)
    cells, signs = tp.ghost_src[stage], tp.ghost_sign[stage].astype(np.float64)
    r = net.reach_index(net.inflow_reach) if net.inflow_reach else 0
    for _ in range(net.n_reaches):
        down = net.reaches[r].down
        if down[0] != "junction":
            break
        nxt = [k for k, q in enumerate(net.reaches) if q.up == down]
        if not nxt:
            break
        r = nxt[0]
    mouth = [k for k, c in enumerate(cells) if tp.reach_of[c] == r]
    return cells, signs, (mouth[0] if mouth else 0)


def sea_discharge(Q: np.ndarray, cells, signs, main: int):
    """(total channel discharge into the sea, the river mouth's) from a member's 1-D discharge."""
    q = Q[cells] * signs
    return float(q.sum()), float(q[main]) if len(q) else 0.0


def run_alone(spec: GridSpec, req: dict, cfg: SolverConfig, device):
    """A request solved by itself (its own engine), with the same outputs the server returns."""
    ter, net = spec.build()
    eng = request_engine(spec, ter, net, req, cfg, device)
    hmax = [None]
    cells, signs, main = outlets(net, eng.one.topo)
    hydro = []

    def snap(s):
        hmax[0] = s.h[0] if hmax[0] is None else np.maximum(hmax[0], s.h[0])
        hydro.append((s.t,) + sea_discharge(eng.one.Q[0].detach().cpu().numpy(), cells, signs, main))
    dur = storm_length(req["storm_s"])
    err = eng.run(dur, out_times=output_times(dur), on_snapshot=snap)
    return dict(h=eng.two.h[0].detach().cpu().numpy(), A=eng.one.A[0].detach().cpu().numpy(),
                depth_max=hmax[0], steps=eng.log.steps, mass_err=float(np.asarray(err).reshape(-1)[0]),
                hydrograph=hydro)


def slot_bytes(spec: GridSpec, precision: str = "fp32bulk") -> int:
    """HBM one busy slot of a grid size holds: 2-D state (5 fields), parameter fields (5), stage
    temporaries (3), the running maximum depth, and the 1-D network's state, sections and workspace."""
    ter, net = spec.build()
    from .solver.swe1d import build_topology
    item = 8 if precision == "fp64" else 4
    nc = build_topology(net).n_cells
    return ter.nx * ter.ny * item * (5 + 5 + 3 + 1) + nc * 8 * 12


class Pool:
    """The slots of one grid size: a B-member engine whose members are overwritten on admission."""

    def __init__(self, spec: GridSpec, slots: int, cfg: SolverConfig, device, seed: int):
        self.spec, self.B = spec, slots
        self.cfg = cfg.replace(slot_width_frac=spec.slot_width_frac)
        self.ter, self.net = spec.build()
        members = W.draw_members(slots, seed)
        zeros = np.zeros((slots, K_KNOTS))
        fo = Forcing(PiecewiseLinear(0.0, STEP_S, zeros), PiecewiseLinear(0.0, STEP_S, zeros + INFLOW_BASE),
                     sine_tide(TIDE_AMP, W.TIDE_PERIOD_S, [m.tide_phase for m in members]), 0.0)
        self.eng = build_engine_for(spec, self.ter, self.net, members, fo, cfg, device)
        self.free = deque(range(slots))
        two = self.eng.two
        self.cells = two.nx * two.ny
        self.hmax = torch.zeros_like(two.h)
        self.slot_bytes = slot_bytes(spec, cfg.precision)
        dev = two.device
        self.sea = torch.as_tensor(self.ter.sea, device=dev)
        self.land = ~self.sea
        tp = self.eng.one.topo
        self.out_cells, self.out_signs, self.out_main = outlets(self.net, tp)
        self.interior = self.eng.one.interior.reshape(1, -1)

    def load_request(self, b: int, req: dict, fo: Forcing) -> float:
        """Load a request into slot b straight from its member draw, with the operations a standalone
        build of it performs (``request_engine``), so the slot starts from the same bits: the 2-D
        parameter fields and state, the channel sections of its conveyance, the channel at normal depth
        (not below the tide).  Returns the storm's initial total volume (float64)."""
        from .domain.channels import apply_gamma
        from .engine import normal_depth_arrays
        from .solver.swe1d import build_topology, section_arrays
        m = req["member"]
        eng = self.eng
        two, one = eng.two, eng.one
        f64 = dict(dtype=torch.float64, device=two.device)
        zero = torch.zeros((), **f64)
        on_land = lambda v: torch.where(self.land, torch.tensor(float(v), **f64), zero)
        two.n[b].fill_(m.manning)
        two.storage[b].copy_(on_land(m.retention_m))
        two.horton.f0[b].copy_(on_land(m.f0_mmh * W.MMH))
        two.horton.fc[b].copy_(on_land(m.fc_mmh * W.MMH))
        two.horton.k[b].copy_(on_land(m.k_per_h / 3600.0))
        eta0 = np.asarray(fo.tide(0.0), dtype=np.float64).reshape(-1)
        eta0_2d = torch.as_tensor(eta0, dtype=two.dtype, device=two.device).reshape(1, 1)
        two.h[b].copy_(torch.where(self.sea, torch.clamp(eta0_2d - two.z, min=0.0), torch.zeros_like(two.z)))
        for name in ("hu", "hv", "s_filled", "infil_cum"):
            getattr(two, name)[b].zero_()
        # the member's network: conveyance-scaled sections, then the level of normal depth or the tide
        tp = build_topology(apply_gamma(self.net, m.gamma))
        z1, b1 = tp.z[None], tp.b[None]
        sa = section_arrays(tp, z1, b1, self.cfg.slot_width_frac)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), **f64)
        one.z[b].copy_(T(z1[0]))
        for sec, fields in ((one.sec, (("b", b1), ("yb", sa["yb"]), ("ts", sa["ts"]))),
                            (one.fsec, (("b", sa["fb"]), ("yb", sa["fyb"]), ("ts", sa["fts"])))):
            for name, val in fields:
                dst = getattr(sec, name)
                if dst.dim() == 2:
                    dst[b].copy_(T(val[0]))
                elif not torch.equal(dst, T(val[0])):
                    raise ValueError(f"section field {name} is shared in the pool but differs for the request")
        yn = normal_depth_arrays(tp, self.net.n_reaches, z1, b1, tp.m, sa["yb"], sa["ts"], tp.n,
                                 fo.inflow(0.0), (1,))
        zb = T(z1)
        level = torch.maximum(zb + T(yn), torch.as_tensor(eta0, **f64).reshape(1, 1).expand_as(zb))
        y = torch.clamp(level - zb, min=0.0)
        low, high = torch.minimum(y, T(sa["yb"])), torch.clamp(y - T(sa["yb"]), min=0.0)
        A = low * (T(b1) + T(tp.m)[None] * low) + T(sa["ts"]) * high
        one.A[b].copy_(A[0])
        one.Q[b].zero_()
        self.hmax[b].copy_(two.h[b])
        vol = two._volumes()
        v1 = (one.A[b:b + 1] * one.dx * self.interior).to(torch.float64).sum(-1)
        return float((vol[0, b:b + 1] + vol[1, b:b + 1] + v1)[0])

    def discharge_to_sea(self, members: list) -> dict:
        """{slot: (total seaward channel discharge, river mouth's)} for the given slots, one copy."""
        Q = self.eng.one.Q.detach().cpu().numpy()
        return {b: sea_discharge(Q[b], self.out_cells, self.out_signs, self.out_main) for b in members}


@dataclass
class Result:
    id: int
    size: int
    arrival_s: float
    admitted_s: float = 0.0
    done_s: float = 0.0
    t_end: float = 0.0
    steps: int = 0
    status: str = "ok"
    mass_err: float = 0.0
    max_depth_m: float = 0.0
    wet_cells: int = 0
    # (t, total channel discharge into the sea, the river mouth's) at every output time, t = 0 included
    hydrograph: list = field(default_factory=list)
    fields: dict = field(default_factory=dict)

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s


class Server:
    """Slot pools for the serving grid sizes in one device engine, fed by a request stream."""

    def __init__(self, slots=None, precision: str = "fp32bulk", hbm_budget_bytes: float | None = None,
                 device=None, grids=SERVING_GRIDS, chunk: int = 32, keep_fields: bool = False, tune: bool = True,
                 max_slots=(24, 16, 16)):
        """``hbm_budget_bytes`` (default: a quarter of the device's memory) bounds the HBM that busy
        slots may hold; a request is admitted only while it fits.  ``slots`` per grid size default to
        what an equal share of the budget holds, at most ``max_slots``.  ``chunk`` coupled steps run
        between two looks at the clocks (one 32-step graph replay by default)."""
        from .batching import GroupedEngine
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.cfg = SolverConfig().replace(precision=precision)
        self.grids = grids
        if hbm_budget_bytes is None:
            hbm_budget_bytes = 0.25 * torch.cuda.get_device_properties(self.device).total_memory
        self.budget = float(hbm_budget_bytes)
        if slots is None:
            share = self.budget / len(grids)
            slots = tuple(max(1, min(int(cap), int(share // slot_bytes(g, precision))))
                          for g, cap in zip(grids, max_slots))
        self.pools = [Pool(g, s, self.cfg, self.device, seed=100 + i) for i, (g, s) in enumerate(zip(grids, slots))]
        if tune:
            from . import tuning
            tuning.tune(self.pools[-1].eng.two)
        self.ge = GroupedEngine([p.eng for p in self.pools], dt_group_sizes=[1] * len(self.pools))
        self.dev = self.ge.dev
        self.m0 = self.ge.m0
        self.chunk = int(chunk)
        self.keep_fields = keep_fields
        Bt = self.dev.Btot
        self.t_limit = np.zeros(Bt)
        self.t_end = np.zeros(Bt)
        self.outs: list = [None] * Bt
        self.busy: dict = {}          # member index -> (pool, slot, Result, V0)
        self.resident = 0.0
        d = self.dev
        d.t.fill_(0.0)
        d.t_end.fill_(0.0)
        d.t_limit.fill_(0.0)
        d.refresh()

    # ------------------------------------------------------------ admission
    def _admit(self, pi: int, req: dict, now: float) -> bool:
        pool = self.pools[pi]
        if not pool.free or self.resident + pool.slot_bytes > self.budget:
            return False
        b = pool.free.popleft()
        m = self.m0[pi] + b
        fo = request_forcing(req)
        v0 = pool.load_request(b, req, fo)
        tab = self.dev.tables[pi]
        tab["rain"]["vals"][b].copy_(torch.as_tensor(fo.rain.values[0], device=self.device))
        tab["inflow"]["vals"][b].copy_(torch.as_tensor(fo.inflow.values[0], device=self.device))
        for key in ("stage", "east"):
            for name in ("base", "shift", "on"):
                src = getattr(fo.tide, name)[0]
                tab[key][name][b] = int(src) if name == "on" else float(src)
        d = self.dev
        for name, val in (("t", 0.0), ("dt", 0.0), ("dt_max_seen", 0.0), ("dt_min_seen", math.inf)):
            getattr(d, name)[m] = val
        for name in ("n_sub", "steps", "clamped", "n_sub_max", "error"):
            getattr(d, name)[m] = 0
        d.acc[:, m] = 0.0
        outs = output_times(fo.t_end)[1:]
        self.outs[m] = deque(outs)
        self.t_end[m] = fo.t_end
        self.t_limit[m] = outs[0]
        d.t_end[m] = fo.t_end
        d.t_limit[m] = outs[0]
        mask = torch.zeros(d.Btot, dtype=torch.int32, device=self.device)
        mask[m] = 1
        d.refresh(mask)
        res = Result(req["id"], pi, req["arrival_s"], admitted_s=now, t_end=fo.t_end)
        res.hydrograph.append((0.0, 0.0, 0.0))     # channels start at rest (Q = 0)
        self.busy[m] = (pi, b, res, v0)
        self.resident += pool.slot_bytes
        return True

    # ------------------------------------------------------------ completion
    def _output(self, members: list, now: float, done: list):
        """Members at an output time: fold depth maxima, check plausibility, finish or move on."""
        pl, qs = {}, {}
        t_now = self.dev.t.cpu().numpy()
        for pi, pool in enumerate(self.pools):
            mine = [self.busy[m][1] for m in members if self.busy[m][0] == pi]
            if mine:
                pl[pi] = pool.eng.plausibility()
                qs[pi] = pool.discharge_to_sea(mine)
        finished = []
        for m in members:
            pi, b, res, v0 = self.busy[m]
            pool = self.pools[pi]
            two = pool.eng.two
            torch.maximum(pool.hmax[b], two.h[b], out=pool.hmax[b])
            res.hydrograph.append((float(t_now[m]),) + qs[pi][b])
            dmax, rise = pl[pi][0][b], pl[pi][1][b]
            if not (np.isfinite(dmax) and np.isfinite(rise) and dmax <= MAX_PLAUSIBLE_DEPTH_2D_M
                    and rise <= MAX_PLAUSIBLE_RISE_1D_M):
                res.status = f"implausible: max depth {dmax:.2f} m, 1-D rise {rise:.2f} m"
                finished.append(m)
                continue
            q = self.outs[m]
            q.popleft()
            if q:
                self.t_limit[m] = q[0]
                self.dev.t_limit[m] = q[0]
            else:
                finished.append(m)
        if finished:
            self._finish(finished, now, done)

    def _finish(self, members: list, now: float, done: list):
        d = self.dev
        acc = d.acc.cpu().numpy()
        steps = d.steps.cpu().numpy()
        vols = {}
        for m in members:
            pi, b, res, v0 = self.busy.pop(m)
            pool = self.pools[pi]
            if pi not in vols:
                e = pool.eng
                vols[pi] = (e.two.volume() + e.two.storage_volume() + e.one.volume()).cpu().numpy()
            a = acc[:, m]
            net = a[0] - a[1] + a[2] - a[3] + a[4] - a[5]
            gross = a[0] + a[1] + a[2] + a[3] + a[4] + a[5]
            res.mass_err = float(abs(vols[pi][b] - v0 - net) / max(gross, 1e-9 * abs(v0) + 1.0))
            if res.status == "ok" and not res.mass_err < self.cfg.mass_tol:
                res.status = f"mass balance {res.mass_err:.2e}"
            res.steps = int(steps[m])
            hm = pool.hmax[b]
            res.max_depth_m = float(hm.max())
            res.wet_cells = int((hm > 0.10).sum())
            if self.keep_fields:
                res.fields = dict(h=pool.eng.two.h[b].detach().cpu().numpy(),
                                  A=pool.eng.one.A[b].detach().cpu().numpy(),
                                  depth_max=hm.detach().cpu().numpy())
            # park the slot: its clock at its end, nothing more to do
            d.t_end[m] = 0.0
            d.t_limit[m] = 0.0
            self.t_end[m] = self.t_limit[m] = 0.0
            res.done_s = now
            pool.free.append(b)
            self.resident -= pool.slot_bytes
            done.append(res)

    # ------------------------------------------------------------ loop
    def serve(self, requests: list, realtime: bool = True, log_every_s: float = 20.0) -> list:
        """Serve requests (dicts with id, arrival_s, size, storm_s, member); returns Results in
        completion order.  ``realtime=False`` treats every request as arrived at t = 0."""
        pending = deque(sorted(requests, key=lambda r: r["arrival_s"]))
        queues = [deque() for _ in self.pools]
        done: list = []
        d = self.dev
        start = time.perf_counter()
        last_log = start
        n = len(requests)
        while len(done) < n:
            now = time.perf_counter() - start
            while pending and (not realtime or pending[0]["arrival_s"] <= now):
                r = pending.popleft()
                queues[r["size"]].append(r)
            for pi, q in enumerate(queues):
                while q and self._admit(pi, q[0], now):
                    q.popleft()
            if not self.busy:
                if not pending:
                    if any(queues):
                        raise RuntimeError("requests are queued but none fits the HBM budget")
                    break
                time.sleep(max(0.0, min(pending[0]["arrival_s"] - (time.perf_counter() - start), 0.05)))
                continue
            d.step(self.chunk)
            t = d.t.cpu().numpy()
            err = d.error.cpu().numpy()
            now = time.perf_counter() - start
            bad = [m for m in self.busy if err[m]]
            for m in bad:
                self.busy[m][2].status = "unstable: " + ERROR_CODES.get(int(err[m]), f"error {int(err[m])}")
            if bad:
                self._finish(bad, now, done)
            at_out = [m for m in self.busy if t[m] >= self.t_limit[m] - 1e-9]
            if at_out:
                self._output(at_out, now, done)
            if log_every_s and time.perf_counter() - last_log >= log_every_s:
                last_log = time.perf_counter()
                print(f"[serve] {now:.0f} s: {len(done)}/{n} done, {len(self.busy)} running, "
                      f"{sum(len(q) for q in queues)} queued", flush=True)
        return done

    def close(self):
        self.ge.close()


def summarize(results: list, wall_s: float, grids=SERVING_GRIDS) -> dict:
    lat = np.array([r.latency_s for r in results])
    cells = np.array([grids[r.size].nx * grids[r.size].ny for r in results])
    steps = np.array([r.steps for r in results])
    ok = [r for r in results if r.status == "ok"]
    return dict(requests=len(results), ok=len(ok), failed=len(results) - len(ok),
                latency_p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                latency_p99_s=float(np.percentile(lat, 99)) if len(lat) else None,
                latency_max_s=float(lat.max()) if len(lat) else None,
                wall_s=wall_s, cell_updates=int((cells * steps).sum()),
                cell_updates_per_s=float((cells * steps).sum() / wall_s) if wall_s > 0 else None,
                worst_mass_err=float(max((r.mass_err for r in ok), default=0.0)),
                per_size={i: int(sum(1 for r in results if r.size == i)) for i in range(len(grids))})


def requests_for(n: int, rate_per_s: float = 4.0, seed: int = 5) -> list:
    return W.serving_requests(n, rate_per_s, seed)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--precision", default="fp32bulk")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    reqs = requests_for(a.requests, a.rate)
    srv = Server(precision=a.precision)
    w0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - w0
    summ = summarize(res, wall)
    summ["rate_per_s"] = a.rate
    summ["precision"] = a.precision
    print(json.dumps(summ), flush=True)
    if a.out:
        from pathlib import Path
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(dict(summary=summ, requests=[
            dict(id=r.id, size=r.size, arrival_s=r.arrival_s, admitted_s=r.admitted_s, done_s=r.done_s,
                 latency_s=r.latency_s, t_end_s=r.t_end, steps=r.steps, status=r.status, mass_err=r.mass_err,
                 max_depth_m=r.max_depth_m, wet_cells=r.wet_cells, hydrograph_samples=len(r.hydrograph),
                 peak_sea_discharge_m3s=max((q[1] for q in r.hydrograph), default=0.0),
                 peak_mouth_discharge_m3s=max((q[2] for q in r.hydrograph), default=0.0))
            for r in res]), indent=0))


if __name__ == "__main__":
    main()
