"""Several uneven grids advanced by one device engine.

The reference batches storms over one terrain (hydrointel/solver/swe2d.py:114-145 with
``batch=B``; hydrointel/engine.py:169-203) and runs different terrains one after the
other. Here up to four grid groups share one ``his_engine``:
* every 2-D stage is ONE kernel launch covering every group's tiles;
* every group keeps its own dt, shared by its members as in the reference, or per
  member with ``dt_group_size=1``;
* one hipGraph replays the whole coupled step of all groups.

Each group's arithmetic is exactly the arithmetic of that group run alone: the grouping
changes which launch a tile belongs to, never what the tile computes.

    ge = GroupedEngine([eng_a, eng_b, eng_c])       # CoupledEngines built as usual
    ge.run(t_end, out_times=..., on_snapshot=lambda gi, snap: ...)
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .solver.coupling import ERROR_CODES, CoupledEngine, DeviceEngine, NumericalInstabilityError
from . import _lib as L


class GroupedEngine:
    def __init__(self, engines: list[CoupledEngine], dt_group_sizes: list[int] | None = None,
                 use_graph: bool = True):
        if not 1 <= len(engines) <= L.MAX_GROUPS:
            raise ValueError(f"1 to {L.MAX_GROUPS} grid groups per engine, got {len(engines)}")
        cfg = engines[0].cfg
        for e in engines[1:]:
            if e.cfg.dtype != cfg.dtype or e.cfg.spatial_order != cfg.spatial_order:
                raise ValueError("every group of an engine must share precision and spatial order")
            if (e.cfg.cfl, e.cfg.dt_min, e.cfg.dt_max) != (cfg.cfl, cfg.dt_min, cfg.dt_max):
                raise ValueError("every group of an engine must share cfl, dt_min and dt_max")
        for e in engines:
            if not e.device_ready():
                raise ValueError("a group's forcing cannot be evaluated on the device")
        self.engines = engines
        self.cfg = cfg
        sizes = dt_group_sizes or [e.two.B for e in engines]
        parts = []
        for e, k in zip(engines, sizes):
            e.two._sync_state()
            if e.one is not None:
                e.one._sync_state()
            parts.append(dict(two=e.two, one=e.one, links=e.links, rain=e.rain, dt_group_size=int(k)))
        self.dev = DeviceEngine(parts, cfg, use_graph=use_graph)
        self.m0 = list(self.dev.member0)
        self.cells = [e.two.B * e.two.nx * e.two.ny for e in engines]

    @property
    def member_cells(self) -> int:
        return sum(self.cells)

    def _fold(self):
        for e, m0 in zip(self.engines, self.m0):
            e._fold_device_ledger(self.dev, m0)

    def _check_errors(self, t: float, label: str):
        err = self.dev.error.cpu().numpy()
        if err.any():
            why = ERROR_CODES
            bad = []
            for gi, m0 in enumerate(self.m0):
                for b in range(self.engines[gi].two.B):
                    if err[m0 + b]:
                        bad.append(f"group {gi} member {b}: {why.get(int(err[m0 + b]), err[m0 + b])}")
            raise NumericalInstabilityError(f"unstable state{(' in ' + label) if label else ''} near "
                                            f"t = {t / 3600:.3f} h: " + ", ".join(bad[:8]) + "; this run cannot be used")

    def steps(self) -> list[int]:
        s = self.dev.steps.cpu().numpy()
        return [int(s[m0]) for m0 in self.m0]

    def run(self, t_end: float, out_times=(), on_snapshot=None, check_every: int = 500, label: str = "",
            chunk: int | None = None) -> list:
        """Advance every group from t = 0 to t_end, stopping each member at every output time;
        ``on_snapshot(group, Snapshot)`` at each.  Returns each group's per-member mass error."""
        dev = self.dev
        dev.t.fill_(0.0)
        dev.t_end.fill_(t_end)
        dev.refresh()
        outs = sorted(float(x) for x in out_times if 0.0 <= x <= t_end + 1e-9)
        if outs and outs[0] < 1e-9:
            if on_snapshot:
                for gi, e in enumerate(self.engines):
                    on_snapshot(gi, e.snapshot(0.0))
            outs.pop(0)
        steps0 = self.steps()
        wall = time.perf_counter()
        taken_check = 0
        t_min = 0.0
        while t_min < t_end - 1e-9:
            limit = outs[0] if outs else t_end
            dev.t_limit.fill_(limit)
            while True:
                n = chunk or 64
                dev.step(n)
                t_all = dev.t.cpu().numpy()
                t_min = float(t_all.min())
                self._check_errors(t_min, label)
                taken = max(a - b for a, b in zip(self.steps(), steps0))
                if taken - taken_check >= check_every:
                    taken_check = taken
                    self._fold()
                    for gi, e in enumerate(self.engines):
                        e.mb.check(e.total_volume(), t_min)
                        e.check_plausible(t_min, f"{label} group {gi}".strip())
                if t_min >= limit - 1e-9:
                    break
            if outs and t_min >= outs[0] - 1e-9:
                self._fold()
                for gi, e in enumerate(self.engines):
                    e.check_plausible(t_min, f"{label} group {gi}".strip())
                    if on_snapshot:
                        on_snapshot(gi, e.snapshot(t_min))
                outs.pop(0)
        self._fold()
        errs = []
        done = self.steps()
        for gi, e in enumerate(self.engines):
            e.log.steps += done[gi] - steps0[gi]
            e.log.wall_s += time.perf_counter() - wall
            e.check_plausible(t_end, f"{label} group {gi}".strip())
            errs.append(e.mb.assert_ok(e.total_volume(), t_end, e.cfg.mass_tol, f"{label} group {gi}".strip()))
        return errs

    def close(self):
        self.dev.close()


def build_batching_form(workloads=None, device=None, use_graph: bool = True, tune: bool = False):
    """The three-resolution batch (100x81 @ 80 m x16, 200x163 @ 40 m x8, 400x325 @ 20 m x8,
    float32 bulk) as one GroupedEngine; returns (grouped engine, forcings).  ``tune`` picks the
    tile shape on the largest grid first (one shape serves every group of a launch)."""
    from . import workloads as W
    from .engine import build_from_workload
    wls = workloads if workloads is not None else W.batching_form()
    dev = torch.device(device) if device is not None else torch.device("cuda")
    built = [build_from_workload(wl, device=dev, batched=True) for wl in wls]
    engines = [b[0] for b in built]
    if tune:
        from . import tuning
        tuning.tune(max(engines, key=lambda e: e.two.B * e.two.nx * e.two.ny).two)
    return GroupedEngine(engines, use_graph=use_graph), [b[1] for b in built]


def member_cell_updates(ge: GroupedEngine, steps: list[int]) -> int:
    return int(sum(c * s for c, s in zip(ge.cells, steps)))

