"""Autotuner for the 2-D kernel shape.

For a grid (dtype, nx, ny, B) it times every compiled kernel shape -- LDS-tiled
workgroups of several tile sizes and column-marching waves of several chunk
heights -- on the fused 2-D step with HIP events, and keeps the fastest.  Results
persist as JSON keyed by "dtype:nx:ny:B" in $HIS_AMD_TUNING (default
~/.cache/his_amd/tuning.json) so a later process reuses them without timing.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np
import torch

from . import _lib as L



def _path() -> Path:
    return Path(os.environ.get("HIS_AMD_TUNING", Path.home() / ".cache" / "his_amd" / "tuning.json"))


def _load() -> dict:
    p = _path()
    try:
        return json.loads(p.read_text())
    except Exception:
        return {}


def _save(tab: dict) -> None:
    p = _path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(tab, indent=1, sort_keys=True))
    except OSError:
        pass


def key(dtype, nx, ny, B) -> str:
    return f"{'f64' if dtype == torch.float64 else 'f32'}:{nx}:{ny}:{B}"


def time_shapes(two, reps: int = 20, passes: int = 2) -> dict:
    """Microseconds of the 2-D step for every shape, on a copy of ``two``'s state: the mean over
    ``reps`` steps, the smallest such mean over ``passes`` sweeps of all shapes (clock and
    neighbour drift between sweeps otherwise decide between shapes a few percent apart)."""
    two._sync_state()
    st = (two.h.clone(), two.hu.clone(), two.hv.clone(), two.s_filled.clone(), two.infil_cum.clone())
    pm = two._pack([0.0] * 9)
    dt_t = two._pack([0.05, 0.0])
    out = {}
    for _ in range(passes):
        for shape in range(L.num_tile_shapes()):
            L.set_tile_shape(two.dtype, shape)
            for _ in range(3):
                two._run(*st, dt_t, pm)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                two._run(*st, dt_t, pm)
            e1.record()
            torch.cuda.synchronize()
            us = 1e3 * e0.elapsed_time(e1) / reps
            out[shape] = min(us, out.get(shape, us))
    return out


def tune(two, reps: int = 20, force: bool = False) -> int:
    """Select (and remember) the fastest shape for this solver's grid; sets it active."""
    k = key(two.dtype, two.nx, two.ny, two.B)
    tab = _load()
    if not force and k in tab:
        best = int(tab[k]["shape"])
    else:
        times = time_shapes(two, reps)
        best = int(min(times, key=times.get))
        tab[k] = {"shape": best, "us": times[best], "all_us": {str(s): t for s, t in times.items()}}
        _save(tab)
    L.set_tile_shape(two.dtype, best)
    return best


def tune_coupled(eng, steps: int = 60, warm: int = 4) -> tuple[int, dict]:
    """Select the shape whose coupled step is fastest: the engine's own graph-replayed step, with
    the 1-D network on its side stream beside the 2-D stages, which holds CUs the 2-D waves would
    otherwise take (a shape whose waves fill the chip in one round alone can need a second round
    beside it).  With a network, both capture orders of the two branches are timed for each shape
    (the branch captured second pays a cross-queue start).  The timing steps are real steps of
    the run: the storm advances ``warm + steps`` steps per candidate, and the engine's ledger and
    step count take them in.  Sets the best shape and order and returns the shape with the
    milliseconds per coupled step of every shape (its better order)."""
    two = eng.two
    dev = eng._device()
    dev.refresh()
    t_end0, t_lim0 = dev.t_end.clone(), dev.t_limit.clone()
    dev.t_end.fill_(1e12)
    dev.t_limit.fill_(1e12)
    s0 = int(dev.steps[0].item())
    orders = (0, 1) if getattr(eng, "one", None) is not None else (-1,)
    times, order_of = {}, {}
    try:
        for shape in range(L.num_tile_shapes()):
            L.set_tile_shape(two.dtype, shape)
            for order in orders:
                dev.set_branch_order(order)
                dev.step(warm)               # the engine takes the new tiles and recaptures its graphs
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                dev.step(steps)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / steps
                if shape not in times or ms < times[shape]:
                    times[shape], order_of[shape] = ms, order
    finally:
        dev.t_end.copy_(t_end0)
        dev.t_limit.copy_(t_lim0)
    eng.log.steps += int(dev.steps[0].item()) - s0
    eng._fold_device_ledger()
    best = int(min(times, key=times.get))
    L.set_tile_shape(two.dtype, best)
    dev.set_branch_order(order_of[best])
    tab = _load()
    tab[key(two.dtype, two.nx, two.ny, two.B)] = {"shape": best, "us": 1e3 * times[best], "coupled": True,
                                                  "branch_order": order_of[best],
                                                  "all_us": {str(s): 1e3 * t for s, t in times.items()}}
    _save(tab)
    eng.branch_order = order_of[best]
    return best, times


def describe(dtype, shape: int) -> str:
    tx, ty, var = L.tile_shape(dtype, shape)
    return f"{('tile', 'march', 'wave')[var]} {tx}x{ty}"
