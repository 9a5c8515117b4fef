"""Accuracy studies of the float32 bulk engine against the float64 engine.

precision_study: the same members run in float64 (the reference's arithmetic,
checked against the reference solver and the CPU oracle) and in float32 bulk, over
the whole storm.  Reported per member, at the 15-minute output times (the
FloodResult.depth_max semantics, hydrointel/api.py:239):
  * max-depth error: max over cells of |peak depth fp32 - peak depth fp64| (all cells
    and land cells), and the maximum where both runs are deeper than 0.10 m
  * peak-time error: largest shift of the output time at which a cell peaks, over cells
    whose fp64 peak exceeds 0.10 m
  * volume error: largest relative difference of surface + storage + channel volume
  * mass balance error of each run (relative, per member)
  * the agreement criteria of hydrointel/criteria.py:40-100 (depth RMSE, p99, ...)

    python -m his_amd.studies precision --members 32 [--hours H] [--out results/precision.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import workloads as W
from .criteria import agreement
from .engine import build_from_workload


def run_storm(wl: W.Workload, precision: str, t_end: float | None = None, interval: float = 900.0):
    """Run a workload at a precision; return peak depth, peak time index, volumes, mass error, log."""
    wl = W.Workload(wl.name, wl.terrain, wl.network, wl.members, wl.storm, precision, wl.slot_width_frac)
    eng, fo = build_from_workload(wl, device=torch.device("cuda"), batched=True)
    from . import tuning
    tuning.tune(eng.two)
    t_end = fo.t_end if t_end is None else t_end
    times = np.arange(0.0, t_end + 1e-9, interval)
    peaks, when, vols, series = None, None, [], []

    w_start = time.perf_counter()

    def snap(s):
        nonlocal peaks, when
        print(f"[{precision}] t = {s.t / 3600:.2f} h, {time.perf_counter() - w_start:.1f} s", file=sys.stderr,
              flush=True)
        h = s.h.astype(np.float64)
        if peaks is None:
            peaks = h.copy()
            when = np.zeros(h.shape, np.int32)
        else:
            up = h > peaks
            peaks = np.where(up, h, peaks)
            when = np.where(up, len(vols), when)
        led = s.ledger
        vols.append(np.asarray(led["v2d"]) + np.asarray(led["storage"]) + np.asarray(led["v1d"]))
        series.append(h[:, ::4, ::4].astype(np.float32))
    w0 = time.perf_counter()
    err = eng.run(t_end, out_times=times, on_snapshot=snap, progress_s=0)
    wall = time.perf_counter() - w0
    return dict(peak=peaks, when=when, vol=np.stack(vols), mass_err=np.asarray(err, dtype=float).reshape(-1),
                steps=eng.log.steps, wall_s=wall, series=np.stack(series), times=times)


def precision_study(members: int = 32, hours: float | None = None, seed: int = 20260917) -> dict:
    wl = W.precision_form(seed=seed, B=members)
    t_end = None if hours is None else hours * 3600.0
    r64 = run_storm(wl, "fp64", t_end)
    r32 = run_storm(wl, "fp32bulk", t_end)
    land = ~wl.terrain.sea
    out = dict(workload=wl.name, members=members, t_end_s=float(r64["times"][-1]), steps_fp64=r64["steps"],
               steps_fp32=r32["steps"], wall_fp64_s=r64["wall_s"], wall_fp32_s=r32["wall_s"], per_member=[])
    cells = wl.terrain.nx * wl.terrain.ny
    out["cell_updates_per_s_fp64"] = members * cells * r64["steps"] / r64["wall_s"]
    out["cell_updates_per_s_fp32"] = members * cells * r32["steps"] / r32["wall_s"]
    for b in range(members):
        p64, p32 = r64["peak"][b], r32["peak"][b]
        d = np.abs(p32 - p64)
        deep = np.minimum(p32, p64) > 0.10
        wet = p64 > 0.10
        dt_peak = np.abs(r32["when"][b] - r64["when"][b])[wet] * 900.0 if wet.any() else np.zeros(1)
        v64, v32 = r64["vol"][:, b], r32["vol"][:, b]
        agr = agreement(r64["series"][:, b], r32["series"][:, b], land[::4, ::4], wl.terrain.dx * 4, 1e-3)
        out["per_member"].append(dict(
            member=b, max_depth_err_m=float(d.max()), max_depth_err_land_m=float(d[land].max()),
            max_depth_err_both_deep_m=float(d[deep].max()) if deep.any() else 0.0,
            p99_depth_err_m=float(np.percentile(d[wet], 99)) if wet.any() else 0.0,
            peak_time_shift_max_s=float(dt_peak.max()), peak_time_shift_p99_s=float(np.percentile(dt_peak, 99)),
            volume_rel_err_max=float(np.max(np.abs(v32 - v64) / np.maximum(np.abs(v64), 1.0))),
            mass_err_fp64=float(r64["mass_err"][b]), mass_err_fp32=float(r32["mass_err"][b]),
            agreement_passed=bool(agr["passed"]),
            agreement={k: v["value"] for k, v in agr["checks"].items()}))
    pm = out["per_member"]
    out["summary"] = dict(
        max_depth_err_m=max(x["max_depth_err_m"] for x in pm),
        max_depth_err_land_m=max(x["max_depth_err_land_m"] for x in pm),
        max_depth_err_both_deep_m=max(x["max_depth_err_both_deep_m"] for x in pm),
        worst_mass_err_fp32=max(x["mass_err_fp32"] for x in pm),
        worst_mass_err_fp64=max(x["mass_err_fp64"] for x in pm),
        worst_volume_rel_err=max(x["volume_rel_err_max"] for x in pm),
        agreement_passed=sum(x["agreement_passed"] for x in pm))
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("precision")
    p.add_argument("--members", type=int, default=32)
    p.add_argument("--hours", type=float, default=None)
    p.add_argument("--out", default=None)
    p.add_argument("--tag", default=None, help="recorded in the output (for example the source revision)")
    a = ap.parse_args()
    if a.cmd == "precision":
        r = precision_study(a.members, a.hours)
        r["tag"] = a.tag
        txt = json.dumps(r, indent=1)
        if a.out:
            from pathlib import Path
            Path(a.out).parent.mkdir(parents=True, exist_ok=True)
            Path(a.out).write_text(txt)
        print(json.dumps(r["summary"]), flush=True)


if __name__ == "__main__":
    main()
