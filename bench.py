"""Throughput of the coupled flood-engine step on one MI355X.

Workload (the headline): 400 x 325 cells at 20 m, 32 storm members sharing terrain
and dt, float32 bulk with float64 accumulators, 1-D network of 60 reaches and weir
exchange coupled in.  A step is one coupled step of the whole batch (CFL dt, exchange,
2-D SSP-RK2 step with sources and friction, 1-D substeps), replayed from a hipGraph;
value = member-cells x steps / wall time, timed from a wet state in the middle of the
storm (t = 3 h, the rain peak; --t-start).  Beside it: the same grid at B=1 in float64
(the plain-GPU path), the batching and serving forms, the 2-D kernel's achieved HBM
bandwidth and VALU issue rate from HIP events, the whole storm in float32 bulk and in
float64 with the float32 errors against float64 (peak depth, peak time, volume, mass),
and the float64 CPU oracle timed on this host's cores on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s at 400x325 B=32, FP32-bulk vs FP64; p50/p99 storm latency"
HBM_PEAK_GBS = 8000.0
# read h, hu, hv, s, infil_cum, n, S, f0, fc, k, z; write h, hu, hv, s, infil_cum: 16 values per
# member-cell-step (infil_cum is state the step reads and writes, so it counts)
ALG_BYTES_F32 = 64
ALG_BYTES_F64 = 128
# VALU issue peak: one wave64 instruction per 4 cycles on each of 1,024 SIMDs at 2.4 GHz
VALU_PEAK_WAVE_INSTR_PER_S = 1024 * 2.4e9 / 4
TUNE_LEAD_S = 1800.0        # the shape is tuned on coupled steps from t_start - 30 min
T_START_S = 3 * 3600.0      # headline window starts at the rain peak, with the plain wet


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def engine_throughput(wl, steps: int, warmup: int, t_start: float):
    """Coupled steps of a workload timed from t_start (the storm run there first, untimed)."""
    import torch
    from his_amd.engine import build_from_workload
    dev = torch.device("cuda")
    eng, fo = build_from_workload(wl, device=dev)
    from his_amd import tuning
    # the kernel shape and capture order are chosen on coupled steps of the wet storm half an hour
    # before the timed window (the timing steps are steps of the storm), then the storm runs on
    t_tune = max(0.0, t_start - TUNE_LEAD_S)
    if t_tune > 0:
        eng.run(t_tune, out_times=[], progress_s=0)
    shape, times = tuning.tune_coupled(eng)
    de = eng._device()
    t_now = float(de.t[0].item())
    log(f"[bench]   kernel shape {shape} ({tuning.describe(eng.two.dtype, shape)}), branch captured first: "
        f"{('2-D', '1-D')[eng.branch_order] if eng.branch_order >= 0 else 'estimate'}, coupled step per shape: "
        + ", ".join(f"{s}:{1e3 * t:.0f}us" for s, t in times.items()) + f"; tuned over t = {t_tune:.0f}..{t_now:.0f} s")
    if t_now < t_start:
        eng.run(t_start, out_times=[], t0=t_now, progress_s=0)
    de.t.fill_(t_start)
    de.t_end.fill_(1e12)
    de.t_limit.fill_(1e12)
    de.refresh()
    s0 = int(de.steps[0].item())
    de.step(warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record()
    de.step(steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_ev = e0.elapsed_time(e1)
    taken = int(de.steps[0].item()) - s0
    if taken != warmup + steps:
        raise RuntimeError(f"engine took {taken} steps, expected {warmup + steps}")
    t_sim = float(de.t[0].item())
    n_sub_max = int(de.n_sub_max.max().item())     # over the warm-up and the timed window
    eng.log.steps += taken
    eng._fold_device_ledger()
    err = eng.mb.check(eng.total_volume(), t_sim)
    eng.check_plausible(t_sim, "bench")
    cells = wl.batch * wl.cells
    return dict(cells=cells, wall_s=wall, event_ms=ms_ev, rate=cells * steps / wall, t_sim=t_sim, t_start=t_start,
                n_sub_max=n_sub_max,
                shape=tuning.describe(eng.two.dtype, shape),
                mass_err=float(max(abs(float(x)) for x in (err if hasattr(err, "__len__") else [err]))),
                ms_per_step=1e3 * wall / steps, eng=eng)


def batching_throughput(steps: int, warmup: int, t_start: float):
    """Three uneven grids (16 x 100x81 @ 80 m, 8 x 200x163 @ 40 m, 8 x 400x325 @ 20 m, float32
    bulk) in one engine: one launch per 2-D stage over all of them, one dt per grid group; timed
    from t_start (every group run there first, untimed)."""
    import torch
    from his_amd.batching import build_batching_form, member_cell_updates
    ge, _ = build_batching_form(device=torch.device("cuda"), tune=True)
    de = ge.dev
    if t_start > 0:
        ge.run(t_start)
    de.t.fill_(t_start)
    de.t_end.fill_(1e12)
    de.t_limit.fill_(1e12)
    de.refresh()
    s0 = ge.steps()
    de.step(warmup)
    torch.cuda.synchronize()
    s1 = ge.steps()
    w0 = time.perf_counter()
    de.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    s2 = ge.steps()
    if any(b - a != warmup for a, b in zip(s0, s1)) or any(b - a != steps for a, b in zip(s1, s2)):
        raise RuntimeError(f"grouped engine took {s1} / {s2} steps from {s0}")
    ge._fold()
    t_sim = de.t.cpu().numpy()
    errs = []
    for gi, e in enumerate(ge.engines):
        m0 = ge.m0[gi]
        t_g = float(t_sim[m0])
        errs.append(float(np.max(np.abs(np.asarray(e.mb.check(e.total_volume(), t_g))))))
        e.check_plausible(t_g, f"bench group {gi}")
    cu = member_cell_updates(ge, [steps] * len(ge.engines))
    ge.close()
    return dict(rate=cu / wall, ms_per_step=1e3 * wall / steps, member_cells=ge.member_cells,
                mass_err=max(errs), groups=[f"{e.two.B}x{e.two.nx}x{e.two.ny}" for e in ge.engines])


def serving_latency(n: int):
    """The first n requests of the seeded Poisson(4/s) stream (grid size uniform over 100x81 @ 80 m,
    200x163 @ 40 m, 400x325 @ 20 m; storm U[0.5, 6] h), served in real time; latency = completion -
    arrival."""
    from his_amd import serving as S
    reqs = S.requests_for(n)
    server = S.Server()
    w0 = time.perf_counter()
    res = server.serve(reqs, log_every_s=0)
    wall = time.perf_counter() - w0
    server.close()
    out = S.summarize(res, wall)
    out["sample"] = f"first {n} of the 2000-request stream (seed 5)"
    return out


def kernel_roofline(eng, reps: int, alg_bytes: int):
    """Time the fused 2-D step (two stage kernels + finalize) with HIP events on the stream the
    kernels are launched on, from the engine's current state (copies of it, so the run is untouched)."""
    import torch
    two = eng.two
    dt = 0.05
    pm = two._pack([0.0] + [0.0] * 8)
    two._sync_state()
    de = eng._dev
    t_now = float(de.t[0].item()) if de is not None else 0.0
    dt_t = two._pack([dt, t_now])
    st = (two.h.clone(), two.hu.clone(), two.hv.clone(), two.s_filled.clone(), two.infil_cum.clone())
    for _ in range(5):
        two._run(*st, dt_t, pm)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        two._run(*st, dt_t, pm)
    e1.record()
    torch.cuda.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / reps
    cells = two.B * two.nx * two.ny
    gbs = alg_bytes * cells / (us * 1e-6) / 1e9
    return us, gbs


def cpu_baseline(budget_s: float = 20.0):
    """The float64 oracle (C, OpenMP over members) on the headline workload, bounded sample."""
    import numpy as np
    from his_amd import workloads as W
    from his_amd.domain.channels import apply_gamma
    from oracle.oracle import Oracle2D, OracleCoupled, OracleNet, oracle_links, normal_depth
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    wl = W.precision_form()
    ter = wl.terrain
    n, S, f0, fc, k = wl.fields()
    rain, tide, inflow = wl.forcing()
    B = wl.batch
    two = Oracle2D(ter.z, n, ter.dx, dict(west="transmissive", east="stage", south="transmissive",
                                          north="transmissive"), B=B, storage=S, horton=(f0, fc, k, 0.0))
    one = OracleNet([apply_gamma(wl.network, m.gamma) for m in wl.members], wl.slot_width_frac)
    orc = OracleCoupled(two, one, oracle_links(one, ter.z, ter.dx), dict(east=tide), rain, inflow, tide)
    eta0 = tide(0.0)
    orc.h = np.where(ter.sea[None], np.clip(eta0.reshape(B, 1, 1) - ter.z[None], 0, None), 0.0)
    yn = normal_depth(one, inflow(0.0))
    orc.A = one.area(np.clip(np.maximum(one.arr["z"] + yn, eta0.reshape(B, 1)) - one.arr["z"], 0, None))
    t, steps = 0.0, 0
    w0 = time.perf_counter()
    while time.perf_counter() - w0 < budget_s:
        t = orc.step(t, 1e12)
        steps += 1
    wall = time.perf_counter() - w0
    return dict(value=B * wl.cells * steps / wall, unit="cell-updates/s", cores=cores, kind="port",
                sample=f"float64 oracle, {steps} coupled steps of the 400x325 B=32 workload from t=0 "
                       f"({wall:.1f} s wall, OpenMP over members with {cores} threads)")


TRAFFIC_FILE = ROOT / "profiles" / "swe2d_step_f32_400x325_b32_traffic.md"
VALU_FILE = ROOT / "profiles" / "swe2d_step_f32_400x325_b32_valu.md"


def _profile_figure(path, key):
    """A figure from a committed rocprofv3 summary (a line 'key: N'), or None."""
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def traffic_from_profile():
    """HBM bytes per 2-D step from the FETCH_SIZE / WRITE_SIZE summary."""
    return _profile_figure(TRAFFIC_FILE, "traffic_bytes_per_step")


def precision_vs_fp64():
    """The whole storm of the headline workload in float32 bulk and in float64 (same members, same
    terrain, each with its own dt): per member, the float32 run's errors against float64 at the
    15-minute output times -- peak depth, peak time, volume -- and both runs' mass balance."""
    import numpy as np
    from his_amd import studies
    r = studies.precision_study(32)
    pm = r["per_member"]
    s = r["summary"]
    return dict(t_end_s=r["t_end_s"], steps_fp32=r["steps_fp32"], steps_fp64=r["steps_fp64"],
                storm_cell_updates_per_s_fp32=r["cell_updates_per_s_fp32"],
                storm_cell_updates_per_s_fp64=r["cell_updates_per_s_fp64"],
                max_depth_err_m=s["max_depth_err_m"], max_depth_err_land_m=s["max_depth_err_land_m"],
                peak_time_shift_max_s=float(max(x["peak_time_shift_max_s"] for x in pm)),
                peak_time_shift_p99_max_s=float(max(x["peak_time_shift_p99_s"] for x in pm)),
                volume_rel_err_max=s["worst_volume_rel_err"], mass_err_fp32_max=s["worst_mass_err_fp32"],
                mass_err_fp64_max=s["worst_mass_err_fp64"], agreement_passed=f"{s['agreement_passed']}/32",
                max_depth_err_per_member_m=[round(x["max_depth_err_m"], 6) for x in pm],
                median_member_max_depth_err_m=float(np.median([x["max_depth_err_m"] for x in pm])))


def kernel_only(reps: int, shape: int | None = None, t_start: float = T_START_S):
    """The 2-D step of the headline workload alone, from the headline's wet state at t_start, then a
    256 MiB copy of known size: the run the FETCH_SIZE / WRITE_SIZE counter passes profile (the
    copy calibrates the counters' units).  ``shape`` fixes the float32 kernel shape (default: the
    autotuner's choice)."""
    import torch
    from his_amd import tuning
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    wl = W.precision_form()
    eng, _ = build_from_workload(wl, device=torch.device("cuda"))
    if shape is not None:
        from his_amd import _lib
        _lib.set_tile_shape(torch.float32, shape)
    eng.run(t_start, out_times=[], progress_s=0)        # the storm to the timed window, untimed
    if shape is None:
        shape = tuning.tune(eng.two, force=True)       # on the wet state, as the benchmark does
    log(f"[kernel-only] shape {shape} ({tuning.describe(eng.two.dtype, shape)}), {reps} steps")
    torch.cuda.synchronize()
    us, gbs = kernel_roofline(eng, reps, ALG_BYTES_F32)
    log(f"[kernel-only] 2-D step {us:.1f} us, {gbs:.0f} GB/s algorithmic")
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.ones_like(a)
    for _ in range(3):
        a.copy_(b)
    torch.cuda.synchronize()
    log("[kernel-only] copy 256 MiB x 3 done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--t-start", type=float, default=T_START_S,
                    help="simulated time at which the timed window starts (the storm is run there first)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-storm", action="store_true", help="skip the whole-storm float32 / float64 comparison")
    ap.add_argument("--fp64-steps", type=int, default=1000)
    ap.add_argument("--serve", type=int, default=120, help="requests of the serving sample (0 = skip)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="run only the 2-D step (and a calibration copy) for counter profiling; prints no JSON")
    ap.add_argument("--shape", type=int, default=None, help="--kernel-only: fixed float32 kernel shape index")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the storms, members and requests of this engine are independent and "
              "would shard across GPUs with no exchange; runs on several GPUs are not supported.",
              file=sys.stderr)
        sys.exit(2)
    if args.kernel_only:
        kernel_only(args.steps, args.shape, args.t_start)
        return
    import torch
    from his_amd import workloads as W

    t0 = args.t_start
    log(f"[bench] fp32-bulk 400x325 B=32 coupled step, timed from t = {t0 / 3600:.2f} h")
    wl = W.precision_form()
    r32 = engine_throughput(wl, args.steps, args.warmup, t0)
    log(f"[bench]   {r32['rate'] / 1e9:.3f} G cell-updates/s, {r32['ms_per_step']:.4f} ms/step, "
        f"t {t0:.0f} -> {r32['t_sim']:.0f} s, n_sub max {r32['n_sub_max']}, mass err {r32['mass_err']:.2e}")
    us, gbs = kernel_roofline(r32["eng"], 200, ALG_BYTES_F32)
    log(f"[bench]   2-D step {us:.1f} us -> {gbs:.0f} GB/s algorithmic")
    del r32["eng"]
    log("[bench] fp64 plain 400x325 B=1 coupled step")
    wl64 = W.plain_gpu()
    r64 = engine_throughput(wl64, args.fp64_steps, 50, t0)
    us64, gbs64 = kernel_roofline(r64["eng"], 100, ALG_BYTES_F64)
    del r64["eng"]
    log(f"[bench]   {r64['rate'] / 1e9:.3f} G cell-updates/s, {r64['ms_per_step']:.4f} ms/step, "
        f"2-D step {us64:.1f} us ({gbs64:.0f} GB/s)")
    log("[bench] batching form: 16x100x81 + 8x200x163 + 8x400x325, one engine")
    rb = batching_throughput(max(args.steps // 2, 1), max(args.warmup // 2, 1), t0)
    log(f"[bench]   {rb['rate'] / 1e9:.3f} G cell-updates/s, {rb['ms_per_step']:.4f} ms/step, "
        f"mass err {rb['mass_err']:.2e}")
    srv = None
    if args.serve:
        log(f"[bench] serving form: {args.serve} Poisson(4/s) requests over 3 grid sizes, continuous batching")
        srv = serving_latency(args.serve)
        log(f"[bench]   p50 {srv['latency_p50_s']:.2f} s, p99 {srv['latency_p99_s']:.2f} s, "
            f"{srv['ok']}/{srv['requests']} ok, worst mass err {srv['worst_mass_err']:.1e}")
    prec = None
    if not args.no_storm:
        log("[bench] whole storm (8.5 h), float32 bulk and float64, 32 members")
        prec = precision_vs_fp64()
        log(f"[bench]   float32 {prec['storm_cell_updates_per_s_fp32'] / 1e9:.2f} G, float64 "
            f"{prec['storm_cell_updates_per_s_fp64'] / 1e9:.2f} G cell-updates/s over the storm; max depth err "
            f"{prec['max_depth_err_m'] * 1e3:.2f} mm, peak time shift {prec['peak_time_shift_max_s']:.0f} s, "
            f"volume err {prec['volume_rel_err_max']:.1e}, mass err {prec['mass_err_fp32_max']:.1e}")
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (float64 oracle)")
        cpu = cpu_baseline(args.cpu_budget)
        log(f"[bench]   {cpu['value'] / 1e6:.2f} M cell-updates/s on {cpu['cores']} threads")
    valu_instr = _profile_figure(VALU_FILE, "valu_wave_instructions_per_step")
    valu_rate = valu_instr / (us * 1e-6) if valu_instr else None
    out = {
        "metric": METRIC, "value": r32["rate"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "precision form: 400x325 @ 20 m, 32 members, float32 bulk with float64 "
                               "accumulators, 1-D network (60 reaches) + weir exchange, shared dt, hipGraph step",
                   "members": wl.batch, "nx": 400, "ny": 325, "cells_per_step": r32["cells"],
                   "window_t_sim_s": [r32["t_start"], r32["t_sim"]], "n_sub_max": r32["n_sub_max"],
                   "mass_err_max": r32["mass_err"],
                   "fp64_plain": {"value": r64["rate"], "ms_per_step": r64["ms_per_step"],
                                  "window_t_sim_s": [r64["t_start"], r64["t_sim"]], "n_sub_max": r64["n_sub_max"],
                                  "kernel_us": us64, "kernel_gbs_alg": gbs64, "mass_err": r64["mass_err"]},
                   "fp32_over_fp64_plain": r32["rate"] / r64["rate"],
                   "batching": {"value": rb["rate"], "ms_per_step": rb["ms_per_step"],
                                "member_cells": rb["member_cells"], "groups": rb["groups"],
                                "mass_err": rb["mass_err"], "window_start_s": t0},
                   "serving": srv,
                   "whole_storm_fp32_vs_fp64": prec},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(),
                     "kernel": f"2-D step ({r32['shape']}): stage-1 + stage-2 kernels + finalize, "
                               f"{ALG_BYTES_F32} B per member-cell",
                     "kernel_us": us,
                     "valu": None if valu_rate is None else {
                         "achieved_wave_instr_per_s": valu_rate, "peak_wave_instr_per_s": VALU_PEAK_WAVE_INSTR_PER_S,
                         "frac": valu_rate / VALU_PEAK_WAVE_INSTR_PER_S, "wave_instr_per_step": valu_instr,
                         "source": VALU_FILE.name}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
