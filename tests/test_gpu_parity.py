"""HIP kernels against the reference's golden vectors and the float64 oracle.

Golden files (tests/golden/*.npz) hold the reference solver's own float64 outputs on
seeded inputs; the oracle (oracle/) is the float64 CPU restatement pinned to them.
Tolerance for float64: every cell within 1e-9 relative (plus 1e-13 absolute for
values at round-off level), the bar the port is held to.
"""
import math

import numpy as np
import pytest
import torch

from _helpers import assert_rel

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-9, 1e-13


def _f64(t):
    return t.detach().double().cpu().numpy()


def _swe2d_mixed(d, device, dtype=torch.float64):
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import EdgeBC
    from his_amd.solver.infiltration import Horton
    from his_amd.solver.swe2d import SWE2D
    B = int(d["B"])
    stage = W.sine_tide(0.8, W.TIDE_PERIOD_S, [0.3, 2.1])
    qin = W.triangle(0.0, 600.0, [0.4, 0.9], base=0.05)
    bcs = {"west": EdgeBC("transmissive"), "east": EdgeBC("stage", stage), "south": EdgeBC("discharge", qin),
           "north": EdgeBC("closed")}
    T = lambda a: torch.as_tensor(a, dtype=dtype, device=device)
    hz = Horton(T(d["f0"]), T(d["fc"]), T(d["k"]), float(d["t0"]))
    return SWE2D(d["z"], d["n"], float(d["dx"]), SolverConfig(), bcs, device, dtype, storage=d["S"], horton=hz,
                 batch=B)


def test_swe2d_step_core_matches_reference(golden, gpu):
    d = golden("swe2d_mixed")
    sim = _swe2d_mixed(d, gpu)
    st = d["states"]
    for s in range(3):
        rate = _f64(sim.member_rates(*(torch.as_tensor(st[s][k], device=gpu) for k in range(3))))
        assert_rel(rate, d[f"rate_member_{s}"], RTOL, 0, f"step {s} CFL rate")
        ins = [torch.as_tensor(st[s][k], device=gpu) for k in range(5)]
        sc = torch.tensor([d["dt"][s], d["t"][s]], dtype=torch.float64)
        pm = torch.as_tensor(d["pm"][s]).reshape(9, -1, 1, 1)
        out = sim._core(*ins, sc, pm)
        for k, name in enumerate(("h", "hu", "hv", "s_filled", "infil_cum")):
            assert_rel(_f64(out[k]), st[s + 1][k], RTOL, ATOL, f"step {s} {name}")
        vols = np.stack([_f64(v) for v in out[5:]])
        assert_rel(vols, d["vols"][s], RTOL, 1e-9, f"step {s} volumes")
        # the inputs are untouched (functional slot)
        assert_rel(_f64(ins[0]), st[s][0], 0, 0, "input h unchanged")


def test_swe2d_order1_hll_step_matches_reference(golden, gpu):
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import EdgeBC
    from his_amd.solver.swe2d import SWE2D
    d = golden("swe2d_o1hll")
    cfg = SolverConfig(spatial_order=1, flux="hll")
    bcs = {"west": EdgeBC("closed"), "east": EdgeBC("transmissive"), "south": EdgeBC("stage", 1.2),
           "north": EdgeBC("discharge", -0.3)}
    sim = SWE2D(d["z"], float(d["manning"]), float(d["dx"]), cfg, bcs, gpu, torch.float64, storage=d["S"],
                rain_weight=d["rw"])
    st = d["states"]
    for s in range(2):
        sim.h, sim.hu, sim.hv, sim.s_filled, sim.infil_cum = (torch.as_tensor(st[s][k], device=gpu) for k in range(5))
        assert abs(sim.max_dt() - d["dt"][s]) <= 1e-12 * d["dt"][s]
        v = sim.step(float(d["dt"][s]), float(d["t"][s]), float(d["rain"]))
        for k, name in enumerate(("h", "hu", "hv", "s_filled", "infil_cum")):
            assert_rel(_f64(getattr(sim, name)), st[s + 1][k], RTOL, ATOL, f"step {s} {name}")
        got = np.array([float(v.bnd_in), float(v.bnd_out), float(v.rain), float(v.infil), float(v.clip)])
        assert_rel(got, d["vols"][s], RTOL, 1e-9, f"step {s} volumes")


@pytest.mark.parametrize("name", ["swe1d_net60", "swe1d_chain"])
def test_swe1d_step_matches_reference(golden, gpu, name):
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.domain.channels import apply_gamma
    from his_amd.solver.swe1d import SWE1D
    d = golden(name)
    B = int(d["B"])
    ter = W.make_terrain(100, 81, 80.0)
    cfg = SolverConfig(slot_width_frac=float(d["slot"]))
    if B:
        net = W.network60(ter)
        one = SWE1D([apply_gamma(net, g) for g in d["gammas"]], cfg, gpu,
                    inflow=W.triangle(0.0, 3600.0, [150.0, 320.0], base=10.0),
                    stage=W.sine_tide(0.8, W.TIDE_PERIOD_S, [0.7, 1.9]))
    else:
        one = SWE1D(W.chain_network(ter, 20), cfg, gpu,
                    inflow=W.triangle(0.0, 3600.0, [150.0], base=10.0, batched=False),
                    stage=W.sine_tide(0.8, W.TIDE_PERIOD_S, [0.7], batched=False))
    assert np.array_equal(one.topo.kind, d["kind"]) and np.array_equal(one.topo.face_l, d["fl"])
    assert np.array_equal(one.topo.face_r, d["fr"])
    for nm in ("sec_b", "sec_yb", "sec_ts", "fsec_b", "fsec_yb", "fsec_ts"):
        sec = one.sec if nm.startswith("sec") else one.fsec
        assert_rel(_f64(getattr(sec, nm.split("_")[1])), d[nm], 0, 0, nm)
    st = d["states"]
    for s in range(3):
        one.A, one.Q = (torch.as_tensor(st[s][k], device=gpu) for k in range(2))
        assert abs(one.max_dt() - d["max_dt"][s]) <= 1e-12 * d["max_dt"][s]
        v = one.step(float(d["dt"][s]), float(d["t"][s]))
        assert_rel(_f64(one.A), st[s + 1][0], RTOL, ATOL, f"step {s} A")
        assert_rel(_f64(one.Q), st[s + 1][1], RTOL, ATOL, f"step {s} Q")
        got = np.stack([_f64(v.bnd_in), _f64(v.bnd_out), _f64(v.clip)]).reshape(d["vols"][s].shape)
        assert_rel(got, d["vols"][s], RTOL, 1e-9, f"step {s} volumes")


def coupled_case(device, dtype_precision="fp64", B=2, order=2):
    """The golden coupled workload: 100x81 @ 80 m, 20-reach chain, two members."""
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.engine import build_from_workload
    wl = W.plain_cpu()
    wl.members = [wl.members[0], W.Member(manning=0.05, retention_m=0.008, f0_mmh=40.0, gamma=1.2,
                                          rain_peak_mmh=110.0, tide_phase=1.3, inflow_peak=220.0)][:B]
    wl.precision = dtype_precision
    eng, fo = build_from_workload(wl, SolverConfig(spatial_order=order), device, batched=True)
    return eng, fo


def test_exchange_matches_reference(golden, gpu):
    from his_amd.solver.coupling import exchange
    d = golden("exchange")
    eng, _ = coupled_case(gpu)
    one, two = eng.one, eng.two
    assert np.array_equal(_f64(eng.links.cell1d).astype(int), d["cell1d"])
    assert np.array_equal(_f64(eng.links.inv).astype(int), d["inv"])
    assert np.array_equal(_f64(eng.links.uniq_dev).astype(int), d["uniq"])
    for s in range(3):
        two.set_state(torch.as_tensor(d["h0"][s]), torch.as_tensor(d["hu0"][s]), torch.as_tensor(d["hv0"][s]))
        one.A = torch.as_tensor(d["A0"][s], device=gpu).clone()
        V = exchange(one, two, eng.links, float(d["dt"]), eng.cfg)
        assert_rel(_f64(V), d["V"][s], RTOL, 1e-12, f"draw {s} V")
        assert_rel(_f64(two.h), d["h1"][s], RTOL, ATOL, f"draw {s} h")
        assert_rel(_f64(two.hu), d["hu1"][s], RTOL, ATOL, f"draw {s} hu")
        assert_rel(_f64(two.hv), d["hv1"][s], RTOL, ATOL, f"draw {s} hv")
        assert_rel(_f64(one.A), d["A1"][s], RTOL, ATOL, f"draw {s} A")


def test_coupled_steps_match_reference(golden, gpu):
    """40 coupled steps chosen by the engine itself: same dt / n_sub sequence, same state."""
    d = golden("coupled")
    eng, _ = coupled_case(gpu)
    assert_rel(_f64(eng.two.h), d["init_h"], 0, 0, "initial h")
    assert_rel(_f64(eng.one.A), d["init_A"], RTOL, ATOL, "initial A")
    t, dts, nsubs = 0.0, [], []
    for s in range(40):
        dt = eng.two.max_dt()
        n_sub = max(1, math.ceil(dt / eng.one.max_dt() - 1e-9))
        t = eng.step(t, 1e9, dt=dt, n_sub=n_sub)
        dts.append(dt)
        nsubs.append(n_sub)
    assert_rel(np.array(dts), d["dt"], RTOL, 0, "dt sequence")
    assert nsubs == list(d["n_sub"])
    for name, key in (("h", "h"), ("hu", "hu"), ("hv", "hv"), ("s_filled", "s"), ("infil_cum", "ic")):
        assert_rel(_f64(getattr(eng.two, name)), d[key], 1e-8, 1e-12, f"2-D {name} after 40 steps")
    assert_rel(_f64(eng.one.A), d["A"], 1e-8, 1e-12, "A after 40 steps")
    assert_rel(_f64(eng.one.Q), d["Q"], 1e-8, 1e-10, "Q after 40 steps")
    src = eng.mb.sources()
    for k in ("rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out"):
        assert_rel(np.asarray(src[k]), d[f"src_{k}"], 1e-9, 1e-6, f"ledger {k}")


def test_device_engine_matches_host_driven_steps(gpu):
    """The graph-captured device loop takes the same steps as the host-driven per-op loop."""
    eng_a, _ = coupled_case(gpu)
    eng_b, _ = coupled_case(gpu)
    t = 0.0
    for s in range(60):
        t = eng_a.step(t, 1e9)
    eng_b.run(t, out_times=[])
    assert eng_b.log.steps == 60
    for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
        assert_rel(_f64(getattr(eng_b.two, name)), _f64(getattr(eng_a.two, name)), 1e-12, 1e-14, f"2-D {name}")
    assert_rel(_f64(eng_b.one.A), _f64(eng_a.one.A), 1e-12, 1e-14, "A")
    assert_rel(_f64(eng_b.one.Q), _f64(eng_a.one.Q), 1e-12, 1e-12, "Q")
    sa, sb = eng_a.mb.sources(), eng_b.mb.sources()
    for k in sa:
        assert_rel(np.asarray(sb[k]), np.asarray(sa[k]), 1e-12, 1e-9, f"ledger {k}")


def test_device_engine_survives_snapshots_and_checks():
    """Snapshots, plausibility and mass checks between device launches leave the engine's buffers bound:
    the device loop, stopped at output times and checked every few steps, takes the host loop's steps.
    The device is named without an index, as callers usually do."""
    dev = torch.device("cuda")
    eng_a, _ = coupled_case(dev)
    eng_b, _ = coupled_case(dev)
    outs = [4.0 * k for k in range(1, 16)]
    t, k = 0.0, 0
    for limit in outs:
        while t < limit - 1e-9:
            t = eng_a.step(t, limit)
            k += 1
    seen = []
    eng_b.run(outs[-1], out_times=[0.0] + outs, on_snapshot=lambda s: seen.append(s.t), check_every=3)
    assert eng_b.log.steps == k
    assert seen == [0.0] + outs
    for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
        assert_rel(_f64(getattr(eng_b.two, name)), _f64(getattr(eng_a.two, name)), 1e-12, 1e-14, f"2-D {name}")
    assert_rel(_f64(eng_b.one.A), _f64(eng_a.one.A), 1e-12, 1e-14, "A")
    assert_rel(_f64(eng_b.one.Q), _f64(eng_a.one.Q), 1e-12, 1e-12, "Q")


@pytest.mark.parametrize("order", [2, 1])
def test_device_engine_matches_oracle(gpu, order):
    """Device engine vs the float64 CPU oracle running the same coupled order, 30 steps.  Second
    order takes the single-phase exchange (one link per 1-D cell); first order keeps the two
    phases, since its one 2-D stage cannot clear the CFL rates the exchange reads."""
    from oracle.oracle import Oracle2D, OracleCoupled, OracleNet, oracle_links
    from his_amd import workloads as W
    from his_amd.domain.channels import apply_gamma
    eng, fo = coupled_case(gpu, order=order)
    wl = W.plain_cpu()
    members = [wl.members[0], W.Member(manning=0.05, retention_m=0.008, f0_mmh=40.0, gamma=1.2,
                                       rain_peak_mmh=110.0, tide_phase=1.3, inflow_peak=220.0)]
    wl.members = members
    n, S, f0, fc, k = wl.fields()
    ter = wl.terrain
    two = Oracle2D(ter.z, n, ter.dx, dict(west="transmissive", east="stage", south="transmissive",
                                          north="transmissive"), B=2, storage=S, horton=(f0, fc, k, 0.0),
                   order=order)
    one = OracleNet([apply_gamma(wl.network, m.gamma) for m in members])
    orc = OracleCoupled(two, one, oracle_links(one, ter.z, ter.dx), dict(east=fo.tide), fo.rain, fo.inflow,
                        fo.tide)
    orc.h = _f64(eng.two.h).copy()
    orc.A = _f64(eng.one.A).copy()
    t = 0.0
    for s in range(30):
        t = orc.step(t, 1e9)
    eng.run(t, out_times=[])
    assert eng.log.steps == 30
    assert_rel(_f64(eng.two.h), orc.h, 1e-9, 1e-13, "h")
    assert_rel(_f64(eng.two.hu), orc.hu, 1e-9, 1e-13, "hu")
    assert_rel(_f64(eng.one.A), orc.A, 1e-9, 1e-13, "A")
    assert_rel(_f64(eng.one.Q), orc.Q, 1e-9, 1e-11, "Q")


def test_lake_at_rest_fp64(gpu):
    """Well-balancedness (benchmarks/suite.py:84-116): 500 steps, closed box, partly dry DEM."""
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import all_edges
    from his_amd.solver.swe2d import SWE2D
    ter = W.make_terrain(100, 81, 80.0)
    sim = SWE2D(ter.z, 0.035, ter.dx, SolverConfig(), all_edges("closed"), gpu, torch.float64)
    sim.set_level(3.0)
    wet0 = (sim.h > 0)
    assert 0.1 < float(wet0.double().mean()) < 0.9
    v0 = float(sim.volume())
    t = 0.0
    for _ in range(500):
        dt = sim.max_dt()
        sim.step(dt, t, 0.0)
        t += dt
    u, v = sim.velocities(sim.h, sim.hu, sim.hv)
    assert float(u.abs().max()) < 1e-10 and float(v.abs().max()) < 1e-10
    assert float(((sim.h + sim.z - 3.0).abs() * wet0).max()) < 1e-10
    assert abs(float(sim.volume()) - v0) <= 1e-12 * v0


def test_infiltration_capped_by_available_water(gpu):
    """tests/test_massbalance.py:61-69 of the reference."""
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import all_edges
    from his_amd.solver.infiltration import Horton
    from his_amd.solver.swe2d import SWE2D
    z = np.zeros((6, 6))
    h = Horton.constant(1.0, z.shape, gpu, torch.float64)
    sim = SWE2D(z, 0.03, 10.0, SolverConfig(), all_edges("closed"), gpu, torch.float64, horton=h)
    sim.set_state(np.full(z.shape, 0.05))
    vol = sim.step(1.0, 0.0, 0.0)
    assert float(sim.h.min()) >= 0.0
    assert abs(float(vol.infil) - 0.05 * 36 * 100.0) < 1e-9


def test_storage_fills_before_runoff_and_never_changes_bed(gpu):
    """tests/test_massbalance.py:72-86 of the reference."""
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import all_edges
    from his_amd.solver.swe2d import SWE2D
    z = np.zeros((5, 5))
    sim = SWE2D(z, 0.03, 10.0, SolverConfig(), all_edges("closed"), gpu, torch.float64, storage=np.full(z.shape, 0.02))
    z0 = sim.z.clone()
    for k in range(10):
        sim.step(1.0, float(k), 1e-3)
    assert float(sim.h.max()) == 0.0
    assert torch.allclose(sim.s_filled, torch.full_like(sim.s_filled, 0.01))
    for k in range(20):
        sim.step(1.0, float(k), 1e-3)
    assert float(sim.h.min()) > 0.0
    assert torch.equal(sim.z, z0)


@pytest.mark.parametrize("store,h_end,s_end", [(0.03, 0.02, 0.03), (0.2, 0.0, 0.05)])
def test_retained_water_closed_box(gpu, store, h_end, s_end):
    """tests/test_retained_water.py:45-70 of the reference: 50 mm/h for 1 h on an 8x8 closed box."""
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import EdgeBC
    from his_amd.solver.massbalance import MassBalance
    from his_amd.solver.swe2d import SWE2D
    bcs = {e: EdgeBC("closed") for e in ("west", "east", "south", "north")}
    sim = SWE2D(np.zeros((8, 8)), 0.03, 10.0, SolverConfig(), bcs, gpu, torch.float64, storage=np.full((8, 8), store))
    rain = 50.0 / 3.6e6
    mb = MassBalance(gpu, 0.0)
    t = 0.0
    while t < 3600.0 - 1e-9:
        dt = min(sim.max_dt(), 30.0, 3600.0 - t)
        mb.add("rain", sim.step(dt, t, rain).rain)
        t += dt
    assert torch.allclose(sim.s_filled, torch.full_like(sim.s_filled, s_end), atol=1e-12)
    assert torch.allclose(sim.h, torch.full_like(sim.h, h_end), atol=1e-12)
    assert abs(mb.sources()["rain"] - rain * 3600.0 * 64 * 100.0) < 1e-9 * rain * 3600.0 * 64 * 100.0
    assert mb.check(float(sim.volume() + sim.storage_volume()), t) < 1e-12


def test_fp32_bulk_step_tracks_fp64(golden, gpu):
    """One step in float32 bulk against the float64 reference output: within float32 round-off."""
    d = golden("swe2d_mixed")
    sim = _swe2d_mixed(d, gpu, torch.float32)
    st = d["states"]
    ins = [torch.as_tensor(st[0][k], dtype=torch.float32, device=gpu) for k in range(5)]
    out = sim._core(*ins, torch.tensor([d["dt"][0], d["t"][0]]), torch.as_tensor(d["pm"][0]).reshape(9, -1, 1, 1))
    h, ref = _f64(out[0]), st[1][0]
    assert np.abs(h - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
    vols = np.stack([_f64(v) for v in out[5:]])
    assert np.all(np.abs(vols - d["vols"][0]) <= 1e-4 * np.abs(d["vols"][0]) + 1e-3)


@pytest.mark.parametrize("shape", range(9))
def test_every_kernel_shape_matches_reference(golden, gpu, shape):
    """Tiled and marching kernel variants (the tuner's choices) all reproduce the reference step."""
    from his_amd import _lib
    d = golden("swe2d_mixed")
    st = d["states"]
    try:
        for dtype, rtol in ((torch.float64, RTOL), (torch.float32, None)):
            _lib.set_tile_shape(dtype, shape)
            sim = _swe2d_mixed(d, gpu, dtype)
            for s in range(2):
                ins = [torch.as_tensor(st[s][k], dtype=dtype, device=gpu) for k in range(5)]
                sc = torch.tensor([d["dt"][s], d["t"][s]], dtype=torch.float64)
                out = sim._core(*ins, sc, torch.as_tensor(d["pm"][s]).reshape(9, -1, 1, 1))
                if rtol is not None:
                    for k in range(5):
                        assert_rel(_f64(out[k]), st[s + 1][k], rtol, ATOL, f"shape {shape} step {s} field {k}")
                    assert_rel(np.stack([_f64(v) for v in out[5:]]), d["vols"][s], RTOL, 1e-9, f"shape {shape} vols")
                else:
                    assert np.abs(_f64(out[0]) - st[s + 1][0]).max() < 1e-5
            rates = _f64(sim.member_rates(*(torch.as_tensor(st[0][k], dtype=dtype, device=gpu) for k in range(3))))
            assert_rel(rates, d["rate_member_0"], RTOL if dtype == torch.float64 else 1e-5, 0, f"shape {shape} rate")
    finally:
        _lib.set_tile_shape(torch.float64, -1)
        _lib.set_tile_shape(torch.float32, -1)


def _small_groups(device, precision="fp64"):
    """Two uneven grids with their own networks and members (100x81 @ 80 m x2, 200x163 @ 40 m x2)."""
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    out = []
    for nx, ny, dx, net, seed in ((100, 81, 80.0, "chain", 11), (200, 163, 40.0, "net60", 12)):
        ter = W.make_terrain(nx, ny, dx)
        netw = W.chain_network(ter, 20) if net == "chain" else W.network60(ter)
        wl = W.Workload(f"{nx}x{ny}", ter, netw, W.draw_members(2, seed), W.Storm(1800.0, 1800.0),
                        precision=precision, slot_width_frac=1.0 if net == "chain" else 0.25)
        out.append(build_from_workload(wl, device=device, batched=True)[0])
    return out


def test_grouped_engine_equals_each_group_alone(gpu):
    """Uneven grids in one engine (one launch per stage over both, one dt per group) take exactly
    the steps and produce exactly the state of each grid run by itself."""
    from his_amd.batching import GroupedEngine
    outs = [60.0, 150.0]
    alone = _small_groups(gpu)
    for e in alone:
        e.run(outs[-1], out_times=outs)
    mixed = _small_groups(gpu)
    ge = GroupedEngine(mixed)
    ge.run(outs[-1], out_times=outs)
    for gi, (a, b) in enumerate(zip(alone, mixed)):
        assert a.log.steps == b.log.steps, f"group {gi}: {a.log.steps} vs {b.log.steps} steps"
        for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
            assert np.array_equal(_f64(getattr(a.two, name)), _f64(getattr(b.two, name))), f"group {gi} {name}"
        assert np.array_equal(_f64(a.one.A), _f64(b.one.A)), f"group {gi} A"
        assert np.array_equal(_f64(a.one.Q), _f64(b.one.Q)), f"group {gi} Q"
        # the ledger is the same device sums, moved to the host at different moments (checks fire on
        # the group with the most steps), so its float64 host totals agree to rounding only
        sa, sb = a.mb.sources(), b.mb.sources()
        for k in sa:
            assert_rel(np.asarray(sb[k]), np.asarray(sa[k]), 1e-13, 1e-12, f"group {gi} ledger {k}")


@pytest.mark.parametrize("precision", ["fp32bulk", "fp64"])
def test_served_request_equals_the_request_alone(gpu, precision):
    """A request served among others -- sharing launches with busy slots of two grid sizes, loaded
    into a slot another request used before -- gives exactly its standalone result."""
    from his_amd import serving as S
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    grids = (S.GridSpec(100, 81, 80.0, "chain"), S.GridSpec(200, 163, 40.0, "net60"))
    ms = W.draw_members(4, 77)
    reqs = [dict(id=0, arrival_s=0.0, size=0, storm_s=1800.0, member=ms[0]),
            dict(id=1, arrival_s=0.0, size=1, storm_s=1200.0, member=ms[1]),
            dict(id=2, arrival_s=0.0, size=0, storm_s=960.0, member=ms[2]),
            dict(id=3, arrival_s=0.0, size=1, storm_s=1800.0, member=ms[3])]
    srv = S.Server(slots=(1, 2), precision=precision, device=gpu, grids=grids, chunk=16, keep_fields=True,
                   tune=False)
    res = {r.id: r for r in srv.serve(reqs, realtime=False, log_every_s=0)}
    srv.close()
    assert sorted(res) == [0, 1, 2, 3]
    cfg = SolverConfig().replace(precision=precision)
    for req in reqs:
        r = res[req["id"]]
        assert r.status == "ok", (req["id"], r.status)
        alone = S.run_alone(grids[req["size"]], req, cfg, gpu)
        assert r.steps == alone["steps"], (req["id"], r.steps, alone["steps"])
        assert np.array_equal(r.fields["h"], alone["h"]), f"request {req['id']} h"
        assert np.array_equal(r.fields["A"], alone["A"]), f"request {req['id']} A"
        assert np.array_equal(r.fields["depth_max"].astype(np.float32), alone["depth_max"]), \
            f"request {req['id']} depth max"
        assert abs(r.mass_err - alone["mass_err"]) <= 1e-12 + 1e-9 * alone["mass_err"]


def test_a_failing_request_fails_alone(gpu):
    """A request whose member the engine flags stops and is reported unstable; the requests
    beside it, in the same pool and the same launches, finish with their standalone results.
    The flag is set from the host right after admission, as the step head sets it on a
    non-finite rate or too many 1-D substeps, so every kernel still runs normally."""
    from his_amd import serving as S
    from his_amd import workloads as W
    from his_amd.config import SolverConfig

    class Flagging(S.Server):
        def _admit(self, pi, req, now):
            ok = super()._admit(pi, req, now)
            if ok and req["id"] == 1:
                m = next(m for m, v in self.busy.items() if v[2].id == 1)
                self.dev.error[m] = 2
            return ok

    grids = (S.GridSpec(100, 81, 80.0, "chain"),)
    ms = W.draw_members(3, 91)
    reqs = [dict(id=i, arrival_s=0.0, size=0, storm_s=s, member=ms[i]) for i, s in enumerate((1200.0, 1800.0, 960.0))]
    srv = Flagging(slots=(3,), precision="fp64", device=gpu, grids=grids, chunk=16, keep_fields=True, tune=False)
    res = {r.id: r for r in srv.serve(reqs, realtime=False, log_every_s=0)}
    srv.close()
    assert sorted(res) == [0, 1, 2]
    assert res[1].status == "unstable: non-finite CFL rate" and res[1].steps == 0, (res[1].status, res[1].steps)
    cfg = SolverConfig().replace(precision="fp64")
    for i in (0, 2):
        r = res[i]
        assert r.status == "ok", (i, r.status)
        alone = S.run_alone(grids[0], reqs[i], cfg, gpu)
        assert r.steps == alone["steps"]
        assert np.array_equal(r.fields["h"], alone["h"]) and np.array_equal(r.fields["A"], alone["A"])


def _small_context(gpu):
    """The API's context over the 100 x 81 @ 80 m plain with the 20-reach chain and its land use."""
    from his_amd import api
    from his_amd import workloads as W
    from his_amd.domain.site import domain_from
    ter = W.make_terrain(100, 81, 80.0)
    return api.configure(domain_from(ter, W.chain_network(ter, 20)), device=gpu)


def test_simulate_and_simulate_batch(gpu):
    """The reference's call surface: ``simulate(site, scenario, ...)`` and ``simulate_batch(sites,
    scenarios)`` called as generate.py:288 calls it.  depth_max is the maximum over the output-time
    snapshots, mass balance holds per member, a batch of one is routed to simulate, the disbenefit of
    a design against itself is zero, and a bad site raises ValueError."""
    from his_amd import api
    from his_amd.scenarios import Scenario
    ctx = _small_context(gpu)
    dom = ctx.domain
    try:
        sc1 = Scenario(50, "RCP4.5", "SSP2", 2050)
        sc2 = Scenario(100, "RCP8.5", "SSP5", 2080, tide_on=False)
        site = api.baseline_site(sc1)
        r = api.simulate(site, sc1)
        nt = len(r.extras["forcing"].times(900.0))
        assert r.depth_series.shape == (nt, dom.ny, dom.nx) and r.times.shape == (nt,)
        assert torch.equal(r.depth_max, r.depth_series.amax(0))
        assert r.mass_balance_error < 1e-3 and float(r.depth_max.max()) > 0.1 and r.infiltrated_volume_m3 > 0
        assert r.channel_eta.shape[1] == int((r.extras["topology"].kind <= 1).sum())
        assert "exposure" in r.extras and r.extras["exposure"].shape == dom.shape
        (r1,) = api.simulate_batch([site], [sc1])
        assert torch.equal(r1.depth_max, r.depth_max)
        wetter = site.replace(storage_depth=site.storage_depth * 0.0)
        rb = api.simulate_batch([site, wetter], [sc1, sc2])
        assert len(rb) == 2 and all(x.mass_balance_error < 1e-3 for x in rb)
        assert rb[1].extras["forcing"].ssp == "SSP5" and rb[0].extras["batch_size"] == 2
        # member 0 of the batch shares dt with member 1, so it is close to, not equal to, the run alone
        assert float((rb[0].depth_max - r.depth_max).abs().max()) < 0.05
        # the wetter, heavier storm floods more
        assert float(rb[1].depth_max.sum()) > float(rb[0].depth_max.sum())
        assert api.disbenefit(r.depth_max, r.depth_max) == {"max_depth_increase_m": 0.0, "area_worsened_ha": 0.0}
        rc = api.simulate(wetter, sc1, compare_to_baseline=True)
        assert rc.max_depth_increase_m >= 0.0 and rc.area_worsened_ha >= 0.0
        with pytest.raises(ValueError):
            api.simulate(site.replace(manning=site.manning * 0.0), sc1)
        with pytest.raises(ValueError):
            api.simulate_batch([site, site], [sc1])
    finally:
        api.configure()


@pytest.mark.parametrize("precision", ["fp32bulk", "fp64"])
def test_non_finite_depth_stops_the_member(gpu, precision):
    """A NaN written into one member's depth from the host is found by the step's float64 depth sum:
    after one step the engine has flagged that member with code 3 (the rate fold alone would not see
    it: a NaN drops out of a max), and ``run`` raises NumericalInstabilityError naming the member.
    Every kernel runs normally on NaN values; no index or loop bound depends on the depth."""
    from his_amd.solver.coupling import NumericalInstabilityError
    eng, _ = coupled_case(gpu, precision)
    t = 20.0
    eng.run(t, out_times=[])
    de = eng._device()
    de.refresh()
    de.t.fill_(t)
    de.t_end.fill_(1e12)
    de.t_limit.fill_(1e12)
    assert de.error.cpu().tolist() == [0, 0]
    eng.two.h[1, 5, 97] = float("nan")
    de.step(1)
    assert de.error.cpu().tolist() == [0, 3]
    with pytest.raises(NumericalInstabilityError, match="member 1"):
        eng.run(60.0, t0=t, out_times=[])


def test_engine_follows_a_tile_shape_change(gpu):
    """The tile shape is a library-wide setting; an engine that has captured its step graphs with one
    shape takes the new one from its next step (new tiles, partial layout and ledger ranges), so the
    state and the ledger match an engine that used the new shape throughout."""
    from his_amd import _lib
    try:
        _lib.set_tile_shape(torch.float64, -1)
        eng_a, _ = coupled_case(gpu)
        eng_a.run(12.0, out_times=[])
        _lib.set_tile_shape(torch.float64, 6)
        eng_a.run(30.0, t0=12.0, out_times=[])
        eng_b, _ = coupled_case(gpu)
        eng_b.run(12.0, out_times=[])
        eng_b.run(30.0, t0=12.0, out_times=[])
    finally:
        _lib.set_tile_shape(torch.float64, -1)
    assert eng_a.log.steps == eng_b.log.steps
    for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
        assert_rel(_f64(getattr(eng_a.two, name)), _f64(getattr(eng_b.two, name)), 1e-12, 1e-14, name)
    sa, sb = eng_a.mb.sources(), eng_b.mb.sources()
    for k in sa:
        assert_rel(np.asarray(sa[k]), np.asarray(sb[k]), 1e-12, 1e-9, f"ledger {k}")
    err = np.asarray(eng_a.mb.check(eng_a.total_volume(), 30.0))
    assert err.max() < 1e-12


def test_branch_order_does_not_change_results(gpu):
    """Capturing the 1-D network or the 2-D stages first in each step's graph changes only which
    hardware queue each branch starts on: states, 1-D state and ledger are bit-identical."""
    runs = []
    for order in (0, 1):
        eng, _ = coupled_case(gpu)
        eng._device().set_branch_order(order)
        eng.run(60.0, out_times=[])
        runs.append(eng)
    a, b = runs
    assert a.log.steps == b.log.steps
    for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
        assert np.array_equal(_f64(getattr(a.two, name)), _f64(getattr(b.two, name))), name
    assert np.array_equal(_f64(a.one.A), _f64(b.one.A)) and np.array_equal(_f64(a.one.Q), _f64(b.one.Q))
    sa, sb = a.mb.sources(), b.mb.sources()
    for k in sa:
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k


def test_coupled_tuner_steps_are_steps_of_the_run(gpu, tmp_path, monkeypatch):
    """tune_coupled times every shape and capture order on the engine's own coupled steps: those
    steps count in the run's step total and ledger (mass balance holds after them), the chosen
    shape and order stay set, and the run continues to the same state as an engine that took the
    same steps with the chosen settings throughout (float64: shape changes move only the order of
    the float64 partial sums)."""
    from his_amd import _lib, tuning
    monkeypatch.setenv("HIS_AMD_TUNING", str(tmp_path / "tuning.json"))
    try:
        eng_a, _ = coupled_case(gpu)
        eng_a.run(6.0, out_times=[])
        shape, times = tuning.tune_coupled(eng_a, steps=3, warm=1)
        assert sorted(times) == list(range(_lib.num_tile_shapes()))
        assert shape == min(times, key=times.get)
        de = eng_a._device()
        t_now = float(de.t[0].item())
        steps_a = eng_a.log.steps
        assert steps_a == int(de.steps[0].item())
        err = np.asarray(eng_a.mb.check(eng_a.total_volume(), t_now))
        assert err.max() < 1e-12
        eng_a.run(t_now + 20.0, t0=t_now, out_times=[])
        # the same number of steps with the chosen shape and order from the start
        _lib.set_tile_shape(torch.float64, shape)
        eng_b, _ = coupled_case(gpu)
        eng_b._device().set_branch_order(eng_a.branch_order)
        eng_b.run(6.0, out_times=[])
        db = eng_b._device()
        db.t_end.fill_(1e12)
        db.t_limit.fill_(1e12)
        db.step(steps_a - eng_b.log.steps)
        eng_b.log.steps = steps_a
        eng_b._fold_device_ledger()
        assert abs(float(db.t[0].item()) - t_now) < 1e-9
        eng_b.run(t_now + 20.0, t0=t_now, out_times=[])
    finally:
        _lib.set_tile_shape(torch.float64, -1)
    assert eng_a.log.steps == eng_b.log.steps
    for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
        assert_rel(_f64(getattr(eng_a.two, name)), _f64(getattr(eng_b.two, name)), 1e-12, 1e-14, name)
    err = np.asarray(eng_a.mb.check(eng_a.total_volume(), t_now + 20.0))
    assert err.max() < 1e-12


def test_async_outputs_equal_sync_outputs(gpu):
    """run(async_outputs=True) delivers every output, in order, with the same arrays and ledger
    as synchronous snapshots of the same run."""
    outs = [0.0, 5.0, 10.0, 15.0, 20.0, 25.0, 30.0]
    got = {}
    for mode in (False, True):
        eng, _ = coupled_case(gpu)
        snaps = []
        eng.run(30.0, out_times=outs, on_snapshot=snaps.append, async_outputs=mode)
        got[mode] = snaps
    a, b = got[False], got[True]
    assert [s.t for s in a] == [s.t for s in b] == outs
    for sa, sb in zip(a, b):
        for k in ("h", "u", "v", "eta1", "q1"):
            assert np.array_equal(getattr(sa, k), getattr(sb, k)), (sa.t, k)
        for k in sa.ledger:
            assert np.array_equal(np.asarray(sa.ledger[k]), np.asarray(sb.ledger[k])), (sa.t, k)


def test_device_guard_flags_an_implausible_2d_depth(gpu):
    """A 2-D depth above 10 m after a step flags that member on the device (error code 4), and the
    run then raises, naming the member; the other member is not flagged."""
    from his_amd.solver.coupling import NumericalInstabilityError
    eng, _ = coupled_case(gpu)
    two = eng.two
    j, i = two.ny // 2, two.nx // 2
    two.h[1, j, i] = 11.0 + float(two.h[1, j, i])     # a column of water the step cannot drain
    dev = eng._device()
    dev.refresh()
    dev.t.fill_(0.0)
    dev.t_end.fill_(1e12)
    dev.t_limit.fill_(1e12)
    dev.step(1)
    assert dev.error.cpu().tolist() == [0, 4]
    with pytest.raises(NumericalInstabilityError, match="member 1"):
        eng._check_device_errors(dev, float(dev.t[0].item()), "guard")


def test_device_guard_flags_an_implausible_channel_level(gpu):
    """A channel level more than 10 m above its bank after a step flags that member on the device
    (error code 5); the other member is not flagged.  The weir coefficient is zero here so that the
    exchange, which would otherwise spill the excess onto the floodplain, leaves it in the channel."""
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.engine import build_from_workload
    wl = W.plain_cpu()
    wl.members = [wl.members[0], wl.members[0]]
    eng, _ = build_from_workload(wl, SolverConfig(weir_cw=0.0), gpu, batched=True)
    one = eng.one
    c = int(np.flatnonzero(one.topo.kind <= 1)[len(one.topo.kind) // 3])
    ys = one.sec.depth(one.A).clone()
    ys[1, c] = float(one.topo.zbank[c]) + 40.0 - float(one.z[1, c])
    one.A[1, c] = one.sec.area(ys)[1, c]
    dev = eng._device()
    dev.refresh()
    dev.t.fill_(0.0)
    dev.t_end.fill_(1e12)
    dev.t_limit.fill_(1e12)
    dev.step(1)
    assert dev.error.cpu().tolist() == [0, 5]

