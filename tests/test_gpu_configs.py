"""The HIP engine at the full-size workloads: 400 x 325 cells at 20 m with the 60-reach network.

* plain float64, B = 1 (Preissmann slot 0.25): from a wet state two hours into the storm,
  20 coupled steps against the float64 oracle -- the same dt and n_sub sequence, every cell
  of h, hu, hv, s_filled, infil_cum, A and Q within 1e-9 relative, the ledger increments too;
* float32 bulk, B = 32: three hours of storm against the same members in float64 -- every
  member's peak depth within 1 cm, volumes close, both runs mass-balanced below 1e-3 -- and two
  members stepped beside the float64 oracle under the same dt, within float32 round-off;
* the batching form (16 x 100x81 + 8 x 200x163 + 8 x 400x325) in one engine against each
  group run alone: bit for bit;
* served requests of all three sizes, 400 x 325 included, against each request run alone:
  bit for bit.

At this grid the register march meets shapes the small cases do not: a last strip of 52
columns (400 = 6 x 58 + 52) and a last chunk of 5 rows (325 = 40 x 8 + 5).
"""
import math

import numpy as np
import pytest
import torch

from _helpers import assert_rel

pytestmark = pytest.mark.gpu
EDGE_KINDS = dict(west="transmissive", east="stage", south="transmissive", north="transmissive")


def _f64(t):
    return t.detach().double().cpu().numpy()


def _oracle_for(wl, members, slot):
    """The float64 oracle of a workload restricted to ``members`` (indices into wl.members)."""
    from oracle.oracle import Oracle2D, OracleCoupled, OracleNet, oracle_links
    from his_amd import workloads as W
    from his_amd.domain.channels import apply_gamma
    sub = W.Workload(wl.name, wl.terrain, wl.network, [wl.members[i] for i in members], wl.storm,
                     "fp64", slot)
    n, S, f0, fc, k = sub.fields()
    rain, tide, inflow = sub.forcing(batched=True)
    ter = wl.terrain
    two = Oracle2D(ter.z, n, ter.dx, EDGE_KINDS, B=len(members), storage=S, horton=(f0, fc, k, 0.0))
    one = OracleNet([apply_gamma(wl.network, m.gamma) for m in sub.members], slot)
    return OracleCoupled(two, one, oracle_links(one, ter.z, ter.dx), dict(east=tide), rain, inflow, tide)


def _load_oracle(orc, eng, members=None):
    """Copy the engine's state (members ``members``, all if None) into the oracle as float64."""
    two, one = eng.two, eng.one
    B = orc.two.B
    pick = (lambda a: a) if members is None else (lambda a: a[members])
    shp2, shp1 = (B, two.ny, two.nx), (B, one.topo.n_cells)
    orc.h, orc.hu, orc.hv, orc.s, orc.ic = (pick(_f64(getattr(two, k)).reshape((-1,) + shp2[1:])).reshape(shp2)
                                            .copy() for k in ("h", "hu", "hv", "s_filled", "infil_cum"))
    orc.A = pick(_f64(one.A).reshape((-1,) + shp1[1:])).reshape(shp1).copy()
    orc.Q = pick(_f64(one.Q).reshape((-1,) + shp1[1:])).reshape(shp1).copy()


def test_plain_gpu_fp64_matches_oracle_from_wet_state(gpu):
    """400 x 325, B = 1, float64, 60 reaches, slot 0.25: 20 steps from t = 2 h against the oracle."""
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    wl = W.plain_gpu()
    eng, _ = build_from_workload(wl, device=gpu)
    assert eng.batch is None and eng.two.dtype == torch.float64
    t0 = 7200.0
    eng.run(t0, out_times=[])
    h0 = _f64(eng.two.h)
    assert float((h0[~wl.terrain.sea] > 0.05).mean()) > 0.02, "the land should be wetting by now"
    orc = _oracle_for(wl, [0], wl.slot_width_frac)
    _load_oracle(orc, eng)
    eng._fold_device_ledger()
    de = eng._device()
    de.refresh()
    de.t.fill_(t0)
    de.t_end.fill_(1e12)
    de.t_limit.fill_(1e12)
    de.acc.zero_()
    t = t0
    for s in range(20):
        t = orc.step(t, 1e12)
        de.step(1)
        dt_dev, ns_dev = float(de.dt[0].item()), int(de.n_sub[0].item())
        dt_orc, ns_orc = orc.last
        assert abs(dt_dev - dt_orc) <= 1e-9 * dt_orc, f"step {s}: dt {dt_dev!r} vs oracle {dt_orc!r}"
        assert ns_dev == ns_orc, f"step {s}: n_sub {ns_dev} vs oracle {ns_orc}"
    assert abs(float(de.t[0].item()) - t) <= 1e-9 * t
    sq = lambda a: a.reshape(orc.h.shape[1:]) if a.ndim == 3 else a
    for name, ref, atol in (("h", orc.h, 1e-13), ("hu", orc.hu, 1e-13), ("hv", orc.hv, 1e-13),
                            ("s_filled", orc.s, 1e-13), ("infil_cum", orc.ic, 1e-13)):
        assert_rel(_f64(getattr(eng.two, name)), sq(ref[0]), 1e-9, atol, f"2-D {name} after 20 steps")
    assert_rel(_f64(eng.one.A), orc.A[0], 1e-9, 1e-13, "A after 20 steps")
    assert_rel(_f64(eng.one.Q), orc.Q[0], 1e-9, 1e-11, "Q after 20 steps")
    acc = _f64(de.acc)[:, 0]
    for row, key in enumerate(("rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out")):
        assert_rel(acc[row:row + 1], orc.acc[key], 1e-9, 1e-9, f"ledger {key}")
    assert_rel(acc[8:9], orc.ex_in, 1e-9, 1e-9, "exchanged 2-D -> 1-D")
    assert_rel(acc[9:10], orc.ex_out, 1e-9, 1e-9, "exchanged 1-D -> 2-D")


def _run_to(eng, t_end, interval=900.0):
    """Run an engine from t = 0 to t_end; per-member peak depth over the output times and the
    surface + storage + channel volume at each, with the mass error the run asserts."""
    peak = [None]
    vols = []

    def snap(s):
        h = s.h.astype(np.float64)
        peak[0] = h if peak[0] is None else np.maximum(peak[0], h)
        led = s.ledger
        vols.append(np.asarray(led["v2d"]) + np.asarray(led["storage"]) + np.asarray(led["v1d"]))
    err = eng.run(t_end, out_times=np.arange(0.0, t_end + 1e-9, interval), on_snapshot=snap, progress_s=0)
    return peak[0], np.stack(vols), np.asarray(err, dtype=np.float64).reshape(-1)


def test_precision_form_fp32_bulk_tracks_fp64(gpu):
    """400 x 325, B = 32, float32 bulk with float64 accumulators, first 3 h of the storm (rain
    rising to its peak), against the same 32 members in float64: peak depth within 1 cm for
    every member (the 1 cm bar), volumes within 1e-3 relative, every member mass-balanced.  The
    float32 run uses the kernel shape the benchmark's coupled-step tuner picks for this workload
    (58-column strips x 19-row chunks)."""
    from his_amd import _lib
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    wl = W.precision_form()
    T = 3 * 3600.0
    e32, _ = build_from_workload(wl, device=gpu)
    assert e32.two.dtype == torch.float32 and e32.two.B == 32
    try:
        _lib.set_tile_shape(torch.float32, 1)
        assert _lib.tile_shape(torch.float32, 1)[:2] == (58, 19)
        p32, v32, m32 = _run_to(e32, T)
    finally:
        _lib.set_tile_shape(torch.float32, -1)
    steps32 = e32.log.steps
    del e32
    wl64 = W.precision_form()
    wl64.precision = "fp64"
    e64, _ = build_from_workload(wl64, device=gpu)
    p64, v64, m64 = _run_to(e64, T)
    assert steps32 > 2000 and e64.log.steps > 2000
    d = np.abs(p32 - p64).reshape(32, -1).max(axis=1)
    assert float((p64 > 0.1).mean()) > 0.05, "a tenth of a metre somewhere on a twentieth of the cells"
    assert d.max() < 0.01, f"peak-depth error per member (m): {np.array2string(d, precision=4)}"
    vrel = np.abs(v32 - v64) / np.maximum(np.abs(v64), 1.0)
    assert vrel.max() < 1e-3, f"worst volume relative error {vrel.max():.3e}"
    assert m32.max() < 1e-3 and m64.max() < 1e-3, (m32.max(), m64.max())


# float32 kernel shapes: the default (58-column strips x 8-row chunks, last chunk 5 rows at 325 rows)
# and the two the coupled-step tuner picks for this workload (19 rows: a last chunk of 2 rows; 12
# rows: 1 row)
@pytest.mark.parametrize("shape", [2, 1, 7])
def test_precision_form_members_track_the_oracle(gpu, shape):
    """Two float32-bulk members of the 400 x 325 workload, wet after 1 h, take 5 coupled steps
    beside the float64 oracle under the same dt and n_sub: every cell of h, hu, hv within
    float32 round-off of the state's scale (5e-5 of max |value|, plus 1e-6), A and Q within 1e-6
    relative of their scale."""
    from his_amd import _lib
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    wl = W.precision_form()
    wl.members = wl.members[:2]
    eng, _ = build_from_workload(wl, device=gpu)
    t = 3600.0
    eng.run(t, out_times=[])
    orc = _oracle_for(wl, [0, 1], wl.slot_width_frac)
    _load_oracle(orc, eng)
    try:
        _lib.set_tile_shape(torch.float32, shape)
        for s in range(5):
            dt = eng.two.max_dt()
            n_sub = max(1, math.ceil(dt / eng.one.max_dt() - 1e-9))
            t_dev = eng.step(t, 1e12, dt=dt, n_sub=n_sub)
            t = orc.step(t, 1e12, dt=dt, n_sub=n_sub)
            assert t_dev == t
    finally:
        _lib.set_tile_shape(torch.float32, -1)
    for name, ref in (("h", orc.h), ("hu", orc.hu), ("hv", orc.hv), ("s_filled", orc.s)):
        got = _f64(getattr(eng.two, name))
        scale = float(np.abs(ref).max())
        err = float(np.abs(got - ref).max())
        assert err <= 5e-5 * scale + 1e-6, f"{name}: max |fp32 - oracle| {err:.3e} at scale {scale:.3e}"
    for name, ref in (("A", orc.A), ("Q", orc.Q)):
        got = _f64(getattr(eng.one, name))
        err = float(np.abs(got - ref).max())
        assert err <= 1e-6 * float(np.abs(ref).max()) + 1e-9, f"{name}: max |got - oracle| {err:.3e}"


def test_batching_form_equals_each_group_alone(gpu):
    """The full batching form in one engine (one launch per 2-D stage over all three grids, one
    dt per group) against each group's engine run by itself: the same steps and the same state,
    bit for bit, after 15 simulated minutes with an output stop at 5 and 10."""
    from his_amd.batching import build_batching_form
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    outs = [300.0, 600.0, 900.0]
    ge, _ = build_batching_form(device=gpu)
    ge.run(outs[-1], out_times=outs)
    assert [e.two.B for e in ge.engines] == [16, 8, 8]
    for gi, wl in enumerate(W.batching_form()):
        alone, _ = build_from_workload(wl, device=gpu, batched=True)
        alone.run(outs[-1], out_times=outs)
        mixed = ge.engines[gi]
        assert alone.log.steps == mixed.log.steps, (gi, alone.log.steps, mixed.log.steps)
        for name in ("h", "hu", "hv", "s_filled", "infil_cum"):
            assert torch.equal(getattr(alone.two, name), getattr(mixed.two, name)), f"group {gi} {name}"
        assert torch.equal(alone.one.A, mixed.one.A) and torch.equal(alone.one.Q, mixed.one.Q), f"group {gi} 1-D"
        del alone
    ge.close()


def test_served_mix_with_full_size_requests_equals_each_alone(gpu):
    """Requests of all three serving sizes, two of them 400 x 325, through pools of two slots
    (so a slot is reused): every result equals the same request run alone, bit for bit."""
    from his_amd import serving as S
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    ms = W.draw_members(7, 1234)
    spec = [(2, 1200.0), (0, 1800.0), (1, 960.0), (2, 720.0), (0, 600.0), (1, 1440.0), (0, 960.0)]
    reqs = [dict(id=i, arrival_s=0.0, size=sz, storm_s=dur, member=ms[i]) for i, (sz, dur) in enumerate(spec)]
    srv = S.Server(slots=(2, 2, 2), device=gpu, chunk=16, keep_fields=True, tune=False)
    res = {r.id: r for r in srv.serve(reqs, realtime=False, log_every_s=0)}
    srv.close()
    assert sorted(res) == list(range(len(reqs)))
    cfg = SolverConfig().replace(precision="fp32bulk")
    for req in reqs:
        r = res[req["id"]]
        assert r.status == "ok", (req["id"], r.status)
        alone = S.run_alone(S.SERVING_GRIDS[req["size"]], req, cfg, gpu)
        assert r.steps == alone["steps"], (req["id"], r.steps, alone["steps"])
        assert np.array_equal(r.fields["h"], alone["h"]), f"request {req['id']} h"
        assert np.array_equal(r.fields["A"], alone["A"]), f"request {req['id']} A"
        assert np.array_equal(r.fields["depth_max"], alone["depth_max"]), f"request {req['id']} depth max"
        assert r.hydrograph == alone["hydrograph"], f"request {req['id']} outlet hydrograph"
        assert abs(r.mass_err - alone["mass_err"]) <= 1e-12 + 1e-9 * alone["mass_err"]
    big = [res[q["id"]] for q in reqs if q["size"] == 2]
    assert all(max(x[1] for x in r.hydrograph) > 1.0 for r in big), "water reaches the sea through the channels"


def test_hbm_budget_holds_a_request_in_the_queue(gpu):
    """A budget with room for one busy 200 x 163 slot: the second request waits, although a slot is
    free, until the first has finished; each still equals its standalone run, hydrograph included."""
    from his_amd import serving as S
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    grids = (S.GridSpec(200, 163, 40.0, "net60"),)
    one_slot = S.slot_bytes(grids[0], "fp64")
    ms = W.draw_members(2, 4321)
    reqs = [dict(id=i, arrival_s=0.0, size=0, storm_s=600.0, member=ms[i]) for i in range(2)]
    srv = S.Server(slots=(2,), precision="fp64", hbm_budget_bytes=1.5 * one_slot, device=gpu, grids=grids,
                   chunk=16, keep_fields=True, tune=False)
    res = {r.id: r for r in srv.serve(reqs, realtime=False, log_every_s=0)}
    srv.close()
    assert res[1].admitted_s >= res[0].done_s > res[0].admitted_s
    cfg = SolverConfig().replace(precision="fp64")
    for req in reqs:
        r, alone = res[req["id"]], S.run_alone(grids[0], req, cfg, gpu)
        assert r.status == "ok" and r.steps == alone["steps"]
        assert np.array_equal(r.fields["h"], alone["h"]) and np.array_equal(r.fields["A"], alone["A"])
        assert r.hydrograph == alone["hydrograph"]
    with pytest.raises(RuntimeError, match="budget"):
        S.Server(slots=(1,), precision="fp64", hbm_budget_bytes=0.5 * one_slot, device=gpu, grids=grids,
                 tune=False).serve(reqs[:1], realtime=False, log_every_s=0)
