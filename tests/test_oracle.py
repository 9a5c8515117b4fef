"""The float64 CPU oracle against the reference's own outputs (golden vectors) and
against what the mathematics fixes.  These pins are what make the oracle a valid
checker for the HIP kernels."""
import math

import numpy as np
import pytest

from _helpers import assert_rel
from his_amd import workloads as W
from his_amd.domain.channels import apply_gamma
from oracle.oracle import Oracle2D, OracleCoupled, OracleNet, exchange, normal_depth, oracle_links

RTOL, ATOL = 1e-13, 1e-15


def test_swe2d_step_matches_reference(golden):
    d = golden("swe2d_mixed")
    o = Oracle2D(d["z"], d["n"], float(d["dx"]), dict(west="transmissive", east="stage", south="discharge",
                                                       north="closed"), B=int(d["B"]), storage=d["S"],
                 horton=(d["f0"], d["fc"], d["k"], float(d["t0"])))
    st = d["states"]
    for s in range(3):
        assert_rel(o.max_rate(*st[s][:3]), d[f"rate_member_{s}"], 1e-15, 0, "rate")
        out, vols = o.step(*st[s], d["dt"][s], d["t"][s], d["pm"][s])
        for k in range(5):
            assert_rel(out[k], st[s + 1][k], RTOL, ATOL, f"step {s} field {k}")
        assert_rel(vols, d["vols"][s], 1e-13, 1e-9, f"step {s} volumes")


def test_swe2d_order1_hll_matches_reference(golden):
    d = golden("swe2d_o1hll")
    o = Oracle2D(d["z"], float(d["manning"]), float(d["dx"]), dict(west="closed", east="transmissive",
                                                                    south="stage", north="discharge"),
                 order=1, flux="hll", storage=d["S"], rain_weight=d["rw"])
    st = d["states"]
    for s in range(2):
        pm = np.array([float(d["rain"]), 0, 0, 1.2, -0.3, 0, 0, 1.2, -0.3])
        out, vols = o.step(*st[s], d["dt"][s], d["t"][s], pm)
        for k in range(5):
            assert_rel(out[k][0], st[s + 1][k], RTOL, ATOL, f"step {s} field {k}")
        assert_rel(vols[:, 0], d["vols"][s], 1e-13, 1e-9, f"step {s} volumes")


@pytest.mark.parametrize("name", ["swe1d_net60", "swe1d_chain"])
def test_swe1d_matches_reference(golden, name):
    d = golden(name)
    B = max(int(d["B"]), 1)
    ter = W.make_terrain(100, 81, 80.0)
    nets = [apply_gamma(W.network60(ter), g) for g in d["gammas"]] if int(d["B"]) else [W.chain_network(ter, 20)]
    net = OracleNet(nets, float(d["slot"]))
    assert np.array_equal(net.tp["kind"], d["kind"]) and np.array_equal(net.tp["fl"], d["fl"])
    tide = W.sine_tide(0.8, W.TIDE_PERIOD_S, [0.7, 1.9][:B])
    inflow = W.triangle(0.0, 3600.0, [150.0, 320.0][:B], base=10.0)
    st = d["states"]
    for s in range(3):
        A, Q = st[s][0].reshape(B, -1), st[s][1].reshape(B, -1)
        assert net.max_dt(A, Q) == pytest.approx(float(d["max_dt"][s]), rel=1e-15)
        t = d["t"][s]
        A1, Q1, v = net.step(A, Q, d["dt"][s], inflow(t), tide(t))
        assert_rel(A1, st[s + 1][0].reshape(B, -1), RTOL, ATOL, f"step {s} A")
        assert_rel(Q1, st[s + 1][1].reshape(B, -1), RTOL, ATOL, f"step {s} Q")
        assert_rel(v, d["vols"][s].reshape(3, B), 1e-13, 1e-12, f"step {s} volumes")


def _coupled_oracle():
    wl = W.plain_cpu()
    members = [wl.members[0], W.Member(manning=0.05, retention_m=0.008, f0_mmh=40.0, gamma=1.2,
                                       rain_peak_mmh=110.0, tide_phase=1.3, inflow_peak=220.0)]
    wl.members = members
    ter = wl.terrain
    rain, tide, inflow = wl.forcing()
    n, S, f0, fc, k = wl.fields()
    two = Oracle2D(ter.z, n, ter.dx, dict(west="transmissive", east="stage", south="transmissive",
                                          north="transmissive"), B=2, storage=S, horton=(f0, fc, k, 0.0))
    one = OracleNet([apply_gamma(wl.network, m.gamma) for m in members])
    links = oracle_links(one, ter.z, ter.dx)
    eng = OracleCoupled(two, one, links, dict(east=tide), rain, inflow, tide)
    eta0 = tide(0.0)
    eng.h = np.where(ter.sea[None], np.clip(eta0.reshape(2, 1, 1) - ter.z[None], 0, None), 0.0)
    yn = normal_depth(one, inflow(0.0))
    eng.A = one.area(np.clip(np.maximum(one.arr["z"] + yn, eta0.reshape(2, 1)) - one.arr["z"], 0, None))
    return eng, one, links, yn


def test_exchange_matches_reference(golden):
    d = golden("exchange")
    eng, one, links, _ = _coupled_oracle()
    assert np.array_equal(links["cell1d"], d["cell1d"]) and np.array_equal(links["inv"], d["inv"])
    assert np.array_equal(links["uniq"], d["uniq"])
    for s in range(3):
        h, hu, hv, A = (np.array(d[k][s]) for k in ("h0", "hu0", "hv0", "A0"))
        V = exchange(one, links, h, hu, hv, A, float(d["dt"]))
        assert_rel(V, d["V"][s], 1e-14, 1e-15, "V")
        assert_rel(h, d["h1"][s], RTOL, ATOL, "h")
        assert_rel(A, d["A1"][s], RTOL, ATOL, "A")


def test_coupled_steps_match_reference(golden):
    d = golden("coupled")
    eng, one, links, yn = _coupled_oracle()
    assert_rel(yn, d["yn"], 1e-15, 0, "normal depth")
    assert_rel(eng.total_volume(), d["v0"], 1e-15, 0, "initial volume")
    t = 0.0
    seq = []
    for _ in range(40):
        t = eng.step(t, 1e9)
        seq.append(eng.last)
    assert_rel(np.array([x[0] for x in seq]), d["dt"], 1e-14, 0, "dt")
    assert [x[1] for x in seq] == list(d["n_sub"])
    for k, key in (("h", "h"), ("hu", "hu"), ("hv", "hv"), ("s", "s"), ("ic", "ic"), ("A", "A"), ("Q", "Q")):
        assert_rel(getattr(eng, k), d[key], 1e-12, 1e-14, key)
    for k in ("rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out"):
        assert_rel(eng.acc[k], d[f"src_{k}"], 1e-13, 1e-9, k)


def test_normal_depth_matches_reference(golden):
    d = golden("normal_depth")
    wl = W.plain_cpu()
    assert_rel(normal_depth(OracleNet([wl.network]), float(d["chain_q"]))[0], d["chain_yn"], 1e-15, 0, "chain")
    t2 = W.make_terrain(100, 81, 80.0)
    n2 = OracleNet([W.network60(t2), apply_gamma(W.network60(t2), 0.8)], 0.25)
    assert_rel(normal_depth(n2, d["net60_q"]), d["net60_yn"], 1e-15, 0, "net60")


# ---------------------------------------------------------------- closed forms

def test_lake_at_rest_is_exact_in_the_oracle():
    """Well-balancing: still water over an irregular bed stays still (hydrostatic reconstruction)."""
    ter = W.make_terrain(40, 33, 200.0, seed=3)
    o = Oracle2D(ter.z, 0.03, ter.dx, {e: "closed" for e in ("west", "east", "south", "north")})
    h = np.clip(3.0 - ter.z, 0, None)[None]
    z = np.zeros_like(h)
    for _ in range(20):
        (h, hu, hv, _, _), _ = o.step(h, z, z, z, z, 5.0, 0.0, np.zeros(9))
    assert np.abs(hu).max() < 1e-12 and np.abs(hv).max() < 1e-12
    wet = np.clip(3.0 - ter.z, 0, None) > 0
    assert np.abs((h[0] + ter.z - 3.0)[wet]).max() < 1e-12


def test_flux_consistency_uniform_flow():
    """HLLC of a uniform state equals the physical flux: a uniform, flat, periodic-like state
    with transmissive edges does not change (F(U,U) = f(U))."""
    nx, ny = 12, 7
    o = Oracle2D(np.zeros((ny, nx)), 0.0, 10.0, {e: "transmissive" for e in ("west", "east", "south", "north")})
    h = np.full((1, ny, nx), 1.3)
    hu, hv = 0.7 * h, -0.2 * h
    (h1, hu1, hv1, _, _), vols = o.step(h, hu, hv, 0 * h, 0 * h, 0.1, 0.0, np.zeros(9))
    assert np.abs(h1 - h).max() < 1e-14 and np.abs(hu1 - hu).max() < 1e-13 and np.abs(hv1 - hv).max() < 1e-13
    # net boundary transport of a uniform flow equals (in - out) = 0 up to round-off
    assert abs(vols[0, 0] - vols[1, 0]) < 1e-9 * max(vols[0, 0], 1.0)


def test_dambreak_mass_conserved_and_mirror_symmetric():
    """Closed 1-row dam break: volume constant to round-off; the solution of a mirrored
    initial state is the mirror of the solution (an index or sign slip breaks this)."""
    n = 40
    o = Oracle2D(np.zeros((1, n)), 0.0, 2.5, {e: "closed" for e in ("west", "east", "south", "north")})
    h = np.where(np.arange(n) < 13, 1.0, 0.1)[None, None].astype(float)
    hm = h[..., ::-1].copy()
    z = np.zeros_like(h)
    a = (h, z, z, z, z)
    b = (hm, z, z, z, z)
    for _ in range(30):
        (ha, hua, hva, _, _), _ = o.step(*a, 0.05, 0.0, np.zeros(9))
        (hb, hub, hvb, _, _), _ = o.step(*b, 0.05, 0.0, np.zeros(9))
        a, b = (ha, hua, hva, z, z), (hb, hub, hvb, z, z)
    assert abs(a[0].sum() - h.sum()) < 1e-12 * h.sum()
    assert np.abs(a[0][..., ::-1] - b[0]).max() < 1e-13
    assert np.abs(a[1][..., ::-1] + b[1]).max() < 1e-13


def _ritter_l1(n, order):
    """Reference benchmark 2 (benchmarks/suite.py:122-136): dry-bed dam break, L = 100 m, t = 6 s."""
    dx = 100.0 / n
    x = (np.arange(n) + 0.5) * dx
    o = Oracle2D(np.zeros((1, n)), 0.0, dx, dict(west="transmissive", east="transmissive", south="closed",
                                                 north="closed"), order=order)
    h = np.where(x < 50.0, 1.0, 0.0)[None, None].astype(float)
    z = np.zeros_like(h)
    st = (h, z, z, z, z)
    t = 0.0
    while t < 6.0 - 1e-12:
        m = float(o.max_rate(*st[:3]).max())
        dt = 0.45 / m if m > 0 else 5.0
        dt = min(min(max(dt, 1e-4), 5.0), 6.0 - t)        # SWE2D.max_dt clamp, then run_2d's t_end cap
        (h1, hu1, hv1, s1, i1), _ = o.step(*st, dt, t, np.zeros(9))
        st = (h1, hu1, hv1, s1, i1)
        t += dt
    c0 = math.sqrt(9.81)
    xi = (x - 50.0) / 6.0
    he = np.where(xi <= -c0, 1.0, np.where(xi >= 2 * c0, 0.0, (2 * c0 - xi) ** 2 / (9 * 9.81)))
    return float(np.abs(st[0][0, 0] - he).sum() / np.abs(he).sum())


def test_ritter_l1_matches_published_value():
    """The reference's published Ritter L1 error at N=200, MUSCL (artifacts/benchmark_results.json,
    key 2, o2_L1[0] = 0.005977452599511156), reproduced by the oracle from scratch."""
    assert _ritter_l1(200, 2) == pytest.approx(0.005977452599511156, rel=1e-9)
