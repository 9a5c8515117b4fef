"""The exact solutions the analytical cases are judged against, pinned by the physics they
must satisfy (CPU only): conservation, Rankine-Hugoniot jump conditions, Riemann invariants,
Bernoulli, continuity."""
import json
import math
from pathlib import Path

import numpy as np

from his_amd import benchmarks as BM

G = 9.81


def test_ritter_conserves_mass_and_has_textbook_front():
    x = np.linspace(-200.0, 300.0, 2_000_001)
    h, u = BM.ritter_exact(x, 6.0, 50.0, 1.0)
    assert abs(np.trapezoid(h, x) - (50.0 + 200.0) * 1.0) < 1e-6     # water left of the dam: 250 m x 1 m
    c0 = math.sqrt(G)
    front = x[np.nonzero(h > 0)[0].max()]
    assert abs(front - (50.0 + 2 * c0 * 6.0)) < 1e-3                  # dry front at x0 + 2 c0 t
    fan = (x > 50 - c0 * 6) & (x < 50 + 2 * c0 * 6)
    # the left-going Riemann invariant u + 2c is constant through the fan: 2 c0
    assert np.allclose(u[fan] + 2 * np.sqrt(G * h[fan]), 2 * c0, atol=1e-12)


def test_wet_riemann_satisfies_the_jump_conditions():
    hL, hR = 1.0, 0.1
    hs, us = BM.star_region(hL, 0.0, hR, 0.0)
    assert hR < hs < hL
    # right shock: mass and momentum flux are continuous in the shock frame
    S = us * hs / (hs - hR)
    mass_l, mass_r = hs * (us - S), hR * (0.0 - S)
    mom_l = hs * (us - S) ** 2 + 0.5 * G * hs * hs
    mom_r = hR * S * S + 0.5 * G * hR * hR
    assert abs(mass_l - mass_r) < 1e-12 and abs(mom_l - mom_r) < 1e-11
    # left rarefaction: u + 2c carried from the left state
    assert abs(us + 2 * math.sqrt(G * hs) - 2 * math.sqrt(G * hL)) < 1e-12
    x = np.linspace(-400.0, 500.0, 900_001)
    h, u = BM.riemann_exact(x, 6.0, 50.0, hL, 0.0, hR, 0.0)
    assert abs(np.trapezoid(h, x) - (450.0 * 1.0 + 450.0 * 0.1)) < 1e-3   # mass conserved (the shock is a step)
    sr = 0.0 + math.sqrt(G * hR) * math.sqrt(0.5 * hs * (hs + hR)) / hR
    assert abs(sr - S) < 1e-12                                        # the sampled shock moves at the RH speed


def test_thacker_is_exact_for_the_continuity_equation():
    a, h0, eta, L = 1.0, 0.1, 0.5, 4.0
    n = 1600
    xc = (np.arange(n) + 0.5) * (L / n)
    X, Y = np.meshgrid(xc, xc)
    vols = [BM.thacker_planar(X, Y, t, a, h0, eta, L)[0].sum() * (L / n) ** 2 for t in (0.0, 0.7, 1.9, 3.3)]
    assert max(vols) - min(vols) < 1e-4 * vols[0]
    T = BM.thacker_planar(X, Y, 0.0, a, h0, eta, L)[4]
    assert abs(T - 2 * math.pi * a / math.sqrt(2 * G * h0)) < 1e-12
    h_T = BM.thacker_planar(X, Y, T, a, h0, eta, L)[0]
    assert np.allclose(h_T, BM.thacker_planar(X, Y, 0.0, a, h0, eta, L)[0], atol=1e-12)


def test_bump_profiles_keep_energy_and_discharge():
    x = np.linspace(0.0, 25.0, 2001)
    z = BM.bump_bed(x)
    for case in ("subcritical", "transcritical"):
        h, q, info = BM.bump_exact(x, case)
        E = q * q / (2 * G * h * h) + h + z
        assert np.ptp(E) < 1e-9, case
    h, q, info = BM.bump_exact(x, "transcritical")
    hc = (q * q / G) ** (1 / 3)
    assert abs(h[np.argmin(np.abs(x - 10.0))] - hc) < 1e-3            # critical at the crest
    h, q, info = BM.bump_exact(x, "shock")
    xs = info["x_shock"]
    zs = BM.bump_bed(np.array([xs]))
    hc = (q * q / G) ** (1 / 3)
    h1 = BM._depth_at_energy(q, 1.5 * hc + 0.2, zs, True)[0]
    h2 = BM._depth_at_energy(q, q * q / (2 * G * 0.33 ** 2) + 0.33, zs, False)[0]
    conj = 0.5 * h1 * (math.sqrt(1 + 8 * q * q / (G * h1 ** 3)) - 1)
    assert 10.0 < xs < 25.0 and abs(conj - h2) < 1e-3                 # Belanger across the jump
    assert h[x < xs - 0.1].min() > 0 and np.all(h[x > xs + 0.1] > h1)


def test_kinematic_plane_is_continuous_at_equilibrium():
    L, S, n, i = 200.0, 0.05, 0.03, 100.0 / 3.6e6
    q, te = BM.kinematic_plane([0.0], L, S, n, i, 1500.0)
    assert abs(te - 478.3) < 0.05                                     # the reference's printed t_equilibrium
    qa, _ = BM.kinematic_plane([te * (1 - 1e-9), te * (1 + 1e-9)], L, S, n, i, 1500.0)
    assert abs(qa[0] - qa[1]) < 1e-8 * i * L


def test_fixture_is_the_published_report():
    """The fixture's values print as the reference's report prints them (benchmark_report.md)."""
    REF = json.loads((Path(__file__).parent / "golden" / "analytical_reference.json").read_text())["metrics"]
    P = BM.PUBLISHED
    for case, key in (("ritter", "2"), ("stoker", "3")):
        for name in ("o1_L1", "o2_L1"):
            got = [float(f"{REF[f'{key}.{name}[{i}]']['local']:.4g}") for i in range(4)]
            assert got == P[case][name]
    assert np.isclose(REF["6.t_equilibrium_s"]["local"], P["plane"]["t_equilibrium_s"], atol=0.05)
