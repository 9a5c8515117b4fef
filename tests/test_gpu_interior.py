"""Every compiled 2-D kernel shape on a grid large enough to hold strips and row chunks that touch
no grid edge, against the float64 oracle.

The register-marching kernels take a leaner path for waves whose strip and rows lie wholly inside
the grid; the reference fixtures (48 x 40) are too small to contain one, so this case supplies it:
190 x 130, two members, wet and dry cells with fronts, rain, Horton infiltration, a retention
store and all four edge kinds, two steps; a second set of edges swaps the kinds between the x and
y sides, so that each kind is met both by strips at a west or east edge and by row chunks at a south
or north edge.  float64 is held to 1e-9 relative of the oracle and float32 to float32 round-off.
"""
import numpy as np
import pytest
import torch

from _helpers import assert_rel

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-9, 1e-13
NY, NX, B = 130, 190, 2
# edge kinds, and the forcing rows (W, E, S, N at t and at t + dt) their valued edges take
BCSETS = {
    "we_transmissive_stage": dict(west="transmissive", east="stage", south="discharge", north="closed"),
    "we_discharge_closed": dict(west="discharge", east="closed", south="transmissive", north="stage"),
}
EDGE_VALUES = {"stage": (2.1, 1.8), "discharge": (0.4, 0.7)}


def _case(bcset):
    rng = np.random.default_rng(20261203)
    yy, xx = np.mgrid[0:NY, 0:NX].astype(np.float64)
    z = 1.0 + 0.012 * xx - 0.006 * yy + 0.6 * np.sin(xx / 9.0) * np.cos(yy / 7.0)
    eta = 1.9 + 0.4 * np.sin(xx / 23.0 + yy / 17.0)
    h = np.stack([np.clip(eta + 0.2 * b - z + 0.05 * rng.standard_normal((NY, NX)), 0.0, None) for b in range(B)])
    h[h < 0.02] = 0.0
    u = 0.8 * rng.standard_normal((B, NY, NX))
    v = 0.8 * rng.standard_normal((B, NY, NX))
    hu, hv = h * u, h * v
    n = 0.02 + 0.04 * rng.random((B, NY, NX))
    S = 0.01 * rng.random((B, NY, NX))
    s = S * rng.random((B, NY, NX))
    f0 = np.full((B, NY, NX), 60.0 / 3.6e6)
    fc = np.full((B, NY, NX), 10.0 / 3.6e6)
    k = np.full((B, NY, NX), 2.0 / 3600.0)
    ic = np.zeros((B, NY, NX))
    # rain, W, E, S, N at t and at t + dt (stage levels in m, inflow per metre in m^2/s)
    bcs = BCSETS[bcset]
    pm = np.zeros((9, 2))
    pm[0] = [40.0 / 3.6e6, 60.0 / 3.6e6]
    for e, name in enumerate(("west", "east", "south", "north")):
        v = EDGE_VALUES.get(bcs[name])
        if v:
            pm[1 + e] = v
            pm[5 + e] = (v[0] + 0.05, v[1] + 0.05)
    return dict(z=z, n=n, S=S, f0=f0, fc=fc, k=k, state=[h, hu, hv, s, ic], pm=pm, dx=20.0, bcs=bcs)


def _sim(c, gpu, dtype):
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import EdgeBC
    from his_amd.solver.infiltration import Horton
    from his_amd.solver.swe2d import SWE2D
    value = {"stage": 2.0, "discharge": 0.5}
    bcs = {e: EdgeBC(kind, value[kind]) if kind in value else EdgeBC(kind) for e, kind in c["bcs"].items()}
    T = lambda a: torch.as_tensor(a, dtype=dtype, device=gpu)
    hz = Horton(T(c["f0"]), T(c["fc"]), T(c["k"]), 0.0)
    return SWE2D(c["z"], c["n"], c["dx"], SolverConfig(), bcs, gpu, dtype, storage=c["S"], horton=hz, batch=B)


@pytest.fixture(scope="module", params=sorted(BCSETS))
def reference(request):
    """Two oracle steps from the case's state: (states[0..2], dt[0..1], vols[0..1], member CFL rates of
    the first state)."""
    from oracle.oracle import Oracle2D
    c = _case(request.param)
    orc = Oracle2D(c["z"], c["n"], c["dx"], c["bcs"], B=B, storage=c["S"], horton=(c["f0"], c["fc"], c["k"], 0.0))
    states, dts, vols = [c["state"]], [], []
    rate0 = orc.max_rate(*c["state"][:3])
    t = 600.0
    for _ in range(2):
        st = states[-1]
        dt = 0.4 / float(orc.max_rate(st[0], st[1], st[2]).max())
        out, vo = orc.step(*st, dt, t, c["pm"])
        states.append(out)
        dts.append(dt)
        vols.append(vo)
        t += dt
    return c, states, dts, vols, rate0


def _f64(t):
    return t.detach().double().cpu().numpy()


@pytest.mark.parametrize("shape", range(9))
def test_every_shape_on_a_grid_with_interior_strips(gpu, reference, shape):
    from his_amd import _lib
    c, states, dts, vols, rate0 = reference
    try:
        for dtype in (torch.float64, torch.float32):
            _lib.set_tile_shape(dtype, shape)
            sim = _sim(c, gpu, dtype)
            t = 600.0
            for s in range(2):
                ins = [torch.as_tensor(states[s][k], dtype=dtype, device=gpu) for k in range(5)]
                sc = torch.tensor([dts[s], t], dtype=torch.float64)
                out = sim._core(*ins, sc, torch.as_tensor(c["pm"]).reshape(9, -1, 1, 1))
                label = f"{dtype} shape {shape} step {s}"
                if dtype == torch.float64:
                    for k in range(5):
                        assert_rel(_f64(out[k]), states[s + 1][k], RTOL, ATOL, f"{label} field {k}")
                    assert_rel(np.stack([_f64(v) for v in out[5:]]), vols[s], RTOL, 1e-9, f"{label} vols")
                else:
                    # depth to float32 round-off; momentum loosely, since cells within round-off of
                    # h_dry may take the other wet/dry branch in float32
                    href = states[s + 1][0]
                    assert np.abs(_f64(out[0]) - href).max() < 1e-5 * max(1.0, np.abs(href).max()), label
                    qmax = np.abs(states[s + 1][1]).max()
                    assert np.abs(_f64(out[1]) - states[s + 1][1]).max() < 1e-3 * qmax, label
                t += dts[s]
            rates = _f64(sim.member_rates(*(torch.as_tensor(states[0][k], dtype=dtype, device=gpu)
                                            for k in range(3))))
            assert_rel(rates, rate0, RTOL if dtype == torch.float64 else 1e-5, 0, f"{dtype} shape {shape} rate")
    finally:
        _lib.set_tile_shape(torch.float64, -1)
        _lib.set_tile_shape(torch.float32, -1)
