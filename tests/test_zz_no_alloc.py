"""his_swe2d_step_noalloc (scratch in a fixed device array) against his_swe2d_step (caller workspace).

Both run the same kernels on the same inputs, so their outputs must agree bit for bit: the five state
fields, the five step volumes and the per-member CFL rates.  Two calls in a row go through the
allocation-free entry point -- a 400 x 325 float32 grid of 4 wet members, then the smaller 48 x 40
float64 reference case -- so the second runs on scratch the first left behind.  A grid larger than the
scratch is refused with HIS_E_WORKSPACE before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from _helpers import assert_rel

pytestmark = pytest.mark.gpu


def _both(two, state, dt_t, pm):
    """(outputs of his_swe2d_step, outputs of his_swe2d_step_noalloc) from copies of one state."""
    from his_amd import _lib as L
    lib = L.load()
    outs = []
    for noalloc in (False, True):
        st = [a.clone() for a in state]
        vols = torch.full((5, two.B), float("nan"), dtype=torch.float64, device=two.device)
        rate = torch.full((two.B,), float("nan"), dtype=torch.float64, device=two.device)
        grid = two._grid()
        if noalloc:
            rc = lib.his_swe2d_step_noalloc(C.byref(grid), two._state(*st), dt_t.data_ptr(), pm.data_ptr(),
                                            vols.data_ptr(), rate.data_ptr(), L.stream(two.device))
        else:
            ws = torch.empty(int(lib.his_swe2d_workspace_bytes(C.byref(grid))), dtype=torch.uint8,
                             device=two.device)
            rc = lib.his_swe2d_step(C.byref(grid), two._state(*st), dt_t.data_ptr(), pm.data_ptr(), vols.data_ptr(),
                                    rate.data_ptr(), ws.data_ptr(), ws.numel(), L.stream(two.device))
        L.check(rc, "noalloc test step")
        torch.cuda.synchronize()
        outs.append((st, vols, rate))
    return outs


def _assert_same(a, b, what):
    (sa, va, ra), (sb, vb, rb) = a, b
    for k, name in enumerate(("h", "hu", "hv", "s_filled", "infil_cum")):
        assert torch.equal(sa[k], sb[k]), f"{what}: {name} differs"
    assert torch.equal(va, vb), f"{what}: volumes differ"
    assert torch.equal(ra, rb), f"{what}: rates differ"
    assert bool(torch.isfinite(va).all()) and bool(torch.isfinite(ra).all()), f"{what}: outputs not written"


def test_noalloc_step_equals_workspace_step_twice(golden, gpu):
    from his_amd import workloads as W
    from his_amd.engine import build_from_workload
    from his_amd.solver.boundaries import edge_values
    from test_gpu_parity import _swe2d_mixed

    # call 1: 400 x 325, float32 bulk, 4 members, wet after half an hour of storm
    wl = W.precision_form(B=4)
    eng, _ = build_from_workload(wl, device=gpu)
    eng.run(1800.0, out_times=[])
    two = eng.two
    state = [getattr(two, k).clone() for k in ("h", "hu", "hv", "s_filled", "infil_cum")]
    assert float(state[0].max()) > 0.5
    dt = two.max_dt()
    dt_t = two._pack([dt, 1800.0])
    pm = two._pack([eng.rain(1800.0)] + edge_values(two.bcs, 1800.0) + edge_values(two.bcs, 1800.0 + dt))
    a, b = _both(two, state, dt_t, pm)
    _assert_same(a, b, "400x325 float32 B=4")
    assert not torch.equal(a[0][0], state[0]), "the step changed nothing"

    # call 2: the smaller 48 x 40 float64 reference case, two members; also against the reference's output
    d = golden("swe2d_mixed")
    sim = _swe2d_mixed(d, gpu)
    st0 = d["states"]
    state = [torch.as_tensor(st0[0][k], device=gpu).contiguous() for k in range(5)]
    dt_t = torch.tensor([[d["dt"][0]] * sim.B, [d["t"][0]] * sim.B], dtype=torch.float64, device=gpu)
    pm = torch.as_tensor(d["pm"][0], dtype=torch.float64).reshape(9, -1).expand(9, sim.B).contiguous().to(gpu)
    a, b = _both(sim, state, dt_t, pm)
    _assert_same(a, b, "48x40 float64 B=2")
    for k in range(5):
        assert_rel(b[0][k].cpu().numpy(), st0[1][k], 1e-9, 1e-13, f"field {k} against the reference")


def test_noalloc_step_refuses_a_grid_beyond_its_scratch(gpu):
    """A float64 400 x 325 grid of 20 members needs more than the scratch holds: HIS_E_WORKSPACE and a
    message naming the sizes; nothing is launched, so the state is untouched."""
    from his_amd import _lib as L
    from his_amd import workloads as W
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import all_edges
    from his_amd.solver.swe2d import SWE2D
    ter = W.make_terrain(400, 325, 20.0)
    two = SWE2D(ter.z, 0.035, ter.dx, SolverConfig(), all_edges("transmissive"), gpu, torch.float64, batch=20)
    lib = L.load()
    assert lib.his_swe2d_workspace_bytes(C.byref(two._grid())) > L.SWE2D_NOALLOC_MAX_BYTES
    two.set_level(1.0)
    h0 = two.h.clone()
    dt_t = two._pack([0.1, 0.0])
    pm = two._pack([0.0] * 9)
    vols = torch.zeros((5, two.B), dtype=torch.float64, device=gpu)
    rc = lib.his_swe2d_step_noalloc(C.byref(two._grid()), two._state(two.h, two.hu, two.hv, two.s_filled,
                                                                        two.infil_cum),
                                    dt_t.data_ptr(), pm.data_ptr(), vols.data_ptr(), None, L.stream(gpu))
    assert rc == L.HIS_E_WORKSPACE
    assert b"scratch" in lib.his_last_error()
    torch.cuda.synchronize()
    assert torch.equal(two.h, h0)
