"""Host-side logic: topology, CSR scatter order, links, forcing series, workloads,
and the C ABI library's exported symbols (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from his_amd import workloads as W
from his_amd.domain.channels import apply_gamma
from his_amd.forcing import Harmonic, PiecewiseConstant, PiecewiseLinear, sine_tide, triangle
from his_amd.solver.swe1d import build_topology, csr

ROOT = Path(__file__).resolve().parents[1]


def test_topology_matches_reference_layout(golden):
    """Cell kinds and faces equal the reference's flattening (golden files carry its arrays)."""
    ter = W.make_terrain(100, 81, 80.0)
    d = golden("swe1d_net60")
    tp = build_topology(apply_gamma(W.network60(ter), 0.8))
    assert np.array_equal(tp.kind, d["kind"]) and np.array_equal(tp.face_l, d["fl"])
    assert np.array_equal(tp.face_r, d["fr"])
    d = golden("swe1d_chain")
    tp = build_topology(W.chain_network(ter, 20))
    assert np.array_equal(tp.kind, d["kind"]) and np.array_equal(tp.face_l, d["fl"])


def test_csr_preserves_index_add_order():
    idx = np.array([3, 1, 3, 0, 1, 3])
    ptr, items = csr(idx, 5)
    assert list(ptr) == [0, 1, 3, 3, 6, 6]
    assert list(items[ptr[3]:ptr[4]]) == [0, 2, 5]        # increasing position, as index_add adds
    vals = np.array([0.1, 1e16, 0.3, 2.0, -1e16, 0.7])
    seq = np.zeros(5)
    for p, t in enumerate(idx):
        seq[t] += vals[p]
    got = np.array([sum(vals[items[ptr[t]:ptr[t + 1]]], 0.0) if ptr[t + 1] > ptr[t] else 0.0 for t in range(5)])
    got2 = np.zeros(5)
    for t in range(5):
        for p in items[ptr[t]:ptr[t + 1]]:
            got2[t] += vals[p]
    assert np.array_equal(got2, seq)


def test_series_forms():
    pl = PiecewiseLinear(10.0, 5.0, [[1.0, 3.0, 2.0], [0.0, 1.0, 0.0]])
    assert np.allclose(pl(0.0), [1.0, 0.0]) and np.allclose(pl(12.5), [2.0, 0.5]) and np.allclose(pl(99), [2, 0])
    pc = PiecewiseConstant(0.0, 10.0, [[1.0, 2.0]])
    assert pc(-1.0)[0] == 0.0 and pc(9.99)[0] == 1.0 and pc(10.0)[0] == 2.0 and pc(20.0)[0] == 0.0
    td = sine_tide(0.8, 3600.0, [0.0, np.pi / 2])
    assert np.allclose(td(900.0), [0.8, 0.0], atol=1e-12)
    tr = triangle(0.0, 100.0, [4.0], base=1.0, batched=False)
    assert tr(50.0) == pytest.approx(4.0) and tr(25.0) == pytest.approx(2.5) and tr(150.0) == 1.0


def test_workload_shapes_and_draws():
    wl = W.precision_form()
    assert wl.terrain.z.shape == (325, 400) and wl.batch == 32
    assert wl.network.n_reaches == 60 and len(wl.network.junctions) == 2
    m = wl.members
    assert all(0.02 <= x.manning <= 0.08 and 0.7 <= x.gamma <= 1.3 and 40 <= x.rain_peak_mmh <= 120 for x in m)
    assert wl.terrain.z[:, -5:].max() < 0 and wl.terrain.z[:, :-5].min() > -4.0
    c1 = W.plain_cpu()
    assert c1.terrain.z.shape == (81, 100) and c1.network.n_reaches == 20
    groups = W.batching_form()
    assert [g.batch for g in groups] == [16, 8, 8]
    assert [g.terrain.z.shape for g in groups] == [(81, 100), (163, 200), (325, 400)]
    reqs = W.serving_requests(50)
    assert all(0.5 * 3600 <= r["storm_s"] <= 6 * 3600 for r in reqs)
    assert np.all(np.diff([r["arrival_s"] for r in reqs]) > 0)


def test_library_exports_every_declared_symbol():
    """libhis_amd.so loads without a GPU and exports each function include/his_amd.h declares."""
    from his_amd import _lib
    lib = _lib.load()
    header = (ROOT / "include" / "his_amd.h").read_text()
    names = set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*\*?\s*(his_\w+)\s*\(", header, re.M))
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(_lib.SYMBOLS) == names
    assert lib.his_version().startswith(b"his_amd")


def test_solvers_refuse_to_run_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from his_amd.config import SolverConfig
    from his_amd.solver.boundaries import all_edges
    from his_amd.solver.swe2d import SWE2D
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        SWE2D(np.zeros((4, 4)), 0.03, 10.0, SolverConfig(), all_edges(), "cpu")


def test_serving_tables_and_output_times():
    """Serving storms sit on the knot grid (peak on a knot), outputs every 15 min plus the end."""
    from his_amd import serving as S
    assert S.storm_length(1801.0) == 1800.0 and S.storm_length(10.0) == 2 * S.STEP_S
    assert S.storm_length(9 * 3600.0) == S.MAX_STORM_S
    v = S.triangle_knots(1800.0, 5.0, 1.0)
    t = np.arange(S.K_KNOTS) * S.STEP_S
    assert v.shape == (S.K_KNOTS,) and v[0] == 1.0 and v[15] == 5.0 and v[30] == 1.0 and (v[31:] == 1.0).all()
    # linear between knots: the closed form of the triangle at every knot
    ref = 1.0 + 4.0 * np.maximum(0.0, 1.0 - np.abs(t - 900.0) / 900.0)
    assert np.allclose(v, ref, rtol=0, atol=1e-15)
    assert S.output_times(1800.0) == [0.0, 900.0, 1800.0]
    assert S.output_times(1000.0) == [0.0, 900.0, 1000.0]
    fo = S.request_forcing(dict(storm_s=1800.0, member=S.W.Member(rain_peak_mmh=72.0, inflow_peak=200.0)))
    assert abs(float(fo.rain(900.0)[0]) - 72.0 / 3.6e6) < 1e-18 and float(fo.rain(1800.0)[0]) == 0.0
    assert abs(float(fo.inflow(450.0)[0]) - 105.0) < 1e-12 and fo.t_end == 1800.0


def test_shape_table_matches_the_shape_tests():
    """Every compiled 2-D kernel shape is exercised by the GPU parity test over shapes."""
    import torch
    from his_amd import _lib as L
    n = L.num_tile_shapes()
    src = open(__file__.replace("test_host.py", "test_gpu_parity.py")).read()
    m = re.search(r'parametrize\("shape", range\((\d+)\)\)', src)
    assert m and int(m.group(1)) == n
    assert [L.tile_shape(torch.float32, k)[2] for k in range(n)].count(2) >= 3
