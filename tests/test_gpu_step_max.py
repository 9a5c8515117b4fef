"""The every-step depth maximum (``his_engine_track_step_max``).

Folded after the 2-D stages of every coupled step, it must equal the maximum of the depth taken on
the host after each single step; an engine stepping through the 32-step, 8-step and single-step
graphs must reach the same maximum and state; and stopping it must leave the maximum as it was.  The maximum is exact
arithmetic, so the comparisons are bit for bit.
"""
import pytest
import torch

from test_gpu_parity import coupled_case

pytestmark = pytest.mark.gpu


def _start(eng, hmax):
    eng.track_depth_max_every_step(hmax)
    dev = eng._device()
    dev.t.fill_(0.0)
    dev.t_end.fill_(1e12)
    dev.t_limit.fill_(1e12)
    dev.refresh()
    return dev


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_step_max_is_the_max_over_every_step(gpu, precision):
    eng, _ = coupled_case(gpu, precision)
    two = eng.two
    hmax = torch.zeros_like(two.h)
    dev = _start(eng, hmax)
    ref = torch.zeros_like(two.h)
    for _ in range(44):
        dev.step(1)
        torch.maximum(ref, two.h, out=ref)
    torch.cuda.synchronize()
    assert bool((ref > 0).any())
    assert torch.equal(hmax, ref)

    # 44 steps in one call: the 32-step graph, the 8-step graph and four single steps; the same
    # maximum and the same state
    eng2, _ = coupled_case(gpu, precision)
    hmax2 = torch.zeros_like(eng2.two.h)
    dev2 = _start(eng2, hmax2)
    dev2.step(44)
    torch.cuda.synchronize()
    assert torch.equal(eng2.two.h, two.h)
    assert torch.equal(hmax2, hmax)

    # stopped: later steps leave it alone
    eng.track_depth_max_every_step(None)
    frozen = hmax.clone()
    dev.step(4)
    torch.cuda.synchronize()
    assert torch.equal(hmax, frozen)


def test_step_max_rejects_a_wrong_buffer(gpu):
    eng, _ = coupled_case(gpu, "fp64")
    dev = eng._device()
    with pytest.raises(ValueError):
        dev.track_step_max(0, torch.zeros(3, 4, dtype=torch.float64, device=gpu))
    with pytest.raises(ValueError):
        dev.track_step_max(0, torch.zeros_like(eng.two.h, dtype=torch.float32))
