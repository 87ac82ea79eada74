"""tfsm_step_noalloc (the allocation-free variant of tfsm_step) against tfsm_step, through the C ABI.

Two calls in a row on one GPU, the second with a smaller plan (fewer samples, smaller grid), each
compared bit for bit with tfsm_step on a caller workspace from the same inputs; a plan whose
workspace exceeds the static scratch is rejected with ValueError before anything runs."""

import ctypes
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _ns_plan(n, B, dev):
    from torchfsm_amd import _lib
    from torchfsm_amd.engine import SpectralPlan

    plan = SpectralPlan([(0, 2 * math.pi, n)] * 2, B, 1, nonlinear="vc", nonlinear_coef=-1.0,
                        linear=[("lap", 1e-3), ("unit", -0.1)], device=dev)
    plan.set_scheme("SETDRK4", 1e-3)
    assert plan.workspace().numel() <= _lib.NOALLOC_SCRATCH_BYTES
    return plan


def _noalloc(plan, uh, nsteps, dev):
    from torchfsm_amd._lib import check, lib

    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    with torch.cuda.device(dev):
        check(lib.tfsm_step_noalloc(plan.handle, ctypes.c_void_p(uh.data_ptr()), int(nsteps), None, None, 0, stream),
              "step_noalloc")


def _case(n, B, seed, steps, dev):
    plan = _ns_plan(n, B, dev)
    g = torch.Generator().manual_seed(seed)
    u0 = torch.randn(B, 1, n, n, generator=g).to(dev)
    a = plan.forward(u0)
    b = a.clone()
    plan.step(a, steps)
    _noalloc(plan, b, steps, dev)
    torch.cuda.synchronize(dev)
    assert torch.isfinite(torch.view_as_real(a)).all()
    assert torch.equal(a, b), (n, B, float((a - b).abs().max()))
    plan.close()


def test_noalloc_step_matches_step_twice_second_smaller(dev):
    # first call: 128^2, 8 samples (the scratch then holds that call's stage arrays and fields);
    # second call: 64^2, 3 samples, reusing the same scratch with different offsets and sizes
    _case(128, 8, 0, 12, dev)
    _case(64, 3, 1, 9, dev)


def test_noalloc_step_rejects_plans_beyond_the_scratch(dev):
    from torchfsm_amd.engine import SpectralPlan

    plan = SpectralPlan([(0, 2 * math.pi, 512)] * 2, 16, 1, nonlinear="vc", nonlinear_coef=-1.0,
                        linear=[("lap", 1e-3)], device=dev)
    plan.set_scheme("SETDRK4", 1e-3)
    uh = plan.empty_spectrum()
    with pytest.raises(ValueError, match="exceeds the static scratch"):
        _noalloc(plan, uh, 1, dev)
    plan.close()
