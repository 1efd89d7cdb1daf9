"""custenCHAdvanceNoAlloc (the allocation-free entry point) against the
existing entry point custenCHRun (mode 0) and the oracle, bit for bit, on two
calls in a row, the second with a smaller grid (the shared work array keeps
the first call's contents), then the refusal of a grid above its maximum."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import mismatches

pytestmark = pytest.mark.gpu


def ic(n, r):
    return synth.uniform(n * n, 7000 + r, -0.1, 0.1).reshape(n, n)


def reference_run(n, steps, c0):
    s = api.CHSolver(api.ch_params(n), c0)
    s.run(steps, mode=0)
    c, co = s.field()
    s.close()
    return c, co


@pytest.mark.parametrize("calls", [[(256, 5), (128, 3)], [(1024, 2), (96, 4)]])
def test_no_alloc_matches_existing_entry_point(gpu, calls):
    torch = gpu
    for k, (n, steps) in enumerate(calls):
        print(f"no-alloc call {k}: n={n} steps={steps}", flush=True)
        c0 = ic(n, k)
        dc = torch.from_numpy(c0.copy()).cuda()
        dco = torch.from_numpy(c0.copy()).cuda()  # cOld starts equal to c
        api.ch_advance_no_alloc(api.ch_params(n), dc, dco, steps,
                                stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got_c, got_co = dc.cpu().numpy(), dco.cpu().numpy()
        ref_c, ref_co = reference_run(n, steps, c0)
        assert mismatches(got_c, ref_c) == 0, f"call {k}: c differs from custenCHRun"
        assert mismatches(got_co, ref_co) == 0, f"call {k}: cOld differs from custenCHRun"
        if n <= 256:
            o = oracle.CH(n)
            rc, rco = c0.ravel().copy(), c0.ravel().copy()
            o.step(rc, rco, steps=steps)
            assert mismatches(got_c, rc.reshape(n, n)) == 0


def test_no_alloc_continues_a_simulation(gpu):
    """Two calls on the same fields equal one call of the total step count."""
    torch = gpu
    n = 160
    c0 = ic(n, 9)
    dc = torch.from_numpy(c0.copy()).cuda()
    dco = torch.from_numpy(c0.copy()).cuda()
    api.ch_advance_no_alloc(api.ch_params(n), dc, dco, 3)
    api.ch_advance_no_alloc(api.ch_params(n), dc, dco, 4)
    torch.cuda.synchronize()
    ref_c, ref_co = reference_run(n, 7, c0)
    assert mismatches(dc.cpu().numpy(), ref_c) == 0
    assert mismatches(dco.cpu().numpy(), ref_co) == 0


def test_no_alloc_refuses_above_maximum(gpu):
    torch = gpu
    n = 1040
    dc = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    dco = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    with pytest.raises(api.CustenError) as ei:
        api.ch_advance_no_alloc(api.ch_params(n), dc, dco, 1)
    assert ei.value.code == 1 and "1024" in str(ei.value)
    api.clear_error()
