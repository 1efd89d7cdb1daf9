"""GPU parity of the batched pentadiagonal API (custenPentFactorBatch /
custenPentSolveBatch, the host counterparts of cuPentBatch.cu:35-198) against
the CPU oracle: factors and solutions bit-identical, solutions also checked by
their residual."""
import numpy as np
import pytest

import oracle
from custen_amd import api
from tests.helpers import bits_equal, mismatches
from tests.test_oracle import _pent_systems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,batch", [(4, 1), (5, 63), (17, 100), (64, 65), (1000, 4096),
                                     (9, 70000)])
def test_factor_and_solve_match_oracle(gpu, m, batch):
    print(f"pent m={m} batch={batch}", flush=True)
    host = _pent_systems(m, batch, 500 + m)
    dev = [gpu.from_numpy(a.copy()).cuda() for a in host]
    api.custenPentFactorBatch(*dev[:5], m, batch)
    api.custenPentSolveBatch(*dev[:6], m, batch)
    gpu.cuda.synchronize()
    fac = oracle.pent_factor_batch(*host[:5], nthreads=8)
    x = oracle.pent_solve_batch(*fac, host[5], nthreads=8)
    for got, ref in zip(dev[:5], fac):
        g = got.cpu().numpy()
        assert bits_equal(g, ref), mismatches(g, ref)
    g = dev[5].cpu().numpy()
    assert bits_equal(g, x), mismatches(g, x)
    # residual of a few systems against the original matrices
    ds, dl, d, du, dw, b = host
    for k in range(0, batch, max(1, batch // 5)):
        y = g[:, k]
        r = d[:, k] * y
        r[1:] += dl[1:, k] * y[:-1]
        r[2:] += ds[2:, k] * y[:-2]
        r[:-1] += du[:-1, k] * y[1:]
        r[:-2] += dw[:-2, k] * y[2:]
        assert np.abs(r - b[:, k]).max() < 1e-13


def test_solve_reuses_factors_on_a_stream(gpu):
    """One factorisation, several right-hand sides, launched on a caller stream."""
    m, batch = 128, 512
    host = _pent_systems(m, batch, 900)
    dev = [gpu.from_numpy(a.copy()).cuda() for a in host[:5]]
    s = gpu.cuda.Stream()
    with gpu.cuda.stream(s):
        api.custenPentFactorBatch(*dev, m, batch, stream=s)
        outs = []
        for j in range(3):
            rhs = host[5] * (j + 1) - j
            db = gpu.from_numpy(rhs.copy()).cuda()
            api.custenPentSolveBatch(*dev, db, m, batch, stream=s)
            outs.append((rhs, db))
    s.synchronize()
    fac = oracle.pent_factor_batch(*host[:5])
    for rhs, db in outs:
        ref = oracle.pent_solve_batch(*fac, rhs)
        assert bits_equal(db.cpu().numpy(), ref)


def test_rejects_host_arrays(gpu):
    m, batch = 8, 4
    host = _pent_systems(m, batch, 1)
    with pytest.raises(api.CustenError, match="device memory"):
        api.custenPentFactorBatch(*host[:5], m, batch)
