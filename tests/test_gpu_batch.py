"""Batched uneven grids: every grid after k steps (fused graph launches and the
per-grid loop) equals k oracle applies with swaps, bit for bit."""
import numpy as np
import pytest

import oracle
from custen_amd import api, synth
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("mode,config", [(0, None), (1, None), (0, (8, 4, False)),
                                         (0, (4, 2, True)), (0, (8, 1, True)),
                                         (0, (4, 3, True)), (1, (4, 2, True))])
def test_batch_matches_oracle(gpu, f32, mode, config):
    """config None: the autotuner's choice; otherwise (rows per wave, chunks,
    band walker).  Grids narrower than 128 columns, of 128 and 256 columns and
    of 200 (one walked strip and one edge strip) exercise both kernels of a
    band-walker step."""
    torch = gpu
    sizes = [64, 128, 40, 256, 128, 96, 200]
    steps = 3
    h = 2 * np.pi / 128
    pat = np.array([[0, 0, 1, 0, 0], [0, 2, -8, 2, 0], [1, -8, 20, -8, 1],
                    [0, 2, -8, 2, 0], [0, 0, 1, 0, 0]], dtype=float)
    w = (pat * 1e-3).ravel()
    dt = np.float32 if f32 else np.float64
    fields = [synth.field(n, n, 1000 + g).astype(dt) for g, n in enumerate(sizes)]
    ins = [torch.from_numpy(f.copy()).cuda() for f in fields]
    outs = [torch.zeros_like(t) for t in ins]
    dw = torch.from_numpy(w).cuda()
    b = api.Batch(outs, ins, dw, f32=f32)
    if config is not None:
        b.set_tile_config(*config)
        assert b.tile_config() == (config[0], config[1], int(config[2]))
    b.run(steps, mode=mode)
    torch.cuda.synchronize()
    assert b.parity == steps % 2
    cur = outs if b.parity == 1 else ins
    wt = w.astype(dt)
    for g, f in enumerate(fields):
        ref = f
        for _ in range(steps):
            ref = oracle.apply(ref, wt, 5, 2, 5, 2, True)
        assert bits_equal(cur[g].cpu().numpy(), ref), f"grid {g} ({sizes[g]}) differs"
    # a second run continues from the current parity
    b.run(2, mode=mode)
    torch.cuda.synchronize()
    cur = outs if b.parity == 1 else ins
    for g, f in enumerate(fields):
        ref = f
        for _ in range(steps + 2):
            ref = oracle.apply(ref, wt, 5, 2, 5, 2, True)
        assert bits_equal(cur[g].cpu().numpy(), ref)
    b.close()
