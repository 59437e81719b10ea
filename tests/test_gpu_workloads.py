"""Full-size workloads on the GPU: parity against the oracle at the sizes the
benchmark uses, and the size-independent properties (round trips)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_headline_workload_parity(gpu, oracle):
    """The benchmark's batched workload (4096 signals, ~6.3e8 samples, random
    wavelets and levels): GPU forward and forward-adjoint equal the oracle per
    signal within 4 ulp of the signal's max |coefficient|."""
    import torch
    import bench
    from lifted_amd.batch import BatchPlan
    sigs, total = bench.make_workload()
    g = torch.Generator(device=gpu)
    g.manual_seed(bench.SEED)
    x = torch.randn(total, generator=g, device=gpu, dtype=torch.float32)
    y = BatchPlan(sigs, 0).execute(x)
    z = BatchPlan(sigs, 2).execute(y)
    xh, yh, zh = x.cpu().numpy(), y.cpu().numpy(), z.cpu().numpy()
    yr = oracle.batch(xh, sigs, oracle.FORWARD, oracle.ZERO, n_threads=16)
    zr = oracle.batch(yh, sigs, oracle.FORWARD_ADJOINT, oracle.ZERO, n_threads=16)
    eps = np.finfo(np.float32).eps
    worst = 0.0
    for off, n, w, L in sigs:
        for got, ref in ((yh, yr), (zh, zr)):
            a, b = got[off:off + n], ref[off:off + n]
            e = float(np.abs(a - b).max() / (eps * max(np.abs(b).max(), 1e-30)))
            worst = max(worst, e)
            assert e <= 4.0, (off, n, w, L, e)
    print(f"max error {worst:.3f} ulp of max|coef| over {len(sigs)} signals")
