"""GPU tests of the MIMO throughput form (fp32 matrix cores): every output
against the fp64 sum of per-pair convolutions, for a full 32x32 case and a
non-square smaller case (zero-padded tiles)."""
import numpy as np
import pytest

from oracle import oracle as O
from partconv_amd.workloads import cfg4_irs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_in,n_out,taps,T,N,P", [
    (32, 32, 20000, 12000, [256, 4096], [30, 0]),
    (5, 3, 20000, 9000, [128, 1024, 8192], [14, 14, 0]),
])
def test_mimo_mfma_matches_fp64(n_in, n_out, taps, T, N, P):
    import torch
    from partconv_amd.mimo import MimoConv
    print(f"case mimo {n_in}x{n_out} taps={taps}", flush=True)
    irs = cfg4_irs(seed=4, n_in=n_in, n_out=n_out, taps=taps)
    h = np.stack([np.stack(irs[o]) for o in range(n_out)]).astype(np.float32)
    rng = np.random.default_rng(n_in * 100 + n_out)
    x = rng.standard_normal((n_in, T)).astype(np.float32)
    mx = MimoConv(N, P, taps, n_in, n_out)
    mx.set_filters(torch.from_numpy(h).cuda())
    y = mx(torch.from_numpy(x).cuda()).cpu().numpy()
    for o in range(n_out):
        r = sum(O.direct_conv_f64(x[i], h[o, i]) for i in range(n_in))
        err = float(np.max(np.abs(y[o] - r)) / np.sqrt(np.mean(r ** 2)))
        assert err < 1e-5, (o, err)
