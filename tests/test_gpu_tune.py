"""GPU half of the partition autotuner: per-level block costs measured on this
GPU (PartConvFilter.bench, HIP events), the partition the dynamic program picks
from them, and that partition running through the engine against fp64."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import rules as R
from partconv_amd import tune
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu


def test_measured_model_and_tuned_partition_run():
    print("case autotune N0=64", flush=True)
    from partconv_amd import PartConvMulti, last_error
    model = tune.measure_model(64, max_block=4096, trials=50, max_time=0.01)
    assert model.sizes() == [64, 128, 256, 512, 1024, 2048, 4096]
    for n in model.sizes():
        t1, t2 = model.samples[n]
        assert 0 < t1 < 1e-2 and 0 < t2 < 1e-2
    taps = 20000
    N, P, ld = tune.optimal_partition(64, taps, model)
    print(f"tuned {N} x {P}: load {ld:.4f}", flush=True)
    R.check_partition(N, P, taps)
    g = R.greedy_partition([64 << i for i in range(7)], taps)
    assert ld <= tune.load(g.N, g.P, model) * (1 + 1e-9)
    h = decaying_noise_ir(taps, 0.5, 48000, 600)
    m = PartConvMulti(0)
    assert m.setup_mem(48000, 64, 1, 1, [dict(N=N, P=P, ir=h)]) == 0, last_error()
    x = np.random.default_rng(6).standard_normal(64 * 400).astype(np.float32)
    y = m.process(x)[:, 0]
    r = O.direct_conv_f64(x, h)
    assert np.max(np.abs(y - r)) / np.sqrt(np.mean(r ** 2)) < 1e-5
    m.cleanup()
