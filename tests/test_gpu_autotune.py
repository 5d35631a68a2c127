"""GPU half of the launch-shape autotuner: a measurement on a cache miss, the
same choice replayed from the cache on a hit (no measurement), and outputs
that do not depend on the choice (bitwise), for the real-time engine and the
differentiable form."""
import numpy as np
import pytest

from partconv_amd import PartConvMulti, autotune, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu


def engine(irs):
    m = PartConvMulti(0)
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=c, out_ch=c, ir=h) for c, h in enumerate(irs)]
    assert m.setup_mem(48000, 64, len(irs), len(irs), insts) == 0, last_error()
    return m


def run(m, xd, nfr):
    import torch
    y = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert m.run_device(y.data_ptr(), xd.data_ptr(), nfr, True, persistent=True) == 0, last_error()
    torch.cuda.synchronize()
    return y


def test_engine_tuning_cached_and_output_unchanged(tmp_path):
    import torch
    cache = tmp_path / "tuning.json"
    irs = [decaying_noise_ir(40000, 0.8, 48000, 70 + c) for c in range(4)]
    nfr = 700
    xd = torch.randn((nfr * 64, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(70))
    ref = engine(irs)
    y_ref = run(ref, xd, nfr)
    ref.cleanup()
    m = engine(irs)
    info = autotune.tune_engine(m, nticks=600, cache=cache)
    print("tuned:", info["workers"], info["bins_per_thread"], info["measured"], flush=True)
    assert not info["from_cache"] and info["workers"] in (0, 1, 2) and info["bins_per_thread"] in (2, 4, 8)
    assert info["release_window"] in autotune.RELEASE_WINDOWS
    y = run(m, xd, nfr)
    assert torch.equal(y, y_ref)
    m2 = engine(irs)
    info2 = autotune.tune_engine(m2, cache=cache)
    assert info2["from_cache"] and (info2["workers"], info2["bins_per_thread"], info2["release_window"]) == (
        info["workers"], info["bins_per_thread"], info["release_window"])
    import ctypes
    from partconv_amd._lib import lib
    w, b = ctypes.c_int(), ctypes.c_int()
    assert lib().pcm_get_tuning(m2._h, ctypes.byref(w), ctypes.byref(b)) == 0
    assert (w.value, b.value) == (info["workers"], info["bins_per_thread"])
    # every width of the delay-line sum gives the same output
    for width in (8, 4):
        assert lib().pcm_set_partial_width(m2._h, width) == 0
        m2.reset()
        assert torch.equal(run(m2, xd, nfr), y_ref), width


def test_offline_tuning_cached(tmp_path):
    import torch
    from partconv_amd.autograd import PartitionedConv
    cache = tmp_path / "tuning.json"
    N, P, taps, T, C = [128, 512, 4096], [14, 14, 0], 20000, 15000, 3
    g = torch.Generator(device="cuda").manual_seed(71)
    x = torch.randn((C, T), device="cuda", generator=g)
    h = torch.from_numpy(decaying_noise_ir(taps, 0.8, 48000, 71)).cuda()
    gy = torch.randn((C, T), device="cuda", generator=g)
    a = PartitionedConv(N, P, taps)
    ia = autotune.tune_offline(a, x, h, cache=cache)
    assert not ia["from_cache"] and len(ia["levels"]) == 3
    b = PartitionedConv(N, P, taps)
    ib = autotune.tune_offline(b, x, h, cache=cache)
    assert ib["from_cache"] and ib["levels"] == ia["levels"]
    outs = []
    for conv in (a, b):
        xt = x.clone().requires_grad_(True)
        ht = h.clone().requires_grad_(True)
        y = conv(xt, ht)
        y.backward(gy)
        torch.cuda.synchronize()
        outs.append((y.detach(), xt.grad, ht.grad))
    # the same choice: forward and both gradients bitwise equal
    for u, v in zip(*outs):
        assert torch.equal(u, v)


def test_mimo_window_tuning_cached_and_output_unchanged(tmp_path):
    """The real-time MIMO engine's release window: measured on a cache miss,
    replayed from the cache, output bitwise that of one tick at a time."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    cache = tmp_path / "tuning.json"
    g = torch.Generator(device="cuda").manual_seed(71)
    n_in, n_out, taps, nfr = 4, 3, 6000, 120
    h = torch.randn((n_out, n_in, taps), device="cuda", generator=g) * 0.05
    x = torch.randn((nfr * 128, n_in), device="cuda", generator=g)

    def mk():
        rt = MimoRealtime(48000, 128, [128, 1024], [14, 0], taps, n_in, n_out)
        rt.set_filters(h)
        return rt
    ref = mk()
    yr = torch.zeros((nfr * 128, n_out), device="cuda")
    ref.run_device(yr, x, nfr)
    rt = mk()
    info = autotune.tune_mimo(rt, nticks=96, cache=cache)
    print("mimo window:", info["release_window"], info["measured"], flush=True)
    assert not info["from_cache"] and info["release_window"] in autotune.MIMO_WINDOWS
    y = torch.zeros_like(yr)
    rt.run_device(y, x, nfr)
    torch.cuda.synchronize()
    assert torch.equal(y, yr)
    info2 = autotune.tune_mimo(mk(), cache=cache)
    assert info2["from_cache"] and info2["release_window"] == info["release_window"]
