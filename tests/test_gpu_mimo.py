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


@pytest.mark.parametrize("n_in,n_out,taps,N,P,nfr", [
    (5, 3, 10000, [128, 1024], [14, 0], 3 * 8 + 77),
    (32, 32, 20000, [256, 4096], [30, 0], 3 * 16 + 7),
    (3, 2, 3000, [128], [0], 50),                              # level 0 only
    (4, 4, 40000, [64, 512, 4096], [14, 14, 0], 5 * 64 + 13), # three levels, buffer 64
    (1, 32, 5000, [512, 2048], [6, 0], 3 * 4 + 5),            # one input, buffer 512
    (32, 1, 5000, [1024], [0], 9),                             # one output, buffer 1024
])
def test_mimo_realtime_matches_fp64(n_in, n_out, taps, N, P, nfr):
    """Real-time MIMO engine (shared input transforms and delay lines, per-bin
    matrix-vector products): every output against the fp64 sum of per-pair
    convolutions (zero added latency); the ticks split over two calls; a reset
    replays the same output bitwise."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    print(f"case mimo real time {n_in}x{n_out} taps={taps} N={N}", flush=True)
    N0 = N[0]
    irs = cfg4_irs(seed=7, n_in=n_in, n_out=n_out, taps=taps)
    h = np.stack([np.stack(irs[o]) for o in range(n_out)]).astype(np.float32)
    rng = np.random.default_rng(n_in * 10 + n_out)
    x = rng.standard_normal((nfr * N0, n_in)).astype(np.float32)
    rt = MimoRealtime(48000, N0, N, P, taps, n_in, n_out)
    rt.set_filters(torch.from_numpy(h).cuda())
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((nfr * N0, n_out), dtype=torch.float32, device="cuda")
    a = nfr // 3
    rt.run_device(yd[:a * N0], xd[:a * N0], a)
    rt.run_device(yd[a * N0:], xd[a * N0:], nfr - a)
    assert rt.stats().ticks == nfr
    y = yd.cpu().numpy()
    for o in range(n_out):
        r = sum(O.direct_conv_f64(x[:, i], h[o, i]) for i in range(n_in))[:nfr * N0]
        err = float(np.max(np.abs(y[:, o] - r)) / np.sqrt(np.mean(r ** 2)))
        assert err < 1e-5, (o, err)
    rt.reset()
    y2 = torch.zeros_like(yd)
    rt.run_device(y2, xd, nfr)
    assert np.array_equal(y2.cpu().numpy(), y)


def test_mimo_realtime_probe():
    """pcxr_probe times the delay-line sum of a level >= 1 on its stream and
    reports that launch's algorithmic bytes; probing leaves the output
    unchanged (bitwise, after a reset)."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    n_in, n_out, taps, N, P, nfr = 5, 3, 10000, [128, 1024], [14, 0], 101
    irs = cfg4_irs(seed=3, n_in=n_in, n_out=n_out, taps=taps)
    h = torch.from_numpy(np.stack([np.stack(irs[o]) for o in range(n_out)]).astype(np.float32)).cuda()
    x = torch.randn((nfr * 128, n_in), device="cuda")
    rt = MimoRealtime(48000, 128, N, P, taps, n_in, n_out)
    rt.set_filters(h)
    y0 = torch.zeros((nfr * 128, n_out), device="cuda")
    rt.run_device(y0, x, nfr)
    rt.reset()
    rt.probe(1)
    y1 = torch.zeros_like(y0)
    rt.run_device(y1, x, nfr)
    avg_us, n, nbytes = rt.probe_read()
    rt.probe(-1)
    Nl, Pl = rt.partition()
    assert n >= nfr // 8 - 1 and avg_us > 0
    assert nbytes == 8.0 * Nl[1] * (Pl[1] - 1) * (n_out * n_in + n_in) + 8.0 * Nl[1] * n_out
    assert torch.equal(y0, y1)
    with pytest.raises(RuntimeError):
        rt.probe(0)


def test_mimo_realtime_host_buffer_ticks():
    """pcxr_run: one tick per call on host buffers (PartConvMulti::run's
    interleaved layout) gives the same output as the device-resident run,
    bitwise; a wrong frameCount is an error."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    n_in, n_out, taps, N, P, nfr, n0 = 5, 3, 10000, [128, 1024], [14, 0], 41, 128
    irs = cfg4_irs(seed=5, n_in=n_in, n_out=n_out, taps=taps)
    h = torch.from_numpy(np.stack([np.stack(irs[o]) for o in range(n_out)]).astype(np.float32)).cuda()
    x = np.random.default_rng(4).standard_normal((nfr * n0, n_in)).astype(np.float32)
    rt = MimoRealtime(48000, n0, N, P, taps, n_in, n_out)
    rt.set_filters(h)
    y_host = np.zeros((nfr * n0, n_out), dtype=np.float32)
    out = np.zeros((n0, n_out), dtype=np.float32)
    for f in range(nfr):
        rt.run(out, np.ascontiguousarray(x[f * n0:(f + 1) * n0]))
        y_host[f * n0:(f + 1) * n0] = out
    assert rt.stats().ticks == nfr
    rt.reset()
    yd = torch.zeros((nfr * n0, n_out), device="cuda")
    rt.run_device(yd, torch.from_numpy(x).cuda(), nfr)
    assert np.array_equal(yd.cpu().numpy(), y_host)
    with pytest.raises(ValueError):
        rt.run(np.zeros((64, n_out), np.float32), np.zeros((64, n_in), np.float32))
