"""GPU parity at the benchmark configurations' stated sizes, through the C ABI.

Every case runs long enough that the longest level's last partition
contributes to the checked output, and compares with the float64 direct
convolution (max |error| <= 1e-5 of output RMS):
  * 16 channels x 262144-tap IRs, partition [64x6, 256x14, 2048x14, 16384x15],
    every channel; the plain (per-instance) engine bitwise equal to the batched;
  * 512 uneven instances, IRs log-uniform in [2^12, 2^20] taps, greedy
    partitions over 128..32768 (up to 32768 x 31), a seeded sample of 32
    channels including the shortest and the longest IR;
  * MIMO 32 x 32 x 65536 taps, [256x30, 4096x15]: the real-time MIMO engine,
    the MFMA throughput form and 1024 pair instances, 4 seeded outputs plus
    every output of the throughput form against the real-time engine;
  * the reference's own pc.config shape with a 65536 level
    ([64x6, 256x6, 1024x6, 4096x6, 16384x6, 65536xP]) and buffer_size 2048.
The differentiable workload at size is in test_gpu_autograd.py."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from partconv_amd import PartConvMulti, greedy_partition, last_error
from partconv_amd.workloads import CFG2, CFG3, CFG4, cfg2_irs, cfg3_irs, cfg3_lengths, cfg4_irs, decaying_noise_ir

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(y, r):
    return float(np.max(np.abs(np.asarray(y, np.float64) - r)) / np.sqrt(np.mean(np.asarray(r) ** 2)))


def setup_multi(insts, irs, N0, n_in, n_out, mode="batched", FS=48000):
    m = PartConvMulti(0, mode=mode)
    d = [dict(ir=h, **i) for i, h in zip(insts, irs)]
    assert m.setup_mem(FS, N0, n_in, n_out, d) == 0, last_error()
    return m


def run_persistent(m, xd, nfr, n_out):
    import torch
    yd = torch.zeros((nfr * m.buffer_size, n_out), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert m.run_device(yd.data_ptr(), xd.data_ptr(), nfr, True, persistent=True) == 0, last_error()
    torch.cuda.synchronize()
    return yd


def test_sixteen_channels_262144_taps_full_size():
    """16 x 262144 taps at buffer 64: P = [6, 14, 14, 15]; 4800 ticks (307200
    samples, past the last partition's start at 262144 - 16384 and over 18
    periods of the 16384 level) through the persistent tick kernel; every
    channel against fp64; the plain engine bitwise equal."""
    import torch
    c = CFG2
    n, nfr = c["channels"], 4800
    irs = cfg2_irs(seed=2)
    insts = [dict(N=c["N"], P=c["P"], in_ch=i, out_ch=i) for i in range(n)]
    rng = np.random.default_rng(202)
    x = rng.standard_normal((nfr * 64, n)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    m = setup_multi(insts, irs, 64, n, n)
    assert m.instance_partition(0) == ([64, 256, 2048, 16384], [6, 14, 14, 15])
    y = run_persistent(m, xd, nfr, n).cpu().numpy()
    st = m.stats()
    say(f"cfg2 full: persistent ticks {st.persistent_ticks}/{nfr}, p50 {st.p50_us:.1f} us")
    worst = 0.0
    for i in range(n):
        worst = max(worst, rel(y[:, i], O.direct_conv_f64(x[:, i], irs[i])))
    say(f"cfg2 full: worst channel error {worst:.2e} of RMS")
    assert worst < TOL
    m.cleanup()
    p = setup_multi(insts, irs, 64, n, n, mode="plain")
    yp = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert p.run_device(yp.data_ptr(), xd.data_ptr(), nfr, True) == 0, last_error()
    assert np.array_equal(yp.cpu().numpy(), y)
    p.cleanup()


@pytest.mark.timeout(900)
def test_512_uneven_instances_to_2_20_taps_full_size():
    """512 instances, IR lengths log-uniform in [2^12, 2^20] (seeded), each with
    its own greedy partition over 128..32768, buffer 128, one persistent tick
    kernel; 8400 ticks (1075200 samples, past the start of the longest IR's
    last 32768 partition).  32 seeded channels, the shortest and the longest
    IR among them, against fp64."""
    import torch
    c = CFG3
    n = c["instances"]
    lens = cfg3_lengths(seed=3)
    assert lens.max() > (1 << 19)
    irs = cfg3_irs(lens, seed=3)
    insts = []
    for i, T in enumerate(lens):
        N, P = greedy_partition(c["block_sizes"], int(T))
        insts.append(dict(N=N, P=P, in_ch=i, out_ch=i))
    longest = int(np.argmax(lens))
    Nl, Pl = insts[longest]["N"], insts[longest]["P"]
    say(f"cfg3 full: longest IR {lens[longest]} taps, partition {list(zip(Nl, Pl))}")
    assert Nl[-1] == 32768
    start_last = sum(a * b for a, b in zip(Nl[:-1], Pl[:-1])) + (Pl[-1] - 1) * Nl[-1]
    nfr = 8400
    assert nfr * 128 > start_last + 32768
    g = torch.Generator(device="cuda").manual_seed(303)
    xd = torch.randn((nfr * 128, n), device="cuda", generator=g)
    m = setup_multi(insts, irs, 128, n, n)
    yd = run_persistent(m, xd, nfr, n)
    st = m.stats()
    say(f"cfg3 full: persistent ticks {st.persistent_ticks}/{nfr}, p50 {st.p50_us:.1f} us, "
        f"p99 {st.p99_us:.1f} us")
    rng = np.random.default_rng(33)
    sample = set(rng.choice(n, size=30, replace=False).tolist()) | {int(np.argmin(lens)), longest}
    worst = 0.0
    for i in sorted(sample):
        xi = xd[:, i].cpu().numpy()
        worst = max(worst, rel(yd[:, i].cpu().numpy(), O.direct_conv_f64(xi, irs[i])))
    say(f"cfg3 full: {len(sample)} channels, worst error {worst:.2e} of RMS")
    assert len(sample) >= 32
    assert worst < TOL
    m.cleanup()


def f64_mimo_outputs(x, irs, outs):
    """fp64 y[o] = sum_i conv(x[:, i], irs[o][i]) for the outputs `outs`."""
    T, n_in = x.shape
    taps = irs[0][0].size
    nfft = 1 << (T + taps - 1).bit_length()
    X = np.fft.rfft(x.astype(np.float64), nfft, axis=0)
    res = {}
    for o in outs:
        acc = np.zeros(X.shape[0], dtype=np.complex128)
        for i in range(n_in):
            acc += X[:, i] * np.fft.rfft(irs[o][i].astype(np.float64), nfft)
        res[o] = np.fft.irfft(acc, nfft)[:T]
    return res


@pytest.fixture(scope="module")
def mimo_full():
    """32 x 32 x 65536-tap network (RT60 1 s), 320 frames of 256 (81920
    samples: past the last 4096 partition's start at 7680 + 14 x 4096)."""
    c = CFG4
    irs = cfg4_irs(seed=4)
    rng = np.random.default_rng(404)
    nfr = 320
    x = rng.standard_normal((nfr * 256, c["n_in"])).astype(np.float32)
    outs = sorted(rng.choice(c["n_out"], size=4, replace=False).tolist())
    ref = f64_mimo_outputs(x, irs, outs)
    return dict(irs=irs, x=x, nfr=nfr, outs=outs, ref=ref)


def test_mimo_32x32_65536_taps_realtime_and_throughput_forms(mimo_full):
    import torch
    from partconv_amd.mimo import MimoConv, MimoRealtime
    c = CFG4
    irs, x, nfr, outs, ref = (mimo_full[k] for k in ("irs", "x", "nfr", "outs", "ref"))
    h = torch.from_numpy(np.stack([np.stack(r) for r in irs])).cuda()   # [n_out][n_in][taps]
    rt = MimoRealtime(48000, 256, c["N"], c["P"], c["taps"], c["n_in"], c["n_out"])
    assert rt.partition() == ([256, 4096], [30, 15])
    rt.set_filters(h)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((nfr * 256, c["n_out"]), dtype=torch.float32, device="cuda")
    rt.run_device(yd, xd, nfr)
    y = yd.cpu().numpy()
    st = rt.stats()
    say(f"cfg4 real-time: p50 {st.p50_us:.1f} us, misses {st.deadline_misses}")
    errs = [rel(y[:, o], ref[o]) for o in outs]
    say(f"cfg4 real-time: outputs {outs}, errors {['%.2e' % e for e in errs]}")
    assert max(errs) < TOL
    mx = MimoConv(c["N"], c["P"], c["taps"], c["n_in"], c["n_out"])
    mx.set_filters(h)
    ym = mx(xd.t().contiguous()).cpu().numpy()           # [n_out][T]
    errs = [rel(ym[o], ref[o]) for o in outs]
    say(f"cfg4 throughput form: errors {['%.2e' % e for e in errs]}")
    assert max(errs) < TOL
    # every output: the two forms agree to the same bar
    for o in range(c["n_out"]):
        assert rel(ym[o], y[:, o].astype(np.float64)) < 2 * TOL, o


def test_mimo_32x32_65536_taps_as_pair_instances(mimo_full):
    """The reference's own formulation: 1024 instances (inCh = i, outCh = o)
    on the batched engine, summed per output, one persistent tick kernel."""
    import torch
    c = CFG4
    irs, x, nfr, outs, ref = (mimo_full[k] for k in ("irs", "x", "nfr", "outs", "ref"))
    insts, hs = [], []
    for o in range(c["n_out"]):
        for i in range(c["n_in"]):
            insts.append(dict(N=c["N"], P=c["P"], in_ch=i, out_ch=o))
            hs.append(irs[o][i])
    m = setup_multi(insts, hs, 256, c["n_in"], c["n_out"])
    xd = torch.from_numpy(x).cuda()
    y = run_persistent(m, xd, nfr, c["n_out"]).cpu().numpy()
    errs = [rel(y[:, o], ref[o]) for o in outs]
    say(f"cfg4 pair instances: errors {['%.2e' % e for e in errs]}, "
        f"persistent ticks {m.stats().persistent_ticks}")
    assert max(errs) < TOL
    m.cleanup()


def test_reference_config_shape_with_65536_level():
    """The level layout of the reference's pc.config (src/pc.config:2):
    64x6, 256x6, 1024x6, 4096x6, 16384x6 and a 65536 level (last count
    recounted: 600000 taps -> 8).  Two instances, 10400 ticks (665600 samples,
    past the end of the last 65536 partition); persistent kernel and per-tick
    launches bitwise equal; both channels against fp64."""
    import torch
    N = [64, 256, 1024, 4096, 16384, 65536]
    P = [6, 6, 6, 6, 6, 0]
    taps = 600000
    irs = [decaying_noise_ir(taps, 4.0, 48000, 61 + i) for i in range(2)]
    insts = [dict(N=N, P=P, in_ch=i, out_ch=i) for i in range(2)]
    nfr = 10400
    rng = np.random.default_rng(65)
    x = rng.standard_normal((nfr * 64, 2)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    m = setup_multi(insts, irs, 64, 2, 2)
    Ni, Pi = m.instance_partition(0)
    assert Ni == N and Pi == [6, 6, 6, 6, 6, 8]
    start_last = sum(a * b for a, b in zip(N[:-1], Pi[:-1])) + 7 * 65536
    assert nfr * 64 > start_last + 65536
    y = run_persistent(m, xd, nfr, 2).cpu().numpy()
    say(f"65536 level: persistent ticks {m.stats().persistent_ticks}/{nfr}")
    m.cleanup()
    q = setup_multi(insts, irs, 64, 2, 2)
    yl = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert q.run_device(yl.data_ptr(), xd.data_ptr(), nfr, True) == 0, last_error()
    assert np.array_equal(yl.cpu().numpy(), y)
    q.cleanup()
    for i in range(2):
        e = rel(y[:, i], O.direct_conv_f64(x[:, i], irs[i]))
        say(f"65536 level: channel {i} error {e:.2e}")
        assert e < TOL


@pytest.mark.parametrize("relaxed", [False, True])
def test_buffer_size_2048(relaxed):
    """buffer_size 2048 (the largest accepted): partition [2048x14, 16384x6,
    65536xP], two instances; host-buffer run() through the persistent session,
    device-resident persistent and per-tick runs, all bitwise equal, and fp64
    (relaxed: delayed by one buffer)."""
    import torch
    from partconv_amd import PartConvMultiRelaxed
    N, P, taps = [2048, 16384, 65536], [14, 6, 0], 300000
    irs = [decaying_noise_ir(taps, 2.0, 48000, 71 + i) for i in range(2)]
    insts = [dict(N=N, P=P, in_ch=i, out_ch=i) for i in range(2)]
    nfr = 190
    rng = np.random.default_rng(2048 + relaxed)
    x = rng.standard_normal((nfr * 2048, 2)).astype(np.float32)

    def mk():
        m = (PartConvMultiRelaxed if relaxed else PartConvMulti)(0)
        assert m.setup_mem(48000, 2048, 2, 2, [dict(ir=h, **i) for i, h in zip(insts, irs)]) == 0, last_error()
        return m
    m = mk()
    yh = m.process(x)
    say(f"buffer 2048 relaxed={relaxed}: host run, persistent ticks {m.stats().persistent_ticks}")
    m.cleanup()
    xd = torch.from_numpy(x).cuda()
    for persistent in (True, False):
        q = mk()
        yd = torch.zeros_like(xd)
        torch.cuda.synchronize()
        assert q.run_device(yd.data_ptr(), xd.data_ptr(), nfr, True, persistent=persistent) == 0, last_error()
        assert np.array_equal(yd.cpu().numpy(), yh), persistent
        q.cleanup()
    for i in range(2):
        r = O.direct_conv_f64(x[:, i], irs[i])
        if relaxed:
            r = np.concatenate([np.zeros(2048), r[:-2048]])
        assert rel(yh[:, i], r) < TOL
