"""Release windows of the persistent device-resident run (pcm_set_release_window):
up to W ticks released ahead of the oldest incomplete one.  The deadline rule
(a tick runs only once the level blocks it reads first are complete,
src/partconvMulti.cpp:886-903) and every output sample are those of one tick
at a time: the outputs must be bitwise those of per-tick launches, for split
runs, strict and relaxed, with the short levels in the kernel's workers, in
self mode or launched from the host (period-2 levels narrow the window), and
a late level block must still be waited for and counted."""
import sys

import numpy as np
import pytest

from partconv_amd import PartConvMulti, PartConvMultiRelaxed, greedy_partition, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(insts, irs, nin, nout, relaxed=False):
    m = (PartConvMultiRelaxed if relaxed else PartConvMulti)(0)
    assert m.setup_mem(48000, 64, nin, nout, [dict(ir=h, **i) for i, h in zip(insts, irs)]) == 0, last_error()
    return m


def run_cuts(m, x, nout, cuts, persistent):
    import torch
    y = torch.zeros((x.shape[0], nout), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rin, rout = 64 * x.shape[1] * 4, 64 * nout * 4
    for a, b in zip(cuts[:-1], cuts[1:]):
        assert m.run_device(y.data_ptr() + a * rout, x.data_ptr() + a * rin, b - a, True,
                            persistent=persistent) == 0, last_error()
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("relaxed", [False, True])
def test_release_window_bitwise_equal_to_tick_launches(relaxed):
    import torch
    rng = np.random.default_rng(31)
    bs = [64 << i for i in range(9)]
    taps = [30000, 9000, 70000]
    insts, irs = [], []
    for i, T in enumerate(taps):
        N, P = greedy_partition(bs, T)
        insts.append(dict(N=N, P=P, in_ch=i, out_ch=i))
        irs.append(decaying_noise_ir(T, 1.0, 48000, 140 + i))
    nfr = 611
    x = torch.from_numpy(rng.standard_normal((nfr * 64, 3)).astype(np.float32)).cuda()
    cuts = [0, 1, 302, 303, 500, nfr]
    say(f"case window relaxed={relaxed}: launch")
    m = make(insts, irs, 3, 3, relaxed)
    ref = run_cuts(m, x, 3, cuts, False)
    m.cleanup()
    for workers in (1, 2, 0):
        for W in (2, 3, 8):
            say(f"case window relaxed={relaxed} workers={workers} W={W}")
            m = make(insts, irs, 3, 3, relaxed)
            m.set_persistent_workers(workers)
            assert m.set_release_window(W) == 0
            y = run_cuts(m, x, 3, cuts, True)
            st = m.stats()
            say(f"  persistent {st.persistent_ticks}, stalls {st.stalls}, p50 {st.p50_us:.1f} us")
            assert st.persistent_ticks == nfr and st.stalls == 0
            assert np.array_equal(y, ref), (workers, W)
            m.cleanup()


def test_release_window_sixteen_level_shapes_and_summed_outputs():
    """Four channels of the 16-channel line's block sizes [64, 256, 2048, 16384]
    over two periods of the 2048 level, and a summed output (two instances into
    one channel: the window does not apply and the run is one tick at a time)."""
    import torch
    rng = np.random.default_rng(32)
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=i, out_ch=i) for i in range(4)]
    irs = [decaying_noise_ir(60000, 1.0, 48000, 150 + i) for i in range(4)]
    nfr = 700
    x = torch.from_numpy(rng.standard_normal((nfr * 64, 4)).astype(np.float32)).cuda()
    say("case window 16384: launch")
    m = make(insts, irs, 4, 4)
    ref = run_cuts(m, x, 4, [0, nfr], False)
    m.cleanup()
    for W in (2, 4, 16):
        say(f"case window 16384: W={W}")
        m = make(insts, irs, 4, 4)
        assert m.set_release_window(W) == 0
        y = run_cuts(m, x, 4, [0, 255, nfr], True)
        assert m.stats().persistent_ticks == nfr
        assert np.array_equal(y, ref), W
        m.cleanup()
    say("case window summed output")
    summed = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=i, out_ch=0) for i in range(2)]
    irs2 = [decaying_noise_ir(12000, 0.5, 48000, 160 + i) for i in range(2)]
    x2 = x[:, :2].contiguous()
    m = make(summed, irs2, 2, 1)
    ref2 = run_cuts(m, x2, 1, [0, nfr], False)
    m.cleanup()
    m = make(summed, irs2, 2, 1)
    assert m.set_release_window(4) == 0
    y2 = run_cuts(m, x2, 1, [0, 300, nfr], True)
    assert np.array_equal(y2, ref2)
    m.cleanup()


def test_release_window_late_level():
    """A level block held 0.3 s past its deadline (pcm_stall_level: a host
    function on the level's stream) while ticks are released ahead: the
    released ticks are served, the kernel stops, the late tick waits for the
    block (a counted miss) and the output is bitwise that of an unstalled run."""
    import torch
    hold = 0.3
    insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=i, out_ch=i) for i in range(4)]
    irs = [decaying_noise_ir(20000, 0.4, 48000, 170 + i) for i in range(4)]
    nfr = 300
    rng = np.random.default_rng(98)
    xd = torch.from_numpy(rng.standard_normal((nfr * 64, 4)).astype(np.float32)).cuda()
    say("case window late level: reference")
    ref = make(insts, irs, 4, 4)
    yr = run_cuts(ref, xd, 4, [0, nfr], True)
    ref.cleanup()
    say("case window late level: held")
    m = make(insts, irs, 4, 4)
    assert m.set_release_window(3) == 0
    y = torch.zeros_like(xd)
    torch.cuda.synchronize()
    row = 64 * 4 * 4
    assert m.run_device(y.data_ptr(), xd.data_ptr(), 96, True, persistent=True) == 0, last_error()
    assert m.stall_level(2, hold) == 0
    m.clear_stats()
    assert m.run_device(y.data_ptr() + 96 * row, xd.data_ptr() + 96 * row, 48, True, persistent=True) == 0, \
        last_error()
    st = m.stats()
    say(f"  ticks {st.ticks}, misses {st.deadline_misses}, stalls {st.stalls}, max {st.max_us:.0f} us")
    assert st.deadline_misses >= 1 and st.stalls >= 1
    assert st.max_us > 0.9 * hold * 1e6
    assert m.run_device(y.data_ptr() + 144 * row, xd.data_ptr() + 144 * row, nfr - 144, True,
                        persistent=True) == 0, last_error()
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), yr)
    m.cleanup()


def test_release_window_argument_bounds():
    m = PartConvMulti(0)
    assert m.set_release_window(0) != 0
    assert m.set_release_window(17) != 0
    assert m.set_release_window(16) == 0


@pytest.mark.parametrize("n_in,n_out,taps,N,P,nfr", [
    (5, 3, 10000, [128, 1024], [14, 0], 3 * 8 + 77),
    (32, 32, 20000, [256, 4096], [30, 0], 3 * 16 + 7),
    (4, 4, 40000, [64, 512, 4096], [14, 14, 0], 5 * 64 + 13),
    (3, 2, 3000, [128, 256], [2, 0], 61),                      # P_0 = 2 caps the window at 2
])
def test_mimo_release_window_bitwise_equal(n_in, n_out, taps, N, P, nfr):
    """Real-time MIMO engine with ticks released ahead (pcxr_set_release_window):
    outputs bitwise those of one tick at a time, over split calls."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    from partconv_amd.workloads import cfg4_irs
    N0 = N[0]
    irs = cfg4_irs(seed=9, n_in=n_in, n_out=n_out, taps=taps)
    h = torch.from_numpy(np.stack([np.stack(irs[o]) for o in range(n_out)]).astype(np.float32)).cuda()
    x = torch.from_numpy(np.random.default_rng(n_in + 7 * n_out).standard_normal((nfr * N0, n_in))
                         .astype(np.float32)).cuda()
    outs = []
    for W in (1, 2, 3):
        say(f"case mimo window {n_in}x{n_out} N={N} W={W}")
        rt = MimoRealtime(48000, N0, N, P, taps, n_in, n_out)
        rt.set_filters(h)
        rt.set_release_window(W)
        y = torch.zeros((nfr * N0, n_out), dtype=torch.float32, device="cuda")
        a = nfr // 3
        rt.run_device(y[:a * N0], x[:a * N0], a)
        rt.run_device(y[a * N0:], x[a * N0:], nfr - a)
        torch.cuda.synchronize()
        st = rt.stats()
        assert st.ticks == nfr and st.stalls == 0
        outs.append(y.cpu().numpy())
    assert np.array_equal(outs[1], outs[0])
    assert np.array_equal(outs[2], outs[0])


def test_mimo_release_window_late_level():
    """A held level block inside a MIMO release window: waited for, counted,
    output bitwise that of an unheld run."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    n_in, n_out, taps, nfr, hold = 2, 3, 3000, 200, 0.3
    g = torch.Generator(device="cuda").manual_seed(48)
    h = torch.randn((n_out, n_in, taps), device="cuda", generator=g) * 0.05
    x = torch.randn((nfr * 64, n_in), device="cuda", generator=g)

    def mk(W):
        rt = MimoRealtime(48000, 64, [64, 512], [14, 0], taps, n_in, n_out)
        rt.set_filters(h)
        rt.set_release_window(W)
        return rt
    ref = mk(1)
    yr = torch.zeros((nfr * 64, n_out), device="cuda")
    ref.run_device(yr, x, nfr)
    say("case mimo window late level")
    rt = mk(3)
    y = torch.zeros_like(yr)
    rt.run_device(y[:40 * 64], x[:40 * 64], 40)
    rt.stall_level(1, hold)
    rt.clear_stats()
    rt.run_device(y[40 * 64:80 * 64], x[40 * 64:80 * 64], 40)
    st = rt.stats()
    say(f"  ticks {st.ticks}, misses {st.deadline_misses}, stalls {st.stalls}, max {st.max_us:.0f} us")
    assert st.deadline_misses >= 1 and st.stalls >= 1 and st.max_us > 0.9 * hold * 1e6
    rt.run_device(y[80 * 64:], x[80 * 64:], nfr - 80)
    torch.cuda.synchronize()
    assert torch.equal(y, yr)


def test_mimo_release_window_bounds():
    from partconv_amd.mimo import MimoRealtime
    rt = MimoRealtime(48000, 64, [64, 512], [14, 0], 3000, 2, 2)
    with pytest.raises(RuntimeError):
        rt.set_release_window(0)
    with pytest.raises(RuntimeError):
        rt.set_release_window(4)
    rt.set_release_window(3)
