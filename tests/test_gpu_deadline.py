"""Deadline handling (the real-time harness's underflow counter,
src/pcRTMulti.cpp:21,138-185; the strict run waits for late levels,
src/partconvMulti.cpp:886-903).  A level block is held past its deadline by a
host function sleeping 0.3 s on the level's stream (pcm_stall_level /
pcxr_stall_level: the stall is on the host, every kernel exits normally).
The engine must notice the late block, leave the persistent kernel, wait for
the block and go on; the late tick counts as a deadline miss (p99 above the
period in a short window), the interruption is counted, and the output is
bitwise that of an unstalled run."""
import sys

import numpy as np
import pytest

from partconv_amd import PartConvMulti, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu

HOLD_S = 0.3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine(insts, irs, n):
    m = PartConvMulti(0)
    assert m.setup_mem(48000, 64, n, n, [dict(ir=h, **i) for i, h in zip(insts, irs)]) == 0, last_error()
    return m


def check_window(st, period_us):
    say(f"window: ticks {st.ticks}, misses {st.deadline_misses}, stalls {st.stalls}, "
        f"p99 {st.p99_us:.0f} us, max {st.max_us:.0f} us")
    assert st.deadline_misses >= 1
    assert st.stalls >= 1
    assert st.max_us > 0.9 * HOLD_S * 1e6
    assert st.p99_us > period_us     # one late tick in <= 50 is the 99th percentile


def test_instance_engine_late_level_device_resident():
    """Four channels, [64x6, 256x14, 2048xP]: the 2048 level's block started at
    tick 96 (deadline tick 127) is held; 48-tick window 96..143."""
    import torch
    insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=i, out_ch=i) for i in range(4)]
    irs = [decaying_noise_ir(20000, 0.4, 48000, 80 + i) for i in range(4)]
    nfr = 300
    rng = np.random.default_rng(96)
    xd = torch.from_numpy(rng.standard_normal((nfr * 64, 4)).astype(np.float32)).cuda()
    ref = engine(insts, irs, 4)
    yr = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert ref.run_device(yr.data_ptr(), xd.data_ptr(), nfr, True, persistent=True) == 0, last_error()
    assert ref.stats().stalls == 0
    ref.cleanup()
    m = engine(insts, irs, 4)
    y = torch.zeros_like(xd)
    torch.cuda.synchronize()
    row = 64 * 4 * 4  # bytes per frame
    assert m.run_device(y.data_ptr(), xd.data_ptr(), 96, True, persistent=True) == 0, last_error()
    assert m.stall_level(2, HOLD_S) == 0
    m.clear_stats()
    assert m.run_device(y.data_ptr() + 96 * row, xd.data_ptr() + 96 * row, 48, True, persistent=True) == 0, \
        last_error()
    check_window(m.stats(), 1e6 * 64 / 48000)
    assert m.run_device(y.data_ptr() + 144 * row, xd.data_ptr() + 144 * row, nfr - 144, True,
                        persistent=True) == 0, last_error()
    assert torch.equal(y, yr)
    m.cleanup()


def test_instance_engine_late_level_host_session():
    """The same hold through run() on host buffers (the persistent session):
    the session ends at the late block, the tick is served by a launch that
    waits for it, and the next call starts a new session."""
    insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=i, out_ch=i) for i in range(2)]
    irs = [decaying_noise_ir(12000, 0.3, 48000, 90 + i) for i in range(2)]
    rng = np.random.default_rng(97)
    x = rng.standard_normal((300 * 64, 2)).astype(np.float32)
    yr = engine(insts, irs, 2).process(x)
    m = engine(insts, irs, 2)
    y1 = m.process(x[:96 * 64])
    assert m.stall_level(2, HOLD_S) == 0
    m.clear_stats()
    y2 = m.process(x[96 * 64:144 * 64])
    check_window(m.stats(), 1e6 * 64 / 48000)
    y3 = m.process(x[144 * 64:])
    assert m.stats().persistent_ticks > 0          # sessions resumed after the stall
    assert np.array_equal(np.concatenate([y1, y2, y3]), yr)
    m.cleanup()


def test_mimo_realtime_late_level():
    """Real-time MIMO engine, 2 -> 3 channels, [64x14, 512xP]: the 512 level's
    block started at tick 40 (deadline 47) is held; the engine stops its
    persistent kernel, waits for the block and relaunches for the rest."""
    import torch
    from partconv_amd.mimo import MimoRealtime
    n_in, n_out, taps, nfr = 2, 3, 3000, 200
    g = torch.Generator(device="cuda").manual_seed(47)
    h = torch.randn((n_out, n_in, taps), device="cuda", generator=g) * 0.05
    x = torch.randn((nfr * 64, n_in), device="cuda", generator=g)

    def mk():
        rt = MimoRealtime(48000, 64, [64, 512], [14, 0], taps, n_in, n_out)
        rt.set_filters(h)
        return rt
    ref = mk()
    yr = torch.zeros((nfr * 64, n_out), device="cuda")
    ref.run_device(yr, x, nfr)
    assert ref.stats().stalls == 0
    rt = mk()
    y = torch.zeros_like(yr)
    rt.run_device(y[:40 * 64], x[:40 * 64], 40)
    rt.stall_level(1, HOLD_S)
    rt.clear_stats()
    rt.run_device(y[40 * 64:80 * 64], x[40 * 64:80 * 64], 40)
    check_window(rt.stats(), 1e6 * 64 / 48000)
    rt.run_device(y[80 * 64:], x[80 * 64:], nfr - 80)
    torch.cuda.synchronize()
    assert torch.equal(y, yr)


def test_stall_hook_rejects_bad_arguments():
    insts = [dict(N=[64, 256], P=[6, 0], in_ch=0, out_ch=0)]
    m = engine(insts, [decaying_noise_ir(2000, 0.1, 48000, 1)], 1)
    assert m.stall_level(0, 0.1) != 0      # level 0 runs in the tick itself
    assert m.stall_level(2, 0.1) != 0      # no such level
    assert m.stall_level(1, 10.0) != 0     # bounded
    m.cleanup()
