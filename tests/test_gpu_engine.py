"""GPU parity tests: the HIP engine (through the C ABI) against the CPU oracle
and the float64 direct convolution on the same seeded inputs.

Bars: max |error| <= 1e-5 x output RMS against the fp32 oracle (the
reference's algorithm restated) and against the fp64 direct convolution.
"""
import sys
import wave

import numpy as np
import pytest

from oracle import oracle as O
from partconv_amd import PartConvFilter, PartConvMulti, PartConvMultiRelaxed, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(y, r):
    return float(np.max(np.abs(np.asarray(y, np.float64) - r)) / np.sqrt(np.mean(np.asarray(r) ** 2)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_multi(insts, irs, N0, n_in, n_out, relaxed=False, mode="batched", FS=48000):
    cls = PartConvMultiRelaxed if relaxed else PartConvMulti
    m = cls(0, mode=mode)
    d = [dict(ir=h, **i) for i, h in zip(insts, irs)]
    assert m.setup_mem(FS, N0, n_in, n_out, d) == 0, last_error()
    return m


PARTS = [
    ([64, 256], [6, 0]),
    ([64, 256, 2048], [6, 14, 0]),
    ([32, 128, 512, 2048], [6, 6, 6, 0]),
    ([64, 512, 4096], [14, 14, 0]),            # first big level (4096 = 64 x 64)
    ([128, 1024, 8192], [14, 14, 0]),          # 8192 = 128 x 64
    ([64, 256, 2048, 16384], [6, 14, 14, 0]),  # config-2 partition
]


@pytest.mark.parametrize("N,P", PARTS)
@pytest.mark.parametrize("relaxed", [False, True])
def test_single_instance_matches_oracle(N, P, relaxed):
    say(f"case N={N} P={P} relaxed={relaxed}")
    rng = np.random.default_rng(sum(N) + relaxed)
    taps = 2 * N[-1] + sum(n * p for n, p in zip(N[:-1], P[:-1])) + 123
    h = decaying_noise_ir(taps, taps / 48000, 48000, 5)
    nfr = 2 * (N[-1] // N[0]) + 40
    x = rng.standard_normal(nfr * N[0]).astype(np.float32)
    m = make_multi([dict(N=N, P=P)], [h], N[0], 1, 1, relaxed=relaxed)
    y = m.process(x)[:, 0]
    yo = O.OracleMulti(48000, N[0], 1, 1, [dict(N=N, P=P)], [h], relaxed=relaxed).process(x)[:, 0]
    r = O.direct_conv_f64(x, h)
    if relaxed:
        r = np.concatenate([np.zeros(N[0]), r[:-N[0]]])
    assert rel(y, yo) < TOL
    assert rel(y, r) < TOL


def test_impulse_reads_back_h():
    say("case impulse")
    N, P = [64, 256, 2048, 16384], [6, 14, 14, 0]
    h = decaying_noise_ir(40000, 0.5, 48000, 9)
    nfr = 800
    x = np.zeros(nfr * 64, dtype=np.float32)
    t0 = 64 * 37 + 11
    x[t0] = 1.0
    m = make_multi([dict(N=N, P=P)], [h], 64, 1, 1)
    y = m.process(x)[:, 0]
    ref = np.zeros_like(y, dtype=np.float64)
    ref[t0:t0 + h.size] = h[:y.size - t0]
    assert np.max(np.abs(y - ref)) < 1e-6 * np.max(np.abs(h))


def test_multi_instance_mix_mute_scale_uneven():
    say("case multi mix")
    rng = np.random.default_rng(3)
    insts = [
        dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=0, out_ch=1, scale=0.5),
        dict(N=[64, 128, 512, 2048], P=[2, 6, 6, 0], in_ch=1, out_ch=1, scale=2.0),
        dict(N=[64, 256], P=[6, 0], in_ch=2, out_ch=0),
        dict(N=[64, 256], P=[6, 0], in_ch=0, out_ch=0, mute=1),
        dict(N=[64, 1024, 8192], P=[30, 15, 0], in_ch=1, out_ch=2),
    ]
    lens = [9000, 12000, 1500, 800, 20000]
    irs = [decaying_noise_ir(n, n / 48000, 48000, 20 + i) for i, n in enumerate(lens)]
    nfr = 3 * 8192 // 64
    x = rng.standard_normal((nfr * 64, 3)).astype(np.float32)
    ref = O.OracleMulti(48000, 64, 3, 3, insts, irs).process(x)
    for mode in ("batched", "plain"):
        m = make_multi(insts, irs, 64, 3, 3, mode=mode)
        y = m.process(x)
        for c in range(3):
            assert rel(y[:, c], ref[:, c]) < TOL, (mode, c)
    d = O.direct_conv_f64
    r1 = d(0.5 * x[:, 0].astype(np.float64), irs[0]) + d(2.0 * x[:, 1].astype(np.float64), irs[1])
    assert rel(y[:, 1], r1) < TOL


def test_plain_and_batched_agree_bitwise_close():
    say("case plain vs batched")
    rng = np.random.default_rng(4)
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=i, out_ch=i) for i in range(4)]
    irs = [decaying_noise_ir(40000, 1.0, 48000, 40 + i) for i in range(4)]
    x = rng.standard_normal((600 * 64, 4)).astype(np.float32)
    yb = make_multi(insts, irs, 64, 4, 4, mode="batched").process(x)
    yp = make_multi(insts, irs, 64, 4, 4, mode="plain").process(x)
    # same kernels, same order of arithmetic per instance
    assert np.array_equal(yb, yp)


def test_filter_api_matches_multi():
    say("case PartConvFilter API")
    rng = np.random.default_rng(5)
    N, P = [64, 256, 2048, 8192], [6, 14, 6, 0]
    h = decaying_noise_ir(30000, 0.6, 48000, 50)
    x = rng.standard_normal(64 * 400).astype(np.float32)
    f = PartConvFilter(0, 48000)
    assert f.setup(4, N, P, h) == 0
    assert f.num_parts[-1] == -(-(30000 - (384 + 3584 + 12288)) // 8192)
    y = f.process(x)
    yo = O.OracleFilter(N, P, h, "f32").process(x)
    assert rel(y, yo) < TOL
    f.reset()
    y2 = f.process(x, relaxed=True)
    yr = O.OracleFilter(N, P, h, "f32").process(x, relaxed=True)
    assert rel(y2, yr) < TOL
    t = f.bench(1, 20, 0, 0.05)
    assert 0 < t < 0.05
    f.cleanup()


def test_filter_setup_errors_like_reference():
    f = PartConvFilter()
    assert f.setup(3, [64, 1024, 8192], [16, 15, 7], np.ones(65536, np.float32)) == 1
    assert "sufficient delay" in last_error()
    assert f.setup(2, [48, 96], [2, 0], np.ones(100, np.float32)) == 1


def test_frame_count_mismatch_is_an_error():
    m = make_multi([dict(N=[64, 256], P=[6, 0])], [np.ones(500, np.float32)], 64, 1, 1)
    buf = np.zeros(32, np.float32)
    assert m.run(buf, buf, 32) == 7
    assert "frameCount" in last_error()


def test_reset_replays_identically():
    say("case reset")
    rng = np.random.default_rng(6)
    h = decaying_noise_ir(20000, 0.5, 48000, 60)
    x = rng.standard_normal(64 * 300).astype(np.float32)
    m = make_multi([dict(N=[64, 256, 4096], P=[6, 30, 0])], [h], 64, 1, 1)
    y1 = m.process(x)
    m.reset()
    y2 = m.process(x)
    assert np.array_equal(y1, y2)
    assert m.frameNum == 300


def test_config_file_with_wav(tmp_path):
    say("case config file")
    rng = np.random.default_rng(7)
    h = decaying_noise_ir(6000, 0.1, 48000, 70)
    pcm = np.clip(np.round(h / np.max(np.abs(h)) * 0.9 * 32767), -32768, 32767).astype("<i2")
    st = np.stack([pcm, -pcm], axis=1)   # stereo: the reader takes channel 0
    wp = tmp_path / "ir.wav"
    with wave.open(str(wp), "wb") as w:
        w.setnchannels(2)
        w.setsampwidth(2)
        w.setframerate(48000)
        w.writeframes(st.tobytes())
    cfg = tmp_path / "pc.config"
    cfg.write_text("48000 64 1 1 1 2\nir.wav 0 0 0 0.25 3 64 6 256 14 2048 5\n# comment ignored\n")
    m = PartConvMulti(0)
    assert m.setup(str(cfg)) == 0, last_error()
    assert m.info().numPCs == 2       # one repeat, muted
    assert m.instance_partition(0) == ([64, 256, 2048], [6, 14, 1])
    hf = pcm.astype(np.float32) / 32768.0
    x = rng.standard_normal(64 * 200).astype(np.float32)
    y = m.process(x)[:, 0]
    r = O.direct_conv_f64(0.25 * x.astype(np.float64), hf)
    assert rel(y, r) < TOL


def test_device_resident_run_matches_host_run():
    say("case device-resident")
    import torch
    rng = np.random.default_rng(8)
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=i, out_ch=i) for i in range(3)]
    irs = [decaying_noise_ir(50000, 1.0, 48000, 80 + i) for i in range(3)]
    nfr = 520
    x = rng.standard_normal((nfr * 64, 3)).astype(np.float32)
    y_host = make_multi(insts, irs, 64, 3, 3).process(x)
    m = make_multi(insts, irs, 64, 3, 3)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((nfr * 64, 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert m.run_device(yd.data_ptr(), xd.data_ptr(), nfr, True) == 0, last_error()
    assert np.array_equal(yd.cpu().numpy(), y_host)
    m.reset()
    yd.zero_()
    torch.cuda.synchronize()
    assert m.run_device(yd.data_ptr(), xd.data_ptr(), nfr, False) == 0, last_error()
    assert np.array_equal(yd.cpu().numpy(), y_host)
    st = m.stats()
    assert st.ticks == nfr and st.p99_us > 0


def test_persistent_tick_kernel_matches_launch_mode():
    say("case persistent")
    import gc
    import torch
    gc.collect()
    rng = np.random.default_rng(9)
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=i, out_ch=i) for i in range(4)]
    irs = [decaying_noise_ir(60000, 1.0, 48000, 90 + i) for i in range(4)]
    nfr = 700
    x = rng.standard_normal((nfr * 64, 4)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    for relaxed in (False, True):
        ref = make_multi(insts, irs, 64, 4, 4, relaxed=relaxed)
        yl = torch.zeros_like(xd)
        torch.cuda.synchronize()
        assert ref.run_device(yl.data_ptr(), xd.data_ptr(), nfr, True) == 0, last_error()
        ref.cleanup()
        del ref
        m = make_multi(insts, irs, 64, 4, 4, relaxed=relaxed)
        yp = torch.zeros_like(xd)
        torch.cuda.synchronize()
        # two calls: the second continues the stream where the first stopped
        assert m.run_device(yp.data_ptr(), xd.data_ptr(), 300, True, persistent=True) == 0, last_error()
        off = 300 * 64 * 4
        assert m.run_device(yp.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, nfr - 300, True,
                            persistent=True) == 0, last_error()
        st = m.stats()
        say(f"persistent ticks served {st.persistent_ticks} of {nfr}; p50 {st.p50_us:.1f} us")
        assert np.array_equal(yp.cpu().numpy(), yl.cpu().numpy())
        r = O.direct_conv_f64(x[:, 2], irs[2])
        if relaxed:
            r = np.concatenate([np.zeros(64), r[:-64]])
        assert rel(yp.cpu().numpy()[:, 2], r) < TOL
        m.cleanup()


@pytest.mark.parametrize("relaxed", [False, True])
def test_persistent_level_workers_match_host_launched_levels(relaxed):
    """Levels with period 2, 4 and 8 ticks run inside the persistent kernel
    (level workers); the result must be bitwise the one with every level
    launched from the host, and the per-tick launch path.  Runs are split at
    ticks that are not multiples of the level periods."""
    say(f"case level workers relaxed={relaxed}")
    import torch
    from partconv_amd import greedy_partition
    rng = np.random.default_rng(12)
    bs = [64 << i for i in range(9)]
    taps = [30000, 9000, 70000]
    insts, irs = [], []
    for i, T in enumerate(taps):
        N, P = greedy_partition(bs, T)
        insts.append(dict(N=N, P=P, in_ch=i, out_ch=i))
        irs.append(decaying_noise_ir(T, 1.0, 48000, 120 + i))
    nfr = 611
    x = rng.standard_normal((nfr * 64, 3)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    outs = []
    for variant in ("launch", "workers", "no-workers"):
        m = make_multi(insts, irs, 64, 3, 3, relaxed=relaxed)
        if variant == "workers":
            assert m.workers == sum(1 for i in insts for n in i["N"][1:] if n <= 8 * 64)
        if variant == "no-workers":
            m.set_persistent_workers(False)
        yd = torch.zeros_like(xd)
        torch.cuda.synchronize()
        cuts = [0, 1, 302, 303, 500, nfr]
        for a, b in zip(cuts[:-1], cuts[1:]):
            off = a * 64 * 3 * 4
            assert m.run_device(yd.data_ptr() + off, xd.data_ptr() + off, b - a, True,
                                persistent=(variant != "launch")) == 0, last_error()
            say(f"  {variant} ticks {a}-{b}: persistent so far {m.stats().persistent_ticks}")
        if variant != "launch":
            assert m.stats().persistent_ticks == nfr
        outs.append(yd.cpu().numpy())
        m.cleanup()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])
    for i in range(3):
        r = O.direct_conv_f64(x[:, i], irs[i])
        if relaxed:
            r = np.concatenate([np.zeros(64), r[:-64]])
        assert rel(outs[1][:, i], r) < TOL, i


WORKER_CASES = [
    (16, [16, 64, 256, 1024], [6, 6, 6, 0]),
    (32, [32, 64, 256, 2048], [2, 6, 14, 0]),
    (128, [128, 512, 1024, 4096], [6, 2, 6, 0]),
    (256, [256, 512, 2048, 8192], [2, 6, 6, 0]),
    (512, [512, 1024, 4096], [2, 6, 0]),
    (1024, [1024, 2048, 8192], [2, 6, 0]),
]


@pytest.mark.parametrize("N0,N,P", WORKER_CASES)
@pytest.mark.parametrize("relaxed", [False, True])
def test_worker_modes_every_buffer_size(N0, N, P, relaxed):
    """Short-period levels (k = 2, 4, 8) at every buffer size, with partition
    counts above 2: worker workgroups and self mode against host-launched
    levels and the per-tick launch path, bitwise; one channel against fp64.
    Runs are split at ticks that are not multiples of the level periods."""
    say(f"case worker modes N0={N0} N={N} relaxed={relaxed}")
    import torch
    taps = 3 * N[-1] + sum(n * p for n, p in zip(N[:-1], P[:-1])) - N0 // 2
    insts = [dict(N=N, P=P, in_ch=i, out_ch=i) for i in range(3)]
    irs = [decaying_noise_ir(taps, 1.0, 48000, 300 + i) for i in range(3)]
    nfr = 3 * (N[-1] // N0) + 7
    rng = np.random.default_rng(N0)
    x = rng.standard_normal((nfr * N0, 3)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    cuts = sorted({0, 1, nfr // 3 + 1, nfr // 2 + 3, nfr})
    outs = []
    for variant in ("launch", 1, 2, 0):
        m = make_multi(insts, irs, N0, 3, 3, relaxed=relaxed)
        if variant != "launch":
            m.set_persistent_workers(variant)
            if variant == 2:
                assert m.self_mode and m.workers == 0
            if variant == 1:
                assert not m.self_mode and m.workers > 0
        yd = torch.zeros_like(xd)
        torch.cuda.synchronize()
        for a, b in zip(cuts[:-1], cuts[1:]):
            off = a * N0 * 3 * 4
            assert m.run_device(yd.data_ptr() + off, xd.data_ptr() + off, b - a, True,
                                persistent=(variant != "launch")) == 0, last_error()
        if variant != "launch":
            assert m.stats().persistent_ticks == nfr
        outs.append(yd.cpu().numpy())
        m.cleanup()
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    r = O.direct_conv_f64(x[:, 1], irs[1])
    if relaxed:
        r = np.concatenate([np.zeros(N0), r[:-N0]])
    assert rel(outs[0][:, 1], r) < TOL


def test_many_instances_run_short_levels_in_tick_workgroups():
    """Many instances: one worker workgroup per job would not fit in half of
    the device, so each tick workgroup runs its own instance's short-period
    level blocks (self mode).  Bitwise equal to the per-tick launch path; spot
    channels against fp64."""
    say("case self mode (many instances)")
    import torch
    from partconv_amd import greedy_partition
    n = 300
    rng = np.random.default_rng(21)
    lens = rng.integers(4096, 8192, size=n)
    bs = [128 << i for i in range(9)]
    insts, irs = [], []
    for i, T in enumerate(lens):
        N, P = greedy_partition(bs, int(T))
        insts.append(dict(N=N, P=P, in_ch=i, out_ch=i))
        irs.append(decaying_noise_ir(int(T), 0.3, 48000, 1000 + i))
    nfr = 97
    x = rng.standard_normal((nfr * 128, n)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    outs = []
    for persistent in (False, True):
        m = make_multi(insts, irs, 128, n, n)
        if persistent:
            say(f"  worker workgroups {m.workers}, self mode {m.self_mode}")
            assert m.workers == 0 and m.self_mode
        yd = torch.zeros_like(xd)
        torch.cuda.synchronize()
        for a, b in ((0, 45), (45, nfr)):
            off = a * 128 * n * 4
            assert m.run_device(yd.data_ptr() + off, xd.data_ptr() + off, b - a, True,
                                persistent=persistent) == 0, last_error()
        if persistent:
            assert m.stats().persistent_ticks == nfr
        outs.append(yd.cpu().numpy())
        m.cleanup()
    assert np.array_equal(outs[0], outs[1])
    for i in (0, 137, n - 1):
        assert rel(outs[1][:, i], O.direct_conv_f64(x[:, i], irs[i])) < TOL, i


@pytest.mark.parametrize("session", [1, 2])
@pytest.mark.parametrize("relaxed", [False, True])
def test_host_run_through_persistent_session(relaxed, session):
    """run() with host buffers through the resident persistent kernel (session
    1: input frames as tagged words in device memory; 2: release word and
    pinned frames) equals one launch per tick, bitwise; the session survives
    a pause longer than its idle limit (restart) and a device-resident call in
    between; plus fp64."""
    say(f"case host session relaxed={relaxed} session={session}")
    import time
    import torch
    insts = [dict(N=[64, 256, 2048, 16384], P=[6, 14, 14, 0], in_ch=i, out_ch=(i + 1) % 3) for i in range(3)]
    irs = [decaying_noise_ir(40000, 1.0, 48000, 500 + i) for i in range(3)]
    nfr = 300
    rng = np.random.default_rng(77)
    x = rng.standard_normal((nfr * 64, 3)).astype(np.float32)
    ref = make_multi(insts, irs, 64, 3, 3, relaxed=relaxed)
    ref.set_host_session(False)
    y_ref = ref.process(x)
    assert ref.stats().persistent_ticks == 0
    ref.cleanup()
    m = make_multi(insts, irs, 64, 3, 3, relaxed=relaxed)
    m.set_host_session(session)
    a = 137 * 64
    y1 = m.process(x[:a])
    time.sleep(1.3)                      # idle past the session limit: restarted on the next call
    b = 201 * 64
    y2 = m.process(x[a:b])
    # device-resident ticks in between (the session stops, then resumes)
    xd = torch.from_numpy(np.ascontiguousarray(x[b:b + 40 * 64])).cuda()
    yd = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert m.run_device(yd.data_ptr(), xd.data_ptr(), 40, True, persistent=True) == 0, last_error()
    y3 = yd.cpu().numpy()
    y4 = m.process(x[b + 40 * 64:])
    y = np.concatenate([y1, y2, y3, y4])
    assert m.stats().persistent_ticks == nfr
    m.cleanup()
    assert np.array_equal(y, y_ref)
    r = O.direct_conv_f64(x[:, 0], irs[0])
    if relaxed:
        r = np.concatenate([np.zeros(64), r[:-64]])
    assert rel(y[:, 1], r) < TOL


@pytest.mark.parametrize("session", [1, 2])
@pytest.mark.parametrize("N0", [64, 128])
def test_host_session_summed_outputs(N0, session):
    """Host session with several sources per output channel (the reduce mix),
    a muted instance and an output channel with no source: bitwise equal to
    one launch per tick, and a reset mid-stream replays the same output."""
    say(f"case host session reduce N0={N0} session={session}")
    N = [N0, 4 * N0, 32 * N0]
    P = [6, 14, 0]
    insts = [dict(N=N, P=P, in_ch=0, out_ch=0), dict(N=N, P=P, in_ch=1, out_ch=0, scale=0.5),
             dict(N=N, P=P, in_ch=1, out_ch=2), dict(N=N, P=P, in_ch=2, out_ch=2, mute=1),
             dict(N=N, P=P, in_ch=2, out_ch=3)]
    irs = [decaying_noise_ir(30000, 0.8, 48000, 900 + i) for i in range(len(insts))]
    nfr = 3 * 32 + 7
    rng = np.random.default_rng(5)
    x = rng.standard_normal((nfr * N0, 3)).astype(np.float32)
    ref = make_multi(insts, irs, N0, 3, 4)
    ref.set_host_session(False)
    y_ref = ref.process(x)
    ref.cleanup()
    m = make_multi(insts, irs, N0, 3, 4)
    m.set_host_session(session)
    y = m.process(x)
    assert m.stats().persistent_ticks == nfr
    assert np.array_equal(y, y_ref)
    assert not np.any(y[:, 1])
    m.reset()
    y2 = m.process(x[:40 * N0])
    m.cleanup()
    assert np.array_equal(y2, y_ref[:40 * N0])
    r = O.direct_conv_f64(x[:, 0], irs[0]) + O.direct_conv_f64(0.5 * x[:, 1].astype(np.float64), irs[1])
    assert rel(y[:, 0], r) < TOL


def test_uneven_batched_instances_greedy_partitions():
    """Configuration-3 shape at reduced size: uneven IR lengths (log-uniform),
    each instance with its own greedy partition over 128..32768, one batched
    engine (persistent tick kernel), every channel against fp64."""
    say("case uneven batched (cfg3 shape)")
    import torch
    from partconv_amd import greedy_partition
    from partconv_amd.workloads import cfg3_irs, cfg3_lengths
    n = 40
    lens = cfg3_lengths(seed=3, instances=n, min_log2=12, max_log2=17)
    irs = cfg3_irs(lens, seed=3)
    bs = [128 << i for i in range(9)]
    insts = []
    for i, T in enumerate(lens):
        N, P = greedy_partition(bs, int(T))
        insts.append(dict(N=N, P=P, in_ch=i, out_ch=i))
    nfr = 2 * 256 + 30
    rng = np.random.default_rng(33)
    x = rng.standard_normal((nfr * 128, n)).astype(np.float32)
    m = make_multi(insts, irs, 128, n, n)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    torch.cuda.synchronize()
    assert m.run_device(yd.data_ptr(), xd.data_ptr(), nfr, True, persistent=True) == 0, last_error()
    y = yd.cpu().numpy()
    worst = 0.0
    for i in range(n):
        r = O.direct_conv_f64(x[:, i], irs[i])
        worst = max(worst, rel(y[:, i], r))
    say(f"worst channel rel err {worst:.2e}; persistent ticks {m.stats().persistent_ticks}")
    assert worst < TOL
    m.cleanup()


@pytest.mark.parametrize("n,m,k", [(3, 3, 3), (3, 3, 1), (2, 6, 3)])
def test_partconvmax_channel_modes(n, m, k):
    """PartConvMax (src/partconv.cpp:653-847, 1219-1303): parallel distinct IRs,
    one IR on every channel, filterbank; per-channel buffers; IRs read with a
    stride from one array; every output channel against fp64."""
    say(f"case PartConvMax n={n} m={m} k={k}")
    from partconv_amd import PartConvMax
    taps, stride, N0 = 9000, 2, 64
    irs = [decaying_noise_ir(taps, 0.5, 48000, 400 + i) for i in range(k)]
    imp = np.zeros(stride * taps * k, dtype=np.float32)
    for i in range(k):
        imp[stride * taps * i:stride * taps * (i + 1):stride] = irs[i]
    pm = PartConvMax(0)
    bs = [64, 256, 1024, 4096, 16384]
    assert pm.setup(48000, n, m, k, imp, taps, stride, bs, len(bs), 1) == 0, last_error()
    assert pm.buffer_size == N0 and pm.numInputChannels == n and pm.numOutputChannels == m
    rng = np.random.default_rng(n * 10 + k)
    nfr = 70
    x = rng.standard_normal((nfr * N0, n)).astype(np.float32)
    y = pm.process(x)
    assert pm.frameNum == nfr
    for o in range(m):
        if n == m == k:
            i_ir, i_in = o, o
        elif k == 1:
            i_ir, i_in = 0, o
        else:
            i_ir, i_in = o // n, o % n
        r = O.direct_conv_f64(x[:, i_in], irs[i_ir])
        assert rel(y[:, o], r) < TOL, o
    pm.cleanup()


def test_partconvmax_rejects_unsupported_channel_layout():
    from partconv_amd import PartConvMax
    pm = PartConvMax(0)
    h = np.ones(1000, dtype=np.float32)
    assert pm.setup(48000, 2, 3, 2, h, 500, 1, [64, 256, 1024], 3, 1) == 1
    assert "channel configuration" in last_error()


def test_mono_reference_configuration_full_length():
    """The reference's own CPU-runnable case: 10 s of white noise at 44.1 kHz,
    a 65536-tap decaying-noise IR (unit L2), block sizes [64, 1024, 8192] with
    the nearest feasible counts [30, 15, 6] (the stated [16, 15, 7] fails the
    delay rule), sample by sample against the fp32 oracle and fp64, strict and
    relaxed; the stated partition is rejected like the reference rejects it."""
    say("case configuration 1 (mono, 10 s)")
    from partconv_amd.workloads import CFG1, CFG1_STATED, cfg1
    x, h = cfg1(seed=1)
    for relaxed in (False, True):
        m = make_multi([dict(N=CFG1["N"], P=CFG1["P"], in_ch=0, out_ch=0)], [h], 64, 1, 1,
                       relaxed=relaxed, FS=CFG1["FS"])
        y = m.process(x)[:x.size, 0]
        m.cleanup()
        ref = O.OracleMulti(CFG1["FS"], 64, 1, 1, [dict(N=CFG1["N"], P=CFG1["P"])], [h],
                            relaxed=relaxed).process(x)[:x.size, 0]
        assert rel(y, ref) < TOL
        r = O.direct_conv_f64(x, h)
        if relaxed:
            r = np.concatenate([np.zeros(64), r[:-64]])
        assert rel(y, r) < TOL
    bad = PartConvMulti(0)
    assert bad.setup_mem(CFG1["FS"], 64, 1, 1, [dict(ir=h, **CFG1_STATED)]) != 0
    assert "partition" in last_error().lower() or "level" in last_error().lower()


def test_mimo_as_pair_instances():
    """MIMO in the reference's own terms: one instance per (input i -> output o)
    pair, outputs summed (src/partconvMulti.cpp:911-921).  8x8 at reduced size;
    every output against the fp64 sum of per-pair convolutions; launch and
    persistent modes bitwise equal."""
    say("case MIMO pairs")
    import torch
    from partconv_amd.workloads import cfg4_irs
    n_in = n_out = 8
    irs = cfg4_irs(seed=4, n_in=n_in, n_out=n_out, taps=9000)
    insts, hs = [], []
    for o in range(n_out):
        for i in range(n_in):
            insts.append(dict(N=[256, 4096], P=[30, 0], in_ch=i, out_ch=o))
            hs.append(irs[o][i])
    nfr = 2 * 16 + 10
    rng = np.random.default_rng(44)
    x = rng.standard_normal((nfr * 256, n_in)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    outs = []
    for persistent in (False, True):
        m = make_multi(insts, hs, 256, n_in, n_out)
        yd = torch.zeros((nfr * 256, n_out), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        assert m.run_device(yd.data_ptr(), xd.data_ptr(), nfr, True, persistent=persistent) == 0, last_error()
        outs.append(yd.cpu().numpy())
        m.cleanup()
    assert np.array_equal(outs[0], outs[1])
    y = outs[0]
    for o in range(n_out):
        r = sum(O.direct_conv_f64(x[:, i], irs[o][i]) for i in range(n_in))
        assert rel(y[:, o], r) < TOL, o
