"""Pins for the CPU oracle (oracle/pc_oracle.c, oracle/rules.py).

The reference ships no golden vectors, known-answer tests or fixtures for this
path (its test*.cpp files are timing loops), and FFTW is absent here, so the
oracle is pinned by mathematics that fixes its output independently of its own
formulas:
  * its real FFT against numpy's float64 rfft/irfft;
  * the filter-spectrum store against the closed form
    H_{L,p}[k] = (-1)^k / (2N) * sum_j h[T_L + pN + j] e^{-i pi k j / N};
  * end-to-end output against float64 direct convolution (strict = zero
    latency, relaxed = N0 samples late), impulses, and brute force on tiny cases;
  * integer rules against hand-derived known answers.
Planted mistakes (oracle mutations) must each fail at least one pin.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import rules as R
from tests.golden import load_golden


def rel_err(y, r):
    return float(np.max(np.abs(y - r)) / np.sqrt(np.mean(r ** 2)))


# ---------------------------------------------------------------- FFT pin
@pytest.mark.parametrize("N", [1, 2, 4, 32, 64, 512, 4096])
def test_rfft_matches_numpy_f64(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(2 * N)
    X = O.rfft(x, "f64")
    assert np.allclose(X, np.fft.rfft(x), atol=1e-12 * np.sqrt(2 * N))
    Y = rng.standard_normal(N + 1) + 1j * rng.standard_normal(N + 1)
    Y[0] = Y[0].real
    Y[-1] = Y[-1].real
    # FFTW c2r is unnormalised: numpy's irfft times 2N
    assert np.allclose(O.irfft(Y, "f64"), np.fft.irfft(Y, 2 * N) * 2 * N, atol=1e-10 * N)


def test_rfft_f32_accuracy():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(2 * 16384)
    X = O.rfft(x.astype(np.float32), "f32")
    R_ = np.fft.rfft(x)
    assert np.max(np.abs(X - R_)) / np.max(np.abs(R_)) < 1e-6


# ---------------------------------------------------------------- H store
def test_filter_spectra_closed_form():
    rng = np.random.default_rng(11)
    N, P = [16, 64, 256], [6, 6, 0]
    h = rng.standard_normal(1000)
    f = O.OracleFilter(N, P, h, "f64")
    hz = np.zeros(sum(n * p for n, p in zip(f.N, f.P)))
    hz[:h.size] = h
    T = 0
    for L, (n, p_cnt) in enumerate(zip(f.N, f.P)):
        k = np.arange(n + 1)
        j = np.arange(n)
        for p in range(p_cnt):
            taps = hz[T + p * n:T + (p + 1) * n]
            ref = ((-1.0) ** k) / (2 * n) * (np.exp(-1j * np.pi * np.outer(k, j) / n) @ taps)
            assert np.allclose(f.H(L, p), ref, atol=1e-12)
        T += n * p_cnt


# ------------------------------------------------------- end-to-end output
PARTS = [
    ([4], [0]),
    ([8, 16], [2, 0]),
    ([16, 64, 256], [6, 6, 0]),
    ([64, 256, 2048], [6, 14, 0]),
    ([32, 64, 128, 256], [2, 2, 2, 0]),
]


@pytest.mark.parametrize("N,P", PARTS)
@pytest.mark.parametrize("relaxed", [False, True])
def test_impulse_gives_h(N, P, relaxed):
    rng = np.random.default_rng(sum(N))
    h = rng.standard_normal(3 * N[-1] + 7)
    N0 = N[0]
    for f0, off in [(0, 0), (3, 1), (17, N0 - 1)]:
        f = O.OracleFilter(N, P, h, "f64")
        n = (f0 + 2) * N0 + h.size + 8 * N[-1]
        x = np.zeros(n)
        t0 = f0 * N0 + off
        x[t0] = 1.0
        y = f.process(x, relaxed=relaxed)
        ref = np.zeros_like(y)
        d = N0 if relaxed else 0
        ref[t0 + d:t0 + d + h.size] = h[:max(0, min(h.size, y.size - t0 - d))]
        assert np.max(np.abs(y - ref)) < 1e-12


@pytest.mark.parametrize("N,P", PARTS)
def test_random_matches_direct_conv(N, P):
    rng = np.random.default_rng(7 + len(N))
    h = rng.standard_normal(4 * N[-1] + 3)
    x = rng.standard_normal(N[0] * 200 + 2 * h.size)
    r = O.direct_conv_f64(x, h)
    y64 = O.OracleFilter(N, P, h, "f64").process(x)[:x.size]
    assert rel_err(y64, r) < 1e-12
    y32 = O.OracleFilter(N, P, h, "f32").process(x)[:x.size]
    assert rel_err(y32, r) < 1e-5
    yr = O.OracleFilter(N, P, h, "f64").process(x, relaxed=True)[:x.size]
    assert np.max(np.abs(yr[N[0]:] - r[:-N[0]])) < 1e-10 * np.sqrt(np.mean(r ** 2))


def test_brute_force_tiny():
    rng = np.random.default_rng(3)
    h = rng.standard_normal(23)
    x = rng.standard_normal(64)
    y = O.OracleFilter([4, 8], [2, 0], h, "f64").process(x)[:64]
    assert np.max(np.abs(y - O.brute_conv(x, h))) < 1e-12


def test_linearity_and_time_invariance():
    rng = np.random.default_rng(9)
    N, P = [16, 64], [6, 0]
    h = rng.standard_normal(300)
    a, b = rng.standard_normal(1600), rng.standard_normal(1600)
    f = lambda s: O.OracleFilter(N, P, h, "f64").process(s)
    assert np.allclose(f(2 * a - 3 * b), 2 * f(a) - 3 * f(b), atol=1e-11)
    s = 16 * 5
    sh = np.concatenate([np.zeros(s), a])[:1600]
    assert np.allclose(f(sh)[s:], f(a)[:1600 - s], atol=1e-11)


def test_reset_restores_initial_state():
    rng = np.random.default_rng(12)
    h = rng.standard_normal(500)
    f = O.OracleFilter([16, 64], [6, 0], h, "f64")
    x = rng.standard_normal(800)
    y1 = f.process(x)
    f.reset()
    y2 = f.process(x)
    assert np.array_equal(y1, y2)


# ------------------------------------------------------------ multi engine
def test_multi_mix_scale_mute_and_level_map():
    rng = np.random.default_rng(21)
    insts = [
        dict(N=[16, 64, 256], P=[6, 6, 0], in_ch=0, out_ch=1, mute=0, scale=0.5),
        dict(N=[16, 32, 128], P=[2, 6, 0], in_ch=1, out_ch=1, mute=0, scale=2.0),
        dict(N=[16, 64], P=[6, 0], in_ch=1, out_ch=0, mute=0, scale=1.0),
        dict(N=[16, 64], P=[6, 0], in_ch=0, out_ch=0, mute=1, scale=1.0),
    ]
    irs = [rng.standard_normal(n) for n in (900, 700, 300, 300)]
    m = O.OracleMulti(48000, 16, 2, 2, insts, irs, prec="f64")
    sizes, lmap = m.level_map()
    assert sizes == [1, 2, 4, 8, 16]
    assert lmap == [[0, -1, 1, -1, 2], [0, 1, -1, 2, -1], [0, -1, 1, -1, -1], [0, -1, 1, -1, -1]]
    assert (sizes, lmap) == R.level_map(16, [d["N"] for d in insts])
    x = rng.standard_normal((16 * 150, 2))
    y = m.process(x)
    r0 = O.direct_conv_f64(x[:, 1], irs[2])
    r1 = O.direct_conv_f64(0.5 * x[:, 0], irs[0]) + O.direct_conv_f64(2.0 * x[:, 1], irs[1])
    assert np.max(np.abs(y[:, 0] - r0)) < 1e-11
    assert np.max(np.abs(y[:, 1] - r1)) < 1e-11
    mr = O.OracleMulti(48000, 16, 2, 2, insts, irs, relaxed=True, prec="f64")
    yr = mr.process(x)
    assert np.max(np.abs(yr[16:] - y[:-16])) < 1e-11


def test_multi_rejects_bad_setup():
    h = [np.ones(100)]
    with pytest.raises(ValueError):
        O.OracleMulti(48000, 32, 1, 1, [dict(N=[16, 64], P=[6, 0])], h)
    with pytest.raises(ValueError):
        O.OracleMulti(48000, 16, 1, 1, [dict(N=[16, 64], P=[6, 0], out_ch=3)], h)


# ----------------------------------------------------------- integer rules
def test_partition_known_answers():
    p = R.check_partition([64, 256, 2048, 16384], [6, 14, 14, 0], 262144)
    assert p.P == [6, 14, 14, 15] and p.delay == [2, 8, 64, 512] and p.total == 278400
    with pytest.raises(R.PartitionError, match="sufficient delay"):
        R.check_partition([64, 1024, 8192], [16, 15, 7], 65536)
    p = R.check_partition([64, 1024, 8192], [30, 15, 6], 65536)
    assert p.P == [30, 15, 6] and p.delay == [2, 32, 272] and p.total == 66432
    p = R.check_partition([256, 4096], [30, 0], 65536)
    assert p.P == [30, 15]
    p = R.check_partition([256, 2048, 16384], [14, 14, 0], 65536)
    assert p.P == [14, 14, 3]
    p = R.check_partition([256, 2048, 16384], [14, 14, 0], 1 << 20)
    assert p.P == [14, 14, 63]
    g = R.greedy_partition([128 << i for i in range(9)], 4096)
    assert g.N == [128, 256, 512, 1024, 2048] and g.P == [2, 2, 2, 2, 1]
    g = R.greedy_partition([128 << i for i in range(9)], 1 << 20)
    assert g.N[-1] == 32768 and g.P == [2] * 8 + [31]
    with pytest.raises(R.PartitionError, match="power of 2"):
        R.check_partition([48, 96], [2, 0], 100)
    with pytest.raises(R.PartitionError, match="<="):
        R.check_partition([64, 64], [2, 0], 100)
    with pytest.raises(R.PartitionError, match="not required"):
        R.check_partition([64, 256], [6, 0], 300)


def test_rules_agree_with_c_oracle_on_random_partitions():
    rng = np.random.default_rng(99)
    for _ in range(300):
        M = int(rng.integers(1, 5))
        lg = sorted(set(int(v) for v in rng.integers(2, 12, size=M)))
        N = [1 << v for v in lg]
        P = [int(v) for v in rng.integers(1, 20, size=len(N))]
        ir = int(rng.integers(1, 6000))
        try:
            py = R.check_partition(N, P, ir)
        except R.PartitionError:
            py = None
        try:
            c = O.OracleFilter(N, P, np.ones(ir), "f64")
        except ValueError:
            c = None
        assert (py is None) == (c is None), (N, P, ir)
        if py is not None:
            assert py.P == c.P and py.delay == c.delay


def test_config_grammar():
    text = """48000 64 2 2 2 2
ir_a.wav 0 0 1 0.25 3 64 6 256 14 2048 99
ir_b.wav 1 1 0 1.5 2 64 6 256
trailing text is ignored
"""
    cfg = R.parse_config(text)
    assert (cfg.FS, cfg.buffer_size, cfg.n_in, cfg.n_out, cfg.num_pcs) == (48000, 64, 2, 2, 2)
    assert len(cfg.instances) == 4
    a, b, ra, rb = cfg.instances
    assert a.N == [64, 256, 2048] and a.P == [6, 14, 0] and a.scale == 0.25 and a.out_ch == 1
    assert b.N == [64, 256] and b.P == [6, 0] and b.mute == 1
    assert ra.mute == 1 and rb.mute == 1 and ra.N == a.N
    with pytest.raises(R.ConfigError):
        R.parse_config("48000 32 1 1 1 1\nx.wav 0 0 0 1.0 2 64 6 256\n")


# ------------------------------------------------------------ golden vectors
def test_golden_vectors():
    """Committed small fixtures (tests/golden/make_golden.py): fp64 direct
    convolution outputs for seeded inputs; the f32 oracle must match them."""
    for case in load_golden():
        f = O.OracleFilter(case["N"], case["P"], case["h"], "f32")
        y = f.process(case["x"], relaxed=case["relaxed"])[:case["y"].size]
        assert rel_err(y, case["y"]) < 1e-5, case["name"]


# --------------------------------------------------------- planted mistakes
def _pins_pass():
    """Run a compact set of the pins above; False as soon as one fails."""
    rng = np.random.default_rng(1234)
    N, P = [16, 64, 256], [6, 6, 0]
    h = rng.standard_normal(1200)
    x = rng.standard_normal(16 * 300)
    r = O.direct_conv_f64(x, h)
    if rel_err(O.OracleFilter(N, P, h, "f64").process(x)[:x.size], r) > 1e-9:
        return False
    f = O.OracleFilter(N, P, h, "f64")
    k = np.arange(17)
    ref = ((-1.0) ** k) / 32 * (np.exp(-1j * np.pi * np.outer(k, np.arange(16)) / 16) @ h[:16])
    return bool(np.allclose(f.H(0, 0), ref, atol=1e-12))


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mutation):
    assert _pins_pass()
    try:
        O.set_mutation(mutation, "f64")
        assert not _pins_pass(), f"mutation {mutation} slipped through the pins"
    finally:
        O.set_mutation(0, "f64")
    assert _pins_pass()


def test_threaded_reference_schedule_matches_restatement():
    """The CPU baseline's threaded runner (the reference's run/worker
    scheduling, src/partconvMulti.cpp:815-1013: level 0 inline, one worker
    group per global block size, a wait at each level's deadline tick) gives
    bitwise the single-threaded restatement's output, for instances with
    different partitions and an input-column map, over several calls."""
    from oracle import oracle as O
    from partconv_amd.workloads import decaying_noise_ir
    insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=1), dict(N=[64, 512], P=[14, 0], in_ch=0),
             dict(N=[64, 256, 1024, 4096], P=[6, 6, 6, 0], in_ch=1)]
    irs = [decaying_noise_ir(t, 0.3, 48000, 40 + i) for i, t in enumerate((9000, 4000, 20000))]
    x = np.random.default_rng(8).standard_normal((400 * 64, 2)).astype(np.float32)
    ref = O.OracleMulti(48000, 64, 2, 3, [dict(out_ch=i, **d) for i, d in enumerate(insts)], irs).process(x)
    fs = [O.OracleFilter(d["N"], d["P"], h) for d, h in zip(insts, irs)]
    chans = [d["in_ch"] for d in insts]
    ys = []
    for a, b in ((0, 150), (150, 400)):        # the stream continues across calls
        y, lat, lv = O.rt_run(fs, x[a * 64:b * 64], 2, in_ch=chans)
        ys.append(y)
        assert lat.shape == (b - a,) and np.all(lat > 0)
    assert [n for n, _ in lv] == [64, 256, 512, 1024, 2048, 4096]
    assert np.array_equal(np.concatenate(ys), ref)
