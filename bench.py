#!/usr/bin/env python3
"""Benchmark: channel-samples/s of the partitioned-convolution engine at p99
block latency below the block period, batched vs plain, on one GPU.

Workload (16-channel drop-in): 16 independent channels at 48 kHz, each with its
own 262144-tap exponentially decaying noise IR (RT60 uniform in [0.5, 3] s),
partition [64x6, 256x14, 2048x14, 16384x15], buffer 64 samples (1.333 ms
period), seeded white-noise input resident in HBM.

A step is one period of the largest level: 256 ticks of 64 samples for all 16
channels (every level runs a whole number of blocks).  Each tick is submitted
and waited for individually (real-time semantics); the per-tick latency is the
host's submit -> output-visible time.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FS = 48000
N0 = 64
CHANNELS = 16
TAPS = 262144
PART_N = [64, 256, 2048, 16384]
PART_P = [6, 14, 14, 0]
TICKS_PER_STEP = 256


def build_engine(mode, irs):
    from partconv_amd import PartConvMulti, last_error
    m = PartConvMulti(0, mode=mode)
    insts = [dict(N=PART_N, P=PART_P, ir=irs[c], in_ch=c, out_ch=c) for c in range(CHANNELS)]
    if m.setup_mem(FS, N0, CHANNELS, CHANNELS, insts) != 0:
        raise RuntimeError(last_error())
    return m


def timed_run(m, xd, yd, warmup, steps, torch, persistent=False):
    """warmup + timed steps of TICKS_PER_STEP ticks, every tick released and
    waited for individually; returns (seconds, stats).  With the persistent
    kernel the warmup and the timed steps are one launch each (a real-time
    stream keeps its kernel running; relaunching it every step would add a
    launch, about 0.1-0.15 ms, per 256 ticks); the per-tick launch engines
    make one call per step."""
    stride = TICKS_PER_STEP * N0 * CHANNELS
    calls = [(0, warmup)] if persistent else [(w, 1) for w in range(warmup)]
    for first, n in calls:
        if n and m.run_device(yd.data_ptr() + 4 * first * stride, xd.data_ptr() + 4 * first * stride,
                              n * TICKS_PER_STEP, True, persistent=persistent) != 0:
            raise RuntimeError("warmup failed")
    m.clear_stats()
    torch.cuda.synchronize()
    calls = [(warmup, steps)] if persistent else [(warmup + s, 1) for s in range(steps)]
    t0 = time.perf_counter()
    for first, n in calls:
        off = first * stride
        if m.run_device(yd.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, n * TICKS_PER_STEP, True,
                        persistent=persistent) != 0:
            raise RuntimeError("step failed")
    torch.cuda.synchronize()
    return time.perf_counter() - t0, m.stats()


def _lib_process(m, y, x, nframes):
    """pcm_process: nframes run() calls over interleaved host buffers."""
    from partconv_amd._lib import lib
    return lib().pcm_process(m._h, y.ctypes.data, x.ctypes.data, int(nframes))


HBM_PEAK_GBPS = 8000.0


def block_bytes(N, P):
    """Canonical algorithmic bytes of one level block: H (P rows) and FDL (P rows
    read, one written) of N+1 complex bins, the 2N-sample window read and the N
    outputs written: 8 (N+1)(2P+1) + 12 N."""
    return 8.0 * (N + 1) * (2 * P + 1) + 12.0 * N


def tick_bytes(parts, n0):
    """Algorithmic bytes per tick of a set of instances: each level of block
    size N runs one block every N / n0 ticks."""
    return sum(block_bytes(n, p) * n0 / n for N, P in parts for n, p in zip(N, P))


def workload_roofline(bytes_per_tick, ticks, secs, ch_per_tick, what, limit):
    """The timed region's roofline: the algorithmic bytes of every block served
    (canonical 8(N+1)(2P+1)+12N per level block, or the engine's own bytes
    where stated) over the wall time of the timed region, against 8 TB/s.
    `limit` says what bounds the region when it is not HBM."""
    gbps = bytes_per_tick * ticks / secs / 1e9
    return {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
            "traffic": None, "kernel": what, "alg_bytes_per_tick": bytes_per_tick,
            "alg_bytes_per_ch_sample": bytes_per_tick / ch_per_tick, "limit": limit,
            "traffic_note": "per-kernel HBM traffic is in kernel_rooflines (counter runs)"}


def diff_step_bytes(C, T, Ns, Ps, dh_chunks):
    """Algorithmic HBM bytes of one forward + dL/dh step of the differentiable
    form: every kernel of the step with each of its inputs read once and each
    output written once (signal samples 4 B, spectra 8 B per bin).  Per level:
    the filter spectra from h (h changes every step), the forward transform of
    the input windows, the spectral product, the inverse transform into y (read
    and written from the second level on), the gradient transform, the dL/dh
    partial sums (G and X read, partials written), their reduction and the
    inverse of the P filter-gradient spectra; plus the loss (y and the target
    read, the gradient written).  Returns (total, {kernel: bytes})."""
    sig = 4.0 * C * T
    parts = {"loss": 3.0 * sig}
    for li, (N, P) in enumerate(zip(Ns, Ps)):
        M = -(-T // N)
        spec, hsp = 8.0 * C * M * N, 8.0 * P * N
        part = 8.0 * P * C * (-(-M // dh_chunks[li])) * N
        for k, v in (("filter_transform", 4.0 * P * N + hsp), ("forward_transform", sig + spec),
                     ("product", 2.0 * spec + hsp), ("inverse_transform", spec + (sig if li == 0 else 2.0 * sig)),
                     ("gradient_transform", sig + spec), ("dh_partial", 2.0 * spec + part),
                     ("dh_reduce", part + hsp), ("dh_inverse", hsp + 4.0 * P * N)):
            parts[k] = parts.get(k, 0.0) + v
    return sum(parts.values()), parts


def mimo_engine_tick_bytes(levels, n_in, n_out, n0):
    """Bytes the real-time MIMO engine moves per tick (its shared-input
    formulation, not the pair instances'): per level block of N bins and P
    partitions, the filter spectra of every pair (level 0's p = 0 block stays in
    the output workgroups' LDS for the whole run), P delay-line rows read and
    one written per input, the delay-line sum written and read per output, the
    2N-sample window read per input and N outputs written per output; a level
    runs one block every N / n0 ticks."""
    tot = 0.0
    for l, (N, P) in enumerate(levels):
        h = 8.0 * N * (P - (1 if l == 0 else 0)) * n_out * n_in
        blk = h + 8.0 * N * (P + 1) * n_in + 16.0 * N * n_out + 8.0 * N * n_in + 4.0 * N * n_out
        tot += blk * n0 / N
    return tot


def kernel_roofline(avg_us, launches, bytes_per, kernel):
    """Roofline of one probed kernel (HIP events on its own stream)."""
    if not avg_us:
        return None
    gbps = bytes_per / (avg_us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
            "traffic": None, "kernel": kernel, "avg_us": avg_us, "launches": launches,
            "alg_bytes_per_launch": bytes_per}


def with_traffic(roof, name):
    """Attach the HBM bytes per launch that FETCH_SIZE / WRITE_SIZE counters
    measured for this kernel (copy-calibrated, profiles/<name>) to a kernel
    roofline, labelled with that file: the counters come from a separate
    profiled run of the same bench command (recorded in the file), not from
    this run."""
    path = os.path.join(ROOT, "profiles", name)
    if roof and os.path.exists(path):
        with open(path) as fh:
            t = json.load(fh)
        roof["traffic"] = 1024.0 * (t["fetch_kib_per_dispatch_scaled"] + t["write_kib_per_dispatch_scaled"])
        roof["traffic_source"] = (f"profiles/{name}: rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                                  f"copy-calibrated, separate profiled run"
                                  + (f" at commit {t['commit']}" if t.get("commit") else ""))
    return roof


def cpu_reference_schedule(specs, in_ch, n_in, n0, budget_s, channels_per_tick, what, out_sum=None):
    """CPU baseline with the reference's scheduling (src/partconvMulti.cpp:815-1013):
    the fp32 CPU restatement of PartConvFilter per instance, level 0 on the
    calling thread, one worker-thread group per global block size (started at
    the level's start tick, waited for at the tick that first reads it), ticks
    back to back over a cyclic seeded white-noise input, for about budget_s.
    specs: (N, P, ir) per instance.  Reports channel-samples/s, per-tick p50 /
    p99 against the period, and per level the reference's own load metric,
    100 t_level / (N_L / FS) for one instance's block (src/partconv.cpp:553-559),
    and the share of its worker threads' time."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    fs = [O.OracleFilter(N, P, h) for N, P, h in specs]
    G = len({n for N, _, _ in specs for n in N[1:]})
    nth = max(1, (cores - 1) // max(1, G))
    chunk = 512
    x = np.random.default_rng(123).standard_normal((chunk * n0, n_in)).astype(np.float32)
    lats, lvl_sum, calls = [], None, 0
    t0 = time.perf_counter()
    while True:
        _, lat, lv = O.rt_run(fs, x, nth, in_ch=in_ch)
        lats.append(lat)
        lvl_sum = [t for _, t in lv] if lvl_sum is None else [a + t for a, (_, t) in zip(lvl_sum, lv)]
        calls += 1
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    for f in fs:
        f.close()
    lat = np.concatenate(lats) * 1e6
    period = 1e6 * n0 / FS
    ticks = lat.size
    levels = []
    for (n, _), tsum in zip(lv, lvl_sum):
        t = tsum / calls
        users = sum(1 for N, _, _ in specs if n in N)
        share = users / (nth if n > n0 else 1)
        levels.append({"block": n, "block_us": 1e6 * t, "pct_of_block_period": 100.0 * t / (n / FS),
                       "thread_load_pct": 100.0 * t * share / (n / FS)})
    return {"value": channels_per_tick * ticks * n0 / wall, "unit": "channel-samples/s",
            "cores": 1 + nth * G, "kind": "port",
            "sample": f"{ticks} ticks back to back ({ticks * n0 / FS:.1f} s of audio) of {what}, "
                      f"{wall:.1f} s of CPU time",
            "schedule": f"reference's: level 0 on the calling thread, {G} worker groups of {nth} threads "
                        f"(one per global block size), a wait at each level's deadline tick",
            "fft_backend": "project-written radix-2 complex FFT + real split, fp32 (FFTW is absent on this "
                           "host; the reference uses FFTW 3.2.2)",
            "latency_us": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max()), "period": period, "misses": int(np.sum(lat > period))},
            "levels": levels}


def cpu_free_running(irs, x_host, budget_s):
    """The fp32 oracle on the same workload with no per-tick synchronisation:
    one instance per channel, the channels spread over the host's cores
    (threads; the oracle runs outside the GIL), every level inline at its start
    tick; ticks chosen for ~budget_s.  An upper bound on CPU throughput, not a
    real-time schedule."""
    import threading
    from oracle import oracle as O
    cores = max(1, min(CHANNELS, len(os.sched_getaffinity(0))))
    # one mono engine per channel (same partition, same IR, same input column)
    ms = [O.OracleMulti(FS, N0, 1, 1, [dict(N=PART_N, P=PART_P, in_ch=0, out_ch=0)], [irs[c]], prec="f32")
          for c in range(CHANNELS)]
    cols = [np.ascontiguousarray(x_host[:, c:c + 1]) for c in range(CHANNELS)]

    def run(nticks):
        def work(w):
            for c in range(w, CHANNELS, cores):
                ms[c].process_frames(cols[c], nticks)
        th = [threading.Thread(target=work, args=(w,)) for w in range(cores)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0

    t_cal = run(TICKS_PER_STEP)
    for m in ms:
        m.reset()
    steps = max(1, int(budget_s / max(t_cal, 1e-6)))
    steps = min(steps, x_host.shape[0] // (TICKS_PER_STEP * N0))
    ticks = steps * TICKS_PER_STEP
    t = run(ticks)
    for m in ms:
        m.close()
    return dict(value=CHANNELS * ticks * N0 / t, unit="channel-samples/s", cores=cores,
                sample=f"{ticks} ticks x {CHANNELS} channels ({ticks * N0 / FS:.2f} s of audio), one instance per "
                       f"channel on {cores} threads, no per-tick synchronisation")


def bench_uneven(args):
    """Configuration 3: 512 instances, IR lengths log-uniform in [2^12, 2^20]
    (decaying noise, -60 dB at the last tap), greedy partition per instance over
    block sizes 128..32768, buffer 128 at 48 kHz (2.667 ms period), one
    batched engine with the persistent tick kernel; one JSON line."""
    import torch
    from partconv_amd import PartConvMulti, greedy_partition, last_error
    from partconv_amd.workloads import cfg3_irs, cfg3_lengths
    n, n0 = args.instances, 128
    lens = cfg3_lengths(seed=3, instances=n)
    irs = cfg3_irs(lens, seed=3)
    bs = [128 << i for i in range(9)]
    insts = []
    for i, T in enumerate(lens):
        N, P = greedy_partition(bs, int(T))
        insts.append(dict(N=N, P=P, ir=irs[i], in_ch=i, out_ch=i))
    m = PartConvMulti(0)
    t_setup = time.perf_counter()
    if m.setup_mem(FS, n0, n, n, insts) != 0:
        raise RuntimeError(last_error())
    t_setup = time.perf_counter() - t_setup
    parts = [m.instance_partition(i) for i in range(n)]
    from partconv_amd import autotune
    tuning = autotune.tune_engine(m)  # launch shapes for this device (cached), before the timed region
    gtop = m.info().num_levels - 1
    m.probe(gtop)  # delay-line partial-sum kernel of the largest level
    tps = 256  # one period of the 32768 level
    total = (args.warmup + args.steps) * tps
    xd = torch.randn((total * n0, n), dtype=torch.float32, device="cuda",
                     generator=torch.Generator(device="cuda").manual_seed(3))
    yd = torch.zeros_like(xd)
    torch.cuda.synchronize()
    stride = tps * n0 * n
    # warmup and timed steps: one persistent launch each, every tick released
    # and waited for individually
    if args.warmup and m.run_device(yd.data_ptr(), xd.data_ptr(), args.warmup * tps, True, persistent=True) != 0:
        raise RuntimeError(last_error())
    m.clear_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    off = args.warmup * stride
    if m.run_device(yd.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, args.steps * tps, True,
                    persistent=True) != 0:
        raise RuntimeError(last_error())
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    st = m.stats()
    avg_us, launches, bytes_per = m.probe_read()
    m.probe(-1)
    # fp64 check of two channels (shortest and longest IR) over the whole run
    x = xd.cpu().numpy()
    y = yd.cpu().numpy()
    from numpy.fft import irfft, rfft
    err = 0.0
    for c in (int(np.argmin(lens)), int(np.argmax(lens))):
        L = 1 << int(np.ceil(np.log2(x.shape[0] + irs[c].size)))
        r = irfft(rfft(x[:, c].astype(np.float64), L) * rfft(irs[c].astype(np.float64), L), L)[:x.shape[0]]
        lo = args.warmup * tps * n0
        err = max(err, float(np.max(np.abs(y[lo:, c] - r[lo:])) / np.sqrt(np.mean(r[lo:] ** 2))))
    # the same ticks with the level blocks launched as per-level HIP graphs, and
    # with one tick-kernel launch per tick (with and without the graphs)
    variants = {}
    vsteps = max(1, args.steps // 4)
    for name, persistent, graphs in (("persistent_level_graphs", True, True),
                                     ("launch_per_tick", False, False),
                                     ("launch_per_tick_level_graphs", False, True)):
        m.reset()
        m.set_level_graphs(graphs)
        if args.warmup and m.run_device(yd.data_ptr(), xd.data_ptr(), args.warmup * tps, True,
                                        persistent=persistent) != 0:
            raise RuntimeError(last_error())
        m.clear_stats()
        torch.cuda.synchronize()
        tv = time.perf_counter()
        if m.run_device(yd.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, vsteps * tps, True,
                        persistent=persistent) != 0:
            raise RuntimeError(last_error())
        torch.cuda.synchronize()
        vsecs = time.perf_counter() - tv
        vst = m.stats()
        variants[name] = {"value": n * vsteps * tps * n0 / vsecs, "p50_us": vst.p50_us, "p99_us": vst.p99_us,
                          "misses": int(vst.deadline_misses), "ticks": int(vst.ticks),
                          "persistent_ticks": int(vst.persistent_ticks)}
    m.set_level_graphs(False)
    m.cleanup()
    value = n * args.steps * tps * n0 / secs
    cpu = None
    if not args.no_cpu:
        # the fp32 oracle on a seeded random sample of 64 of the instances (mono
        # engines, the same IRs, partitions and input columns), over the host's cores
        from oracle import oracle as O
        pick = np.random.default_rng(33).choice(n, size=min(64, n), replace=False)
        cpu = cpu_reference_schedule([(insts[i]["N"], insts[i]["P"], irs[i]) for i in pick], None, len(pick), n0,
                                     args.cpu_budget, len(pick),
                                     f"{len(pick)} of the {n} uneven instances (seeded random pick)")
    print(json.dumps({
        "metric": "channel-samples/s at p99 block latency < block period (uneven batched instances)",
        "value": value, "unit": "channel-samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded white noise, seeded decaying-noise IRs)",
        "config": {"workload": f"{n} instances, IR lengths log-uniform in [2^12, 2^20], greedy partitions "
                               "over 128..32768, buffer 128 at 48 kHz, batched engine, persistent tick kernel",
                   "instances": n, "total_taps": int(np.sum(lens)), "setup_s": t_setup},
        "tuning": {k: tuning[k] for k in ("workers", "bins_per_thread", "release_window", "from_cache", "measured")
                   if k in tuning},
        "latency_us": {"p50": st.p50_us, "p99": st.p99_us, "max": st.max_us, "period": st.period_us,
                       "misses": int(st.deadline_misses), "ticks": int(st.ticks),
                       "persistent_ticks": int(st.persistent_ticks)},
        "p99_below_period": bool(st.p99_us < st.period_us),
        "max_rel_err_vs_fp64": err,
        "frame_variants": dict(variants, note="level blocks as per-level HIP graphs (arguments updated in "
                                              "place, one graph launch per level start) and one tick-kernel "
                                              "launch per tick, same engine and frames"),
        "roofline": workload_roofline(tick_bytes(parts, n0), args.steps * tps, secs, n * n0,
                                      "whole timed region: k_persist<7> resident for every tick (level 0 and, in "
                                      "self mode, the 256-1024 levels in each tick workgroup) + the longer "
                                      "levels' kernels beside it",
                                      "serial host round trip per tick and the tick workgroups' level-0 latency"),
        "kernel_rooflines": {"k_partial": with_traffic(kernel_roofline(
            avg_us, launches, bytes_per, f"k_partial (largest level, "
                                         f"{int(parts and max(max(N) for N, _ in parts))}, "
                                         "delay-line sum of every instance that has it)"),
            "k_partial_cfg3_512_traffic.json")},
        "cpu_baseline": cpu}), flush=True)


def bench_mimo(args):
    """Configuration 4: MIMO 32x32, 65536-tap IRs (RT60 1 s), partition
    [256x30, 4096x15], buffer 256 at 48 kHz (5.333 ms period).
    Real time (value): the real-time MIMO engine (shared input transforms and
    delay lines, per-bin matrix-vector products).  Also 1024 pair instances
    (the reference's own formulation) on the batched engine, persistent and
    with one launch per tick.  Throughput: the same filters over the whole
    input at once on the fp32 matrix cores."""
    import torch
    from partconv_amd import PartConvMulti, last_error
    from partconv_amd.mimo import MimoConv, MimoRealtime
    from partconv_amd.workloads import cfg4_irs
    n, n0, taps = 32, 256, 65536
    irs = cfg4_irs(seed=4, n_in=n, n_out=n, taps=taps)
    insts = [dict(N=[256, 4096], P=[30, 0], ir=irs[o][i], in_ch=i, out_ch=o) for o in range(n) for i in range(n)]
    m = PartConvMulti(0)
    if m.setup_mem(FS, n0, n, n, insts) != 0:
        raise RuntimeError(last_error())
    parts = [m.instance_partition(i) for i in range(len(insts))]
    m.probe(1)  # delay-line partial-sum kernel of the 4096 level
    tps = 16
    total = (args.warmup + args.steps) * tps
    g = torch.Generator(device="cuda").manual_seed(4)
    xd = torch.randn((total * n0, n), dtype=torch.float32, device="cuda", generator=g)
    yd = torch.zeros_like(xd)
    torch.cuda.synchronize()
    stride = tps * n0 * n
    # warmup and timed steps: one persistent launch each, every tick released
    # and waited for individually
    if args.warmup and m.run_device(yd.data_ptr(), xd.data_ptr(), args.warmup * tps, True, persistent=True) != 0:
        raise RuntimeError(last_error())
    m.clear_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    off = args.warmup * stride
    if m.run_device(yd.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, args.steps * tps, True,
                    persistent=True) != 0:
        raise RuntimeError(last_error())
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    st = m.stats()
    avg_us, launches, bytes_per = m.probe_read()
    m.probe(-1)
    # the same ticks with one tick-kernel launch per tick, for comparison
    m.reset()
    m.clear_stats()
    yl = torch.zeros_like(xd)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if m.run_device(yl.data_ptr() + 4 * off, xd.data_ptr() + 4 * off, args.steps * tps, True) != 0:
        raise RuntimeError(last_error())
    torch.cuda.synchronize()
    lsecs = time.perf_counter() - t1
    lst = m.stats()
    m.cleanup()
    # the real-time MIMO engine on the same frames: warmup and timed steps one call
    # each, with the release window the tuner chose (before the timed region);
    # then one tick at a time
    from partconv_amd import autotune
    h = torch.from_numpy(np.stack([np.stack(irs[o]) for o in range(n)]).astype(np.float32)).cuda()
    rt = MimoRealtime(FS, n0, [256, 4096], [30, 0], taps, n, n)
    rt.set_filters(h)
    mtune = autotune.tune_mimo(rt)
    yr = torch.zeros_like(xd)
    roff = args.warmup * tps * n0  # rows (frames x buffer size) of the warmup ticks
    if args.warmup:
        rt.run_device(yr[:roff], xd[:roff], args.warmup * tps)
    rt.clear_stats()
    rt.probe(1)  # delay-line sum of the 4096 level
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    rt.run_device(yr[roff:], xd[roff:], args.steps * tps)
    torch.cuda.synchronize()
    rsecs = time.perf_counter() - t2
    rst = rt.stats()
    r_avg_us, r_launches, r_bytes = rt.probe_read()
    rt.probe(-1)
    rt1 = MimoRealtime(FS, n0, [256, 4096], [30, 0], taps, n, n)
    rt1.set_filters(h)
    y1 = torch.zeros_like(xd)
    if args.warmup:
        rt1.run_device(y1[:roff], xd[:roff], args.warmup * tps)
    rt1.clear_stats()
    torch.cuda.synchronize()
    t21 = time.perf_counter()
    rt1.run_device(y1[roff:], xd[roff:], args.steps * tps)
    torch.cuda.synchronize()
    r1secs = time.perf_counter() - t21
    r1st = rt1.stats()
    one_at_a_time = {"value": n * n0 * args.steps * tps / r1secs, "p50_us": r1st.p50_us, "p99_us": r1st.p99_us,
                     "misses": int(r1st.deadline_misses), "bitwise_equal_to_windowed": bool(torch.equal(y1, yr))}
    del rt1, y1
    # host buffers, one call per tick (frames cross the bus; one launch per call)
    xh = xd[:tps * n0].cpu().numpy()
    oh = np.zeros((n0, n), dtype=np.float32)
    hcalls = 4 * tps
    hlat = []
    rt.reset()
    t3 = time.perf_counter()
    for c in range(hcalls):
        tc = time.perf_counter()
        rt.run(oh, np.ascontiguousarray(xh[(c % tps) * n0:(c % tps + 1) * n0]))
        hlat.append(1e6 * (time.perf_counter() - tc))
    hsecs = time.perf_counter() - t3
    host_rt = {"value": n * n0 * hcalls / hsecs, "p50_us": float(np.percentile(hlat, 50)),
               "p99_us": float(np.percentile(hlat, 99)), "calls": hcalls,
               "note": "pcxr_run per tick: frame into pinned memory, one persistent launch of one tick, "
                       "output from pinned memory"}
    del rt
    # throughput form on the matrix cores: same filters, same input
    mx = MimoConv([256, 4096], [30, 0], taps, n, n)
    mx.set_filters(h)
    xt = xd.t().contiguous()
    ymx = mx(xt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    reps = 3
    for _ in range(reps):
        ymx = mx(xt)
    e1.record()
    torch.cuda.synchronize()
    tput_s = e0.elapsed_time(e1) / 1e3 / reps
    T = xt.shape[1]
    flops = 0.0
    for Nl, Pl in ((256, 30), (4096, 15)):
        M = -(-T // Nl)
        flops += Nl * Pl * M * 8.0 * n * n
    # fp64 check of one output over the run: real-time and throughput forms
    x = xd.cpu().numpy().astype(np.float64)
    o = 7
    from numpy.fft import irfft, rfft
    L = 1 << int(np.ceil(np.log2(x.shape[0] + taps)))
    r = np.zeros(x.shape[0])
    for i in range(n):
        r += irfft(rfft(x[:, i], L) * rfft(irs[o][i].astype(np.float64), L), L)[:x.shape[0]]
    y = yd.cpu().numpy()
    lo = args.warmup * tps * n0
    err_pairs = float(np.max(np.abs(y[lo:, o] - r[lo:])) / np.sqrt(np.mean(r[lo:] ** 2)))
    y_rt = yr.cpu().numpy()
    err_rt = float(np.max(np.abs(y_rt[:, o] - r)) / np.sqrt(np.mean(r ** 2)))
    err_mx = float(np.max(np.abs(ymx[o].cpu().numpy() - r)) / np.sqrt(np.mean(r ** 2)))
    cpu = None
    if not args.no_cpu:
        # the fp32 oracle as the reference formulates MIMO (pair instances summed
        # per output): 16 of the 32 outputs, each an engine of its 32 pairs
        no = 8
        specs = [([256, 4096], [30, 0], irs[oo][i]) for oo in range(no) for i in range(n)]
        cpu = cpu_reference_schedule(specs, [i for oo in range(no) for i in range(n)], n, n0, args.cpu_budget, no,
                                     f"{no} of the {n} outputs as the reference formulates MIMO: {no * n} pair "
                                     f"instances (input i -> output o), outputs per instance (the per-output sum "
                                     f"of {n} frames is not timed)")
    print(json.dumps({
        "metric": "channel-samples/s at p99 block latency < block period (MIMO 32x32)",
        "value": n * args.steps * tps * n0 / rsecs, "unit": "channel-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * rsecs / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded white noise, seeded decaying-noise IRs)",
        "config": {"workload": "MIMO 32x32, 65536-tap IRs, [256x30, 4096x15], buffer 256 at 48 kHz; real-time "
                               "MIMO engine (shared input transforms and delay lines, per-bin matrix-vector "
                               "products; level 0 in one persistent kernel)",
                   "release_window": mtune["release_window"]},
        "tuning": {k: mtune[k] for k in ("release_window", "from_cache", "measured") if k in mtune},
        "latency_us": {"p50": rst.p50_us, "p99": rst.p99_us, "max": rst.max_us, "period": rst.period_us,
                       "misses": int(rst.deadline_misses), "ticks": int(rst.ticks)},
        "one_tick_released_at_a_time": one_at_a_time,
        "p99_below_period": bool(rst.p99_us < rst.period_us),
        "max_rel_err_vs_fp64": err_rt,
        "host_buffers_pcie_inclusive": host_rt,
        "roofline": workload_roofline(mimo_engine_tick_bytes([(256, 30), (4096, 15)], n, n, n0), args.steps * tps,
                                      rsecs, n * n0,
                                      "whole timed region of the real-time MIMO engine (its own bytes): "
                                      "k_mxr_persist<8> resident for every tick + the 4096 level's kernels",
                                      "level 0's per-tick critical path (host round trip, input transforms, "
                                      "counter hand-offs) beside the filter-spectrum stream"),
        "kernel_rooflines": {"k_mxr_part": with_traffic(kernel_roofline(
            r_avg_us, r_launches, r_bytes, "k_mxr_part (4096 level: the next block's delay-line sum, 32x32 pairs x "
                                           "14 partitions, one workgroup per bin), beside the persistent kernel"),
            "k_mxr_part_cfg4_traffic.json")},
        "pair_instances": {"value": n * args.steps * tps * n0 / secs, "p50_us": st.p50_us, "p99_us": st.p99_us,
                           "misses": int(st.deadline_misses), "persistent_ticks": int(st.persistent_ticks),
                           "max_rel_err_vs_fp64": err_pairs,
                           "roofline": kernel_roofline(avg_us, launches, bytes_per,
                                                       "k_partial (4096 level, delay-line sum of the 1024 pair "
                                                       "instances)")},
        "pair_instances_launch_per_tick": {"value": n * args.steps * tps * n0 / lsecs, "p50_us": lst.p50_us,
                                           "p99_us": lst.p99_us, "misses": int(lst.deadline_misses)},
        "cpu_baseline": cpu,
        "throughput_form": {"output_channel_samples_per_s": n * T / tput_s, "seconds_per_call": tput_s,
                            "samples": T, "matrix_core_tflops_effective": flops / tput_s / 1e12,
                            "mfma_peak_tflops_fp32": 157.3, "max_rel_err_vs_fp64": err_mx}}), flush=True)


def bench_diff(args):
    """Configuration 5: 64 clips x 131072 samples, learnable 65536-tap IR
    (RT60 0.8 s), partition [256x14, 2048x14, 16384x3], MSE loss against targets
    from a hidden IR (RT60 1.5 s) + N(0, (1e-3)^2); forward + hand-written
    adjoint per step.  Also the 2^20-tap IR in recompute-spectra mode."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    from partconv_amd.workloads import decaying_noise_ir
    C, T, taps = 64, 131072, 65536
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn((C, T), dtype=torch.float32, device="cuda", generator=g)
    h0 = torch.from_numpy(decaying_noise_ir(taps, 0.8, FS, 50)).cuda().requires_grad_(True)
    hs = torch.from_numpy(decaying_noise_ir(taps, 1.5, FS, 51)).cuda()
    conv = PartitionedConv([256, 2048, 16384], [14, 14, 0], taps)
    with torch.no_grad():
        tgt = conv(x, hs) + 1e-3 * torch.randn((C, T), dtype=torch.float32, device="cuda", generator=g)

    # chunk sizes of the spectral products and dL/dh chosen by timing on these
    # clips (or replayed from the tuning cache)
    from partconv_amd import autotune
    tinfo = autotune.tune_offline(conv, x, h0.detach())
    tuning = tinfo["levels"]

    def step():
        # the fused loss: pcc_mse forms the loss and its gradient in one pass
        # over y and the target, pcc_backward_scaled applies the upstream factor
        h0.grad = None
        loss = conv.mse_loss(x, h0, tgt)
        loss.backward()
        return loss

    def step_torch_loss():
        h0.grad = None
        loss = torch.nn.functional.mse_loss(conv(x, h0), tgt)
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if args.profile_step:
        # one step between two marker copies (4 MiB device copies: a copy
        # kernel of a grid no other launch of the step uses)
        msrc = torch.ones(1 << 20, dtype=torch.float32, device="cuda")
        mdst = torch.empty_like(msrc)
        mdst.copy_(msrc)
        step()
        mdst.copy_(msrc)
        torch.cuda.synchronize()
        print(json.dumps({"profile_step": "cfg5", "marker_bytes": 4 << 20, "tuning": tuning}), flush=True)
        return
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    cgraphs = conv.call_graph_stats()
    # each kernel of each level, timed by HIP events its own launch stamps,
    # over a few more steps, with the levels on one stream (each kernel alone
    # on the device)
    probes = {}
    conv.set_level_streams("off")
    for li, nl in enumerate(conv.N):
        for kind in ("forward_transform", "product", "inverse_transform", "dh_partial"):
            conv.probe(li, kind)
            for _ in range(4):
                step()
            torch.cuda.synchronize()
            probes[(kind, nl)] = conv.probe_read()
            conv.probe(-1)
    # the same step with torch's mse_loss over the convolution's output
    for _ in range(2):
        step_torch_loss()
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    for _ in range(args.steps):
        step_torch_loss()
    torch.cuda.synchronize()
    secs_tl = time.perf_counter() - t5
    # without the library's call graphs: every launch eager, the levels on one
    # stream, then each level on its own stream
    conv.set_call_graphs(False)
    step()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    secs_eager = time.perf_counter() - t3
    conv.set_level_streams("on")
    step()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    secs_par = time.perf_counter() - t2
    conv.set_level_streams("captured")
    conv.set_call_graphs(True)
    # the same step (fused loss) captured once in a HIP graph and replayed
    def gstep():
        conv.mse_loss(x, h0, tgt).backward()

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            h0.grad = None
            gstep()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    h0.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gstep()
    graph.replay()
    torch.cuda.synchronize()
    g_graph = h0.grad.clone()
    h0.grad = None
    gstep()  # eager, same formulation
    torch.cuda.synchronize()
    graph_ok = bool(torch.equal(g_graph, h0.grad))
    t4 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    gsecs = time.perf_counter() - t4
    del graph
    # long IR, spectra recomputed in backward
    taps2 = 1 << 20
    conv2 = PartitionedConv([256, 2048, 16384], [14, 14, 0], taps2, budget_bytes=0)
    h2 = torch.from_numpy(decaying_noise_ir(taps2, 3.0, FS, 52)).cuda().requires_grad_(True)
    for _ in range(2):
        h2.grad = None
        torch.mean(conv2(x, h2) ** 2).backward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        h2.grad = None
        torch.mean(conv2(x, h2) ** 2).backward()
    torch.cuda.synchronize()
    secs2 = time.perf_counter() - t1
    # fp64 check of one clip's forward
    from numpy.fft import irfft, rfft
    xc = x[3].cpu().numpy().astype(np.float64)
    hc = h0.detach().cpu().numpy().astype(np.float64)
    L = 1 << int(np.ceil(np.log2(T + taps)))
    r = irfft(rfft(xc, L) * rfft(hc, L), L)[:T]
    with torch.no_grad():
        yc = conv(x, h0)[3].cpu().numpy()
    err = float(np.max(np.abs(yc - r)) / np.sqrt(np.mean(r ** 2)))
    # roofline of the timed region: the algorithmic bytes of every kernel of
    # the step (each input read once, each output written once) over the
    # measured step time; per-kernel figures from the probes beside it
    step_b, step_parts = diff_step_bytes(C, T, conv.N, conv.P, [t[1] for t in tuning])
    gbps = step_b * args.steps / secs / 1e9
    roof = {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
            "traffic": None, "kernel": "whole timed step: forward (3 levels) + MSE loss + dL/dh, every kernel's "
                                    "algorithmic bytes",
            "alg_bytes_per_step": step_b, "alg_bytes_by_kernel": step_parts,
            "traffic_note": "per-kernel HBM traffic is in kernel_rooflines (counter runs)"}
    step_traffic = os.path.join(ROOT, "profiles", "kernels_cfg5_step_traffic.json")
    tr = None
    if os.path.exists(step_traffic):
        with open(step_traffic) as fh:
            tr = json.load(fh)
        src = ("profiles/kernels_cfg5_step_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, "
               "separate profiled run" + (f" at commit {tr['commit']}" if tr.get("commit") else ""))
        roof["traffic"] = tr["bytes_per_step"]
        roof["traffic_source"] = src + " (every kernel of one step)"
    kroofs = {}
    names = {"forward_transform": "k_off_rfft_reg (input windows -> X)", "product": "k_off_mac_w (Y = sum X H)",
             "inverse_transform": "k_off_irfft_reg (Y -> y)", "dh_partial": "k_off_dh_part (dL/dh partial sums)"}
    for (kind, nl), (avg_us, launches, nbytes) in probes.items():
        key = f"{kind}@{nl}"
        kroofs[key] = kernel_roofline(avg_us, launches, nbytes, f"{names[kind]}, {nl} level, 64 clips x 131072")
        if kroofs[key] and tr and key in tr.get("by_kernel", {}):
            kroofs[key]["traffic"] = tr["by_kernel"][key]
            kroofs[key]["traffic_source"] = src
    # the reference's canonical per-block bytes (8(N+1)(2P+1)+12N per block) describe
    # the real-time formulation, which streams H and the delay line per block;
    # this form shares H across the clips and reads each spectrum once, so the
    # canonical figure is not a traffic count here
    Ns, Ps = conv.N, conv.P
    canon = sum(block_bytes(nl, pl) / nl for nl, pl in zip(Ns, Ps))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_diff(x[:8].cpu().numpy(), h0.detach().cpu().numpy(), tgt[:8].cpu().numpy(),
                                args.cpu_budget)
    print(json.dumps({
        "metric": "clip-samples/s, forward + adjoint (differentiable partitioned convolution)",
        "value": C * T * args.steps / secs, "unit": "clip-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * secs / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded white noise, seeded decaying-noise IRs)",
        "config": {"workload": "64 clips x 131072, 65536-tap learnable IR, [256x14, 2048x14, 16384x3], MSE"},
        "max_rel_err_vs_fp64": err,
        "long_ir_recompute": {"taps": taps2, "partition": "[256x14, 2048x14, 16384x63]",
                              "ms_per_step": 1e3 * secs2 / args.steps,
                              "recomputed_levels": int(conv2.recomputed_levels)},
        "call_graphs": {"forward_replays_captures": list(cgraphs["forward"]),
                        "backward_replays_captures": list(cgraphs["backward"]),
                        "note": "value: the library's default, each pcc_forward / pcc_backward call replayed "
                                "from a HIP graph of its own launches (levels as parallel branches) once its "
                                "arguments repeat; counts over warm-up + timed steps"},
        "torch_mse_loss": {"ms_per_step": 1e3 * secs_tl / args.steps,
                           "note": "the same step with torch.nn.functional.mse_loss over the convolution's output "
                                   "instead of the fused loss (pcc_mse)"},
        "eager_no_call_graphs": {"ms_per_step": 1e3 * secs_eager / args.steps,
                                 "note": "the same step with call graphs off: every launch eager, every level "
                                         "on the caller's stream; bitwise the same results"},
        "eager_level_streams": {"ms_per_step": 1e3 * secs_par / args.steps,
                                "note": "call graphs off, each level on its own stream; bitwise the same "
                                        "results"},
        "tuned_chunks": {"levels": list(conv.N), "mac_dh_blocks_per_chunk": tuning,
                         "from_cache": tinfo["from_cache"]},
        "hip_graph_replay": {"ms_per_step": 1e3 * gsecs / args.steps, "value": C * T * args.steps / gsecs,
                             "grad_equals_eager": graph_ok,
                             "note": "forward + fused loss + adjoint captured once (torch.cuda.graph) and replayed, "
                                     "the levels as parallel branches"},
        "roofline": roof, "kernel_rooflines": kroofs,
        "canonical_bytes_per_clip_sample": {"value": canon, "use": "reference's per-block formula for one forward "
                                            "pass; not a traffic count for this form (H shared by the clips, each "
                                            "spectrum read once)"},
        "cpu_baseline": cpu}), flush=True)


def cpu_baseline_diff(x, h, tgt, budget_s):
    """The same step on the host: forward y = x * h (causal, truncated), MSE
    loss gradient g, dL/dx = correlation of g with h, dL/dh = sum over clips of
    the correlation of g with x; whole-signal fp32 FFTs (scipy.fft, all host
    cores).  No reference counterpart exists (the reference has no adjoint);
    this restates the mathematics on the CPU as a baseline."""
    import scipy.fft as sf
    cores = len(os.sched_getaffinity(0))
    C, T = x.shape
    taps = h.size
    L = sf.next_fast_len(T + taps)

    def step():
        X = sf.rfft(x, L, axis=1, workers=cores)
        H = sf.rfft(h, L, workers=cores)
        y = sf.irfft(X * H, L, axis=1, workers=cores)[:, :T]
        g = (2.0 / (C * T)) * (y - tgt)
        G = sf.rfft(g.astype(np.float32), L, axis=1, workers=cores)
        dx = sf.irfft(G * np.conj(H), L, axis=1, workers=cores)[:, :T]
        dh = sf.irfft(np.sum(G * np.conj(X), axis=0), L, workers=cores)[:taps]
        return dx, dh

    step()
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < budget_s or reps < 2:
        step()
        reps += 1
    t = time.perf_counter() - t0
    return {"value": C * T * reps / t, "unit": "clip-samples/s", "cores": cores, "kind": "port",
            "sample": f"{reps} steps over {C} of the 64 clips: forward + dL/dx + dL/dh as whole-signal fp32 "
                      f"FFTs (scipy.fft, {cores} workers)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 30 s of input per channel for cfg2, the configuration's "
                         "length; 40 for the others)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--workload", choices=["cfg2", "cfg3", "cfg4", "cfg5"], default="cfg2",
                    help="cfg2: 16-channel drop-in (default, headline); cfg3: uneven batched instances; "
                         "cfg4: MIMO 32x32; cfg5: differentiable forward + adjoint")
    ap.add_argument("--instances", type=int, default=512, help="cfg3 instance count")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 64 MiB device copy first (HBM-counter calibration under a profiler)")
    ap.add_argument("--profile-step", action="store_true",
                    help="cfg5: after set-up and warm-up run one step between two 4 MiB device copies (markers "
                         "for per-step counter sums, prof_summary step) and stop")
    args = ap.parse_args()
    if args.steps is None:
        # cfg2 states 30 s of input per channel at 48 kHz: 88 steps of 256 ticks x 64 samples
        args.steps = -(-30 * FS // (TICKS_PER_STEP * N0)) if args.workload == "cfg2" else 40
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: each GPU runs its own share of channels independently "
              "(no collective on this path); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy:
        # a 64 MiB device copy first: the counter runs scale FETCH_SIZE / WRITE_SIZE by it
        import torch
        src = torch.ones(16 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "cfg3":
        bench_uneven(args)
        return
    if args.workload == "cfg4":
        bench_mimo(args)
        return
    if args.workload == "cfg5":
        bench_diff(args)
        return
    import torch
    from partconv_amd.workloads import cfg2_irs
    irs = cfg2_irs(seed=2, channels=CHANNELS, taps=TAPS)
    total_steps = args.warmup + max(args.steps, args.plain_steps)
    n = total_steps * TICKS_PER_STEP * N0
    rng = np.random.default_rng(2 + 77)
    x_host = rng.standard_normal((n, CHANNELS)).astype(np.float32)
    xd = torch.from_numpy(x_host).cuda()
    yd = torch.zeros_like(xd)
    torch.cuda.synchronize()

    # ---- batched engine, persistent tick kernel (headline)
    mb = build_engine("batched", irs)
    parts = [mb.instance_partition(i) for i in range(CHANNELS)]
    from partconv_amd import autotune
    tuning = autotune.tune_engine(mb)  # launch shapes for this device (cached), before the timed region
    mb.probe(3)  # delay-line partial-sum kernel of the 16384 level
    secs, st = timed_run(mb, xd, yd, args.warmup, args.steps, torch, persistent=True)
    release = "device memory via the PCIe BAR" if mb.release_in_device_memory else "pinned host memory"
    avg_us, launches, bytes_per = mb.probe_read()
    mb.probe(-1)
    ticks = args.steps * TICKS_PER_STEP
    value = CHANNELS * ticks * N0 / secs
    # parity spot check of the timed output against fp64 direct convolution
    y = yd.cpu().numpy()
    c = 5
    hi = (args.warmup + args.steps) * TICKS_PER_STEP * N0
    from numpy.fft import irfft, rfft
    L = 1 << int(np.ceil(np.log2(hi + TAPS)))
    r = irfft(rfft(x_host[:hi, c].astype(np.float64), L) * rfft(irs[c].astype(np.float64), L), L)[:hi]
    lo = args.warmup * TICKS_PER_STEP * N0
    err = float(np.max(np.abs(y[lo:hi, c] - r[lo:hi])) / np.sqrt(np.mean(r[lo:hi] ** 2)))

    # the same persistent kernel with one tick released at a time (no release window)
    window = int(tuning.get("release_window", 1))
    serial = None
    if window > 1:
        mb.reset()
        mb.set_release_window(1)
        ssteps = max(1, args.steps // 4)
        ssecs, sst = timed_run(mb, xd, yd, 1, ssteps, torch, persistent=True)
        serial = {"value": CHANNELS * ssteps * TICKS_PER_STEP * N0 / ssecs, "p50_us": sst.p50_us,
                  "p99_us": sst.p99_us, "misses": int(sst.deadline_misses), "steps": ssteps}
        mb.set_release_window(window)
    # batched engine with one tick-kernel launch per tick (same outputs)
    mb.reset()
    lsecs, lst = timed_run(mb, xd, yd, 1, max(1, args.steps // 4), torch, persistent=False)
    lval = CHANNELS * max(1, args.steps // 4) * TICKS_PER_STEP * N0 / lsecs
    # tick-kernel probe (separate short pass, not part of the timed region)
    mb.reset()
    mb.probe(0)
    timed_run(mb, xd, yd, 0, 2, torch)
    tick_us, _, tick_b = mb.probe_read()
    mb.probe(-1)
    mb.cleanup()

    out = {
        "metric": "channel-samples/s at p99 block latency < block period; batched vs plain, same GPU",
        "value": value,
        "unit": "channel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded white noise, seeded decaying-noise IRs)",
        "config": {"workload": "16 channels x 262144-tap IRs, 48 kHz, partition [64x6,256x14,2048x14,16384x15], "
                               "buffer 64, inputs resident in HBM, batched engine, persistent tick kernel",
                   "channels": CHANNELS, "taps": TAPS, "buffer_size": N0,
                   "ticks_per_step": TICKS_PER_STEP, "release_word": release, "release_window": window},
        "tuning": {k: tuning[k] for k in ("workers", "bins_per_thread", "release_window", "from_cache", "measured")
                   if k in tuning},
        "latency_us": {"p50": st.p50_us, "p99": st.p99_us, "max": st.max_us, "period": st.period_us,
                       "misses": int(st.deadline_misses), "ticks": int(st.ticks),
                       "persistent_ticks": int(st.persistent_ticks)},
        "persistent_one_tick_released_at_a_time": serial,
        "batched_launch_per_tick": {"value": lval, "p50_us": lst.p50_us, "p99_us": lst.p99_us,
                                    "misses": int(lst.deadline_misses)},
        "p99_below_period": bool(st.p99_us < st.period_us),
        "max_rel_err_vs_fp64": err,
        "roofline": workload_roofline(tick_bytes(parts, N0), ticks, secs, CHANNELS * N0,
                                      "whole timed region: k_persist<6> resident for every tick (level 0 and the "
                                      "256 level inside it) + the 2048 and 16384 levels' kernels beside it",
                                      "per-tick latency chain, not HBM: one tick moves about 0.9 MB; with the "
                                      "release window the host round trip overlaps the device's ticks and each "
                                      "tick workgroup's own chain (frame and level loads, the 64-point transforms, "
                                      "the handshake with the 256 level's worker every 4 ticks) bounds the rate"),
        "kernel_rooflines": {
            "k_partial": with_traffic(kernel_roofline(avg_us, launches, bytes_per,
                                                      "k_partial (16384-level delay-line sum over 14 partitions, "
                                                      "16 jobs)"), "k_partial_cfg2_16ch_traffic.json"),
            "tick_kernel": kernel_roofline(tick_us, 2, tick_b, "k_tick<6> (one launch per tick, level 0 of 16 "
                                                               "instances; the persistent kernel's tick work)")},
    }

    # ---- host buffers (PCIe-inclusive: frames and output in pinned host memory),
    # batched engine through run(): resident persistent session vs one launch per tick
    hticks = 4 * TICKS_PER_STEP
    xh = np.ascontiguousarray(x_host[:hticks * N0])
    yh = np.zeros_like(xh)
    host = {}
    for name, sess in (("persistent_session", True), ("launch_per_tick", False)):
        mh = build_engine("batched", irs)
        mh.set_host_session(sess)
        mh.process(xh[:TICKS_PER_STEP * N0])      # warmup
        mh.clear_stats()
        t0 = time.perf_counter()
        if _lib_process(mh, yh, xh, hticks) != 0:
            raise RuntimeError("host run failed")
        hs = time.perf_counter() - t0
        st_h = mh.stats()
        host[name] = {"value": CHANNELS * hticks * N0 / hs, "p50_us": st_h.p50_us, "p99_us": st_h.p99_us,
                      "misses": int(st_h.deadline_misses), "ticks": hticks}
        mh.cleanup()
    out["host_buffers_pcie_inclusive"] = host

    # ---- plain engine (same GPU, per-instance launches)
    if not args.no_plain:
        yd.zero_()
        mp = build_engine("plain", irs)
        psecs, pst = timed_run(mp, xd, yd, 1, args.plain_steps, torch)
        mp.cleanup()
        pval = CHANNELS * args.plain_steps * TICKS_PER_STEP * N0 / psecs
        out["plain"] = {"value": pval, "p50_us": pst.p50_us, "p99_us": pst.p99_us,
                        "misses": int(pst.deadline_misses), "steps": args.plain_steps}
        out["batched_over_plain"] = value / pval

    if not args.no_cpu:
        cpu = cpu_reference_schedule([(PART_N, PART_P, irs[c]) for c in range(CHANNELS)], None, CHANNELS, N0,
                                     args.cpu_budget, CHANNELS, "the same 16-channel workload (same IRs and "
                                                                "partition), strict run()")
        cpu["free_running"] = cpu_free_running(irs, x_host, min(3.0, args.cpu_budget))
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
