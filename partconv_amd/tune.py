"""Partition autotuner: choose the multi-level partition of an IR from costs
measured on this GPU instead of a fixed rule.

The reference tunes partitions by benchmarking candidate levels with
PartConvFilter::bench and comparing the summed load (src/tune.py:20-154), and
its prototype finds the optimum with a Viterbi search over block positions
under an analytical cost model (matlab/optimal_part.m:1-125,
matlab/part_conv_work.m:8-13).  Here both halves are joined: per-level block
costs are measured on the GPU (PartConvFilter.bench, HIP events) and fitted as

    cost(N, P) = f(N) + m(N) * P        seconds per block of size N,

and an exact dynamic program over covered taps finds the partition with the
least load,

    load = sum_L cost(N_L, P_L) * FS / N_L        (fraction of real time),

subject to the partition rules of src/partconv.cpp:86-182: block sizes are
powers of two, N_0 = buffer size, strictly increasing; a level L+1 may start
only when 2 N_0 + T_L >= 2 N_{L+1} (T_L = taps covered before it); the last
level covers the rest of the IR.

    from partconv_amd import tune
    model = tune.measure_model(64, max_block=16384)      # GPU
    N, P, load = tune.optimal_partition(64, 262144, model)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


@dataclass
class CostModel:
    """Per-block cost of a level of block size N with P partitions:
    f[N] + m[N] * P seconds.  `samples` keeps the measurements the fit came from."""
    N0: int
    FS: int = 48000
    f: dict = field(default_factory=dict)
    m: dict = field(default_factory=dict)
    samples: dict = field(default_factory=dict)

    def cost(self, N, P):
        return self.f[N] + self.m[N] * P

    def sizes(self):
        return sorted(self.f)


def _bench_level(N0, N, P, FS, trials, max_time):
    """Seconds per block of a level (N, P), measured on the GPU: a filter whose
    level 1 is (N, P) behind the smallest level 0 that allows it (or level 0
    itself when N == N0)."""
    from .partconv import PartConvFilter, intArray
    if N == N0:
        Ns, Ps, L = [N0], [P], 0
    else:
        k = N // N0
        Ns, Ps, L = [N0, N], [2 * k - 2, P], 1
    taps = sum(n * p for n, p in zip(Ns, Ps))
    import numpy as np
    h = (np.random.default_rng(N + P).standard_normal(taps) * 1e-3).astype(np.float32)
    pc = PartConvFilter(0, FS)
    a, b = intArray(len(Ns)), intArray(len(Ns))
    for i, (n, p) in enumerate(zip(Ns, Ps)):
        a[i], b[i] = n, p
    if pc.setup(len(Ns), a, b, h) != 0:
        raise RuntimeError(f"autotune: setup of {Ns} x {Ps} failed")
    pc.bench(L, 5, 0, 0.01)                    # warm
    t = pc.bench(L, trials, 0, max_time)
    pc.cleanup()
    return t


def measure_model(N0, max_block=32768, FS=48000, P_hi=8, trials=200, max_time=0.05):
    """Measure f(N), m(N) for every power-of-two block size N0..max_block on the
    GPU (two partition counts per size: 1 and P_hi) and return the CostModel."""
    model = CostModel(N0=N0, FS=FS)
    N = N0
    while N <= max_block:
        t1 = _bench_level(N0, N, 1, FS, trials, max_time)
        t2 = _bench_level(N0, N, P_hi, FS, trials, max_time)
        slope = max(0.0, (t2 - t1) / (P_hi - 1))
        model.m[N] = slope
        model.f[N] = max(0.0, t1 - slope)
        model.samples[N] = (t1, t2)
        N *= 2
    return model


def load(N, P, model):
    """Load of a partition: sum_L cost(N_L, P_L) * FS / N_L (1.0 = real time)."""
    return sum(model.cost(n, p) * model.FS / n for n, p in zip(N, P))


def optimal_partition(N0, taps, model, max_block=None):
    """Exact minimum-load partition for an IR of `taps` taps under `model`.

    Dynamic program over (covered length t in units of N0, current block size
    2^b N0): adding a partition costs m(N)/N per sample-of-time, starting a
    level costs f(N)/N; a level 2^b' N0 may start at t only if
    2 + t >= 2^(b'+1) and the IR is not yet covered.  Returns (N, P, load)."""
    if taps < 1:
        raise ValueError("taps must be positive")
    sizes = [n for n in model.sizes() if max_block is None or n <= max_block]
    if not sizes or sizes[0] != N0:
        raise ValueError("the model must include the buffer size")
    B = len(sizes)
    T = -(-taps // N0)                      # units of N0 to cover
    span = [n // N0 for n in sizes]         # block length in units
    per_part = [model.m[n] * model.FS / n for n in sizes]
    per_start = [model.f[n] * model.FS / n for n in sizes]
    tmax = T + span[-1]
    INF = math.inf
    dist = [[INF] * B for _ in range(tmax + 1)]
    back = [[None] * B for _ in range(tmax + 1)]
    dist[span[0]][0] = per_start[0] + per_part[0]
    for t in range(tmax + 1):
        for b in range(B):
            d = dist[t][b]
            if d == INF or t >= T:
                continue
            # one more partition of the current size
            t2 = t + span[b]
            if t2 <= tmax and d + per_part[b] < dist[t2][b]:
                dist[t2][b] = d + per_part[b]
                back[t2][b] = (t, b)
            # start a larger level (delay rule, src/partconv.cpp:148-166)
            for b2 in range(b + 1, B):
                if 2 + t < 2 * span[b2]:
                    break
                t2 = t + span[b2]
                if t2 <= tmax:
                    c = d + per_start[b2] + per_part[b2]
                    if c < dist[t2][b2]:
                        dist[t2][b2] = c
                        back[t2][b2] = (t, b)
    best, arg = INF, None
    for t in range(T, tmax + 1):
        for b in range(B):
            if dist[t][b] < best:
                best, arg = dist[t][b], (t, b)
    if arg is None:
        raise ValueError("no partition covers the IR within the block sizes given")
    # backtrack into per-level counts
    levels = []
    t, b = arg
    while True:
        levels.append(b)
        prev = back[t][b]
        if prev is None:
            break
        t, b = prev
    levels.reverse()
    N, P = [], []
    for b in levels:
        if N and N[-1] == sizes[b]:
            P[-1] += 1
        else:
            N.append(sizes[b])
            P.append(1)
    return N, P, best


def tune(N0, taps, FS=48000, max_block=32768, **measure):
    """Measure the cost model on this GPU and return (N, P, load, model)."""
    model = measure_model(N0, max_block=max_block, FS=FS, **measure)
    N, P, ld = optimal_partition(N0, taps, model)
    return N, P, ld, model
