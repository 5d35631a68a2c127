"""The partition autotuner's search (partconv_amd/tune.py) on the CPU: the
dynamic program against brute-force enumeration of every valid partition
under synthetic cost models, validity under the partition rules
(src/partconv.cpp:86-182, restated in oracle/rules.py), and no worse than the
greedy rule.  The GPU measurement half is in tests/test_gpu_tune.py."""
import itertools
import math

import pytest

from oracle import rules as R
from partconv_amd import tune


def model(N0, max_block, f, m):
    M = tune.CostModel(N0=N0, FS=48000)
    N = N0
    while N <= max_block:
        M.f[N] = f(N)
        M.m[N] = m(N)
        N *= 2
    return M


def all_partitions(N0, taps, sizes):
    """Every partition the rules allow (strictly increasing power-of-two sizes
    from sizes, counts >= 1, delay rule, last level recounted, last level
    required)."""
    out = []
    T = taps

    def rec(N, P, covered):
        if covered >= T:
            return
        last = N[-1]
        # close with the current level as the last one
        extra = -(-(T - covered) // last)
        out.append((N[:-1] + [last], P[:-1] + [P[-1] + extra]))
        # or add partitions to the current level, then start a larger one
        for add in range(0, -(-(T - covered) // last)):
            c = covered + add * last
            if c >= T:
                break
            for n2 in sizes:
                if n2 <= last:
                    continue
                if 2 * N0 + c < 2 * n2:
                    continue
                rec(N + [n2], P[:-1] + [P[-1] + add] + [1], c + n2)

    rec([N0], [1], N0)
    return out


MODELS = [
    ("fft-heavy", lambda n: 2e-6 + 1e-9 * n * math.log2(2 * n), lambda n: 5e-10 * n),
    ("mac-heavy", lambda n: 1e-7 + 1e-10 * n * math.log2(2 * n), lambda n: 4e-9 * n),
    ("launch-bound", lambda n: 6e-6, lambda n: 2e-7 + 1e-11 * n),
]


@pytest.mark.parametrize("name,f,m", MODELS)
@pytest.mark.parametrize("N0,taps", [(64, 700), (64, 2000), (32, 1500), (128, 3000)])
def test_dp_matches_brute_force(name, f, m, N0, taps):
    mb = 1024 if N0 < 128 else 2048
    M = model(N0, mb, f, m)
    N, P, ld = tune.optimal_partition(N0, taps, M)
    # valid under the reference's rules
    R.check_partition(N, P, taps)
    assert abs(tune.load(N, P, M) - ld) < 1e-9 * max(1.0, ld)
    best = min(tune.load(n, p, M) for n, p in all_partitions(N0, taps, M.sizes()))
    assert ld <= best * (1 + 1e-9) + 1e-15
    assert ld >= best * (1 - 1e-9) - 1e-15


@pytest.mark.parametrize("name,f,m", MODELS)
def test_never_worse_than_greedy(name, f, m):
    N0 = 64
    M = model(N0, 32768, f, m)
    bs = [N0 << i for i in range(10)]
    for taps in (5000, 65536, 262144):
        g = R.greedy_partition(bs, taps)
        N, P, ld = tune.optimal_partition(N0, taps, M)
        R.check_partition(N, P, taps)
        assert ld <= tune.load(g.N, g.P, M) * (1 + 1e-9)


def test_rejects_bad_models():
    M = tune.CostModel(N0=64)
    with pytest.raises(ValueError):
        tune.optimal_partition(64, 1000, M)
