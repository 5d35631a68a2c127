"""Integer rules of partconv, restated in pure Python (TEST INFRASTRUCTURE ONLY).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module; it is the checker for the product's C++ setup code, never part of it.

Restated reference logic:
  * partition validation, per-level delays, last-level recount
        PartConvFilter::setup   src/partconv.cpp:86-182
  * greedy partition rule P_i = 2 N_{i+1}/N_i - 2 with early stop
        PartConvMax::setup      src/partconv.cpp:712-728
  * global level set and level_map
        PartConvMulti::setup    src/partconvMulti.cpp:139-244
  * pc.config grammar
        PartConvMulti::setup    src/partconvMulti.cpp:36-134
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


class PartitionError(ValueError):
    pass


class ConfigError(ValueError):
    pass


def _is_pow2(x: int) -> bool:
    return x > 0 and (x & (x - 1)) == 0


@dataclass
class Partition:
    N: list          # block size per level
    P: list          # partitions per level (last one recounted)
    delay: list      # d_L in frames of N[0]  (src/partconv.cpp:148-167)
    T: list          # taps before level L
    total: int       # taps covered (zero-padded IR length)


def check_partition(N, P, ir_len=None) -> Partition:
    """PartConvFilter::setup validation (src/partconv.cpp:86-182).

    ir_len=None mirrors impulse==NULL: an all-ones IR of sum(N*P) taps
    (src/partconv.cpp:132-142)."""
    N = [int(n) for n in N]
    P = [int(p) for p in P]
    M = len(N)
    if M < 1 or len(P) != M:
        raise PartitionError("need M >= 1 levels with one count each")
    for i, n in enumerate(N):
        if not _is_pow2(n):
            raise PartitionError(f"block size of level {i} must be a power of 2")
    for i in range(1, M):
        if N[i] <= N[i - 1]:
            raise PartitionError(f"N[{i}]={N[i]} <= N[{i-1}]={N[i-1]}")
    length = sum(n * p for n, p in zip(N, P)) if ir_len is None else int(ir_len)
    position = 2 * N[0]
    delay = [0] * M
    for i in range(M - 1):
        delay[i] = position // N[0]
        position += N[i] * P[i]
        if position < 2 * N[i + 1]:
            raise PartitionError(
                f"sufficient delay not allowed for block size of level {i + 2}")
    delay[M - 1] = position // N[0]
    rest = length - (position - 2 * N[0])
    if rest < 0:
        raise PartitionError(f"last partitioning level ({M}) not required")
    # float32 division then ceil, as the reference computes it
    import numpy as np
    plast = int(math.ceil(float(np.float32(rest) / np.float32(N[M - 1]))))
    if plast == 0:
        raise PartitionError(f"last partitioning level ({M}) not required")
    P = P[:M - 1] + [plast]
    T, t = [], 0
    for n, p in zip(N, P):
        T.append(t)
        t += n * p
    return Partition(N=N, P=P, delay=delay, T=T, total=t)


def greedy_partition(block_sizes, filter_len):
    """PartConvMax partition rule (src/partconv.cpp:712-728): P_i = 2N_{i+1}/N_i - 2,
    stop at the first level whose cumulative coverage reaches the IR; the last
    level's count is then recounted by check_partition."""
    bs = [int(b) for b in block_sizes]
    levels = len(bs)
    P = [0] * levels
    total = 0
    for i in range(levels - 1):
        P[i] = 2 * bs[i + 1] // bs[i] - 2
        total += bs[i] * P[i]
        if total >= filter_len:
            levels = i + 1
            break
    N = bs[:levels]
    P = P[:levels]
    return check_partition(N, P, filter_len)


def level_map(buffer_size, partitions):
    """Global level set and level_map (src/partconvMulti.cpp:139-244).

    partitions: list of per-instance block-size lists.  Returns
    (level_sizes, lmap) where level_sizes[g] is the global level's block size in
    frames and lmap[i][g] the instance's local level or -1."""
    max_block = max(max(n) for n in partitions)
    temp = int(math.log2(max_block)) - int(math.log2(buffer_size)) + 1
    sizes, lmap = [], [[] for _ in partitions]
    for g in range(temp):
        bs = buffer_size << g
        if not any(bs in n for n in partitions):
            continue
        sizes.append(1 << g)
        for i, n in enumerate(partitions):
            lmap[i].append(n.index(bs) if bs in n else -1)
    return sizes, lmap


@dataclass
class ConfigInstance:
    ir_file: str
    mute: int
    in_ch: int
    out_ch: int
    scale: float
    N: list
    P: list


@dataclass
class Config:
    FS: int
    buffer_size: int
    n_in: int
    n_out: int
    num_pcs: int
    num_repeats: int
    instances: list = field(default_factory=list)


def parse_config(text: str) -> Config:
    """pc.config grammar (src/partconvMulti.cpp:36-134).

    Line 1: FS buffer_size nIn nOut numPCs numRepeats.  Then numPCs instance
    records `irfile mute inCh outCh scale M N0 P0 ... N_{M-1} [P_{M-1}]`; the
    last count is optional and ignored (set to 0, recounted at setup).  Tokens
    after the numPCs records are ignored.  Repeats are muted copies
    (src/partconvMulti.cpp:118-133)."""
    toks = text.split()
    pos = 0

    def take(conv, what):
        nonlocal pos
        if pos >= len(toks):
            raise ConfigError(f"Error reading config file! ({what})")
        try:
            v = conv(toks[pos])
        except ValueError:
            raise ConfigError(f"Error reading config file! ({what}: {toks[pos]!r})")
        pos += 1
        return v

    FS = take(int, "FS")
    bsz = take(int, "buffer_size")
    nin = take(int, "numInputChannels")
    nout = take(int, "numOutputChannels")
    npcs = take(int, "numPCs")
    nrep = take(int, "numRepeats")
    cfg = Config(FS, bsz, nin, nout, npcs, nrep)
    base = []
    for i in range(npcs):
        fn = take(str, "filename")
        mute = take(int, "mute")
        ic = take(int, "inputChannel")
        oc = take(int, "outputChannel")
        sc = take(float, "scaleFactor")
        M = take(int, "M")
        N, P = [], []
        for j in range(M):
            N.append(take(int, "N"))
            if j < M - 1:
                P.append(take(int, "num_parts"))
            else:
                # optional trailing count: consumed only if it parses as an int
                if pos < len(toks):
                    try:
                        int(toks[pos])
                        pos += 1
                    except ValueError:
                        pass
                P.append(0)
        if N[0] != bsz:
            raise ConfigError(f"line {i + 1}: first block size must equal buffer_size")
        base.append(ConfigInstance(fn, mute, ic, oc, sc, N, P))
    cfg.instances = list(base)
    for r in range(1, nrep):
        for b in base:
            cfg.instances.append(ConfigInstance(b.ir_file, 1, b.in_ch, b.out_ch,
                                                b.scale, list(b.N), list(b.P)))
    return cfg
