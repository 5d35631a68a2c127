"""Autotuner for the engines' launch shapes, with an on-disk cache.

The reference measures its own level costs before it distributes work
(PartConvFilter::bench, src/partconv.cpp:508-566; the work split of
PartConvMulti::setup, src/partconvMulti.cpp:290-408) and keeps its FFT plans as
wisdom on disk (src/partconv.cpp:255-285, src/wisdom.wis).  Here the choices
are the GPU engines' launch shapes, measured on this device and kept in a JSON
cache keyed by (device, workload shape):

  * real-time engine (PartConvMulti): who runs the short levels in persistent
    runs (pcm_set_persistent_workers: 1 separate worker workgroups when they
    fit, 2 every tick workgroup runs its own instance's blocks, 0 host-launched
    kernels) -- chosen by the mean tick latency of persistent device-resident
    ticks; then the delay-line-sum kernel's width (pcm_set_partial_width: 2, 4
    or 8 bins per thread) -- chosen by its HIP-event time on the longest level;
    then the release window of persistent device-resident runs
    (pcm_set_release_window: 1 to 4 ticks released ahead) -- chosen by the wall
    time per tick, so the smallest window within the tie of the fastest wins;
  * real-time MIMO engine (MimoRealtime): its release window
    (pcxr_set_release_window: 1 to 3 ticks), chosen by the wall time per tick;
  * differentiable form (PartitionedConv): each level's chunk sizes of the
    spectral products and of dL/dh (pcc_tune), replayed from the cache with
    pcc_set_level_tuning.

Ties are deterministic: candidates within TIE (5%) of the fastest count as
equal, and then the default wins, else the first in candidate order.  A cached
entry is applied without measuring, so a choice, once made on a device, is the
same in every later run.  None of the choices changes an engine's output.

    from partconv_amd import autotune
    info = autotune.tune_engine(m)                  # PartConvMulti after setup
    info = autotune.tune_mimo(rt)                   # MimoRealtime after set_filters
    info = autotune.tune_offline(conv, x, h)        # PartitionedConv + clips
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import pathlib
import tempfile
import time

from ._lib import last_error, lib

TIE = 0.05
WORKER_MODES = (1, 2, 0)      # default first
PARTIAL_WIDTHS = (2, 4, 8)    # default first
RELEASE_WINDOWS = (1, 2, 3, 4)  # default first
VERSION = 2


def default_cache_path():
    """$PARTCONV_AMD_TUNE_CACHE, else ~/.cache/partconv_amd/tuning.json."""
    p = os.environ.get("PARTCONV_AMD_TUNE_CACHE")
    if p:
        return pathlib.Path(p)
    return pathlib.Path(os.path.expanduser("~")) / ".cache" / "partconv_amd" / "tuning.json"


def choose(candidates, costs, default, tie=TIE):
    """Deterministic pick: every candidate whose cost is within `tie` of the
    lowest counts as equal; among those the default wins, else the first in
    candidate order."""
    best = min(costs)
    near = [c for c, v in zip(candidates, costs) if v <= best * (1.0 + tie)]
    return default if default in near else near[0]


def load_cache(path):
    try:
        with open(path) as fh:
            d = json.load(fh)
        if d.get("version") == VERSION and isinstance(d.get("entries"), dict):
            return d
    except (OSError, ValueError):
        pass
    return {"version": VERSION, "entries": {}}


def store_entry(path, key, value):
    """Add one entry to the cache file (read-modify-write, atomic replace)."""
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    d = load_cache(path)
    d["entries"][key] = value
    fd, tmp = tempfile.mkstemp(dir=str(path.parent), prefix=".tuning")
    with os.fdopen(fd, "w") as fh:
        json.dump(d, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def device_name():
    buf = ctypes.create_string_buffer(256)
    if lib().pc_device_name(buf, 256) != 0:
        raise RuntimeError(last_error())
    return buf.value.decode()


def engine_key(m, device=None):
    """(device, buffer size, relaxed, instance count, multiset of partitions)."""
    info = m.info()
    parts = {}
    for i in range(info.numPCs):
        N, P = m.instance_partition(i)
        k = ",".join(f"{a}x{b}" for a, b in zip(N, P))
        parts[k] = parts.get(k, 0) + 1
    shape = ";".join(f"{k}*{v}" for k, v in sorted(parts.items()))
    digest = hashlib.sha1(shape.encode()).hexdigest()[:16]
    return (f"engine|{device or device_name()}|N0={info.buffer_size}|relaxed={int(m._relaxed)}"
            f"|inst={info.numPCs}|parts={digest}")


def _apply_engine(m, workers, width, window=1):
    if lib().pcm_set_persistent_workers(m._h, int(workers)) != 0:
        raise RuntimeError(last_error())
    if lib().pcm_set_partial_width(m._h, int(width)) != 0:
        raise RuntimeError(last_error())
    if lib().pcm_set_release_window(m._h, int(window)) != 0:
        raise RuntimeError(last_error())


def tune_engine(m, nticks=None, cache=None, force=False):
    """Choose the persistent-run workers mode, the delay-line-sum width and the
    release window of a set-up PartConvMulti (batched engine) on this device;
    apply them; return {"workers": w, "bins_per_thread": b, "release_window":
    r, "from_cache": bool, "key": ..., "measured": {...}}.  A measurement runs persistent device-resident ticks
    of zero input and resets the engine afterwards (state and frame counter)."""
    import torch
    path = default_cache_path() if cache is None else pathlib.Path(cache)
    key = engine_key(m)
    ent = load_cache(path)["entries"].get(key)
    if ent and not force:
        _apply_engine(m, ent["workers"], ent["bins_per_thread"], ent["release_window"])
        return dict(ent, from_cache=True, key=key)
    info = m.info()
    n0, G = info.buffer_size, info.num_levels
    if nticks is None:  # two periods of the longest level, at least 256 ticks
        nticks = max(256, 2 * (info.max_block_size // n0))
    # the width is timed over at least four blocks of the longest level
    nwide = max(nticks, 4 * (info.max_block_size // n0))
    x = torch.zeros((nwide * n0, info.numInputChannels), dtype=torch.float32, device="cuda")
    y = torch.zeros((nwide * n0, info.numOutputChannels), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def run(n=nticks):
        m.reset()
        m.clear_stats()
        if m.run_device(y.data_ptr(), x.data_ptr(), n, True, persistent=True) != 0:
            raise RuntimeError(last_error())
        return m.stats()

    measured = {"workers_mean_us": {}, "partial_us": {}, "window_us_per_tick": {}}
    costs = []
    for w in WORKER_MODES:
        _apply_engine(m, w, PARTIAL_WIDTHS[0])
        run()                                   # warm-up (plans the workers, sizes launches)
        st = run()
        costs.append(st.mean_us)
        measured["workers_mean_us"][str(w)] = st.mean_us
    workers = choose(list(WORKER_MODES), costs, WORKER_MODES[0])
    costs = []
    for b in PARTIAL_WIDTHS:
        _apply_engine(m, workers, b)
        if G > 1:
            run()                               # warm-up
            m.probe(G - 1)                      # the delay-line sum of the longest level
            run(nwide)
            avg_us, n, _ = m.probe_read()
            m.probe(-1)
        else:
            avg_us, n = 0.0, 0
        costs.append(avg_us if n else 0.0)
        measured["partial_us"][str(b)] = avg_us if n else None
    width = choose(list(PARTIAL_WIDTHS), costs, PARTIAL_WIDTHS[0]) if any(costs) else PARTIAL_WIDTHS[0]
    # release window: wall time per tick of persistent runs (the engine's reset
    # outside the timed call; every call pays the same launch)
    costs = []
    for r in RELEASE_WINDOWS:
        _apply_engine(m, workers, width, r)
        run(nwide)                              # warm-up
        m.reset()
        m.clear_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if m.run_device(y.data_ptr(), x.data_ptr(), nwide, True, persistent=True) != 0:
            raise RuntimeError(last_error())
        torch.cuda.synchronize()
        us = 1e6 * (time.perf_counter() - t0) / nwide
        costs.append(us)
        measured["window_us_per_tick"][str(r)] = us
    window = choose(list(RELEASE_WINDOWS), costs, RELEASE_WINDOWS[0])
    _apply_engine(m, workers, width, window)
    m.reset()
    m.clear_stats()
    ent = {"workers": workers, "bins_per_thread": width, "release_window": window, "measured": measured,
           "ticks": [nticks, nwide], "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    store_entry(path, key, ent)
    return dict(ent, from_cache=False, key=key)


MIMO_WINDOWS = (1, 2, 3)  # default first


def mimo_key(rt, device=None):
    N, P = rt.partition()
    parts = ",".join(f"{a}x{b}" for a, b in zip(N, P))
    return (f"mimo|{device or device_name()}|N0={rt.buffer_size}|in={rt.n_in}|out={rt.n_out}"
            f"|taps={rt.taps}|parts={parts}")


def tune_mimo(rt, nticks=None, cache=None, force=False):
    """Choose the release window of a MimoRealtime with its filters set, on
    this device: the wall time per tick of device-resident runs of zero input
    for each window (the engine reset before each timed call, and after);
    apply it and return {"release_window": w, "from_cache": bool, "key": ...,
    "measured": {...}}."""
    import torch
    path = default_cache_path() if cache is None else pathlib.Path(cache)
    key = mimo_key(rt)
    ent = load_cache(path)["entries"].get(key)
    if ent and not force:
        rt.set_release_window(ent["release_window"])
        return dict(ent, from_cache=True, key=key)
    N, _ = rt.partition()
    n0 = rt.buffer_size
    if nticks is None:  # two blocks of the longest level, at least 128 ticks
        nticks = max(128, 2 * (max(N) // n0))
    x = torch.zeros((nticks * n0, rt.n_in), dtype=torch.float32, device="cuda")
    y = torch.zeros((nticks * n0, rt.n_out), dtype=torch.float32, device="cuda")
    costs, measured = [], {}
    for w in MIMO_WINDOWS:
        rt.set_release_window(w)
        rt.reset()
        rt.run_device(y, x, nticks)              # warm-up
        rt.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rt.run_device(y, x, nticks)
        torch.cuda.synchronize()
        us = 1e6 * (time.perf_counter() - t0) / nticks
        costs.append(us)
        measured[str(w)] = us
    window = choose(list(MIMO_WINDOWS), costs, MIMO_WINDOWS[0])
    rt.set_release_window(window)
    rt.reset()
    rt.clear_stats()
    ent = {"release_window": window, "measured": {"window_us_per_tick": measured}, "ticks": nticks,
           "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    store_entry(path, key, ent)
    return dict(ent, from_cache=False, key=key)


def offline_key(conv, C, T, device=None):
    parts = ",".join(f"{a}x{b}" for a, b in zip(conv.N, conv.P))
    return f"offline|{device or device_name()}|taps={conv.taps}|parts={parts}|C={C}|T={T}"


def tune_offline(conv, x, h, cache=None, force=False):
    """Chunk sizes of a PartitionedConv for clips x [C, T] (pcc_tune on a cache
    miss, pcc_set_level_tuning on a hit); returns {"levels": [(mac, dh), ...],
    "from_cache": bool, "key": ...}."""
    path = default_cache_path() if cache is None else pathlib.Path(cache)
    key = offline_key(conv, int(x.shape[0]), int(x.shape[1]))
    ent = load_cache(path)["entries"].get(key)
    if ent and not force:
        for lvl, (mac, dh) in enumerate(ent["levels"]):
            if lib().pcc_set_level_tuning(conv._h, lvl, int(mac), int(dh)) != 0:
                raise RuntimeError(last_error())
        return dict(ent, from_cache=True, key=key)
    levels = [list(v) for v in conv.tune(x, h)]
    ent = {"levels": levels, "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    store_entry(path, key, ent)
    return dict(ent, from_cache=False, key=key)
