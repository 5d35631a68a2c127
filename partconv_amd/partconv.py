"""Python surface of the engine, mirroring the reference's SWIG module `partconv`
(src/partconv.i:3-11: PartConvFilter, intArray, doubleArray) and adding
PartConvMulti / PartConvMultiRelaxed, which the reference does not wrap.

Method names, argument meaning and return conventions follow the C++ classes
(src/partconv.h:40-114, src/partconvMulti.h:52-148): setup/run/reset/cleanup
return 0 on success and a nonzero status otherwise (the message is printed to
stderr, as the reference does, and kept in partconv_amd.last_error()).  Where
the reference calls exit(-1) (config/IR/frameCount errors) these return the
status instead.

All compute runs in libpartconv_amd.so on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import PcInstance, PcMultiInfo, PcStats, lib, last_error


def _report(rc, what):
    if rc != 0:
        print(f"{what}: {last_error()}", file=sys.stderr)
    return rc


class intArray:
    """SWIG %array_class(int, intArray) work-alike (src/partconv.i:10)."""

    def __init__(self, nelements):
        self._a = (ctypes.c_int * int(nelements))()

    def __getitem__(self, i):
        return self._a[i]

    def __setitem__(self, i, v):
        self._a[i] = int(v)

    def __len__(self):
        return len(self._a)

    def cast(self):
        return self._a

    @staticmethod
    def frompointer(p):
        raise NotImplementedError("raw pointers are not exposed")


class doubleArray:
    """SWIG %array_class(double, doubleArray) work-alike (src/partconv.i:11)."""

    def __init__(self, nelements):
        self._a = (ctypes.c_double * int(nelements))()

    def __getitem__(self, i):
        return self._a[i]

    def __setitem__(self, i, v):
        self._a[i] = float(v)

    def __len__(self):
        return len(self._a)

    def cast(self):
        return self._a


def _as_int_list(x, n=None):
    if isinstance(x, intArray):
        x = list(x._a)
    x = [int(v) for v in x]
    return x if n is None else x[:n]


def readImpulseResponse(filename, verbosity=0):
    """readImpulseResponse (src/sndtools.cpp:3-47): the first channel of a
    RIFF/WAVE file as float32 (integer PCM scaled by 2^(bits-1)); returns
    (samples, sample_rate).  Raises OSError for a missing, unsupported or
    malformed file."""
    n = ctypes.c_longlong()
    sr = ctypes.c_int()
    path = str(filename).encode()
    if lib().pc_read_ir(path, None, 0, ctypes.byref(n), ctypes.byref(sr), int(verbosity)) != 0:
        raise OSError(last_error())
    out = np.zeros(n.value, dtype=np.float32)
    if n.value and lib().pc_read_ir(path, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n.value,
                                    ctypes.byref(n), None, 0) != 0:
        raise OSError(last_error())
    return out, sr.value


def greedy_partition(block_sizes, filter_length):
    """PartConvMax's partition rule (src/partconv.cpp:712-728): returns (N, P)
    with the last count recounted for filter_length taps."""
    bs = [int(b) for b in block_sizes]
    n = len(bs)
    B = (ctypes.c_int * n)(*bs)
    N = (ctypes.c_int * n)()
    P = (ctypes.c_int * n)()
    M = ctypes.c_int()
    rc = lib().pc_greedy_partition(B, n, int(filter_length), N, P, ctypes.byref(M))
    if rc != 0:
        raise ValueError(last_error())
    return list(N)[:M.value], list(P)[:M.value]


class PartConvFilter:
    """PartConvFilter (src/partconv.h:40-114): one instance, caller-driven levels."""

    def __init__(self, inVerbosity=0, inSampleRate=44100):
        self._h = lib().pcf_create(int(inVerbosity), int(inSampleRate))
        self.verbosity = int(inVerbosity)
        self._M = 0
        self._N = []
        self._P = []

    # public members (src/partconv.h:66-71)
    @property
    def FS(self):
        return lib().pcf_get_fs(self._h)

    @FS.setter
    def FS(self, v):
        lib().pcf_set_fs(self._h, int(v))

    @property
    def M(self):
        return self._M

    @property
    def N(self):
        return list(self._N)

    @property
    def num_parts(self):
        return list(self._P)

    def setup(self, M_in, N_in, num_parts_in, impulse=None, fftwOptLevel=1, wisdom=None):
        """PartConvFilter::setup (src/partconv.cpp:70-306); 0 on success, 1 on error."""
        M = int(M_in)
        N = (ctypes.c_int * M)(*_as_int_list(N_in, M))
        P = (ctypes.c_int * M)(*_as_int_list(num_parts_in, M))
        if impulse is None:
            ir, n = None, 0
        else:
            self._ir = np.ascontiguousarray(impulse, dtype=np.float32)
            ir, n = self._ir.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self._ir.size
        rc = lib().pcf_setup(self._h, M, N, P, ir, n, int(fftwOptLevel),
                             None if wisdom is None else str(wisdom).encode())
        if rc != 0:
            _report(rc, "PartConvFilter ERROR")
            return 1
        # the recounted partition lives in the filter's members; the caller's
        # arrays are left as they were (src/partconv.cpp:83-84, 169)
        Mo = ctypes.c_int()
        No = (ctypes.c_int * M)()
        Po = (ctypes.c_int * M)()
        lib().pcf_get_levels(self._h, ctypes.byref(Mo), No, Po, M)
        self._M = M
        self._N = list(No)
        self._P = list(Po)
        if self.verbosity > 0:
            print("partitioning: " + ", ".join(f"{a}x{b}" for a, b in zip(self._N, self._P)))
        return 0

    def runLevel(self, L):
        return _report(lib().pcf_run_level(self._h, int(L)), "runLevel")

    def sync_levels(self):
        lib().pcf_sync_levels(self._h)

    def sync_levels_relaxed(self):
        lib().pcf_sync_levels_relaxed(self._h)

    def input(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.size < self._N[0]:
            raise ValueError("input needs N[0] samples")
        return _report(lib().pcf_input(self._h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "input")

    def _out(self, fn):
        y = np.zeros(self._N[0], dtype=np.float32)
        rc = fn(self._h, y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        if rc != 0:
            raise RuntimeError(last_error())
        return y

    def output(self):
        return self._out(lib().pcf_output)

    def output_relaxed(self):
        return self._out(lib().pcf_output_relaxed)

    def reset(self):
        lib().pcf_reset(self._h)

    def bench(self, L, trials=100, pollute_every=0, max_time=0.1, benchData=None):
        """PartConvFilter::bench (src/partconv.cpp:508-566): mean seconds per level-L block."""
        bd = None
        if benchData is not None:
            bd = benchData._a if isinstance(benchData, doubleArray) else None
        t = lib().pcf_bench(self._h, int(L), int(trials), int(pollute_every), float(max_time), bd)
        if t < 0:
            raise RuntimeError(last_error())
        return t

    def cleanup(self):
        return lib().pcf_cleanup(self._h)

    def process(self, x, relaxed=False):
        """Convenience: run a whole mono signal through the tick sequence that
        PartConvMulti::run uses for one instance; returns whole frames."""
        N0 = self._N[0]
        nfr = -(-len(x) // N0)
        xin = np.zeros(nfr * N0, dtype=np.float32)
        xin[:len(x)] = x
        y = np.zeros_like(xin)
        for t in range(nfr):
            if relaxed:
                self.sync_levels_relaxed()
                y[t * N0:(t + 1) * N0] = self.output_relaxed()
                self.input(xin[t * N0:(t + 1) * N0])
                for L in range(lib().pcf_start_sync_depth(self._h) + 1):
                    self.runLevel(L)
            else:
                self.sync_levels()
                self.input(xin[t * N0:(t + 1) * N0])
                for L in range(lib().pcf_start_sync_depth(self._h) + 1):
                    self.runLevel(L)
                y[t * N0:(t + 1) * N0] = self.output()
        return y

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcf_destroy", h)
        except Exception:
            pass


class PartConvMulti:
    """PartConvMulti (src/partconvMulti.h:52-140) on the GPU engine."""

    _relaxed = 0

    def __init__(self, inVerbosity=0, mode="batched"):
        self._h = lib().pcm_create(int(inVerbosity), self._relaxed)
        self.verbosity = int(inVerbosity)
        self.set_mode(mode)

    def set_mode(self, mode):
        m = {"batched": _lib.PC_BATCHED, "plain": _lib.PC_PLAIN}[mode]
        return _report(lib().pcm_set_mode(self._h, m), "set_mode")

    def setup(self, config_file_name, max_threads_per_level=1, max_level0_threads=0, min_work=0.0):
        """PartConvMulti::setup (src/partconvMulti.cpp:23-718): pc.config + WAV IRs."""
        rc = lib().pcm_setup(self._h, str(config_file_name).encode(), int(max_threads_per_level),
                             int(max_level0_threads), float(min_work))
        return _report(rc, "PartConvMulti setup")

    def setup_mem(self, FS, buffer_size, numInputChannels, numOutputChannels, instances):
        """In-memory setup.  instances: dicts with N, P, ir and optional in_ch,
        out_ch, mute, scale (one config line each)."""
        n = len(instances)
        arr = (PcInstance * n)()
        self._keep = []
        for i, d in enumerate(instances):
            N = (ctypes.c_int * len(d["N"]))(*[int(v) for v in d["N"]])
            P = (ctypes.c_int * len(d["P"]))(*[int(v) for v in d["P"]])
            ir = np.ascontiguousarray(d["ir"], dtype=np.float32)
            self._keep += [N, P, ir]
            arr[i] = PcInstance(int(d.get("mute", 0)), int(d.get("in_ch", 0)), int(d.get("out_ch", 0)),
                                float(d.get("scale", 1.0)), len(d["N"]), N, P,
                                ir.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ir.size)
        rc = lib().pcm_setup_mem(self._h, int(FS), int(buffer_size), int(numInputChannels),
                                 int(numOutputChannels), n, arr)
        self._keep = []
        return _report(rc, "PartConvMulti setup")

    def run(self, output, input, frameCount):
        """PartConvMulti::run (src/partconvMulti.cpp:815-928): one tick on host
        float32 buffers (interleaved)."""
        if output.dtype != np.float32 or input.dtype != np.float32:
            raise TypeError("float32 buffers required")
        if not (output.flags.c_contiguous and input.flags.c_contiguous):
            raise ValueError("contiguous buffers required")
        return _report(lib().pcm_run(self._h, output.ctypes.data, input.ctypes.data, int(frameCount)), "run")

    def run_device(self, out_ptr, in_ptr, nframes, sync_each_tick=True, persistent=False):
        """Device-resident ticks (pcm_run_device): integer device pointers to
        interleaved float32 frames.  persistent=True serves the ticks with one
        persistent tick kernel (real-time semantics kept)."""
        mode = 2 if persistent else (1 if sync_each_tick else 0)
        return _report(lib().pcm_run_device(self._h, ctypes.c_void_p(int(out_ptr)),
                                            ctypes.c_void_p(int(in_ptr)), int(nframes), mode),
                       "run_device")

    def process(self, x):
        """Convenience: run interleaved (n, numInputChannels) host samples tick by
        tick through run(); returns (frames*buffer_size, numOutputChannels)."""
        info = self.info()
        N0 = info.buffer_size
        x = np.asarray(x, dtype=np.float32)
        if x.ndim == 1:
            x = x[:, None]
        nfr = -(-x.shape[0] // N0)
        xin = np.zeros((nfr * N0, info.numInputChannels), dtype=np.float32)
        xin[:x.shape[0]] = x
        y = np.zeros((nfr * N0, info.numOutputChannels), dtype=np.float32)
        if lib().pcm_process(self._h, y.ctypes.data, xin.ctypes.data, nfr) != 0:
            raise RuntimeError(last_error())
        return y

    def reset(self):
        lib().pcm_reset(self._h)

    def cleanup(self):
        return lib().pcm_cleanup(self._h)

    def info(self):
        i = PcMultiInfo()
        if lib().pcm_get_info(self._h, ctypes.byref(i)) != 0:
            raise RuntimeError(last_error())
        return i

    def stats(self):
        s = PcStats()
        if lib().pcm_get_stats(self._h, ctypes.byref(s)) != 0:
            raise RuntimeError(last_error())
        return s

    def clear_stats(self):
        lib().pcm_clear_stats(self._h)

    def latencies(self, cap=1 << 20):
        buf = np.zeros(cap, dtype=np.float64)
        n = lib().pcm_get_latencies(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap)
        return buf[:max(0, n)]

    def probe(self, global_level):
        """Time the multiply-accumulate kernel of one global level (0 = tick
        kernel) with HIP events from now on; -1 disables."""
        return _report(lib().pcm_probe(self._h, int(global_level)), "probe")

    def probe_read(self):
        """(avg_us, launches, algorithmic_bytes_per_launch) of the probed kernel."""
        a, b = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        if lib().pcm_probe_read(self._h, ctypes.byref(a), ctypes.byref(n), ctypes.byref(b)) != 0:
            raise RuntimeError(last_error())
        return a.value, n.value, b.value

    def trace(self, nticks):
        """Record GPU timestamps of the next persistent run (see pcm_trace)."""
        return _report(lib().pcm_trace(self._h, int(nticks)), "trace")

    def set_host_session(self, on):
        """run() through a persistent tick kernel kept resident between calls
        (True / 1, default; input frames as tagged words in device memory where
        they fit), 2 (a session with a release word and the frame read from
        pinned host memory),
        or one launch per tick (False / 0); identical output."""
        mode = 1 if on is True else (0 if on is False else int(on))
        return _report(lib().pcm_set_host_session(self._h, mode), "set_host_session")

    def set_release_window(self, ticks):
        """Ticks the persistent device-resident run (run_device(..., persistent=True))
        may release ahead of the oldest incomplete tick (1 .. 16, default 1),
        where each instance has its own output channel; identical output."""
        return _report(lib().pcm_set_release_window(self._h, int(ticks)), "set_release_window")

    def set_persistent_workers(self, mode):
        """Short-period levels in persistent runs: True/1 inside the kernel (worker
        workgroups, or self mode when they do not fit), 2 always self mode,
        False/0 host launches.  Output is identical."""
        return _report(lib().pcm_set_persistent_workers(self._h, int(mode)), "set_persistent_workers")

    @property
    def workers(self):
        """Number of level-worker workgroups a persistent run adds (0 if none)."""
        return lib().pcm_workers_info(self._h, None)

    @property
    def self_mode(self):
        """True when the tick workgroups run the short-period levels themselves."""
        v = ctypes.c_int()
        lib().pcm_workers_info(self._h, ctypes.byref(v))
        return bool(v.value)

    @property
    def release_in_device_memory(self):
        """True when the persistent kernel's release counter is in device memory
        written through the PCIe BAR (polled locally), False when it is in
        pinned host memory (polled across the bus)."""
        return lib().pcm_release_info(self._h) == 1

    def trace_read(self, cap=1 << 20):
        """Per tick 8 GPU clock values (100 MHz): release seen / published by
        workgroups 0, njobs/2, njobs-1; completion raised; workgroup 0 ready."""
        buf = (ctypes.c_ulonglong * (8 * cap))()
        n = lib().pcm_trace_read(self._h, buf, cap)
        return np.frombuffer(buf, dtype=np.uint64)[:8 * max(n, 0)].reshape(-1, 8)

    def instance_partition(self, i):
        N = (ctypes.c_int * 16)()
        P = (ctypes.c_int * 16)()
        M = lib().pcm_instance_partition(self._h, int(i), N, P, 16)
        return list(N)[:M], list(P)[:M]

    def set_level_graphs(self, on):
        """Launch level blocks as per-level HIP graphs (pcm_set_level_graphs)."""
        return _report(lib().pcm_set_level_graphs(self._h, int(bool(on))), "set_level_graphs")

    def stall_level(self, global_level, seconds):
        """Deadline-handling test hook (pcm_stall_level): the next block of a
        global level waits `seconds` behind a host function on its stream."""
        return _report(lib().pcm_stall_level(self._h, int(global_level), float(seconds)), "stall_level")

    # public fields (src/partconvMulti.h:64-79)
    FS = property(lambda self: self.info().FS)
    buffer_size = property(lambda self: self.info().buffer_size)
    max_block_size = property(lambda self: self.info().max_block_size)
    numInputChannels = property(lambda self: self.info().numInputChannels)
    numOutputChannels = property(lambda self: self.info().numOutputChannels)
    frameNum = property(lambda self: self.info().frameNum)

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcm_destroy", h)
        except Exception:
            pass


class PartConvMultiRelaxed(PartConvMulti):
    """PartConvMultiRelaxed (src/partconvMulti.h:142-148): output delayed by one
    buffer, every level gets a full period."""

    _relaxed = 1


class PartConvMax:
    """PartConvMax (src/partconv.h, src/partconv.cpp:653-1313), the reference's
    in-memory front end: channel modes chosen from (n, m, k), IRs from one
    array, one greedy partition, per-channel buffers.  Runs on the batched GPU
    engine (strict semantics, zero added latency)."""

    def __init__(self, inVerbosity=0):
        self._h = lib().pcm_create(int(inVerbosity), 0)
        self.verbosity = int(inVerbosity)

    def setup(self, inSampleRate, n, m, k, impulses, filterLength, stride, blockSizes, levels,
              fftwOptLevel=1, wisdom=None, max_threads_per_level=1, max_level0_threads=0):
        """PartConvMax::setup (src/partconv.cpp:653-847).  impulses: float32
        array holding IR i, tap l at impulses[stride*(filterLength*i + l)];
        blockSizes: the candidate block sizes (first = buffer size).  The FFTW
        and thread arguments are accepted and have no effect.  Returns 0 / 1."""
        imp = np.ascontiguousarray(impulses, dtype=np.float32).ravel()
        need = stride * (filterLength * k - 1) + 1
        if imp.size < need:
            print(f"PartConvMax setup: impulses holds {imp.size} values, {need} needed", file=sys.stderr)
            return 1
        bs = (ctypes.c_int * int(levels))(*[int(v) for v in list(blockSizes)[:int(levels)]])
        rc = lib().pcm_setup_max(self._h, int(inSampleRate), int(n), int(m), int(k),
                                 ctypes.c_void_p(imp.ctypes.data), int(filterLength), int(stride), bs, int(levels))
        return 1 if _report(rc, "PartConvMax setup") else 0

    def run(self, output, input):
        """PartConvMax::run (src/partconv.cpp:1219-1303): one tick.  output and
        input are sequences of float32 arrays of buffer_size samples (one per
        channel), or 2-D (channels, buffer_size) arrays."""
        outs = [np.asarray(o) for o in output]
        ins = [np.ascontiguousarray(x, dtype=np.float32) for x in input]
        for o in outs:
            if o.dtype != np.float32 or not o.flags.c_contiguous:
                raise TypeError("output channels must be contiguous float32 arrays")
        po = (ctypes.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        pi = (ctypes.c_void_p * len(ins))(*[x.ctypes.data for x in ins])
        return _report(lib().pcm_run_channels(self._h, po, pi), "PartConvMax run")

    def process(self, x):
        """Convenience: (n_samples, numInputChannels) host samples, tick by tick
        through run(); returns (frames*buffer_size, numOutputChannels)."""
        info = self.info()
        N0, ni, no = info.buffer_size, info.numInputChannels, info.numOutputChannels
        x = np.asarray(x, dtype=np.float32)
        if x.ndim == 1:
            x = x[:, None]
        nfr = -(-x.shape[0] // N0)
        xin = np.zeros((ni, nfr * N0), dtype=np.float32)
        xin[:, :x.shape[0]] = x.T
        y = np.zeros((no, nfr * N0), dtype=np.float32)
        fo = np.zeros((no, N0), dtype=np.float32)
        for t in range(nfr):
            if self.run(fo, xin[:, t * N0:(t + 1) * N0]) != 0:
                raise RuntimeError(last_error())
            y[:, t * N0:(t + 1) * N0] = fo
        return y.T.copy()

    def reset(self):
        lib().pcm_reset(self._h)

    def cleanup(self):
        return lib().pcm_cleanup(self._h)

    def info(self):
        i = PcMultiInfo()
        if lib().pcm_get_info(self._h, ctypes.byref(i)) != 0:
            raise RuntimeError(last_error())
        return i

    def instance_partition(self, i):
        N = (ctypes.c_int * 16)()
        P = (ctypes.c_int * 16)()
        M = lib().pcm_instance_partition(self._h, int(i), N, P, 16)
        return list(N)[:M], list(P)[:M]

    # public fields (src/partconv.h:117-175)
    FS = property(lambda self: self.info().FS)
    buffer_size = property(lambda self: self.info().buffer_size)
    max_block_size = property(lambda self: self.info().max_block_size)
    numInputChannels = property(lambda self: self.info().numInputChannels)
    numOutputChannels = property(lambda self: self.info().numOutputChannels)
    num_levels = property(lambda self: self.info().num_levels)
    frameNum = property(lambda self: self.info().frameNum)

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcm_destroy", h)
        except Exception:
            pass
