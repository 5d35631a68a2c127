"""MIMO forms: the throughput form on the matrix cores (pcx_* entry points) and
the real-time engine (pcxr_*).

    mx = MimoConv([256, 4096], [30, 0], taps=65536, n_in=32, n_out=32)
    mx.set_filters(h)        # h: [n_out, n_in, taps] float32 cuda
    y = mx(x)                # x: [n_in, T] -> y: [n_out, T]

y[o] = sum_i conv(x[i], h[o][i]) over whole device-resident signals: every
frequency bin's complex n_out x n_in product, for all output blocks at once,
runs on v_mfma_f32_32x32x2f32.

    rt = MimoRealtime(48000, 256, [256, 4096], [30, 0], taps=65536, n_in=32, n_out=32)
    rt.set_filters(h)
    rt.run_device(y, x, nframes)   # x: [nframes*256, n_in], y: [nframes*256, n_out] cuda, interleaved
    rt.run(out, inp)               # or one tick on host arrays [256, n_in] -> [256, n_out]

runs the same network tick by tick in real time: the output equals one
PartConvMulti instance per (input, output) pair summed per output (the
reference's formulation), with one transform and one delay line per input
and a per-bin matrix-vector product instead.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import last_error, lib


class MimoConv:
    def __init__(self, N, P, taps, n_in, n_out):
        n = len(N)
        Na = (ctypes.c_int * n)(*[int(v) for v in N])
        Pa = (ctypes.c_int * n)(*[int(v) for v in P])
        self._h = lib().pcx_create(n, Na, Pa, int(taps), int(n_in), int(n_out))
        if not self._h:
            raise ValueError(last_error())
        self.taps, self.n_in, self.n_out = int(taps), int(n_in), int(n_out)

    def _s(self):
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def set_filters(self, h):
        h = h.contiguous()
        if tuple(h.shape) != (self.n_out, self.n_in, self.taps) or h.dtype != torch.float32 or not h.is_cuda:
            raise ValueError("h must be float32 cuda [n_out, n_in, taps]")
        self._hkeep = h
        if lib().pcx_set_filters(self._h, ctypes.c_void_p(h.data_ptr()), self._s()) != 0:
            raise RuntimeError(last_error())

    def __call__(self, x):
        x = x.contiguous()
        if x.dim() != 2 or x.shape[0] != self.n_in or x.dtype != torch.float32 or not x.is_cuda:
            raise ValueError("x must be float32 cuda [n_in, T]")
        y = torch.empty((self.n_out, x.shape[1]), dtype=torch.float32, device=x.device)
        if lib().pcx_forward(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                             int(x.shape[1]), self._s()) != 0:
            raise RuntimeError(last_error())
        return y

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcx_destroy", h)
        except Exception:
            pass


class MimoRealtime:
    def __init__(self, FS, buffer_size, N, P, taps, n_in, n_out):
        n = len(N)
        Na = (ctypes.c_int * n)(*[int(v) for v in N])
        Pa = (ctypes.c_int * n)(*[int(v) for v in P])
        self._h = lib().pcxr_create(int(FS), int(buffer_size), n, Na, Pa, int(taps), int(n_in), int(n_out))
        if not self._h:
            raise ValueError(last_error())
        self.buffer_size, self.taps, self.n_in, self.n_out = int(buffer_size), int(taps), int(n_in), int(n_out)

    def set_filters(self, h):
        h = h.contiguous()
        if tuple(h.shape) != (self.n_out, self.n_in, self.taps) or h.dtype != torch.float32 or not h.is_cuda:
            raise ValueError("h must be float32 cuda [n_out, n_in, taps]")
        torch.cuda.synchronize()
        if lib().pcxr_set_filters(self._h, ctypes.c_void_p(h.data_ptr())) != 0:
            raise RuntimeError(last_error())

    def run_device(self, y, x, nframes):
        """nframes ticks over device-resident interleaved frames (x: [nframes *
        buffer_size, n_in], y: [nframes * buffer_size, n_out], float32 cuda)."""
        n = int(nframes) * self.buffer_size
        if x.shape[0] < n or x.shape[1] != self.n_in or y.shape[0] < n or y.shape[1] != self.n_out:
            raise ValueError("frame buffers too small")
        if not (x.is_contiguous() and y.is_contiguous()):
            raise ValueError("frame buffers must be contiguous")
        if x.dtype != torch.float32 or y.dtype != torch.float32 or not (x.is_cuda and y.is_cuda):
            raise ValueError("frame buffers must be float32 cuda tensors")
        torch.cuda.synchronize()
        if lib().pcxr_run_device(self._h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                 int(nframes)) != 0:
            raise RuntimeError(last_error())

    def run(self, out, inp):
        """One tick on host buffers: inp float32 [buffer_size, n_in], out float32
        [buffer_size, n_out], C-contiguous numpy arrays (PartConvMulti.run's
        interleaved layout)."""
        import numpy as np
        if (not isinstance(inp, np.ndarray) or not isinstance(out, np.ndarray) or inp.dtype != np.float32
                or out.dtype != np.float32 or inp.shape != (self.buffer_size, self.n_in)
                or out.shape != (self.buffer_size, self.n_out) or not inp.flags.c_contiguous
                or not out.flags.c_contiguous):
            raise ValueError("run takes float32 C-contiguous [buffer_size, n_in] / [buffer_size, n_out] arrays")
        if lib().pcxr_run(self._h, out.ctypes.data, inp.ctypes.data, self.buffer_size) != 0:
            raise RuntimeError(last_error())

    def reset(self):
        if lib().pcxr_reset(self._h) != 0:
            raise RuntimeError(last_error())

    def stats(self):
        from ._lib import PcStats
        s = PcStats()
        if lib().pcxr_get_stats(self._h, ctypes.byref(s)) != 0:
            raise RuntimeError(last_error())
        return s

    def clear_stats(self):
        lib().pcxr_clear_stats(self._h)

    def probe(self, level):
        """Time the delay-line-sum kernel of one level (>= 1; -1 disables)."""
        if lib().pcxr_probe(self._h, int(level)) != 0:
            raise RuntimeError(last_error())

    def probe_read(self):
        """(avg_us, launches, algorithmic_bytes_per_launch) of the probed kernel."""
        a, b = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        if lib().pcxr_probe_read(self._h, ctypes.byref(a), ctypes.byref(n), ctypes.byref(b)) != 0:
            raise RuntimeError(last_error())
        return a.value, n.value, b.value

    def set_release_window(self, ticks):
        """Ticks run_device may release ahead of the oldest incomplete one
        (pcxr_set_release_window: 1 .. 3, default 1); identical output."""
        if lib().pcxr_set_release_window(self._h, int(ticks)) != 0:
            raise RuntimeError(last_error())

    def stall_level(self, level, seconds):
        """Deadline-handling test hook (pcxr_stall_level): the next block of
        level `level` waits `seconds` behind a host function on its stream."""
        if lib().pcxr_stall_level(self._h, int(level), float(seconds)) != 0:
            raise RuntimeError(last_error())

    def partition(self):
        N = (ctypes.c_int * 16)()
        P = (ctypes.c_int * 16)()
        M = lib().pcxr_partition(self._h, N, P, 16)
        return list(N)[:M], list(P)[:M]

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcxr_destroy", h)
        except Exception:
            pass


def _throughput_demo():
    """The MFMA throughput form at 32 x 32 pairs, 65536-tap filters,
    [256 x 30, 4096 x 15], 172032 samples per call (three calls): a fixed
    workload for hardware counters (rocprofv3 --pmc MfmaUtil MfmaFlopsF32 --
    python -u -m partconv_amd.mimo)."""
    import time
    n, taps, T = 32, 65536, 172032
    g = torch.Generator(device="cuda").manual_seed(4)
    h = torch.randn((n, n, taps), device="cuda", generator=g) * (1.0 / 256)
    x = torch.randn((n, T), device="cuda", generator=g)
    mx = MimoConv([256, 4096], [30, 0], taps, n, n)
    mx.set_filters(h)
    mx(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        mx(x)
    torch.cuda.synchronize()
    print(f"throughput form: {1e3 * (time.perf_counter() - t0) / 3:.3f} ms per call of {T} samples x {n} outputs",
          flush=True)


if __name__ == "__main__":
    _throughput_demo()
