"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Wraps oracle/lib/libpco_{f32,f64}.so (built from oracle/pc_oracle.c by
oracle/Makefile) and adds the fp64 direct-convolution reference.  Only tests,
the smoke check and the benchmark's CPU-baseline leg may import this module.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIBDIR = HERE / "lib"

_c_int_p = ctypes.POINTER(ctypes.c_int)
_libs = {}


def build():
    """Compile the C restatement (both precisions) with make."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib(prec="f32"):
    if prec not in _libs:
        path = LIBDIR / f"libpco_{prec}.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        vp = ctypes.c_void_p
        L.pco_filter_setup.restype = vp
        L.pco_filter_setup.argtypes = [ctypes.c_int, _c_int_p, _c_int_p, vp, ctypes.c_int]
        for name in ("pco_filter_run_level",):
            getattr(L, name).argtypes = [vp, ctypes.c_int]
        for name in ("pco_filter_sync_levels", "pco_filter_sync_levels_relaxed",
                     "pco_filter_reset", "pco_filter_free"):
            getattr(L, name).argtypes = [vp]
        for name in ("pco_filter_input", "pco_filter_output", "pco_filter_output_relaxed"):
            getattr(L, name).argtypes = [vp, vp]
        L.pco_filter_levels.argtypes = [vp]
        L.pco_filter_partition.argtypes = [vp, _c_int_p, _c_int_p, _c_int_p]
        L.pco_filter_get_H.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.pco_filter_process.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int]
        L.pco_filter_start_depth.argtypes = [vp]
        L.pco_filter_end_depth.argtypes = [vp]
        L.pco_multi_setup.restype = vp
        L.pco_multi_setup.argtypes = [ctypes.c_int] * 5 + [_c_int_p, _c_int_p, _c_int_p,
                                                           _c_int_p, _c_int_p, _c_int_p,
                                                           vp, vp, _c_int_p, ctypes.c_int]
        L.pco_multi_run.argtypes = [vp, vp, vp, ctypes.c_int]
        L.pco_multi_process.argtypes = [vp, vp, vp, ctypes.c_int]
        L.pco_multi_reset.argtypes = [vp]
        L.pco_multi_free.argtypes = [vp]
        L.pco_multi_num_levels.argtypes = [vp]
        L.pco_multi_level_map.argtypes = [vp, _c_int_p, vp]
        L.pco_rfft.argtypes = [ctypes.c_int, vp, vp]
        L.pco_irfft.argtypes = [ctypes.c_int, vp, vp]
        L.pco_last_error.restype = ctypes.c_char_p
        L.pco_set_mutation.argtypes = [ctypes.c_int]
        L.pco_now.restype = ctypes.c_double
        L.pco_rt_run.argtypes = [vp, ctypes.c_int, _c_int_p, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int,
                                 vp, vp, vp]
        _libs[prec] = L
    return _libs[prec]


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _ints(seq):
    a = (ctypes.c_int * max(1, len(seq)))(*[int(v) for v in seq])
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_mutation(m, prec="f32"):
    """Plant a known mistake in the oracle (0 = none).  1: FDL slot off by one;
    2: conjugated filter spectrum in the MAC; 3: H scaled by 1/N not 1/2N;
    4: strict output reads sub-block ind instead of ind+1; 5: taps placed at
    offset 0 of the 2N row instead of offset N."""
    lib(prec).pco_set_mutation(int(m))


def rfft(x, prec="f64"):
    x = np.ascontiguousarray(x, dtype=_dt(prec))
    N = x.size // 2
    X = np.zeros(2 * (N + 1), dtype=_dt(prec))
    lib(prec).pco_rfft(N, _ptr(x), _ptr(X))
    return X[0::2] + 1j * X[1::2]


def irfft(X, prec="f64"):
    X = np.asarray(X)
    N = X.size - 1
    Xi = np.zeros(2 * (N + 1), dtype=_dt(prec))
    Xi[0::2] = X.real
    Xi[1::2] = X.imag
    x = np.zeros(2 * N, dtype=_dt(prec))
    lib(prec).pco_irfft(N, _ptr(Xi), _ptr(x))
    return x


class OracleFilter:
    """Restated PartConvFilter (src/partconv.cpp:70-648)."""

    def __init__(self, N, P, impulse=None, prec="f32"):
        self.prec = prec
        self.L = lib(prec)
        self._h = None if impulse is None else np.ascontiguousarray(impulse, dtype=_dt(prec))
        ptr = None if self._h is None else _ptr(self._h)
        n = 0 if self._h is None else self._h.size
        self.h = self.L.pco_filter_setup(len(N), _ints(N), _ints(P), ptr, n)
        if not self.h:
            raise ValueError(self.L.pco_last_error().decode())
        M = self.L.pco_filter_levels(self.h)
        a, b, c = _ints([0] * M), _ints([0] * M), _ints([0] * M)
        self.L.pco_filter_partition(self.h, a, b, c)
        self.N, self.P, self.delay = list(a), list(b), list(c)
        self.M = M

    def H(self, L, p):
        C = self.N[L] + 1
        out = np.zeros(2 * C, dtype=_dt(self.prec))
        self.L.pco_filter_get_H(self.h, L, p, _ptr(out))
        return out[0::2] + 1j * out[1::2]

    def process(self, x, relaxed=False):
        N0 = self.N[0]
        nfr = -(-len(x) // N0)
        xin = np.zeros(nfr * N0, dtype=_dt(self.prec))
        xin[:len(x)] = x
        y = np.zeros_like(xin)
        if self.L.pco_filter_process(self.h, _ptr(xin), _ptr(y), nfr, int(relaxed)):
            raise RuntimeError(self.L.pco_last_error().decode())
        return y

    def reset(self):
        self.L.pco_filter_reset(self.h)

    def close(self):
        if self.h:
            self.L.pco_filter_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rt_run(filters, x, threads_per_level, in_ch=None):
    """The reference's threaded schedule over OracleFilter instances with one
    buffer size (pco_rt_run): instance i reads column in_ch[i] (default i) of
    x [nticks*N0, nIn] and writes column i of the result.  Returns (y, per-tick
    seconds, [(block size, mean seconds of one runLevel of one instance)] per
    global level).  Timing leg of the CPU baseline; the output equals the
    single-threaded restatement's."""
    L = filters[0].L
    n, N0 = len(filters), filters[0].N[0]
    dt = _dt(filters[0].prec)
    x = np.ascontiguousarray(x, dtype=dt)
    if x.ndim == 1:
        x = x[:, None]
    nticks = x.shape[0] // N0
    y = np.zeros((nticks * N0, n), dtype=dt)
    lat = np.zeros(max(1, nticks))
    gN = np.zeros(24, dtype=np.int32)
    lvl = np.zeros(24)
    hs = (ctypes.c_void_p * n)(*[f.h for f in filters])
    ch = None if in_ch is None else _ints(in_ch)
    G = L.pco_rt_run(ctypes.cast(hs, ctypes.c_void_p), n, ch, x.shape[1], _ptr(x), _ptr(y), nticks,
                     int(threads_per_level), _ptr(lat), _ptr(gN), _ptr(lvl))
    if G < 0:
        raise RuntimeError("pco_rt_run: bad instance set (one buffer size, input columns in range)")
    return y, lat[:nticks], [(int(gN[g]), float(lvl[g])) for g in range(G)]


class OracleMulti:
    """Restated PartConvMulti / PartConvMultiRelaxed (src/partconvMulti.cpp)."""

    def __init__(self, FS, buffer_size, n_in, n_out, instances, irs, relaxed=False, prec="f32"):
        """instances: list of dicts with keys N, P, in_ch, out_ch, mute, scale."""
        self.prec = prec
        self.L = lib(prec)
        dt = _dt(prec)
        self.n_in, self.n_out, self.N0 = n_in, n_out, buffer_size
        self._irs = [np.ascontiguousarray(h, dtype=dt) for h in irs]
        M = [len(d["N"]) for d in instances]
        Nf = [n for d in instances for n in d["N"]]
        Pf = [p for d in instances for p in d["P"]]
        ptrs = (ctypes.c_void_p * len(self._irs))(*[h.ctypes.data for h in self._irs])
        scale = np.array([d.get("scale", 1.0) for d in instances], dtype=dt)
        self._scale = scale
        self.h = self.L.pco_multi_setup(
            FS, buffer_size, n_in, n_out, len(instances), _ints(M), _ints(Nf), _ints(Pf),
            _ints([d.get("in_ch", 0) for d in instances]),
            _ints([d.get("out_ch", 0) for d in instances]),
            _ints([d.get("mute", 0) for d in instances]),
            _ptr(scale), ptrs, _ints([h.size for h in self._irs]), int(relaxed))
        if not self.h:
            raise ValueError(self.L.pco_last_error().decode())
        self.num_levels = self.L.pco_multi_num_levels(self.h)

    def level_map(self):
        n = len(self._irs)
        m = _ints([0] * (n * self.num_levels))
        s = (ctypes.c_uint * self.num_levels)()
        self.L.pco_multi_level_map(self.h, m, s)
        return list(s), [list(m)[i * self.num_levels:(i + 1) * self.num_levels] for i in range(n)]

    def process(self, x):
        """x: (n_samples, n_in) -> (n_frames*N0, n_out), zero-padded to whole frames."""
        x = np.asarray(x)
        if x.ndim == 1:
            x = x[:, None]
        nfr = -(-x.shape[0] // self.N0)
        xin = np.zeros((nfr * self.N0, self.n_in), dtype=_dt(self.prec))
        xin[:x.shape[0]] = x
        y = np.zeros((nfr * self.N0, self.n_out), dtype=_dt(self.prec))
        if self.L.pco_multi_process(self.h, _ptr(y), _ptr(xin), nfr):
            raise RuntimeError(self.L.pco_last_error().decode())
        return y

    def process_frames(self, x, nframes):
        """Time nframes ticks over interleaved x (nframes*N0, n_in); returns
        (y, seconds)."""
        y = np.zeros((nframes * self.N0, self.n_out), dtype=_dt(self.prec))
        xin = np.ascontiguousarray(x[:nframes * self.N0], dtype=_dt(self.prec))
        t0 = self.L.pco_now()
        if self.L.pco_multi_process(self.h, _ptr(y), _ptr(xin), nframes):
            raise RuntimeError(self.L.pco_last_error().decode())
        return y, self.L.pco_now() - t0

    def reset(self):
        self.L.pco_multi_reset(self.h)

    def close(self):
        if self.h:
            self.L.pco_multi_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def direct_conv_f64(x, h, n_out=None):
    """Linear convolution in float64 (zero-padded FFT; exact to ~1e-15 relative).
    Same mathematics as the reference's naive_convolution (src/sigproc_util.c:464-482)."""
    x = np.asarray(x, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    n = x.shape[0] + h.shape[0] - 1
    nfft = 1 << (n - 1).bit_length()
    y = np.fft.irfft(np.fft.rfft(x, nfft, axis=0) * (np.fft.rfft(h, nfft) if x.ndim == 1
                                                    else np.fft.rfft(h, nfft)[:, None]),
                     nfft, axis=0)
    n_out = x.shape[0] if n_out is None else n_out
    return y[:n_out]


def brute_conv(x, h, n_out=None):
    """Direct O(len(x)*len(h)) float64 convolution (tiny inputs only)."""
    y = np.convolve(np.asarray(x, dtype=np.float64), np.asarray(h, dtype=np.float64))
    n_out = len(x) if n_out is None else n_out
    return y[:n_out]
