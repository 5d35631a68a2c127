"""MIMO throughput form on the matrix cores (pcx_* entry points).

    mx = MimoConv([256, 4096], [30, 0], taps=65536, n_in=32, n_out=32)
    mx.set_filters(h)        # h: [n_out, n_in, taps] float32 cuda
    y = mx(x)                # x: [n_in, T] -> y: [n_out, T]

y[o] = sum_i conv(x[i], h[o][i]) over whole device-resident signals: every
frequency bin's complex n_out x n_in product, for all output blocks at once,
runs on v_mfma_f32_32x32x2f32.  The real-time form of the same computation is
PartConvMulti with one instance per (input, output) pair.
"""
from __future__ import annotations

import ctypes

import torch

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
        try:
            if self._h:
                lib().pcx_destroy(self._h)
                self._h = None
        except Exception:
            pass
