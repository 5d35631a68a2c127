"""Differentiable partitioned convolution on the GPU (torch.autograd.Function
over the pcc_* entry points of libpartconv_amd.so).

    conv = PartitionedConv([256, 2048, 16384], [14, 14, 0], taps=65536)
    y = conv(x, h)          # x: [C, T] float32 cuda, h: [taps] float32 cuda
    loss = ((y - target) ** 2).mean(); loss.backward()
    loss = conv.mse_loss(x, h, target); loss.backward()   # the same, fused

Forward: y[c] = conv(x[c], h) truncated to T samples, computed as the
reference's multi-level partitioned convolution (partition rules of
src/partconv.cpp:86-182, filter spectra of src/partconv.cpp:199-230) over every
block of every clip at once.  Backward is the hand-written adjoint (pcc_backward):
dL/dx is the delay-line product with conjugated partition spectra run in
reverse block order, dL/dh_p the conjugated input spectrum times the output-
gradient spectrum accumulated over blocks and clips.  Input spectra are kept
from forward for backward while they fit `budget_bytes`, else recomputed from
the stored time-domain input.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import last_error, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class PartitionedConv:
    """Handle on one partition / tap count; call it like a function."""

    def __init__(self, N, P, taps, budget_bytes=1 << 30):
        n = len(N)
        Na = (ctypes.c_int * n)(*[int(v) for v in N])
        Pa = (ctypes.c_int * n)(*[int(v) for v in P])
        self._h = lib().pcc_create(n, Na, Pa, int(taps))
        if not self._h:
            raise ValueError(last_error())
        self.taps = int(taps)
        self.set_budget(budget_bytes)
        No = (ctypes.c_int * 16)()
        Po = (ctypes.c_int * 16)()
        M = lib().pcc_levels(self._h, No, Po, 16)
        self.N, self.P = list(No)[:M], list(Po)[:M]
        # generation of the filter spectra held by the handle: bumped by every
        # set_filter, so a backward can tell whether the spectra are still the
        # ones its own forward computed
        self._filter_gen = 0

    def set_budget(self, nbytes):
        lib().pcc_set_budget(self._h, int(nbytes))

    @property
    def recomputed_levels(self):
        return lib().pcc_recomputed_levels(self._h)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(last_error())

    def tune(self, x, h):
        """Choose each level's chunk sizes by timing candidates on clips x [C, T]
        with filter h; returns [(mac_chunk, dh_chunk)] per level."""
        x = _prep(x, "x")
        self.set_filter(_prep(h, "h"))
        torch.cuda.synchronize()
        self._check(lib().pcc_tune(self._h, _p(x), x.shape[0], x.shape[1], _stream()))
        return self.tuning()

    LEVEL_STREAMS = {"off": 0, "on": 1, "captured": 2}

    def set_level_streams(self, mode):
        """Run each level on its own stream: "on", "off" (every level on the
        caller's stream) or "captured" (default: only while the caller's
        stream is captured into a graph).  True / False mean "on" / "off".
        Results are bitwise the same in every mode."""
        if isinstance(mode, bool):
            mode = "on" if mode else "off"
        self._check(lib().pcc_set_level_streams(self._h, self.LEVEL_STREAMS[mode]))

    def set_call_graphs(self, on):
        """Replay each forward / backward call with the same arguments as an
        earlier one from a HIP graph of its launches (default on; bitwise the
        same results)."""
        self._check(lib().pcc_set_call_graphs(self._h, 1 if on else 0))

    def call_graph_stats(self):
        """{"forward": (replays, captures), "backward": (replays, captures)}."""
        v = [ctypes.c_longlong() for _ in range(4)]
        self._check(lib().pcc_call_graph_stats(self._h, *[ctypes.byref(x) for x in v]))
        return {"forward": (v[0].value, v[1].value), "backward": (v[2].value, v[3].value)}

    def tuning(self):
        """[(mac_chunk, dh_chunk)] per level: the chunk sizes in use."""
        out = []
        for l in range(len(self.N)):
            a, b = ctypes.c_int(), ctypes.c_int()
            self._check(lib().pcc_level_tuning(self._h, l, ctypes.byref(a), ctypes.byref(b)))
            out.append((a.value, b.value))
        return out

    PROBE_KINDS = {"product": 0, "forward_transform": 1, "inverse_transform": 2, "dh_partial": 3}

    def probe(self, level, kind="product"):
        """Time one kernel of one level with HIP events stamped by its launch
        (-1 disables): the forward spectral product ("product"), the forward's
        input transform ("forward_transform"), its inverse transform into y
        ("inverse_transform"), or the backward's dL/dh partial sums
        ("dh_partial")."""
        self._check(lib().pcc_probe_kernel(self._h, int(level), self.PROBE_KINDS[kind]))

    def probe_read(self):
        """(avg_us, launches, algorithmic_bytes_per_launch) of the probed kernel."""
        a, b = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(lib().pcc_probe_read(self._h, ctypes.byref(a), ctypes.byref(n), ctypes.byref(b)))
        return a.value, n.value, b.value

    def set_filter(self, h):
        """Filter spectra from h (always recomputed: a tensor's address and
        version do not identify its contents, since the caching allocator hands
        a freed block to the next tensor of the same size).  Returns the new
        spectra generation."""
        self._check(lib().pcc_set_filter(self._h, _p(h), _stream()))
        self._filter_gen += 1
        return self._filter_gen

    def forward_raw(self, x, h, y):
        """Forward with fresh filter spectra; returns their generation."""
        gen = self.set_filter(h)
        self._check(lib().pcc_forward(self._h, _p(x), _p(y), x.shape[0], x.shape[1], _stream()))
        return gen

    def backward_raw(self, x, h, g, dx, dh, gen=None, scale=None):
        """Adjoint for upstream gradient g (times the one-element device tensor
        `scale` when given).  The spectra are recomputed unless they are still
        generation `gen` (the forward that saved h computed them and no
        set_filter ran since; autograd rejects a saved h that was modified in
        place).  Under graph capture they are always recomputed, so a replay
        sees the filter's values at replay time."""
        if gen is None or gen != self._filter_gen or torch.cuda.is_current_stream_capturing():
            self.set_filter(h)
        odx = None if dx is None else _p(dx)
        odh = None if dh is None else _p(dh)
        if scale is None:
            self._check(lib().pcc_backward(self._h, _p(x), _p(g), odx, odh, x.shape[0], x.shape[1], _stream()))
        else:
            self._check(lib().pcc_backward_scaled(self._h, _p(x), _p(g), _p(scale), odx, odh, x.shape[0],
                                                  x.shape[1], _stream()))

    def mse_loss(self, x, h, target):
        """mean((conv(x, h) - target) ** 2) as one differentiable call: the loss
        and its gradient come from one fused pass over y and the target
        (pcc_mse), and the backward scales that gradient by the loss's upstream
        gradient inside its transforms (pcc_backward_scaled)."""
        return PartConvMSEFunction.apply(x, h, target, self)

    def __call__(self, x, h):
        return PartConvFunction.apply(x, h, self)

    def __del__(self):
        try:  # module globals may already be gone at interpreter exit
            h, self._h = self.__dict__.get("_h"), None
            _lib.destroy("pcc_destroy", h)
        except Exception:
            pass


def _prep(t, shape_msg):
    if not (t.is_cuda and t.dtype == torch.float32):
        raise TypeError(f"{shape_msg}: float32 CUDA tensor required")
    return t.contiguous()


class PartConvFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, h, conv):
        x = _prep(x, "x")
        h = _prep(h, "h")
        if x.dim() != 2 or h.dim() != 1 or h.numel() != conv.taps:
            raise ValueError("x must be [C, T] and h [taps]")
        y = torch.empty_like(x)
        ctx.filter_gen = conv.forward_raw(x, h, y)
        ctx.save_for_backward(x, h)
        ctx.conv = conv
        return y

    @staticmethod
    def backward(ctx, g):
        x, h = ctx.saved_tensors
        conv = ctx.conv
        g = g.contiguous()
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        dh = torch.empty_like(h) if ctx.needs_input_grad[1] else None
        if dx is not None or dh is not None:
            conv.backward_raw(x, h, g, dx, dh, ctx.filter_gen)
        return dx, dh, None


class PartConvMSEFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, h, target, conv):
        x = _prep(x, "x")
        h = _prep(h, "h")
        target = _prep(target, "target")
        if x.dim() != 2 or h.dim() != 1 or h.numel() != conv.taps or target.shape != x.shape:
            raise ValueError("x and target must be [C, T] and h [taps]")
        y = torch.empty_like(x)
        ctx.filter_gen = conv.forward_raw(x, h, y)
        g = torch.empty_like(x)
        loss = torch.empty((), dtype=torch.float32, device=x.device)
        conv._check(lib().pcc_mse(conv._h, _p(y), _p(target), _p(g), _p(loss), x.shape[0], x.shape[1],
                                  _stream()))
        ctx.save_for_backward(x, h, g)
        ctx.conv = conv
        return loss

    @staticmethod
    def backward(ctx, gl):
        x, h, g = ctx.saved_tensors
        conv = ctx.conv
        gl = gl.reshape(1).contiguous().float()
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        dh = torch.empty_like(h) if ctx.needs_input_grad[1] else None
        if dx is not None or dh is not None:
            conv.backward_raw(x, h, g, dx, dh, ctx.filter_gen, scale=gl)
        dt = -(g * gl) if ctx.needs_input_grad[2] else None
        return dx, dh, dt, None
