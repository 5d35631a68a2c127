"""GPU tests of the differentiable form (forward + hand-written adjoint):
forward against fp64 direct convolution, gradients against the fp64 direct
gradient (dL/dx = correlation of g with h, dL/dh = sum_c correlation of g_c
with x_c) and against fp64 finite differences of the MSE loss on random taps.
Bars: forward 1e-5 of output RMS; gradients 1e-4 relative."""
import numpy as np
import pytest

from oracle import oracle as O
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu


def f64_corr_x(g, h):
    """dx[s] = sum_j h[j] g[s + j] (per clip)."""
    C, T = g.shape
    L = 1 << int(np.ceil(np.log2(T + h.size)))
    Gf = np.fft.rfft(g.astype(np.float64), L, axis=1)
    Hf = np.fft.rfft(h.astype(np.float64), L)
    return np.fft.irfft(Gf * np.conj(Hf)[None, :], L, axis=1)[:, :T]


def f64_corr_h(g, x, taps):
    """dh[j] = sum_c sum_t g[c,t] x[c, t - j]."""
    C, T = g.shape
    L = 1 << int(np.ceil(np.log2(2 * T + taps)))
    Gf = np.fft.rfft(g.astype(np.float64), L, axis=1)
    Xf = np.fft.rfft(x.astype(np.float64), L, axis=1)
    r = np.fft.irfft(np.sum(Gf * np.conj(Xf), axis=0), L)
    return r[:taps]


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("N,P,taps,T,C", [
    ([256, 2048, 16384], [14, 14, 0], 65536, 40000, 3),
    ([64, 512, 4096], [14, 14, 0], 20000, 15000, 4),
    ([256, 2048], [14, 0], 9000, 8000, 2),
    ([128, 512], [6, 0], 4000, 7000, 2),        # 5..8 partitions per level
])
def test_forward_and_gradients_match_fp64(N, P, taps, T, C):
    import torch
    from partconv_amd.autograd import PartitionedConv
    print(f"case N={N} taps={taps} T={T} C={C}", flush=True)
    rng = np.random.default_rng(taps)
    x = rng.standard_normal((C, T)).astype(np.float32)
    h = decaying_noise_ir(taps, 0.8, 48000, 7)
    g = rng.standard_normal((C, T)).astype(np.float32)
    conv = PartitionedConv(N, P, taps)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    ht = torch.from_numpy(h).cuda().requires_grad_(True)
    y = conv(xt, ht)
    y.backward(torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    yr = np.stack([O.direct_conv_f64(x[c], h) for c in range(C)])
    err = float(np.max(np.abs(y.detach().cpu().numpy() - yr)) / np.sqrt(np.mean(yr ** 2)))
    assert err < 1e-5, err
    dx = xt.grad.cpu().numpy()
    dh = ht.grad.cpu().numpy()
    assert relmax(dx, f64_corr_x(g, h)) < 1e-4
    assert relmax(dh, f64_corr_h(g, x, taps)) < 1e-4


def test_mse_gradient_finite_differences_and_recompute_mode():
    """Config-5 shape at reduced batch: learnable IR (RT60 0.8 s), hidden IR
    (RT60 1.5 s), targets = conv(x, hidden) + N(0, 1e-3^2) noise, MSE loss.
    dL/dh on 64 random taps against central fp64 finite differences of the
    fp64 loss, and the recompute-spectra mode against the kept-spectra mode."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps = [256, 2048, 16384], [14, 14, 0], 65536
    C, T = 4, 32768
    rng = np.random.default_rng(5)
    x = rng.standard_normal((C, T)).astype(np.float32)
    h0 = decaying_noise_ir(taps, 0.8, 48000, 11)
    hs = decaying_noise_ir(taps, 1.5, 48000, 12)
    tgt = (np.stack([O.direct_conv_f64(x[c], hs) for c in range(C)])
           + 1e-3 * rng.standard_normal((C, T))).astype(np.float32)

    def loss64(h):
        y = np.stack([O.direct_conv_f64(x[c], h) for c in range(C)])
        return np.mean((y - tgt.astype(np.float64)) ** 2)

    grads = []
    for budget in (1 << 30, 0):
        conv = PartitionedConv(N, P, taps, budget_bytes=budget)
        xt = torch.from_numpy(x).cuda()
        ht = torch.from_numpy(h0).cuda().requires_grad_(True)
        loss = torch.mean((conv(xt, ht) - torch.from_numpy(tgt).cuda()) ** 2)
        loss.backward()
        torch.cuda.synchronize()
        grads.append(ht.grad.cpu().numpy().astype(np.float64))
        if budget == 0:
            assert conv.recomputed_levels == len(conv.N)
    assert np.array_equal(grads[0], grads[1])
    # fp64 direct gradient of the same loss
    y64 = np.stack([O.direct_conv_f64(x[c], h0) for c in range(C)])
    g64 = 2.0 * (y64 - tgt) / (C * T)
    gd = f64_corr_h(g64, x, taps)
    assert relmax(grads[0], gd) < 1e-4
    # central finite differences on 64 random taps (fp64 loss)
    idx = rng.choice(taps, size=64, replace=False)
    eps = 1e-3
    fd = []
    for j in idx:
        hp = h0.astype(np.float64).copy()
        hm = hp.copy()
        hp[j] += eps
        hm[j] -= eps
        fd.append((loss64(hp) - loss64(hm)) / (2 * eps))
    fd = np.array(fd)
    assert np.max(np.abs(grads[0][idx] - fd)) / np.max(np.abs(fd)) < 1e-4


def test_long_ir_recompute_mode():
    """2^20-tap IR, partition [256x14, 2048x14, 16384x63], spectra recomputed in
    backward (budget 0): forward and dL/dh against fp64."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps = [256, 2048, 16384], [14, 14, 0], 1 << 20
    C, T = 2, 65536
    rng = np.random.default_rng(6)
    x = rng.standard_normal((C, T)).astype(np.float32)
    h = decaying_noise_ir(taps, 3.0, 48000, 13)
    g = rng.standard_normal((C, T)).astype(np.float32)
    conv = PartitionedConv(N, P, taps, budget_bytes=0)
    assert conv.P[-1] == 63
    xt = torch.from_numpy(x).cuda()
    ht = torch.from_numpy(h).cuda().requires_grad_(True)
    y = conv(xt, ht)
    y.backward(torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    yr = np.stack([O.direct_conv_f64(x[c], h) for c in range(C)])
    assert float(np.max(np.abs(y.detach().cpu().numpy() - yr)) / np.sqrt(np.mean(yr ** 2))) < 1e-5
    # taps beyond T never meet the input: their gradient is exactly zero
    dh = ht.grad.cpu().numpy()
    assert relmax(dh, f64_corr_h(g, x, taps)) < 1e-4
    assert conv.recomputed_levels == 3


def test_probe_times_the_forward_product():
    """pcc_probe: one timed launch per forward of the probed level, the unique
    bytes of that product (block spectra in, products out, filter spectra),
    and a probed forward equal to an unprobed one."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    C, T, taps = 3, 20000, 9000
    conv = PartitionedConv([256, 2048], [14, 0], taps)
    x = torch.randn((C, T), device="cuda")
    h = torch.from_numpy(decaying_noise_ir(taps, 0.3, 48000, 9)).cuda()
    with torch.no_grad():
        y0 = conv(x, h)
        conv.probe(0)
        for _ in range(3):
            y1 = conv(x, h)
        avg_us, n, nbytes = conv.probe_read()
        conv.probe(-1)
        y2 = conv(x, h)
    assert n == 3 and avg_us > 0
    N, P = conv.N[0], conv.P[0]
    M = -(-T // N)  # blocks of the 256 level covering the clip
    assert nbytes == 2 * 8 * C * M * N + 8 * P * N, (nbytes, M)
    assert torch.equal(y0, y1) and torch.equal(y0, y2)
    with pytest.raises(RuntimeError):
        conv.probe(5)


def test_probe_kernel_kinds():
    """pcc_probe_kernel: the forward transform, the inverse transform and the
    dL/dh partial sums of a level are timed once per call, with their
    algorithmic bytes (signal samples and spectra read or written once); the
    four-step level (32768) is timed across its two kernels; probing changes
    no result."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    C, T, taps = 2, 70000, 70000
    conv = PartitionedConv([256, 2048, 32768], [14, 30, 0], taps)
    assert conv.N[2] == 32768
    x = torch.randn((C, T), device="cuda")
    h = torch.from_numpy(decaying_noise_ir(taps, 0.3, 48000, 10)).cuda().requires_grad_(True)
    g = torch.randn((C, T), device="cuda")

    def step():
        h.grad = None
        y = conv(x, h)
        y.backward(g)
        return y.detach().clone(), h.grad.clone()

    y0, d0 = step()
    for level in (0, 1, 2):
        N, P = conv.N[level], conv.P[level]
        M = -(-T // N)
        sig, spec = 4 * C * T, 8 * C * M * N
        for kind, want in (("forward_transform", sig + spec),
                           ("inverse_transform", spec + (sig if level == 0 else 2 * sig)),
                           ("product", 2 * spec + 8 * P * N),
                           ("dh_partial", None)):
            conv.probe(level, kind)
            for _ in range(3):
                y1, d1 = step()
            avg_us, n, nbytes = conv.probe_read()
            conv.probe(-1)
            assert n == 3 and avg_us > 0, (level, kind, n)
            if want is None:
                dh_mc = conv.tuning()[level][1]
                want = 2 * spec + 8 * P * C * (-(-M // dh_mc)) * N
            assert nbytes == want, (level, kind, nbytes, want)
            assert torch.equal(y0, y1) and torch.equal(d0, d1), (level, kind)
    with pytest.raises(KeyError):
        conv.probe(0, "no such kernel")


@pytest.mark.parametrize("budget", [1 << 30, 0])
def test_level_streams_bitwise_equal_to_one_stream(budget):
    """pcc_set_level_streams: levels on their own streams give y, dL/dx and
    dL/dh bitwise equal to every level on the caller's stream (and to the
    default mode, which is one stream outside graph capture), with kept or
    recomputed spectra and with a four-step level (32768)."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    C, T, taps = 3, 90000, 70000
    conv = PartitionedConv([256, 2048, 32768], [14, 30, 0], taps, budget_bytes=budget)
    x = torch.randn((C, T), device="cuda")
    g = torch.randn((C, T), device="cuda")
    h = torch.from_numpy(decaying_noise_ir(taps, 0.3, 48000, 11)).cuda()
    out = {}
    for on in ("off", "on", "off", "on", "captured"):
        conv.set_level_streams(on)
        xg = x.clone().requires_grad_(True)
        hg = h.clone().requires_grad_(True)
        y = conv(xg, hg)
        y.backward(g)
        torch.cuda.synchronize()
        r = (y.detach().clone(), xg.grad.clone(), hg.grad.clone())
        if on in out:
            assert all(torch.equal(a, b) for a, b in zip(out[on], r)), on
        out[on] = r
    assert all(torch.equal(a, b) for a, b in zip(out["off"], out["on"]))
    assert all(torch.equal(a, b) for a, b in zip(out["off"], out["captured"]))
    with pytest.raises(KeyError):
        conv.set_level_streams("sometimes")
    # and against FP64
    xn, hn, gn = x.cpu().numpy(), h.cpu().numpy(), g.cpu().numpy()
    yr = np.stack([O.direct_conv_f64(xn[c], hn) for c in range(C)])
    y = out["on"][0].cpu().numpy()
    assert float(np.max(np.abs(y - yr)) / np.sqrt(np.mean(yr ** 2))) < 1e-5
    assert relmax(out["on"][1].cpu().numpy(), f64_corr_x(gn, hn)) < 1e-4
    assert relmax(out["on"][2].cpu().numpy(), f64_corr_h(gn, xn, taps)) < 1e-4


@pytest.mark.parametrize("budget", [1 << 30, 0])
def test_call_graphs_bitwise_equal_to_eager(budget):
    """pcc_set_call_graphs: forward and adjoint calls replayed from the
    library's own graphs (captured on the second call with the same
    arguments) give y, dL/dx and dL/dh bitwise equal to eager calls on the
    same values, with new values written into the same tensors between calls,
    with pointers alternating between two sets (two graphs per direction),
    after a level-stream mode change, a smaller clip count and a larger one
    (buffers grown: the old graphs are never replayed)."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    taps = 70000
    N, P = [256, 2048, 32768], [14, 30, 0]
    conv = PartitionedConv(N, P, taps, budget_bytes=budget)
    ref = PartitionedConv(N, P, taps, budget_bytes=budget)
    ref.set_call_graphs(False)

    def call(cv, x, h, g, y, dx, dh):
        gen = cv.forward_raw(x, h, y)
        cv.backward_raw(x, h, g, dx, dh, gen)

    held = []  # every buffer set stays alive, so no address is reused

    def bufs(C, T):
        held.append([torch.empty((C, T), device="cuda") for _ in range(4)] + [torch.empty(taps, device="cuda")])
        return held[-1]

    def check(C, T, sets, calls):
        h = torch.empty(taps, device="cuda")
        for k in range(calls):
            x, g, y, dx, dh = sets[k % len(sets)]
            x.copy_(torch.randn((C, T), device="cuda"))
            g.copy_(torch.randn((C, T), device="cuda"))
            h.copy_(torch.from_numpy(decaying_noise_ir(taps, 0.3, 48000, 11 + k)).cuda())
            call(conv, x, h, g, y, dx, dh)
            r = [torch.empty_like(t) for t in (y, dx, dh)]
            call(ref, x, h, g, *r)
            torch.cuda.synchronize()
            for a, b, name in zip((y, dx, dh), r, ("y", "dx", "dh")):
                assert torch.equal(a, b), (C, T, k, name)

    C, T = 3, 90000
    check(C, T, [bufs(C, T)], 5)
    st = conv.call_graph_stats()
    assert st == {"forward": (3, 1), "backward": (3, 1)}, st
    check(C, T, [bufs(C, T), bufs(C, T)], 8)  # alternating: each set captured on its second call
    st = conv.call_graph_stats()
    assert st == {"forward": (7, 3), "backward": (7, 3)}, st
    conv.set_level_streams("on")
    check(C, T, [bufs(C, T)], 3)
    conv.set_level_streams("captured")
    check(2, 50000, [bufs(2, 50000)], 3)
    check(4, 120000, [bufs(4, 120000)], 3)
    conv.set_call_graphs(False)
    before = conv.call_graph_stats()
    check(C, T, [bufs(C, T)], 3)
    assert conv.call_graph_stats() == before


def test_forward_and_adjoint_replay_in_a_cuda_graph():
    """A step (forward + adjoint, dL/dx and dL/dh) captured once in a HIP graph
    and replayed with new inputs, output gradients and filter values written
    into the captured tensors between replays: every replay equals the eager
    computation on the same values, bitwise (the library makes no allocation,
    copy or synchronisation once its buffers are sized, and clears with
    kernels, not memsets)."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    C, T, taps = 2, 6000, 3000
    conv = PartitionedConv([128, 512], [6, 0], taps)
    rng = torch.Generator(device="cuda").manual_seed(11)
    xs = torch.randn((C, T), device="cuda", generator=rng).requires_grad_(True)
    hs = torch.randn(taps, device="cuda", generator=rng).requires_grad_(True)
    gs = torch.randn((C, T), device="cuda", generator=rng)

    def step():
        y = conv(xs, hs)
        y.backward(gs)
        return y

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):  # sizes the work space outside the capture
            xs.grad = hs.grad = None
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    xs.grad = hs.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        ys = step()
    ref = PartitionedConv([128, 512], [6, 0], taps)
    for r in range(3):
        with torch.no_grad():
            xs.copy_(torch.randn((C, T), device="cuda", generator=rng))
            hs.copy_(torch.randn(taps, device="cuda", generator=rng))
            gs.copy_(torch.randn((C, T), device="cuda", generator=rng))
        graph.replay()
        torch.cuda.synchronize()
        xe = xs.detach().clone().requires_grad_(True)
        he = hs.detach().clone().requires_grad_(True)
        ye = ref(xe, he)
        ye.backward(gs)
        torch.cuda.synchronize()
        assert torch.equal(ys, ye), r
        assert torch.equal(xs.grad, xe.grad), r
        assert torch.equal(hs.grad, he.grad), r


def test_fused_mse_loss_matches_fp64_and_replays_in_a_graph():
    """conv.mse_loss (pcc_mse + pcc_backward_scaled): the loss against the fp64
    loss; dL/dx, dL/dh and dL/dtarget against the fp64 direct gradient of
    3 * MSE (the factor exercises the device-side upstream scale); agreement
    with the unfused path (torch's mse_loss over conv); and a captured step,
    replayed with new values written into its tensors, bitwise equal to the
    eager fused step every time."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps = [256, 2048], [14, 0], 9000
    C, T = 3, 12000
    rng = np.random.default_rng(21)
    x = rng.standard_normal((C, T)).astype(np.float32)
    h = decaying_noise_ir(taps, 0.8, 48000, 22)
    hid = decaying_noise_ir(taps, 1.5, 48000, 23)
    t = (np.stack([O.direct_conv_f64(x[c], hid) for c in range(C)])
         + 1e-3 * rng.standard_normal((C, T))).astype(np.float32)
    conv = PartitionedConv(N, P, taps)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    ht = torch.from_numpy(h).cuda().requires_grad_(True)
    tt = torch.from_numpy(t).cuda().requires_grad_(True)
    loss = conv.mse_loss(xt, ht, tt)
    (3.0 * loss).backward()
    torch.cuda.synchronize()
    y64 = np.stack([O.direct_conv_f64(x[c], h) for c in range(C)])
    r64 = y64 - t.astype(np.float64)
    loss64 = float(np.mean(r64 ** 2))
    g64 = 3.0 * 2.0 * r64 / (C * T)
    lf = float(loss.detach())
    assert abs(lf - loss64) / loss64 < 1e-4, (lf, loss64)
    assert relmax(xt.grad.cpu().numpy(), f64_corr_x(g64, h)) < 1e-4
    assert relmax(ht.grad.cpu().numpy(), f64_corr_h(g64, x, taps)) < 1e-4
    assert relmax(tt.grad.cpu().numpy(), -g64) < 1e-4
    # the unfused path: torch's loss over the convolution
    xu = torch.from_numpy(x).cuda().requires_grad_(True)
    hu = torch.from_numpy(h).cuda().requires_grad_(True)
    lu = torch.nn.functional.mse_loss(conv(xu, hu), torch.from_numpy(t).cuda())
    (3.0 * lu).backward()
    torch.cuda.synchronize()
    assert abs(float(lu.detach()) - lf) / lf < 1e-5
    assert relmax(ht.grad.cpu().numpy(), hu.grad.cpu().numpy()) < 1e-4
    assert relmax(xt.grad.cpu().numpy(), xu.grad.cpu().numpy()) < 1e-4
    # a captured step against eager fused steps on the same values
    gen = torch.Generator(device="cuda").manual_seed(24)
    xs = torch.randn((C, T), device="cuda", generator=gen).requires_grad_(True)
    hs = (0.01 * torch.randn(taps, device="cuda", generator=gen)).requires_grad_(True)
    ts = torch.randn((C, T), device="cuda", generator=gen)

    def step():
        lo = conv.mse_loss(xs, hs, ts)
        (3.0 * lo).backward()
        return lo

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):  # sizes the work space outside the capture
            xs.grad = hs.grad = None
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    xs.grad = hs.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        ls = step()
    ref = PartitionedConv(N, P, taps)
    for r in range(3):
        with torch.no_grad():
            xs.copy_(torch.randn((C, T), device="cuda", generator=gen))
            hs.copy_(0.01 * torch.randn(taps, device="cuda", generator=gen))
            ts.copy_(torch.randn((C, T), device="cuda", generator=gen))
        graph.replay()
        torch.cuda.synchronize()
        xe = xs.detach().clone().requires_grad_(True)
        he = hs.detach().clone().requires_grad_(True)
        le = ref.mse_loss(xe, he, ts)
        (3.0 * le).backward()
        torch.cuda.synchronize()
        assert torch.equal(ls, le), r
        assert torch.equal(xs.grad, xe.grad), r
        assert torch.equal(hs.grad, he.grad), r


def test_fused_mse_kernel_unaligned_and_ragged():
    """pcc_mse on arrays that are not 16-byte aligned (the scalar path) and on
    a sample count that is not a multiple of 4 or of the block count: the loss
    and the gradient against fp64, and the aligned path equal to the scalar one
    on the same values within fp32 rounding of the loss's sum."""
    import ctypes
    import torch
    from partconv_amd import lib
    from partconv_amd.autograd import PartitionedConv
    conv = PartitionedConv([64], [2], 100)
    rng = np.random.default_rng(31)
    s = torch.cuda.current_stream().cuda_stream
    for C, T, off in ((1, 12345, 1), (3, 1001, 3), (2, 4096, 0), (1, 7, 2)):
        n = C * T
        yb = torch.from_numpy(rng.standard_normal(n + 4).astype(np.float32)).cuda()
        tb = torch.from_numpy(rng.standard_normal(n + 4).astype(np.float32)).cuda()
        gb = torch.zeros(n + 4, device="cuda")
        loss = torch.zeros(1, device="cuda")
        y, t, g = yb[off:off + n], tb[off:off + n], gb[off:off + n]
        rc = lib().pcc_mse(conv._h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(t.data_ptr()),
                           ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(loss.data_ptr()), C, T, ctypes.c_void_p(s))
        assert rc == 0
        torch.cuda.synchronize()
        d = y.cpu().numpy().astype(np.float64) - t.cpu().numpy().astype(np.float64)
        l64 = float(np.mean(d ** 2))
        assert abs(float(loss) - l64) / l64 < 1e-5, (C, T, off)
        assert relmax(g.cpu().numpy(), 2.0 * d / n) < 1e-6, (C, T, off)
        # nothing outside the gradient's range was written
        assert float(gb[:off].abs().sum()) == 0.0 and float(gb[off + n:].abs().sum()) == 0.0


def test_tuned_chunk_sizes_keep_results():
    """pcc_tune picks each level's chunk sizes from its candidates by timing;
    the forward output is bitwise that of the untuned form, and both gradients
    stay within the fp64 bars."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps, T, C = [128, 512, 4096], [14, 14, 0], 20000, 15000, 3
    rng = np.random.default_rng(21)
    x = rng.standard_normal((C, T)).astype(np.float32)
    h = decaying_noise_ir(taps, 0.8, 48000, 8)
    g = rng.standard_normal((C, T)).astype(np.float32)
    xt0 = torch.from_numpy(x).cuda()
    ht0 = torch.from_numpy(h).cuda()
    plain = PartitionedConv(N, P, taps)
    tuned = PartitionedConv(N, P, taps)
    choice = tuned.tune(xt0, ht0)
    assert len(choice) == 3
    for lvl, (mac, dh) in enumerate(choice):
        assert dh in (16, 32, 64, 128, 256), choice
        assert mac in (8, 16, 32, 64, 128), choice
    with torch.no_grad():
        assert torch.equal(plain(xt0, ht0), tuned(xt0, ht0))
    xt = xt0.clone().requires_grad_(True)
    ht = ht0.clone().requires_grad_(True)
    tuned(xt, ht).backward(torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    assert relmax(xt.grad.cpu().numpy(), f64_corr_x(g, h)) < 1e-4
    assert relmax(ht.grad.cpu().numpy(), f64_corr_h(g, x, taps)) < 1e-4


def test_filter_rebuilt_every_step_matches_fp64():
    """A filter rebuilt each step (h = p * gain) is a fresh tensor that the
    caching allocator places at the freed block of the previous step's h, with
    the same version counter: every step's forward and gradients must still use
    that step's filter values (fp64 direct forward and gradients)."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps, T, C = [128, 1024], [14, 0], 6000, 9000, 2
    rng = np.random.default_rng(31)
    x = rng.standard_normal((C, T)).astype(np.float32)
    g = rng.standard_normal((C, T)).astype(np.float32)
    p0 = decaying_noise_ir(taps, 0.3, 48000, 14)
    conv = PartitionedConv(N, P, taps)
    p = torch.from_numpy(p0).cuda().requires_grad_(True)
    xt = torch.from_numpy(x).cuda()
    gt = torch.from_numpy(g).cuda()
    addrs = []
    for gain in (1.0, -2.5, 0.5):
        h = p * gain                      # new tensor each step
        addrs.append(h.data_ptr())
        h.retain_grad()
        y = conv(xt, h)
        y.backward(gt)
        torch.cuda.synchronize()
        hf = p0.astype(np.float64) * gain
        yr = np.stack([O.direct_conv_f64(x[c], hf.astype(np.float32)) for c in range(C)])
        err = float(np.max(np.abs(y.detach().cpu().numpy() - yr)) / np.sqrt(np.mean(yr ** 2)))
        assert err < 1e-5, (gain, err)
        assert relmax(h.grad.cpu().numpy(), f64_corr_h(g, x, taps)) < 1e-4, gain
        del y, h
    print("filter addresses per step:", addrs, flush=True)
    # the gain flows into p's gradient: sum of gains times dL/dh
    want = (1.0 - 2.5 + 0.5) * f64_corr_h(g, x, taps)
    assert relmax(p.grad.cpu().numpy(), want) < 1e-4


def test_differentiable_workload_full_size():
    """The differentiable workload at its stated size: 64 clips x 131072
    samples, 65536-tap learnable IR (RT60 0.8 s), partition [256, 2048, 16384]
    (16384 x 3 after the recount), forward + dL/dx + dL/dh.  Forward against
    fp64 direct convolution on 8 seeded clips (1e-5 of RMS); dL/dx on the same
    clips and dL/dh (a sum over all 64 clips) against the fp64 direct gradient
    (1e-4 relative)."""
    import torch
    from partconv_amd.autograd import PartitionedConv
    N, P, taps, T, C = [256, 2048, 16384], [14, 14, 0], 65536, 131072, 64
    rng = np.random.default_rng(64)
    x = rng.standard_normal((C, T)).astype(np.float32)
    g = rng.standard_normal((C, T)).astype(np.float32)
    h = decaying_noise_ir(taps, 0.8, 48000, 15)
    conv = PartitionedConv(N, P, taps)
    assert conv.P == [14, 14, 3]
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    ht = torch.from_numpy(h).cuda().requires_grad_(True)
    y = conv(xt, ht)
    y.backward(torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    yg = y.detach().cpu().numpy()
    dx = xt.grad.cpu().numpy()
    sample = sorted(rng.choice(C, size=8, replace=False).tolist())
    print("checked clips", sample, flush=True)
    yr = np.stack([O.direct_conv_f64(x[c], h) for c in sample])
    err = float(np.max(np.abs(yg[sample] - yr)) / np.sqrt(np.mean(yr ** 2)))
    assert err < 1e-5, err
    assert relmax(dx[sample], f64_corr_x(g[sample], h)) < 1e-4
    dh_ref = np.zeros(taps)
    for c0 in range(0, C, 8):  # fp64 sum over clips, in slices to bound host memory
        dh_ref += f64_corr_h(g[c0:c0 + 8], x[c0:c0 + 8], taps)
    assert relmax(ht.grad.cpu().numpy(), dh_ref) < 1e-4
