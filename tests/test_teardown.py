"""Process teardown with live handles.  A process that exits without cleanup
must still exit with status 0: pc_shutdown (run from the Python loader's
atexit hook and from a C atexit handler registered after the library's first
HIP call) destroys the live handles -- resident persistent kernels stopped,
launcher threads joined -- before the HIP runtime's own exit-time teardown,
and destructors that run later are no-ops."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_child(code, timeout=240):
    p = subprocess.run([sys.executable, "-c", textwrap.dedent(code)], cwd=ROOT, capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout, p.stderr


def test_exit_with_live_handles_without_gpu():
    rc, out, err = run_child("""
        import partconv_amd as P
        m = P.PartConvMulti(0)
        r = P.PartConvMultiRelaxed(0)
        f = P.PartConvFilter(0)
        x = P.PartConvMax(0)
        keep = [m, r, f, x]          # alive at exit, never cleaned up
        print("handles live", flush=True)
    """)
    assert rc == 0, (rc, out, err)
    assert "handles live" in out


def test_shutdown_is_terminal_and_idempotent():
    rc, out, err = run_child("""
        import partconv_amd as P
        from partconv_amd import _lib
        m = P.PartConvMulti(0)
        f = P.PartConvFilter(0)
        L = _lib.lib()
        L.pc_shutdown()
        L.pc_shutdown()                      # idempotent
        assert L.pcm_create(0, 0) is None    # creation fails afterwards
        assert "shut down" in _lib.last_error()
        assert L.pcf_create(0, 44100) is None
        L.pcm_destroy(m._h)                  # destroyed by the shutdown: no-op
        L.pcf_destroy(f._h)
        del m, f                             # __del__ after shutdown: no-op
        print("ok", flush=True)
    """)
    assert rc == 0, (rc, out, err)
    assert "ok" in out


def hold_cooperative_queue():
    """Make this process take the device's cooperative launch (a persistent
    run), which it then holds while alive: a child's cooperative launches are
    refused and its persistent kernels must go through plain launches."""
    import numpy as np
    import torch
    import partconv_amd as P
    m = P.PartConvMulti(0)
    h = np.zeros(500, dtype=np.float32)
    h[0] = 1.0
    assert m.setup_mem(48000, 64, 1, 1, [dict(N=[64, 256], P=[6, 0], in_ch=0, out_ch=0, ir=h)]) == 0
    x = torch.randn(64 * 20, device="cuda")
    y = torch.zeros_like(x)
    torch.cuda.synchronize()
    assert m.run_device(y.data_ptr(), x.data_ptr(), 20, True, persistent=True) == 0
    assert m.stats().persistent_ticks == 20
    m.cleanup()


@pytest.mark.gpu
def test_persistent_kernels_without_the_cooperative_queue():
    """While another process holds the device's cooperative launch, the
    persistent tick kernels of both real-time engines are launched plainly
    after a residency check: the instance engine (device-resident ticks and
    the host-buffer session) and the real-time MIMO engine must serve their
    ticks persistently, with outputs equal to fp64 and to per-tick launches."""
    hold_cooperative_queue()
    rc, out, err = run_child("""
        import numpy as np, torch
        import partconv_amd as P
        from oracle import oracle as O
        from partconv_amd.mimo import MimoRealtime
        from partconv_amd.workloads import decaying_noise_ir
        rng = np.random.default_rng(1)
        irs = [decaying_noise_ir(20000, 0.5, 48000, 30 + c) for c in range(3)]
        insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=c, out_ch=c, ir=irs[c]) for c in range(3)]
        x = rng.standard_normal((300 * 64, 3)).astype(np.float32)
        m = P.PartConvMulti(0)
        assert m.setup_mem(48000, 64, 3, 3, insts) == 0
        y = m.process(x)
        assert m.stats().persistent_ticks == 300, m.stats().persistent_ticks
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros_like(xd)
        m.reset()
        m.clear_stats()
        assert m.run_device(yd.data_ptr(), xd.data_ptr(), 300, True, persistent=True) == 0, P.last_error()
        assert m.stats().persistent_ticks == 300
        assert np.array_equal(yd.cpu().numpy(), y)
        for c in range(3):
            r = O.direct_conv_f64(x[:, c], irs[c])
            assert np.max(np.abs(y[:, c] - r)) / np.sqrt(np.mean(r ** 2)) < 1e-5
        rt = MimoRealtime(48000, 64, [64, 512], [14, 0], 3000, 2, 3)
        h = torch.randn((3, 2, 3000), device="cuda") * 0.05
        rt.set_filters(h)
        xm = torch.randn((100 * 64, 2), device="cuda")
        ym = torch.zeros((100 * 64, 3), device="cuda")
        rt.run_device(ym, xm, 100)
        hc, xc, yc = h.cpu().numpy(), xm.cpu().numpy(), ym.cpu().numpy()
        for o in range(3):
            r = sum(O.direct_conv_f64(xc[:, i], hc[o, i]) for i in range(2))
            assert np.max(np.abs(yc[:, o] - r)) / np.sqrt(np.mean(r ** 2)) < 1e-5
        print("plain persistent launches ok", flush=True)
    """)
    print(out, err[-2000:])
    assert rc == 0, (rc, out, err[-4000:])
    assert "plain persistent launches ok" in out


@pytest.mark.gpu
def test_exit_with_live_gpu_work():
    """Every engine kind alive at exit, with device state: a host-buffer
    session whose persistent tick kernel is resident, a launcher thread, the
    differentiable form's work space, the real-time MIMO engine and the MIMO
    throughput form; the child exits without cleanup and must return 0."""
    hold_cooperative_queue()
    rc, out, err = run_child("""
        import numpy as np, torch
        import partconv_amd as P
        from partconv_amd.autograd import PartitionedConv
        from partconv_amd.mimo import MimoConv, MimoRealtime
        from partconv_amd.workloads import decaying_noise_ir
        rng = np.random.default_rng(0)
        m = P.PartConvMulti(0)
        insts = [dict(N=[64, 256, 2048], P=[6, 14, 0], in_ch=c, out_ch=c,
                      ir=decaying_noise_ir(20000, 0.5, 48000, c)) for c in range(4)]
        assert m.setup_mem(48000, 64, 4, 4, insts) == 0
        y = m.process(rng.standard_normal((200 * 64, 4)).astype(np.float32))
        print("session ticks", m.stats().persistent_ticks, flush=True)
        assert m.stats().persistent_ticks == 200
        conv = PartitionedConv([128, 1024], [14, 0], 5000)
        x = torch.randn((2, 8000), device="cuda", requires_grad=True)
        h = torch.randn(5000, device="cuda", requires_grad=True)
        conv(x, h).sum().backward()
        rt = MimoRealtime(48000, 64, [64, 512], [14, 0], 3000, 2, 3)
        rt.set_filters(torch.randn((3, 2, 3000), device="cuda"))
        rt.run_device(torch.empty((50 * 64, 3), device="cuda"), torch.randn((50 * 64, 2), device="cuda"), 50)
        mx = MimoConv([128, 1024], [14, 0], 3000, 2, 3)
        mx.set_filters(torch.randn((3, 2, 3000), device="cuda"))
        ymx = mx(torch.randn((2, 5000), device="cuda"))
        torch.cuda.synchronize()
        keep = [m, conv, rt, mx]
        print("exiting with live handles", flush=True)
    """)
    print(out, err[-2000:])
    assert rc == 0, (rc, out, err[-4000:])
    assert "exiting with live handles" in out
