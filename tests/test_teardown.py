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


@pytest.mark.gpu
def test_exit_with_live_gpu_work():
    """Every engine kind alive at exit, with device state: a host-buffer
    session whose persistent tick kernel is resident, a launcher thread, the
    differentiable form's work space, the real-time MIMO engine and the MIMO
    throughput form; the child exits without cleanup and must return 0."""
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
