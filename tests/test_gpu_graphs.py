"""Level blocks launched as per-level HIP graphs (pcm_set_level_graphs): every
graph is replayed once per block of its level with the tick-dependent
arguments updated in place; the outputs of per-tick and persistent runs, in the
batched and the plain layout, strict and relaxed, must equal the directly
launched levels' bitwise, and fp64."""
import numpy as np
import pytest

from oracle import oracle as O
from partconv_amd import PartConvMulti, PartConvMultiRelaxed, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("relaxed", [False, True])
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_level_graphs_equal_direct_launches(relaxed, mode):
    import torch
    print(f"case level graphs relaxed={relaxed} mode={mode}", flush=True)
    # small levels (256, 1024 inside some instances), four-step levels (4096, 8192)
    insts = [dict(N=[64, 256, 2048, 8192], P=[6, 14, 6, 0], in_ch=0, out_ch=0),
             dict(N=[64, 1024, 4096], P=[30, 6, 0], in_ch=1, out_ch=1),
             dict(N=[64, 256, 1024], P=[6, 6, 0], in_ch=2, out_ch=1)]
    irs = [decaying_noise_ir(t, 0.4, 48000, 80 + i) for i, t in enumerate((40000, 30000, 5000))]
    nfr = 1100
    rng = np.random.default_rng(81 + relaxed)
    x = rng.standard_normal((nfr * 64, 3)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()

    def make(graphs):
        m = (PartConvMultiRelaxed if relaxed else PartConvMulti)(0, mode=mode)
        assert m.setup_mem(48000, 64, 3, 2, [dict(ir=h, **d) for d, h in zip(insts, irs)]) == 0, last_error()
        assert m.set_level_graphs(graphs) == 0
        return m

    outs = {}
    for graphs in (False, True):
        for persistent in (False, True):
            if persistent and mode == "plain":
                continue
            m = make(graphs)
            y = torch.zeros((nfr * 64, 2), device="cuda")
            torch.cuda.synchronize()
            # two calls: the graphs are reused across them
            assert m.run_device(y.data_ptr(), xd.data_ptr(), 500, True, persistent=persistent) == 0, last_error()
            off = 500 * 64 * 3 * 4
            assert m.run_device(y.data_ptr() + 500 * 64 * 2 * 4, xd.data_ptr() + off, nfr - 500, True,
                                persistent=persistent) == 0, last_error()
            torch.cuda.synchronize()
            outs[(graphs, persistent)] = y.cpu().numpy()
            m.cleanup()
    ref = outs[(False, False)]
    for k, y in outs.items():
        assert np.array_equal(y, ref), k
    r = O.direct_conv_f64(x[:, 1], irs[1]) + O.direct_conv_f64(x[:, 2], irs[2])
    if relaxed:
        r = np.concatenate([np.zeros(64), r[:-64]])
    assert np.max(np.abs(ref[:, 1] - r)) / np.sqrt(np.mean(r ** 2)) < 1e-5


def test_level_graphs_follow_the_partial_width():
    """Changing the delay-line-sum width rebuilds the graphs; output unchanged."""
    import torch
    from partconv_amd._lib import lib
    insts = [dict(N=[64, 512, 4096], P=[14, 14, 0], in_ch=0, out_ch=0)]
    irs = [decaying_noise_ir(30000, 0.5, 48000, 90)]
    x = torch.randn((600 * 64, 1), device="cuda")
    ys = []
    for width in (2, 8, 4):
        m = PartConvMulti(0)
        assert m.setup_mem(48000, 64, 1, 1, [dict(ir=irs[0], **insts[0])]) == 0
        m.set_level_graphs(True)
        y = torch.zeros_like(x)
        assert m.run_device(y.data_ptr(), x.data_ptr(), 300, True) == 0
        assert lib().pcm_set_partial_width(m._h, width) == 0
        assert m.run_device(y.data_ptr() + 300 * 64 * 4, x.data_ptr() + 300 * 64 * 4, 300, True) == 0
        torch.cuda.synchronize()
        ys.append(y.clone())
        m.cleanup()
    assert torch.equal(ys[0], ys[1]) and torch.equal(ys[0], ys[2])
