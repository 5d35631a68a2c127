"""Randomised parity sweep: seeded engine configurations (buffer size, channel
counts, instance count, per-instance greedy or widened partitions, mute,
scale, channel routing with several or no sources per output, strict or
relaxed, batched or plain) driven through every entry path in turn -- host
run() through the persistent session, device-resident persistent ticks,
device-resident per-tick launches, host run() with one launch per tick -- and
compared with the fp32 oracle (src/partconvMulti.cpp:815-928, 1017-1134
restated) at 1e-5 of each output channel's RMS; outputs without a source must
be exactly zero."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import rules as R
from partconv_amd import PartConvMulti, PartConvMultiRelaxed, last_error
from partconv_amd.workloads import decaying_noise_ir

pytestmark = pytest.mark.gpu

TOL = 1e-5


def random_config(seed):
    rng = np.random.default_rng(1000 + seed)
    N0 = int(rng.choice([16, 32, 64, 128, 256]))
    n_in, n_out = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    n_inst = int(rng.integers(1, 13)) if seed % 5 else int(rng.integers(30, 45))
    insts, irs = [], []
    for i in range(n_inst):
        taps = int(rng.integers(N0, min(60000, 400 * N0)))
        js = sorted({0} | {int(j) for j in rng.choice(np.arange(1, 10), size=int(rng.integers(1, 5)), replace=False)})
        bs = [N0 << j for j in js if (N0 << j) <= 16384]
        g = R.greedy_partition(bs, taps)
        N, P = list(g.N), list(g.P)
        if len(N) > 1 and rng.random() < 0.4:
            # widen a non-last level; keep it only if the last level is still needed
            L = int(rng.integers(0, len(N) - 1))
            P2 = list(P)
            P2[L] += int(rng.integers(1, 4))
            try:
                R.check_partition(N, P2, taps)
                P = P2
            except R.PartitionError:
                pass
        insts.append(dict(N=N, P=P, in_ch=int(rng.integers(0, n_in)), out_ch=int(rng.integers(0, n_out)),
                          mute=int(rng.random() < 0.15), scale=float(rng.uniform(0.25, 2.0))))
        irs.append(decaying_noise_ir(taps, max(0.05, taps / 48000), 48000, 7000 + 31 * seed + i))
    relaxed = bool(rng.random() < 0.4)
    mode = "plain" if seed % 4 == 3 else "batched"
    max_k = max(max(d["N"]) for d in insts) // N0
    nfr = int(min(600, 3 * max_k + 37))
    x = rng.standard_normal((nfr * N0, n_in)).astype(np.float32)
    return N0, n_in, n_out, insts, irs, relaxed, mode, x


@pytest.mark.parametrize("seed", range(16))
def test_random_configuration_all_paths(seed):
    import torch
    N0, n_in, n_out, insts, irs, relaxed, mode, x = random_config(seed)
    print(f"case random seed={seed} N0={N0} in={n_in} out={n_out} inst={len(insts)} relaxed={relaxed} "
          f"mode={mode}", flush=True)
    ref = O.OracleMulti(48000, N0, n_in, n_out, insts, irs, relaxed=relaxed).process(x)
    cls = PartConvMultiRelaxed if relaxed else PartConvMulti
    m = cls(0, mode=mode)
    assert m.setup_mem(48000, N0, n_in, n_out, [dict(ir=h, **d) for d, h in zip(insts, irs)]) == 0, last_error()
    nfr = x.shape[0] // N0
    cuts = sorted({0, nfr // 5, 2 * nfr // 5 + 1, 3 * nfr // 5 + 2, nfr})
    ys = []
    for part, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
        seg = x[a * N0:b * N0]
        if part in (0, 3):
            m.set_host_session(part == 0)
            ys.append(m.process(seg))
        else:
            xd = torch.from_numpy(np.ascontiguousarray(seg)).cuda()
            yd = torch.zeros((seg.shape[0], n_out), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            assert m.run_device(yd.data_ptr(), xd.data_ptr(), b - a, True, persistent=(part == 1)) == 0, last_error()
            ys.append(yd.cpu().numpy())
    y = np.concatenate(ys)
    m.cleanup()
    for c in range(n_out):
        r = ref[:, c].astype(np.float64)
        rms = np.sqrt(np.mean(r ** 2))
        if rms == 0.0:
            assert np.all(y[:, c] == 0.0), c
        else:
            assert float(np.max(np.abs(y[:, c] - r)) / rms) < TOL, c
