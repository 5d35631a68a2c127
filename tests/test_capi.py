"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/partconv_amd.h declares, and reports errors without a GPU."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "partconv_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pc[a-z]*_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_expected_api():
    names = declared_functions()
    for n in ["pcm_create", "pcm_setup", "pcm_setup_mem", "pcm_run", "pcm_reset", "pcm_cleanup",
              "pcf_setup", "pcf_run_level", "pcf_input", "pcf_output", "pcf_bench", "pc_last_error"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from partconv_amd import _lib
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python signature table covers them all
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_errors_are_returned_not_fatal():
    from partconv_amd import _lib
    L = _lib.lib()
    h = L.pcm_create(0, 0)
    assert h
    import numpy as np
    y = np.zeros(64, dtype=np.float32)
    # run before setup -> PC_ERR_STATE, message set
    assert L.pcm_run(h, y.ctypes.data, y.ctypes.data, 64) == 6
    assert "before setup" in _lib.last_error()
    # bad mode
    assert L.pcm_set_mode(h, 7) == 1
    # missing config file -> PC_ERR_CONFIG
    assert L.pcm_setup(h, b"/nonexistent/pc.config", 1, 0, 0.0) == 2
    assert "Unable to open config file" in _lib.last_error()
    # release window: 1 .. 16 ticks
    assert L.pcm_set_release_window(h, 0) == 1 and "release window" in _lib.last_error()
    assert L.pcm_set_release_window(h, 17) == 1
    assert L.pcm_set_release_window(h, 16) == 0
    assert L.pcm_set_release_window(None, 2) == 1
    assert L.pcxr_set_release_window(None, 2) == 1
    L.pcm_destroy(h)


def test_mimo_realtime_rejects_bad_shapes_without_gpu():
    """pcxr_create validates sizes and the partition before touching the GPU
    (the partition rules are the reference's, src/partconv.cpp:94-158)."""
    from partconv_amd import _lib
    L = _lib.lib()

    def create(FS, bs, N, P, taps, n_in, n_out):
        Na = (ctypes.c_int * len(N))(*N)
        Pa = (ctypes.c_int * len(P))(*P)
        return L.pcxr_create(FS, bs, len(N), Na, Pa, taps, n_in, n_out)

    assert not create(48000, 256, [256, 4096], [30, 0], 65536, 33, 32)
    assert "1..32" in _lib.last_error()
    assert not create(48000, 256, [256, 4096], [30, 0], 65536, 32, 0)
    assert not create(48000, 128, [256, 4096], [30, 0], 65536, 4, 4)
    assert "buffer_size" in _lib.last_error()
    assert not create(48000, 32, [32, 256], [8, 0], 4096, 4, 4)
    assert not create(48000, 256, [256, 8192], [64, 0], 65536, 4, 4)
    assert "4096" in _lib.last_error()
    # a later block smaller than an earlier one violates the reference's rules
    assert not create(48000, 256, [256, 128], [4, 0], 4096, 4, 4)
    assert not create(0, 256, [256], [0], 4096, 4, 4)


def test_config_parse_errors_without_gpu(tmp_path):
    from partconv_amd import _lib
    L = _lib.lib()
    cfg = tmp_path / "bad.config"
    cfg.write_text("48000 32 1 1 1 1\nx.wav 0 0 0 1.0 2 64 6 256\n")
    h = L.pcm_create(0, 0)
    assert L.pcm_setup(h, str(cfg).encode(), 1, 0, 0.0) == 2
    assert "first block size must equal buffer_size" in _lib.last_error()
    cfg.write_text("48000 64 1 1 1 1\nmissing.wav 0 0 0 1.0 2 64 6 256\n")
    assert L.pcm_setup(h, str(cfg).encode(), 1, 0, 0.0) == 4
    assert "Can't open audio file" in _lib.last_error()
    L.pcm_destroy(h)


def test_python_surface_shapes():
    import partconv_amd as pc
    a = pc.intArray(3)
    a[0], a[1], a[2] = 64, 256, 1024
    assert [a[i] for i in range(3)] == [64, 256, 1024]
    d = pc.doubleArray(4)
    d[3] = 2.5
    assert d[3] == 2.5


def test_greedy_partition_matches_rules():
    """The product's greedy rule (C ABI) against the oracle's restatement."""
    import numpy as np
    from oracle import rules as R
    from partconv_amd import greedy_partition
    bs = [128 << i for i in range(9)]
    rng = np.random.default_rng(31)
    for T in [4096, 5000, 1 << 20, 70000] + [int(v) for v in rng.integers(300, 1 << 20, size=200)]:
        N, P = greedy_partition(bs, T)
        g = R.greedy_partition(bs, T)
        assert (N, P) == (g.N, g.P), T
    assert greedy_partition(bs, 4096) == ([128, 256, 512, 1024, 2048], [2, 2, 2, 2, 1])


def test_cxx_dropin_header_compiles_and_links(tmp_path):
    """include/partconvMulti.h: the reference's class names and signatures
    compile with g++ against the C ABI and link against the library; without a
    GPU, setup reports an error instead of exiting."""
    import shutil
    import subprocess
    from partconv_amd import _lib
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "dropin.cpp"
    src.write_text(r"""
#include "partconvMulti.h"
int main()
{
    PartConvMulti pc(0);
    PartConvMultiRelaxed pr(0);
    PartConvFilter f;
    PartConvMax mx(0);
    int N[2] = {64, 256}, P[2] = {6, 0};
    float h[500] = {1.f};
    int bs[3] = {64, 256, 1024};
    int rc1 = pc.setup("/nonexistent/pc.config");
    int rc2 = mx.setup(48000, 2, 2, 1, h, 500, 1, bs, 3, 1, nullptr);
    float a[64] = {0}, b[64] = {0};
    float *outs[2] = {a, b};
    const float *ins_c[2] = {a, b};
    float **ins = const_cast<float **>(ins_c);
    int rc3 = mx.frameNum == 0 ? 0 : 1;
    PartConvMimo mimo;
    int Nm[2] = {256, 4096}, Pm[2] = {30, 0};
    int rc4 = mimo.setup(48000, 256, 2, Nm, Pm, 65536, 33, 32);  // 33 inputs: rejected
    if (rc4 == 0) return 2;
    (void)f;
    (void)pr;
    (void)outs;
    (void)ins;
    printf("%d %d %d\n", rc1 != 0, rc2 == 0 || rc2 == 1, rc3);
    return 0;
}
""")
    exe = tmp_path / "dropin"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-variable", "-I", str(ROOT / "include"),
                    str(src), "-o", str(exe), "-L", str(_lib.LIB_PATH.parent), "-lpartconv_amd",
                    f"-Wl,-rpath,{_lib.LIB_PATH.parent}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["1", "1", "0"]


def test_prof_summary_splits_calls_by_phase(tmp_path, capsys):
    """prof_summary trace --split: launches of one kernel shape are grouped by
    call index mod K in time order (one row per level of a per-level pass)."""
    from partconv_amd import prof_summary
    rows = ["Kernel_Name,Grid_Size_X,Workgroup_Size_X,Start_Timestamp,End_Timestamp"]
    t = 0
    for call in range(6):  # durations 1, 2, 3, 1, 2, 3 us
        d = 1000 * (call % 3 + 1)
        rows.append(f"k_mac(int),256,256,{t},{t + d}")
        t += d + 100
    rows.append(f"k_other(int),64,64,{t},{t + 500}")
    p = tmp_path / "trace.csv"
    p.write_text("\n".join(rows) + "\n")
    prof_summary.trace(str(p), ("k_mac", 3))
    out = capsys.readouterr().out
    means = {}
    for line in out.splitlines():
        if line.startswith("[call%3="):
            means[int(line[8])] = float(line.split()[-3])
    assert means == {0: 1.0, 1: 2.0, 2: 3.0}
    assert "k_other(int)" in out


def test_prof_summary_step_sums_between_markers(tmp_path):
    """prof_summary step: only the dispatches between the last two 4 MiB
    copies count (a copy of another size inside the step is a kernel of the
    step); bytes are scaled by the calibration copy (65536 KiB); the
    transforms are keyed by level and order (filter, forward, gradient,
    inverse), the products and dL/dh kernels by their order (levels
    ascending), other kernels as the loss."""
    import json
    from partconv_amd import prof_summary
    seq = [("__amd_rocclr_copyBuffer", 131072, 2 * 65536),   # calibration: reads half -> scale 0.5
           ("void k_off_rfft_reg<8>(WinSrc)", 1024, 999),      # before the markers: ignored
           ("__amd_rocclr_copyBuffer", 131072, 8192),         # marker: 4 MiB after scaling
           ("void k_off_rfft_reg<8>(WinSrc)", 64, 2),          # filter transform (smaller grid)
           ("void k_off_rfft_reg<8>(WinSrc)", 1024, 10),       # forward transform
           ("void k_off_mac_w<16, 1>(float2 const*)", 512, 20),
           ("void k_off_irfft_reg<8>(float2 const*)", 1024, 30),
           ("void at::native::reduce_kernel<512>(int)", 512, 4),
           ("__amd_rocclr_copyBuffer", 131072, 16384),        # a copy inside the step (loss)
           ("void k_off_rfft_reg<8>(WinSrc)", 1024, 40),       # gradient transform
           ("void k_off_dh_part_b<16>(float2 const*)", 256, 50),
           ("k_off_dh_reduce(float2 const*)", 256, 6),
           ("void k_off_irfft_reg<8>(float2 const*)", 64, 8),  # dL/dh inverse
           ("__amd_rocclr_copyBuffer", 131072, 8192)]         # marker
    for cname in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = ["Dispatch_Id,Kernel_Name,Grid_Size,Counter_Name,Counter_Value"]
        for i, (n, g, v) in enumerate(seq):
            rows.append(f'{i},"{n}",{g},{cname},{v}')
        (tmp_path / f"{cname}.csv").write_text("\n".join(rows) + "\n")
    out = tmp_path / "step.json"
    prof_summary.step(str(tmp_path / "FETCH_SIZE.csv"), str(tmp_path / "WRITE_SIZE.csv"), 131072, str(out), "abc")
    d = json.loads(out.read_text())
    kib = 1024 * 0.5 * 2  # both counters, scale 0.5
    want = {"filter_transform@256": 2, "forward_transform@256": 10, "product@256": 20, "inverse_transform@256": 30,
            "loss": 4 + 16384, "gradient_transform@256": 40, "dh_partial@256": 50, "dh_reduce@256": 6, "dh_inverse@256": 8}
    assert d["by_kernel"] == {k: v * kib for k, v in want.items()}
    assert d["bytes_per_step"] == sum(want.values()) * kib and d["commit"] == "abc"


def test_cxx_reference_call_shapes(tmp_path):
    """The reference's own call shapes compile and behave without a GPU:
    PartConvFilter::setup(M, N, P, Vector*, fftwOptLevel, wisdom) with a
    Vector, with NULL (src/tune.py:31 calls setup(M, N, P, None, 2,
    'wisdom.wis')) and with defaults; readImpulseResponse(Vector&, file,
    verbosity) (src/sndtools.h:16) and Vector's members; a partition error
    returns 1 and leaves the caller's arrays unchanged (the reference copies
    them, src/partconv.cpp:83-84); PartConvMulti::run counts frameNum only for
    ticks that ran."""
    import shutil
    import struct
    import subprocess
    import numpy as np
    from partconv_amd import _lib
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    pcm = (np.arange(300) - 150).astype("<i2")
    fmt = struct.pack("<HHIIHH", 1, 1, 44100, 88200, 2, 16)
    body = b"WAVE" + b"fmt " + struct.pack("<I", 16) + fmt + b"data" + struct.pack("<I", pcm.nbytes) + pcm.tobytes()
    wav = tmp_path / "ir.wav"
    wav.write_bytes(b"RIFF" + struct.pack("<I", len(body)) + body)
    src = tmp_path / "shapes.cpp"
    src.write_text(r"""
#include "partconvMulti.h"
#include "sndtools.h"
int main(int argc, char **argv)
{
    Vector ir;
    if (readImpulseResponse(ir, argv[1], 0) != 0) return 10;
    if (ir.getSize() != 300 || ir[0] != -150.f / 32768.f || *ir(299) != 149.f / 32768.f) return 11;
    Vector copy = ir;
    if (copy.zeroPad(400, 10) != 0 || copy.getSize() != 400 || copy[10] != ir[0] || copy[0] != 0.f) return 12;
    if (copy.zeroPad(100) != 1) return 13;
    Vector missing;
    if (readImpulseResponse(missing, "/nonexistent.wav") == 0) return 14;

    PartConvFilter pc;
    int M = 2;
    int blocksize[2] = {64, 256}, numblocks[2] = {6, 7};
    int bad_blocks[2] = {64, 100}, bad_counts[2] = {6, 7};
    if (pc.setup(M, bad_blocks, bad_counts, NULL, 2, "wisdom.wis") != 1) return 20;   /* not a power of 2 */
    if (bad_counts[1] != 7 || bad_blocks[1] != 100) return 21;                         /* caller's arrays kept */
    int short_blocks[2] = {64, 256}, short_counts[2] = {1, 7};
    if (pc.setup(M, short_blocks, short_counts, &ir) != 1) return 22;                  /* delay rule */
    if (short_counts[1] != 7) return 23;
    /* the valid shapes: without a GPU they fail at the device, after validation */
    int r1 = pc.setup(M, blocksize, numblocks, NULL, 2, "wisdom.wis");
    int r2 = pc.setup(M, blocksize, numblocks, &ir, 2, "wisdom.wis");
    int r3 = pc.setup(M, blocksize, numblocks);
    if (numblocks[1] != 7) return 24;
    PartConvMulti multi;
    float in[64] = {0}, out[64] = {0};
    int rr = multi.run(out, in, 64);                                                   /* before setup: error */
    if (rr == 0 || multi.frameNum != 0) return 25;
    Waiter *w = multi.doneWaiter;
    (void)w;
    printf("%d %d %d\n", r1 == 0 || r1 == 1, r2 == 0 || r2 == 1, r3 == 0 || r3 == 1);
    return 0;
}
""")
    exe = tmp_path / "shapes"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-variable", "-I", str(ROOT / "include"),
                    str(src), "-o", str(exe), "-L", str(_lib.LIB_PATH.parent), "-lpartconv_amd", f"-Wl,-rpath,{_lib.LIB_PATH.parent}"],
                   check=True)
    out = subprocess.run([str(exe), str(wav)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stderr)
    assert out.stdout.split() == ["1", "1", "1"]
    assert "power of 2" in out.stderr and "sufficient delay" in out.stderr
