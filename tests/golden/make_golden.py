"""Writes the golden fixtures under tests/golden/.

The reference holds no golden vectors for this path, so these fixtures are
float64 direct-convolution outputs (no FFT, no partitioning: numpy.convolve in
float64) for seeded white-noise inputs and seeded decaying-noise IRs.  Strict
mode expects y = conv(x, h); relaxed mode expects the same delayed by N0.

Run: python tests/golden/make_golden.py
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from partconv_amd.workloads import decaying_noise_ir, white_noise  # noqa: E402

CASES = [
    ("case_a_two_level", [16, 64], [6, 0], False, 300, 16 * 120),
    ("case_b_three_level", [32, 128, 512], [6, 6, 0], False, 2000, 32 * 150),
    ("case_c_relaxed", [32, 128, 512], [6, 6, 0], True, 1500, 32 * 120),
    ("case_d_cfg2_shape", [64, 256, 2048], [6, 14, 0], False, 6000, 64 * 160),
    ("case_e_single_level", [64], [0], False, 500, 64 * 40),
]


def main():
    for i, (name, N, P, relaxed, taps, n) in enumerate(CASES):
        x = white_noise(n, 500 + i).astype(np.float64)
        h = decaying_noise_ir(taps, taps / 48000.0, 48000, 600 + i).astype(np.float64)
        y = np.convolve(x, h)[:n]
        if relaxed:
            y = np.concatenate([np.zeros(N[0]), y[:-N[0]]])
        np.savez_compressed(HERE / f"{name}.npz", N=np.array(N), P=np.array(P),
                            relaxed=np.array(relaxed), x=x.astype(np.float32),
                            h=h.astype(np.float32), y=y)
        print(name, x.size, h.size)


if __name__ == "__main__":
    main()
