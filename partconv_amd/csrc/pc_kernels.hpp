// pc_kernels.hpp -- gfx950 kernels of the partitioned-convolution engine.
//
// One "level job" is one block of one partition level of one instance: the
// reference's PartConvFilter::runLevel (src/partconv.cpp:354-490):
//     window (2N real samples) -> r2c -> FDL slot -> sum_p X_{m-p} . H_p -> c2r
//     -> first N samples into the level's double buffer.
// A launch processes a whole batch of jobs of one block size (blockIdx picks
// the job), so hundreds of uneven instances share one launch per level tick.
//
// Data layout in HBM (per instance, per local level with block N, P parts):
//   H    : P rows of N float2, already scaled by 1/(2N).  Bin layout "packed":
//          slot 0 holds (X[0], X[N]) (both real), slots 1..N-1 hold X[k].
//          Small levels keep natural bin order; big levels keep the four-step
//          order (see k_big2), identical for H and the FDL.
//   FDL  : P rows of N float2, row m % P holds the spectrum of block m.
//   ybuf : 2 x N floats, block m writes half (m & 1).
//   ring : per-instance input history, R (power of two) samples written twice
//          (at t mod R and R + t mod R) so any window of <= R samples is
//          contiguous.  Sample t of the (scaled) input stream lives at t mod R.
//
// Window of block m at a level with tap offset T_L (taps before the level):
//     samples [m*N - N0 - T_L, m*N - N0 - T_L + 2N)
// which is exactly what the reference's DelayBuffer(N0, k, d_L) hands to
// runLevel (ring position r+f for frame f, r = d_L - k - 1,
// src/partconv.cpp:1401-1457).  Block m's first N output samples are output
// samples [(m+1)N - N0, (m+2)N - N0).
#pragma once
#include "pc_fft.hpp"

#define PC_MAX_LEVELS 16

enum { PC_MODE_RUN = 0, PC_MODE_SPECTRUM = 1 };

struct LevelJob {
    const float *ring;   // RUN: instance input ring (2R floats); SPECTRUM: IR taps
    cf *fdl;             // RUN: P x N FDL; SPECTRUM: destination H row
    const cf *H;         // P x N filter spectra
    float *ybuf;         // 2 x N output double buffer
    cf *scratch;         // big levels: N float2 work space
    int P;
    int T;               // RUN: taps before this level; SPECTRUM: first tap of the row
    int ring_mask;       // RUN: R - 1; SPECTRUM: number of IR taps
    int pad;
};

struct TickJob {
    LevelJob l0;
    int in_ch, out_ch, mute;
    float scale;
    int nlev;                              // levels >= 1 of this instance
    int kshift[PC_MAX_LEVELS];             // log2(N_L / N0)
    int NL[PC_MAX_LEVELS];
    const float *ybuf[PC_MAX_LEVELS];
};

struct TickArgs {
    const TickJob *jobs;
    int njobs;
    int N0;
    uint64_t f;            // tick index (frame number since setup/reset)
    const float *in;       // interleaved input frame(s)
    long long in_stride;   // floats between consecutive frames (0 = same buffer)
    long long fbase;       // frame index of in[0]
    int nIn;
    float *out;            // interleaved output
    long long out_stride;
    int nOut;
    int relaxed;
    int mix;               // 0: write y_inst only (plain mode), 1: direct, 2: reduce
    float *y_inst;         // njobs x N0 per-instance outputs (mix == 0 or 2)
    const int *src_off;    // reduce: per output channel offsets into src_list (nOut+1)
    const int *src_list;   // reduce: job indices per output channel
    unsigned *ticket;      // reduce/direct: arrival counter (reset by the last block)
    unsigned long long *done_flag;  // host-visible completion flag (may be null)
};

// ---------------------------------------------------------------------------
// Small levels: one workgroup runs the whole job in LDS (N <= 2048).
// ---------------------------------------------------------------------------

template <int LOGN>
struct SmallCfg {
    static constexpr int N = 1 << LOGN;
    static constexpr int NT = (N / 4 < 64) ? 64 : ((N / 4 > 256) ? 256 : N / 4);
};

// Forward real FFT of the 2N-sample window already packed in bufA as N float2;
// result X (packed) in the returned buffer.
template <int LOGN>
__device__ cf *small_forward(cf *bufA, cf *bufB, const cf *__restrict__ tw, int tid)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    cf *Z = fft_lds<false>(bufA, bufB, LOGN, 1, N, tw, tid, NT);
    // split pairs (k, N-k) in place, k in [0, N/2]
    for (int k = tid; k <= N / 2; k += NT) {
        if (k == 0) {
            const cf z = Z[0];
            Z[0] = make_float2(z.x + z.y, z.x - z.y);
        } else {
            cf xk, xnk;
            rsplit_fwd(Z[k], Z[N - k], tw[2 * N + k], &xk, &xnk);
            Z[k] = xk;
            Z[N - k] = xnk;
        }
    }
    __syncthreads();
    return Z;
}

// Inverse: packed half spectrum Y in `buf` -> first N real samples written as
// N/2 float2 to `out` (LDS or global).  Returns the LDS buffer holding z'.
template <int LOGN>
__device__ cf *small_inverse(cf *Y, cf *other, const cf *__restrict__ tw, int tid)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    for (int k = tid; k <= N / 2; k += NT) {
        if (k == 0) {
            const cf y = Y[0];
            Y[0] = make_float2(y.x + y.y, y.x - y.y);
        } else {
            cf ck, cnk;
            rsplit_inv(Y[k], Y[N - k], tw[2 * N + k], &ck, &cnk);
            Y[k] = ck;
            Y[N - k] = cnk;
        }
    }
    __syncthreads();
    return fft_lds<true>(Y, other, LOGN, 1, N, tw, tid, NT);
}

// MAC over the FDL: Y[k] = X[k] H_0[k] + sum_{p>=1} FDL[(b-p)%P][k] H_p[k],
// accumulated p = 0..P-1 as the reference does; stores X into FDL slot b%P.
// X and Y may alias (each thread owns its bins).
template <int LOGN>
__device__ void small_mac(cf *X, const LevelJob &J, uint64_t bcur, int tid)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    const int P = J.P;
    const int slot = (int)(bcur % (uint64_t)P);
    cf *fdl_w = J.fdl + (size_t)slot * N;
    for (int k = tid; k < N; k += NT) {
        const cf x = X[k];
        fdl_w[k] = x;
        cf acc;
        const cf h0 = J.H[k];
        if (k == 0) {
            acc = make_float2(x.x * h0.x, x.y * h0.y);
            for (int p = 1; p < P; p++) {
                int s = slot - p;
                s += (s < 0) ? P : 0;
                const cf a = J.fdl[(size_t)s * N];
                const cf h = J.H[(size_t)p * N];
                acc.x += a.x * h.x;
                acc.y += a.y * h.y;
            }
        } else {
            acc = cmul(x, h0);
            int s = slot;
            int p = 1;
            for (; p + 3 < P; p += 4) {
                int s1 = s - 1; s1 += (s1 < 0) ? P : 0;
                int s2 = s1 - 1; s2 += (s2 < 0) ? P : 0;
                int s3 = s2 - 1; s3 += (s3 < 0) ? P : 0;
                int s4 = s3 - 1; s4 += (s4 < 0) ? P : 0;
                const cf a1 = J.fdl[(size_t)s1 * N + k], h1 = J.H[(size_t)p * N + k];
                const cf a2 = J.fdl[(size_t)s2 * N + k], h2 = J.H[(size_t)(p + 1) * N + k];
                const cf a3 = J.fdl[(size_t)s3 * N + k], h3 = J.H[(size_t)(p + 2) * N + k];
                const cf a4 = J.fdl[(size_t)s4 * N + k], h4 = J.H[(size_t)(p + 3) * N + k];
                acc = cmac(acc, a1, h1);
                acc = cmac(acc, a2, h2);
                acc = cmac(acc, a3, h3);
                acc = cmac(acc, a4, h4);
                s = s4;
            }
            for (; p < P; p++) {
                int s1 = s - 1; s1 += (s1 < 0) ? P : 0;
                acc = cmac(acc, J.fdl[(size_t)s1 * N + k], J.H[(size_t)p * N + k]);
                s = s1;
            }
        }
        X[k] = acc;
    }
    __syncthreads();
}

// Load the 2N-sample window of a RUN job (ring) or SPECTRUM job (taps) as N float2.
template <int LOGN>
__device__ void small_load_window(cf *buf, const LevelJob &J, int mode, long long s, int tid)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    if (mode == PC_MODE_RUN) {
        const cf *w = reinterpret_cast<const cf *>(J.ring + (s & (long long)J.ring_mask));
        for (int n = tid; n < N; n += NT) buf[n] = w[n];
    } else {
        // [0_N | taps T .. T+N) (splitVector, src/buffers.cpp:193-217)
        for (int n = tid; n < N; n += NT) {
            float v0 = 0.f, v1 = 0.f;
            const int t0 = 2 * n - N;
            if (t0 >= 0) {
                const int i0 = J.T + t0;
                v0 = (i0 < J.ring_mask) ? J.ring[i0] : 0.f;
                v1 = (i0 + 1 < J.ring_mask) ? J.ring[i0 + 1] : 0.f;
            }
            buf[n] = make_float2(v0, v1);
        }
    }
    __syncthreads();
}

// Level job kernel for levels >= 1 with N <= 2048 (and filter-spectrum setup).
//   m     : block index whose window is read
//   bcur  : FDL block counter (slot = bcur % P)
//   ypar  : double-buffer half written
template <int LOGN>
__global__ void __launch_bounds__(SmallCfg<LOGN>::NT)
k_level_small(const LevelJob *__restrict__ jobs, int N0, uint64_t m, uint64_t bcur, int ypar,
              const cf *__restrict__ tw, int mode)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    __shared__ cf bufA[N];
    __shared__ cf bufB[N];
    const int tid = threadIdx.x;
    const LevelJob J = jobs[blockIdx.x];
    const long long s = (long long)m * N - N0 - J.T;
    small_load_window<LOGN>(bufA, J, mode, s, tid);
    cf *X = small_forward<LOGN>(bufA, bufB, tw, tid);
    cf *other = (X == bufA) ? bufB : bufA;
    if (mode == PC_MODE_SPECTRUM) {
        const float sc = 1.0f / (float)(2 * N);
        for (int k = tid; k < N; k += NT) J.fdl[k] = cscale(X[k], sc);
        return;
    }
    small_mac<LOGN>(X, J, bcur, tid);
    cf *z = small_inverse<LOGN>(X, other, tw, tid);
    cf *yo = reinterpret_cast<cf *>(J.ybuf + (size_t)ypar * N);
    for (int n = tid; n < N / 2; n += NT) yo[n] = z[n];
}

// ---------------------------------------------------------------------------
// Tick kernel: input scatter + level 0 (N = N0) + output sum/mix, one
// workgroup per instance.  Mirrors PartConvMulti::run / Relaxed::run
// (src/partconvMulti.cpp:815-928, 1017-1134) for the work that is due every
// tick; levels >= 1 run as separate level-job launches on their own streams.
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ void __launch_bounds__(SmallCfg<LOGN>::NT)
k_tick(TickArgs A, const cf *__restrict__ tw)
{
    constexpr int N = SmallCfg<LOGN>::N, NT = SmallCfg<LOGN>::NT;
    __shared__ cf bufA[N];
    __shared__ cf bufB[N];
    __shared__ float xin[N];
    __shared__ float acc[N];
    __shared__ unsigned s_last;
    const int tid = threadIdx.x;
    const int i = blockIdx.x;
    const TickJob &TJ = A.jobs[i];
    const LevelJob &J = TJ.l0;
    const uint64_t f = A.f;
    const long long mask = J.ring_mask;
    const long long R = mask + 1;

    // 1. de-interleave x scale (src/partconvMulti.cpp:839-847) and ring write
    const float *inf = A.in + (long long)(f - A.fbase) * A.in_stride;
    const long long t0 = (long long)f * N;
    float *ring = const_cast<float *>(J.ring);
    for (int j = tid; j < N; j += NT) {
        const float v = inf[TJ.in_ch + (long long)j * A.nIn] * TJ.scale;
        xin[j] = v;
        const long long p = (t0 + j) & mask;
        ring[p] = v;
        ring[p + R] = v;
    }
    // 2. window = [frame f-1 from the ring | frame f]
    const float *prev = J.ring + ((t0 - N) & mask);
    __syncthreads();
    for (int n = tid; n < N; n += NT) {
        const int a = 2 * n;
        if (a < N) bufA[n] = make_float2(prev[a], prev[a + 1]);
        else bufA[n] = make_float2(xin[a - N], xin[a - N + 1]);
    }
    __syncthreads();
    // 3. level 0 of this tick
    cf *X = small_forward<LOGN>(bufA, bufB, tw, tid);
    cf *other = (X == bufA) ? bufB : bufA;
    small_mac<LOGN>(X, J, f, tid);
    cf *z = small_inverse<LOGN>(X, other, tw, tid);
    // keep y0 for relaxed output of the next tick
    float *y0w = J.ybuf + (size_t)(f & 1) * N;
    for (int n = tid; n < N / 2; n += NT) reinterpret_cast<cf *>(y0w)[n] = z[n];

    // 4. output = level 0 + completed blocks of levels >= 1
    //    strict : level L reads block m = (f+1)/k - 1, sub-block (f+1)%k
    //    relaxed: level L reads block m = f/k - 1, sub-block f%k; level 0 = tick f-1
    const float *zr = reinterpret_cast<const float *>(z);
    for (int j = tid; j < N; j += NT) {
        float a;
        if (!A.relaxed) {
            a = zr[j];
        } else {
            a = (f > 0) ? J.ybuf[(size_t)((f - 1) & 1) * N + j] : 0.f;
        }
        const uint64_t fr = A.relaxed ? f : f + 1;
        for (int L = 0; L < TJ.nlev; L++) {
            const int ks = TJ.kshift[L];
            const uint64_t q = fr >> ks;
            if (q == 0) continue;
            const uint64_t mb = q - 1;
            const uint64_t sub = fr & ((1ull << ks) - 1);
            a += TJ.ybuf[L][(size_t)(mb & 1) * TJ.NL[L] + sub * N + j];
        }
        acc[j] = a;
    }
    __syncthreads();

    // 5. mix (src/partconvMulti.cpp:906-921)
    if (A.mix == 1) {
        if (!TJ.mute) {
            float *o = A.out + (long long)(f - A.fbase) * A.out_stride + TJ.out_ch;
            for (int j = tid; j < N; j += NT) o[(long long)j * A.nOut] = acc[j];
        }
    } else {
        float *yi = A.y_inst + (size_t)i * N;
        for (int j = tid; j < N; j += NT) yi[j] = acc[j];
    }
    if (A.mix == 0) return;

    // arrival ticket: the last block finishes the mix and raises the flag
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        const unsigned t = atomicAdd(A.ticket, 1u);
        s_last = (t == (unsigned)(A.njobs - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (A.mix == 2) {
        for (int e = tid; e < A.nOut * N; e += NT) {
            const int c = e / N, j = e - c * N;
            float sum = 0.f;
            for (int q = A.src_off[c]; q < A.src_off[c + 1]; q++)
                sum += __builtin_nontemporal_load(&A.y_inst[(size_t)A.src_list[q] * N + j]);
            A.out[(long long)(f - A.fbase) * A.out_stride + c + (long long)j * A.nOut] = sum;
        }
    }
    __syncthreads();
    if (tid == 0) {
        *A.ticket = 0u;
        if (A.done_flag) {
            __threadfence_system();
            __hip_atomic_store(A.done_flag, (unsigned long long)(f + 1), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Plain-mode mix: out[c + j nOut] = sum over unmuted instances feeding c, in
// instance order (src/partconvMulti.cpp:906-921).  One launch per tick; the
// last block to finish (device counter `cnt`) raises the host flag.
__global__ void k_mix(const float *__restrict__ y_inst, const int *__restrict__ src_off,
                      const int *__restrict__ src_list, int N0, int nOut, float *out,
                      long long out_off, unsigned *cnt, unsigned long long *done_flag, uint64_t f)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nOut * N0) {
        const int c = e / N0, j = e - c * N0;
        float sum = 0.f;
        for (int q = src_off[c]; q < src_off[c + 1]; q++)
            sum += y_inst[(size_t)src_list[q] * N0 + j];
        out[out_off + c + (long long)j * nOut] = sum;
    }
    if (done_flag) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const unsigned t = atomicAdd(cnt, 1u);
            if (t == gridDim.x - 1) {
                *cnt = 0u;
                __threadfence_system();
                __hip_atomic_store(done_flag, (unsigned long long)(f + 1), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Big levels (N = A*B >= 4096): four-step transform over three launches with
// the FDL multiply-accumulate fused into the middle one.
//   z[n] = x[2n] + i x[2n+1],  n = a + A b,  bins k = c + B d.
//   k_big1: for each a, length-B DFT over b, twiddle W_N^{a c}     -> S[c][a]
//   k_big2: per column pair {c, B-c} (u = 0 takes {0, B/2}): length-A DFT over a
//           -> Z[c + B d]; real split (pairs k, N-k stay inside the pair);
//           X -> FDL slot; MAC over partitions; inverse split; inverse
//           length-A DFT over d; twiddle W_N^{-a c}                -> S[c][a]
//   k_big3: for each a, inverse length-B DFT over c -> z'[a + A b]; keep the
//           first N real samples (b < B/2).
// Spectrum slot of bin (c + B d) for workgroup u: u*2A + d (first column) or
// u*2A + A + d (second column); H is produced by the same kernels, so the FDL
// and H share this order.
// ---------------------------------------------------------------------------
#define BIG_G 16      // sequences per workgroup in k_big1 / k_big3
#define BIG_NT 256

struct BigShape {
    int logA, logB;
};

__global__ void __launch_bounds__(BIG_NT)
k_big1(const LevelJob *__restrict__ jobs, BigShape sh, int N0, uint64_t m,
       const cf *__restrict__ tw, int mode)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = B + 1;
    cf *bufA = lds, *bufB = lds + BIG_G * sst;
    const int groups = A / BIG_G;
    const int jb = blockIdx.x / groups, g = blockIdx.x - jb * groups;
    const LevelJob J = jobs[jb];
    const int tid = threadIdx.x;
    const int a0 = g * BIG_G;
    if (mode == PC_MODE_RUN) {
        const long long s = (long long)m * N - N0 - J.T;
        const cf *w = reinterpret_cast<const cf *>(J.ring + (s & (long long)J.ring_mask));
        for (int e = tid; e < BIG_G * B; e += BIG_NT) {
            const int b = e / BIG_G, ai = e - b * BIG_G;
            bufA[ai * sst + b] = w[a0 + ai + A * b];
        }
    } else {
        for (int e = tid; e < BIG_G * B; e += BIG_NT) {
            const int b = e / BIG_G, ai = e - b * BIG_G;
            const int n = a0 + ai + A * b;
            float v0 = 0.f, v1 = 0.f;
            const int t0 = 2 * n - N;
            if (t0 >= 0) {
                const int i0 = J.T + t0;
                v0 = (i0 < J.ring_mask) ? J.ring[i0] : 0.f;
                v1 = (i0 + 1 < J.ring_mask) ? J.ring[i0 + 1] : 0.f;
            }
            bufA[ai * sst + b] = make_float2(v0, v1);
        }
    }
    __syncthreads();
    cf *R = fft_lds<false>(bufA, bufB, sh.logB, BIG_G, sst, tw, tid, BIG_NT);
    cf *S = J.scratch;
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int c = e / BIG_G, ai = e - c * BIG_G;
        const int a = a0 + ai;
        const cf v = cmul(R[ai * sst + c], tw[N + ((a * c) & (N - 1))]);
        S[(size_t)c * A + a] = v;
    }
}

__global__ void __launch_bounds__(BIG_NT)
k_big2(const LevelJob *__restrict__ jobs, BigShape sh, uint64_t bcur, const cf *__restrict__ tw,
       int mode)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = A + 1;
    cf *bufA = lds, *bufB = lds + 2 * sst;
    const int pairs = B / 2;
    const int jb = blockIdx.x / pairs, u = blockIdx.x - jb * pairs;
    const LevelJob J = jobs[jb];
    const int tid = threadIdx.x;
    const int c1 = u, c2 = u ? (B - u) : (B / 2);
    cf *S = J.scratch;
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, a = e & (A - 1);
        bufA[q * sst + a] = S[(size_t)(q ? c2 : c1) * A + a];
    }
    __syncthreads();
    cf *Z = fft_lds<false>(bufA, bufB, sh.logA, 2, sst, tw, tid, BIG_NT);
    cf *W = (Z == bufA) ? bufB : bufA;   // X in slot order (padded by column)
    // real split; results into W at slot positions (col0: d, col1: sst + d)
    if (u != 0) {
        for (int d = tid; d < A; d += BIG_NT) {
            const int k = c1 + B * d;
            cf xk, xnk;
            rsplit_fwd(Z[d], Z[sst + (A - 1 - d)], tw[2 * N + k], &xk, &xnk);
            W[d] = xk;
            W[sst + (A - 1 - d)] = xnk;
        }
    } else {
        for (int e = tid; e < A / 2 + 1 + A / 2; e += BIG_NT) {
            if (e <= A / 2) {
                const int d = e;
                if (d == 0) {
                    const cf zz = Z[0];
                    W[0] = make_float2(zz.x + zz.y, zz.x - zz.y);
                } else {
                    const int k = B * d;
                    cf xk, xnk;
                    rsplit_fwd(Z[d], Z[A - d], tw[2 * N + k], &xk, &xnk);
                    W[d] = xk;
                    W[A - d] = xnk;
                }
            } else {
                const int d = e - (A / 2 + 1);
                const int k = B / 2 + B * d;
                cf xk, xnk;
                rsplit_fwd(Z[sst + d], Z[sst + (A - 1 - d)], tw[2 * N + k], &xk, &xnk);
                W[sst + d] = xk;
                W[sst + (A - 1 - d)] = xnk;
            }
        }
    }
    __syncthreads();
    const size_t base = (size_t)u * 2 * A;
    if (mode == PC_MODE_SPECTRUM) {
        const float sc = 1.0f / (float)(2 * N);
        for (int e = tid; e < 2 * A; e += BIG_NT) {
            const int q = e >> sh.logA, d = e & (A - 1);
            J.fdl[base + e] = cscale(W[q * sst + d], sc);
        }
        return;
    }
    // FDL store + MAC (slot order, contiguous per workgroup)
    const int P = J.P;
    const int slot = (int)(bcur % (uint64_t)P);
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, d = e & (A - 1);
        const cf x = W[q * sst + d];
        const size_t bin = base + e;
        J.fdl[(size_t)slot * N + bin] = x;
        const cf h0 = J.H[bin];
        cf acc;
        if (bin == 0) {
            acc = make_float2(x.x * h0.x, x.y * h0.y);
            for (int p = 1; p < P; p++) {
                int s = slot - p;
                s += (s < 0) ? P : 0;
                const cf a = J.fdl[(size_t)s * N];
                const cf h = J.H[(size_t)p * N];
                acc.x += a.x * h.x;
                acc.y += a.y * h.y;
            }
        } else {
            acc = cmul(x, h0);
            int s = slot;
            int p = 1;
            for (; p + 3 < P; p += 4) {
                int s1 = s - 1; s1 += (s1 < 0) ? P : 0;
                int s2 = s1 - 1; s2 += (s2 < 0) ? P : 0;
                int s3 = s2 - 1; s3 += (s3 < 0) ? P : 0;
                int s4 = s3 - 1; s4 += (s4 < 0) ? P : 0;
                const cf a1 = J.fdl[(size_t)s1 * N + bin], h1 = J.H[(size_t)p * N + bin];
                const cf a2 = J.fdl[(size_t)s2 * N + bin], h2 = J.H[(size_t)(p + 1) * N + bin];
                const cf a3 = J.fdl[(size_t)s3 * N + bin], h3 = J.H[(size_t)(p + 2) * N + bin];
                const cf a4 = J.fdl[(size_t)s4 * N + bin], h4 = J.H[(size_t)(p + 3) * N + bin];
                acc = cmac(acc, a1, h1);
                acc = cmac(acc, a2, h2);
                acc = cmac(acc, a3, h3);
                acc = cmac(acc, a4, h4);
                s = s4;
            }
            for (; p < P; p++) {
                int s1 = s - 1; s1 += (s1 < 0) ? P : 0;
                acc = cmac(acc, J.fdl[(size_t)s1 * N + bin], J.H[(size_t)p * N + bin]);
                s = s1;
            }
        }
        W[q * sst + d] = acc;
    }
    __syncthreads();
    // inverse split back into Z (column sequences over d)
    if (u != 0) {
        for (int d = tid; d < A; d += BIG_NT) {
            const int k = c1 + B * d;
            cf ck, cnk;
            rsplit_inv(W[d], W[sst + (A - 1 - d)], tw[2 * N + k], &ck, &cnk);
            Z[d] = ck;
            Z[sst + (A - 1 - d)] = cnk;
        }
    } else {
        for (int e = tid; e < A / 2 + 1 + A / 2; e += BIG_NT) {
            if (e <= A / 2) {
                const int d = e;
                if (d == 0) {
                    const cf y = W[0];
                    Z[0] = make_float2(y.x + y.y, y.x - y.y);
                } else {
                    const int k = B * d;
                    cf ck, cnk;
                    rsplit_inv(W[d], W[A - d], tw[2 * N + k], &ck, &cnk);
                    Z[d] = ck;
                    Z[A - d] = cnk;
                }
            } else {
                const int d = e - (A / 2 + 1);
                const int k = B / 2 + B * d;
                cf ck, cnk;
                rsplit_inv(W[sst + d], W[sst + (A - 1 - d)], tw[2 * N + k], &ck, &cnk);
                Z[sst + d] = ck;
                Z[sst + (A - 1 - d)] = cnk;
            }
        }
    }
    __syncthreads();
    cf *U = fft_lds<true>(Z, W, sh.logA, 2, sst, tw, tid, BIG_NT);
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, a = e & (A - 1);
        const int c = q ? c2 : c1;
        const cf v = cmulc(U[q * sst + a], tw[N + ((a * c) & (N - 1))]);
        S[(size_t)c * A + a] = v;
    }
}

__global__ void __launch_bounds__(BIG_NT)
k_big3(const LevelJob *__restrict__ jobs, BigShape sh, int ypar, const cf *__restrict__ tw)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB;
    const int N = A * B;
    const int sst = B + 1;
    cf *bufA = lds, *bufB = lds + BIG_G * sst;
    const int groups = A / BIG_G;
    const int jb = blockIdx.x / groups, g = blockIdx.x - jb * groups;
    const LevelJob J = jobs[jb];
    const int tid = threadIdx.x;
    const int a0 = g * BIG_G;
    const cf *S = J.scratch;
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int c = e / BIG_G, ai = e - c * BIG_G;
        bufA[ai * sst + c] = S[(size_t)c * A + a0 + ai];
    }
    __syncthreads();
    cf *R = fft_lds<true>(bufA, bufB, sh.logB, BIG_G, sst, tw, tid, BIG_NT);
    cf *yo = reinterpret_cast<cf *>(J.ybuf + (size_t)ypar * N);
    for (int e = tid; e < BIG_G * (B / 2); e += BIG_NT) {
        const int b = e / BIG_G, ai = e - b * BIG_G;
        yo[a0 + ai + A * b] = R[ai * sst + b];
    }
}
