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

// Global-memory view of a pointer.  Job descriptors are read from LDS, where
// the compiler can no longer see that their pointers address global memory;
// without the cast it emits flat loads / stores, which also count against the
// LDS wait counter and so hold up the LDS work that follows them.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// float2 loads / stores through a global-memory pointer (the HIP float2 class
// has no address-space-qualified operators; the built-in vector type does).
typedef float pc_v2f __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 gload2(const float2 *p)
{
    const pc_v2f v = *(const __attribute__((address_space(1))) pc_v2f *)p;
    return make_float2(v.x, v.y);
}

__device__ __forceinline__ void gstore2(float2 *p, float2 v)
{
    *(__attribute__((address_space(1))) pc_v2f *)p = pc_v2f{v.x, v.y};
}

enum { PC_MODE_RUN = 0, PC_MODE_SPECTRUM = 1 };

struct LevelJob {
    const float *ring;   // RUN: instance input ring (2R floats); SPECTRUM: IR taps
    cf *fdl;             // RUN: P x N FDL; SPECTRUM: destination H row
    const cf *H;         // P x N filter spectra
    float *ybuf;         // 2 x N output double buffer
    cf *scratch;         // big levels: N float2 work space
    cf *S;               // N float2: delay-line sum of partitions 1..P-1 for the next block
    int P;
    int T;               // RUN: taps before this level; SPECTRUM: first tap of the row
    int ring_mask;       // RUN: R - 1; SPECTRUM: number of IR taps
};

struct TickJob {
    LevelJob l0;
    int in_ch, out_ch, mute;
    float scale;
    int nlev;                              // levels >= 1 of this instance
    int kshift[PC_MAX_LEVELS];             // log2(N_L / N0)
    int NL[PC_MAX_LEVELS];
    const float *ybuf[PC_MAX_LEVELS];
    int wslot[PC_MAX_LEVELS];              // persistent runs: level-worker slot, -1 = host-launched level
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
    int in_cs, in_ss;      // sample j of channel c at in[c * in_cs + j * in_ss] (interleaved: 1, nIn)
    float *out;            // output frame(s)
    long long out_stride;
    int nOut;
    int out_cs, out_ss;    // as in_cs, in_ss (interleaved: 1, nOut; host staging is planar: N0, 1)
    int relaxed;
    int mix;               // 0: write y_inst only (plain mode), 1: direct, 2: reduce
    float *y_inst;         // njobs x N0 per-instance outputs (mix == 0 or 2)
    const int *src_off;    // reduce: per output channel offsets into src_list (nOut+1)
    const int *src_list;   // reduce: job indices per output channel
    unsigned *ticket;      // reduce: arrival counter (reset by the last block)
    unsigned long long *done_flag;  // reduce: host-visible completion flag (may be null)
    unsigned long long *wg_flags;   // direct: one host-visible completion flag per workgroup,
    int wg_stride;                  // wg_stride words apart
    int fence_sys;                  // output lives in host memory: system-scope release per tick
    unsigned long long fence_mask;  // else: device-scope release at ticks with (f & mask) == 0
    unsigned *ticket_ch;            // reduce: arrival counter per output channel
    unsigned *ticket_grp;           // direct: arrival counters per group of PC_TICKET_GROUP workgroups
    int nactive;                    // reduce: channels with at least one unmuted source
    const int *empty_ch;            // reduce: channels with no source (zeroed each tick)
    int nempty;
    unsigned long long *trace;      // optional timestamps (persistent runs), PC_TRACE_W per tick
    long long trace_cap;
    // host sessions with tagged input words (ll != 0): the input frame as
    // planar [channel][N0] words (sample bits | tag << 32), see ll_tag; they
    // are the release (no release word, no system-scope acquire)
    int ll;
    const unsigned long long *ll_in;
};

// Tag of tick f in the tagged input words of host sessions.  A word carries
// one sample in its low half and the tick's tag in its high half, written by
// one 8-byte store, so the kernel, which polls the words themselves, sees the
// sample whenever it sees the tag: no separate release word to wait for before
// the frame can be fetched.  Tags run 1 .. 2^32 - 2: consecutive ticks never
// share one, and 0 marks a cleared word.
__host__ __device__ inline unsigned ll_tag(uint64_t f) { return (unsigned)(f % 4294967294ull) + 1u; }
__host__ __device__ inline unsigned long long ll_word(float v, unsigned tag)
{
    return ((unsigned long long)tag << 32) | (unsigned long long)__builtin_bit_cast(unsigned, v);
}

// Completion of many workgroups: arrivals are counted per group of
// PC_TICKET_GROUP workgroups, and the last of each group arrives at a global
// counter (counters PC_TICKET_STRIDE words apart: own cache lines; the global
// one after the groups).  The counters only grow, by a power of two per tick
// (PC_TICKET_GROUP per group: the first member of a short last group adds
// the missing weight; the next power of two >= groups at the global counter,
// the first group adding the missing weight), so the arrival that completes a
// tick is the one that makes its counter a multiple of that power of two --
// modulo 2^32 as well -- and nothing is reset between ticks.
#define PC_TICKET_GROUP 16
#define PC_TICKET_STRIDE 32
// Per-workgroup completion flags of persistent runs in pinned host memory, one
// 64-byte line each (in words): the workgroups' writes then never land in the
// same host line (measured 0.3-0.8 us less per host round trip at 16
// workgroups).  Per-tick launches keep them packed (measured faster there).
#define PC_WG_FLAG_STRIDE 8

// Persistent-run trace, per tick (s_memrealtime, 100 MHz): release seen and
// output published by workgroups 0, njobs/2 and njobs-1; completion raised;
// workgroup 0 ready for the next release.
#define PC_TRACE_W 8

__device__ __forceinline__ void trace_stamp(const TickArgs &A, uint64_t f, int slot)
{
    const long long t = (long long)(f - (uint64_t)A.fbase);
    if (A.trace && t < A.trace_cap) A.trace[PC_TRACE_W * t + slot] = __builtin_amdgcn_s_memrealtime();
}

// ---------------------------------------------------------------------------
// Small levels: one workgroup runs the whole job in LDS (N <= 2048).
// ---------------------------------------------------------------------------

template <int LOGN>
struct SmallCfg {
    static constexpr int N = 1 << LOGN;
    static constexpr int NT = (N / 4 < 64) ? 64 : ((N / 4 > 512) ? 512 : N / 4);  // threads
    static constexpr int IPT = (N + NT - 1) / NT;          // bins / samples per thread
};

// Forward real FFT of the 2N-sample window already packed in bufA as N float2;
// returns the buffer holding X (packed: X[0] = (X_0, X_N)).
template <int LOGN, int NT>
__device__ cf *small_forward(cf *bufA, cf *bufB, const cf *__restrict__ tw, int tid)
{
    constexpr int N = 1 << LOGN;
    cf *Z = fft_lds<false>(bufA, bufB, LOGN, 1, N, tw, tid, NT);
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

// Inverse: packed half spectrum in Y -> z' with z'[n] = (x[2n], x[2n+1]);
// returns the LDS buffer holding z' (the first N/2 entries are the first N
// output samples).
template <int LOGN, int NT>
__device__ cf *small_inverse(cf *Y, cf *other, const cf *__restrict__ tw, int tid)
{
    constexpr int N = 1 << LOGN;
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

// acc += a*b for a packed bin: bin 0 holds the two real bins (DC, Nyquist).
__device__ __forceinline__ cf pmac(cf acc, cf a, cf b, bool bin0)
{
    if (bin0) {
        acc.x += a.x * b.x;
        acc.y += a.y * b.y;
        return acc;
    }
    return cmac(acc, a, b);
}
__device__ __forceinline__ cf pmul(cf a, cf b, bool bin0)
{
    return bin0 ? make_float2(a.x * b.x, a.y * b.y) : cmul(a, b);
}

// Delay-line sum step for two bins held as one float4 (bins 2k, 2k+1):
// acc += a * h, with bin 0 of a row packing the real DC and Nyquist values
// (bin0: componentwise products for the first pair).  Written as a fixed
// sequence of fused multiply-adds -- re: fma(a.x, h.x, acc) then
// fma(-a.y, h.y, .), im: fma(a.x, h.y, acc) then fma(a.y, h.x, .) -- so every
// kernel that forms a delay-line sum (k_partial of any width, the persistent
// kernel's level workers) rounds identically, whatever contraction the
// compiler would choose for the plain expression in each of them.
__device__ __forceinline__ float4 dl_mac2(float4 acc, float4 a, float4 h, bool bin0)
{
    if (bin0) {
        acc.x = fmaf(a.x, h.x, acc.x);
        acc.y = fmaf(a.y, h.y, acc.y);
    } else {
        acc.x = fmaf(-a.y, h.y, fmaf(a.x, h.x, acc.x));
        acc.y = fmaf(a.y, h.x, fmaf(a.x, h.y, acc.y));
    }
    acc.z = fmaf(-a.w, h.w, fmaf(a.z, h.z, acc.z));
    acc.w = fmaf(a.w, h.z, fmaf(a.z, h.w, acc.w));
    return acc;
}

// One bin of dl_mac2: the same fused multiply-adds in the same order, so a sum
// formed bin by bin rounds exactly as the float4 form.
__device__ __forceinline__ cf dl_mac1(cf acc, cf a, cf h, bool bin0)
{
    if (bin0) {
        acc.x = fmaf(a.x, h.x, acc.x);
        acc.y = fmaf(a.y, h.y, acc.y);
    } else {
        acc.x = fmaf(-a.y, h.y, fmaf(a.x, h.x, acc.x));
        acc.y = fmaf(a.y, h.x, fmaf(a.x, h.y, acc.y));
    }
    return acc;
}

// Block output spectrum Y[k] = X[k] H_0[k] + S[k], where S (written by
// k_partial after the previous block) holds sum_{p=1}^{P-1} X_{b-p}[k] H_p[k].
// Stores X into FDL row b % P for later blocks.  X and Y alias.
template <int LOGN, int NT>
__device__ __forceinline__ void small_prefetch(const LevelJob &J, int tid, cf (&h)[(((1 << LOGN) + NT - 1) / NT)],
                                               cf (&sv)[(((1 << LOGN) + NT - 1) / NT)])
{
    constexpr int N = 1 << LOGN, IPT = (N + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        if (k < N) {
            h[q] = J.H[k];
            sv[q] = J.S[k];
        }
    }
}

template <int LOGN, int NT>
__device__ void small_mac(cf *X, const LevelJob &J, uint64_t bcur, int tid,
                          const cf (&h)[(((1 << LOGN) + NT - 1) / NT)], const cf (&sv)[(((1 << LOGN) + NT - 1) / NT)])
{
    constexpr int N = 1 << LOGN, IPT = (N + NT - 1) / NT;
    const int slot = (int)(bcur % (uint64_t)J.P);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        if (k < N) {
            const cf x = X[k];
            J.fdl[(size_t)slot * N + k] = x;
            cf y = pmul(x, h[q], k == 0);
            y.x += sv[q].x;
            y.y += sv[q].y;
            X[k] = y;
        }
    }
    __syncthreads();
}

// Load the 2N-sample window of a RUN job (ring) or SPECTRUM job (taps) as N float2.
template <int LOGN, int NT>
__device__ void small_load_window(cf *buf, const LevelJob &J, int mode, long long s, int tid)
{
    constexpr int N = 1 << LOGN;
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
    constexpr int IPT = (N + NT - 1) / NT;
    const LevelJob J = jobs[blockIdx.x];
    const long long s = (long long)m * N - N0 - J.T;
    cf h[IPT], sv[IPT];
    if (mode == PC_MODE_RUN) small_prefetch<LOGN, NT>(J, tid, h, sv);
    small_load_window<LOGN, NT>(bufA, J, mode, s, tid);
    cf *X = small_forward<LOGN, NT>(bufA, bufB, tw, tid);
    cf *other = (X == bufA) ? bufB : bufA;
    if (mode == PC_MODE_SPECTRUM) {
        const float sc = 1.0f / (float)(2 * N);
        for (int k = tid; k < N; k += NT) J.fdl[k] = cscale(X[k], sc);
        return;
    }
    small_mac<LOGN, NT>(X, J, bcur, tid, h, sv);
    cf *z = small_inverse<LOGN, NT>(X, other, tw, tid);
    cf *yo = reinterpret_cast<cf *>(J.ybuf + (size_t)ypar * N);
    for (int n = tid; n < N / 2; n += NT) yo[n] = z[n];
}

// ---------------------------------------------------------------------------
// Tick: input scatter + level 0 (N = N0) + output sum/mix, one workgroup per
// instance.  Mirrors PartConvMulti::run / Relaxed::run
// (src/partconvMulti.cpp:815-928, 1017-1134) for the work due every tick;
// levels >= 1 run as level-job launches on their own streams.
//
// Latency split: phase A (on the tick's critical path) needs only the new
// frame, the previous frame, H_0 and S, the delay-line sum of partitions
// 1..P-1, which does not depend on the new block.  Phase B (after the tick's
// output is published) stores X into the FDL and computes S for the next tick:
//   S_{f+1}[k] = sum_{p=1}^{P-1} X_{f+1-p}[k] H_p[k],   Y_f = X_f H_0 + S_f.
// S is kept in the level-0 job's scratch row (N0 float2).
// ---------------------------------------------------------------------------
template <int LOGN>
struct TickCfg {
    static constexpr int N = 1 << LOGN;
    // 64 threads for N = 64 (the register-resident 64-point form); 256 for
    // N = 128 (the tick itself uses half of them, but the short-period level
    // blocks that tick workgroups run in self mode go twice as wide: measured
    // 6% more throughput at 512 instances) and from N = 512; 128 otherwise (two
    // wavefronts: more instances' tick workgroups fit beside the level kernels)
    static constexpr int NT = (N <= 64) ? 64 : ((N >= 512 || N == 128) ? 256 : 128);
    static constexpr int IPT = (N + NT - 1) / NT;
    // persistent kernel: waves per SIMD it must keep (<= 512 / MINW VGPRs).  At
    // N = 128 four 4-wave tick workgroups per CU keep 512 instances within half
    // of the device (the other half is left to the host-launched levels)
    static constexpr int MINW = (N == 128) ? 4 : 1;
};

// Per-thread state of one instance's level 0 that does not depend on the new
// frame: H_0 (constant), S (from the previous tick's phase B) and the previous
// frame (left half of the window).  k_tick loads it each launch; k_persist
// keeps it in registers across ticks.
// A level of block size 2^LOGNL runs in the persistent kernel's workgroups
// (NT threads) only if each thread holds at most 8 of its bins: the larger
// blocks would raise the kernel's register count for every workgroup.
template <int LOGNL, int NT>
constexpr bool worker_fits() { return LOGNL <= 11 && (1 << LOGNL) / NT <= 8; }

template <int LOGN>
struct TickState {
    cf h0[TickCfg<LOGN>::IPT];
    cf sp[TickCfg<LOGN>::IPT];
    float pv[TickCfg<LOGN>::IPT];
};

template <int LOGN>
__device__ __forceinline__ void tick_load_state(const LevelJob &J, uint64_t f, int tid, TickState<LOGN> &st)
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, IPT = TickCfg<LOGN>::IPT;
    const float *prev = J.ring + (((long long)f * N - N) & (long long)J.ring_mask);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int j = tid + q * NT;
        if (j < N) {
            st.h0[q] = J.H[j];
            st.sp[q] = J.S[j];
            st.pv[q] = prev[j];
        }
    }
}

// Phase A (the tick's critical path): new frame -> ring, window
// [previous frame | frame f] -> r2c -> Y = X H_0 + S -> c2r -> this instance's
// output = level 0 + the level blocks due (strict: block (f+1)/k - 1,
// sub-block (f+1) % k; relaxed: block f/k - 1, sub-block f % k, and level 0
// of tick f-1).  Leaves the output in acc, X_f in xr, and the scaled frame in
// st.pv (the next tick's previous frame).
template <int LOGN>
__device__ __forceinline__ void tick_phaseA(const TickJob &TJ, const TickArgs &A, uint64_t f, cf *bufA,
                                            cf *bufB, const cf *__restrict__ tw, int tid, TickState<LOGN> &st,
                                            float (&acc)[TickCfg<LOGN>::IPT], cf (&xr)[TickCfg<LOGN>::IPT],
                                            const float *llx = nullptr)
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, IPT = TickCfg<LOGN>::IPT;
    const LevelJob &J = TJ.l0;
    const long long mask = J.ring_mask;
    const long long R = mask + 1;
    const long long t0 = (long long)f * N;
    const float *inf = A.in + (long long)(f - A.fbase) * A.in_stride;
    float xv[IPT], lev[IPT], y0p[IPT];
    const float scale = TJ.scale;
    const int inch = TJ.in_ch;
    const uint64_t fr = A.relaxed ? f : f + 1;
    // ---- every load first: the frame and the level blocks due
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int j = tid + q * NT;
        if (j < N) {
            xv[q] = llx ? llx[q] : inf[(long long)inch * A.in_cs + (long long)j * A.in_ss];
            y0p[q] = (A.relaxed && f > 0) ? gmem(J.ybuf)[(size_t)((f - 1) & 1) * N + j] : 0.f;
            lev[q] = 0.f;
        }
    }
    // level blocks CH at a time: all loads of a group in flight together, then
    // summed in level order
    constexpr int CH = (IPT >= 8) ? 2 : 16 / IPT;
    for (int L0 = 0; L0 < TJ.nlev; L0 += CH) {
        float v[CH][IPT];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int L = L0 + c;
            const uint64_t qb = (L < TJ.nlev) ? (fr >> TJ.kshift[L]) : 0;
            __attribute__((address_space(1))) const float *src = nullptr;
            if (qb != 0) {
                const uint64_t sub = fr & ((1ull << TJ.kshift[L]) - 1);
                src = gmem(TJ.ybuf[L]) + (size_t)((qb - 1) & 1) * TJ.NL[L] + sub * N;
            }
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const int j = tid + q * NT;
                v[c][q] = (src && j < N) ? src[j] : 0.f;
            }
        }
#pragma unroll
        for (int c = 0; c < CH; c++)
#pragma unroll
            for (int q = 0; q < IPT; q++) lev[q] += v[c][q];
    }
    // ---- scale (src/partconvMulti.cpp:839-847), window in LDS
    float *wf = reinterpret_cast<float *>(bufA);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int j = tid + q * NT;
        if (j < N) {
            const float v = xv[q] * scale;
            wf[j] = st.pv[q];
            wf[N + j] = v;
            st.pv[q] = v;
        }
    }
    __syncthreads();
    cf *X = small_forward<LOGN, NT>(bufA, bufB, tw, tid);
    cf *other = (X == bufA) ? bufB : bufA;
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        if (k < N) {
            const cf x = X[k];
            xr[q] = x;
            cf y = pmul(x, st.h0[q], k == 0);
            y.x += st.sp[q].x;
            y.y += st.sp[q].y;
            X[k] = y;
        }
    }
    __syncthreads();
    cf *z = small_inverse<LOGN, NT>(X, other, tw, tid);
    const float *zf = reinterpret_cast<const float *>(z);
    // ring write of the scaled frame (st.pv), issued after the transforms so the
    // stores do not hold up their barriers
    auto ring = gmem(const_cast<float *>(J.ring));
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int j = tid + q * NT;
        if (j < N) {
            const long long p = (t0 + j) & mask;
            ring[p] = st.pv[q];
            ring[p + R] = st.pv[q];
            if (A.relaxed) {
                gmem(J.ybuf)[(size_t)(f & 1) * N + j] = zf[j];   // read back at tick f+1
                acc[q] = y0p[q] + lev[q];
            } else {
                acc[q] = zf[j] + lev[q];
            }
        }
    }
}

// Phase B (after the output is published): FDL row f % P <- X_f and
//   S <- sum_{p=1}^{P-1} X_{f+1-p} H_p   (for tick f+1; kept in st.sp and J.S).
template <int LOGN>
__device__ __forceinline__ void tick_phaseB(const TickJob &TJ, uint64_t f, int tid,
                                            const cf (&xr)[TickCfg<LOGN>::IPT], TickState<LOGN> &st)
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, IPT = TickCfg<LOGN>::IPT;
    const LevelJob &J = TJ.l0;
    const int P = J.P;
    const int slot = (int)(f % (uint64_t)P);
    cf s[IPT];
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        if (k < N) {
            J.fdl[(size_t)slot * N + k] = xr[q];
            s[q] = (P > 1) ? pmul(xr[q], J.H[(size_t)N + k], k == 0) : make_float2(0.f, 0.f);
        }
    }
    // partitions 2..P-1 in batches of B: all loads of a batch in flight
    // together, accumulated in partition order
    constexpr int B = (IPT >= 4) ? 2 : 8 / IPT;
    for (int p0 = 2; p0 < P; p0 += B) {
        cf av[B][IPT], hv[B][IPT];
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int p = p0 + b;
            int r = slot + 1 - p;               // (f + 1 - p) mod P
            if (r < 0) r += P;
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const int k = tid + q * NT;
                if (p < P && k < N) {
                    av[b][q] = gload2(J.fdl + (size_t)r * N + k);
                    hv[b][q] = gload2(J.H + (size_t)p * N + k);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const int k = tid + q * NT;
                if (p0 + b < P && k < N) s[q] = pmac(s[q], av[b][q], hv[b][q], k == 0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        if (k < N) {
            J.S[k] = s[q];
            st.sp[q] = s[q];
        }
    }
}

// Every thread's outstanding vector stores complete (reached L2) before the
// barrier, so one thread's device-scope release after it covers the whole
// workgroup's stores.
__device__ __forceinline__ void wg_stores_done()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

// Publish this instance's output.
//   direct mix (each output channel fed by exactly one unmuted instance): every
//     workgroup writes its channel and then its own completion flag
//     wg_flags[i] = f + 1 (system-scope release); no cross-workgroup step.
//   reduce / plain: the per-instance output goes to y_inst; in reduce mode the
//     last workgroup to arrive (device ticket) sums the channels in instance
//     order (src/partconvMulti.cpp:906-921) and raises done_flag.
template <int LOGN>
__device__ __forceinline__ void tick_arrive(const TickJob &TJ, const TickArgs &A, uint64_t f, int tid,
                                            const float (&acc)[TickCfg<LOGN>::IPT], unsigned *s_last)
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, IPT = TickCfg<LOGN>::IPT;
    const int i = blockIdx.x;
    if (A.mix == 1) {
        if (!TJ.mute) {
            float *o = A.out + (long long)(f - A.fbase) * A.out_stride + (long long)TJ.out_ch * A.out_cs;
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const int j = tid + q * NT;
                if (j < N) o[(long long)j * A.out_ss] = acc[q];
            }
        }
        if (A.wg_flags) {
            __syncthreads();
            if (tid == 0) {
                // What the flag publishes: host-memory output (system scope), or the
                // ring frames that level jobs launched after this tick will read
                // from other XCDs (device scope, at level-start ticks); otherwise
                // only this tick's completion (workgroup scope: the stores issued).
                if (A.fence_sys) __threadfence_system();
                else if ((f & A.fence_mask) == 0) __threadfence();
                else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __hip_atomic_store(A.wg_flags + (size_t)i * A.wg_stride, (unsigned long long)(f + 1), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            }
            return;
        }
    } else {
        float *yi = A.y_inst + (size_t)i * N;
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int j = tid + q * NT;
            if (j < N) yi[j] = acc[q];
        }
        if (A.mix == 0) return;
    }
    if (A.mix == 1) {
        // direct mix without per-workgroup flags: device ticket, last one raises done_flag.
        // Device-scope release only where a reader on another XCD follows: the
        // ring at level-start ticks, or host-memory output
        wg_stores_done();
        if (tid == 0) {
            if (A.fence_sys) __threadfence_system();
            else if ((f & A.fence_mask) == 0) __threadfence();
            const int grp = i / PC_TICKET_GROUP;
            const int ngrp = (A.njobs + PC_TICKET_GROUP - 1) / PC_TICKET_GROUP;
            const int gsz = min(PC_TICKET_GROUP, A.njobs - grp * PC_TICKET_GROUP);
            const unsigned wg = (i == grp * PC_TICKET_GROUP) ? (unsigned)(PC_TICKET_GROUP - gsz + 1) : 1u;
            const unsigned tg = atomicAdd(A.ticket_grp + (size_t)grp * PC_TICKET_STRIDE, wg) + wg;
            int last = 0;
            if ((tg & (PC_TICKET_GROUP - 1)) == 0) {
                const unsigned p2 = 1u << (32 - __clz(ngrp - 1));   // next power of two >= ngrp (ngrp > 1)
                const unsigned pw = (ngrp == 1) ? 1u : p2;
                const unsigned w = (grp == 0) ? pw - (unsigned)ngrp + 1u : 1u;
                const unsigned t = atomicAdd(A.ticket_grp + (size_t)ngrp * PC_TICKET_STRIDE, w) + w;
                last = ((t & (pw - 1)) == 0);
            }
            *s_last = last;
        }
        __syncthreads();
        if (!*s_last) return;
        if (tid == 0) {
            trace_stamp(A, f, 6);
            if (A.done_flag)  // the release store is itself the system-scope release
                __hip_atomic_store(A.done_flag, (unsigned long long)(f + 1), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    // reduce mix: the last source of each output channel sums that channel in
    // instance order (src/partconvMulti.cpp:906-921); the last channel to be
    // finished zeroes the channels without a source and raises done_flag
    if (TJ.mute) return;
    const int c = TJ.out_ch;
    const int cnt = A.src_off[c + 1] - A.src_off[c];
    wg_stores_done();  // every wave's y_inst stores, not only thread 0's
    if (tid == 0) {
        __threadfence();
        const unsigned t = atomicAdd(A.ticket_ch + c, 1u);
        *s_last = (t == (unsigned)(cnt - 1));
    }
    __syncthreads();
    if (!*s_last) return;
    __threadfence();
    float *oc = A.out + (long long)(f - A.fbase) * A.out_stride + (long long)c * A.out_cs;
    for (int j = tid; j < N; j += NT) {
        float sum = 0.f;
        for (int q = A.src_off[c]; q < A.src_off[c + 1]; q++) sum += A.y_inst[(size_t)A.src_list[q] * N + j];
        oc[(long long)j * A.out_ss] = sum;
    }
    wg_stores_done();  // every wave's output stores before the last arrival publishes them
    if (tid == 0) {
        A.ticket_ch[c] = 0u;
        __threadfence();
        const unsigned t = atomicAdd(A.ticket, 1u);
        *s_last = (t == (unsigned)(A.nactive - 1));
    }
    __syncthreads();
    if (!*s_last) return;
    for (int e = tid; e < A.nempty * N; e += NT) {
        const int ce = A.empty_ch[e / N], j = e % N;
        A.out[(long long)(f - A.fbase) * A.out_stride + (long long)ce * A.out_cs + (long long)j * A.out_ss] = 0.f;
    }
    wg_stores_done();
    if (tid == 0) {
        *A.ticket = 0u;
        trace_stamp(A, f, 6);
        if (A.done_flag)  // the release store is itself the system-scope release
            __hip_atomic_store(A.done_flag, (unsigned long long)(f + 1), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// buffer_size 64: level 0 of a tick as one wavefront with the 64-point complex
// transform held in registers (one value per lane; the radix-2 stages exchange
// values with DPP / permlane moves in the VALU, no LDS passes or barriers).
// The forward transform is decimation in frequency (natural input, bit-reversed
// output: lane l holds bin k = brev6(l)); the inverse is decimation in time
// (bit-reversed input, natural output).  Every lane keeps its stage twiddles,
// its split twiddle and its bin's H_0 / S in registers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int brev6(int x) { return (int)(__builtin_bitreverse32((unsigned)x) >> 26); }
__device__ __forceinline__ cf shfl_cf(cf v, int src) { return make_float2(__shfl(v.x, src), __shfl(v.y, src)); }

// Lane exchange v <- v[lane ^ D] in the VALU, without the LDS crossbar:
// D = 1, 2: DPP quad permutations; D = 4: quad xor 3 then row half-mirror
// (xor 7); D = 8: row half-mirror then row mirror (xor 15); D = 16, 32: the
// gfx950 row / half-wave swaps (v_permlane16_swap, v_permlane32_swap).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

template <int D>
__device__ __forceinline__ float xor_lane(float v, int lane)
{
    if constexpr (D == 1) {
        return dpp_mov<0xB1>(v);                 // quad_perm [1,0,3,2]
    } else if constexpr (D == 2) {
        return dpp_mov<0x4E>(v);                 // quad_perm [2,3,0,1]
    } else if constexpr (D == 4) {
        return dpp_mov<0x141>(dpp_mov<0x1B>(v)); // quad_perm [3,2,1,0], row_half_mirror
    } else if constexpr (D == 8) {
        return dpp_mov<0x140>(dpp_mov<0x141>(v)); // row_half_mirror, row_mirror
    } else if constexpr (D == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, v),
                                                        __builtin_bit_cast(unsigned, v), false, false);
        return __builtin_bit_cast(float, (lane & 16) ? p[0] : p[1]);
    } else {
        static_assert(D == 32, "lane exchange distance");
        const auto p = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, v),
                                                        __builtin_bit_cast(unsigned, v), false, false);
        return __builtin_bit_cast(float, (lane & 32) ? p[0] : p[1]);
    }
}

template <int D>
__device__ __forceinline__ cf xor_cf(cf v, int lane)
{
    return make_float2(xor_lane<D>(v.x, lane), xor_lane<D>(v.y, lane));
}

struct W64 {
    cf tw[6];       // stage d = 32,16,..,1: W_{2d}^{lane & (d-1)} on lanes with bit d set, 1 elsewhere
    float sg[6];    // stage d: -1 on lanes with bit d set, +1 elsewhere
    cf ws;          // split twiddle exp(-i pi k / 64) of this lane's bin k
    int k;          // bin held by this lane after the forward transform
    int partner;    // lane holding bin (64 - k) mod 64
};

__device__ __forceinline__ void w64_setup(W64 &w, const cf *__restrict__ tw, int lane)
{
    int s = 0;
    for (int d = 32; d >= 1; d >>= 1, s++) {
        const bool hi = (lane & d) != 0;
        w.tw[s] = hi ? tw[2 * d + (lane & (d - 1))] : make_float2(1.f, 0.f);
        w.sg[s] = hi ? -1.f : 1.f;
    }
    w.k = brev6(lane);
    w.partner = brev6((64 - w.k) & 63);
    w.ws = tw[128 + w.k];
}

// One radix-2 stage at distance D (stage index S), without branches: the
// lower lane of a pair computes v + o (twiddle 1), the upper (o - v) W.
template <int D, int S>
__device__ __forceinline__ cf w64_fwd_stage(cf v, const W64 &w, int lane)
{
    const cf o = xor_cf<D>(v, lane);
    const cf u = make_float2(o.x + w.sg[S] * v.x, o.y + w.sg[S] * v.y);
    return cmul(u, w.tw[S]);
}

// Inverse stage: t = v conj(W) on the upper lane (v on the lower), then
// t + o / o - t.
template <int D, int S>
__device__ __forceinline__ cf w64_inv_stage(cf v, const W64 &w, int lane)
{
    const cf t = cmulc(v, w.tw[S]);
    const cf o = xor_cf<D>(t, lane);
    return make_float2(o.x + w.sg[S] * t.x, o.y + w.sg[S] * t.y);
}

__device__ __forceinline__ cf w64_fwd(cf v, const W64 &w, int lane)
{
    v = w64_fwd_stage<32, 0>(v, w, lane);
    v = w64_fwd_stage<16, 1>(v, w, lane);
    v = w64_fwd_stage<8, 2>(v, w, lane);
    v = w64_fwd_stage<4, 3>(v, w, lane);
    v = w64_fwd_stage<2, 4>(v, w, lane);
    return w64_fwd_stage<1, 5>(v, w, lane);
}

__device__ __forceinline__ cf w64_inv(cf v, const W64 &w, int lane)
{
    v = w64_inv_stage<1, 5>(v, w, lane);
    v = w64_inv_stage<2, 4>(v, w, lane);
    v = w64_inv_stage<4, 3>(v, w, lane);
    v = w64_inv_stage<8, 2>(v, w, lane);
    v = w64_inv_stage<16, 1>(v, w, lane);
    return w64_inv_stage<32, 0>(v, w, lane);
}

// Level 0's partitions 1 .. T64_R of the 64-point tick live in registers: a
// lane's delay-line sum only reads its own bin's filter spectra and its own
// bin's spectra of the last ticks (which it wrote itself), so they stay in the
// lane between ticks; partitions past T64_R are read from memory.
#define T64_R 8
struct TickState64 {
    W64 w;
    cf h0, sp;       // H_0 and S of bin w.k
    float2 pv;       // previous frame samples 2*lane, 2*lane+1 (lanes < 32)
    cf hr[T64_R];    // H_1 .. H_T64_R of bin w.k (those with p < P)
    cf xq[T64_R - 1];  // X_{f-1}, X_{f-2}, ... of bin w.k (delay-line rows of the last ticks)
};

__device__ __forceinline__ void tick64_load_state(const LevelJob &J, uint64_t f, int lane, const cf *__restrict__ tw,
                                                  TickState64 &st)
{
    w64_setup(st.w, tw, lane);
    const int k = st.w.k, P = J.P;
    st.h0 = J.H[k];
    st.sp = J.S[k];
    const float *prev = J.ring + (((long long)f * 64 - 64) & (long long)J.ring_mask);
    st.pv = (lane < 32) ? make_float2(prev[2 * lane], prev[2 * lane + 1]) : make_float2(0.f, 0.f);
    const int slot = (int)(f % (uint64_t)P);
#pragma unroll
    for (int j = 0; j < T64_R; j++) st.hr[j] = (j + 1 < P) ? J.H[(size_t)(j + 1) * 64 + k] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < T64_R - 1; j++) {
        int r = slot - 1 - j;  // row of tick f - 1 - j
        if (r < 0) r += P;
        st.xq[j] = (j + 2 < P) ? J.fdl[(size_t)r * 64 + k] : make_float2(0.f, 0.f);
    }
}

// Level-block samples of the first four levels for one tick, read ahead of
// the tick (bit c of mask: level c's pair is in v[c]).
struct Lev64Pre {
    float2 v[4];
    unsigned mask;
};

// Reads ahead the level-block samples tick f will add, for the levels whose
// block tick f does not read first: that block was complete, and acquired,
// when an earlier tick of this workgroup first read it, and the half of the
// level's double buffer it sits in is not written again before every tick
// that reads it has run.  So the values are those tick f would load itself.
// A persistent tick workgroup issues these loads before it waits for tick f's
// release, where they overlap the host's round trip.
__device__ __forceinline__ Lev64Pre tick64_prefetch(const TickJob &TJ, const TickArgs &A, uint64_t f, int lane)
{
    Lev64Pre p;
    p.mask = 0;
    const uint64_t fr = A.relaxed ? f : f + 1;
    const int lo = lane & 31;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        p.v[c] = make_float2(0.f, 0.f);
        if (c >= TJ.nlev) continue;
        const int ks = TJ.kshift[c];
        const uint64_t qb = fr >> ks, sub = fr & ((1ull << ks) - 1);
        if (qb == 0 || sub == 0) continue;
        if (lane < 32)
            p.v[c] = gload2(reinterpret_cast<const float2 *>(TJ.ybuf[c] + (size_t)((qb - 1) & 1) * TJ.NL[c] +
                                                             sub * 64) + lo);
        p.mask |= 1u << c;
    }
    return p;
}

// Phase A for buffer_size 64; leaves the instance's output sample `lane` in
// *acc (through LDS) and X_f of bin st.w.k in *xr.  Levels in pre->mask take
// their samples from *pre (tick64_prefetch) instead of loading them.
__device__ __forceinline__ void tick64_phaseA(const TickJob &TJ, const TickArgs &A, uint64_t f, float *obuf, int lane,
                                              TickState64 &st, float *acc, cf *xr, const float *llx = nullptr,
                                              const Lev64Pre *pre = nullptr)
{
    const LevelJob &J = TJ.l0;
    const long long mask = J.ring_mask;
    const long long R = mask + 1;
    const long long t0 = (long long)f * 64;
    const float *inf = A.in + (long long)(f - A.fbase) * A.in_stride;
    const uint64_t fr = A.relaxed ? f : f + 1;
    const int lo = lane & 31;
    // loads: lanes 32..63 fetch the new frame's pairs, lanes 0..31 the level blocks due
    float2 xv = make_float2(0.f, 0.f), lev = make_float2(0.f, 0.f), y0p = make_float2(0.f, 0.f);
    if (lane >= 32) {
        if (llx) {
            xv = make_float2(llx[0], llx[1]);
        } else {
            const float *ic = inf + (long long)TJ.in_ch * A.in_cs;
            xv.x = ic[(long long)(2 * lo) * A.in_ss];
            xv.y = ic[(long long)(2 * lo + 1) * A.in_ss];
        }
    } else {
        if (A.relaxed && f > 0) y0p = gload2(reinterpret_cast<const float2 *>(J.ybuf + (size_t)((f - 1) & 1) * 64) + lo);
        // level blocks four at a time: loads in flight together, summed in level order
        for (int L0 = 0; L0 < TJ.nlev; L0 += 4) {
            float2 v[4];
            const unsigned pm = (pre && L0 == 0) ? pre->mask : 0u;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int L = L0 + c;
                const uint64_t qb = (L < TJ.nlev) ? (fr >> TJ.kshift[L]) : 0;
                v[c] = make_float2(0.f, 0.f);
                if ((pm >> c) & 1u) {
                    v[c] = pre->v[c];
                } else if (qb != 0) {
                    const uint64_t sub = fr & ((1ull << TJ.kshift[L]) - 1);
                    v[c] = gload2(reinterpret_cast<const float2 *>(TJ.ybuf[L] + (size_t)((qb - 1) & 1) * TJ.NL[L] +
                                                                   sub * 64) + lo);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                lev.x += v[c].x;
                lev.y += v[c].y;
            }
        }
    }
    // scale (src/partconvMulti.cpp:839-847); window pair of this lane
    cf z;
    if (lane >= 32) z = make_float2(xv.x * TJ.scale, xv.y * TJ.scale);
    else z = make_float2(st.pv.x, st.pv.y);
    // the new frame is the next tick's previous frame
    {
        const cf nz = xor_cf<32>(z, lane);
        if (lane < 32) st.pv = make_float2(nz.x, nz.y);
    }
    cf Z = w64_fwd(z, st.w, lane);
    const cf Zp = shfl_cf(Z, st.w.partner);
    cf X;
    if (st.w.k == 0) {
        X = make_float2(Z.x + Z.y, Z.x - Z.y);
    } else {
        cf xnk;
        rsplit_fwd(Z, Zp, st.w.ws, &X, &xnk);
    }
    *xr = X;
    cf Y = pmul(X, st.h0, st.w.k == 0);
    Y.x += st.sp.x;
    Y.y += st.sp.y;
    const cf Yp = shfl_cf(Y, st.w.partner);
    cf C;
    if (st.w.k == 0) {
        C = make_float2(Y.x + Y.y, Y.x - Y.y);
    } else {
        cf cnk;
        rsplit_inv(Y, Yp, st.w.ws, &C, &cnk);
    }
    const cf zo = w64_inv(C, st.w, lane);   // lane n: z'[n] = (y[2n], y[2n+1])
    if (lane >= 32) {
        // ring write of the scaled frame, after the transforms
        float *ring = const_cast<float *>(J.ring);
        const long long p = (t0 + 2 * lo) & mask;
        gstore2(reinterpret_cast<float2 *>(ring + p), make_float2(z.x, z.y));
        gstore2(reinterpret_cast<float2 *>(ring + p + R), make_float2(z.x, z.y));
    }
    if (lane < 32) {
        float2 o;
        if (A.relaxed) {
            gstore2(reinterpret_cast<float2 *>(J.ybuf + (size_t)(f & 1) * 64) + lo, make_float2(zo.x, zo.y));
            o = make_float2(y0p.x + lev.x, y0p.y + lev.y);
        } else {
            o = make_float2(zo.x + lev.x, zo.y + lev.y);
        }
        reinterpret_cast<float2 *>(obuf)[lo] = o;
    }
    __syncthreads();
    *acc = obuf[lane];
}

// Phase B for buffer_size 64: FDL row f % P <- X_f; S for tick f+1 (bin st.w.k).
__device__ __forceinline__ void tick64_phaseB(const TickJob &TJ, uint64_t f, cf xr, TickState64 &st)
{
    const LevelJob &J = TJ.l0;
    const int P = J.P;
    const int k = st.w.k;
    const int slot = (int)(f % (uint64_t)P);
    J.fdl[(size_t)slot * 64 + k] = xr;
    // partitions 1 .. T64_R from registers, then the rest from memory in batches
    // of 8 (loads in flight together), every bin summed in partition order
    cf s = (P > 1) ? pmul(xr, st.hr[0], k == 0) : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < T64_R - 1; j++)
        if (j + 2 < P) s = pmac(s, st.xq[j], st.hr[j + 1], k == 0);
    for (int p0 = T64_R + 1; p0 < P; p0 += 8) {
        cf av[8], hv[8];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int p = p0 + b;
            int r = slot + 1 - p;               // (f + 1 - p) mod P
            if (r < 0) r += P;
            if (p < P) {
                av[b] = gload2(J.fdl + (size_t)r * 64 + k);
                hv[b] = gload2(J.H + (size_t)p * 64 + k);
            }
        }
#pragma unroll
        for (int b = 0; b < 8; b++)
            if (p0 + b < P) s = pmac(s, av[b], hv[b], k == 0);
    }
    J.S[k] = s;
    st.sp = s;
#pragma unroll
    for (int j = T64_R - 2; j > 0; j--) st.xq[j] = st.xq[j - 1];
    st.xq[0] = xr;
}

// ---------------------------------------------------------------------------
// buffer_size 128: level 0 of a tick in one wavefront with the 128-point complex
// transform in registers, two values per lane (window elements l and l + 64):
// one decimation-in-frequency radix-2 stage inside the lane, then the 64-point
// lane-exchange transform on each half, so lane l holds bins 2k and 2k + 1 with
// k = brev6(l).  The real split pairs bin 2k with 128 - 2k (the 64-point
// partner lane) and 2k + 1 with 128 - 2k - 1 (a second partner lane).  The
// inverse runs the 64-point inverse on each half and a final in-lane stage that
// forms only the kept half of the window (samples 2l, 2l + 1 of the block).
// H_0, S and the delay line stay in natural bin order, as the other tick
// kernels and the level kernels lay them out.  Waves 1.. of the workgroup wait
// at the barrier that hands the output to the publishing threads.
// ---------------------------------------------------------------------------
struct TickState128 {
    W64 w;           // 64-point stage twiddles; w.k = k, w.partner: lane of bin 64 - k
    cf w1;           // W_128^lane: the in-lane stage
    cf wsa, wsb;     // split twiddles exp(-i pi b / 128) of bins b = 2k, 2k + 1
    int partner_b;   // lane holding bin 128 - (2k + 1)
    cf h0a, h0b, spa, spb;
    float2 pv;       // previous frame samples 2 lane, 2 lane + 1
};

// h1l (LDS, 128 float2, or null): this lane's bins of H_1 are copied there, so
// that phase B reads no global memory when P = 2 (the lane that stores them is
// the lane that reads them).
__device__ __forceinline__ void tick128_load_state(const LevelJob &J, uint64_t f, int lane, const cf *__restrict__ tw,
                                                   TickState128 &st, cf *h1l = nullptr)
{
    if (lane >= 64) return;
    w64_setup(st.w, tw, lane);
    const int ka = 2 * st.w.k, kb = ka + 1;
    st.w1 = tw[128 + lane];
    st.wsa = tw[256 + ka];
    st.wsb = tw[256 + kb];
    st.partner_b = brev6((63 - st.w.k) & 63);
    st.h0a = J.H[ka];
    st.h0b = J.H[kb];
    st.spa = J.S[ka];
    st.spb = J.S[kb];
    if (h1l && J.P > 1) {
        h1l[ka] = J.H[128 + ka];
        h1l[kb] = J.H[128 + kb];
    }
    const float *prev = J.ring + (((long long)f * 128 - 128) & (long long)J.ring_mask);
    st.pv = make_float2(prev[2 * lane], prev[2 * lane + 1]);
}

// Phase A for buffer_size 128 (wave 0 computes; every thread of the workgroup
// calls it): leaves output sample tid in acc[0] for tid < 128 and X_f of bins
// 2k, 2k + 1 in xr[0], xr[1] (wave 0).
__device__ __forceinline__ void tick128_phaseA(const TickJob &TJ, const TickArgs &A, uint64_t f, float *obuf,
                                               int tid, TickState128 &st, float *acc, cf *xr,
                                               const float *llx = nullptr)
{
    if (tid < 64) {
        const int lane = tid;
        const LevelJob &J = TJ.l0;
        const long long mask = J.ring_mask;
        const long long R = mask + 1;
        const long long t0 = (long long)f * 128;
        const float *inf = A.in + (long long)(f - A.fbase) * A.in_stride;
        const uint64_t fr = A.relaxed ? f : f + 1;
        // loads: the new frame's samples 2 lane, 2 lane + 1 and the level blocks due
        float2 xv, lev = make_float2(0.f, 0.f), y0p = make_float2(0.f, 0.f);
        if (llx) {
            xv = make_float2(llx[0], llx[1]);
        } else {
            const float *ic = inf + (long long)TJ.in_ch * A.in_cs;
            xv.x = ic[(long long)(2 * lane) * A.in_ss];
            xv.y = ic[(long long)(2 * lane + 1) * A.in_ss];
        }
        if (A.relaxed && f > 0)
            y0p = gload2(reinterpret_cast<const float2 *>(J.ybuf + (size_t)((f - 1) & 1) * 128) + lane);
        for (int L0 = 0; L0 < TJ.nlev; L0 += 4) {
            float2 v[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int L = L0 + c;
                const uint64_t qb = (L < TJ.nlev) ? (fr >> TJ.kshift[L]) : 0;
                v[c] = make_float2(0.f, 0.f);
                if (qb != 0) {
                    const uint64_t sub = fr & ((1ull << TJ.kshift[L]) - 1);
                    v[c] = gload2(reinterpret_cast<const float2 *>(TJ.ybuf[L] + (size_t)((qb - 1) & 1) * TJ.NL[L] +
                                                                   sub * 128) + lane);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                lev.x += v[c].x;
                lev.y += v[c].y;
            }
        }
        // window [previous frame | frame f * scale] (src/partconvMulti.cpp:839-847)
        const cf z0 = make_float2(st.pv.x, st.pv.y);
        const cf z1 = make_float2(xv.x * TJ.scale, xv.y * TJ.scale);
        st.pv = make_float2(z1.x, z1.y);
        // in-lane DIF stage, then the two 64-point halves
        const cf a = w64_fwd(cadd(z0, z1), st.w, lane);
        const cf b = w64_fwd(cmul(csub(z0, z1), st.w1), st.w, lane);
        const int k = st.w.k;
        const cf ap = shfl_cf(a, st.w.partner);
        const cf bp = shfl_cf(b, st.partner_b);
        cf Xa, Xb, dummy;
        if (k == 0) Xa = make_float2(a.x + a.y, a.x - a.y);
        else rsplit_fwd(a, ap, st.wsa, &Xa, &dummy);
        rsplit_fwd(b, bp, st.wsb, &Xb, &dummy);
        xr[0] = Xa;
        xr[1] = Xb;
        cf Ya = pmul(Xa, st.h0a, k == 0);
        Ya.x += st.spa.x;
        Ya.y += st.spa.y;
        cf Yb = cmul(Xb, st.h0b);
        Yb.x += st.spb.x;
        Yb.y += st.spb.y;
        const cf yap = shfl_cf(Ya, st.w.partner);
        const cf ybp = shfl_cf(Yb, st.partner_b);
        cf Ca, Cb;
        if (k == 0) Ca = make_float2(Ya.x + Ya.y, Ya.x - Ya.y);
        else rsplit_inv(Ya, yap, st.wsa, &Ca, &dummy);
        rsplit_inv(Yb, ybp, st.wsb, &Cb, &dummy);
        const cf ai = w64_inv(Ca, st.w, lane);
        const cf bi = w64_inv(Cb, st.w, lane);
        const cf zo = cadd(ai, cmulc(bi, st.w1));   // z'[lane] = (y[2 lane], y[2 lane + 1])
        // ring write of the scaled frame, after the transforms
        float *ring = const_cast<float *>(J.ring);
        const long long p = (t0 + 2 * lane) & mask;
        gstore2(reinterpret_cast<float2 *>(ring + p), make_float2(z1.x, z1.y));
        gstore2(reinterpret_cast<float2 *>(ring + p + R), make_float2(z1.x, z1.y));
        float2 o;
        if (A.relaxed) {
            gstore2(reinterpret_cast<float2 *>(J.ybuf + (size_t)(f & 1) * 128) + lane, make_float2(zo.x, zo.y));
            o = make_float2(y0p.x + lev.x, y0p.y + lev.y);
        } else {
            o = make_float2(zo.x + lev.x, zo.y + lev.y);
        }
        reinterpret_cast<float2 *>(obuf)[lane] = o;
    }
    __syncthreads();
    if (tid < 128) *acc = obuf[tid];
}

// Phase B for buffer_size 128 (wave 0): FDL row f % P <- X_f; S for tick f+1
// (bins 2k, 2k + 1).
__device__ __forceinline__ void tick128_phaseB(const TickJob &TJ, uint64_t f, int tid, const cf *xr,
                                               TickState128 &st, const cf *h1l = nullptr)
{
    if (tid >= 64) return;
    const LevelJob &J = TJ.l0;
    const int P = J.P;
    const int ka = 2 * st.w.k, kb = ka + 1;
    const int slot = (int)(f % (uint64_t)P);
    J.fdl[(size_t)slot * 128 + ka] = xr[0];
    J.fdl[(size_t)slot * 128 + kb] = xr[1];
    const cf *h1 = h1l ? h1l : J.H + 128;
    const cf h1a = (P > 1) ? h1[ka] : make_float2(0.f, 0.f);
    const cf h1b = (P > 1) ? h1[kb] : make_float2(0.f, 0.f);
    cf sa = (P > 1) ? pmul(xr[0], h1a, ka == 0) : make_float2(0.f, 0.f);
    cf sb = (P > 1) ? cmul(xr[1], h1b) : make_float2(0.f, 0.f);
    // partitions 2..P-1 in batches of 4: loads in flight together, summed in order
    for (int p0 = 2; p0 < P; p0 += 4) {
        cf av[4], bv[4], ha[4], hb[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int pp = p0 + c;
            int r = slot + 1 - pp;  // (f + 1 - p) mod P
            if (r < 0) r += P;
            if (pp < P) {
                av[c] = gload2(J.fdl + (size_t)r * 128 + ka);
                bv[c] = gload2(J.fdl + (size_t)r * 128 + kb);
                ha[c] = gload2(J.H + (size_t)pp * 128 + ka);
                hb[c] = gload2(J.H + (size_t)pp * 128 + kb);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (p0 + c < P) {
                sa = pmac(sa, av[c], ha[c], ka == 0);
                sb = pmac(sb, bv[c], hb[c], false);
            }
    }
    J.S[ka] = sa;
    J.S[kb] = sb;
    st.spa = sa;
    st.spb = sb;
}

template <int LOGN>
__global__ void __launch_bounds__(TickCfg<LOGN>::NT)
k_tick(TickArgs A, const cf *__restrict__ tw)
{
    constexpr int N = TickCfg<LOGN>::N, IPT = TickCfg<LOGN>::IPT;
    __shared__ cf bufA[N];
    __shared__ cf bufB[N];
    __shared__ unsigned s_last;
    const int tid = threadIdx.x;
    const TickJob &TJ = A.jobs[blockIdx.x];
    float acc[IPT];
    if constexpr (LOGN == 6) {
        TickState64 st;
        tick64_load_state(TJ.l0, A.f, tid, tw, st);
        cf xr;
        tick64_phaseA(TJ, A, A.f, reinterpret_cast<float *>(bufA), tid, st, &acc[0], &xr);
        tick_arrive<LOGN>(TJ, A, A.f, tid, acc, &s_last);
        tick64_phaseB(TJ, A.f, xr, st);
    } else if constexpr (LOGN == 7) {
        TickState128 st;
        tick128_load_state(TJ.l0, A.f, tid, tw, st);
        cf xr[2];
        tick128_phaseA(TJ, A, A.f, reinterpret_cast<float *>(bufA), tid, st, &acc[0], xr);
        tick_arrive<LOGN>(TJ, A, A.f, tid, acc, &s_last);
        tick128_phaseB(TJ, A.f, tid, xr, st);
    } else {
        TickState<LOGN> st;
        tick_load_state<LOGN>(TJ.l0, A.f, tid, st);
        cf xr[IPT];
        tick_phaseA<LOGN>(TJ, A, A.f, bufA, bufB, tw, tid, st, acc, xr);
        tick_arrive<LOGN>(TJ, A, A.f, tid, acc, &s_last);
        tick_phaseB<LOGN>(TJ, A.f, tid, xr, st);
    }
}

// ---------------------------------------------------------------------------
// Persistent tick kernel (device-resident runs): one launch serves nticks
// consecutive ticks.  The host releases tick f by writing go = f + 1 into
// pinned memory (after it has checked that the level blocks tick f reads are
// complete); each workgroup waits for the release, runs the same phases as
// k_tick with its per-instance state kept in registers and its job descriptor
// in LDS, and publishes its output.  Every wait is bounded: a release that
// does not come within `spin_limit` s_memrealtime ticks (100 MHz) ends the
// kernel, and go = PC_GO_STOP | r ends it at the first tick >= r (r = 0: at
// once; r > 0: the ticks already released, f < r, are served first).
// ---------------------------------------------------------------------------
#define PC_GO_STOP (1ull << 62)
// Tick f may run under release word g (stop bit or not).
__device__ __forceinline__ bool go_released(unsigned long long g, uint64_t f)
{
    return (g & ~PC_GO_STOP) >= f + 1;
}

// Level workers: workgroups of the persistent kernel that run the blocks of
// the short-period levels (k = N_L / N0 <= 8, N_L <= 2048), so those blocks
// need no host launch between the tick that completes their input and the
// tick that reads their output.  Each worker owns a run of jobs (instances) of
// one level.  Handshake (device memory, agent scope; every wait bounded):
//   tick workgroup i, at ticks f = 0 mod k_min:  ring stores -> release ->
//       dev_tick[i] = f + 1
//   worker, at its level's start ticks f = m k, for each of its jobs: wait
//       dev_tick[i] >= f + 1, acquire, block m (window -> r2c -> FDL slot ->
//       X H_0 + S -> c2r -> double-buffer half m & 1), release,
//       lvl_done[job] = 2m + 1; then, for each job, the delay-line sum S for
//       block m + 1 and lvl_done[job] = 2m + 2
//   tick workgroup i, at the first tick that reads block mb of the level:
//       wait lvl_done[job] >= 2 mb + 1, acquire.
// A tick workgroup that stops (PC_GO_STOP or a timeout) writes the stop bit
// into dev_tick[i] so its workers end too.
struct WorkJob {
    LevelJob J;
    int inst;      // tick workgroup (instance) whose frames feed the level
    int kshift;    // log2(N_L / N0)
    int logN;
    int slot;      // index into lvl_done
};

struct PersistArgs {
    TickArgs A;
    uint64_t f0;
    long long nticks;
    const unsigned long long *go;  // host-written release counter (pinned)
    unsigned long long *dev_go;    // relay: workgroup 0 copies go here for the others (or null)
    unsigned long long spin_limit;
    unsigned long long acq_mask;   // device-scope acquire at ticks whose (f [+1]) & mask == 0
    unsigned long long pub_mask;   // workers: frames handed over at ticks f & mask == 0
    const WorkJob *wjobs;          // worker-level jobs (lvl_done slot = index)
    const int2 *wrange;            // per worker workgroup: first job, job count (one level)
    int nworkers;                  // workgroups njobs .. njobs + nworkers - 1
    int self_levels;               // no workers: each tick workgroup runs its own worker-level blocks
    unsigned long long *dev_tick;  // per tick workgroup: frames published to the workers
    unsigned long long *lvl_done;  // per worker slot: 2m + 1 (block m out), 2m + 2 (S of m + 1 too)
    unsigned *present;             // plain (non-cooperative) launch: workgroup w stores present_tag
    unsigned present_tag;          // into present[w] (host-visible) on entry; null otherwise
    int tw_lds;                    // twiddle-table entries tw[0 .. tw_lds) copied into LDS at entry (0: none)
    int tw_lds_off;                // their offset (float2) in the dynamic LDS, after the block buffers
    int prefetch;                  // 64-point ticks: launch k_persist<6, true> (reads ahead of each release)
};

// Residency stamp of a persistent kernel launched without the cooperative
// launch's guarantee: the host releases the first tick only once every
// workgroup has stamped (each one polls the release from then on).
__device__ __forceinline__ void stamp_present(unsigned *present, unsigned tag, int tid)
{
    if (present && tid == 0)
        __hip_atomic_store(present + blockIdx.x, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}


// S[k] = sum_{p=1}^{P-1} FDL[(b - p) % P][k] H_p[k] for one job, by one workgroup
// (two bins per thread and step, as k_partial).
template <int NT>
__device__ __forceinline__ void wg_partial(const LevelJob &J, int N, uint64_t b, int tid)
{
    const int P = J.P;
    const float4 *fdl = reinterpret_cast<const float4 *>(J.fdl);
    const float4 *H = reinterpret_cast<const float4 *>(J.H);
    const size_t row = (size_t)(N / 2);
    const int s0 = (int)(b % (uint64_t)P);
    for (int k2 = tid; k2 < N / 2; k2 += NT) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        int s = s0;
#pragma unroll 4
        for (int p = 1; p < P; p++) {
            s = (s == 0) ? P - 1 : s - 1;
            const float4 a = fdl[(size_t)s * row + k2];
            const float4 h = H[(size_t)p * row + k2];
            acc = dl_mac2(acc, a, h, k2 == 0);
        }
        reinterpret_cast<float4 *>(J.S)[k2] = acc;
    }
}

// Bounded spin of one thread until *p >= target.  Returns 1 when reached, 0 on
// the stop bit or the time limit.
__device__ __forceinline__ int spin_geq(const unsigned long long *p, unsigned long long target,
                                        unsigned long long limit, int sleep_rounds)
{
    const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + limit;
    for (;;) {
        const unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v & PC_GO_STOP) return 0;
        if (v >= target) return 1;
        if (__builtin_amdgcn_s_memrealtime() > t_end) return 0;
        for (int r = 0; r < sleep_rounds; r++) __builtin_amdgcn_s_sleep(2);
    }
}

// Block m of one worker-level job (window -> r2c -> FDL slot -> X H_0 + S ->
// c2r -> double-buffer half m & 1), all in LDS, by one workgroup.
//
// With P <= 2 the delay-line sum of block m + 1 is X_m H_1 (P = 2) or zero
// (P = 1): it is formed here from X_m in registers, with the float4 kernels'
// rounding (dl_mac1), instead of reading X_m back from the delay line after
// the block -- where H_1's registers fit beside the persistent kernel's tick
// state (at most 2 bins per thread).  Returns true when S for block m + 1 was
// written.
template <int LOGNL, int NT>
__device__ __forceinline__ constexpr bool block_forms_partial(int P)
{
    return ((1 << LOGNL) + NT - 1) / NT <= 2 && P <= 2;
}

template <int LOGNL, int NT>
__device__ __forceinline__ bool level_block(const LevelJob &J, uint64_t m, int N0, cf *bufA, cf *bufB, const cf *__restrict__ tw,
                            int tid)
{
    constexpr int N = 1 << LOGNL, IPT = (N + NT - 1) / NT;
    cf h[IPT], sv[IPT], h1[IPT];
    const int P = J.P;
    const bool fuse = block_forms_partial<LOGNL, NT>(P);
    small_prefetch<LOGNL, NT>(J, tid, h, sv);
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const int k = tid + q * NT;
        h1[q] = (IPT <= 2 && P == 2 && k < N) ? J.H[N + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();  // the LDS buffers may still be read by the previous user
    small_load_window<LOGNL, NT>(bufA, J, PC_MODE_RUN, (long long)m * N - N0 - J.T, tid);
    cf *X = small_forward<LOGNL, NT>(bufA, bufB, tw, tid);
    cf *other = (X == bufA) ? bufB : bufA;
    if (fuse) {
        const int slot = (int)(m % (uint64_t)P);
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const int k = tid + q * NT;
            if (k < N) {
                const cf x = X[k];
                J.fdl[(size_t)slot * N + k] = x;
                cf y = pmul(x, h[q], k == 0);
                y.x += sv[q].x;
                y.y += sv[q].y;
                X[k] = y;
                J.S[k] = dl_mac1(make_float2(0.f, 0.f), x, h1[q], k == 0);
            }
        }
        __syncthreads();
    } else {
        small_mac<LOGNL, NT>(X, J, m, tid, h, sv);
    }
    cf *z = small_inverse<LOGNL, NT>(X, other, tw, tid);
    cf *yo = reinterpret_cast<cf *>(J.ybuf + (size_t)(m & 1) * N);
    for (int n = tid; n < N / 2; n += NT) yo[n] = z[n];
    return fuse;
}

// One level worker: every block of its jobs whose start tick lies in the run.
template <int LOGNL, int NT>
__device__ __forceinline__ void level_worker(const PersistArgs &P, int2 range, cf *bufA, cf *bufB, const cf *__restrict__ tw,
                             int tid, int *s_ok)
{
    constexpr int N = 1 << LOGNL;
    const int ks = P.wjobs[range.x].kshift;
    const uint64_t k = 1ull << ks;
    const uint64_t fend = P.f0 + (uint64_t)P.nticks;
    for (uint64_t f = (P.f0 + k - 1) & ~(k - 1); f < fend; f += k) {
        const uint64_t m = f >> ks;
        for (int q = range.x; q < range.x + range.y; q++) {
            const WorkJob &W = P.wjobs[q];
            if (tid == 0) {
                const int ok = spin_geq(P.dev_tick + W.inst, f + 1, P.spin_limit, 4);
                if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                *s_ok = ok;
            }
            __syncthreads();
            if (!*s_ok) return;
            level_block<LOGNL, NT>(W.J, m, P.A.N0, bufA, bufB, tw, tid);
            wg_stores_done();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(P.lvl_done + q, 2 * m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // delay-line sums for the next blocks, off the deadline path
        for (int q = range.x; q < range.x + range.y; q++) {
            if (!block_forms_partial<LOGNL, NT>(P.wjobs[q].J.P)) wg_partial<NT>(P.wjobs[q].J, N, m + 1, tid);
            wg_stores_done();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(P.lvl_done + q, 2 * m + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Self mode: a tick workgroup runs block m of its own worker-level job q after
// one of its ticks (the consumer is itself, so no handshake is needed; the
// lvl_done stores only report progress to the host after the kernel ends).
template <int LOGN>
__device__ __forceinline__ void self_level(const PersistArgs &P, int q, uint64_t m, cf *lds, const cf *__restrict__ tw, int tid)
{
    constexpr int NT = TickCfg<LOGN>::NT;
    const WorkJob &W = P.wjobs[q];
    const int d = W.logN - LOGN;
    bool formed = false;  // S for block m + 1 written by the block itself
    if constexpr (worker_fits<LOGN + 1, TickCfg<LOGN>::NT>()) {
        if (d == 1) {
            formed = level_block<LOGN + 1, NT>(W.J, m, P.A.N0, lds, lds + (1 << (LOGN + 1)), tw, tid);
        }
    }
    if constexpr (worker_fits<LOGN + 2, TickCfg<LOGN>::NT>()) {
        if (d == 2) {
            formed = level_block<LOGN + 2, NT>(W.J, m, P.A.N0, lds, lds + (1 << (LOGN + 2)), tw, tid);
        }
    }
    if constexpr (worker_fits<LOGN + 3, TickCfg<LOGN>::NT>()) {
        if (d == 3) {
            formed = level_block<LOGN + 3, NT>(W.J, m, P.A.N0, lds, lds + (1 << (LOGN + 3)), tw, tid);
        }
    }
    if (!formed) {
        if (tid == 0) __hip_atomic_store(P.lvl_done + q, 2 * m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        wg_partial<NT>(W.J, 1 << W.logN, m + 1, tid);
    }
    __syncthreads();
    if (tid == 0) __hip_atomic_store(P.lvl_done + q, 2 * m + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int LOGN>
__device__ __forceinline__ void level_worker_dispatch(const PersistArgs &P, int2 range, cf *bufA, cf *bufB,
                                      const cf *__restrict__ tw, int tid, int *s_ok)
{
    constexpr int NT = TickCfg<LOGN>::NT;
    const int d = P.wjobs[range.x].logN - LOGN;
    if constexpr (worker_fits<LOGN + 1, TickCfg<LOGN>::NT>()) {
        if (d == 1) return level_worker<LOGN + 1, NT>(P, range, bufA, bufB, tw, tid, s_ok);
    }
    if constexpr (worker_fits<LOGN + 2, TickCfg<LOGN>::NT>()) {
        if (d == 2) return level_worker<LOGN + 2, NT>(P, range, bufA, bufB, tw, tid, s_ok);
    }
    if constexpr (worker_fits<LOGN + 3, TickCfg<LOGN>::NT>()) {
        if (d == 3) return level_worker<LOGN + 3, NT>(P, range, bufA, bufB, tw, tid, s_ok);
    }
}

// Before a persistent run from tick f0: every worker-level block whose start
// tick precedes f0 is complete (the host launched it), so job q starts at
// lvl_done = 2 ceil(f0 / k_q).  Also clears the frame hand-over counters.
__global__ void k_worker_init(const WorkJob *__restrict__ wjobs, int nwjobs, uint64_t f0,
                              unsigned long long *lvl_done, unsigned long long *dev_tick, int ninst)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nwjobs) {
        const uint64_t k = 1ull << wjobs[q].kshift;
        lvl_done[q] = 2 * ((f0 + k - 1) / k);
    }
    if (q < ninst) dev_tick[q] = 0;
}

// After a persistent run that served ticks up to fend - 1: job q (one
// workgroup) completes every block with start tick < fend that the run did not
// finish (a stopped run, or blocks whose turn came after the last tick) -- its
// delay-line sum first if lvl_done is odd -- so that what follows (another
// run, or per-tick launches) finds the state of a fully served history.
template <int LOGNL, int NT>
__device__ __forceinline__ void worker_fix(const WorkJob &W, unsigned long long *done, uint64_t fend, int N0, cf *lds,
                           const cf *__restrict__ tw, int tid)
{
    constexpr int N = 1 << LOGNL;
    const uint64_t k = 1ull << W.kshift;
    const uint64_t m_end = (fend + k - 1) / k;
    const unsigned long long d = *done;
    uint64_t next = (d + 1) / 2;
    if (d & 1) {
        wg_partial<NT>(W.J, N, next, tid);
        __syncthreads();
    }
    for (uint64_t m = next; m < m_end; m++) {
        if (!level_block<LOGNL, NT>(W.J, m, N0, lds, lds + N, tw, tid)) {
            __syncthreads();
            wg_partial<NT>(W.J, N, m + 1, tid);
        }
        __syncthreads();
    }
    if (tid == 0 && (next < m_end || (d & 1))) *done = 2 * m_end;
}

template <int LOGN>
__global__ void __launch_bounds__(TickCfg<LOGN>::NT)
k_worker_fix(const WorkJob *__restrict__ wjobs, unsigned long long *lvl_done, uint64_t fend, int N0,
             const cf *__restrict__ tw)
{
    constexpr int NT = TickCfg<LOGN>::NT;
    extern __shared__ cf pc_lds[];
    const WorkJob &W = wjobs[blockIdx.x];
    unsigned long long *done = lvl_done + blockIdx.x;
    const int d = W.logN - LOGN;
    if constexpr (worker_fits<LOGN + 1, TickCfg<LOGN>::NT>()) {
        if (d == 1) worker_fix<LOGN + 1, NT>(W, done, fend, N0, pc_lds, tw, threadIdx.x);
    }
    if constexpr (worker_fits<LOGN + 2, TickCfg<LOGN>::NT>()) {
        if (d == 2) worker_fix<LOGN + 2, NT>(W, done, fend, N0, pc_lds, tw, threadIdx.x);
    }
    if constexpr (worker_fits<LOGN + 3, TickCfg<LOGN>::NT>()) {
        if (d == 3) worker_fix<LOGN + 3, NT>(W, done, fend, N0, pc_lds, tw, threadIdx.x);
    }
}

// Wait for the host's release counter, one system-scope load at a time.  The
// counter is in fine-grained device memory that the host writes through the
// BAR (a local load), or in pinned host memory (a load across the bus, about
// 1.2 us on MI355X; more loads in flight per workgroup made the whole round
// trip slower).  Returns the value that ended the wait (>= target, or with the
// stop bit), 0 on timeout.
__device__ __forceinline__ unsigned long long poll_host(const unsigned long long *p, unsigned long long target,
                                                        unsigned long long t_end)
{
    for (;;) {
        const unsigned long long g = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if ((g & PC_GO_STOP) || g >= target) return g;
        if (__builtin_amdgcn_s_memrealtime() > t_end) return 0;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Samples per thread that phase A takes from the tagged input words: two for
// lanes 32..63 of the 64-point form, else IPT.
template <int LOGN>
struct LLCnt {
    static constexpr int v = (LOGN == 6 || LOGN == 7) ? 2 : TickCfg<LOGN>::IPT;
};

// Host sessions with tagged words: wait until this workgroup's input words of
// tick f carry the tick's tag and return the thread's samples in x.  Every
// wave polls its own words (system-scope loads: device memory written through
// the BAR, or pinned host memory); lane 0 also reads the release word for the
// stop bit.  Returns 1 when the frame is in, 0 on stop or timeout (uniform over
// the workgroup).
template <int LOGN>
__device__ __forceinline__ int ll_wait_frame(const TickJob &TJ, const TickArgs &A, const unsigned long long *go,
                                             uint64_t f, int tid, unsigned long long limit,
                                             float (&x)[LLCnt<LOGN>::v])
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, C = LLCnt<LOGN>::v;
    const unsigned want = ll_tag(f);
    const unsigned long long *src = A.ll_in + (size_t)TJ.in_ch * N;
    const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + limit;
    int stop = 0;
    for (;;) {
        int mine = 1;
#pragma unroll
        for (int q = 0; q < C; q++) {
            int j;
            bool valid;
            if constexpr (LOGN == 6) {
                valid = tid >= 32;
                j = 2 * (tid & 31) + q;
            } else if constexpr (LOGN == 7) {
                valid = tid < 64;
                j = 2 * tid + q;
            } else {
                j = tid + q * NT;
                valid = j < N;
            }
            if (valid) {
                const unsigned long long w = __hip_atomic_load(src + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                x[q] = __uint_as_float((unsigned)w);
                mine &= ((unsigned)(w >> 32) == want) ? 1 : 0;
            }
        }
        if (__all(mine)) break;
        int st = 0;
        if ((tid & 63) == 0) {
            const unsigned long long g = __hip_atomic_load(go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            st = ((g & PC_GO_STOP) != 0) || __builtin_amdgcn_s_memrealtime() > t_end;
        }
        if (__any(st)) {
            stop = 1;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if constexpr (NT > 64) stop = __syncthreads_or(stop);
    return !stop;
}

// AHEAD (64-point ticks, one tick released at a time): before it waits for a
// tick's release, each tick workgroup reads that tick's level samples and the
// worker-level counter it will wait on (a separate instantiation, so the
// windowed runs keep the plain kernel's code and registers)
template <int LOGN, bool AHEAD = false>
__global__ void __launch_bounds__(TickCfg<LOGN>::NT, TickCfg<LOGN>::MINW)
k_persist(PersistArgs P, const cf *__restrict__ tw)
{
    constexpr int N = TickCfg<LOGN>::N, NT = TickCfg<LOGN>::NT, IPT = TickCfg<LOGN>::IPT;
    // 2 x max(N0, largest worker block) float2, sized at launch; then, with
    // level workers, the twiddle-table prefix the level blocks read
    extern __shared__ cf pc_lds[];
    __shared__ unsigned s_last;
    __shared__ int s_go;
    __shared__ TickJob sTJ;
    __shared__ cf s_h1[(LOGN == 7) ? 128 : 1];  // level 0's H_1 (buffer 128)
    const int tid = threadIdx.x;
    stamp_present(P.present, P.present_tag, tid);
    // The level blocks' transform passes each read their twiddles after a
    // barrier; from global memory every pass waited for an L2 (after a
    // device-scope acquire, a memory) round trip.  Copied once per launch into
    // LDS they are the same values, so results do not change.
    const cf *twl = tw;
    if (P.tw_lds > 0) {
        cf *t = pc_lds + P.tw_lds_off;
        for (int i = tid; i < P.tw_lds; i += NT) t[i] = tw[i];
        __syncthreads();
        twl = t;
    }
    if ((int)blockIdx.x >= P.A.njobs) {
        const int2 range = P.wrange[blockIdx.x - P.A.njobs];
        const int lw = P.wjobs[range.x].logN;
        level_worker_dispatch<LOGN>(P, range, pc_lds, pc_lds + (1 << lw), twl, tid, &s_go);
        return;
    }
    cf *bufA = pc_lds;
    cf *bufB = pc_lds + N;
    {
        const unsigned *src = reinterpret_cast<const unsigned *>(P.A.jobs + blockIdx.x);
        unsigned *dst = reinterpret_cast<unsigned *>(&sTJ);
        for (int w = tid; w < (int)(sizeof(TickJob) / 4); w += NT) dst[w] = src[w];
    }
    __syncthreads();
    TickState<LOGN> st;
    TickState64 st64;
    TickState128 st128;
    if constexpr (LOGN == 6) tick64_load_state(sTJ.l0, P.f0, tid, tw, st64);
    else if constexpr (LOGN == 7) tick128_load_state(sTJ.l0, P.f0, tid, tw, st128, s_h1);
    else tick_load_state<LOGN>(sTJ.l0, P.f0, tid, st);
    // release word last read by thread 0: when the host releases ticks ahead
    // (a release window), the ticks it already covers need no new read
    unsigned long long g_seen = 0;
    // level samples read ahead for the next tick (64-point ticks)
    Lev64Pre pre;
    pre.mask = 0;
    // thread 0, one tick at a time: the worker-level counter lvl_done[ld_slot]
    // as read before this tick's release, for the block this tick reads first
    // (-1: none read)
    unsigned long long ld_pre = 0;
    int ld_slot = -1;
    for (long long t = 0; t < P.nticks; t++) {
        const uint64_t f = P.f0 + (uint64_t)t;
        float llx[LLCnt<LOGN>::v];
        int ll_in = 0;
        if (P.A.ll) ll_in = ll_wait_frame<LOGN>(sTJ, P.A, P.go, f, tid, P.spin_limit, llx);
        if (tid == 0) {
            int go = 0;
            // with many workgroups only workgroup 0 polls host memory over the bus;
            // it relays the counter through device memory to the others
            const bool host_poll = (P.dev_go == nullptr) || blockIdx.x == 0;
            if (P.A.ll) {
                go = ll_in;  // the tagged input words were the release
            } else if (go_released(g_seen, f)) {
                go = 1;
            } else if (host_poll) {
                const unsigned long long g = poll_host(P.go, f + 1, __builtin_amdgcn_s_memrealtime() + P.spin_limit);
                if (P.dev_go && g) __hip_atomic_store(P.dev_go, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                go = go_released(g, f);
                g_seen = g;
            } else {
                const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + P.spin_limit;
                for (;;) {
                    const unsigned long long g = __hip_atomic_load(P.dev_go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (go_released(g, f)) {
                        go = 1;
                        g_seen = g;
                        break;
                    }
                    if (g & PC_GO_STOP) break;
                    if (__builtin_amdgcn_s_memrealtime() > t_end) break;
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            const uint64_t fr = P.A.relaxed ? f : f + 1;
            // blocks of worker levels first read at this tick
            bool waited = false;
            for (int L = 0; go && P.nworkers > 0 && L < sTJ.nlev; L++) {
                const int slot = sTJ.wslot[L];
                if (slot < 0) continue;
                const int ks = sTJ.kshift[L];
                if ((fr & ((1ull << ks) - 1)) != 0 || (fr >> ks) == 0) continue;
                const unsigned long long need = 2 * ((fr >> ks) - 1) + 1;
                waited = true;
                // the counter only grows: a value read before the release that
                // already covers the block ends the wait (the acquire below
                // orders this tick's loads after that read)
                if (AHEAD && slot == ld_slot && !(ld_pre & PC_GO_STOP) && ld_pre >= need) continue;
                go = spin_geq(P.lvl_done + slot, need, P.spin_limit, 0);
            }
            ld_slot = -1;
            // level blocks written by other workgroups or kernels (other XCDs) are
            // first read at deadline ticks: acquire at device scope there
            if (go && (waited || (fr & P.acq_mask) == 0)) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            // frames in pinned host memory, written before the release (tagged
            // input words need none: each word carries its tick)
            if (go && P.A.fence_sys && !P.A.ll) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            if (!go && P.dev_tick)
                __hip_atomic_store(P.dev_tick + blockIdx.x, PC_GO_STOP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;
        const int tslot = (blockIdx.x == 0) ? 0 : (blockIdx.x == P.A.njobs / 2) ? 2 : (blockIdx.x == P.A.njobs - 1) ? 4 : -1;
        const bool tr = P.A.trace && tslot >= 0 && tid == 0;
        if (tr) trace_stamp(P.A, f, tslot);
        float acc[IPT];
        if constexpr (LOGN == 6) {
            cf xr;
            tick64_phaseA(sTJ, P.A, f, reinterpret_cast<float *>(bufA), tid, st64, &acc[0], &xr,
                          P.A.ll ? llx : nullptr, AHEAD ? &pre : nullptr);
            pre.mask = 0;
            tick_arrive<LOGN>(sTJ, P.A, f, tid, acc, &s_last);
            if (tr) trace_stamp(P.A, f, tslot + 1);
            tick64_phaseB(sTJ, f, xr, st64);
        } else if constexpr (LOGN == 7) {
            cf xr[2];
            tick128_phaseA(sTJ, P.A, f, reinterpret_cast<float *>(bufA), tid, st128, &acc[0], xr,
                           P.A.ll ? llx : nullptr);
            tick_arrive<LOGN>(sTJ, P.A, f, tid, acc, &s_last);
            if (tr) trace_stamp(P.A, f, tslot + 1);
            tick128_phaseB(sTJ, f, tid, xr, st128, s_h1);
        } else {
            cf xr[IPT];
            tick_phaseA<LOGN>(sTJ, P.A, f, bufA, bufB, twl, tid, st, acc, xr, P.A.ll ? llx : nullptr);
            tick_arrive<LOGN>(sTJ, P.A, f, tid, acc, &s_last);
            if (tr) trace_stamp(P.A, f, tslot + 1);
            tick_phaseB<LOGN>(sTJ, f, tid, xr, st);
        }
        if (P.self_levels) {
            // block m of a level with period k starts at tick m k and runs at tick
            // m k + d_k, d_k = k/2 - 1 (0, 1, 3 for k = 2, 4, 8): at most one
            // block per tick, each done before its first reader (m k + k - 1)
            for (int L = 0; L < sTJ.nlev; L++) {
                const int q = sTJ.wslot[L];
                if (q < 0) continue;
                const int ks = sTJ.kshift[L];
                const uint64_t d = (1ull << (ks - 1)) - 1;
                if (f < P.f0 + d || ((f - d) & ((1ull << ks) - 1)) != 0) continue;
                self_level<LOGN>(P, q, (f - d) >> ks, pc_lds, twl, tid);
            }
        }
        if (P.nworkers > 0 && (f & P.pub_mask) == 0) {
            // the frame is in the ring: hand it to this instance's level workers
            wg_stores_done();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(P.dev_tick + blockIdx.x, f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (tr && tslot == 0) trace_stamp(P.A, f, 7);
        __syncthreads();
        if constexpr (LOGN == 6 && AHEAD) {
            // the next tick is not released yet as far as this workgroup knows:
            // read its level samples now, while the release makes its round trip
            // (thread 0 holds the release word; the workgroup is one wavefront)
            if (t + 1 < P.nticks && __shfl((int)!go_released(g_seen, f + 1), 0)) {
                pre = tick64_prefetch(sTJ, P.A, f + 1, tid);
                // so is the counter of a worker-level block the next tick reads
                // first (a device-scope read of a line another XCD writes)
                const uint64_t fr1 = (P.A.relaxed ? f : f + 1) + 1;
                for (int L = 0; tid == 0 && P.nworkers > 0 && L < sTJ.nlev; L++) {
                    const int slot = sTJ.wslot[L];
                    if (slot < 0) continue;
                    const int ks = sTJ.kshift[L];
                    if ((fr1 & ((1ull << ks) - 1)) != 0 || (fr1 >> ks) == 0) continue;
                    ld_pre = __hip_atomic_load(P.lvl_done + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ld_slot = slot;
                    break;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Delay-line partial sum for block b of a level (every job of one launch has
// the same N):  S[k] = sum_{p=1}^{P-1} FDL[(b - p) % P][k] * H_p[k],
// accumulated p = 1..P-1.  It runs on the level's stream right after block
// b-1 (whose spectrum is FDL row (b-1) % P), i.e. off the path from a block's
// input to its output: that path then only needs X_b H_0 + S.  This is the
// kernel that streams the filter spectra and the delay line: each thread owns
// two bins (one float4 of FDL and H per partition).
// ---------------------------------------------------------------------------
#define PART_NT 256

// Q float4 groups (2 Q bins) per thread, Q in {1, 2, 4}: every bin is summed
// over p = 1 .. P-1 in the same order whatever Q is, so the result does not
// depend on it; a larger Q keeps more loads in flight per thread on fewer
// workgroups (a tuning choice).
template <int Q>
__global__ void __launch_bounds__(PART_NT)
k_partial(const LevelJob *__restrict__ jobs, int N, int wg_per_job, uint64_t b)
{
    const int jb = blockIdx.x / wg_per_job, w = blockIdx.x - jb * wg_per_job;
    const LevelJob J = jobs[jb];
    const int P = J.P;
    const int base = w * PART_NT * Q + threadIdx.x;  // float4 index of group 0; group q at + q * PART_NT
    if (2 * base >= N) return;
    float4 acc[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) acc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    int s = (int)(b % (uint64_t)P);
    const float4 *fdl = reinterpret_cast<const float4 *>(J.fdl);
    const float4 *H = reinterpret_cast<const float4 *>(J.H);
    const size_t row = (size_t)(N / 2);
#pragma unroll 4
    for (int p = 1; p < P; p++) {
        s = (s == 0) ? P - 1 : s - 1;
        float4 a[Q], h[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int k2 = base + q * PART_NT;
            if (2 * k2 < N) {
                a[q] = fdl[(size_t)s * row + k2];
                h[q] = H[(size_t)p * row + k2];
            }
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int k2 = base + q * PART_NT;
            if (2 * k2 >= N) continue;
            acc[q] = dl_mac2(acc[q], a[q], h[q], k2 == 0);
        }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k2 = base + q * PART_NT;
        if (2 * k2 < N) reinterpret_cast<float4 *>(J.S)[k2] = acc[q];
    }
}

// Plain-mode mix: out[c cs + j ss] = sum over unmuted instances feeding c, in
// instance order (src/partconvMulti.cpp:906-921).  One launch per tick; the
// last block to finish (device counter `cnt`) raises the host flag.
__global__ void k_mix(const float *__restrict__ y_inst, const int *__restrict__ src_off,
                      const int *__restrict__ src_list, int N0, int nOut, float *out, int cs, int ss,
                      long long out_off, unsigned *cnt, unsigned long long *done_flag, uint64_t f)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nOut * N0) {
        const int c = e / N0, j = e - c * N0;
        float sum = 0.f;
        for (int q = src_off[c]; q < src_off[c + 1]; q++)
            sum += y_inst[(size_t)src_list[q] * N0 + j];
        out[out_off + (long long)c * cs + (long long)j * ss] = sum;
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
    const size_t base = (size_t)u * 2 * A;
    // H_0 and S of this workgroup's bins, fetched before the transform
    constexpr int BQ = 4;   // 2A <= 4 * BIG_NT (A <= 512)
    cf h0v[BQ], sv[BQ];
    if (mode == PC_MODE_RUN) {
#pragma unroll
        for (int q = 0; q < BQ; q++) {
            const int e = tid + q * BIG_NT;
            if (e < 2 * A) {
                h0v[q] = J.H[base + e];
                sv[q] = J.S[base + e];
            }
        }
    }
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
    if (mode == PC_MODE_SPECTRUM) {
        const float sc = 1.0f / (float)(2 * N);
        for (int e = tid; e < 2 * A; e += BIG_NT) {
            const int q = e >> sh.logA, d = e & (A - 1);
            J.fdl[base + e] = cscale(W[q * sst + d], sc);
        }
        return;
    }
    // FDL store; Y = X H_0 + S (slot order, contiguous per workgroup)
    const int slot = (int)(bcur % (uint64_t)J.P);
#pragma unroll
    for (int qq = 0; qq < BQ; qq++) {
        const int e = tid + qq * BIG_NT;
        if (e >= 2 * A) break;
        const int q = e >> sh.logA, d = e & (A - 1);
        const size_t bin = base + e;
        const cf x = W[q * sst + d];
        J.fdl[(size_t)slot * N + bin] = x;
        cf y = pmul(x, h0v[qq], bin == 0);
        y.x += sv[qq].x;
        y.y += sv[qq].y;
        W[q * sst + d] = y;
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
