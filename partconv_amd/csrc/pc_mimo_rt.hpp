// pc_mimo_rt.hpp -- real-time MIMO engine: n_in inputs, n_out outputs, one
// filter per (output, input) pair, all with the same partition.
//
// The reference formulates a MIMO network as n_in x n_out independent
// PartConvFilter instances whose outputs are summed per output channel
// (src/partconvMulti.cpp:815-928): every input spectrum is computed, stored
// in a delay line and streamed once per pair.  Here the input side is shared:
// per level and block, each input's window is transformed once (one delay
// line per input), and per frequency bin the delay-line product is a complex
// (n_out x n_in P) matrix times (n_in P) vector,
//     Y_o[k] = sum_p sum_i H_p[o][i][k] X_i^(m-p)[k],
// followed by one inverse transform per output.  The product is split as in
// the instance engine: Y = (p = 0 term) + S, S = sum_{p>=1} precomputed after
// block m-1, off the path from a block's input to its output.
//
// Level 0 runs in one persistent cooperative kernel (k_mxr_persist) whose
// workgroups have roles -- input, output, delay-line sum -- and synchronise
// through two monotonic counters instead of grid barriers (see the kernel).
// Levels >= 1 are host-launched per block on their own streams: k_mxr_fwd
// (one workgroup per input), k_mxr_comb (per bin), k_mxr_inv (per output),
// k_mxr_part (per bin, the next block's S).
//
// Layouts (device, bins in natural order, packed bin 0 = (X_0, X_N)):
//   ring[i]   : 2R floats (doubled ring, as the instance engine's)
//   fdl       : [P][N][ni]       H : [N][P][n_out][ni] (scaled by 1/2N; ni = n_in
//               rounded up to 4, zero-padded, so a lane's four inputs or four
//               filters are one aligned 32-byte load; a bin's inputs share two
//               cache lines)
//   S, Y      : [n_out][N] (level 0: S twice, by tick parity)   ybuf : [n_out][2][N] (levels >= 1)
// Frames in and out are interleaved as PartConvMulti::run's buffers.
#pragma once
#include "pc_kernels.hpp"

#define MXR_NT 256          // threads of the level kernels
#define MXR_NTT 256         // threads of the persistent tick kernel's workgroups
#define MXR_MAX_IO 32       // inputs and outputs (32 outputs x NT/32 lanes per bin)

struct MxrLevel {
    int N, P, kshift;
    int ni;                 // fdl and H row stride: n_in rounded up to 4 (zero-padded)
    long long T;            // taps before this level
    cf *fdl;
    const cf *H;
    cf *S;
    cf *Y;
    float *ybuf;
};

// Level 0's delay-line sums S are kept for MXR_SLOTS ticks: tick f reads slot
// f % MXR_SLOTS, its delay-line-sum workgroups write slot (f + 1) % MXR_SLOTS,
// whose last reader (tick f + 1 - MXR_SLOTS) is complete whenever the host
// releases at most MXR_SLOTS - 1 ticks ahead of the oldest incomplete one.
#define MXR_SLOTS 4
struct MxrArgs {
    int n_in, n_out, N0, M;
    float *ring;            // [n_in][2R]
    long long R, ring_mask;
    MxrLevel lv[PC_MAX_LEVELS];
    const MxrLevel *lvd;    // the same levels in device memory: the kernel indexes
                            // levels >= 1 with a loop variable, which on a by-value
                            // argument array would move the array into registers
    const float *in;
    long long in_stride;
    float *out;
    long long out_stride;
    long long fbase;
    uint64_t f0;
    long long nticks;
    const unsigned long long *go;
    unsigned long long *flags;  // host-visible, one per output workgroup, 8 words apart
    unsigned long long *bar;    // the tick counters: cX, then cS in 8 groups, 8 words apart (zeroed
                                // before each launch)
    unsigned long long spin_limit;
    int hcache;             // output workgroups keep H_0[o] in LDS (after the FFT buffers and twiddles)
    int window;             // ticks the host may release ahead (release window, 1 .. MXR_SLOTS - 1)
    int lpb;                // lanes per bin of level 0's p = 0 product: next power of two >= ni / 2
    unsigned *present;      // plain launch: residency stamps (see stamp_present), else null
    unsigned present_tag;
};

// Bin k of level L, for every output: the sum over partitions p in [p_lo, p_hi)
// and inputs of H_p[o][i][k] X_i^(b-p)[k], plus base[o][k] if given, into
// dst[o][k].  With NT threads, NT/32 lanes share an output (lane t: output
// t / (NT/32)) and each takes II = 1024/NT consecutive inputs.  The loads of
// PB partitions are issued together before they are summed (in partition,
// then input order): the loop is otherwise one memory round trip per
// partition.  The partial sums of an output are added in a fixed tree over
// its consecutive lanes, so the result does not depend on timing.
template <int NT, int PB>
__device__ __forceinline__ void mxr_bin_acc(const MxrLevel &L, int n_in, int n_out, int k, uint64_t b, int p_lo,
                                            int p_hi, cf &acc, int tid)
{
    constexpr int LPO = NT / 32, II = 32 / LPO;  // lanes per output, inputs per lane
    static_assert(II == 2 || II == 4, "256 or 512 threads");
    const int o = tid / LPO, ig = tid % LPO;
    const bool bin0 = (k == 0);
    const bool act = (o < n_out) && (II * ig < n_in);
    if (!act) return;
    const int bm = (int)(b % (uint64_t)L.P);
    for (int p0 = p_lo; p0 < p_hi; p0 += PB) {
        float4 hv[PB][II / 2], xv[PB][II / 2];
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int p = p0 + j;
            if (p < p_hi) {
                int r = bm - p;
                if (r < 0) r += L.P;
                const float4 *h = reinterpret_cast<const float4 *>(
                    L.H + (((size_t)k * L.P + p) * n_out + o) * L.ni + II * ig);
                const float4 *x = reinterpret_cast<const float4 *>(
                    L.fdl + ((size_t)r * L.N + k) * L.ni + II * ig);
#pragma unroll
                for (int q = 0; q < II / 2; q++) {
                    hv[j][q] = h[q];
                    xv[j][q] = x[q];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < PB; j++) {
            if (p0 + j < p_hi) {
#pragma unroll
                for (int q = 0; q < II / 2; q++) {
                    acc = pmac(acc, make_float2(xv[j][q].x, xv[j][q].y), make_float2(hv[j][q].x, hv[j][q].y),
                               bin0);
                    acc = pmac(acc, make_float2(xv[j][q].z, xv[j][q].w), make_float2(hv[j][q].z, hv[j][q].w),
                               bin0);
                }
            }
        }
    }
}

template <int NT>
__device__ __forceinline__ void mxr_bin_finish(const MxrLevel &L, int n_out, int k, cf acc, const cf *base, cf *dst,
                                               int tid)
{
    constexpr int LPO = NT / 32;
    const int o = tid / LPO, ig = tid % LPO;
#pragma unroll
    for (int d = 1; d < LPO; d <<= 1) {
        acc.x += __shfl_xor(acc.x, d);
        acc.y += __shfl_xor(acc.y, d);
    }
    if (o < n_out && ig == 0) {
        if (base) {
            const cf s0 = gload2(base + (size_t)o * L.N + k);
            acc.x += s0.x;
            acc.y += s0.y;
        }
        dst[(size_t)o * L.N + k] = acc;
    }
}

template <int NT, int PB>
__device__ __forceinline__ void mxr_bin(const MxrLevel &L, int n_in, int n_out, int k, uint64_t b, int p_lo,
                                        int p_hi, const cf *base, cf *dst, int tid)
{
    cf acc = make_float2(0.f, 0.f);
    mxr_bin_acc<NT, PB>(L, n_in, n_out, k, b, p_lo, p_hi, acc, tid);
    mxr_bin_finish<NT>(L, n_out, k, acc, base, dst, tid);
}

// Workgroup-wide bounded wait until *ctr >= target (thread 0 polls), then a
// device-scope acquire.  Returns 0 on timeout.
__device__ __forceinline__ int mxr_wait(const unsigned long long *ctr, unsigned long long target,
                                        unsigned long long limit, int tid, int *s_ok)
{
    if (tid == 0) {
        const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + limit;
        int ok = 1;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__builtin_amdgcn_s_memrealtime() > t_end) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok;
}

// The same for counter c1 and the MXR_SG group counters c2[8 g] at once
// (every load in flight per poll).
#define MXR_SG 8
__device__ __forceinline__ int mxr_wait2(const unsigned long long *c1, unsigned long long t1,
                                         const unsigned long long *c2, unsigned long long t2,
                                         unsigned long long limit, int tid, int *s_ok)
{
    if (tid == 0) {
        const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + limit;
        int ok = 1;
        for (;;) {
            const unsigned long long v1 = __hip_atomic_load(c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned long long v2 = ~0ull;
#pragma unroll
            for (int g = 0; g < MXR_SG; g++)
                v2 = min(v2, __hip_atomic_load(c2 + 8 * g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (v1 >= t1 && v2 >= t2) break;
            if (__builtin_amdgcn_s_memrealtime() > t_end) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok;
}

// Every wave's stores complete, device-scope release, then count one arrival.
__device__ __forceinline__ void mxr_arrive(unsigned long long *ctr, int tid)
{
    wg_stores_done();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Level 0's p = 0 term for output o over every bin, without S:
//   Y[k] = sum_i H_0[o][i][k] X_i[k]   (X: delay-line row r0).
// A bin's ni inputs are 8 ni contiguous bytes of the delay line, so lpb lanes
// (the next power of two >= ni / 2) share a bin and each takes one input pair:
// every load instruction of a wave then reads 64 / lpb whole bins, one
// contiguous run, instead of one bin per lane 8 ni bytes apart (which touched
// 64 lines per instruction and fetched each line again for every 16 bytes).
// The pair products are summed over the bin's lanes in a fixed xor tree, so
// the result does not depend on timing.  Hc: H_0[o] bin-major (Hc[k ni + i])
// in LDS, or null to read H from memory.  B steps' delay-line loads (and,
// from memory, filter loads) are in flight together.
template <int B, bool HC>
__device__ __forceinline__ void mxr_out_product(const MxrLevel &L0, int o, int n_out, int N0, int r0, const cf *Hc,
                                                cf *Y, int lpb, int tid)
{
    const int bpi = 64 / lpb;           // bins per load instruction of a wave
    const int steps = lpb;              // a wave takes 64 bins: 64 / bpi steps
    const int lane = tid & 63, wv = tid >> 6;
    const int q = lane & (lpb - 1), bsub = lane / lpb;
    const int ni = L0.ni, np = ni / 2;  // inputs per bin row (padded to 4), input pairs
    const bool lq = q < np;
    const float4 *X4 = reinterpret_cast<const float4 *>(L0.fdl + (size_t)r0 * N0 * ni);
    const float4 *H4 = reinterpret_cast<const float4 *>(L0.H);
    const float4 *Hc4 = reinterpret_cast<const float4 *>(Hc);
#pragma unroll 1
    for (int k0 = wv * 64; k0 < N0; k0 += MXR_NTT) {
#pragma unroll 1
        for (int j0 = 0; j0 < steps; j0 += B) {
            float4 xv[B], hv[HC ? 1 : B];
#pragma unroll
            for (int b = 0; b < B; b++) {
                const int k = k0 + (j0 + b) * bpi + bsub;
                xv[b] = make_float4(0.f, 0.f, 0.f, 0.f);
                if constexpr (!HC) hv[b] = xv[b];
                if (j0 + b < steps && k < N0 && lq) {
                    xv[b] = X4[(size_t)k * np + q];
                    if constexpr (!HC) hv[b] = H4[((size_t)k * L0.P * n_out + o) * np + q];
                }
            }
#pragma unroll
            for (int b = 0; b < B; b++) {
                if (j0 + b >= steps) break;
                const int k = k0 + (j0 + b) * bpi + bsub;
                const bool bin0 = (k == 0);
                float4 h = make_float4(0.f, 0.f, 0.f, 0.f);
                if constexpr (HC) {
                    if (k < N0 && lq) h = Hc4[(size_t)k * np + q];
                } else {
                    h = hv[b];
                }
                cf a = pmul(make_float2(xv[b].x, xv[b].y), make_float2(h.x, h.y), bin0);
                a = pmac(a, make_float2(xv[b].z, xv[b].w), make_float2(h.z, h.w), bin0);
                // fixed xor tree over the bin's lanes (lpb is uniform)
                a = cadd(a, xor_cf<1>(a, lane));
                if (lpb > 2) a = cadd(a, xor_cf<2>(a, lane));
                if (lpb > 4) a = cadd(a, xor_cf<4>(a, lane));
                if (lpb > 8) a = cadd(a, xor_cf<8>(a, lane));
                if (q == 0 && k < N0) Y[k] = a;
            }
        }
    }
}

// Workgroups by role: [0, n_in) input, [n_in, n_in + n_out) output, the rest
// (G4 of them) the delay-line sum.  Per tick t of the launch (f = f0 + t):
//   input:  release; frame -> ring, window -> r2c -> delay-line row f % P;
//           arrive at cX;
//   output: release; fetch the level blocks due; wait cX >= (t+1) n_in
//           (this tick's rows) and every cS group >= t G4 / MXR_SG (S for
//           tick f complete); Y_o = S_o + p = 0 product over every bin, c2r
//           + the level blocks -> output, flag;
//   sum:    release; wait cX >= (t+1) n_in; S for tick f + 1 of its bins;
//           arrive at its cS group (workgroup index mod MXR_SG: the 256
//           arrivals of a tick on one address serialise, 8 groups on their
//           own lines measured 28.0 -> 23.5 us per tick at 32 x 32).
// S has MXR_SLOTS slots (tick f reads slot f % MXR_SLOTS; its sum writes slot
// (f+1) % MXR_SLOTS), so the sum for the next tick runs while this tick's
// outputs are formed and while the host's round trip and the next tick's input
// transforms proceed: the delay-line sum is off the critical path.  With a
// release window the roles may be at different ticks: input workgroups first
// wait for the sums that still read the row they overwrite.  (A grid
// barrier per tick instead, with every workgroup summing after it, measured
// 43 us per tick against 32 x 32 channels.)
template <int LOGN>
__global__ void __launch_bounds__(MXR_NTT) k_mxr_persist(MxrArgs A, const cf *__restrict__ tw)
{
    constexpr int N0 = 1 << LOGN;
    extern __shared__ cf mxr_lds[];
    __shared__ int s_ok;
    cf *bufA = mxr_lds, *bufB = mxr_lds + N0;
    const int tid = threadIdx.x, w = blockIdx.x;
    const int n_io = A.n_in + A.n_out, G4 = (int)gridDim.x - n_io;
    const MxrLevel &L0 = A.lv[0];
    unsigned long long *cX = A.bar, *cS = A.bar + 8;
    stamp_present(A.present, A.present_tag, tid);
    // input and output workgroups: the transforms' twiddles (entries below
    // 2 N0 + N0 / 2 + 1) in LDS, so the transforms make no global loads;
    // output workgroups also keep their partition-0 filter spectra, bin-major
    // as in H (Hc[k ni + i], padded inputs included)
    cf *twl = mxr_lds + 2 * N0;
    cf *Hc = mxr_lds + 5 * N0;
    if (w < n_io) {
        for (int e = tid; e <= 2 * N0 + N0 / 2; e += MXR_NTT) twl[e] = tw[e];
        if (A.hcache && w >= A.n_in) {
            const int o = w - A.n_in;
            for (int e = tid; e < N0 * L0.ni; e += MXR_NTT) {
                const int k = e / L0.ni, i = e - k * L0.ni;
                Hc[e] = L0.H[((size_t)k * L0.P * A.n_out + o) * L0.ni + i];
            }
        }
        __syncthreads();
    }
    // release word last read by thread 0 (a release window covers later ticks too)
    unsigned long long g_seen = 0;
    for (long long t = 0; t < A.nticks; t++) {
        const uint64_t f = A.f0 + (uint64_t)t;
        // ---- release
        if (tid == 0) {
            int go = go_released(g_seen, f);
            if (!go) {
                const unsigned long long t_end = __builtin_amdgcn_s_memrealtime() + A.spin_limit;
                for (;;) {
                    const unsigned long long g = __hip_atomic_load(A.go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
            s_ok = go;
        }
        __syncthreads();
        if (!s_ok) return;
        if (w < A.n_in) {
            // ---- input: frame -> ring, window [frame f-1 | frame f] -> r2c -> delay line
            float *ring = A.ring + (size_t)w * 2 * A.R;
            const float *fr = A.in + (long long)(f - A.fbase) * A.in_stride + w;
            float *win = reinterpret_cast<float *>(bufA);
            const long long t0 = (long long)f * N0;
            for (int j = tid; j < N0; j += MXR_NTT) {
                const float x = fr[(long long)j * A.n_in];
                const long long p = (t0 + j) & A.ring_mask;
                ring[p] = x;
                ring[p + A.R] = x;
                win[j] = ring[(t0 - N0 + j) & A.ring_mask];
                win[N0 + j] = x;
            }
            __syncthreads();
            const cf *X = small_forward<LOGN, MXR_NTT>(bufA, bufB, twl, tid);
            // With ticks released ahead the roles' workgroups may be at different
            // ticks.  The counters count arrivals, so an input workgroup arrives for
            // tick f only once every input has arrived for tick f - 1; and row f % P
            // (tick f - P's) may still be read by the delay-line sums of tick f - 2
            // (partition P - 1 of S_{f-1}): wait for them too.  One tick at a time,
            // both hold already (tick f - 1's outputs waited for them).
            if (A.window > 1 && t >= 1 &&
                !mxr_wait2(cX, (unsigned long long)t * A.n_in, cS,
                           (t >= 2) ? (unsigned long long)(t - 1) * (G4 / MXR_SG) : 0ull, A.spin_limit, tid, &s_ok))
                return;
            cf *row = L0.fdl + (size_t)(f % (uint64_t)L0.P) * N0 * L0.ni + w;
            for (int k = tid; k < N0; k += MXR_NTT) row[(size_t)k * L0.ni] = X[k];
            mxr_arrive(cX, tid);
        } else if (w < n_io) {
            // ---- output o: Y_o = S_o + sum_i H_0[o][i] X_i (inputs in order), c2r + levels
            const int o = w - A.n_in;
            const uint64_t fr = f + 1;  // strict: block fr/k - 1, sub-block fr % k
            // the level blocks' outputs were complete before the release (the
            // host waits for them): fetch them while the inputs transform
            constexpr int JO = (N0 + MXR_NTT - 1) / MXR_NTT;
            float ylv[JO];
            if (A.M > 1) {
                if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < JO; q++) {
                const int j = tid + q * MXR_NTT;
                ylv[q] = 0.f;
                if (j < N0)
                    for (int l = 1; l < A.M; l++) {
                        const MxrLevel &L = A.lvd[l];
                        const uint64_t qb = fr >> L.kshift;
                        if (qb == 0) continue;
                        const uint64_t sub = fr & ((1ull << L.kshift) - 1);
                        ylv[q] += gmem(L.ybuf)[((size_t)o * 2 + (size_t)((qb - 1) & 1)) * L.N + sub * N0 + j];
                    }
            }
            // this tick's delay-line rows and S for this tick
            if (!mxr_wait2(cX, (unsigned long long)(t + 1) * A.n_in, cS, (unsigned long long)t * (G4 / MXR_SG), A.spin_limit,
                           tid, &s_ok))
                return;
            // S for this tick, read coalesced now; added after the product
            // (sk + a: the same sum as before, addition being commutative)
            const cf *S = L0.S + (size_t)(f % MXR_SLOTS) * A.n_out * N0 + (size_t)o * N0;
            constexpr int JS = (N0 + MXR_NTT - 1) / MXR_NTT;
            cf sk[JS];
#pragma unroll
            for (int q = 0; q < JS; q++) {
                const int k = tid + q * MXR_NTT;
                sk[q] = (k < N0) ? gload2(S + k) : make_float2(0.f, 0.f);
            }
            const int r0 = (int)(f % (uint64_t)L0.P);
            if (A.hcache) mxr_out_product<8, true>(L0, o, A.n_out, N0, r0, Hc, bufA, A.lpb, tid);
            else mxr_out_product<4, false>(L0, o, A.n_out, N0, r0, nullptr, bufA, A.lpb, tid);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < JS; q++) {
                const int k = tid + q * MXR_NTT;
                if (k < N0) {
                    const cf a = bufA[k];
                    bufA[k] = make_float2(sk[q].x + a.x, sk[q].y + a.y);
                }
            }
            __syncthreads();
            const cf *z = small_inverse<LOGN, MXR_NTT>(bufA, bufB, twl, tid);
            const float *zf = reinterpret_cast<const float *>(z);
            float *out = A.out + (long long)(f - A.fbase) * A.out_stride + o;
#pragma unroll
            for (int q = 0; q < JO; q++) {
                const int j = tid + q * MXR_NTT;
                if (j < N0) out[(long long)j * A.n_out] = zf[j] + ylv[q];
            }
            wg_stores_done();
            if (tid == 0)
                __hip_atomic_store(A.flags + (size_t)o * 8, (unsigned long long)(f + 1), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
        } else {
            // ---- delay-line sum for tick f + 1 of this workgroup's bins.  Partitions
            // 2 .. P-1 read delay-line rows of earlier ticks (visible since this
            // workgroup's acquire at the previous tick), so they are summed right
            // after the release; partition 1, this tick's row, once every input has
            // arrived, then the lanes' sums are combined.
            constexpr int KB = (N0 + MXR_NTT - 1) / MXR_NTT;  // bins per workgroup at most (G4 >= N0 / KB)
            cf acc[KB];
#pragma unroll
            for (int j = 0; j < KB; j++) {
                acc[j] = make_float2(0.f, 0.f);
                const int k = w - n_io + j * G4;
                if (k < N0) mxr_bin_acc<MXR_NTT, 4>(L0, A.n_in, A.n_out, k, f + 1, 2, L0.P, acc[j], tid);
            }
            // this tick's rows, and every sum of tick f - 1: tick f is released once
            // tick f - 1's outputs are out, which needs only the sums of tick f - 2,
            // so a sum workgroup may start tick f while another still works on tick
            // f - 1; the group counters count arrivals, and an arrival for tick f
            // before all of tick f - 1's would let tick f's outputs read S early
            if (!mxr_wait2(cX, (unsigned long long)(t + 1) * A.n_in, cS, (unsigned long long)t * (G4 / MXR_SG),
                           A.spin_limit, tid, &s_ok))
                return;
            cf *Sn = L0.S + (size_t)((f + 1) % MXR_SLOTS) * A.n_out * N0;
#pragma unroll
            for (int j = 0; j < KB; j++) {
                const int k = w - n_io + j * G4;
                if (k < N0) {
                    mxr_bin_acc<MXR_NTT, 4>(L0, A.n_in, A.n_out, k, f + 1, 1, min(2, L0.P), acc[j], tid);
                    mxr_bin_finish<MXR_NTT>(L0, A.n_out, k, acc[j], nullptr, Sn, tid);
                }
            }
            mxr_arrive(cS + 8 * ((w - n_io) % MXR_SG), tid);  // group counters: 8 lines
        }
        __syncthreads();
    }
}

// Level >= 1, block m: input workgroup i transforms its window into delay-line row m % P.
template <int LOGN>
__global__ void __launch_bounds__(MXR_NT) k_mxr_fwd(MxrLevel L, const float *__restrict__ ring_all, long long R,
                                                     long long ring_mask, int N0, uint64_t m,
                                                     const cf *__restrict__ tw)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ cf mxr_lds[];
    cf *bufA = mxr_lds, *bufB = mxr_lds + N;
    const int i = blockIdx.x, tid = threadIdx.x;
    const float *ring = ring_all + (size_t)i * 2 * R;
    const long long s = (long long)m * N - N0 - L.T;
    const cf *wsrc = reinterpret_cast<const cf *>(ring + (s & ring_mask));
    for (int n = tid; n < N; n += MXR_NT) bufA[n] = wsrc[n];
    __syncthreads();
    const cf *X = small_forward<LOGN, MXR_NT>(bufA, bufB, tw, tid);
    cf *row = L.fdl + (size_t)(m % (uint64_t)L.P) * N * L.ni + i;
    for (int k = tid; k < N; k += MXR_NT) row[(size_t)k * L.ni] = X[k];
}

// Level >= 1, block b: p = 0 term + S -> Y (one workgroup per bin).
__global__ void __launch_bounds__(MXR_NT) k_mxr_comb(MxrLevel L, int n_in, int n_out, uint64_t b)
{
    mxr_bin<MXR_NT, 8>(L, n_in, n_out, blockIdx.x, b, 0, 1, L.S, L.Y, threadIdx.x);
}

// Level >= 1, S for block b (partitions 1 .. P-1; one workgroup per bin).
// It runs beside the persistent kernel, so both batch four partitions per
// round of loads: 84 and about 156 VGPRs, and several of its workgroups fit
// on a CU that holds a persistent one.  (Eight per round in both: 155 and
// 245 VGPRs, 1.45 TB/s here against 2.6 TB/s, measured at 32 x 32 pairs.)
__global__ void __launch_bounds__(MXR_NT) k_mxr_part(MxrLevel L, int n_in, int n_out, uint64_t b)
{
    mxr_bin<MXR_NT, 4>(L, n_in, n_out, blockIdx.x, b, 1, L.P, nullptr, L.S, threadIdx.x);
}

// Level >= 1, block m: output workgroup o: c2r of Y_o -> double-buffer half m & 1.
template <int LOGN>
__global__ void __launch_bounds__(MXR_NT) k_mxr_inv(MxrLevel L, uint64_t m, const cf *__restrict__ tw)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ cf mxr_lds[];
    cf *bufA = mxr_lds, *bufB = mxr_lds + N;
    const int o = blockIdx.x, tid = threadIdx.x;
    for (int k = tid; k < N; k += MXR_NT) bufA[k] = L.Y[(size_t)o * N + k];
    __syncthreads();
    const cf *z = small_inverse<LOGN, MXR_NT>(bufA, bufB, tw, tid);
    const float *zf = reinterpret_cast<const float *>(z);
    float *dst = L.ybuf + ((size_t)o * 2 + (size_t)(m & 1)) * N;
    for (int j = tid; j < N; j += MXR_NT) dst[j] = zf[j];
}

// Filter spectra of one level: workgroup (p, pair) transforms
// [0_N | h_pair[T + pN .. T + pN + N)] (zero past the taps), scales by 1/2N
// and stores H[k][p][o][i] (pair = o n_in + i; h is [n_out][n_in][taps]).
template <int LOGN>
__global__ void __launch_bounds__(MXR_NT) k_mxr_spec(MxrLevel L, const float *__restrict__ h, int taps, int n_in,
                                                      int n_out, const cf *__restrict__ tw)
{
    constexpr int N = 1 << LOGN;
    extern __shared__ cf mxr_lds[];
    cf *bufA = mxr_lds, *bufB = mxr_lds + N;
    const int p = blockIdx.x, pair = blockIdx.y, tid = threadIdx.x;
    const int o = pair / n_in, i = pair - o * n_in;
    const float *hp = h + (size_t)pair * taps;
    const long long t0 = L.T + (long long)p * N;
    float *win = reinterpret_cast<float *>(bufA);
    for (int n = tid; n < 2 * N; n += MXR_NT) {
        const long long t = t0 + n - N;
        win[n] = (n >= N && t < taps) ? hp[t] : 0.f;
    }
    __syncthreads();
    const cf *X = small_forward<LOGN, MXR_NT>(bufA, bufB, tw, tid);
    const float sc = 1.0f / (float)(2 * N);
    cf *H = const_cast<cf *>(L.H);
    for (int k = tid; k < N; k += MXR_NT)
        H[(((size_t)k * L.P + p) * n_out + o) * L.ni + i] = make_float2(X[k].x * sc, X[k].y * sc);
}
