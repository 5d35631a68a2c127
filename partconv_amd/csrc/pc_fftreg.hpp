// pc_fftreg.hpp -- batched real FFTs for the offline (whole-clip) transforms
// with the data in registers.
//
// Each thread holds 16 complex values of one length-N transform (N = 2^LOGN,
// 16 <= N <= 16384) and runs radix-16 Stockham passes on them, plus one
// radix-2, 4 or 8 pass when LOGN is not a multiple of 4.  Between passes the
// values go through LDS once (N = 2048: 3 passes, 2 exchanges; the LDS
// radix-4 form of fft_lds makes 6).  The forward's first pass loads its
// inputs straight from the signal, and the inverse's last pass stores its
// outputs straight to the signal.  A workgroup is one wave holding 64 / (N/16)
// transforms for N <= 1024, two transforms for N = 2048, else one transform
// (N/16 threads); N = 16384 takes 139 KB of LDS, one workgroup per CU, and
// replaces the four-step pair with its scratch round trip.  Every pass's
// twiddles are requested at the start (w^j, w^2j, w^4j, w^8j per butterfly;
// the other powers are products of these), so no global load waits behind a
// barrier.
//
// Pass s (span NS = 16^s, radix R) is the decimation-in-time Stockham step
// fft_lds uses: butterfly j reads elements j + q N/R (q < R), multiplies
// element q by exp(-+2 pi i (j mod NS) q / (NS R)) from the twiddle table
// (tw[n + t] = exp(-2 pi i t / n)), takes the R-point DFT in registers, and
// writes output q to (j - j mod NS) R + (j mod NS) + q NS.  LDS indices skip
// one slot every 16 (i + i/16), so both the strided writes of the first pass
// and the row writes of later passes stay free of bank conflicts.
#pragma once
#include "pc_fft.hpp"

template <int LOGN>
struct RegFft {
    static_assert(LOGN >= 4 && LOGN <= 14, "register FFT covers 16 .. 16384 points");
    static constexpr int N = 1 << LOGN;
    static constexpr int TT = N / 16;                     // threads per transform
    static constexpr int NP16 = LOGN / 4;                 // radix-16 passes
    static constexpr int RL = 1 << (LOGN % 4);            // radix of the last pass (1: none)
    static constexpr int NPASS = NP16 + (RL > 1 ? 1 : 0);
    static constexpr int TPB = TT >= 256 ? 1 : TT >= 128 ? 2 : 64 / TT;  // transforms per workgroup
    static constexpr int NT = TT * TPB;                   // threads per workgroup
    static constexpr int LDSN = N + N / 16;               // padded slots per transform
};

__device__ __forceinline__ int rpad(int i) { return i + (i >> 4); }

// a * exp(-+ i pi/2): -i forward, +i inverse
template <bool INV>
__device__ __forceinline__ cf rot4(cf a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// a * exp(-+ i pi/4)
template <bool INV>
__device__ __forceinline__ cf rot8(cf a)
{
    const float r = 0.70710678118654752f;
    return INV ? make_float2(r * (a.x - a.y), r * (a.x + a.y)) : make_float2(r * (a.x + a.y), r * (a.y - a.x));
}
// a * exp(-+ 3 i pi/4)
template <bool INV>
__device__ __forceinline__ cf rot8_3(cf a)
{
    const float r = 0.70710678118654752f;
    return INV ? make_float2(-r * (a.x + a.y), r * (a.x - a.y)) : make_float2(r * (a.y - a.x), -r * (a.x + a.y));
}
// a * exp(-+ i theta), cos theta = c, sin theta = s
template <bool INV>
__device__ __forceinline__ cf rotc(cf a, float c, float s)
{
    return INV ? make_float2(a.x * c - a.y * s, a.y * c + a.x * s) : make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
}

// R-point DFTs of v[0 .. R) in place, natural order in and out.
template <bool INV>
__device__ __forceinline__ void dft2r(cf *v)
{
    const cf a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4r(cf *v)
{
    dft4<INV>(v[0], v[1], v[2], v[3]);
}

// n = 2 n1 + n2: four-point DFTs over n1, twiddles exp(-+2 pi i n2 k1 / 8),
// two-point DFTs over n2; X[k1 + 4 k2].
template <bool INV>
__device__ __forceinline__ void dft8r(cf *v)
{
    dft4<INV>(v[0], v[2], v[4], v[6]);
    dft4<INV>(v[1], v[3], v[5], v[7]);
    v[3] = rot8<INV>(v[3]);
    v[5] = rot4<INV>(v[5]);
    v[7] = rot8_3<INV>(v[7]);
    cf o[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
        o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = o[k];
}

// n = 4 n1 + n2: four-point DFTs over n1, twiddles exp(-+2 pi i n2 k1 / 16),
// four-point DFTs over n2; X[k1 + 4 k2].
template <bool INV>
__device__ __forceinline__ void dft16r(cf *v)
{
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f;  // cos, sin of pi/8
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4<INV>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
    // v[n2 + 4 k1] *= w^(n2 k1)
    v[5] = rotc<INV>(v[5], c1, s1);         // w^1
    v[9] = rot8<INV>(v[9]);                 // w^2
    v[13] = rotc<INV>(v[13], s1, c1);       // w^3
    v[6] = rot8<INV>(v[6]);                 // w^2
    v[10] = rot4<INV>(v[10]);               // w^4
    v[14] = rot8_3<INV>(v[14]);             // w^6
    v[7] = rotc<INV>(v[7], s1, c1);         // w^3
    v[11] = rot8_3<INV>(v[11]);             // w^6
    {
        const cf t = rotc<INV>(v[15], c1, s1);  // w^9 = -w^1
        v[15] = make_float2(-t.x, -t.y);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    cf o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = o[k];
}

template <int R, bool INV>
__device__ __forceinline__ void dftr(cf *v)
{
    if constexpr (R == 16) dft16r<INV>(v);
    else if constexpr (R == 8) dft8r<INV>(v);
    else if constexpr (R == 4) dft4r<INV>(v);
    else dft2r<INV>(v);
}

// Radix, span and butterflies per thread of pass PASS; LB = log2 R twiddle
// powers w^(jm), w^(2 jm), ... are loaded per butterfly.
template <int LOGN, int PASS>
struct PassShape {
    static constexpr int R = (PASS < RegFft<LOGN>::NP16) ? 16 : RegFft<LOGN>::RL;
    static constexpr int NS = 1 << (4 * PASS);
    static constexpr int NB = 16 / R;
    static constexpr int LB = (R == 16) ? 4 : (R == 8) ? 3 : (R == 4) ? 2 : 1;
};
#define REG_TWB 8  // base twiddle slots per pass (NB * LB <= 8)

// Every pass's base twiddles, loaded once at the start of the kernel so that
// their latency overlaps the first loads instead of following each barrier:
// for pass s >= 1 and butterfly b, w^(2^k jm) with w = exp(-+2 pi i / (NS R)),
// jm = (t + b TT) mod NS, k < LB.
template <int LOGN, bool INV, int PASS>
__device__ __forceinline__ void reg_twiddles(cf *twb, int t, const cf *__restrict__ tw)
{
    if constexpr (PASS < RegFft<LOGN>::NPASS) {
        using S = PassShape<LOGN, PASS>;
#pragma unroll
        for (int b = 0; b < S::NB; b++) {
            const int jm = (t + b * RegFft<LOGN>::TT) & (S::NS - 1);
#pragma unroll
            for (int k = 0; k < S::LB; k++) twb[PASS * REG_TWB + b * S::LB + k] = twiddle<INV>(tw, S::NS * S::R, jm << k);
        }
        reg_twiddles<LOGN, INV, PASS + 1>(twb, t, tw);
    }
}

// w[q] = w^q for q = 1 .. R-1 from (w, w^2, w^4, w^8): at most three complex
// products each.
template <int R>
__device__ __forceinline__ void tw_expand(const cf *base, cf *w)
{
    w[1] = base[0];
    if constexpr (R > 2) {
        w[2] = base[1];
        w[3] = cmul(w[1], w[2]);
    }
    if constexpr (R > 4) {
        w[4] = base[2];
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = base[3];
#pragma unroll
        for (int q = 1; q < 8; q++) w[8 + q] = cmul(w[8], w[q]);
    }
}

// exp(-i pi q / 16), q = 0 .. 15: the real-split twiddle of bin t + q N/16 is
// exp(-i pi t / N) times this.
__device__ __forceinline__ cf split_rot(int q)
{
    // cos, sin of pi q / 16 for q = 0 .. 8
    constexpr float c[9] = {1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
                            0.70710678118654752f, 0.55557023301960218f, 0.38268343236508978f,
                            0.19509032201612825f, 0.0f};
    return q <= 8 ? make_float2(c[q], -c[8 - q]) : make_float2(-c[16 - q], -c[q - 8]);
}

// Pass PASS of a transform whose pass-PASS inputs are in v (PASS 0) or in L
// (later passes), with the base twiddles of reg_twiddles in twb.  Runs this
// pass and every later one; on return v holds the
// last pass's outputs: butterfly b's output q is element
// t + b TT + q N / RL_last (the last pass has span N / R).  Contains barriers:
// every thread of the workgroup calls it.
template <int LOGN, bool INV, int PASS>
__device__ __forceinline__ void reg_passes(cf *v, cf *L, int t, const cf *twb)
{
    using F = RegFft<LOGN>;
    using S = PassShape<LOGN, PASS>;
    constexpr int N = F::N, TT = F::TT;
    constexpr int R = S::R, NS = S::NS, NB = S::NB;
    if constexpr (PASS > 0) {
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int q = 0; q < R; q++) v[b * R + q] = L[rpad(t + b * TT + q * (N / R))];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            cf w[R];
            tw_expand<R>(twb + PASS * REG_TWB + b * S::LB, w);
#pragma unroll
            for (int q = 1; q < R; q++) v[b * R + q] = cmul(v[b * R + q], w[q]);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) dftr<R, INV>(v + b * R);
    if constexpr (PASS + 1 < F::NPASS) {
        __syncthreads();  // every input of this pass has been read from L
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + b * TT, jm = j & (NS - 1);
#pragma unroll
            for (int q = 0; q < R; q++) L[rpad((j - jm) * R + jm + q * NS)] = v[b * R + q];
        }
        __syncthreads();
        reg_passes<LOGN, INV, PASS + 1>(v, L, t, twb);
    }
}

// Last-pass layout: radix and butterflies per thread.
template <int LOGN>
struct RegLast {
    static constexpr int R = (RegFft<LOGN>::RL > 1) ? RegFft<LOGN>::RL : 16;
    static constexpr int NB = 16 / R;
    static constexpr int SPAN = RegFft<LOGN>::N / R;  // element of output q of butterfly j: j + q SPAN
};

// Pass-0 inputs of a forward transform: v[q] = (w[2n], w[2n+1]), n = t + q TT,
// from the window of transform (r, m) (zero outside [ulo, uhi) and the signal).
template <int LOGN>
__device__ __forceinline__ void reg_load_window(const WinSrc &W, int r, int m, int t, bool live, cf *v)
{
    constexpr int TT = RegFft<LOGN>::TT, N = RegFft<LOGN>::N;
    if (!live) {
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = make_float2(0.f, 0.f);
        return;
    }
    const int nz = min(W.uhi, 2 * N);
    const long long i0 = W.base0 + (long long)m * W.base_step;
    const float *p = W.src + (long long)r * W.row_stride + i0;
    if (W.ulo == 0 && !(nz & 1) && i0 >= 0 && i0 + nz <= W.len && !(reinterpret_cast<uintptr_t>(p) & 7)) {
        const float2 *p2 = reinterpret_cast<const float2 *>(p);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int n = t + q * TT;
            v[q] = (2 * n < nz) ? p2[n] : make_float2(0.f, 0.f);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int n = t + q * TT;
            v[q] = make_float2(win_sample(W, r, m, 2 * n), win_sample(W, r, m, 2 * n + 1));
        }
    }
}

// X[t] = scale * r2c(window t), packed (X[0] = (X_0, X_N)), N float2 per transform.
template <int LOGN, int TPB = RegFft<LOGN>::TPB>
__global__ void __launch_bounds__(RegFft<LOGN>::TT * TPB)
k_off_rfft_reg(WinSrc W, int ntrans, cf *__restrict__ X, const cf *__restrict__ tw)
{
    using F = RegFft<LOGN>;
    using LP = RegLast<LOGN>;
    constexpr int N = F::N, TT = F::TT;
    __shared__ cf lds[TPB * F::LDSN];
    const int sub = threadIdx.x / TT, t = threadIdx.x - sub * TT;
    const int tr = blockIdx.x * TPB + sub;
    const bool live = tr < ntrans;
    const int r = live ? tr / W.nb : 0, m = live ? tr - r * W.nb : 0;
    cf *L = lds + sub * F::LDSN;
    cf twb[F::NPASS * REG_TWB];
    reg_twiddles<LOGN, false, 1>(twb, t, tw);
    const cf wsplit = tw[2 * N + t];  // exp(-i pi t / N)
    cf v[16];
    reg_load_window<LOGN>(W, r, m, t, live, v);
    reg_passes<LOGN, false, 0>(v, L, t, twb);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < LP::NB; b++)
#pragma unroll
        for (int q = 0; q < LP::R; q++) L[rpad(t + b * TT + q * LP::SPAN)] = v[b * LP::R + q];
    __syncthreads();
    if (!live) return;
    cf *o = X + (size_t)tr * N;
    const float sc = win_scale(W);
#pragma unroll
    for (int i = 0; i <= 8; i++) {
        const int k = t + i * TT;
        if (k > N / 2) break;
        if (k == 0) {
            const cf z = L[0];
            o[0] = cscale(make_float2(z.x + z.y, z.x - z.y), sc);
        } else {
            cf xk, xnk;
            rsplit_fwd(L[rpad(k)], L[rpad(N - k)], i ? cmul(wsplit, split_rot(i)) : wsplit, &xk, &xnk);
            o[k] = cscale(xk, sc);
            o[N - k] = cscale(xnk, sc);
        }
    }
}

// c2r of spectra Y (packed, N float2 each) into D: samples u = 2n, 2n+1 of
// z'[n] for u in [ulo, uhi), scaled, stored or added.  ntrans transforms
// (all blocks, or those of D.parity).
template <int LOGN, int TPB = RegFft<LOGN>::TPB>
__global__ void __launch_bounds__(RegFft<LOGN>::TT * TPB)
k_off_irfft_reg(const cf *__restrict__ Y, InvDst D, int ntrans, const cf *__restrict__ tw)
{
    using F = RegFft<LOGN>;
    using LP = RegLast<LOGN>;
    constexpr int N = F::N, TT = F::TT;
    __shared__ cf lds[TPB * F::LDSN];
    const int sub = threadIdx.x / TT, t = threadIdx.x - sub * TT;
    const int ti = blockIdx.x * TPB + sub;
    const bool live = ti < ntrans;
    int tr = live ? ti : 0;
    if (D.parity >= 0) {
        // ti enumerates (row, m) with m of the requested parity
        const int half = (D.nb + 1 - D.parity) / 2;
        const int r = tr / half, mi = tr - r * half;
        tr = r * D.nb + 2 * mi + D.parity;
    }
    const int r = tr / D.nb, m = tr - r * D.nb;
    cf *L = lds + sub * F::LDSN;
    cf twb[F::NPASS * REG_TWB];
    reg_twiddles<LOGN, true, 1>(twb, t, tw);
    const cf wsplit = tw[2 * N + t];  // exp(-i pi t / N)
    if (live) {
        const cf *y = Y + (size_t)tr * N;
#pragma unroll
        for (int q = 0; q < 16; q++) L[rpad(t + q * TT)] = y[t + q * TT];
    }
    __syncthreads();
    cf v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int k = t + q * TT;
        if (k == 0) {
            const cf y0 = L[0];
            v[q] = make_float2(y0.x + y0.y, y0.x - y0.y);
        } else {
            cf ck, cnk;
            rsplit_inv(L[rpad(k)], L[rpad(N - k)], q ? cmul(wsplit, split_rot(q)) : wsplit, &ck, &cnk);
            v[q] = ck;
        }
    }
    reg_passes<LOGN, true, 0>(v, L, t, twb);
    if (!live) return;
    const long long i0 = D.base0 + (long long)m * D.base_step;
    float *row = D.dst + (long long)r * D.row_stride;
    if (!(D.ulo & 1) && !(D.uhi & 1) && i0 + D.ulo >= 0 && i0 + D.uhi <= D.len &&
        !(reinterpret_cast<uintptr_t>(row + i0) & 7)) {
        float2 *d2 = reinterpret_cast<float2 *>(row + i0);
#pragma unroll
        for (int b = 0; b < LP::NB; b++)
#pragma unroll
            for (int q = 0; q < LP::R; q++) {
                const int n = t + b * TT + q * LP::SPAN;
                if (2 * n < D.ulo || 2 * n >= D.uhi) continue;
                const cf z = v[b * LP::R + q];
                float2 o = make_float2(z.x * D.scale, z.y * D.scale);
                if (D.accumulate) {
                    const float2 a = d2[n];
                    o = make_float2(a.x + o.x, a.y + o.y);
                }
                d2[n] = o;
            }
    } else {
#pragma unroll
        for (int b = 0; b < LP::NB; b++)
#pragma unroll
            for (int q = 0; q < LP::R; q++) {
                const int n = t + b * TT + q * LP::SPAN;
                const cf z = v[b * LP::R + q];
                inv_store(D, r, m, 2 * n, z.x);
                inv_store(D, r, m, 2 * n + 1, z.y);
            }
    }
}
