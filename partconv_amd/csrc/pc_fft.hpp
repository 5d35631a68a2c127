// pc_fft.hpp -- LDS-resident Stockham FFT building blocks for gfx950.
//
// Complex values are float2.  Twiddles come from a device table built once in
// double precision (host) and rounded to float:  tw[n + t] = exp(-2*pi*i*t/n)
// for every power of two n up to the table size, so a transform of length n
// reads the contiguous slice tw[n .. 2n).
//
// fft_lds() runs `nseq` independent transforms of length 2^logn that live in
// LDS at stride `sst` (sst >= n; an odd padding avoids bank conflicts when
// another stage walks the sequences transposed).  It ping-pongs between two
// LDS buffers with radix-4 passes (plus one radix-2 pass for odd logn) in the
// Stockham autosort form, so the result is in natural order; it returns the
// buffer that holds the result.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float2 cf;

__device__ __forceinline__ cf cadd(cf a, cf b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cf csub(cf a, cf b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cf cconj(cf a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ cf cscale(cf a, float s) { return make_float2(a.x * s, a.y * s); }
// (ar*br - ai*bi, ar*bi + ai*br)
__device__ __forceinline__ cf cmul(cf a, cf b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ cf cmulc(cf a, cf b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// acc += a*b  (complex multiply-accumulate)
__device__ __forceinline__ cf cmac(cf acc, cf a, cf b)
{
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
    return acc;
}

template <bool INV>
__device__ __forceinline__ cf twiddle(const cf *__restrict__ tw, int n, int t)
{
    cf w = tw[n + t];
    if (INV) w.y = -w.y;
    return w;
}

// One radix-4 butterfly in place on v[0..3]; forward uses -i, inverse +i.
template <bool INV>
__device__ __forceinline__ void dft4(cf &a0, cf &a1, cf &a2, cf &a3)
{
    const cf t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const cf t2 = cadd(a1, a3), t3 = csub(a1, a3);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    if (!INV) {
        a1 = make_float2(t1.x + t3.y, t1.y - t3.x);  // t1 - i t3
        a3 = make_float2(t1.x - t3.y, t1.y + t3.x);  // t1 + i t3
    } else {
        a1 = make_float2(t1.x - t3.y, t1.y + t3.x);
        a3 = make_float2(t1.x + t3.y, t1.y - t3.x);
    }
}

// Batched Stockham FFT in LDS.  Threads tid in [0, nthr) of the workgroup take
// part; every thread must call it (it contains barriers).  Input in `a`.
template <bool INV>
__device__ cf *fft_lds(cf *a, cf *b, int logn, int nseq, int sst, const cf *__restrict__ tw,
                       int tid, int nthr)
{
    const int n = 1 << logn;
    int lgNs = 0;
    cf *src = a, *dst = b;
    while (lgNs < logn) {
        const int Ns = 1 << lgNs;
        if (logn - lgNs >= 2) {
            const int nb = n >> 2;
            const int total = nb * nseq;
            const int tstep = n >> (lgNs + 2);  // n / (Ns*4)
            for (int g = tid; g < total; g += nthr) {
                const int s = g >> (logn - 2);
                const int j = g & (nb - 1);
                const cf *sp = src + s * sst;
                cf v0 = sp[j], v1 = sp[j + nb], v2 = sp[j + 2 * nb], v3 = sp[j + 3 * nb];
                const int jm = j & (Ns - 1);
                if (lgNs > 0) {
                    const int t = jm * tstep;
                    v1 = cmul(v1, twiddle<INV>(tw, n, t));
                    v2 = cmul(v2, twiddle<INV>(tw, n, 2 * t));
                    v3 = cmul(v3, twiddle<INV>(tw, n, 3 * t));
                }
                dft4<INV>(v0, v1, v2, v3);
                cf *dp = dst + s * sst + ((j - jm) << 2) + jm;
                dp[0] = v0;
                dp[Ns] = v1;
                dp[2 * Ns] = v2;
                dp[3 * Ns] = v3;
            }
            lgNs += 2;
        } else {
            const int nb = n >> 1;
            const int total = nb * nseq;
            const int tstep = n >> (lgNs + 1);
            for (int g = tid; g < total; g += nthr) {
                const int s = g >> (logn - 1);
                const int j = g & (nb - 1);
                const cf *sp = src + s * sst;
                cf v0 = sp[j], v1 = sp[j + nb];
                const int jm = j & (Ns - 1);
                if (lgNs > 0) v1 = cmul(v1, twiddle<INV>(tw, n, jm * tstep));
                cf *dp = dst + s * sst + ((j - jm) << 1) + jm;
                dp[0] = cadd(v0, v1);
                dp[Ns] = csub(v0, v1);
            }
            lgNs += 1;
        }
        __syncthreads();
        cf *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Real-FFT split for one bin pair (k, N-k) of a length-2N real transform whose
// even/odd samples were packed as z[n] = x[2n] + i x[2n+1] and transformed to
// Z (length N).  a = Z[k], b = Z[N-k], w = exp(-i*pi*k/N).
// Returns X[k] in *xk and X[N-k] in *xnk.
__device__ __forceinline__ void rsplit_fwd(cf a, cf b, cf w, cf *xk, cf *xnk)
{
    const cf bc = cconj(b);
    const cf e = cscale(cadd(a, bc), 0.5f);
    const cf d = csub(a, bc);
    const cf o = make_float2(0.5f * d.y, -0.5f * d.x);  // (a - conj b) / (2i)
    const cf wo = cmul(w, o);
    *xk = cadd(e, wo);
    *xnk = cconj(csub(e, wo));
}

// Inverse split for bin pair (k, N-k) of a half spectrum Y (Hermitian of a real
// 2N signal): returns C[k], C[N-k] with C = A + iB, A = Y[k] + conj Y[N-k],
// B = (Y[k] - conj Y[N-k]) * exp(+i*pi*k/N).  w = exp(-i*pi*k/N) (conjugated here).
__device__ __forceinline__ void rsplit_inv(cf a, cf b, cf w, cf *ck, cf *cnk)
{
    const cf bc = cconj(b);
    const cf A = cadd(a, bc);
    const cf D = csub(a, bc);
    const cf B = cmulc(D, w);  // D * conj(w) = D * exp(+i pi k / N)
    *ck = make_float2(A.x - B.y, A.y + B.x);
    *cnk = make_float2(A.x + B.y, -A.y + B.x);
}
