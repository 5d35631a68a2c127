// pc_offline.hpp -- batched offline partitioned convolution and its adjoint
// (the differentiable form): whole clips [C, T] are resident, every block of
// every level is computed at once, with the partition rules of the real-time
// engine.  The block transforms are the register-resident kernels of
// pc_fftreg.hpp up to N = 16384 and a four-step pair above.
//
// Per level L (block N, P partitions, T_L taps before it), output block m
// covers samples [mN, (m+1)N):
//   X_{c,m} = r2c( x[c, (m-1)N - T_L + u], u in [0,2N) )          (zero outside [0,T))
//   Y_{c,m} = sum_{p=0}^{min(P-1,m)} X_{c,m-p} H_p                 (H_p as the engine's)
//   y[c, mN+n] += c2r(Y_{c,m})[n],  n < N
// Adjoint, with G_{c,m} = r2c([ g[c, mN .. mN+N) | 0_N ]):
//   dx: W_{c,m} = sum_{p: m+p<M} G_{c,m+p} conj(H_p);
//       dx[c, (m-1)N - T_L + u] += c2r(W_{c,m})[u], u in [0,2N)     (overlap-add)
//   dh: D_p = sum_{c} sum_{m>=p} G_{c,m} conj(X_{c,m-p});
//       dh[T_L + pN + j] = c2r(D_p)[N + j] / (2N)
// (circular convolution with [0_N | h_p] and its two adjoints: circular
// correlation with the kernel, and with the window.)
#pragma once
#include "pc_kernels.hpp"

// Window source of a batched r2c: transform t = r * nb + m reads
//   w[u] = src[r*row_stride + base0 + m*base_step + u]  for ulo <= u < uhi and
//          0 <= base0 + m*base_step + u < len, else 0;   X *= scale (* *dscale).
struct WinSrc {
    const float *src;
    long long row_stride;
    long long len;
    long long base0, base_step;
    int nb;
    int ulo, uhi;
    float scale;
    const float *dscale;  // device scalar the spectra are also scaled by (null: none)
};

// The spectra's scale factor of a window source: scale, times *dscale if set.
__device__ __forceinline__ float win_scale(const WinSrc &W) { return W.dscale ? W.scale * W.dscale[0] : W.scale; }

// Destination of a batched c2r: transform t = r * nb + m writes samples u of
// [ulo, uhi) (scaled) to dst[r*row_stride + base0 + m*base_step + u] when that
// index lies in [0, len); accumulate adds.  parity >= 0 restricts the launch
// to blocks m with m % 2 == parity (overlapping windows never race).
struct InvDst {
    float *dst;
    long long row_stride;
    long long len;
    long long base0, base_step;
    int nb;
    int ulo, uhi;
    float scale;
    int accumulate;
    int parity;
};

__device__ __forceinline__ float win_sample(const WinSrc &W, int r, int m, int u)
{
    if (u < W.ulo || u >= W.uhi) return 0.f;
    const long long i = W.base0 + (long long)m * W.base_step + u;
    return (i >= 0 && i < W.len) ? W.src[(long long)r * W.row_stride + i] : 0.f;
}

// The window of transform (r, m) as contiguous floats, when its nonzero
// samples are [0, nz) with nz = min(uhi, n) a multiple of 4, lie inside the
// signal and start on a 16-byte boundary (every block but the first and last
// few), else null.
__device__ __forceinline__ const float4 *win_fast(const WinSrc &W, int r, int m, int n, int *nz)
{
    *nz = min(W.uhi, n);
    if (W.ulo != 0 || (*nz & 3)) return nullptr;
    const long long i0 = W.base0 + (long long)m * W.base_step;
    if (i0 < 0 || i0 + *nz > W.len) return nullptr;
    const float *p = W.src + (long long)r * W.row_stride + i0;
    return (reinterpret_cast<uintptr_t>(p) & 15) ? nullptr : reinterpret_cast<const float4 *>(p);
}

// The destination of samples [ulo, uhi) of transform (r, m) as contiguous
// floats, when they all lie inside the signal, start on a 16-byte boundary
// and number a multiple of 4, else null.
__device__ __forceinline__ float4 *inv_fast(const InvDst &D, int r, int m)
{
    if ((D.uhi - D.ulo) & 3) return nullptr;
    const long long i0 = D.base0 + (long long)m * D.base_step + D.ulo;
    if (i0 < 0 || i0 + (D.uhi - D.ulo) > D.len) return nullptr;
    float *p = D.dst + (long long)r * D.row_stride + i0;
    return (reinterpret_cast<uintptr_t>(p) & 15) ? nullptr : reinterpret_cast<float4 *>(p);
}

__device__ __forceinline__ void inv_store(const InvDst &D, int r, int m, int u, float v)
{
    if (u < D.ulo || u >= D.uhi) return;
    const long long i = D.base0 + (long long)m * D.base_step + u;
    if (i < 0 || i >= D.len) return;
    float *p = D.dst + (long long)r * D.row_stride + i;
    v *= D.scale;
    if (D.accumulate) *p += v;
    else *p = v;
}

#include "pc_fftreg.hpp"

// ---- N > 16384 (four-step, as k_big1..3) -----------------------------------
// S (scratch) holds N float2 per transform.  Block sizes up to 16384 use the
// register-resident kernels of pc_fftreg.hpp.
__global__ void __launch_bounds__(BIG_NT)
k_off_big_fwd1(WinSrc W, BigShape sh, cf *__restrict__ S, const cf *__restrict__ tw)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = B + 1;
    cf *bufA = lds, *bufB = lds + BIG_G * sst;
    const int groups = A / BIG_G;
    const int t = blockIdx.x / groups, g = blockIdx.x - t * groups;
    const int r = t / W.nb, m = t - r * W.nb;
    const int tid = threadIdx.x;
    const int a0 = g * BIG_G;
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int b = e / BIG_G, ai = e - b * BIG_G;
        const int n = a0 + ai + A * b;
        bufA[ai * sst + b] = make_float2(win_sample(W, r, m, 2 * n), win_sample(W, r, m, 2 * n + 1));
    }
    __syncthreads();
    cf *R = fft_lds<false>(bufA, bufB, sh.logB, BIG_G, sst, tw, tid, BIG_NT);
    cf *St = S + (size_t)t * N;
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int c = e / BIG_G, ai = e - c * BIG_G;
        const int a = a0 + ai;
        St[(size_t)c * A + a] = cmul(R[ai * sst + c], tw[N + ((a * c) & (N - 1))]);
    }
}

// column-pair DFTs + real split -> X (four-step slot order), scaled
__global__ void __launch_bounds__(BIG_NT)
k_off_big_fwd2(BigShape sh, const cf *__restrict__ S, cf *__restrict__ X, float scale, const float *dscale,
               const cf *__restrict__ tw)
{
    const float sc = dscale ? scale * dscale[0] : scale;
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = A + 1;
    cf *bufA = lds, *bufB = lds + 2 * sst;
    const int pairs = B / 2;
    const int t = blockIdx.x / pairs, u = blockIdx.x - t * pairs;
    const int tid = threadIdx.x;
    const int c1 = u, c2 = u ? (B - u) : (B / 2);
    const cf *St = S + (size_t)t * N;
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, a = e & (A - 1);
        bufA[q * sst + a] = St[(size_t)(q ? c2 : c1) * A + a];
    }
    __syncthreads();
    cf *Z = fft_lds<false>(bufA, bufB, sh.logA, 2, sst, tw, tid, BIG_NT);
    cf *W = (Z == bufA) ? bufB : bufA;
    if (u != 0) {
        for (int d = tid; d < A; d += BIG_NT) {
            cf xk, xnk;
            rsplit_fwd(Z[d], Z[sst + (A - 1 - d)], tw[2 * N + c1 + B * d], &xk, &xnk);
            W[d] = xk;
            W[sst + (A - 1 - d)] = xnk;
        }
    } else {
        for (int e = tid; e < A + 1; e += BIG_NT) {
            if (e <= A / 2) {
                const int d = e;
                if (d == 0) {
                    const cf zz = Z[0];
                    W[0] = make_float2(zz.x + zz.y, zz.x - zz.y);
                } else {
                    cf xk, xnk;
                    rsplit_fwd(Z[d], Z[A - d], tw[2 * N + B * d], &xk, &xnk);
                    W[d] = xk;
                    W[A - d] = xnk;
                }
            } else {
                const int d = e - (A / 2 + 1);
                cf xk, xnk;
                rsplit_fwd(Z[sst + d], Z[sst + (A - 1 - d)], tw[2 * N + B / 2 + B * d], &xk, &xnk);
                W[sst + d] = xk;
                W[sst + (A - 1 - d)] = xnk;
            }
        }
    }
    __syncthreads();
    cf *Xt = X + (size_t)t * N + (size_t)u * 2 * A;
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, d = e & (A - 1);
        Xt[e] = cscale(W[q * sst + d], sc);
    }
}

// inverse split + inverse column DFTs + twiddle -> S
__global__ void __launch_bounds__(BIG_NT)
k_off_big_inv1(BigShape sh, const cf *__restrict__ Y, cf *__restrict__ S, const cf *__restrict__ tw, int parity,
               int nb)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = A + 1;
    cf *bufA = lds, *bufB = lds + 2 * sst;
    const int pairs = B / 2;
    int t = blockIdx.x / pairs;
    const int u = blockIdx.x - t * pairs;
    if (parity >= 0) {
        const int half = (nb + 1 - parity) / 2;
        const int r = t / half, mi = t - r * half;
        t = r * nb + 2 * mi + parity;
    }
    const int tid = threadIdx.x;
    const int c1 = u, c2 = u ? (B - u) : (B / 2);
    const cf *Yt = Y + (size_t)t * N + (size_t)u * 2 * A;
    cf *W = bufB;
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, d = e & (A - 1);
        W[q * sst + d] = Yt[e];
    }
    __syncthreads();
    cf *Z = bufA;
    if (u != 0) {
        for (int d = tid; d < A; d += BIG_NT) {
            cf ck, cnk;
            rsplit_inv(W[d], W[sst + (A - 1 - d)], tw[2 * N + c1 + B * d], &ck, &cnk);
            Z[d] = ck;
            Z[sst + (A - 1 - d)] = cnk;
        }
    } else {
        for (int e = tid; e < A + 1; e += BIG_NT) {
            if (e <= A / 2) {
                const int d = e;
                if (d == 0) {
                    const cf y = W[0];
                    Z[0] = make_float2(y.x + y.y, y.x - y.y);
                } else {
                    cf ck, cnk;
                    rsplit_inv(W[d], W[A - d], tw[2 * N + B * d], &ck, &cnk);
                    Z[d] = ck;
                    Z[A - d] = cnk;
                }
            } else {
                const int d = e - (A / 2 + 1);
                cf ck, cnk;
                rsplit_inv(W[sst + d], W[sst + (A - 1 - d)], tw[2 * N + B / 2 + B * d], &ck, &cnk);
                Z[sst + d] = ck;
                Z[sst + (A - 1 - d)] = cnk;
            }
        }
    }
    __syncthreads();
    cf *U = fft_lds<true>(Z, W, sh.logA, 2, sst, tw, tid, BIG_NT);
    cf *St = S + (size_t)t * N;
    for (int e = tid; e < 2 * A; e += BIG_NT) {
        const int q = e >> sh.logA, a = e & (A - 1);
        const int c = q ? c2 : c1;
        St[(size_t)c * A + a] = cmulc(U[q * sst + a], tw[N + ((a * c) & (N - 1))]);
    }
}

// inverse row DFTs -> samples u = 2n, 2n+1 of z'[n], n = a + A b
__global__ void __launch_bounds__(BIG_NT)
k_off_big_inv2(BigShape sh, const cf *__restrict__ S, InvDst D, const cf *__restrict__ tw)
{
    extern __shared__ cf lds[];
    const int A = 1 << sh.logA, B = 1 << sh.logB, N = A * B;
    const int sst = B + 1;
    cf *bufA = lds, *bufB = lds + BIG_G * sst;
    const int groups = A / BIG_G;
    int t = blockIdx.x / groups;
    const int g = blockIdx.x - t * groups;
    if (D.parity >= 0) {
        const int half = (D.nb + 1 - D.parity) / 2;
        const int r = t / half, mi = t - r * half;
        t = r * D.nb + 2 * mi + D.parity;
    }
    const int r = t / D.nb, m = t - r * D.nb;
    const int tid = threadIdx.x;
    const int a0 = g * BIG_G;
    const cf *St = S + (size_t)t * N;
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int c = e / BIG_G, ai = e - c * BIG_G;
        bufA[ai * sst + c] = St[(size_t)c * A + a0 + ai];
    }
    __syncthreads();
    cf *R = fft_lds<true>(bufA, bufB, sh.logB, BIG_G, sst, tw, tid, BIG_NT);
    for (int e = tid; e < BIG_G * B; e += BIG_NT) {
        const int b = e / BIG_G, ai = e - b * BIG_G;
        const int n = a0 + ai + A * b;
        const cf v = R[ai * sst + b];
        inv_store(D, r, m, 2 * n, v.x);
        inv_store(D, r, m, 2 * n + 1, v.y);
    }
}

// ---- spectral products ------------------------------------------------------
// Y[c,m] = sum_p X[c, m - p] H_p          (dir = +1, causal)
// Y[c,m] = sum_p X[c, m + p] conj(H_p)    (dir = -1, adjoint)
// spectra laid out [C][M][N] float2; each thread owns two bins of one (c, m).
__global__ void __launch_bounds__(256)
k_off_mac(const cf *__restrict__ X, const cf *__restrict__ H, cf *__restrict__ Y, int C, int M, int N, int P,
          int dir)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int half = N / 2;
    if (i >= (long long)C * M * half) return;
    const int k2 = (int)(i % half);
    const long long cm = i / half;
    const int m = (int)(cm % M), c = (int)(cm / M);
    const float4 *Xr = reinterpret_cast<const float4 *>(X) + (size_t)c * M * half;
    const float4 *Hr = reinterpret_cast<const float4 *>(H);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = 0; p < P; p++) {
        const int mm = m - dir * p;
        if (mm < 0 || mm >= M) break;
        const float4 a = Xr[(size_t)mm * half + k2];
        float4 h = Hr[(size_t)p * half + k2];
        if (dir < 0) {
            h.y = -h.y;
            h.w = -h.w;
        }
        if (k2 == 0) {
            acc.x += a.x * h.x;
            acc.y += a.y * (dir < 0 ? -h.y : h.y);
        } else {
            acc.x += a.x * h.x - a.y * h.y;
            acc.y += a.x * h.y + a.y * h.x;
        }
        acc.z += a.z * h.z - a.w * h.w;
        acc.w += a.z * h.w + a.w * h.z;
    }
    reinterpret_cast<float4 *>(Y)[i] = acc;
}

// The same products for P <= PW partitions, one chunk of MC blocks per
// workgroup column: a thread owns one (clip, bin pair), keeps the level's PW
// filter spectra of its bins in registers and shifts the block spectra
// through a register window (causal: xw[t] = X[c, m-t], m ascending;
// adjoint: xw[t] = X[c, m+t], m descending), so each block spectrum is read
// once (plus the P-1 blocks before a chunk) instead of once per partition.
// Terms are summed in partition order, as k_off_mac does.  Grid:
// (ceil(C N/2 / 256), ceil(M / MC)).
template <int PW, int DIR, int U = 8>
__global__ void __launch_bounds__(256)
k_off_mac_w(const cf *__restrict__ X, const cf *__restrict__ H, cf *__restrict__ Y, int C, int M, int N, int P,
            int MC)
{
    const int half = N / 2;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)C * half) return;
    const int c = (int)(e / half), k2 = (int)(e - (long long)c * half);
    const float4 *Xr = reinterpret_cast<const float4 *>(X) + (size_t)c * M * half + k2;
    float4 *Yr = reinterpret_cast<float4 *>(Y) + (size_t)c * M * half + k2;
    const float4 *Hr = reinterpret_cast<const float4 *>(H) + k2;
    const bool bin0 = (k2 == 0);
    float4 hr[PW], xw[PW];
#pragma unroll
    for (int t = 0; t < PW; t++) {
        hr[t] = (t < P) ? Hr[(size_t)t * half] : make_float4(0.f, 0.f, 0.f, 0.f);
        xw[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int m0 = blockIdx.y * MC, m1 = min(M, m0 + MC);
    const int mfirst = (DIR > 0) ? max(0, m0 - (P - 1)) : min(M - 1, m1 - 1 + (P - 1));
    const int nstep = (DIR > 0) ? (m1 - mfirst) : (mfirst - m0 + 1);
    // U block spectra are requested together, then shifted through the window
    // one by one: U loads in flight per thread instead of one (with cold
    // caches 63 -> 46 us at the 256 and 2048 levels of 64 clips x 131072)
    for (int s0 = 0; s0 < nstep; s0 += U) {
        float4 xl[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            xl[u] = (s0 + u < nstep) ? Xr[(size_t)(mfirst + DIR * (s0 + u)) * half] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int s = s0 + u;
            if (s >= nstep) break;
            const int m = mfirst + DIR * s;
#pragma unroll
            for (int t = PW - 1; t > 0; t--) xw[t] = xw[t - 1];
            xw[0] = xl[u];
            if ((DIR > 0) ? (m < m0) : (m >= m1)) continue;
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int t = 0; t < PW; t++) {
                // hr[t] is zero for t >= P: those terms add exact zeros
                const float4 a = xw[t], h = hr[t];
                if (bin0) {
                    acc.x += a.x * h.x;
                    acc.y += a.y * h.y;
                } else if (DIR > 0) {
                    acc.x += a.x * h.x - a.y * h.y;
                    acc.y += a.x * h.y + a.y * h.x;
                } else {
                    acc.x += a.x * h.x + a.y * h.y;
                    acc.y += a.y * h.x - a.x * h.y;
                }
                if (DIR > 0) {
                    acc.z += a.z * h.z - a.w * h.w;
                    acc.w += a.z * h.w + a.w * h.z;
                } else {
                    acc.z += a.z * h.z + a.w * h.w;
                    acc.w += a.w * h.z - a.z * h.w;
                }
            }
            Yr[(size_t)m * half] = acc;
        }
    }
}

// D partial sums, one chunk of MC blocks per workgroup column:
//   part[p][c nch + j][k] = sum_{m in chunk j, m >= p} G[c,m,k] conj(X[c,m-p,k])
// (nch = ceil(M / MC) chunks, m ascending).  A thread owns one (clip, bin
// pair) and the PT partitions p0 .. p0+PT-1 (p0 = PT blockIdx.y): each step of
// m loads G[c,m] and X[c,m-p0] once and shifts X through a register window
// (xw[t] = X[c,m-p0-t]; zero before block 0), so a spectrum is read once per
// group of PT partitions, not once per partition.  A chunk first shifts in the
// PT-1 blocks before it.  Grid: (ceil(C N/2 / 256), ceil(P / PT), nch).
template <int PT, int U = 4>
__global__ void __launch_bounds__(256)
k_off_dh_part(const cf *__restrict__ G, const cf *__restrict__ X, cf *__restrict__ part, int C, int M, int N,
              int P, int MC)
{
    const int half = N / 2;
    const int nch = (M + MC - 1) / MC;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)C * half) return;
    const int c = (int)(e / half), k2 = (int)(e - (long long)c * half);
    const int p0 = blockIdx.y * PT, j = blockIdx.z;
    const float4 *Gr = reinterpret_cast<const float4 *>(G) + (size_t)c * M * half + k2;
    const float4 *Xr = reinterpret_cast<const float4 *>(X) + (size_t)c * M * half + k2;
    const int m0 = j * MC, m1 = min(M, m0 + MC);
    const bool bin0 = (k2 == 0);
    float4 xw[PT], acc[PT];
#pragma unroll
    for (int t = 0; t < PT; t++) {
        xw[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // U steps' spectra are requested together, then used one step at a time
    for (int mb = max(p0, m0 - (PT - 1)); mb < m1; mb += U) {
        float4 xl[U], gl[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = mb + u;
            xl[u] = (m < m1) ? Xr[(size_t)(m - p0) * half] : make_float4(0.f, 0.f, 0.f, 0.f);
            gl[u] = (m >= m0 && m < m1) ? Gr[(size_t)m * half] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = mb + u;
            if (m >= m1) break;
#pragma unroll
            for (int t = PT - 1; t > 0; t--) xw[t] = xw[t - 1];
            xw[0] = xl[u];
            if (m < m0) continue;
            const float4 g = gl[u];
#pragma unroll
            for (int t = 0; t < PT; t++) {
                const float4 x = xw[t];
                if (bin0) {
                    acc[t].x += g.x * x.x;
                    acc[t].y += g.y * x.y;
                } else {
                    acc[t].x += g.x * x.x + g.y * x.y;
                    acc[t].y += g.y * x.x - g.x * x.y;
                }
                acc[t].z += g.z * x.z + g.w * x.w;
                acc[t].w += g.w * x.z - g.z * x.w;
            }
        }
    }
    float4 *dst = reinterpret_cast<float4 *>(part) + ((size_t)c * nch + j) * half + k2;
#pragma unroll
    for (int t = 0; t < PT; t++)
        if (p0 + t < P) dst[(size_t)(p0 + t) * C * nch * half] = acc[t];
}

// D[p][k] = sum_j part[p][j][k] over the J = C nch partial sums (j: clip-major,
// then chunk).  A workgroup takes 16 consecutive outputs; its 16 thread rows
// each sum one contiguous segment of j in order, and the segments are added
// in order: a fixed order, independent of timing.  Grid: ceil(P N / 16).
__global__ void __launch_bounds__(256)
k_off_dh_reduce(const cf *__restrict__ part, cf *__restrict__ D, int J, int N, int P)
{
    __shared__ cf sp[16][16];
    const int kk = threadIdx.x & 15, sg = threadIdx.x >> 4;
    const long long i = (long long)blockIdx.x * 16 + kk;
    const bool ok = i < (long long)P * N;
    const int p = ok ? (int)(i / N) : 0, k = ok ? (int)(i % N) : 0;
    const int len = (J + 15) / 16, j0 = sg * len, j1 = min(J, j0 + len);
    cf acc = make_float2(0.f, 0.f);
    if (ok) {
        const cf *src = part + (size_t)p * J * N + k;
#pragma unroll 8
        for (int j = j0; j < j1; j++) {
            const cf v = src[(size_t)j * N];
            acc.x += v.x;
            acc.y += v.y;
        }
    }
    sp[sg][kk] = acc;
    __syncthreads();
    if (sg == 0 && ok) {
        cf t = sp[0][kk];
        for (int q = 1; q < 16; q++) {
            t.x += sp[q][kk].x;
            t.y += sp[q][kk].y;
        }
        D[i] = t;
    }
}

// ---------------------------------------------------------------------------
// MIMO (throughput form): for every bin k, with M output blocks at once,
//   Y_k[o][m] = sum_p sum_i H_k,p[o][i] X_k[i][m - p]
// is a complex (32 x 32P) x (32P x M) product, run on the fp32 matrix cores
// (v_mfma_f32_32x32x2f32) as two real 32x32 tiles per wave:
//   Yr = sum_i Hr Xr - Hi Xi,  Yi = sum_i Hr Xi + Hi Xr,
// with K ordered (i, re/im): lane l supplies A[o = l%32][k = l/32] and
// B[k = l/32][m = l%32]; accumulator register v of lane l holds
// C[8(v/4) + 4(l/32) + v%4][l%32].  Packed bin 0 holds the real DC and
// Nyquist bins, which multiply separately (Yr = sum Hr Xr, Yi = sum Hi Xi).
// Layouts: Ht[k][p][i][o] (o contiguous, 32 wide), Xt[k][i][m] (m contiguous),
// Y[o][m][k] (bins contiguous, the inverse transforms' input).
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(64)
k_mimo_mfma(const cf *__restrict__ Ht, const cf *__restrict__ Xt, cf *__restrict__ Y, int N, int M, int P,
            int n_in, int n_out)
{
    const int k = blockIdx.x;
    const int m0 = blockIdx.y * 32;
    const int l = threadIdx.x;
    const int lo = l & 31, hi = l >> 5;
    f32x16 accr, acci;
    for (int v = 0; v < 16; v++) {
        accr[v] = 0.f;
        acci[v] = 0.f;
    }
    const bool bin0 = (k == 0);
    const int m = m0 + lo;
    for (int p = 0; p < P; p++) {
        const cf *Hp = Ht + ((size_t)k * P + p) * n_in * 32;
        const int mm = m - p;
        const bool mv = (mm >= 0 && mm < M);
        const cf *Xk = Xt + (size_t)k * n_in * M;
        for (int j = 0; j < n_in; j++) {
            const cf h = Hp[j * 32 + lo];
            const cf x = mv ? Xk[(size_t)j * M + mm] : make_float2(0.f, 0.f);
            float ar, ai;
            if (!bin0) {
                ar = hi ? -h.y : h.x;
                ai = hi ? h.x : h.y;
            } else {
                ar = hi ? 0.f : h.x;
                ai = hi ? h.y : 0.f;
            }
            const float b = hi ? x.y : x.x;
            accr = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, b, accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, b, acci, 0, 0, 0);
        }
    }
    if (m >= M) return;
    for (int v = 0; v < 16; v++) {
        const int o = 8 * (v >> 2) + 4 * hi + (v & 3);
        if (o < n_out) Y[((size_t)o * M + m) * N + k] = make_float2(accr[v], acci[v]);
    }
}

// The same product with MT waves per workgroup, one m-tile each: the
// partition-p filter block H_k,p (n_in x 32 complex) is staged in LDS once for
// the workgroup (double-buffered, the next partition's loads in flight during
// this one's MFMAs) instead of being read from memory by every m-tile, and the
// workgroup's input window (blocks m0 - P + 1 .. m0 + 32 MT - 1 of every input,
// zero outside 0 .. M-1) is staged once, since partition p reads the blocks of
// partition p-1 shifted by one.  Each wave runs exactly k_mimo_mfma's sequence
// of MFMAs, so the results are the same bit for bit.  Grid: (N, ceil(ceil(M /
// 32) / MT)), block 64 MT, dynamic LDS (2 * 32 * 32 + n_in * (32 MT + P - 1))
// complex values.
template <int MT>
__global__ void __launch_bounds__(64 * MT)
k_mimo_mfma_lds(const cf *__restrict__ Ht, const cf *__restrict__ Xt, cf *__restrict__ Y, int N, int M, int P,
                int n_in, int n_out)
{
    constexpr int NT = 64 * MT, HS = 32 * 32;  // n_in <= 32
    constexpr int PER = (HS + NT - 1) / NT;
    extern __shared__ cf mm_lds[];
    cf *const hs0 = mm_lds, *const hs1 = mm_lds + HS;  // filter blocks, by partition parity
    cf *const xs = mm_lds + 2 * HS;                      // input window [n_in][WL]
    const int k = blockIdx.x;
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
    const int m0 = (blockIdx.y * MT + wv) * 32;
    const bool wave_on = m0 < M;
    const int lo = l & 31, hi = l >> 5;
    const int nh = n_in * 32;
    f32x16 accr, acci;
    for (int v = 0; v < 16; v++) {
        accr[v] = 0.f;
        acci[v] = 0.f;
    }
    const bool bin0 = (k == 0);
    const int m = m0 + lo;
    const cf *Xk = Xt + (size_t)k * n_in * M;
    const cf *Hk = Ht + (size_t)k * P * n_in * 32;
    const int mlo = blockIdx.y * MT * 32 - (P - 1), WL = 32 * MT + P - 1;
    for (int e = tid; e < nh; e += NT) hs0[e] = Hk[e];
    for (int e = tid; e < n_in * WL; e += NT) {
        const int i = e / WL, r = mlo + (e - i * WL);
        xs[e] = (r >= 0 && r < M) ? Xk[(size_t)i * M + r] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int p = 0; p < P; p++) {
        cf pre[PER];
        const cf *Hn = Hk + (size_t)(p + 1) * n_in * 32;
        if (p + 1 < P) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = tid + q * NT;
                if (e < nh) pre[q] = Hn[e];
            }
        }
        if (wave_on) {
            const cf *hp = (p & 1) ? hs1 : hs0;
            const cf *xp = xs + (m - p - mlo);  // block m - p, zero outside the clip
            for (int j = 0; j < n_in; j++) {
                const cf h = hp[j * 32 + lo];
                const cf x = xp[j * WL];
                float ar, ai;
                if (!bin0) {
                    ar = hi ? -h.y : h.x;
                    ai = hi ? h.x : h.y;
                } else {
                    ar = hi ? 0.f : h.x;
                    ai = hi ? h.y : 0.f;
                }
                const float b = hi ? x.y : x.x;
                accr = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, b, accr, 0, 0, 0);
                acci = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, b, acci, 0, 0, 0);
            }
        }
        if (p + 1 < P) {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = tid + q * NT;
                if (e < nh) ((p & 1) ? hs0 : hs1)[e] = pre[q];
            }
        }
        __syncthreads();
    }
    if (!wave_on || m >= M) return;
    for (int v = 0; v < 16; v++) {
        const int o = 8 * (v >> 2) + 4 * hi + (v & 3);
        if (o < n_out) Y[((size_t)o * M + m) * N + k] = make_float2(accr[v], acci[v]);
    }
}

// [R][M][N] -> [N][R][M] (spectra of R signals, M blocks, N bins)
__global__ void k_spec_transpose(const cf *__restrict__ src, cf *__restrict__ dst, int R, int M, int N)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)R * M * N) return;
    const int k = (int)(i % N);
    const long long rm = i / N;
    const int m = (int)(rm % M), r = (int)(rm / M);
    dst[((size_t)k * R + r) * M + m] = src[i];
}

// H [(o*n_in + i)][p][k] -> Ht[k][p][i][o32] (o padded to 32 with zeros)
__global__ void k_h_transpose(const cf *__restrict__ src, cf *__restrict__ dst, int n_in, int n_out, int P, int N)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)N * P * n_in * 32) return;
    const int o = (int)(i % 32);
    long long r = i / 32;
    const int ii = (int)(r % n_in);
    r /= n_in;
    const int p = (int)(r % P);
    const int k = (int)(r / P);
    dst[i] = (o < n_out) ? src[(((size_t)o * n_in + ii) * P + p) * N + k] : make_float2(0.f, 0.f);
}

// ---- dL/dh partial sums with one bin per thread (float2) ----
// As k_off_dh_part, with a thread owning one (clip, bin) instead of a bin
// pair: half the registers per partition, so PT = 16 partitions fit one group
// and G and X are read once for P <= 16 (the bin-pair form with PT = 8 reads
// them twice for P = 14), with twice the waves in flight.  Packed bin 0 holds
// the real DC and Nyquist bins.  Same partial-sum layout and summation order
// as k_off_dh_part.  Grid: (ceil(C N / 256), ceil(P / PT), nch).
template <int PT, int U = 4>
__global__ void __launch_bounds__(256)
k_off_dh_part_b(const cf *__restrict__ G, const cf *__restrict__ X, cf *__restrict__ part, int C, int M, int N,
                int P, int MC)
{
    const int nch = (M + MC - 1) / MC;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)C * N) return;
    const int c = (int)(e / N), k = (int)(e - (long long)c * N);
    const int p0 = blockIdx.y * PT, j = blockIdx.z;
    const cf *Gr = G + (size_t)c * M * N + k;
    const cf *Xr = X + (size_t)c * M * N + k;
    const int m0 = j * MC, m1 = min(M, m0 + MC);
    const bool bin0 = (k == 0);
    cf xw[PT], acc[PT];
#pragma unroll
    for (int t = 0; t < PT; t++) {
        xw[t] = make_float2(0.f, 0.f);
        acc[t] = make_float2(0.f, 0.f);
    }
    for (int mb = max(p0, m0 - (PT - 1)); mb < m1; mb += U) {
        cf xl[U], gl[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = mb + u;
            xl[u] = (m < m1) ? Xr[(size_t)(m - p0) * N] : make_float2(0.f, 0.f);
            gl[u] = (m >= m0 && m < m1) ? Gr[(size_t)m * N] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int m = mb + u;
            if (m >= m1) break;
#pragma unroll
            for (int t = PT - 1; t > 0; t--) xw[t] = xw[t - 1];
            xw[0] = xl[u];
            if (m < m0) continue;
            const cf g = gl[u];
#pragma unroll
            for (int t = 0; t < PT; t++) {
                const cf x = xw[t];
                if (bin0) {
                    acc[t].x += g.x * x.x;
                    acc[t].y += g.y * x.y;
                } else {
                    acc[t].x += g.x * x.x + g.y * x.y;
                    acc[t].y += g.y * x.x - g.x * x.y;
                }
            }
        }
    }
    cf *dst = part + ((size_t)c * nch + j) * N + k;
#pragma unroll
    for (int t = 0; t < PT; t++)
        if (p0 + t < P) dst[(size_t)(p0 + t) * C * nch * N] = acc[t];
}

// ---------------------------------------------------------------------------
// Mean squared error of y against t over n samples, with its gradient:
//   g = (2 / n) (y - t),  loss = sum (y - t)^2 / n.
// k_mse_grad: block b covers the contiguous samples [b chunk, (b + 1) chunk),
// each thread a strided subset in order, then a fixed tree over the block, so
// part[b] does not depend on timing; k_mse_final sums part[0 .. nb) the same
// way in one block.  Fused, the loss reads y and t once and writes g once.
// ---------------------------------------------------------------------------
#define PC_MSE_NT 256
#define PC_MSE_BLOCKS 1024

__device__ __forceinline__ float block_sum_fixed(float v, float *red, int tid)
{
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int d = PC_MSE_NT / 2; d > 0; d >>= 1) {
        if (tid < d) red[tid] += red[tid + d];
        __syncthreads();
    }
    return red[0];
}

__global__ void __launch_bounds__(PC_MSE_NT)
k_mse_grad(const float *__restrict__ y, const float *__restrict__ t, float *__restrict__ g, float gfac, long long n,
           long long chunk, int vec, float *__restrict__ part)
{
    __shared__ float red[PC_MSE_NT];
    const int tid = threadIdx.x;
    const long long lo = (long long)blockIdx.x * chunk, hi = min(n, lo + chunk);
    float acc = 0.f;
    // float4 groups (vec: 16-byte aligned arrays and chunk a multiple of 4), the rest by scalars
    const long long hi4 = vec ? lo + ((hi - lo) & ~3ll) : lo;
    const float4 *y4 = reinterpret_cast<const float4 *>(y), *t4 = reinterpret_cast<const float4 *>(t);
    float4 *g4 = reinterpret_cast<float4 *>(g);
#pragma unroll 4
    for (long long i = lo / 4 + tid; i < hi4 / 4; i += PC_MSE_NT) {
        const float4 a = y4[i], b = t4[i];
        const float4 d = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
        g4[i] = make_float4(gfac * d.x, gfac * d.y, gfac * d.z, gfac * d.w);
        acc = fmaf(d.x, d.x, acc);
        acc = fmaf(d.y, d.y, acc);
        acc = fmaf(d.z, d.z, acc);
        acc = fmaf(d.w, d.w, acc);
    }
    for (long long i = hi4 + tid; i < hi; i += PC_MSE_NT) {
        const float d = y[i] - t[i];
        g[i] = gfac * d;
        acc = fmaf(d, d, acc);
    }
    const float sum = block_sum_fixed(acc, red, tid);
    if (tid == 0) part[blockIdx.x] = sum;
}

__global__ void __launch_bounds__(PC_MSE_NT)
k_mse_final(const float *__restrict__ part, int nb, float inv_n, float *__restrict__ loss)
{
    __shared__ float red[PC_MSE_NT];
    const int tid = threadIdx.x;
    float acc = 0.f;
    for (int i = tid; i < nb; i += PC_MSE_NT) acc += part[i];
    const float sum = block_sum_fixed(acc, red, tid);
    if (tid == 0) *loss = sum * inv_n;
}
