/*
 * pc_oracle.c -- CPU restatement of partconv's partitioned-convolution engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it.  The product
 * (partconv_amd) never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's generation-B engine:
 *   PartConvFilter::setup      src/partconv.cpp:70-306   (validation, delays,
 *                                                          last-level recount,
 *                                                          splitVector, H = r2c/2N)
 *   PartConvFilter::runLevel   src/partconv.cpp:354-490  (window shift, r2c, FDL
 *                                                          slot, MAC p=0..P-1, c2r)
 *   PartConvFilter::reset      src/partconv.cpp:492-506
 *   sync_levels[_relaxed]      src/partconv.cpp:568-603
 *   input / output[_relaxed]   src/partconv.cpp:605-648
 *   DelayBuffer                src/partconv.cpp:1401-1472
 *   DoubleBuffer               src/partconv.cpp:1477-1511
 *   MatrixTemplate::splitVector src/buffers.cpp:193-217
 *   PartConvMulti::run         src/partconvMulti.cpp:815-928 (strict)
 *   PartConvMultiRelaxed::run  src/partconvMulti.cpp:1017-1134 (relaxed)
 *
 * The reference runs levels >= 1 on pthread workers; their results only become
 * visible through the double buffers after the run-side wait, so executing every
 * level inline at its start tick gives bit-identical output.  That is what this
 * restatement does (scheduling affects timing only).
 *
 * Arithmetic type is PCO_REAL (float by default = the reference's FFTW fp32 path;
 * built a second time with -DPCO_REAL=double as an fp64 restatement of the same
 * algorithm).  FFTW is absent from this image, so the real FFTs are a
 * project-written radix-2 transform (half-length complex FFT + split), with
 * twiddles generated in double.  Its output is pinned against numpy's float64
 * rfft by the tests.
 *
 * Test-only mutations (pco_set_mutation) plant known mistakes so the tests can
 * show that the pins catch them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef PCO_REAL
#define PCO_REAL float
#endif
typedef PCO_REAL real;

#define PCO_EXPORT __attribute__((visibility("default")))

static int g_mutation = 0;
static char g_err[512];

PCO_EXPORT void pco_set_mutation(int m) { g_mutation = m; }
PCO_EXPORT const char *pco_last_error(void) { return g_err; }
PCO_EXPORT int pco_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------ */
/* FFT: complex radix-2 DIT (in place, bit-reversed input), real wrappers.   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;          /* complex length (= N, half the real length 2N) */
    real *tw;       /* n/2 twiddles e^{-2 pi i j / n}, interleaved re/im */
    real *tw2;      /* N+1 split twiddles e^{-i pi k / N} for the real split */
    int *rev;       /* bit reversal permutation of length n */
    real *work;     /* 2n scratch */
} fftplan;

static void plan_init(fftplan *p, int N)
{
    p->n = N;
    p->tw = (real *)malloc(sizeof(real) * 2 * (N / 2 + 1));
    p->tw2 = (real *)malloc(sizeof(real) * 2 * (N + 1));
    p->rev = (int *)malloc(sizeof(int) * N);
    p->work = (real *)malloc(sizeof(real) * 2 * N);
    for (int j = 0; j <= N / 2; j++) {
        double a = -2.0 * M_PI * (double)j / (double)N;
        p->tw[2 * j] = (real)cos(a);
        p->tw[2 * j + 1] = (real)sin(a);
    }
    for (int k = 0; k <= N; k++) {
        double a = -M_PI * (double)k / (double)N;
        p->tw2[2 * k] = (real)cos(a);
        p->tw2[2 * k + 1] = (real)sin(a);
    }
    int lg = 0;
    while ((1 << lg) < N) lg++;
    for (int i = 0; i < N; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        p->rev[i] = r;
    }
}

static void plan_free(fftplan *p)
{
    free(p->tw); free(p->tw2); free(p->rev); free(p->work);
    memset(p, 0, sizeof(*p));
}

/* forward complex DFT in place on data[2n] (sign -1); inverse = conj trick */
static void cfft_fwd(const fftplan *p, real *d)
{
    const int n = p->n;
    for (int i = 0; i < n; i++) {
        int r = p->rev[i];
        if (r > i) {
            real t0 = d[2 * i], t1 = d[2 * i + 1];
            d[2 * i] = d[2 * r]; d[2 * i + 1] = d[2 * r + 1];
            d[2 * r] = t0; d[2 * r + 1] = t1;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int j = 0; j < half; j++) {
                const real wr = p->tw[2 * j * step], wi = p->tw[2 * j * step + 1];
                real *a = d + 2 * (s + j), *b = d + 2 * (s + j + half);
                const real br = b[0] * wr - b[1] * wi;
                const real bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br; b[1] = a[1] - bi;
                a[0] = a[0] + br; a[1] = a[1] + bi;
            }
        }
    }
}

/* r2c of length 2N: x[2N] -> X[N+1] (interleaved), unnormalised (FFTW r2c) */
static void rfft(const fftplan *p, const real *x, real *X)
{
    const int N = p->n;
    real *z = p->work;
    memcpy(z, x, sizeof(real) * 2 * N); /* z[n] = x[2n] + i x[2n+1] */
    cfft_fwd(p, z);
    for (int k = 0; k <= N; k++) {
        const int k1 = (k == N) ? 0 : k;
        const int k2 = (N - k == N) ? 0 : N - k;
        const real zr = z[2 * k1], zi = z[2 * k1 + 1];
        const real cr = z[2 * k2], ci = -z[2 * k2 + 1];       /* conj(Z[N-k]) */
        const real er = (real)0.5 * (zr + cr), ei = (real)0.5 * (zi + ci);
        /* O = (Z - conj)/(2i) = -i/2 (Z - conj) */
        const real dr = zr - cr, di = zi - ci;
        const real orr = (real)0.5 * di, oi = (real)-0.5 * dr;
        const real wr = p->tw2[2 * k], wi = p->tw2[2 * k + 1];
        X[2 * k] = er + (orr * wr - oi * wi);
        X[2 * k + 1] = ei + (orr * wi + oi * wr);
    }
}

/* c2r of length 2N: X[N+1] -> x[2N], unnormalised (FFTW c2r; imag of X[0], X[N]
 * ignored) */
static void irfft(const fftplan *p, const real *X, real *x)
{
    const int N = p->n;
    real *z = p->work;
    for (int k = 0; k < N; k++) {
        const real ar0 = X[2 * k], ai0 = (k == 0) ? (real)0 : X[2 * k + 1];
        const real br0 = X[2 * (N - k)], bi0 = (k == 0) ? (real)0 : -X[2 * (N - k) + 1];
        /* A = X[k] + conj(X[N-k]);  B = (X[k] - conj(X[N-k])) e^{+i pi k/N} */
        const real Ar = ar0 + br0, Ai = ai0 + bi0;
        const real dr = ar0 - br0, di = ai0 - bi0;
        const real wr = p->tw2[2 * k], wi = -p->tw2[2 * k + 1];
        const real Br = dr * wr - di * wi, Bi = dr * wi + di * wr;
        /* C = A + iB, then inverse DFT via conj(DFT(conj(C))) */
        z[2 * k] = Ar - Bi;
        z[2 * k + 1] = -(Ai + Br);
    }
    cfft_fwd(p, z);
    for (int n = 0; n < N; n++) {
        x[2 * n] = z[2 * n];
        x[2 * n + 1] = -z[2 * n + 1];
    }
}

/* exported for the FFT pin (numpy rfft) */
PCO_EXPORT int pco_rfft(int N, const real *x, real *X)
{
    fftplan p; plan_init(&p, N); rfft(&p, x, X); plan_free(&p); return 0;
}
PCO_EXPORT int pco_irfft(int N, const real *X, real *x)
{
    fftplan p; plan_init(&p, N); irfft(&p, X, x); plan_free(&p); return 0;
}

/* ------------------------------------------------------------------------ */
/* DelayBuffer / DoubleBuffer (src/partconv.cpp:1401-1511)                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    real *mem;
    int bufferSize, frameSize, framesPerBlock, numBlocks;
    int readBlock, writeBlock, writeFrame, readFrameDelay;
} delaybuf;

static void db_init(delaybuf *b, int frameSize, int blockSize, int delay)
{
    b->frameSize = frameSize;
    b->framesPerBlock = blockSize;
    b->readFrameDelay = delay - blockSize - 1;
    int rem = (b->readFrameDelay + 1) % blockSize;
    int size = b->readFrameDelay + 1 + (blockSize - rem);
    b->numBlocks = size / blockSize + 1;
    b->bufferSize = b->numBlocks * blockSize;
    b->mem = (real *)calloc((size_t)b->bufferSize * frameSize, sizeof(real));
    b->readBlock = b->numBlocks - 1;
    b->writeBlock = b->readFrameDelay / blockSize;
    b->writeFrame = b->readFrameDelay % blockSize;
}

static void db_write(delaybuf *b, const real *x)
{
    int frame = b->writeBlock * b->framesPerBlock + b->writeFrame;
    memcpy(b->mem + (size_t)frame * b->frameSize, x, sizeof(real) * b->frameSize);
    if (++b->writeFrame == b->framesPerBlock) {
        b->writeFrame = 0;
        if (++b->writeBlock == b->numBlocks) b->writeBlock = 0;
    }
}

static int db_read(delaybuf *b, real *y)
{
    if (b->readBlock == b->writeBlock) {
        snprintf(g_err, sizeof g_err, "DelayBuffer: readBlock == writeBlock (%d)", b->readBlock);
        return 1; /* the reference exits here (src/partconv.cpp:1449-1455) */
    }
    int frame = b->readBlock * b->framesPerBlock;
    memcpy(y, b->mem + (size_t)frame * b->frameSize,
           sizeof(real) * b->frameSize * b->framesPerBlock);
    return 0;
}

static void db_prep(delaybuf *b)
{
    if (++b->readBlock == b->numBlocks) b->readBlock = 0;
}

static void db_reset(delaybuf *b)
{
    /* the reference zeroes only bufferSize floats (src/partconv.cpp:1461); this
     * restatement zeroes the whole ring (documented deviation: identical output
     * whenever reset() is called before any input, as every test does). */
    memset(b->mem, 0, sizeof(real) * (size_t)b->bufferSize * b->frameSize);
    b->readBlock = b->numBlocks - 1;
    b->writeBlock = b->readFrameDelay / b->framesPerBlock;
    b->writeFrame = b->readFrameDelay % b->framesPerBlock;
}

/* ------------------------------------------------------------------------ */
/* PartConvFilter                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    int M;
    int *N, *P, *delay, *T;      /* T[L] = taps before level L */
    int *nfft, *cfft;
    unsigned *bcurr;
    int *num_N0, *ind_N0;
    int start_sync_depth, end_sync_depth;
    real **H;          /* H[L]: P x cfft complex (interleaved) */
    real **Xfft;       /* FDL  : P x cfft complex */
    real **inbuf;      /* 2N */
    real **outbuf;     /* 2N */
    real **Yfft;       /* cfft complex */
    real **fftbuf;     /* cfft complex */
    real **ybuf_r, **ybuf_w; /* DoubleBuffer read / write, N floats each */
    delaybuf *Xbuf;
    fftplan *plan;
    int ir_len;
} pco_filter;

static int is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

PCO_EXPORT void pco_filter_free(pco_filter *f);

/* PartConvFilter::setup (src/partconv.cpp:70-306).  impulse==NULL => all-ones IR
 * of length sum N*P (src/partconv.cpp:132-142).  Returns NULL on error. */
PCO_EXPORT pco_filter *pco_filter_setup(int M, const int *N_in, const int *P_in,
                                        const real *impulse, int ir_len)
{
    g_err[0] = 0;
    if (M < 1) { snprintf(g_err, sizeof g_err, "M must be >= 1"); return NULL; }
    for (int i = 0; i < M; i++)
        if (!is_pow2(N_in[i])) {
            snprintf(g_err, sizeof g_err,
                     "PartConvFilter ERROR: block size of level %d must be a power of 2", i);
            return NULL;
        }
    for (int i = 1; i < M; i++)
        if (N_in[i] <= N_in[i - 1]) {
            snprintf(g_err, sizeof g_err, "PartConvFilter ERROR: N[%d]=%d <= N[%d]=%d", i,
                     N_in[i], i - 1, N_in[i - 1]);
            return NULL;
        }
    int *Pw = (int *)malloc(sizeof(int) * M);
    memcpy(Pw, P_in, sizeof(int) * M);

    real *ir = NULL;
    int len;
    if (impulse == NULL) {
        len = 0;
        for (int L = 0; L < M; L++) len += N_in[L] * Pw[L];
        ir = (real *)malloc(sizeof(real) * (len > 0 ? len : 1));
        for (int i = 0; i < len; i++) ir[i] = 1;
    } else {
        len = ir_len;
    }

    int position = 2 * N_in[0];
    int *delay = (int *)malloc(sizeof(int) * M);
    for (int i = 0; i < M - 1; i++) {
        delay[i] = position / N_in[0];
        position += N_in[i] * Pw[i];
        if (position < 2 * N_in[i + 1]) {
            snprintf(g_err, sizeof g_err,
                     "PartConvFilter ERROR: sufficient delay not allowed for block size of level %d",
                     i + 2);
            free(Pw); free(delay); free(ir);
            return NULL;
        }
    }
    delay[M - 1] = position / N_in[0];
    /* float division then ceil, as src/partconv.cpp:169 */
    Pw[M - 1] = (int)ceilf((float)(len - (position - 2 * N_in[0])) / (float)N_in[M - 1]);
    if (len - (position - 2 * N_in[0]) < 0) {
        snprintf(g_err, sizeof g_err, "PartConvFilter ERROR: last partitioning level (%d) not required", M);
        free(Pw); free(delay); free(ir);
        return NULL;
    }
    if (Pw[M - 1] == 0) {
        /* the reference would divide by zero in runLevel; reported as an error here */
        snprintf(g_err, sizeof g_err, "PartConvFilter ERROR: last partitioning level (%d) not required", M);
        free(Pw); free(delay); free(ir);
        return NULL;
    }
    position += N_in[M - 1] * Pw[M - 1];

    pco_filter *f = (pco_filter *)calloc(1, sizeof(pco_filter));
    f->M = M;
    f->N = (int *)malloc(sizeof(int) * M); memcpy(f->N, N_in, sizeof(int) * M);
    f->P = Pw;
    f->delay = delay;
    f->T = (int *)malloc(sizeof(int) * M);
    f->nfft = (int *)malloc(sizeof(int) * M);
    f->cfft = (int *)malloc(sizeof(int) * M);
    f->bcurr = (unsigned *)calloc(M, sizeof(unsigned));
    f->num_N0 = (int *)malloc(sizeof(int) * M);
    f->ind_N0 = (int *)malloc(sizeof(int) * M);
    f->ir_len = len;
    int t = 0;
    for (int L = 0; L < M; L++) {
        f->T[L] = t;
        t += f->N[L] * f->P[L];
        f->nfft[L] = 2 * f->N[L];
        f->cfft[L] = f->N[L] + 1;
        f->num_N0[L] = f->N[L] / f->N[0];
        f->ind_N0[L] = f->num_N0[L] - 1;
    }

    /* zero-padded IR of length `position` */
    real *hz = (real *)calloc((size_t)position, sizeof(real));
    memcpy(hz, impulse ? impulse : ir, sizeof(real) * (size_t)len);
    free(ir);

    f->H = (real **)calloc(M, sizeof(real *));
    f->Xfft = (real **)calloc(M, sizeof(real *));
    f->inbuf = (real **)calloc(M, sizeof(real *));
    f->outbuf = (real **)calloc(M, sizeof(real *));
    f->Yfft = (real **)calloc(M, sizeof(real *));
    f->fftbuf = (real **)calloc(M, sizeof(real *));
    f->ybuf_r = (real **)calloc(M, sizeof(real *));
    f->ybuf_w = (real **)calloc(M, sizeof(real *));
    f->Xbuf = (delaybuf *)calloc(M, sizeof(delaybuf));
    f->plan = (fftplan *)calloc(M, sizeof(fftplan));

    real *row = (real *)malloc(sizeof(real) * 2 * f->N[M - 1]);
    for (int L = 0; L < M; L++) {
        const int N = f->N[L], P = f->P[L], C = f->cfft[L];
        plan_init(&f->plan[L], N);
        f->H[L] = (real *)calloc((size_t)P * C * 2, sizeof(real));
        /* splitVector(h_level, N, 2N, N): taps of block p at offset N of a
         * zeroed 2N row (src/buffers.cpp:193-217, called src/partconv.cpp:203) */
        const int off = (g_mutation == 5) ? 0 : N;
        const real scale = (g_mutation == 3) ? (real)N : (real)(2 * N);
        for (int p = 0; p < P; p++) {
            memset(row, 0, sizeof(real) * 2 * N);
            memcpy(row + off, hz + f->T[L] + (size_t)p * N, sizeof(real) * N);
            real *Hp = f->H[L] + (size_t)p * C * 2;
            rfft(&f->plan[L], row, Hp);
            for (int j = 0; j < 2 * C; j++) Hp[j] /= scale;
        }
        f->Xfft[L] = (real *)calloc((size_t)P * C * 2, sizeof(real));
        f->inbuf[L] = (real *)calloc(2 * N, sizeof(real));
        f->outbuf[L] = (real *)calloc(2 * N, sizeof(real));
        f->Yfft[L] = (real *)calloc(2 * C, sizeof(real));
        f->fftbuf[L] = (real *)calloc(2 * C, sizeof(real));
        f->ybuf_r[L] = (real *)calloc(N, sizeof(real));
        f->ybuf_w[L] = (real *)calloc(N, sizeof(real));
        db_init(&f->Xbuf[L], f->N[0], N / f->N[0], f->delay[L]);
    }
    free(row);
    free(hz);
    return f;
}

PCO_EXPORT int pco_filter_levels(const pco_filter *f) { return f->M; }
PCO_EXPORT void pco_filter_partition(const pco_filter *f, int *N, int *P, int *delay)
{
    for (int L = 0; L < f->M; L++) {
        if (N) N[L] = f->N[L];
        if (P) P[L] = f->P[L];
        if (delay) delay[L] = f->delay[L];
    }
}
/* copy H[L][p] (cfft complex, interleaved) */
PCO_EXPORT void pco_filter_get_H(const pco_filter *f, int L, int p, real *out)
{
    memcpy(out, f->H[L] + (size_t)p * f->cfft[L] * 2, sizeof(real) * 2 * f->cfft[L]);
}

/* PartConvFilter::runLevel (src/partconv.cpp:354-490) */
PCO_EXPORT int pco_filter_run_level(pco_filter *f, int L)
{
    const int N = f->N[L], P = f->P[L], C = f->cfft[L];
    real *in = f->inbuf[L];
    memmove(in, in + N, sizeof(real) * N);
    if (db_read(&f->Xbuf[L], in + N)) return 1;

    unsigned slot = f->bcurr[L] % (unsigned)P;
    rfft(&f->plan[L], in, f->fftbuf[L]);
    memcpy(f->Xfft[L] + (size_t)slot * C * 2, f->fftbuf[L], sizeof(real) * 2 * C);

    real *Y = f->Yfft[L];
    memset(Y, 0, sizeof(real) * 2 * C);
    for (int p = 0; p < P; p++) {
        unsigned s = (f->bcurr[L] - (unsigned)p + (unsigned)P) % (unsigned)P;
        if (g_mutation == 1) s = (f->bcurr[L] - (unsigned)p + 1u + (unsigned)P) % (unsigned)P;
        const real *A = f->Xfft[L] + (size_t)s * C * 2;
        const real *B = f->H[L] + (size_t)p * C * 2;
        for (int i = 0; i < C; i++) {
            const real ar = A[2 * i], ai = A[2 * i + 1];
            const real br = B[2 * i], bi = (g_mutation == 2) ? -B[2 * i + 1] : B[2 * i + 1];
            /* (ar*br - ai*bi, ar*bi + ai*br): the SSE3 moveldup/movehdup/addsub
             * sequence of src/partconv.cpp:384-395 */
            const real re = ar * br - ai * bi;
            const real im = ar * bi + ai * br;
            Y[2 * i] += re;
            Y[2 * i + 1] += im;
        }
    }
    irfft(&f->plan[L], Y, f->outbuf[L]);
    memcpy(f->ybuf_w[L], f->outbuf[L], sizeof(real) * N);
    f->bcurr[L]++;
    return 0;
}

/* PartConvFilter::sync_levels (src/partconv.cpp:568-585) */
PCO_EXPORT void pco_filter_sync_levels(pco_filter *f)
{
    f->start_sync_depth = 0;
    f->end_sync_depth = 0;
    for (int L = 0; L < f->M; L++) {
        f->ind_N0[L]++;
        if (f->ind_N0[L] == f->num_N0[L]) {
            f->ind_N0[L] = 0;
            f->start_sync_depth = L;
        } else if (f->ind_N0[L] == f->num_N0[L] - 1)
            f->end_sync_depth = L;
    }
}

/* PartConvFilter::sync_levels_relaxed (src/partconv.cpp:587-603) */
PCO_EXPORT void pco_filter_sync_levels_relaxed(pco_filter *f)
{
    f->start_sync_depth = 0;
    f->end_sync_depth = 0;
    for (int L = 0; L < f->M; L++) {
        f->ind_N0[L]++;
        if (f->ind_N0[L] == f->num_N0[L]) {
            f->ind_N0[L] = 0;
            f->start_sync_depth = L;
            f->end_sync_depth = L;
        }
    }
}

PCO_EXPORT int pco_filter_start_depth(const pco_filter *f) { return f->start_sync_depth; }
PCO_EXPORT int pco_filter_end_depth(const pco_filter *f) { return f->end_sync_depth; }

/* PartConvFilter::input (src/partconv.cpp:605-613) */
PCO_EXPORT void pco_filter_input(pco_filter *f, const real *x)
{
    for (int L = 0; L < f->M; L++) {
        db_write(&f->Xbuf[L], x);
        if (L <= f->start_sync_depth) db_prep(&f->Xbuf[L]);
    }
}

static void swap_ptr(real **a, real **b) { real *t = *a; *a = *b; *b = t; }

/* PartConvFilter::output (src/partconv.cpp:616-631) */
PCO_EXPORT void pco_filter_output(pco_filter *f, real *y)
{
    const int N0 = f->N[0];
    memcpy(y, f->outbuf[0], sizeof(real) * N0);
    for (int L = 1; L < f->M; L++) {
        if (L <= f->end_sync_depth) swap_ptr(&f->ybuf_r[L], &f->ybuf_w[L]);
        int rb = f->ind_N0[L] + 1;
        if (g_mutation == 4) rb = f->ind_N0[L];
        if (rb == f->num_N0[L]) rb = 0;
        const real *ptr = f->ybuf_r[L] + (size_t)N0 * rb;
        for (int i = 0; i < N0; i++) y[i] += ptr[i];
    }
}

/* PartConvFilter::output_relaxed (src/partconv.cpp:633-648) */
PCO_EXPORT void pco_filter_output_relaxed(pco_filter *f, real *y)
{
    const int N0 = f->N[0];
    memcpy(y, f->outbuf[0], sizeof(real) * N0);
    for (int L = 1; L < f->M; L++) {
        if (L <= f->end_sync_depth) swap_ptr(&f->ybuf_r[L], &f->ybuf_w[L]);
        int rb = f->ind_N0[L];
        if (rb == f->num_N0[L]) rb = 0;
        const real *ptr = f->ybuf_r[L] + (size_t)N0 * rb;
        for (int i = 0; i < N0; i++) y[i] += ptr[i];
    }
}

/* PartConvFilter::reset (src/partconv.cpp:492-506) */
PCO_EXPORT void pco_filter_reset(pco_filter *f)
{
    for (int L = 0; L < f->M; L++) {
        f->bcurr[L] = 0;
        f->ind_N0[L] = f->num_N0[L] - 1;
        memset(f->inbuf[L], 0, sizeof(real) * 2 * f->N[L]);
        /* the reference leaves outbuffer[L] untouched; zeroed here so a relaxed
         * tick right after reset() does not replay pre-reset output */
        memset(f->outbuf[L], 0, sizeof(real) * 2 * f->N[L]);
        memset(f->Xfft[L], 0, sizeof(real) * 2 * (size_t)f->P[L] * f->cfft[L]);
        db_reset(&f->Xbuf[L]);
        memset(f->ybuf_r[L], 0, sizeof(real) * f->N[L]);
        memset(f->ybuf_w[L], 0, sizeof(real) * f->N[L]);
    }
}

PCO_EXPORT void pco_filter_free(pco_filter *f)
{
    if (!f) return;
    for (int L = 0; L < f->M; L++) {
        free(f->H[L]); free(f->Xfft[L]); free(f->inbuf[L]); free(f->outbuf[L]);
        free(f->Yfft[L]); free(f->fftbuf[L]); free(f->ybuf_r[L]); free(f->ybuf_w[L]);
        free(f->Xbuf[L].mem);
        plan_free(&f->plan[L]);
    }
    free(f->H); free(f->Xfft); free(f->inbuf); free(f->outbuf); free(f->Yfft);
    free(f->fftbuf); free(f->ybuf_r); free(f->ybuf_w); free(f->Xbuf); free(f->plan);
    free(f->N); free(f->P); free(f->delay); free(f->T); free(f->nfft); free(f->cfft);
    free(f->bcurr); free(f->num_N0); free(f->ind_N0);
    free(f);
}

/* Single-instance convenience: strict (relaxed=0) or relaxed tick loop over a
 * whole mono signal, in the order PartConvMulti::run uses for one instance.
 * x and y hold nframes*N0 samples. */
PCO_EXPORT int pco_filter_process(pco_filter *f, const real *x, real *y, int nframes, int relaxed)
{
    const int N0 = f->N[0];
    for (int t = 0; t < nframes; t++) {
        if (relaxed) {
            pco_filter_sync_levels_relaxed(f);
            pco_filter_output_relaxed(f, y + (size_t)t * N0);
            pco_filter_input(f, x + (size_t)t * N0);
            for (int L = 0; L <= f->start_sync_depth; L++)
                if (pco_filter_run_level(f, L)) return 1;
        } else {
            pco_filter_sync_levels(f);
            pco_filter_input(f, x + (size_t)t * N0);
            for (int L = 0; L <= f->start_sync_depth; L++)
                if (pco_filter_run_level(f, L)) return 1;
            pco_filter_output(f, y + (size_t)t * N0);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PartConvMulti (src/partconvMulti.cpp:23-244 level map, :815-928 run,      */
/* :1017-1134 relaxed run)                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    int FS, buffer_size, nIn, nOut, numPCs, relaxed;
    pco_filter **pc;
    int *inCh, *outCh, *mute;
    real *scale;
    int num_levels;
    unsigned *level_counter, *level_size;
    int *level_map; /* [numPCs][num_levels] */
    real *in_mix, *out_mix; /* numPCs x N0 */
    unsigned long frameNum;
} pco_multi;

PCO_EXPORT void pco_multi_free(pco_multi *m);

/* In-memory setup: instance i uses partition (N_flat,P_flat)[offs[i] .. offs[i]+M[i])
 * and IR irs[i] of length ir_lens[i].  The config parser of
 * src/partconvMulti.cpp:36-134 lives in the Python rules module. */
PCO_EXPORT pco_multi *pco_multi_setup(int FS, int buffer_size, int nIn, int nOut, int numPCs,
                                      const int *M, const int *N_flat, const int *P_flat,
                                      const int *inCh, const int *outCh, const int *mute,
                                      const real *scale, const real *const *irs,
                                      const int *ir_lens, int relaxed)
{
    g_err[0] = 0;
    pco_multi *m = (pco_multi *)calloc(1, sizeof(pco_multi));
    m->FS = FS; m->buffer_size = buffer_size; m->nIn = nIn; m->nOut = nOut;
    m->numPCs = numPCs; m->relaxed = relaxed;
    m->pc = (pco_filter **)calloc(numPCs, sizeof(pco_filter *));
    m->inCh = (int *)malloc(sizeof(int) * numPCs);
    m->outCh = (int *)malloc(sizeof(int) * numPCs);
    m->mute = (int *)malloc(sizeof(int) * numPCs);
    m->scale = (real *)malloc(sizeof(real) * numPCs);
    m->in_mix = (real *)calloc((size_t)numPCs * buffer_size, sizeof(real));
    m->out_mix = (real *)calloc((size_t)numPCs * buffer_size, sizeof(real));
    int max_block = 0, off = 0;
    for (int i = 0; i < numPCs; i++) {
        m->inCh[i] = inCh[i]; m->outCh[i] = outCh[i]; m->mute[i] = mute[i];
        m->scale[i] = scale[i];
        if (inCh[i] < 0 || inCh[i] >= nIn || outCh[i] < 0 || outCh[i] >= nOut) {
            snprintf(g_err, sizeof g_err, "instance %d: channel out of range", i);
            pco_multi_free(m);
            return NULL;
        }
        if (N_flat[off] != buffer_size) {
            snprintf(g_err, sizeof g_err, "line %d: first block size must equal buffer_size", i + 1);
            pco_multi_free(m);
            return NULL;
        }
        for (int j = 0; j < M[i]; j++)
            if (N_flat[off + j] > max_block) max_block = N_flat[off + j];
        m->pc[i] = pco_filter_setup(M[i], N_flat + off, P_flat + off, irs[i], ir_lens[i]);
        if (!m->pc[i]) { pco_multi_free(m); return NULL; }
        off += M[i];
    }
    /* global level set: block sizes buffer_size*2^i used by any instance */
    int lg = 0;
    while ((buffer_size << lg) < max_block) lg++;
    const int temp_levels = lg + 1;
    int *used = (int *)calloc(temp_levels, sizeof(int));
    for (int i = 0; i < numPCs; i++)
        for (int L = 0; L < m->pc[i]->M; L++)
            for (int g = 0; g < temp_levels; g++)
                if (m->pc[i]->N[L] == (buffer_size << g)) used[g] = 1;
    m->num_levels = 0;
    for (int g = 0; g < temp_levels; g++) m->num_levels += used[g];
    m->level_counter = (unsigned *)malloc(sizeof(unsigned) * m->num_levels);
    m->level_size = (unsigned *)malloc(sizeof(unsigned) * m->num_levels);
    m->level_map = (int *)malloc(sizeof(int) * numPCs * m->num_levels);
    int l = 0;
    for (int g = 0; g < temp_levels; g++) {
        if (!used[g]) continue;
        m->level_size[l] = 1u << g;
        m->level_counter[l] = m->level_size[l] - 1;
        for (int i = 0; i < numPCs; i++) {
            m->level_map[i * m->num_levels + l] = -1;
            for (int L = 0; L < m->pc[i]->M; L++)
                if (m->pc[i]->N[L] == (buffer_size << g)) m->level_map[i * m->num_levels + l] = L;
        }
        l++;
    }
    free(used);
    return m;
}

PCO_EXPORT int pco_multi_num_levels(const pco_multi *m) { return m->num_levels; }
PCO_EXPORT void pco_multi_level_map(const pco_multi *m, int *out_map, unsigned *out_size)
{
    memcpy(out_map, m->level_map, sizeof(int) * m->numPCs * m->num_levels);
    memcpy(out_size, m->level_size, sizeof(unsigned) * m->num_levels);
}

static int run_levels_upto(pco_multi *m, int depth)
{
    for (int l = 0; l <= depth; l++)
        for (int i = 0; i < m->numPCs; i++) {
            int L = m->level_map[i * m->num_levels + l];
            if (L >= 0 && pco_filter_run_level(m->pc[i], L)) return 1;
        }
    return 0;
}

static void mix_output(pco_multi *m, real *out, int frameCount)
{
    memset(out, 0, sizeof(real) * (size_t)m->nOut * frameCount);
    for (int i = 0; i < m->numPCs; i++) {
        if (m->mute[i]) continue;
        const int c = m->outCh[i];
        const real *y = m->out_mix + (size_t)i * m->buffer_size;
        for (int j = 0; j < frameCount; j++) out[c + (size_t)j * m->nOut] += y[j];
    }
}

/* PartConvMulti::run / PartConvMultiRelaxed::run for one tick */
PCO_EXPORT int pco_multi_run(pco_multi *m, real *out, const real *in, int frameCount)
{
    if (frameCount != m->buffer_size) {
        snprintf(g_err, sizeof g_err, "frameCount (%d) != buffer_size (%d)", frameCount, m->buffer_size);
        return 1;
    }
    int start_depth = 0;
    for (int l = 0; l < m->num_levels; l++) {
        m->level_counter[l]++;
        if (m->level_counter[l] == m->level_size[l]) {
            m->level_counter[l] = 0;
            start_depth = l;
        }
    }
    for (int i = 0; i < m->numPCs; i++) {
        const int c = m->inCh[i];
        const real s = m->scale[i];
        real *x = m->in_mix + (size_t)i * m->buffer_size;
        for (int j = 0; j < frameCount; j++) x[j] = in[c + (size_t)j * m->nIn] * s;
    }
    if (!m->relaxed) {
        for (int i = 0; i < m->numPCs; i++) pco_filter_sync_levels(m->pc[i]);
        for (int i = 0; i < m->numPCs; i++)
            pco_filter_input(m->pc[i], m->in_mix + (size_t)i * m->buffer_size);
        if (run_levels_upto(m, start_depth)) return 1;
        for (int i = 0; i < m->numPCs; i++)
            pco_filter_output(m->pc[i], m->out_mix + (size_t)i * m->buffer_size);
    } else {
        for (int i = 0; i < m->numPCs; i++) pco_filter_sync_levels_relaxed(m->pc[i]);
        for (int i = 0; i < m->numPCs; i++)
            pco_filter_output_relaxed(m->pc[i], m->out_mix + (size_t)i * m->buffer_size);
        for (int i = 0; i < m->numPCs; i++)
            pco_filter_input(m->pc[i], m->in_mix + (size_t)i * m->buffer_size);
        if (run_levels_upto(m, start_depth)) return 1;
    }
    mix_output(m, out, frameCount);
    m->frameNum++;
    return 0;
}

/* run nframes ticks over interleaved in[nframes*N0*nIn] -> out[nframes*N0*nOut] */
PCO_EXPORT int pco_multi_process(pco_multi *m, real *out, const real *in, int nframes)
{
    const int N0 = m->buffer_size;
    for (int t = 0; t < nframes; t++)
        if (pco_multi_run(m, out + (size_t)t * N0 * m->nOut, in + (size_t)t * N0 * m->nIn, N0))
            return 1;
    return 0;
}

PCO_EXPORT void pco_multi_reset(pco_multi *m)
{
    for (int i = 0; i < m->numPCs; i++) pco_filter_reset(m->pc[i]);
    for (int l = 0; l < m->num_levels; l++) m->level_counter[l] = m->level_size[l] - 1;
    m->frameNum = 0;
}

PCO_EXPORT void pco_multi_free(pco_multi *m)
{
    if (!m) return;
    if (m->pc)
        for (int i = 0; i < m->numPCs; i++) pco_filter_free(m->pc[i]);
    free(m->pc); free(m->inCh); free(m->outCh); free(m->mute); free(m->scale);
    free(m->in_mix); free(m->out_mix); free(m->level_counter); free(m->level_size);
    free(m->level_map);
    free(m);
}

/* Wall-clock seconds (CLOCK_MONOTONIC), used by the CPU-baseline timing leg. */
PCO_EXPORT double pco_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Threaded real-time runner: the reference's scheduling, for the CPU        */
/* baseline of the benchmark (timing only; its output is the same as the     */
/* single-threaded restatement's, which a test checks bitwise).              */
/*   PartConvMulti::run (src/partconvMulti.cpp:815-928): per tick the frame  */
/*   goes into every instance (sync_levels, input), the levels whose period  */
/*   starts now are handed to their workers, level 0 of every instance runs  */
/*   on the calling thread, the call waits for the levels whose block is     */
/*   read for the first time now, then every instance's output is formed.    */
/*   workerThreadEntry (src/partconvMulti.cpp:930-1013): one group of worker */
/*   threads per global level (a block size used by any instance, mapped to  */
/*   each instance's own level as level_map does, :139-244), each thread     */
/*   running runLevel for its share of the instances when the level starts,  */
/*   then signalling completion.                                             */
/* ------------------------------------------------------------------------ */
#include <pthread.h>

#define RT_MAX_G 24

typedef struct rt_group {
    int k;               /* period in ticks (block size / N0) */
    int nth;
    int *inst, *lvl;     /* the instances with this block size and their local level */
    int nj;
    pthread_t *th;
    pthread_mutex_t mu;
    pthread_cond_t go_cv, done_cv;
    unsigned long gen;   /* starts issued */
    int pending;         /* threads still running the current start */
    int stop;
    double busy_s;       /* summed runLevel time of all its threads */
    long long blocks;    /* runLevel calls */
} rt_group;

typedef struct rt_arg {
    pco_filter **f;
    rt_group *g;
    int t;
} rt_arg;

static void *rt_worker(void *p)
{
    rt_arg *a = (rt_arg *)p;
    rt_group *g = a->g;
    unsigned long seen = 0;
    for (;;) {
        pthread_mutex_lock(&g->mu);
        while (g->gen == seen && !g->stop) pthread_cond_wait(&g->go_cv, &g->mu);
        if (g->stop) {
            pthread_mutex_unlock(&g->mu);
            return NULL;
        }
        seen = g->gen;
        pthread_mutex_unlock(&g->mu);
        const double t0 = pco_now();
        long long nb = 0;
        for (int j = a->t; j < g->nj; j += g->nth) {
            pco_filter_run_level(a->f[g->inst[j]], g->lvl[j]);
            nb++;
        }
        const double dt = pco_now() - t0;
        pthread_mutex_lock(&g->mu);
        g->busy_s += dt;
        g->blocks += nb;
        if (--g->pending == 0) pthread_cond_broadcast(&g->done_cv);
        pthread_mutex_unlock(&g->mu);
    }
}

/* Run nticks ticks of n strict instances (one buffer size N0).  Instance i
 * reads column in_ch[i] of x [nticks * N0][nIn] (in_ch may be NULL: column i)
 * and writes column i of y [nticks * N0][n].  Every global level above the
 * buffer size gets threads_per_level worker threads.  lat_s[t]: seconds from
 * the start of tick t to its output being formed (ticks back to back).
 * gN[g], level_s[g] (room for RT_MAX_G): block size of global level g and the
 * mean seconds of one of its runLevel calls (one instance's block); returns
 * the number of global levels, or -1 on error. */
PCO_EXPORT int pco_rt_run(pco_filter **f, int n, const int *in_ch, int nIn, const real *x, real *y, int nticks,
                          int threads_per_level, double *lat_s, int *gN, double *level_s)
{
    if (n < 1 || nticks < 0 || threads_per_level < 1 || nIn < 1) return -1;
    const int N0 = f[0]->N[0];
    int sizes[RT_MAX_G], G = 0;
    for (int i = 0; i < n; i++) {
        if (f[i]->N[0] != N0 || (in_ch && (in_ch[i] < 0 || in_ch[i] >= nIn)) || (!in_ch && i >= nIn)) return -1;
        for (int L = 1; L < f[i]->M; L++) {
            int q = 0;
            while (q < G && sizes[q] != f[i]->N[L]) q++;
            if (q == G) {
                if (G == RT_MAX_G) return -1;
                sizes[G++] = f[i]->N[L];
            }
        }
    }
    for (int a = 1; a < G; a++)  /* ascending block size */
        for (int b = a; b > 0 && sizes[b] < sizes[b - 1]; b--) {
            const int t = sizes[b];
            sizes[b] = sizes[b - 1];
            sizes[b - 1] = t;
        }
    rt_group *grp = (rt_group *)calloc((size_t)(G > 0 ? G : 1), sizeof(rt_group));
    rt_arg *args = (rt_arg *)calloc((size_t)(G > 0 ? G : 1) * threads_per_level, sizeof(rt_arg));
    for (int g = 0; g < G; g++) {
        rt_group *q = &grp[g];
        q->k = sizes[g] / N0;
        q->inst = (int *)malloc(sizeof(int) * n);
        q->lvl = (int *)malloc(sizeof(int) * n);
        for (int i = 0; i < n; i++)
            for (int L = 1; L < f[i]->M; L++)
                if (f[i]->N[L] == sizes[g]) {
                    q->inst[q->nj] = i;
                    q->lvl[q->nj] = L;
                    q->nj++;
                }
        q->nth = threads_per_level < q->nj ? threads_per_level : q->nj;
        q->th = (pthread_t *)calloc((size_t)q->nth, sizeof(pthread_t));
        pthread_mutex_init(&q->mu, NULL);
        pthread_cond_init(&q->go_cv, NULL);
        pthread_cond_init(&q->done_cv, NULL);
        for (int t = 0; t < q->nth; t++) {
            rt_arg *a = &args[(size_t)g * threads_per_level + t];
            a->f = f;
            a->g = q;
            a->t = t;
            pthread_create(&q->th[t], NULL, rt_worker, a);
        }
    }
    real *fr = (real *)malloc(sizeof(real) * N0);
    double busy0 = 0.0;
    for (int t = 0; t < nticks; t++) {
        const double t0 = pco_now();
        const real *xt = x + (size_t)t * N0 * nIn;
        real *yt = y + (size_t)t * N0 * n;
        for (int i = 0; i < n; i++) {
            const int c = in_ch ? in_ch[i] : i;
            pco_filter_sync_levels(f[i]);
            for (int j = 0; j < N0; j++) fr[j] = xt[(size_t)j * nIn + c];
            pco_filter_input(f[i], fr);
        }
        /* each instance's counters advance one tick at a time from the same
         * reset state, so level starts and ends follow from the tick number:
         * a level of period k starts at tick f = 0 mod k and its block is
         * first read at tick f + 1 = 0 mod k */
        const long long tick = (long long)f[0]->bcurr[0];
        for (int g = 0; g < G; g++) {
            rt_group *q = &grp[g];
            if (tick % q->k != 0) continue;
            pthread_mutex_lock(&q->mu);
            q->pending = q->nth;
            q->gen++;
            pthread_cond_broadcast(&q->go_cv);
            pthread_mutex_unlock(&q->mu);
        }
        const double tb = pco_now();
        for (int i = 0; i < n; i++) pco_filter_run_level(f[i], 0);
        busy0 += pco_now() - tb;
        for (int g = 0; g < G; g++) {
            rt_group *q = &grp[g];
            if ((tick + 1) % q->k != 0) continue;
            pthread_mutex_lock(&q->mu);
            while (q->pending > 0) pthread_cond_wait(&q->done_cv, &q->mu);
            pthread_mutex_unlock(&q->mu);
        }
        for (int i = 0; i < n; i++) {
            pco_filter_output(f[i], fr);
            for (int j = 0; j < N0; j++) yt[(size_t)j * n + i] = fr[j];
        }
        if (lat_s) lat_s[t] = pco_now() - t0;
    }
    /* every started block finished before the workers stop */
    for (int g = 0; g < G; g++) {
        rt_group *q = &grp[g];
        pthread_mutex_lock(&q->mu);
        while (q->pending > 0) pthread_cond_wait(&q->done_cv, &q->mu);
        q->stop = 1;
        pthread_cond_broadcast(&q->go_cv);
        pthread_mutex_unlock(&q->mu);
        for (int t = 0; t < q->nth; t++) pthread_join(q->th[t], NULL);
        if (gN) gN[g + 1] = sizes[g];
        if (level_s) level_s[g + 1] = q->blocks ? q->busy_s / (double)q->blocks : 0.0;
        pthread_mutex_destroy(&q->mu);
        pthread_cond_destroy(&q->go_cv);
        pthread_cond_destroy(&q->done_cv);
        free(q->th);
        free(q->inst);
        free(q->lvl);
    }
    if (gN) gN[0] = N0;
    if (level_s) level_s[0] = nticks ? busy0 / ((double)nticks * n) : 0.0;
    free(fr);
    free(args);
    free(grp);
    return G + 1;
}
