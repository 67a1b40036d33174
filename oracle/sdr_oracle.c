/*
 * sdr_oracle.c - CPU oracle for the PocketSDR search / correlate hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  A plain-C
 * restatement of the reference algorithms; each function cites the reference
 * file:line it follows.  Built with the reference's own SIMD flags
 * (-O3 -mavx2 -mfma, lib/build/libsdr.mk:62-73) but with -ffp-contract=off:
 * every place where the reference build fuses a multiply-add is written out
 * here as an explicit fmaf() so the arithmetic is fixed by the source, not by
 * the compiler.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <dlfcn.h>
#include "sdr_oracle.h"

#define PI_D 3.1415926535897932          /* src/pocket_sdr.h:63 */
#define CSCALE (1 / 11.2f)                /* SDR_CSCALE, src/pocket_sdr.h:61 */
#define NTBL 256                          /* src/sdr_func.c:53 */
#define ROUND(x) floor((x) + 0.5)         /* src/sdr_func.c:66 */

/* cpx8 nibble unpack, src/pocket_sdr.h:107-108 */
static inline int cpx8_I(uint8_t x) { return (int8_t)(uint8_t)(x << 4) >> 4; }
static inline int cpx8_Q(uint8_t x) { return (int8_t)x >> 4; }

/* ---------------------------------------------------------------------------
 * carrier-mixed LUT: src/sdr_func.c:121-131
 * mix_tbl[(j << 8) + i] = cpx8(j) * carrier(i), carrier = round(11.2 e^{-2 pi i/256})
 * products stored as int8 (two's complement wrap, as the reference's int8_t store).
 */
static int8_t mix_tbl[65536 * 2];
static int mix_init = 0;

void orc_func_init(void)
{
    for (int i = 0; i < NTBL; i++) {
        int8_t cI = (int8_t)ROUND(cos(-2.0 * PI_D * i / NTBL) / CSCALE);
        int8_t cQ = (int8_t)ROUND(sin(-2.0 * PI_D * i / NTBL) / CSCALE);
        for (int j = 0; j < 256; j++) {
            int I = cpx8_I((uint8_t)j), Q = cpx8_Q((uint8_t)j);
            mix_tbl[2 * ((j << 8) + i)]     = (int8_t)(I * cI - Q * cQ);
            mix_tbl[2 * ((j << 8) + i) + 1] = (int8_t)(I * cQ + Q * cI);
        }
    }
    mix_init = 1;
}

const int8_t *orc_mix_tbl(void)
{
    if (!mix_init) orc_func_init();
    return mix_tbl;
}

/* ---------------------------------------------------------------------------
 * mix_carr: src/sdr_func.c:772-807; split at the ring-buffer end: 824-836.
 * p = (u32)((phi - floor(phi)) 2^32), s = (u32)(int)(fc/fs 2^32),
 * IQ[i] = mix_tbl[(data[i] << 8) + ((p + i s) >> 24)] (mod-2^32 phase).
 */
static void mix_seg(const uint8_t *data, int N, double phi, double step,
                    int8_t *IQ)
{
    double scale = (double)(1 << 24) * NTBL;
    uint32_t p = (uint32_t)((phi - floor(phi)) * scale);
    uint32_t s = (uint32_t)(int)(step * scale);
    for (int i = 0; i < N; i++, p += s) {
        int idx = ((int)data[i] << 8) + (int)(p >> 24);
        IQ[2 * i]     = mix_tbl[2 * idx];
        IQ[2 * i + 1] = mix_tbl[2 * idx + 1];
    }
}

void orc_mix_carr(const uint8_t *buff, int buffN, int ix, int N, double fs,
                  double fc, double phi, int8_t *IQ)
{
    double step = fc / fs;
    if (!mix_init) orc_func_init();
    if (ix + N <= buffN) {
        mix_seg(buff + ix, N, phi, step, IQ);
    } else {
        int n = buffN - ix;
        mix_seg(buff + ix, n, phi, step, IQ);
        /* the reference build fuses phi + step * n (gcc -mfma) */
        mix_seg(buff, N - n, fma((double)n, step, phi), step, IQ + 2 * n);
    }
}

/* ---------------------------------------------------------------------------
 * dot products: src/sdr_func.c:861-912 (dot_IQ_code) and 915-956
 * (dot_IQ_cpx_code), AVX2 path.  The first 16*floor(N/16) samples are summed
 * in int32 with _mm256_sign_epi8 semantics (x * sign(c), with -(-128) kept as
 * -128), the int32 total is converted to float, then the remaining samples
 * are added one by one in float with a true int product, then scaled by
 * (1.0f * SDR_CSCALE).
 */
static inline int sgn8(int x, int c)
{
    if (c < 0) return (int8_t)(uint8_t)(-x);
    return c == 0 ? 0 : x;
}

static void dot_IQ_code(const int8_t *IQ, const int8_t *code, int N, float s,
                        float *c)
{
    int V = N >= 16 ? N / 16 * 16 : 0;
    int32_t sI = 0, sQ = 0;
    for (int i = 0; i < V; i++) {
        sI += sgn8(IQ[2 * i], code[2 * i]);
        sQ += sgn8(IQ[2 * i + 1], code[2 * i + 1]);
    }
    float cI = (float)sI, cQ = (float)sQ;
    for (int i = V; i < N; i++) {
        cI += (float)(IQ[2 * i] * code[2 * i]);
        cQ += (float)(IQ[2 * i + 1] * code[2 * i + 1]);
    }
    c[0] = cI * (s * CSCALE);
    c[1] = cQ * (s * CSCALE);
}

static void dot_IQ_cpx_code(const int8_t *IQ, const int8_t *code, int N,
                            float s, float *c)
{
    int V = N >= 16 ? N / 16 * 16 : 0;
    int32_t sR = 0, sI = 0;
    for (int i = 0; i < V; i++) {
        int I = IQ[2 * i], Q = IQ[2 * i + 1];
        int cI = code[2 * i], cQ = code[2 * i + 1];
        int ncQ = (int8_t)(uint8_t)(-cQ);
        sR += sgn8(I, cI) + sgn8(Q, cQ);
        sI += sgn8(Q, cI) + sgn8(I, ncQ);
    }
    float cr = (float)sR, ci = (float)sI;
    for (int i = V; i < N; i++) {
        int I = IQ[2 * i], Q = IQ[2 * i + 1];
        int cI = code[2 * i], cQ = code[2 * i + 1];
        cr += (float)(I * cI + Q * cQ);
        ci += (float)(Q * cI - I * cQ);
    }
    c[0] = cr * (s * CSCALE);
    c[1] = ci * (s * CSCALE);
}

/* ---------------------------------------------------------------------------
 * standard correlator: src/sdr_func.c:981-1008 (real code), 1029-1057
 * (complex code).  Tap i: p = coff + pos[i], j = floor(p),
 * k = (int)((p - j) * 10), j wrapped once into [0, N);
 * corr1 = dot(IQ[0:j], bank_k[N-j:N]), corr2 = dot(IQ[j:N], bank_k[0:N-j]).
 * Polarity: the reference build contracts the EPL sum to
 * d += fmaf(c1.re, c2.re, c1.im * c2.im).  The polarity sum always runs over
 * taps 0-2 (src/sdr_func.c:994-997): with n < 3 the reference reads
 * uninitialised corr1 / corr2 entries, so its result is undefined there; the
 * restatement (and the GPU) take the missing taps as zero.
 */
typedef void (*dot_fn)(const int8_t *, const int8_t *, int, float, float *);

static void corr_std_gen(dot_fn dot, const int8_t *IQ, const int8_t *code,
                         int N, double coff, const double *pos, int n,
                         float *corr, float *C)
{
    float corr1[ORC_MAX_CORR][2] = {{0.0f}}, corr2[ORC_MAX_CORR][2] = {{0.0f}};
    float dot_EPL = 0.0f, sign;

    for (int i = 0; i < n; i++) {
        double p = coff + pos[i];
        int j = (int)floor(p), k = (int)((p - j) * ORC_N_CODES);
        if (j < 0) j += N; else if (j >= N) j -= N;
        dot(IQ, code + 2 * ((size_t)k * N + N - j), j, 1.0f, corr1[i]);
        dot(IQ + 2 * j, code + 2 * (size_t)k * N, N - j, 1.0f, corr2[i]);
    }
    for (int i = 0; i < 3; i++) {
        dot_EPL += fmaf(corr1[i][0], corr2[i][0], corr1[i][1] * corr2[i][1]);
    }
    sign = dot_EPL < 0.0f ? -1.0f : 1.0f;
    for (int i = 0; i < n; i++) {
        corr[2 * i]     = (corr1[i][0] + sign * corr2[i][0]) / (float)N;
        corr[2 * i + 1] = (corr1[i][1] + sign * corr2[i][1]) / (float)N;
    }
    C[0] = corr1[0][0]; C[1] = corr1[0][1];
    C[2] = corr2[0][0]; C[3] = corr2[0][1];
}

void orc_corr_std(const int8_t *IQ, const int8_t *code, int N, double coff,
                  const double *pos, int n, float *corr, float *C)
{
    corr_std_gen(dot_IQ_code, IQ, code, N, coff, pos, n, corr, C);
}

void orc_corr_std_cpx_code(const int8_t *IQ, const int8_t *code, int N,
                           double coff, const double *pos, int n, float *corr,
                           float *C)
{
    corr_std_gen(dot_IQ_cpx_code, IQ, code, N, coff, pos, n, corr, C);
}

/* ---------------------------------------------------------------------------
 * complex multiply c = a b s: src/sdr_func.c:197-233.  AVX2 lanes:
 * re = fl(ar br) + fl(-ai bi), im = fl(ar bi) + fl(ai br), then * s.
 * The scalar tail (N % 4 elements) is contracted by the reference build:
 * re = fmaf(ar, br, -(ai bi)), im = fmaf(ar, bi, ai br).
 */
void orc_cpx_mul(const float *a, const float *b, int N, float s, float *c)
{
    int i = 0, V = N >= 4 ? N / 4 * 4 : 0;
    for (; i < V; i++) {
        float ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        float p0 = ar * br, p1 = ai * -bi, p2 = ar * bi, p3 = ai * br;
        c[2 * i] = (p0 + p1) * s;
        c[2 * i + 1] = (p2 + p3) * s;
    }
    for (; i < N; i++) {
        float ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        c[2 * i] = fmaf(ar, br, -(ai * bi)) * s;
        c[2 * i + 1] = fmaf(ar, bi, ai * br) * s;
    }
}

/* ---------------------------------------------------------------------------
 * FFTs.  FP64: Stockham autosort, one pass per prime factor, generic radix-r
 * DFT inside each pass; twiddles from cos/sin of the exact angle.  Used for
 * the FP64 oracle and, rounded to float, as the FP32 FFT when the reference
 * PocketFFT build is not loaded.
 */
static void fft_f64_impl(const double *in, int N, int sign, double *out)
{
    if (N <= 0) return;
    double *x = (double *)malloc(sizeof(double) * 2 * N);
    double *y = (double *)malloc(sizeof(double) * 2 * N);
    memcpy(x, in, sizeof(double) * 2 * N);
    int rem = N, Ns = 1;
    while (rem > 1) {
        int r = 2;
        while (rem % r) r++;
        double a[2 * 64], b[2 * 64], wt[2 * 64];
        double *ar = r <= 64 ? a : (double *)malloc(sizeof(double) * 2 * r);
        double *br = r <= 64 ? b : (double *)malloc(sizeof(double) * 2 * r);
        /* w_r^t, t < r: the r-point DFT's coefficients (index q m mod r
         * taken in 64 bits: q m overflows an int for r > 46340) */
        double *wr = r <= 64 ? wt : (double *)malloc(sizeof(double) * 2 * r);
        for (int t = 0; t < r; t++) {
            double ang = sign * 2.0 * PI_D * (double)t / r;
            wr[2 * t] = cos(ang);
            wr[2 * t + 1] = sin(ang);
        }
        for (int j = 0; j < N / r; j++) {
            int k = j % Ns;
            for (int q = 0; q < r; q++) {
                double xr = x[2 * (j + q * (N / r))], xi = x[2 * (j + q * (N / r)) + 1];
                double ang = sign * 2.0 * PI_D * (double)q * k / ((double)Ns * r);
                double wr = cos(ang), wi = sin(ang);
                ar[2 * q] = xr * wr - xi * wi;
                ar[2 * q + 1] = xr * wi + xi * wr;
            }
            for (int m = 0; m < r; m++) {
                double sr = 0.0, si = 0.0;
                long long t = 0;            /* q m mod r, stepped by m */
                for (int q = 0; q < r; q++) {
                    double cr = wr[2 * t], ci = wr[2 * t + 1];
                    sr += ar[2 * q] * cr - ar[2 * q + 1] * ci;
                    si += ar[2 * q] * ci + ar[2 * q + 1] * cr;
                    t += m;
                    if (t >= r) t -= r;
                }
                br[2 * m] = sr; br[2 * m + 1] = si;
            }
            int o = (j / Ns) * Ns * r + k;
            for (int q = 0; q < r; q++) {
                y[2 * (o + q * Ns)] = br[2 * q];
                y[2 * (o + q * Ns) + 1] = br[2 * q + 1];
            }
        }
        if (r > 64) { free(ar); free(br); free(wr); }
        double *t = x; x = y; y = t;
        Ns *= r;
        rem /= r;
    }
    memcpy(out, x, sizeof(double) * 2 * N);
    free(x);
    free(y);
}

void orc_fft_f64(const double *in, int N, int dir, double *out)
{
    fft_f64_impl(in, N, dir == 0 ? -1 : 1, out);
}

/* reference PocketFFT (lib/pocketfft/pocketfft_wrap.cpp:35-55), loaded from
 * oracle/_ref when it has been built from the reference sources */
typedef void *(*plan_fn)(int, float (*)[2], float (*)[2], int, unsigned);
typedef void (*exec_fn)(void *, float (*)[2], float (*)[2]);
typedef void (*destroy_fn)(void *);
static plan_fn ref_plan = NULL;
static exec_fn ref_exec = NULL;
static destroy_fn ref_destroy = NULL;

int orc_load_ref_fft(const char *so_path)
{
    void *h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    ref_plan = (plan_fn)dlsym(h, "fftwf_plan_dft_1d");
    ref_exec = (exec_fn)dlsym(h, "fftwf_execute_dft");
    ref_destroy = (destroy_fn)dlsym(h, "fftwf_destroy_plan");
    if (!ref_plan || !ref_exec || !ref_destroy) {
        ref_plan = NULL; ref_exec = NULL; ref_destroy = NULL;
        return 0;
    }
    return 1;
}

int orc_ref_fft_loaded(void) { return ref_plan != NULL; }

int orc_fft_f32(const float *in, int N, int dir, float *out)
{
    if (dir != 0 && dir != 1) return 0;
    if (ref_plan) {
        /* FFTW_FORWARD = -1, FFTW_BACKWARD = +1 (lib/pocketfft/pocketfft.h:14-15) */
        float (*tmp)[2] = (float (*)[2])malloc(sizeof(float) * 2 * N);
        memcpy(tmp, in, sizeof(float) * 2 * N);
        void *p = ref_plan(N, tmp, (float (*)[2])out, dir == 0 ? -1 : 1, 0);
        ref_exec(p, tmp, (float (*)[2])out);
        ref_destroy(p);
        free(tmp);
        return 1;
    }
    double *x = (double *)malloc(sizeof(double) * 2 * N);
    double *y = (double *)malloc(sizeof(double) * 2 * N);
    for (int i = 0; i < 2 * N; i++) x[i] = in[i];
    fft_f64_impl(x, N, dir == 0 ? -1 : 1, y);
    for (int i = 0; i < 2 * N; i++) out[i] = (float)y[i];
    free(x);
    free(y);
    return 1;
}

/* ---------------------------------------------------------------------------
 * FFT correlator: src/sdr_func.c:1088-1102
 * corr = IFFT(FFT(IQ * CSCALE) * code_fft * (1/N/N)), unnormalised FFTs.
 */
void orc_corr_fft(const int8_t *IQ, const float *code_fft, int N, float *corr)
{
    float *cpx = (float *)malloc(sizeof(float) * 4 * N);
    for (int i = 0; i < N; i++) {
        cpx[2 * i] = IQ[2 * i] * CSCALE;
        cpx[2 * i + 1] = IQ[2 * i + 1] * CSCALE;
    }
    orc_fft_f32(cpx, N, 0, cpx + 2 * N);
    orc_cpx_mul(cpx + 2 * N, code_fft, N, 1.0f / N / N, cpx);
    orc_fft_f32(cpx, N, 1, corr);
    free(cpx);
}

/* ---------------------------------------------------------------------------
 * parallel code search: src/sdr_func.c:664-687.  For each Doppler bin:
 * mix at fi + fds[i] with phi = 0, FFT-correlate, P[i N + j] += |C[j]|^2 for
 * all N lags.  The reference build contracts the power to
 * P + fmaf(C0, C0, C1 * C1).  The 1 ms sleep every 22 bins is omitted.
 */
void orc_search_code(const float *code_fft, double T, const uint8_t *buff,
                     int buffN, int ix, int N, double fs, double fi,
                     const float *fds, int len_fds, float *P)
{
    (void)T;
    float *C = (float *)malloc(sizeof(float) * 2 * N);
    int8_t *data = (int8_t *)malloc(2 * (size_t)N);
    for (int i = 0; i < len_fds; i++) {
        orc_mix_carr(buff, buffN, ix, N, fs, fi + fds[i], 0.0, data);
        orc_corr_fft(data, code_fft, N, C);
        for (int j = 0; j < N; j++) {
            P[(size_t)i * N + j] += fmaf(C[2 * j], C[2 * j], C[2 * j + 1] * C[2 * j + 1]);
        }
    }
    free(C);
    free(data);
}

void orc_search_code_f64(const double *code_fft, double T, const uint8_t *buff,
                         int buffN, int ix, int N, double fs, double fi,
                         const float *fds, int len_fds, double *P)
{
    (void)T;
    double *x = (double *)malloc(sizeof(double) * 2 * N);
    double *X = (double *)malloc(sizeof(double) * 2 * N);
    int8_t *data = (int8_t *)malloc(2 * (size_t)N);
    double s = 1.0 / ((double)N * N);
    for (int i = 0; i < len_fds; i++) {
        orc_mix_carr(buff, buffN, ix, N, fs, fi + fds[i], 0.0, data);
        for (int j = 0; j < 2 * N; j++) x[j] = data[j] * (double)CSCALE;
        fft_f64_impl(x, N, -1, X);
        for (int j = 0; j < N; j++) {
            double ar = X[2 * j], ai = X[2 * j + 1];
            double br = code_fft[2 * j], bi = code_fft[2 * j + 1];
            x[2 * j] = (ar * br - ai * bi) * s;
            x[2 * j + 1] = (ar * bi + ai * br) * s;
        }
        fft_f64_impl(x, N, 1, X);
        for (int j = 0; j < N; j++) {
            P[(size_t)i * N + j] += X[2 * j] * X[2 * j] + X[2 * j + 1] * X[2 * j + 1];
        }
    }
    free(x);
    free(X);
    free(data);
}

/* ---------------------------------------------------------------------------
 * peak search and C/N0: src/sdr_func.c:690-706 (strict '>' so the first
 * maximum in row-major order wins; running mean in double)
 */
float orc_corr_max(const float *P, int N, int Nmax, int M, double T, int *ix)
{
    double P_max = 0.0, P_ave = 0.0;
    int n = 0;
    for (int i = 0; i < M; i++) {
        for (int j = 0; j < Nmax; j++) {
            P_ave += (P[(size_t)i * N + j] - P_ave) / ++n;
            if (P[(size_t)i * N + j] <= P_max) continue;
            P_max = P[(size_t)i * N + j];
            ix[0] = i;
            ix[1] = j;
        }
    }
    return (P_ave > 0.0) ?
        (float)(10.0 * log10((P_max - P_ave) / P_ave / T)) : 0.0f;
}

/* fine Doppler: src/sdr_func.c:709-744.  The reference fits a quadratic
 * through the three bins around the peak by least squares (RTKLIB lsq); with
 * three points the fit interpolates, so the vertex is solved here from the
 * 3x3 Vandermonde system by Cramer's rule. */
double orc_fine_dop(const float *P, int N, const float *fds, int len_fds,
                    const int *ix)
{
    if (ix[0] == 0 || ix[0] == len_fds - 1) return fds[ix[0]];
    double x[3], y[3];
    for (int i = 0; i < 3; i++) {
        x[i] = fds[ix[0] - 1 + i];
        y[i] = P[(size_t)(ix[0] - 1 + i) * N + ix[1]];
    }
    /* y = p0 + p1 x + p2 x^2 */
    double d01 = x[0] - x[1], d02 = x[0] - x[2], d12 = x[1] - x[2];
    double den = d01 * d02 * d12;
    if (den == 0.0) return fds[ix[0]];
    double p2 = (x[2] * (y[1] - y[0]) + x[1] * (y[0] - y[2]) + x[0] * (y[2] - y[1])) / den;
    double p1 = (x[2] * x[2] * (y[0] - y[1]) + x[1] * x[1] * (y[2] - y[0]) +
                 x[0] * x[0] * (y[1] - y[2])) / den;
    if (p2 == 0.0) return fds[ix[0]];
    return -p1 / (2.0 * p2);
}

/* Doppler bins: src/sdr_func.c:758-769 (float step and float bin values) */
int orc_dop_bins(double T, float dop, float max_dop, float *fds)
{
    float step = (float)(0.5 / T);
    int len = (int)(2.0 * max_dop / step) + 1;
    if (fds) {
        for (int i = 0; i < len; i++) fds[i] = dop - max_dop + i * step;
    }
    return len;
}

/* ---------------------------------------------------------------------------
 * code resampling: src/sdr_code.c:2556-2569 and 2589-2608 (code_Q = NULL)
 * j = (int)((coff fs + i) dx) % len_code, dx = len_code / T / fs.
 */
void orc_res_code(const int8_t *code, int len_code, double T, double coff,
                  double fs, int N, int Nz, int8_t *code_res)
{
    double dx = len_code / T / fs;
    memset(code_res, 0, 2 * (size_t)(N + Nz));
    for (int i = 0; i < N; i++) {
        int j = (int)((coff * fs + i) * dx) % len_code;
        code_res[2 * i] = code[j];
        code_res[2 * i + 1] = code[j];
    }
}

void orc_gen_code_fft(const int8_t *code, int len_code, double T, double coff,
                      double fs, int N, int Nz, float *code_fft)
{
    double dx = len_code / T / fs;
    float *res = (float *)calloc(2 * (size_t)(N + Nz), sizeof(float));
    for (int i = 0; i < N; i++) {
        int j = (int)((coff * fs + i) * dx) % len_code;
        res[2 * i] = code[j];
    }
    orc_fft_f32(res, N + Nz, 0, code_fft);
    for (int i = 0; i < N + Nz; i++) code_fft[2 * i + 1] = -code_fft[2 * i + 1];
    free(res);
}

void orc_gen_code_fft_f64(const int8_t *code, int len_code, double T,
                          double coff, double fs, int N, int Nz,
                          double *code_fft)
{
    double dx = len_code / T / fs;
    double *res = (double *)calloc(2 * (size_t)(N + Nz), sizeof(double));
    for (int i = 0; i < N; i++) {
        int j = (int)((coff * fs + i) * dx) % len_code;
        res[2 * i] = code[j];
    }
    fft_f64_impl(res, N + Nz, -1, code_fft);
    for (int i = 0; i < N + Nz; i++) code_fft[2 * i + 1] = -code_fft[2 * i + 1];
    free(res);
}

/* ---------------------------------------------------------------------------
 * GPS L1 C/A: src/sdr_code.c:776-812.  G1 = LFSR(R=0x3FF, tap 0x081),
 * G2 = LFSR(R=0x3FF, tap 0x197), code[i] = G1[i] G2[(i + 1023 - delay) % 1023]
 * with the IS-GPS-200 G2 delays for PRN 1-32.
 */
static const int L1CA_G2_DELAY[32] = {
    5, 6, 7, 8, 17, 18, 139, 140, 141, 251, 252, 254, 255, 256, 257, 258,
    469, 470, 471, 472, 473, 474, 509, 512, 513, 514, 515, 516, 859, 860, 861, 862
};

static void lfsr10(int32_t R, int32_t tap, int8_t *out)
{
    for (int i = 0; i < 1023; i++) {
        out[i] = (R & 1) ? -1 : 1;        /* CHIP[] = {1, -1} */
        int fb = __builtin_parity((unsigned)(R & tap));
        R = (fb << 9) | (R >> 1);
    }
}

int orc_gen_code_L1CA(int prn, int8_t *code)
{
    if (prn < 1 || prn > 32) return 0;
    int8_t g1[1023], g2[1023];
    lfsr10(0x3FF, 0x081, g1);
    lfsr10(0x3FF, 0x197, g2);
    int d = L1CA_G2_DELAY[prn - 1];
    for (int i = 0; i < 1023; i++) code[i] = g1[i] * g2[(i + 1023 - d) % 1023];
    return 1023;
}

/* ---------------------------------------------------------------------------
 * open-loop tracking: src/sdr_ch.c:674-703 (NCO / code advance, adj_coff) and
 * 722-740 (mix + standard correlator).  No loop filters: the Doppler follows
 * fd(t) = fd0 + drift t with t the end time of the epoch block.
 * state (optional, 3 doubles per epoch): phi, coff (s), fc used for the mix.
 */
void orc_trk_run(const orc_trk_ch_t *ch, const int8_t *code_bank,
                 const uint8_t *buff, int buffN, int n_epoch,
                 const double *pos, int n_pos, float *corr, float *C,
                 double *state)
{
    double time = 0.0, phi = ch->phi0, coff = ch->coff0;
    int N = ch->N;
    int8_t *IQ = (int8_t *)malloc(2 * (size_t)N);
    for (int e = 0; e < n_epoch; e++) {
        double t = (e + 1) * ch->T;
        double tau = t - time;
        double fd = ch->fd0 + ch->drift * t;
        double fc = ch->fi + fd;
        /* fused as in the reference build of src/sdr_ch.c:698-700 */
        coff = fma(-(fd / ch->fc_sig), tau, coff);
        phi = fma(fc, tau, phi);
        phi -= floor(phi);
        time = t;
        if (coff >= ch->T) coff -= ch->T;
        else if (coff < 0.0) coff += ch->T;
        int ix = (int)(((long long)e * N) % buffN);
        orc_mix_carr(buff, buffN, ix, N, ch->fs, fc, phi, IQ);
        orc_corr_std(IQ, code_bank, N, coff * ch->fs, pos, n_pos,
                     corr + (size_t)e * 2 * n_pos, C + (size_t)e * 4);
        if (state) {
            state[3 * e] = phi;
            state[3 * e + 1] = coff;
            state[3 * e + 2] = fc;
        }
    }
    free(IQ);
}


/* ---------------------------------------------------------------------------
 * closed-loop tracking, src/sdr_ch.c:488-624 (FLL, PLL, DLL, CN0, test_lost)
 * and 694-763 (track_sig), for signals with no secondary code, bump-jump or
 * L6 branch, from the state start_track leaves (src/sdr_ch.c:364-394), the
 * loop constants at their defaults (src/sdr_ch.c:29-49).  Epoch e runs at
 * time (e + 1) T on ring index (e N) mod buffN.  The reference build's
 * contractions (gcc -O3 -mavx2 -mfma on replicas of these lines): FLL
 * dot = fma(IP1, IP2, QP1 QP2), cross = fma(-QP1, IP2, IP1 QP2); PLL
 * phas_acc = fma(W W W err, T, phas_acc), fd += fma(T, phas_acc,
 * fma(2.4 W, err - err_prev, 1.1 W W err T)); DLL sums += fmaf(re, re, im im)
 * (float, then double), code_int = fma(W W err, dt, code_int),
 * coff = fma(-fma(1.414 W, err, code_int), dt, coff); C/N0 power sums as
 * fmaf(re, re, +-im im), cn0 = fma(fma(log10(.), 10, -cn0), 0.5, cn0).
 * Outputs per epoch: state[6] = {fd, coff, phi, cn0, lock, idle}, corr[n_pos]
 * (cpx), P (cpx, the stitched prompt); epochs after a loss: idle = 1.
 */
#define TL_DPI (2.0 * PI_D)

/* FLL (src/sdr_ch.c:488-506), the pull-in branch of track_sig: from the
 * second locked epoch, the phase change of the prompt since the previous
 * epoch, atan2 (atan with a Costas discriminator) of cross and dot, steers
 * fd with the wide (5 Hz, lock T < 0.5 s) or narrow (2 Hz) band; C is the
 * epoch's correlators, the prompt first.  Called after lock is counted. */
void orc_tl_fll(orc_tl_t *s, const float *C)
{
    if (s->lock >= 2) {
        double IP1 = C[0], QP1 = C[1], IP2 = s->C0[0], QP2 = s->C0[1];
        double dot = fma(IP1, IP2, QP1 * QP2);
        double cross = fma(-QP1, IP2, IP1 * QP2);
        if (dot != 0.0) {
            double B = s->lock * s->T < 0.5 ? 5.0 : 2.0;   /* B_FLL_W, B_FLL_N */
            double err = s->costas ? atan(cross / dot) : atan2(cross, dot);
            s->fd -= B / 0.25 * err / TL_DPI;
        }
    }
    s->C0[0] = C[0]; s->C0[1] = C[1];
}

/* PLL (src/sdr_ch.c:508-521): third order, W = B_PLL / 0.7845 */
void orc_tl_pll(orc_tl_t *s, const float *C)
{
    double IP = C[0], QP = C[1];
    if (IP != 0.0) {
        double err = (s->costas ? atan(QP / IP) : atan2(QP, IP)) / TL_DPI;
        double W = 5.0 / 0.7845;                      /* B_PLL */
        s->phas_acc = fma(W * W * W * err, s->T, s->phas_acc);
        double u = 1.1 * W * W * err * s->T;
        s->fd += fma(s->T, s->phas_acc, fma(2.4 * W, err - s->err_phas, u));
        s->err_phas = err;
    }
}

/* DLL (src/sdr_ch.c:523-550): early (C[2..3]) and late (C[4..5]) powers
 * summed over T_DLL, then the second-order filter, W = B_DLL / 0.53 */
void orc_tl_dll(orc_tl_t *s, const float *C)
{
    int ND = (int)(0.02 / s->T);                          /* T_DLL */
    if (ND < 1) ND = 1;
    s->sumE += fmaf(C[2], C[2], C[3] * C[3]);
    s->sumL += fmaf(C[4], C[4], C[5] * C[5]);
    if (s->lock % ND == 0) {
        double E = sqrt(s->sumE), L = sqrt(s->sumL);
        if (E + L > 0.0) {
            double err = (E - L) / (E + L) * 0.5 * s->T / s->len_code;
            double W = 0.25 / 0.53;                       /* B_DLL */
            double dt = s->T * ND;
            s->code_int = fma(W * W * err, dt, s->code_int);
            s->coff = fma(-fma(1.414 * W, err, s->code_int), dt, s->coff);
        }
        s->sumE = s->sumL = 0.0;
    }
}

/* C/N0, carrier lock and loss (src/sdr_ch.c:574-624): prompt power P (the
 * stitched prompt), IP^2 - QP^2 and the noise tap's power (C[6..7]) summed
 * over T_CN0; at each window PLI = sumD / sumP, the C/N0 estimate filtered
 * by FILT_CN0 = 0.5, and four bad windows in a row (C/N0 < 30 dB-Hz, or a
 * Costas loop's PLI < 0.25) declare the signal lost.  Returns 1 at a window. */
int orc_tl_cn0(orc_tl_t *s, const float *C, float Px, float Py)
{
    s->sumP += fmaf(Px, Px, Py * Py);
    s->sumD += fmaf(Px, Px, -(Py * Py));
    s->sumN += fmaf(C[6], C[6], C[7] * C[7]);
    if (s->bump_step > 0.0) {                             /* npos >= 6: VE, VL */
        s->sumVE += fmaf(C[8], C[8], C[9] * C[9]);
        s->sumVL += fmaf(C[10], C[10], C[11] * C[11]);
    }
    if (s->lock % (int)(0.5 / s->T) != 0) return 0;       /* T_CN0 */
    if (s->sumP > 0.0) { s->pli = s->sumD / s->sumP; s->pli_valid = 1; }
    if (s->sumN > 0.0) s->cn0 = fma(fma(log10(s->sumP / s->sumN / s->T), 10.0, -s->cn0), 0.5, s->cn0);
    if (s->bump_step > 0.0) {
        /* bump-jump (src/sdr_ch.c:552-572): a very-late (very-early) power
         * above BUMP_K = 1.3 times the prompt's, itself above 1.3 times the
         * other side's, moves the code one step later (earlier) */
        if (s->sumVL > 1.3 * s->sumP && s->sumP > 1.3 * s->sumVE) s->coff += s->bump_step;
        else if (s->sumVE > 1.3 * s->sumP && s->sumP > 1.3 * s->sumVL) s->coff -= s->bump_step;
        s->sumVE = s->sumVL = 0.0;
    }
    s->sumP = s->sumN = s->sumD = 0.0;
    int bad = s->cn0 < 30.0 || (s->costas && s->pli_valid && s->pli < 0.25);
    s->lost_cnt = bad ? s->lost_cnt + 1 : 0;
    if (s->lost_cnt >= 4) { s->idle = 1; s->lock = 0; }
    return 1;
}

/* the prompt history's sdr_add_buff (src/sdr_ch.c:740): drop the oldest,
 * append (Px, Py) */
void orc_tl_push(orc_tl_t *s, float Px, float Py)
{
    memmove(s->hist, s->hist + 2, sizeof(float) * 2 * (ORC_N_HIST - 1));
    s->hist[2 * (ORC_N_HIST - 1)] = Px;
    s->hist[2 * (ORC_N_HIST - 1) + 1] = Py;
}

/* sync_sec_code (src/sdr_ch.c:453-486) over the last N = len_sec prompts:
 * unsynchronised, the in-phase prompts correlated with the secondary code,
 * P = sum x_i c_i / N and R = sum |x_i| / N in float, declare sync at this
 * lock count with the sign of P when |P| >= THRES_SEC_RATIO (0.8) R and
 * R >= THRES_SYNC (0.02); synchronised, every N epochs after the sync the
 * plain mean of the last N prompts below THRES_LOST (0.002) drops it.
 * While synchronised the epoch's taps (C, n_tap (re, im) pairs) and the
 * newest prompt are multiplied by the code chip of this epoch times the
 * polarity.  Returns that multiplier (1 when not synchronised).  A lock
 * count that fell more than N below the sync point (the reference's index
 * then goes negative) wraps into the code instead. */
int orc_tl_sec(orc_tl_t *s, float *C, int n_tap)
{
    const int N = s->len_sec;
    const float *P = s->hist + 2 * (ORC_N_HIST - N);
    if (s->sec_sync == 0) {
        float Pc = 0.0f, R = 0.0f;
        for (int i = 0; i < N; i++) {
            const int j = i % s->len_sec;
            Pc += P[2 * i] * (float)s->sec_code[j] / (float)N;
            R += fabsf(P[2 * i]) / (float)N;
        }
        if ((double)fabsf(Pc) >= 0.8 * (double)R && (double)R >= 0.02) {
            s->sec_sync = s->lock;
            s->sec_pol = Pc > 0.0f ? 1 : -1;
        }
    } else if ((s->lock - s->sec_sync) % N == 0) {
        float Pm = 0.0f;
        for (int i = 0; i < N; i++) Pm += P[2 * i] / (float)N;
        if ((double)fabsf(Pm) < 0.002) s->sec_sync = s->sec_pol = 0;
    }
    if (s->sec_sync <= 0) return 1;
    int idx = (s->lock - s->sec_sync - 1 + N) % N % s->len_sec;
    if (idx < 0) idx += s->len_sec;
    const int m = s->sec_code[idx] * s->sec_pol;
    for (int i = 0; i < 2 * n_tap; i++) C[i] *= (float)m;
    s->hist[2 * (ORC_N_HIST - 1)] *= (float)m;
    s->hist[2 * (ORC_N_HIST - 1) + 1] *= (float)m;
    return m;
}

/* one epoch's loop update after the correlators (src/sdr_ch.c:736-763): the
 * stitched prompt (Px, Py) appended to the history (when kept), lock
 * counted, the secondary code synchronised and wiped off (from T_NPULLIN =
 * 1.5 s), FLL during the pull-in (T_FPULLIN = 1 s) else PLL, DLL, C/N0 with
 * bump-jump and the loss decision.  C: the loop's taps P, E, L, N (and VE,
 * VL with bump-jump) as (re, im) pairs, wiped in place.  Returns the
 * secondary-code multiplier applied to this epoch (1: none). */
int orc_tl_update(orc_tl_t *s, float *C, float Px, float Py)
{
    int m = 1;
    if (s->hist) orc_tl_push(s, Px, Py);
    s->lock++;
    if (s->len_sec >= 2 && s->hist && s->lock * s->T >= 1.5) {   /* T_NPULLIN */
        m = orc_tl_sec(s, C, s->bump_step > 0.0 ? 6 : 4);
        Px = s->hist[2 * (ORC_N_HIST - 1)];
        Py = s->hist[2 * (ORC_N_HIST - 1) + 1];
    }
    if (s->lock * s->T <= 1.0) orc_tl_fll(s, C);           /* T_FPULLIN */
    else orc_tl_pll(s, C);
    orc_tl_dll(s, C);
    orc_tl_cn0(s, C, Px, Py);
    return m;
}

void orc_trkl_run(const orc_trk_ch_t *ch, int len_code, int costas, double cn0_0,
                  const int8_t *code_bank, const uint8_t *buff, int buffN, int n_epoch,
                  const double *pos, int n_pos, double *state, float *corr, float *P,
                  int len_sec, const int8_t *sec_code, double bump_step)
{
    const int N = ch->N;
    orc_tl_t s;
    memset(&s, 0, sizeof(s));
    s.T = ch->T; s.len_code = len_code; s.costas = costas;
    s.fd = ch->fd0; s.coff = ch->coff0; s.cn0 = cn0_0; s.pli = 1.0;
    s.len_sec = len_sec; s.sec_code = sec_code; s.bump_step = bump_step;
    s.hist = (float *)calloc(2 * ORC_N_HIST, sizeof(float));   /* trk->P = 0, src/sdr_ch.c:369 */
    const int nlt = bump_step > 0.0 ? 6 : 4;                   /* the loop's taps */
    float Cl[12];
    double time = 0.0, phi = ch->phi0;
    float C1p[2] = {0.0f, 0.0f};
    int8_t *IQ = (int8_t *)malloc(2 * (size_t)N);
    float *C = (float *)malloc(sizeof(float) * 2 * (size_t)(n_pos > 0 ? n_pos : 1));
    float C1[4];
    for (int e = 0; e < n_epoch; e++) {
        double *st = state + 6 * (size_t)e;
        float *pe = P + 2 * (size_t)e;
        if (s.idle) {
            st[0] = s.fd; st[1] = s.coff; st[2] = phi; st[3] = s.cn0; st[4] = s.lock; st[5] = 1;
            pe[0] = pe[1] = 0.0f;
            memset(corr + 2 * (size_t)e * n_pos, 0, sizeof(float) * 2 * n_pos);
            continue;
        }
        double t = (e + 1) * ch->T;
        double tau = t - time;
        double fc = ch->fi + s.fd;
        s.coff = fma(-(s.fd / ch->fc_sig), tau, s.coff);
        phi = fma(fc, tau, phi);
        phi -= floor(phi);
        time = t;
        /* adj_coff (src/sdr_ch.c:674-689), with its shift of the history */
        if (s.coff >= ch->T) {
            s.coff -= ch->T; s.lock--;
            memmove(s.hist + 2, s.hist, sizeof(float) * 2 * (ORC_N_HIST - 1));
        } else if (s.coff < 0.0) {
            s.coff += ch->T; s.lock++;
            memmove(s.hist, s.hist + 2, sizeof(float) * 2 * (ORC_N_HIST - 1));
        }
        int ix = (int)(((long long)e * N) % buffN);
        orc_mix_carr(buff, buffN, ix, N, ch->fs, fc, phi, IQ);
        orc_corr_std(IQ, code_bank, N, s.coff * ch->fs, pos, n_pos, C, C1);
        float Px = (C1[0] + C1p[0]) / N, Py = (C1[1] + C1p[1]) / N;
        C1p[0] = C1[2]; C1p[1] = C1[3];
        memcpy(Cl, C, sizeof(float) * 2 * nlt);
        const float m = (float)orc_tl_update(&s, Cl, Px, Py);
        /* the taps as trk->C holds them: the secondary code wiped off */
        for (int i = 0; i < 2 * n_pos; i++) corr[2 * (size_t)e * n_pos + i] = C[i] * m;
        st[0] = s.fd; st[1] = s.coff; st[2] = phi; st[3] = s.cn0; st[4] = s.lock; st[5] = 0;
        pe[0] = s.hist[2 * (ORC_N_HIST - 1)]; pe[1] = s.hist[2 * (ORC_N_HIST - 1) + 1];
    }
    free(s.hist);
    free(C);
    free(IQ);
}

/* ---------------------------------------------------------------------------
 * prompt history, src/sdr_ch.c:736-740: P[e] = (C[e][0] + C[e-1][1]) / N in
 * float, with trk->C1 = 0 before the first epoch (src/sdr_ch.c:368).  C holds
 * n_epoch x 2 cpx (orc_trk_run's C), P n_epoch cpx.
 */
void orc_trk_stitch(const float *C, int n_epoch, int N, float *P)
{
    float prev[2] = {0.0f, 0.0f};
    for (int e = 0; e < n_epoch; e++) {
        for (int i = 0; i < 2; i++) {
            P[2 * e + i] = (C[4 * e + i] + prev[i]) / N;
            prev[i] = C[4 * e + 2 + i];
        }
    }
}

/* ---------------------------------------------------------------------------
 * L6D / L6E CSK decoding, src/sdr_ch.c:627-664 (interp_corr, CSK).
 * R = N / (len_code / 2) samples per chip, n = (int)(280 R); the reference
 * copies C[k] = corr[(k - n) mod N] for 0 <= k < 2n (lag 0 at index n).
 * Indices outside [0, 2n), which the reference reads past its array when
 * n + ix R + pos < 0 (small R, far negative taps), take the same circular
 * extension here.  The reference build (gcc -O3 -mavx2 -mfma) contracts
 * n + i R to fma(i, R, n), a0 C[i] + a1 C[i+1] to fma(a0, C[i], a1 C[i+1])
 * in double, and sdr_cpx_abs (src/sdr_func.c:179-182) to
 * sqrtf(fmaf(re, re, im * im)).  Returns the chip shift ix of the peak
 * (-255..255; 0 when no power is > 0); C receives the n_pos taps.
 */
static int circ(long long k, int N)
{
    long long r = k % N;
    return (int)(r < 0 ? r + N : r);
}

static void orc_interp_corr(const float *corr, int N, int n, double x, float *c)
{
    int i = (int)x;
    double a1 = x - i, a0 = 1.0 - a1;
    const float *c0 = corr + 2 * (size_t)circ((long long)i - n, N);
    const float *c1 = corr + 2 * (size_t)circ((long long)i + 1 - n, N);
    c[0] = (float)fma(a0, (double)c0[0], a1 * (double)c1[0]);
    c[1] = (float)fma(a0, (double)c0[1], a1 * (double)c1[1]);
}

int orc_csk(const float *corr, int N, int len_code, const double *pos, int n_pos,
            float *C)
{
    double R = (double)N / (len_code / 2);
    int n = (int)(280 * R);
    double P_max = 0.0;
    int ix = 0;
    for (int i = -255; i <= 255; i++) {
        float c[2];
        orc_interp_corr(corr, N, n, fma((double)i, R, (double)n), c);
        double P = sqrtf(fmaf(c[0], c[0], c[1] * c[1]));
        if (P <= P_max) continue;
        P_max = P;
        ix = i;
    }
    double base = fma((double)ix, R, (double)n);
    for (int t = 0; t < n_pos; t++) {
        orc_interp_corr(corr, N, n, base + pos[t], C + 2 * t);
    }
    return ix;
}

/* ---------------------------------------------------------------------------
 * L6D / L6E FFT tracking epochs, src/sdr_ch.c:694-720: the open-loop advance
 * of orc_trk_run, then i = floor(coff fs), bank j = (int)((coff fs - i) 10)
 * (coff fs is one rounded product used twice, as the reference build
 * computes it), sdr_mix_carr from ring index (e N) mod buffN + i at phase
 * phi + fc i / fs, sdr_corr_fft with code_fft + j N, then CSK.  The ring
 * start is taken mod buffN (the reference receiver passes at most buffN,
 * where a start of buffN mixes the same samples as 0).  code_fft: 10 banks of
 * N complex (src/sdr_ch.c:256-262).  Outputs per epoch: syms (uint8
 * 255 - ix % 256), C[n_pos] taps, corr[N] when corr_out is not NULL,
 * state[4] = {phi, coff, i, j}.
 */
void orc_trk_l6_run(const orc_trk_ch_t *ch, int len_code, const float *code_fft,
                    const uint8_t *buff, int buffN, int n_epoch, const double *pos,
                    int n_pos, uint8_t *syms, float *C, float *corr_out, double *state)
{
    double time = 0.0, phi = ch->phi0, coff = ch->coff0;
    int N = ch->N;
    int8_t *IQ = (int8_t *)malloc(2 * (size_t)N);
    float *corr = (float *)malloc(sizeof(float) * 2 * (size_t)N);
    for (int e = 0; e < n_epoch; e++) {
        double t = (e + 1) * ch->T;
        double tau = t - time;
        double fd = ch->fd0 + ch->drift * t;
        double fc = ch->fi + fd;
        coff = fma(-(fd / ch->fc_sig), tau, coff);
        phi = fma(fc, tau, phi);
        phi -= floor(phi);
        time = t;
        if (coff >= ch->T) coff -= ch->T;
        else if (coff < 0.0) coff += ch->T;
        double cs = coff * ch->fs;
        int i = (int)floor(cs);
        int j = (int)((cs - i) * 10);
        int ix = (int)(((long long)e * N) % buffN);
        int start = circ((long long)ix + i, buffN);
        orc_mix_carr(buff, buffN, start, N, ch->fs, fc, phi + fc * i / ch->fs, IQ);
        orc_corr_fft(IQ, code_fft + 2 * (size_t)j * N, N, corr);
        int x = orc_csk(corr, N, len_code, pos, n_pos, C + 2 * (size_t)e * n_pos);
        syms[e] = (uint8_t)(255 - x % 256);
        if (corr_out) memcpy(corr_out + 2 * (size_t)e * N, corr, sizeof(float) * 2 * (size_t)N);
        if (state) {
            state[4 * e] = phi;
            state[4 * e + 1] = coff;
            state[4 * e + 2] = i;
            state[4 * e + 3] = j;
        }
    }
    free(corr);
    free(IQ);
}

/* L6D / L6E closed-loop tracking, src/sdr_ch.c:694-720 with the loop of
 * :741-763: per epoch the NCO and code offset advance by the loop's
 * Doppler (adj_coff moves the lock count), the epoch is mixed from the
 * code-offset sample, FFT-correlated and CSK-decoded as in orc_trk_l6_run,
 * the prompt history takes the CSK's P tap (C[0]), and the loop update
 * (FLL / PLL / DLL / C/N0 / loss, orc_tl_update) reads the CSK's taps.
 * L6D / L6E have no secondary code and are not BOC; the reference runs
 * them with a plain PLL (ch->costas = 0, src/sdr_ch.c:326).  Per epoch:
 * state[6] = {fd, coff, phi, cn0, lock, idle}, syms[e], C[n_pos] (cpx);
 * idle epochs (after a loss) hold zeros. */
void orc_l6l_run(const orc_trk_ch_t *ch, int len_code, int costas, double cn0_0, const float *code_fft,
                 const uint8_t *buff, int buffN, int n_epoch, const double *pos, int n_pos,
                 double *state, uint8_t *syms, float *C)
{
    const int N = ch->N;
    orc_tl_t s;
    memset(&s, 0, sizeof(s));
    s.T = ch->T; s.len_code = len_code; s.costas = costas;
    s.fd = ch->fd0; s.coff = ch->coff0; s.cn0 = cn0_0; s.pli = 1.0;
    double time = 0.0, phi = ch->phi0;
    int8_t *IQ = (int8_t *)malloc(2 * (size_t)N);
    float *corr = (float *)malloc(sizeof(float) * 2 * (size_t)N);
    float Cl[12];
    for (int e = 0; e < n_epoch; e++) {
        double *st = state + 6 * (size_t)e;
        float *Ce = C + 2 * (size_t)e * n_pos;
        if (s.idle) {
            st[0] = s.fd; st[1] = s.coff; st[2] = phi; st[3] = s.cn0; st[4] = s.lock; st[5] = 1;
            syms[e] = 0;
            memset(Ce, 0, sizeof(float) * 2 * n_pos);
            continue;
        }
        double t = (e + 1) * ch->T;
        double tau = t - time;
        double fc = ch->fi + s.fd;
        s.coff = fma(-(s.fd / ch->fc_sig), tau, s.coff);
        phi = fma(fc, tau, phi);
        phi -= floor(phi);
        time = t;
        if (s.coff >= ch->T) { s.coff -= ch->T; s.lock--; }          /* adj_coff */
        else if (s.coff < 0.0) { s.coff += ch->T; s.lock++; }
        double cs = s.coff * ch->fs;
        int i = (int)floor(cs);
        int j = (int)((cs - i) * 10);
        int ix = (int)(((long long)e * N) % buffN);
        int start = circ((long long)ix + i, buffN);
        orc_mix_carr(buff, buffN, start, N, ch->fs, fc, phi + fc * i / ch->fs, IQ);
        orc_corr_fft(IQ, code_fft + 2 * (size_t)j * N, N, corr);
        int x = orc_csk(corr, N, len_code, pos, n_pos, Ce);
        syms[e] = (uint8_t)(255 - x % 256);
        memset(Cl, 0, sizeof(Cl));
        memcpy(Cl, Ce, sizeof(float) * 2 * (n_pos < 6 ? n_pos : 6));
        orc_tl_update(&s, Cl, Ce[0], Ce[1]);
        st[0] = s.fd; st[1] = s.coff; st[2] = phi; st[3] = s.cn0; st[4] = s.lock; st[5] = 0;
    }
    free(corr);
    free(IQ);
}

/* ---------------------------------------------------------------------------
 * front-end raw data to cpx8, src/sdr_rcv.c:776-936 (write_buff and its LUT
 * generators); the LPF and array-combine steps after it are not included
 * ------------------------------------------------------------------------- */
#define FMT_INT8 1      /* src/pocket_sdr.h:85-92 */
#define FMT_INT8X2 2
#define FMT_RAW8 3
#define FMT_RAW16 4
#define FMT_RAW16I 5
#define FMT_RAW32 6
#define FMT_CS8 7
#define FMT_CS16 8

/* SDR_CPX8, src/pocket_sdr.h:105 */
static inline uint8_t mk_cpx8(int re, int im)
{
    return (uint8_t)(((int8_t)im << 4) | (((int8_t)(re << 4) >> 4) & 0xF));
}

/* src/sdr_rcv.c:776-799 */
void orc_gen_lutc(int fmt, int rfch, int IQ, int bits, uint8_t *lut)
{
    static const int8_t vi2[] = {1, 3, -1, -3}, vq2[] = {-1, -3, 1, 3};
    static const int8_t vi3[] = {1, 3, 5, 7, -1, -3, -5, -7};
    for (int i = 0; i < 256; i++) {
        int I, Q = 0;
        if (fmt == FMT_RAW16I) {
            I = vi2[(i >> ((rfch % 4) * 2)) & 3];
        } else if (IQ == 1 && bits == 2) {
            I = vi2[(i >> (rfch % 2 ? 4 : 0)) & 3];
        } else if (IQ == 1 && bits == 3) {
            int v = (i >> (rfch % 2 ? 4 : 0)) & 0xF;
            I = vi3[((v << 1) & 6) + ((v >> 3) & 1)];
        } else {
            I = vi2[(i >> (rfch % 2 ? 4 : 0)) & 3];
            Q = vq2[(i >> (rfch % 2 ? 6 : 2)) & 3];
        }
        lut[i] = mk_cpx8(I, Q);
    }
}

/* src/sdr_rcv.c:817-831 */
void orc_gen_lutc_rtoc(int rfch, uint8_t *lut)
{
    static const int8_t vi3[] = {1, 3, 5, 7, -1, -3, -5, -7};
    for (int i = 0; i < 256; i++) {
        int v = (i >> (rfch % 2 ? 4 : 0)) & 0xF;
        int I = vi3[((v << 1) & 6) + ((v >> 3) & 1)];
        lut[0 * 256 + i] = mk_cpx8(I, 0);
        lut[1 * 256 + i] = mk_cpx8(0, -I);
        lut[2 * 256 + i] = mk_cpx8(-I, 0);
        lut[3 * 256 + i] = mk_cpx8(0, I);
    }
}

/* src/sdr_rcv.c:802-814 */
void orc_gen_lut8(int fmt, double off, double scale, int8_t *lut)
{
    int N = fmt == FMT_CS16 ? 65536 : 256;
    if (scale <= 0.0) scale = 1.0;
    for (int i = -N / 2; i < N / 2; i++) {
        int val = (int)floor((i - off) / scale + 0.5);
        int j = fmt == FMT_CS16 ? (uint16_t)i : (uint8_t)i;
        lut[j] = (int8_t)(val < -7 ? -7 : (val > 7 ? 7 : val));
    }
}

/* src/sdr_rcv.c:834-914 */
int orc_write_buff(int fmt, const uint8_t *raw, int N, long long ix, const int *rtoc,
                   const void *const *lut, uint8_t *const *out)
{
    const int8_t *L8 = (const int8_t *)lut[0];
    const int ph0 = (int)((ix * N) & 3);
#define WR(k, byte, ph) (out[k][j] = rtoc[k] ? ((const uint8_t *)lut[k])[((ph) << 8) | (byte)] \
                                             : ((const uint8_t *)lut[k])[(byte)])
    if (fmt == FMT_INT8) {
        for (int j = 0; j < N; j++) {
            int I = L8[raw[j]];
            if (rtoc[0]) {
                int ph = (ph0 + j) & 3;
                out[0][j] = ph == 0 ? mk_cpx8(I, 0) : ph == 1 ? mk_cpx8(0, -I) :
                            ph == 2 ? mk_cpx8(-I, 0) : mk_cpx8(0, I);
            } else {
                out[0][j] = mk_cpx8(I, 0);
            }
        }
        return 1;
    }
    if (fmt == FMT_INT8X2 || fmt == FMT_CS8) {
        for (int j = 0; j < N; j++) {
            int I = L8[raw[2 * j]], Q = L8[raw[2 * j + 1]];
            out[0][j] = mk_cpx8(I, fmt == FMT_INT8X2 ? -Q : Q);
        }
        return 1;
    }
    if (fmt == FMT_CS16) {
        const int8_t *LQ = lut[1] ? (const int8_t *)lut[1] : L8;
        for (int j = 0; j < N; j++) {
            uint16_t wi = (uint16_t)(raw[4 * j] | (raw[4 * j + 1] << 8));
            uint16_t wq = (uint16_t)(raw[4 * j + 2] | (raw[4 * j + 3] << 8));
            out[0][j] = mk_cpx8(L8[wi], LQ[wq]);
        }
        return 1;
    }
    if (fmt == FMT_RAW8) {
        for (int j = 0; j < N; j++) {
            int ph = (ph0 + j) & 3;
            WR(0, raw[j], ph); WR(1, raw[j], ph);
        }
        return 2;
    }
    if (fmt == FMT_RAW16) {
        for (int j = 0; j < N; j++) {
            int ph = (ph0 + j) & 3;
            WR(0, raw[2 * j], ph); WR(1, raw[2 * j], ph);
            WR(2, raw[2 * j + 1], ph); WR(3, raw[2 * j + 1], ph);
        }
        return 4;
    }
    if (fmt == FMT_RAW16I) {
        for (int j = 0; j < N; j++) {
            for (int k = 0; k < 8; k++) out[k][j] = ((const uint8_t *)lut[k])[raw[2 * j + (k >> 2)]];
        }
        return 8;
    }
    if (fmt == FMT_RAW32) {
        for (int j = 0; j < N; j++) {
            int ph = (ph0 + j) & 3;
            for (int k = 0; k < 8; k++) WR(k, raw[4 * j + (k >> 1)], ph);
        }
        return 8;
    }
#undef WR
    return 0;
}

/* src/sdr_array.c:486-498 (build_lut) and 553-579 (sdr_arch_combine) */
void orc_arch_combine(int nrfch, const uint8_t *const *in, int N, const int16_t *w, uint8_t *out)
{
    int32_t lut[8][256][2];
    for (int j = 0; j < nrfch; j++) {
        int32_t wr = w[j * 2], wi = w[j * 2 + 1];
        for (int b = 0; b < 256; b++) {
            int32_t I = cpx8_I((uint8_t)b), Q = cpx8_Q((uint8_t)b);
            lut[j][b][0] = I * wr - Q * wi;
            lut[j][b][1] = I * wi + Q * wr;
        }
    }
    for (int i = 0; i < N; i++) {
        int32_t I = 0, Q = 0;
        for (int j = 0; j < nrfch; j++) {
            I += lut[j][in[j][i]][0];
            Q += lut[j][in[j][i]][1];
        }
        I += (I >= 0) ? 128 : -128;          /* ARRAY_W_SCALE / 2, src/sdr_array.c:27 */
        Q += (Q >= 0) ? 128 : -128;
        I /= 256;
        Q /= 256;
        out[i] = mk_cpx8(I < -8 ? -8 : (I > 7 ? 7 : I), Q < -8 ? -8 : (Q > 7 ? 7 : Q));
    }
}

