// psa_search.hip - FFT correlator and parallel code search on gfx950.
//
// Drop-in entry points: sdr_cpx_fft, sdr_cpx_mul, sdr_mix_carr, sdr_corr_fft,
// sdr_corr_fft_cpx, sdr_search_code (src/sdr_func.c:197-264, 664-687,
// 772-848, 1088-1113).  Batched entry points: the sdr_search_plan_* family.
//
// Correlator dataflow (per Doppler bin b and block):
//   fwd column pass  : load IF bytes, carrier-mix (LUT semantics), * CSCALE,
//                      length-M1 FFTs, twiddle         -> tmp
//   fwd row pass     : length-M2 FFTs                  -> X (digit-transposed)
//   inv row pass     : (X * Cf) * (1/M/M) with the sdr_cpx_mul arithmetic,
//                      length-M2 inverse FFTs, twiddle -> Z
//   inv column pass  : length-M1 inverse FFTs, P += re^2 + im^2 (natural lag)
// The code spectrum Cf is held in the same digit-transposed order as X, so no
// transpose of data is ever made.  In the batched plan the two forward passes
// run once per (bin, block) and are shared by every PRN of the signal.
#include <math.h>
#include <string.h>
#include <string>
#include <algorithm>
#include <unordered_map>
#include "psa_internal.h"

namespace psa {

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void fill_kernel(float *p, long long n, float v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static void fill(float *p, long long n, float v, hipStream_t st)
{
    if (n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    PSA_CHECK(hipGetLastError());
}

// P[row][j] += re^2 + im^2 (reference contraction P + fmaf(re, re, im * im))
__global__ void power_kernel(const float2 *__restrict__ C, int M, int nlag, long long nrow,
                             float *__restrict__ P, long long pstride)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = nrow * nlag;
    if (t >= total) return;
    long long r = t / nlag;
    int j = (int)(t - r * nlag);
    float2 v = C[r * M + j];
    float *pp = P + r * pstride + j;
    *pp = __fadd_rn(*pp, __fmaf_rn(v.x, v.x, __fmul_rn(v.y, v.y)));
}

// Peak search and C/N0 (src/sdr_func.c:690-706) plus fine Doppler
// (src/sdr_func.c:728-744) over each PRN's surface P[len_fds][N], in two
// stages: partial (max, first index, FP64 sum) per chunk of PEAK_CHUNK values,
// then one workgroup per PRN combines its chunks.  The first maximum in
// row-major order wins (strict '>'); values <= 0 never become the peak.  The
// mean is an FP64 sum divided by the count.
#define PEAK_THREADS 256
#define PEAK_CHUNK 16384

__device__ __forceinline__ void peak_merge(float &bm, long long &bi, float om, long long oi)
{
    if (oi >= 0 && (bi < 0 || om > bm || (om == bm && oi < bi))) { bm = om; bi = oi; }
}

__device__ __forceinline__ void peak_block_reduce(float &bm, long long &bi, double &s, PeakPart *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(bm, o, 64);
        const long long oi = __shfl_xor(bi, o, 64);
        s += __shfl_xor(s, o, 64);
        peak_merge(bm, bi, om, oi);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w].m = bm; red[w].i = bi; red[w].s = s; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < PEAK_THREADS / 64; i++) {
            s += red[i].s;
            peak_merge(bm, bi, red[i].m, red[i].i);
        }
    }
}

__global__ void __launch_bounds__(PEAK_THREADS) peak_part_kernel(const float *__restrict__ Pall,
                                                                 long long tot, int nchunk,
                                                                 PeakPart *__restrict__ part)
{
    __shared__ PeakPart red[PEAK_THREADS / 64];
    const int prn = blockIdx.x / nchunk, c = blockIdx.x - prn * nchunk;
    const float *P = Pall + (long long)prn * tot;
    const long long i0 = (long long)c * PEAK_CHUNK;
    const long long i1 = min(tot, i0 + PEAK_CHUNK);
    float bm = 0.0f;
    long long bi = -1;
    double s = 0.0;
    for (long long i = i0 + threadIdx.x; i < i1; i += PEAK_THREADS) {
        const float v = P[i];
        s += (double)v;
        if (v > bm) { bm = v; bi = i; }
    }
    peak_block_reduce(bm, bi, s, red);
    if (threadIdx.x == 0) {
        PeakPart q;
        q.m = bm; q.pad = 0; q.i = bi; q.s = s;
        part[blockIdx.x] = q;
    }
}

__global__ void __launch_bounds__(PEAK_THREADS) peak_final_kernel(const float *__restrict__ Pall, int N,
                                                                  int len_fds, double T, int nchunk,
                                                                  const PeakPart *__restrict__ part,
                                                                  const float *__restrict__ fds,
                                                                  float *__restrict__ cn0,
                                                                  int *__restrict__ ixo,
                                                                  double *__restrict__ dop)
{
    __shared__ PeakPart red[PEAK_THREADS / 64];
    const int prn = blockIdx.x;
    const long long tot = (long long)len_fds * N;
    const float *P = Pall + (long long)prn * tot;
    float bm = 0.0f;
    long long bi = -1;
    double s = 0.0;
    for (int c = threadIdx.x; c < nchunk; c += PEAK_THREADS) {
        const PeakPart q = part[(long long)prn * nchunk + c];
        s += q.s;
        peak_merge(bm, bi, q.m, q.i);
    }
    peak_block_reduce(bm, bi, s, red);
    if (threadIdx.x == 0) {
        const double Pave = s / (double)tot;
        const double Pmax = bi >= 0 ? (double)bm : 0.0;
        const int i0 = bi >= 0 ? (int)(bi / N) : 0, i1 = bi >= 0 ? (int)(bi % N) : 0;
        cn0[prn] = Pave > 0.0 ? (float)(10.0 * log10((Pmax - Pave) / Pave / T)) : 0.0f;
        ixo[2 * prn] = i0;
        ixo[2 * prn + 1] = i1;
        double d = fds[i0];
        if (i0 > 0 && i0 < len_fds - 1) {
            const double x0 = fds[i0 - 1], x1 = fds[i0], x2 = fds[i0 + 1];
            const double y0 = P[(long long)(i0 - 1) * N + i1], y1 = P[(long long)i0 * N + i1];
            const double y2 = P[(long long)(i0 + 1) * N + i1];
            const double d01 = x0 - x1, d02 = x0 - x2, d12 = x1 - x2;
            const double den = d01 * d02 * d12;
            if (den != 0.0) {
                const double p2 = (x2 * (y1 - y0) + x1 * (y0 - y2) + x0 * (y2 - y1)) / den;
                const double p1 = (x2 * x2 * (y0 - y1) + x1 * x1 * (y2 - y0) + x0 * x0 * (y1 - y2)) / den;
                if (p2 != 0.0) d = -p1 / (2.0 * p2);
            }
        }
        dop[prn] = d;
    }
}

// Per-thread cache of device code spectra in digit-transposed order, keyed
// on the spectrum's address, length and a 64-entry fingerprint: the
// reference's code spectra are made once per channel and never modified
// (src/sdr_ch.c:197-201), so no call compares the whole spectrum.
struct SpecCache {
    const void *host = nullptr;
    int M = 0;
    uint64_t fp = 0;
    float2 *dev = nullptr;
    ~SpecCache() { if (dev) (void)hipFree(dev); }
    static uint64_t fingerprint(const sdr_cpx_t *c, int M)
    {
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < 64; i++) {
            uint32_t w[2];
            memcpy(w, c[(size_t)M * i / 64], sizeof(w));
            h = (h ^ w[0]) * 1099511628211ull;
            h = (h ^ w[1]) * 1099511628211ull;
        }
        return h;
    }
    const float2 *get(const sdr_cpx_t *code_fft, const FFTPlan *pl, hipStream_t st)
    {
        size_t bytes = sizeof(float2) * (size_t)pl->M;
        const uint64_t f = fingerprint(code_fft, pl->M);
        if (host == code_fft && M == pl->M && dev && fp == f) return dev;
        if (dev && M != pl->M) { PSA_CHECK(hipFree(dev)); dev = nullptr; }
        if (!dev) PSA_CHECK(hipMalloc(&dev, bytes * 2));
        host = code_fft;
        M = pl->M;
        fp = f;
        PSA_CHECK(hipMemcpyAsync(dev + pl->M, code_fft, bytes, hipMemcpyHostToDevice, st));
        fft_to_transposed(pl, dev + pl->M, dev, 1, st);
        return dev;
    }
};

static SpecCache &spec_cache()
{
    thread_local SpecCache c;
    return c;
}

// copy N samples of the ring starting at ix into contiguous pinned memory
static const uint8_t *gather_ring(const sdr_buff_t *buff, int ix, int N, uint8_t *dst)
{
    if (ix + N <= buff->N) {
        memcpy(dst, buff->data + ix, N);
    } else {
        int n = buff->N - ix;
        memcpy(dst, buff->data + ix, n);
        memcpy(dst + n, buff->data, N - n);
    }
    return dst;
}

// The drop-ins' dataflow of src/sdr_func.c:1088-1102 for FFT lengths with a
// prime factor above 5: IQ * CSCALE, any-length forward DFT, sdr_cpx_mul by
// the code spectrum with 1 / N^2, any-length backward DFT.  Exits, as the
// reference's allocation failures do, when no transform length exists.
static const AnyPlan *need_any(int N, const char *who)
{
    const AnyPlan *ap = any_plan(N);
    if (!ap) PSA_FAIL("%s: no FFT for length %d", who, N);
    return ap;
}

static void corr_fft_any(const sdr_cpx16_t *IQ, const sdr_cpx_t *code_fft, int N, sdr_cpx_t *corr)
{
    const AnyPlan *ap = need_any(N, "sdr_corr_fft");
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    const size_t bytes = sizeof(float2) * (size_t)N;
    int8_t *diq = (int8_t *)tc.s[SC_IQ].get((size_t)N * 2);
    float2 *x = (float2 *)tc.s[SC_A].get(bytes);
    float2 *cf = (float2 *)tc.s[SC_C].get(bytes);
    float2 *w = (float2 *)tc.s[SC_B].get(2 * sizeof(float2) * (size_t)ap->L);
    uint8_t *hp = (uint8_t *)tc.host_pinned(bytes);
    memcpy(hp, IQ, (size_t)N * 2);
    PSA_CHECK(hipMemcpyAsync(diq, hp, (size_t)N * 2, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(hp, code_fft, bytes);
    PSA_CHECK(hipMemcpyAsync(cf, hp, bytes, hipMemcpyHostToDevice, st));
    launch_cpx16_to_float(diq, N, x, st);
    fft_any(ap, x, x, w, 1, 0, st);
    launch_cpx_mul(x, cf, N, 1.0f / N / N, x, st);
    fft_any(ap, x, x, w, 1, 1, st);
    PSA_CHECK(hipMemcpyAsync(hp, x, bytes, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(corr, hp, bytes);
}

// sdr_search_code (src/sdr_func.c:664-687) for lengths with a prime factor
// above 5: per Doppler bin mix_carr at phase 0, the correlator above, P +=
// |C|^2; the bins of a chunk are batched
static void search_code_any(const sdr_cpx_t *code_fft, const sdr_buff_t *buff, int ix, int N, double fs,
                            double fi, const float *fds, int len_fds, float *P)
{
    const AnyPlan *ap = need_any(N, "sdr_search_code");
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    // bins per pass: work buffers of at most about 1 GiB
    const int chunk = (int)std::max<long long>(1, std::min<long long>(len_fds,
                                                   (1LL << 30) / (16LL * ap->L + 10LL * N)));
    const size_t bytes = sizeof(float2) * (size_t)N;
    uint8_t *hp = (uint8_t *)tc.host_pinned(std::max(bytes, sizeof(float) * (size_t)N * chunk) + (size_t)N);
    uint8_t *din = (uint8_t *)tc.s[SC_IN].get((size_t)N);
    float2 *cf = (float2 *)tc.s[SC_C].get(bytes);
    int8_t *diq = (int8_t *)tc.s[SC_IQ].get((size_t)N * 2 * chunk);
    float2 *x = (float2 *)tc.s[SC_A].get(bytes * chunk);
    float2 *w = (float2 *)tc.s[SC_B].get(2 * sizeof(float2) * (size_t)ap->L * chunk);
    float *dP = (float *)tc.s[SC_P].get(sizeof(float) * (size_t)N * chunk);
    gather_ring(buff, ix, N, hp);
    PSA_CHECK(hipMemcpyAsync(din, hp, N, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(hp, code_fft, bytes);
    PSA_CHECK(hipMemcpyAsync(cf, hp, bytes, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipStreamSynchronize(st));
    for (int b0 = 0; b0 < len_fds; b0 += chunk) {
        const int nb = std::min(chunk, len_fds - b0);
        const size_t pb = sizeof(float) * (size_t)N * nb;
        memcpy(hp, P + (size_t)b0 * N, pb);
        PSA_CHECK(hipMemcpyAsync(dP, hp, pb, hipMemcpyHostToDevice, st));
        for (int b = 0; b < nb; b++) {
            uint32_t p0, p1, s;
            int nseg;
            mix_params(buff->N, ix, N, fs, fi + fds[b0 + b], 0.0, &p0, &p1, &nseg, &s);
            launch_mix(din, N, 0, N, p0, p1, nseg, s, diq + (size_t)b * N * 2, st, 1);
        }
        launch_cpx16_to_float(diq, (long long)N * nb, x, st);
        fft_any(ap, x, x, w, nb, 0, st);
        launch_cpx_mul(x, cf, N, 1.0f / N / N, x, st, nb);
        fft_any(ap, x, x, w, nb, 1, st);
        const long long total = (long long)nb * N;
        hipLaunchKernelGGL(power_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           x, N, N, (long long)nb, dP, (long long)N);
        PSA_CHECK(hipGetLastError());
        PSA_CHECK(hipMemcpyAsync(hp, dP, pb, hipMemcpyDeviceToHost, st));
        PSA_CHECK(hipStreamSynchronize(st));
        memcpy(P + (size_t)b0 * N, hp, pb);
    }
}

}  // namespace psa

using namespace psa;

// ===========================================================================
// drop-in entry points
// ===========================================================================
extern "C" {

int sdr_cpx_fft(sdr_cpx_t *cpx1, int N, int dir, sdr_cpx_t *cpx2)
{
    if (dir != SDR_FFT_FORWARD && dir != SDR_FFT_BACKWARD) return 0;
    if (N <= 0) return 0;
    const FFTPlan *pl = fft_plan(N);
    const AnyPlan *ap = pl ? nullptr : any_plan(N);
    if (!pl && !ap) {
        fprintf(stderr, "pocketsdr_amd: sdr_cpx_fft: unsupported FFT length N=%d\n", N);
        return 0;
    }
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    size_t bytes = sizeof(float2) * (size_t)N;
    float2 *a = (float2 *)tc.s[SC_A].get(bytes);
    float2 *c = (float2 *)tc.s[SC_C].get(bytes);
    void *hp = tc.host_pinned(bytes);
    memcpy(hp, cpx1, bytes);
    PSA_CHECK(hipMemcpyAsync(a, hp, bytes, hipMemcpyHostToDevice, st));
    if (pl) {
        float2 *b = (float2 *)tc.s[SC_B].get(bytes);
        fft_c2c(pl, a, c, b, 1, dir, st);
    } else {
        float2 *b = (float2 *)tc.s[SC_B].get(2 * sizeof(float2) * (size_t)ap->L);
        fft_any(ap, a, c, b, 1, dir, st);
    }
    PSA_CHECK(hipMemcpyAsync(hp, c, bytes, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(cpx2, hp, bytes);
    return 1;
}

void sdr_cpx_mul(const sdr_cpx_t *a, const sdr_cpx_t *b, int N, float s, sdr_cpx_t *c)
{
    if (N <= 0) return;
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    size_t bytes = sizeof(float2) * (size_t)N;
    float2 *da = (float2 *)tc.s[SC_A].get(bytes);
    float2 *db = (float2 *)tc.s[SC_B].get(bytes);
    float2 *dc = (float2 *)tc.s[SC_C].get(bytes);
    uint8_t *hp = (uint8_t *)tc.host_pinned(2 * bytes);
    memcpy(hp, a, bytes);
    memcpy(hp + bytes, b, bytes);
    PSA_CHECK(hipMemcpyAsync(da, hp, bytes, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipMemcpyAsync(db, hp + bytes, bytes, hipMemcpyHostToDevice, st));
    launch_cpx_mul(da, db, N, s, dc, st);
    PSA_CHECK(hipMemcpyAsync(hp, dc, bytes, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(c, hp, bytes);
}

void sdr_mix_carr(const sdr_buff_t *buff, int ix, int N, double fs, double fc,
    double phi, sdr_cpx16_t *IQ)
{
    if (N <= 0) return;
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    uint32_t p0, p1, s;
    int nseg;
    mix_params(buff->N, ix, N, fs, fc, phi, &p0, &p1, &nseg, &s);
    uint8_t *hp = (uint8_t *)tc.host_pinned((size_t)N * 3);
    gather_ring(buff, ix, N, hp);
    uint8_t *din = (uint8_t *)tc.s[SC_IN].get(N);
    int8_t *dout = (int8_t *)tc.s[SC_IQ].get((size_t)N * 2);
    PSA_CHECK(hipMemcpyAsync(din, hp, N, hipMemcpyHostToDevice, st));
    // contiguous copy: the second segment follows the first directly
    launch_mix(din, N, 0, N, p0, p1, nseg, s, dout, st, 1);
    PSA_CHECK(hipMemcpyAsync(hp + N, dout, (size_t)N * 2, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(IQ, hp + N, (size_t)N * 2);
}

void sdr_corr_fft(const sdr_cpx16_t *IQ, const sdr_cpx_t *code_fft, int N, sdr_cpx_t *corr)
{
    if (N <= 0) return;
    const FFTPlan *pl = fft_plan(N);
    if (!pl) {
        corr_fft_any(IQ, code_fft, N, corr);
        return;
    }
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    const float2 *cft = spec_cache().get(code_fft, pl, st);
    size_t bytes = sizeof(float2) * (size_t)N;
    int8_t *diq = (int8_t *)tc.s[SC_IQ].get((size_t)N * 2);
    float2 *ta = (float2 *)tc.s[SC_A].get(bytes);
    float2 *tb = (float2 *)tc.s[SC_B].get(bytes);
    uint8_t *hp = (uint8_t *)tc.host_pinned(bytes);
    memcpy(hp, IQ, (size_t)N * 2);
    PSA_CHECK(hipMemcpyAsync(diq, hp, (size_t)N * 2, hipMemcpyHostToDevice, st));

    PassArgs a = {};
    args_fwd_col(pl, a);
    a.iq16 = diq; a.in_bstride = N; a.out = ta; a.out_bstride = N;
    launch_pass(LD_CPX16, ST_CPX, a, 1, st);
    PassArgs r = {};
    args_fwd_row(pl, r);
    r.in = ta; r.out = tb; r.in_bstride = r.out_bstride = N;
    launch_pass(LD_CPX, ST_CPX, r, 1, st);
    PassArgs ir = {};
    args_inv_row(pl, ir);
    ir.in = tb; ir.in_bstride = N; ir.cf = cft; ir.cf_bstride = N;
    ir.prod_xdiv = 1; ir.prod_xmod = 1; ir.prod_cdiv = 1;
    set_prod(pl, ir, 1.0f / N / N);
    ir.out = ta; ir.out_bstride = N;
    launch_pass(LD_PROD, ST_CPX, ir, 1, st);
    PassArgs ic = {};
    args_inv_col(pl, ic);
    ic.in = ta; ic.in_bstride = N; ic.out = tb; ic.out_bstride = N;
    launch_pass(LD_CPX, ST_CPX, ic, 1, st);
    PSA_CHECK(hipMemcpyAsync(hp, tb, bytes, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(corr, hp, bytes);
}

// src/sdr_func.c:1105-1113: float complex buffer -> cpx8, mix, FFT correlate
void sdr_corr_fft_cpx(const sdr_cpx_t *buff, int len_buff, int ix, int N,
    double fs, double fc, double phi, const sdr_cpx_t *code_fft, sdr_cpx_t *corr)
{
    std::vector<uint8_t> b8(N);
    for (int i = 0, j = ix; i < N; i++, j = (j + 1) % len_buff) {
        int8_t re = (int8_t)buff[j][0], im = (int8_t)buff[j][1];
        b8[i] = (uint8_t)(((int8_t)im << 4) | (((int8_t)(re << 4) >> 4) & 0xF));
    }
    sdr_buff_t tmp;
    tmp.data = b8.data();
    tmp.IQ = 2;
    tmp.N = N;
    std::vector<sdr_cpx16_t> IQ(N);
    sdr_mix_carr(&tmp, 0, N, fs, fc, phi, IQ.data());
    sdr_corr_fft(IQ.data(), code_fft, N, corr);
}

void sdr_search_code(const sdr_cpx_t *code_fft, double T, const sdr_buff_t *buff, int ix,
    int N, double fs, double fi, const float *fds, int len_fds, float *P)
{
    (void)T;
    if (N <= 0 || len_fds <= 0) return;
    const FFTPlan *pl = fft_plan(N);
    if (!pl) {
        search_code_any(code_fft, buff, ix, N, fs, fi, fds, len_fds, P);
        return;
    }
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    const float2 *cft = spec_cache().get(code_fft, pl, st);
    const size_t nb = (size_t)len_fds;
    const size_t bytes = sizeof(float2) * (size_t)N * nb;
    const size_t pbytes = sizeof(float) * (size_t)N * nb;
    // mixing parameters per bin: phi = 0 (src/sdr_func.c:674)
    std::vector<int> hix(nb), hnseg(nb);
    std::vector<uint32_t> hp0(nb), hp1(nb), hs(nb);
    for (size_t i = 0; i < nb; i++) {
        mix_params(buff->N, ix, N, fs, fi + fds[i], 0.0, &hp0[i], &hp1[i], &hnseg[i], &hs[i]);
        hix[i] = 0;
    }
    const size_t prm = nb * sizeof(uint32_t);
    uint8_t *hp = (uint8_t *)tc.host_pinned(std::max<size_t>((size_t)N + 5 * prm + 64, pbytes));
    gather_ring(buff, ix, N, hp);
    uint8_t *q = hp + ((N + 63) / 64) * 64;
    memcpy(q, hix.data(), prm); memcpy(q + prm, hnseg.data(), prm);
    memcpy(q + 2 * prm, hp0.data(), prm); memcpy(q + 3 * prm, hp1.data(), prm);
    memcpy(q + 4 * prm, hs.data(), prm);
    uint8_t *din = (uint8_t *)tc.s[SC_IN].get((size_t)N + 5 * prm + 64);
    uint8_t *dq = din + ((N + 63) / 64) * 64;
    PSA_CHECK(hipMemcpyAsync(din, hp, (q - hp) + 5 * prm, hipMemcpyHostToDevice, st));
    float2 *ta = (float2 *)tc.s[SC_A].get(bytes);
    float2 *tb = (float2 *)tc.s[SC_B].get(bytes);
    float *dP = (float *)tc.s[SC_P].get(pbytes);
    // P is accumulated in place: upload it after the parameter copy has been
    // staged (the pinned buffer is reused for P)
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(hp, P, pbytes);
    PSA_CHECK(hipMemcpyAsync(dP, hp, pbytes, hipMemcpyHostToDevice, st));

    PassArgs a = {};
    args_fwd_col(pl, a);
    a.buf = din; a.buffN = N; a.mix_compact = 1;
    a.mix_ix = (const int *)dq; a.mix_nseg = (const int *)(dq + prm);
    a.mix_p0 = (const uint32_t *)(dq + 2 * prm); a.mix_p1 = (const uint32_t *)(dq + 3 * prm);
    a.mix_s = (const uint32_t *)(dq + 4 * prm);
    a.out = ta; a.out_bstride = N;
    launch_pass(LD_MIX, ST_CPX, a, (int)nb, st);
    PassArgs r = {};
    args_fwd_row(pl, r);
    r.in = ta; r.out = tb; r.in_bstride = r.out_bstride = N;
    launch_pass(LD_CPX, ST_CPX, r, (int)nb, st);
    PassArgs ir = {};
    args_inv_row(pl, ir);
    ir.in = tb; ir.in_bstride = N; ir.cf = cft; ir.cf_bstride = 0;
    ir.prod_xdiv = 1; ir.prod_xmod = (int)nb; ir.prod_cdiv = (int)nb;
    set_prod(pl, ir, 1.0f / N / N);
    ir.out = ta; ir.out_bstride = N;
    launch_pass(LD_PROD, ST_CPX, ir, (int)nb, st);
    PassArgs ic = {};
    args_inv_col(pl, ic);
    ic.in = ta; ic.in_bstride = N;
    ic.P = dP; ic.p_bstride = N; ic.p_div = 1; ic.p_off1 = 0; ic.p_mul = 1; ic.p_off2 = 0;
    ic.nlag = N;
    launch_pass(LD_CPX, ST_POW, ic, (int)nb, st);
    PSA_CHECK(hipMemcpyAsync(hp, dP, pbytes, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(P, hp, pbytes);
}

}  // extern "C"

// ===========================================================================
// batched search plan
// ===========================================================================
namespace psa {

// host copy of the per-(block, bin) mixing parameters
struct PlainMix {
    std::vector<int> ix, nseg;
    std::vector<uint32_t> p0, p1, s;
};

struct SigDev {
    sdr_search_sig_t h;
    PlainMix hm;
    const FFTPlan *pl = nullptr;
    float2 *cft = nullptr;      // n_prn x M, digit-transposed (batched)
    float2 *cfn = nullptr;      // n_prn x M, natural order (plain)
    const HPlan *hp = nullptr;  // precision form (FP16 / BF16 MFMA passes)
    uint32_t *cfh = nullptr;    // n_prn x M words, digit-transposed, scaled
    float kappa = 1.0f;         // power scale of the precision form
    int hvar_col = 0, hvar_row = 0;   // workgroup shapes of its passes
    float2 *cf_nat = nullptr;   // natural-order FP32 spectra (until set up)
    float *P = nullptr;         // n_prn x len_fds x N
    float *fds = nullptr;       // device copy of the bins
    int *mix_ix = nullptr, *mix_nseg = nullptr;   // [n_blk x len_fds]
    uint32_t *mix_p0 = nullptr, *mix_p1 = nullptr, *mix_s = nullptr;
    float *cn0 = nullptr;
    int *ix = nullptr;
    double *dop = nullptr;
    PeakPart *part = nullptr;   // [n_prn x nchunk] partial peak reductions
    int nchunk = 0;             // partials per PRN of the separate peak pass
    int nchunk_fused = 0;       // partials per PRN emitted by the last column pass
    bool fused_peak = false;    // the last column pass emits the peak partials
    bool xcd_order = true;      // XCD-aware workgroup order of the inverse row pass
    bool pair_blocks = false;   // the power pass of its column shape takes block pairs
    bool cell = false;          // FP16 inverse as fused cell kernels (launch_cells)
    int lane = 0;               // the lane (stream and work buffers) the group runs on
    int bin_chunk = 1, prn_chunk = 1;
    float tuned_ms = 0.0f;      // tuner: best first-block time (0: not tuned)
};

enum { K_FWD_COL = 0, K_FWD_ROW, K_INV_ROW, K_INV_COL, K_MIX, K_CVT, K_MUL, K_POW, K_PEAK, K_FILL, K_CELL,
       K_NCLS };
static const char *k_names[K_NCLS] = {"fwd_col", "fwd_row", "inv_row", "inv_col", "mix",
                                      "cvt", "mul", "pow", "peak", "fill", "cell"};

}  // namespace psa

// Signal groups run on up to PSA_LANES lanes at once: a lane is a stream
// with its own work buffers (sized for the largest of its groups); the groups
// of one lane run in order.  The plan's stream forks to the lanes and joins
// them through events, so a captured run is one graph with parallel branches.
#define PSA_LANES 4

struct SearchLane {
    hipStream_t st = nullptr;
    float2 *tmp = nullptr, *X = nullptr, *Z = nullptr;
    int8_t *iq16 = nullptr;     // plain path mixed samples
    size_t tmp_elems = 0, x_elems = 0, z_elems = 0;
    hipEvent_t done = nullptr;  // end of the lane's groups in a run
};

struct sdr_search_plan_tag {
    int flags = 0;
    std::vector<uint8_t *> d_buf;
    bool own_buf = true;        // false: IF buffers in the caller's device memory
    std::vector<int> buffN;
    std::vector<psa::SigDev> sig;
    std::vector<SearchLane> lanes;
    // per IF stream: a packed raw ring (bps > 0, read through the channel's
    // table by the mixing loaders) or cpx8 samples (bps 0)
    struct RawDesc { int bps = 0, off = 0, rtoc = 0, ph0 = 0; uint8_t *dlut = nullptr; };
    std::vector<RawDesc> raw;
    hipEvent_t fork = nullptr;  // start of a run on the plan's stream
    std::vector<hipEvent_t> gdone;   // end of each group (timed runs chain them)
    hipStream_t st = nullptr;
    hipStream_t cur_st = nullptr;    // stream of the group being enqueued
    hipGraphExec_t gexec = nullptr;
    bool timing = false;
    struct Ev { int cls, sig; hipEvent_t a, b; double elems, bytes, grid; };
    int cur_sig = 0;
    std::vector<Ev> evs;
    size_t ev_used = 0;
    // per (kernel class, signal group) sums over the timed runs
    struct Stat { double ms = 0, elems = 0, bytes = 0, grid = 0; long long cnt = 0; };
    std::vector<Stat> stats;     // [K_NCLS * n_sig]
};

namespace psa {

// Device memory of the search plans goes through pmalloc / pfree, which
// account it, so that sdr_search_plan_bytes reports what live plans hold.
static std::mutex g_pm_mtx;
static std::unordered_map<void *, size_t> g_pm;
static long long g_pm_live = 0;

template <class T>
static void pmalloc(T **p, size_t bytes)
{
    PSA_CHECK(hipMalloc((void **)p, bytes));
    std::lock_guard<std::mutex> lk(g_pm_mtx);
    g_pm[(void *)*p] = bytes;
    g_pm_live += (long long)bytes;
}

static void pfree(void *p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_pm_mtx);
        auto it = g_pm.find(p);
        if (it != g_pm.end()) {
            g_pm_live -= (long long)it->second;
            g_pm.erase(it);
        }
    }
    (void)hipFree(p);
}

// records a timed launch when timing is on; `bytes` is the launch's
// algorithmic HBM traffic (compulsory reads + writes of its operands)
struct Tick {
    sdr_search_plan_t *p;
    int cls;
    double elems;
    size_t slot;
    Tick(sdr_search_plan_t *p_, int c, double e, double bytes = 0.0) : p(p_), cls(c), elems(e), slot(0)
    {
        if (!p->timing) return;
        if (p->ev_used == p->evs.size()) {
            sdr_search_plan_t::Ev ev;
            PSA_CHECK(hipEventCreate(&ev.a));
            PSA_CHECK(hipEventCreate(&ev.b));
            p->evs.push_back(ev);
        }
        slot = p->ev_used++;
        p->evs[slot].cls = cls;
        p->evs[slot].sig = p->cur_sig;
        p->evs[slot].elems = elems;
        p->evs[slot].bytes = bytes;
        pass_grid = 0;
        PSA_CHECK(hipEventRecord(p->evs[slot].a, p->cur_st));
    }
    ~Tick()
    {
        if (!p->timing) return;
        p->evs[slot].grid = (double)pass_grid;
        PSA_CHECK(hipEventRecord(p->evs[slot].b, p->cur_st));
    }
};

// the raw-ring parameters of IF stream i for the mixing loaders
static void set_raw(const sdr_search_plan_t *p, int i, PassArgs &a)
{
    if (p->raw.empty() || !p->raw[i].bps) return;
    const sdr_search_plan_t::RawDesc &r = p->raw[i];
    a.raw_bps = r.bps; a.raw_off = r.off; a.raw_rtoc = r.rtoc; a.raw_ph0 = r.ph0; a.raw_lut = r.dlut;
}

// the batched, shared-forward-FFT schedule of one signal group
static void run_batched(sdr_search_plan_t *p, SigDev &s)
{
    SearchLane &ln = p->lanes[s.lane];
    hipStream_t st = ln.st;
    const FFTPlan *pl = s.pl;
    const int M = pl->M, N = s.h.N, nbin = s.h.len_fds, nprn = s.h.n_prn;
    const uint8_t *buf = p->d_buf[s.h.buff];
    const int buffN = p->buffN[s.h.buff];
    // P is not cleared: the first block stores it, later blocks accumulate
    for (int k = 0; k < s.h.n_blk; k++) {
        for (int b0 = 0; b0 < nbin; b0 += s.bin_chunk) {
            const int nbc = std::min(s.bin_chunk, nbin - b0);
            const long long mo = (long long)k * nbin + b0;
            PassArgs a = {};
            args_fwd_col(pl, a);
            a.buf = buf; a.buffN = buffN; a.mix_compact = 0;
            set_raw(p, s.h.buff, a);
            a.mix_ix = s.mix_ix + mo; a.mix_nseg = s.mix_nseg + mo;
            a.mix_p0 = s.mix_p0 + mo; a.mix_p1 = s.mix_p1 + mo; a.mix_s = s.mix_s + mo;
            a.out = ln.tmp; a.out_bstride = M;
            {
                Tick t(p, K_FWD_COL, (double)nbc * M, (double)M + 8.0 * nbc * M);
                launch_pass(LD_MIX, ST_CPX, a, nbc, st);
            }
            PassArgs r = {};
            args_fwd_row(pl, r);
            r.in = ln.tmp; r.out = ln.X; r.in_bstride = r.out_bstride = M;
            {
                Tick t(p, K_FWD_ROW, (double)nbc * M, 16.0 * nbc * M);
                launch_pass(LD_CPX, ST_CPX, r, nbc, st);
            }
            for (int q0 = 0; q0 < nprn; q0 += s.prn_chunk) {
                const int nqc = std::min(s.prn_chunk, nprn - q0);
                const int nbatch = nqc * nbc;      // b = prn_local * nbc + bin_local
                PassArgs ir = {};
                args_inv_row(pl, ir);
                ir.in = ln.X; ir.in_bstride = M;
                ir.cf = s.cft + (long long)q0 * M; ir.cf_bstride = M;
                ir.prod_xdiv = 1; ir.prod_xmod = nbc; ir.prod_cdiv = nbc;
                set_prod(pl, ir, 1.0f / M / M);
                ir.out = ln.Z; ir.out_bstride = M;
                if (s.xcd_order && spec_length(pl->M2)) {
                    ir.xcd_nbin = nbc;
                    ir.xcd_nprn = nqc;
                }
                {
                    Tick t(p, K_INV_ROW, (double)nbatch * M, 8.0 * M * (nbc + nqc + (double)nbatch));
                    launch_pass(LD_PROD, ST_CPX, ir, nbatch, st);
                }
                PassArgs ic = {};
                args_inv_col(pl, ic);
                ic.in = ln.Z; ic.in_bstride = M;
                ic.P = s.P; ic.p_bstride = N;
                ic.p_div = nbc; ic.p_off1 = q0; ic.p_mul = nbin; ic.p_off2 = b0;
                ic.nlag = N;
                ic.p_nbin = nbin;
                ic.p_init = k == 0;
                if (k == s.h.n_blk - 1 && s.fused_peak) ic.p_part = s.part;
                {
                    // P: written on the first block, read and written after
                    Tick t(p, K_INV_COL, (double)nbatch * M,
                           (double)nbatch * (8.0 * M + (k == 0 ? 4.0 : 8.0) * N));
                    launch_pass(LD_CPX, ST_POW, ic, nbatch, st);
                }
            }
        }
    }
}

// precision-form groups whose forward passes run for all blocks and bins at
// once (spectra of every block held): cell groups, and groups whose bins are
// not chunked
static bool fwd_all_blocks(const SigDev &s)
{
    return s.hp && (s.cell || s.bin_chunk == s.h.len_fds);
}

// The precision form of run_batched: FP16 / BF16 words between the passes,
// MFMA DFT-matrix stages.  Scales (powers of two): forward column pass output
// osc_fc (folded into its twiddle table), forward row pass osc_fr, code
// spectra 2^-ceil(log2 N) (|Cf| <= 1), inverse row pass osc_ir; the inverse
// column pass multiplies by kappa, which undoes them together with the
// reference's 1/11.2 sample scale and 1/M^2 product scale.
static void run_batched_h(sdr_search_plan_t *p, SigDev &s)
{
    SearchLane &ln = p->lanes[s.lane];
    hipStream_t st = ln.st;
    const HPlan *hp = s.hp;
    const int dt = hp->dt;
    const int M = hp->M, N = s.h.N, nbin = s.h.len_fds, nprn = s.h.n_prn;
    const uint8_t *buf = p->d_buf[s.h.buff];
    const int buffN = p->buffN[s.h.buff];
    uint32_t *tmp = reinterpret_cast<uint32_t *>(ln.tmp);
    uint32_t *X = reinterpret_cast<uint32_t *>(ln.X);
    uint32_t *Z = reinterpret_cast<uint32_t *>(ln.Z);
    // unchunked bins (and cell groups): the forward passes of every (block,
    // bin) in one launch each, before the inverse passes
    const bool fwd_all = fwd_all_blocks(s);
    if (fwd_all) {
        const int nall = s.h.n_blk * nbin;
        PassArgs a = {};
        h_col_args(hp, 0, a, s.hvar_col);
        a.tfull = hp->tfull[0];
        a.buf = buf; a.buffN = buffN; a.mix_compact = 0;
        set_raw(p, s.h.buff, a);
        a.mix_ix = s.mix_ix; a.mix_nseg = s.mix_nseg;
        a.mix_p0 = s.mix_p0; a.mix_p1 = s.mix_p1; a.mix_s = s.mix_s;
        a.hout = tmp; a.out_bstride = M;
        {
            Tick t(p, K_FWD_COL, (double)nall * M, (double)s.h.n_blk * M + 4.0 * nall * M);
            launch_hpass(dt, LD_MIX, ST_CPX, a, nall, st);
        }
        PassArgs r = {};
        h_row_args(hp, 0, r, s.hvar_row);
        r.osc = hp->osc_fr;
        r.hin = tmp; r.hout = X; r.in_bstride = r.out_bstride = M;
        {
            Tick t(p, K_FWD_ROW, (double)nall * M, 8.0 * nall * M);
            launch_hpass(dt, LD_CPX, ST_CPX, r, nall, st);
        }
    }
    if (s.cell) {
        // one cell kernel over all (PRN, bin) cells and blocks
        const int nall = s.h.n_blk * nbin;
        CellArgs c = {};
        c.X = X; c.xblk = (long long)nbin * M;
        c.C = s.cfh;
        c.P = s.P; c.part = s.part;
        c.nprn = nprn; c.nbin = nbin; c.nblk = s.h.n_blk; c.N = N;
        c.kap2 = s.kappa * s.kappa;
        {
            // X of every block and bin, the code spectra, P once
            Tick t(p, K_CELL, (double)nall * nprn * M,
                   4.0 * M * ((double)nall + nprn) + 4.0 * (double)nprn * nbin * N);
            launch_cells(hp, c, st);
        }
        return;
    }
    // blocks in pairs (one power pass per pair, P read and written once)
    // when the spectra of every block are resident and the power pass takes
    // two inputs; the pair's intermediates then live in tmp, free after the
    // forward passes
    const size_t zc = (size_t)s.bin_chunk * s.prn_chunk * M;    // words of one Z chunk
    const bool in_tmp = 2 * zc <= 2 * ln.tmp_elems, in_z = 2 * zc <= 2 * ln.z_elems;
    const bool pairs = fwd_all && s.pair_blocks && s.h.n_blk >= 2 && (in_tmp || in_z);
    uint32_t *Z2 = nullptr;
    if (pairs) {
        if (in_tmp) Z = tmp;
        Z2 = Z + zc;
    }
    for (int k = 0; k < s.h.n_blk; k += pairs ? 2 : 1) {
        const bool two = pairs && k + 1 < s.h.n_blk;
        for (int b0 = 0; b0 < nbin; b0 += s.bin_chunk) {
            const int nbc = std::min(s.bin_chunk, nbin - b0);
            const long long mo = (long long)k * nbin + b0;
            if (!fwd_all) {
                PassArgs a = {};
                h_col_args(hp, 0, a, s.hvar_col);
                a.tfull = hp->tfull[0];
                a.buf = buf; a.buffN = buffN; a.mix_compact = 0;
                set_raw(p, s.h.buff, a);
                a.mix_ix = s.mix_ix + mo; a.mix_nseg = s.mix_nseg + mo;
                a.mix_p0 = s.mix_p0 + mo; a.mix_p1 = s.mix_p1 + mo; a.mix_s = s.mix_s + mo;
                a.hout = tmp; a.out_bstride = M;
                {
                    Tick t(p, K_FWD_COL, (double)nbc * M, (double)M + 4.0 * nbc * M);
                    launch_hpass(dt, LD_MIX, ST_CPX, a, nbc, st);
                }
                PassArgs r = {};
                h_row_args(hp, 0, r, s.hvar_row);
                r.osc = hp->osc_fr;
                r.hin = tmp; r.hout = X; r.in_bstride = r.out_bstride = M;
                {
                    Tick t(p, K_FWD_ROW, (double)nbc * M, 8.0 * nbc * M);
                    launch_hpass(dt, LD_CPX, ST_CPX, r, nbc, st);
                }
            }
            // this block's spectra (X of every block when computed at once)
            const uint32_t *Xk = fwd_all ? X + mo * M : X;
            for (int q0 = 0; q0 < nprn; q0 += s.prn_chunk) {
                const int nqc = std::min(s.prn_chunk, nprn - q0);
                const int nbatch = nqc * nbc;      // b = prn_local * nbc + bin_local
                PassArgs ir = {};
                h_row_args(hp, 1, ir, s.hvar_row);
                ir.tfull = hp->tfull[1];
                ir.tfullh = hp->tfullh;
                ir.hin = Xk; ir.in_bstride = M;
                ir.hcf = s.cfh + (long long)q0 * M; ir.cf_bstride = M;
                ir.prod_xdiv = 1; ir.prod_xmod = nbc; ir.prod_cdiv = nbc;
                ir.hout = Z; ir.out_bstride = M;
                if (s.xcd_order) {
                    ir.xcd_nbin = nbc;
                    ir.xcd_nprn = nqc;
                }
                {
                    Tick t(p, K_INV_ROW, (double)nbatch * M, 4.0 * M * (nbc + nqc + (double)nbatch));
                    launch_hpass(dt, LD_PROD, ST_CPX, ir, nbatch, st);
                }
                if (two) {
                    // the next block's product row pass into Z2
                    ir.hin = Xk + (long long)nbin * M;
                    ir.hout = Z2;
                    Tick t(p, K_INV_ROW, (double)nbatch * M, 4.0 * M * (nbc + nqc + (double)nbatch));
                    launch_hpass(dt, LD_PROD, ST_CPX, ir, nbatch, st);
                }
                PassArgs ic = {};
                h_col_args(hp, 1, ic, s.hvar_col);
                ic.kappa = s.kappa;
                ic.hin = Z; ic.in_bstride = M;
                ic.hin2 = two ? Z2 : nullptr;
                ic.P = s.P; ic.p_bstride = N;
                ic.p_div = nbc; ic.p_off1 = q0; ic.p_mul = nbin; ic.p_off2 = b0;
                ic.nlag = N;
                ic.p_nbin = nbin;
                ic.p_init = k == 0;
                const int klast = two ? k + 1 : k;
                if (klast == s.h.n_blk - 1 && s.fused_peak) ic.p_part = s.part;
                {
                    Tick t(p, K_INV_COL, (double)nbatch * M * (two ? 2 : 1),
                           (double)nbatch * ((two ? 8.0 : 4.0) * M + (k == 0 ? 4.0 : 8.0) * N));
                    launch_hpass(dt, LD_CPX, ST_POW, ic, nbatch, st);
                }
            }
        }
    }
}

// the reference dataflow, one sdr_search_code-equivalent per (PRN, block):
// mix -> cpx16, forward FFT, sdr_cpx_mul, inverse FFT, power accumulate,
// each its own kernel, nothing shared across PRNs
static void run_plain(sdr_search_plan_t *p, SigDev &s)
{
    SearchLane &ln = p->lanes[s.lane];
    hipStream_t st = ln.st;
    const FFTPlan *pl = s.pl;
    const int M = pl->M, N = s.h.N, nbin = s.h.len_fds, nprn = s.h.n_prn;
    const uint8_t *buf = p->d_buf[s.h.buff];
    const int buffN = p->buffN[s.h.buff];
    {
        Tick t(p, K_FILL, (double)nprn * nbin * N);
        fill(s.P, (long long)nprn * nbin * N, 0.0f, st);
    }
    for (int q = 0; q < nprn; q++) {
        for (int k = 0; k < s.h.n_blk; k++) {
            for (int b0 = 0; b0 < nbin; b0 += s.bin_chunk) {
                const int nbc = std::min(s.bin_chunk, nbin - b0);
                {
                    Tick t(p, K_MIX, (double)nbc * M);
                    for (int b = 0; b < nbc; b++) {
                        size_t i = (size_t)k * nbin + b0 + b;
                        launch_mix(buf, buffN, s.hm.ix[i], M, s.hm.p0[i], s.hm.p1[i], s.hm.nseg[i],
                                   s.hm.s[i], ln.iq16 + (size_t)b * M * 2, st);
                    }
                }
                PassArgs a = {};
                args_fwd_col(pl, a);
                a.iq16 = ln.iq16; a.in_bstride = M; a.out = ln.tmp; a.out_bstride = M;
                {
                    Tick t(p, K_FWD_COL, (double)nbc * M);
                    launch_pass(LD_CPX16, ST_CPX, a, nbc, st);
                }
                PassArgs r = {};
                args_fwd_row(pl, r);
                r.in = ln.tmp; r.out = ln.X; r.in_bstride = r.out_bstride = M;
                r.out_nstride = pl->M1; r.out_gstride = 1;    // natural order
                {
                    Tick t(p, K_FWD_ROW, (double)nbc * M);
                    launch_pass(LD_CPX, ST_CPX, r, nbc, st);
                }
                {
                    Tick t(p, K_MUL, (double)nbc * M);
                    for (int b = 0; b < nbc; b++) {
                        launch_cpx_mul(ln.X + (size_t)b * M, s.cfn + (size_t)q * M, M,
                                       1.0f / M / M, ln.Z + (size_t)b * M, st);
                    }
                }
                {
                    Tick t(p, K_INV_ROW, (double)nbc * M);
                    fft_c2c(pl, ln.Z, ln.X, ln.tmp, nbc, 1, st);
                }
                {
                    Tick t(p, K_POW, (double)nbc * N);
                    long long total = (long long)nbc * N;
                    hipLaunchKernelGGL(power_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                                       ln.X, M, N, (long long)nbc,
                                       s.P + ((long long)q * nbin + b0) * N, (long long)N);
                    PSA_CHECK(hipGetLastError());
                }
            }
        }
    }
}

// Precision-form group set up for split M1 (column length) and workgroup
// shapes vc, vr: code spectra in that split's digit-transposed order, scaled
// by 2^-ceil(log2 N) (|code_fft| <= N for a zero-padded code of N samples in
// [-1, 1]), the power scale kappa, the peak partial buffer.
static void h_setup(sdr_search_plan_t *p, SigDev &s, int M1, int vc, int vr)
{
    const int M = s.h.M, nprn = s.h.n_prn;
    s.hp = hfft_plan(M, s.hp->dt, M1);
    s.hvar_col = vc;
    s.hvar_row = vr;
    s.pair_blocks = h_col_pairs(s.hp->dt, M1, vc) && !getenv("PSA_NO_PAIR");
    const int e = (int)ceil(log2((double)s.h.N));
    const double sc = ldexp(1.0, -e);
    h_code_spec(s.hp, s.cf_nat, s.cfh, nprn, (float)sc, p->st);
    PSA_CHECK(hipStreamSynchronize(p->st));
    s.kappa = (float)((double)PSA_CSCALE / ((double)M * M * s.hp->osc_fc * s.hp->osc_fr * sc * s.hp->osc_ir));
    PassArgs tmp = {};
    h_col_args(s.hp, 1, tmp, vc);
    // one peak partial per workgroup of the last column pass (cell kernels:
    // one per cell)
    const int nf = s.cell ? s.h.len_fds : s.h.len_fds * pass_ngroups(tmp);
    if (s.cell) launch_cells(s.hp, CellArgs{}, p->st, true);
    if (nf != s.nchunk_fused || !s.part) {
        if (s.part) pfree(s.part);
        s.nchunk_fused = nf;
        pmalloc(&s.part, sizeof(PeakPart) * (size_t)nprn * std::max(s.nchunk, s.nchunk_fused));
    }
}

// FP32 batched group set up for split M1: code spectra in its
// digit-transposed order, the peak partial buffer.
static void f_setup(sdr_search_plan_t *p, SigDev &s, int M1)
{
    s.pl = fft_plan(s.h.M, M1);
    fft_to_transposed(s.pl, s.cf_nat, s.cft, s.h.n_prn, p->st);
    PSA_CHECK(hipStreamSynchronize(p->st));
    s.fused_peak = spec_length(s.pl->M1) && !getenv("PSA_NO_FUSED_PEAK");
    PassArgs tmp = {};
    col_pass_args(s.pl, 1, tmp);
    const int nf = s.h.len_fds * pass_ngroups(tmp);
    if (nf != s.nchunk_fused || !s.part) {
        if (s.part) pfree(s.part);
        s.nchunk_fused = nf;
        pmalloc(&s.part, sizeof(PeakPart) * (size_t)s.h.n_prn * std::max(s.nchunk, s.nchunk_fused));
    }
}

// PSA_TILES fixes the tiling of FFT length M instead of timing it:
// comma-separated entries [h|f]M=M1/col_shape/row_shape, the prefix h or f
// restricting an entry to the precision form or to FP32 (the shapes apply to
// the precision form), e.g. "h96000=300/0/0,f96000=200".  Returns false when
// the variable has no valid entry for the group.
static bool fixed_tiles(const SigDev &s, int *m1, int *vc, int *vr)
{
    const char *env = getenv("PSA_TILES");
    if (!env) return false;
    int cand[8];
    const int ns = s.hp ? h_splits(s.h.M, cand, 8) : fft_splits(s.h.M, cand, 8);
    for (const char *q = env; *q;) {
        int M = 0, a = 0, c = 0, r = 0;
        bool form_ok = true;
        if (*q == 'h' || *q == 'f') form_ok = (*q++ == 'h') == (s.hp != nullptr);
        const int n = sscanf(q, "%d=%d/%d/%d", &M, &a, &c, &r);
        if (form_ok && n >= 2 && M == s.h.M && std::find(cand, cand + ns, a) != cand + ns &&
            (!s.hp || h_shape(M, a, c, r))) {
            *m1 = a; *vc = s.hp ? c : 0; *vr = s.hp ? r : 0;
            return true;
        }
        const char *comma = strchr(q, ',');
        if (!comma) break;
        q = comma + 1;
    }
    return false;
}

// SDR_SEARCH_TUNE: time every candidate tiling of the group's passes on its
// own shapes - the first two blocks of the search, best of two runs after a
// warm-up - and keep the fastest.  Precision form: (split, column shape, row
// shape); FP32: the split.
static void tune_group(sdr_search_plan_t *p, SigDev &s)
{
    int m1[8];
    const bool h = s.hp != nullptr;
    const int ns = h ? h_splits(s.h.M, m1, 8) : fft_splits(s.h.M, m1, 8);
    const int nv = h ? PSA_HVARS : 1;
    const int nblk = s.h.n_blk;
    hipEvent_t e0, e1;
    PSA_CHECK(hipEventCreate(&e0));
    PSA_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    int bm = m1[0], bc = 0, br = 0;
    // cell groups have the one split of their cell kernel
    const bool fixed = fixed_tiles(s, &bm, &bc, &br);
    if (s.cell || (!fixed && !(p->flags & SDR_SEARCH_TUNE))) {
        PSA_CHECK(hipEventDestroy(e0));
        PSA_CHECK(hipEventDestroy(e1));
        return;
    }
    // the first two blocks (one pair for the shapes whose power pass pairs
    // blocks)
    s.h.n_blk = std::min(nblk, 2);
    for (int i = 0; i < (fixed ? 0 : ns); i++) {
        for (int vc = 0; vc < nv; vc++) {
            for (int vr = 0; vr < nv; vr++) {
                if (h && !h_shape(s.h.M, m1[i], vc, vr)) continue;
                if (h) h_setup(p, s, m1[i], vc, vr);
                else f_setup(p, s, m1[i]);
                float t = 1e30f;
                for (int rep = 0; rep < 3; rep++) {
                    hipStream_t lst = p->lanes[s.lane].st;
                    p->cur_st = lst;
                    PSA_CHECK(hipEventRecord(e0, lst));
                    if (h) run_batched_h(p, s);
                    else run_batched(p, s);
                    PSA_CHECK(hipEventRecord(e1, lst));
                    PSA_CHECK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    PSA_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (rep > 0) t = std::min(t, ms);
                }
                if (t < best) { best = t; bm = m1[i]; bc = vc; br = vr; }
            }
        }
    }
    s.h.n_blk = nblk;
    if (h) s.hp = hfft_plan(s.h.M, s.hp->dt, bm);
    else s.pl = fft_plan(s.h.M, bm);
    s.hvar_col = bc;
    s.hvar_row = br;
    s.pair_blocks = h && h_col_pairs(s.hp->dt, s.hp->M1, bc) && !getenv("PSA_NO_PAIR");
    s.tuned_ms = fixed ? 0.0f : best;
    PSA_CHECK(hipEventDestroy(e0));
    PSA_CHECK(hipEventDestroy(e1));
}

// One run: the plan's stream forks to the lanes, every group is enqueued on
// its lane (its passes, then the peak search), and the lanes join back.
// Timed runs chain the groups one after the other, so that the per-kernel
// event times are those of kernels running alone.
static void run_all(sdr_search_plan_t *p)
{
    PSA_CHECK(hipEventRecord(p->fork, p->st));
    for (auto &ln : p->lanes) PSA_CHECK(hipStreamWaitEvent(ln.st, p->fork, 0));
    for (auto &s : p->sig) {
        const int gi = (int)(&s - p->sig.data());
        p->cur_sig = gi;
        hipStream_t st = p->lanes[s.lane].st;
        p->cur_st = st;
        if (p->timing && gi > 0) PSA_CHECK(hipStreamWaitEvent(st, p->gdone[gi - 1], 0));
        if (p->flags & SDR_SEARCH_PLAIN) run_plain(p, s);
        else if (s.hp) run_batched_h(p, s);
        else run_batched(p, s);
        {
            Tick t(p, K_PEAK, (double)s.h.n_prn * s.h.len_fds * s.h.N,
                   4.0 * s.h.n_prn * s.h.len_fds * s.h.N);
            const long long tot = (long long)s.h.len_fds * s.h.N;
            const bool fused = s.fused_peak && !(p->flags & SDR_SEARCH_PLAIN);
            if (!fused) {
                hipLaunchKernelGGL(peak_part_kernel, dim3(s.h.n_prn * s.nchunk), dim3(PEAK_THREADS), 0, st,
                                   s.P, tot, s.nchunk, s.part);
                PSA_CHECK(hipGetLastError());
            }
            const int nch = fused ? s.nchunk_fused : s.nchunk;
            hipLaunchKernelGGL(peak_final_kernel, dim3(s.h.n_prn), dim3(PEAK_THREADS), 0, st, s.P, s.h.N,
                               s.h.len_fds, s.h.T, nch, s.part, s.fds, s.cn0, s.ix, s.dop);
            PSA_CHECK(hipGetLastError());
        }
        PSA_CHECK(hipEventRecord(p->gdone[gi], st));
    }
    for (auto &ln : p->lanes) {
        PSA_CHECK(hipEventRecord(ln.done, ln.st));
        PSA_CHECK(hipStreamWaitEvent(p->st, ln.done, 0));
    }
    p->cur_st = p->st;
}

}  // namespace psa

// IF buffers: host data (copied to the plan's own device memory) or device
// data (read in place at every run; not owned)
static sdr_search_plan_t *plan_create(int n_buff, const uint8_t *const *data, const int *bN, bool dev,
                                      int n_sig, const sdr_search_sig_t *sigs, int flags)
{
    if (n_buff <= 0 || !data || !bN || n_sig <= 0 || !sigs) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new: invalid arguments\n");
        return nullptr;
    }
    for (int i = 0; i < n_buff; i++) {
        if (!data[i] || bN[i] <= 0) {
            fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new: invalid buffer %d\n", i);
            return nullptr;
        }
    }
    for (int i = 0; i < n_sig; i++) {
        const sdr_search_sig_t &g = sigs[i];
        if (g.buff < 0 || g.buff >= n_buff || g.M <= 0 || g.N <= 0 || g.N > g.M || g.n_prn <= 0 ||
            !g.code_fft || g.len_fds <= 0 || !g.fds || g.n_blk <= 0 || !fft_supported(g.M) ||
            g.M > bN[g.buff]) {
            fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new: invalid signal group %d\n", i);
            return nullptr;
        }
        for (int k = 0; k < g.n_blk; k++) {
            long long ix = (long long)g.ix0 + (long long)k * g.blk_step;
            if (ix < 0 || ix >= bN[g.buff]) {
                fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new: block %d of group %d outside the buffer\n", k, i);
                return nullptr;
            }
        }
    }
    ensure_init();
    sdr_search_plan_t *p = new sdr_search_plan_t();
    p->flags = flags;
    p->own_buf = !dev;
    PSA_CHECK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    for (int i = 0; i < n_buff; i++) {
        uint8_t *d = const_cast<uint8_t *>(data[i]);
        if (!dev) {
            pmalloc(&d, bN[i]);
            PSA_CHECK(hipMemcpy(d, data[i], bN[i], hipMemcpyHostToDevice));
        }
        p->d_buf.push_back(d);
        p->buffN.push_back(bN[i]);
    }
    // bytes per intermediate array (the inverse passes' Z is the largest);
    // PSA_CHUNK_MB overrides
    size_t budget = (size_t)4 << 30;
    if (const char *e = getenv("PSA_CHUNK_MB")) budget = (size_t)atol(e) << 20;
    const bool xcd_order = !getenv("PSA_NO_XCD_ORDER");

    p->sig.resize(n_sig);
    for (int i = 0; i < n_sig; i++) {
        SigDev &s = p->sig[i];
        s.h = sigs[i];
        s.h.code_fft = nullptr;
        s.h.fds = nullptr;
        s.pl = fft_plan(sigs[i].M);
        const int M = s.h.M, nprn = s.h.n_prn, nbin = s.h.len_fds, nblk = s.h.n_blk;
        const size_t spec = sizeof(float2) * (size_t)M;
        // code spectra: natural (plain) and digit-transposed (batched)
        float2 *nat;
        pmalloc(&nat, spec * nprn);
        PSA_CHECK(hipMemcpy(nat, sigs[i].code_fft, spec * nprn, hipMemcpyHostToDevice));
        // the precision form's last pass skips lags >= M / 2
        if (!(flags & SDR_SEARCH_PLAIN) && (flags & (SDR_SEARCH_FP16 | SDR_SEARCH_BF16)) && 2 * s.h.N <= M) {
            s.hp = hfft_plan(M, (flags & SDR_SEARCH_BF16) ? HDT_BF16 : HDT_F16);
            // FP16 lengths with a cell kernel take its split; PSA_NO_CELL
            // keeps the two-pass inverse
            if (s.hp && s.hp->dt == HDT_F16 && cell_split(M) && !getenv("PSA_NO_CELL")) {
                s.cell = true;
                s.hp = hfft_plan(M, HDT_F16, cell_split(M));
            }
        }
        if (flags & SDR_SEARCH_PLAIN) {
            s.cfn = nat;
        } else if (s.hp) {
            // converted for the chosen split below
            s.cf_nat = nat;
            pmalloc(&s.cfh, sizeof(uint32_t) * (size_t)M * nprn);
        } else {
            s.cf_nat = nat;
            pmalloc(&s.cft, spec * nprn);
        }
        pmalloc(&s.P, sizeof(float) * (size_t)nprn * nbin * s.h.N);
        pmalloc(&s.fds, sizeof(float) * nbin);
        PSA_CHECK(hipMemcpy(s.fds, sigs[i].fds, sizeof(float) * nbin, hipMemcpyHostToDevice));
        // mixing parameters per (block, bin), phi = 0 (src/sdr_func.c:674)
        PlainMix &pm = s.hm;
        const size_t nm = (size_t)nblk * nbin;
        pm.ix.resize(nm); pm.nseg.resize(nm); pm.p0.resize(nm); pm.p1.resize(nm); pm.s.resize(nm);
        for (int k = 0; k < nblk; k++) {
            int ix = s.h.ix0 + k * s.h.blk_step;
            for (int b = 0; b < nbin; b++) {
                size_t j = (size_t)k * nbin + b;
                pm.ix[j] = ix;
                mix_params(p->buffN[s.h.buff], ix, M, s.h.fs, s.h.fi + sigs[i].fds[b], 0.0,
                           &pm.p0[j], &pm.p1[j], &pm.nseg[j], &pm.s[j]);
            }
        }
        pmalloc(&s.mix_ix, sizeof(int) * nm);
        pmalloc(&s.mix_nseg, sizeof(int) * nm);
        pmalloc(&s.mix_p0, sizeof(uint32_t) * nm);
        pmalloc(&s.mix_p1, sizeof(uint32_t) * nm);
        pmalloc(&s.mix_s, sizeof(uint32_t) * nm);
        PSA_CHECK(hipMemcpy(s.mix_ix, pm.ix.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
        PSA_CHECK(hipMemcpy(s.mix_nseg, pm.nseg.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
        PSA_CHECK(hipMemcpy(s.mix_p0, pm.p0.data(), sizeof(uint32_t) * nm, hipMemcpyHostToDevice));
        PSA_CHECK(hipMemcpy(s.mix_p1, pm.p1.data(), sizeof(uint32_t) * nm, hipMemcpyHostToDevice));
        PSA_CHECK(hipMemcpy(s.mix_s, pm.s.data(), sizeof(uint32_t) * nm, hipMemcpyHostToDevice));
        pmalloc(&s.cn0, sizeof(float) * nprn);
        pmalloc(&s.ix, sizeof(int) * 2 * nprn);
        pmalloc(&s.dop, sizeof(double) * nprn);
        s.nchunk = (int)(((long long)nbin * s.h.N + PEAK_CHUNK - 1) / PEAK_CHUNK);
        // the last column pass emits one partial per workgroup when its
        // length has a specialised kernel
        s.fused_peak = (s.hp || spec_length(s.pl->M1)) && !(flags & SDR_SEARCH_PLAIN) &&
                       !getenv("PSA_NO_FUSED_PEAK");
        s.xcd_order = xcd_order;
        if (flags & SDR_SEARCH_PLAIN) {
            pmalloc(&s.part, sizeof(PeakPart) * (size_t)nprn * s.nchunk);
        }
        // chunking of the intermediates (one word per complex in the
        // precision form)
        const size_t zspec = s.hp ? spec / 2 : spec;
        // (equal chunks: no small remainder launch)
        auto balance = [](int n, int cmax) { const int k = (n + cmax - 1) / cmax; return (n + k - 1) / k; };
        s.bin_chunk = balance(nbin, (int)std::max<size_t>(1, std::min<size_t>(nbin, budget / zspec)));
        size_t pairs = std::max<size_t>(1, budget / zspec);
        s.prn_chunk = balance(nprn, (int)std::max<size_t>(1, std::min<size_t>(nprn, pairs / s.bin_chunk)));
        if (flags & SDR_SEARCH_PLAIN) s.prn_chunk = 1;
    }
    // lanes: the groups in decreasing order of work (FFT points over all
    // cells and blocks) each go to the lane with the least work so far;
    // PSA_LANES_N sets the lane count (1: every group in order on one stream)
    int nl = std::min(PSA_LANES, n_sig);
    if (const char *e = getenv("PSA_LANES_N")) nl = std::max(1, std::min(n_sig, atoi(e)));
    p->lanes.resize(nl);
    {
        std::vector<int> order(n_sig);
        std::vector<double> load(nl, 0.0);
        for (int i = 0; i < n_sig; i++) order[i] = i;
        auto work = [&](int i) {
            const sdr_search_sig_t &g = p->sig[i].h;
            return (double)g.M * g.n_prn * g.len_fds * g.n_blk;
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
        for (int i : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            p->sig[i].lane = l;
            load[l] += work(i);
        }
    }
    for (auto &s : p->sig) {
        SearchLane &ln = p->lanes[s.lane];
        const size_t M = (size_t)s.h.M;
        // (float2 elements; a word of the precision form is half of one)
        const size_t wd = s.hp ? 2 : 1;
        // forward spectra of every block and bin (cell groups need no Z)
        const size_t nfwd = fwd_all_blocks(s) ? (size_t)s.h.n_blk * s.h.len_fds : (size_t)s.bin_chunk;
        ln.tmp_elems = std::max(ln.tmp_elems, (nfwd * M + wd - 1) / wd);
        ln.x_elems = std::max(ln.x_elems, (nfwd * M + wd - 1) / wd);
        if (s.cell) continue;
        // room for two blocks' intermediates where the power pass may pair
        // blocks and the forward scratch (tmp) cannot hold them
        const size_t zc = (size_t)s.bin_chunk * s.prn_chunk * M;
        const bool may_pair = s.hp && s.h.n_blk >= 2 && fwd_all_blocks(s) && h_any_pairs(s.hp->dt, s.h.M);
        const size_t zn = may_pair && 2 * zc > nfwd * M ? 2 * zc : zc;
        ln.z_elems = std::max(ln.z_elems, (zn + wd - 1) / wd);
    }
    for (auto &ln : p->lanes) {
        PSA_CHECK(hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking));
        PSA_CHECK(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
        pmalloc(&ln.tmp, sizeof(float2) * ln.tmp_elems);
        pmalloc(&ln.X, sizeof(float2) * ln.x_elems);
        if (ln.z_elems) pmalloc(&ln.Z, sizeof(float2) * ln.z_elems);
        if (flags & SDR_SEARCH_PLAIN) pmalloc(&ln.iq16, 2 * ln.tmp_elems);
    }
    PSA_CHECK(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
    p->gdone.resize(n_sig);
    for (auto &e : p->gdone) PSA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    p->cur_st = p->st;
    // batched forms: tile selection (SDR_SEARCH_TUNE), then the code spectra
    // in the chosen split's order
    for (auto &s : p->sig) {
        if (!s.cf_nat) continue;
        p->cur_sig = (int)(&s - p->sig.data());
        if ((flags & SDR_SEARCH_TUNE) || getenv("PSA_TILES")) tune_group(p, s);
        if (s.hp) h_setup(p, s, s.hp->M1, s.hvar_col, s.hvar_row);
        else f_setup(p, s, s.pl->M1);
        pfree(s.cf_nat);
        s.cf_nat = nullptr;
    }
    return p;
}

extern "C" {

sdr_search_plan_t *sdr_search_plan_new(int n_buff, const sdr_buff_t *const *buffs,
    int n_sig, const sdr_search_sig_t *sigs, int flags)
{
    if (n_buff <= 0 || !buffs) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new: invalid arguments\n");
        return nullptr;
    }
    std::vector<const uint8_t *> data(n_buff);
    std::vector<int> bN(n_buff);
    for (int i = 0; i < n_buff; i++) {
        if (!buffs[i]) return plan_create(0, nullptr, nullptr, false, n_sig, sigs, flags);
        data[i] = buffs[i]->data;
        bN[i] = buffs[i]->N;
    }
    return plan_create(n_buff, data.data(), bN.data(), false, n_sig, sigs, flags);
}

sdr_search_plan_t *sdr_search_plan_new_dev(int n_buff, const uint8_t *const *dbuf, const int *buffN,
    int n_sig, const sdr_search_sig_t *sigs, int flags)
{
    return plan_create(n_buff, dbuf, buffN, true, n_sig, sigs, flags);
}

sdr_search_plan_t *sdr_search_plan_new_raw(int n_stream, const sdr_raw_stream_t *streams, int n_sig,
    const sdr_search_sig_t *sigs, int flags)
{
    if (n_stream <= 0 || !streams) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new_raw: invalid arguments\n");
        return nullptr;
    }
    if (flags & SDR_SEARCH_PLAIN) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new_raw: raw rings need a batched form\n");
        return nullptr;
    }
    std::vector<const uint8_t *> data(n_stream);
    std::vector<int> bN(n_stream);
    std::vector<sdr_search_plan_t::RawDesc> rd(n_stream);
    for (int i = 0; i < n_stream; i++) {
        const sdr_raw_stream_t &r = streams[i];
        // bytes per sample and the byte holding channel rfch (write_buff,
        // src/sdr_rcv.c:860-905): RAW8 2 ch / 1 B, RAW16 4 ch / 2 B (2 per
        // byte), RAW16I 8 ch / 2 B (4 per byte), RAW32 8 ch / 4 B (2 per byte)
        int bps = 0, nch = 0, per = 1;
        switch (r.fmt) {
        case 3: bps = 1; nch = 2; per = 2; break;
        case 4: bps = 2; nch = 4; per = 2; break;
        case 5: bps = 2; nch = 8; per = 4; break;
        case 6: bps = 4; nch = 8; per = 2; break;
        }
        if (!bps || r.rfch < 0 || r.rfch >= nch || !r.raw || !r.lut || r.N <= 0 ||
            (long long)r.N * bps > 0x7FFFFFFFLL || (r.rtoc && r.fmt == 5)) {
            fprintf(stderr, "pocketsdr_amd: sdr_search_plan_new_raw: invalid raw stream %d\n", i);
            return nullptr;
        }
        data[i] = r.raw;
        bN[i] = r.N;
        rd[i].bps = bps;
        rd[i].off = r.rfch / per;
        rd[i].rtoc = r.rtoc ? 1 : 0;
        rd[i].ph0 = r.ph0 & 3;
    }
    sdr_search_plan_t *p = plan_create(n_stream, data.data(), bN.data(), true, n_sig, sigs, flags);
    if (!p) return nullptr;
    for (int i = 0; i < n_stream; i++) {
        uint8_t lut[1024] = {0};
        memcpy(lut, streams[i].lut, streams[i].rtoc ? 1024 : 256);
        pmalloc(&rd[i].dlut, sizeof(lut));
        PSA_CHECK(hipMemcpy(rd[i].dlut, lut, sizeof(lut), hipMemcpyHostToDevice));
    }
    p->raw = rd;
    return p;
}

int sdr_search_plan_run(sdr_search_plan_t *p)
{
    if (!p) return 0;
    const bool graph = (p->flags & SDR_SEARCH_GRAPH) && !p->timing;
    if (graph) {
        if (!p->gexec) {
            hipGraph_t g;
            PSA_CHECK(hipStreamBeginCapture(p->st, hipStreamCaptureModeThreadLocal));
            run_all(p);
            PSA_CHECK(hipStreamEndCapture(p->st, &g));
            PSA_CHECK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
            PSA_CHECK(hipGraphDestroy(g));
        }
        PSA_CHECK(hipGraphLaunch(p->gexec, p->st));
        PSA_CHECK(hipStreamSynchronize(p->st));
        return 1;
    }
    p->ev_used = 0;
    run_all(p);
    PSA_CHECK(hipStreamSynchronize(p->st));
    if (p->timing) {
        if (p->stats.empty()) p->stats.resize((size_t)K_NCLS * p->sig.size());
        for (size_t i = 0; i < p->ev_used; i++) {
            float ms = 0;
            PSA_CHECK(hipEventElapsedTime(&ms, p->evs[i].a, p->evs[i].b));
            auto &q = p->stats[(size_t)p->evs[i].cls * p->sig.size() + p->evs[i].sig];
            q.ms += ms;
            q.cnt++;
            q.elems += p->evs[i].elems;
            q.bytes += p->evs[i].bytes;
            q.grid += p->evs[i].grid;
        }
    }
    return 1;
}

int sdr_search_plan_result(const sdr_search_plan_t *p, int sig, float *cn0, int *ix, double *dop)
{
    if (!p || sig < 0 || sig >= (int)p->sig.size()) return 0;
    const SigDev &s = p->sig[sig];
    const int n = s.h.n_prn;
    if (cn0) PSA_CHECK(hipMemcpy(cn0, s.cn0, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (ix) PSA_CHECK(hipMemcpy(ix, s.ix, sizeof(int) * 2 * n, hipMemcpyDeviceToHost));
    if (dop) PSA_CHECK(hipMemcpy(dop, s.dop, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 1;
}

int sdr_search_plan_power(const sdr_search_plan_t *p, int sig, int prn_idx, float *P)
{
    if (!p || sig < 0 || sig >= (int)p->sig.size()) return 0;
    const SigDev &s = p->sig[sig];
    if (prn_idx < 0 || prn_idx >= s.h.n_prn) return 0;
    size_t n = (size_t)s.h.len_fds * s.h.N;
    PSA_CHECK(hipMemcpy(P, s.P + (size_t)prn_idx * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    return 1;
}

long long sdr_search_plan_bytes(void)
{
    std::lock_guard<std::mutex> lk(g_pm_mtx);
    return g_pm_live;
}

int sdr_search_plan_form(const sdr_search_plan_t *p, int sig)
{
    if (!p || sig < 0 || sig >= (int)p->sig.size()) return -1;
    if (p->flags & SDR_SEARCH_PLAIN) return SDR_FORM_PLAIN;
    const SigDev &s = p->sig[sig];
    if (!s.hp) return SDR_FORM_FP32;
    return s.hp->dt == HDT_BF16 ? SDR_FORM_BF16 : SDR_FORM_FP16;
}

int sdr_search_plan_tiles(const sdr_search_plan_t *p, int sig, int *tiles)
{
    if (!p || sig < 0 || sig >= (int)p->sig.size() || !tiles) return 0;
    const SigDev &s = p->sig[sig];
    tiles[0] = s.hp ? s.hp->M1 : s.pl->M1;
    tiles[1] = s.hp ? s.hp->M2 : s.pl->M2;
    tiles[2] = s.hvar_col;
    tiles[3] = s.hvar_row;
    return 1;
}

void sdr_search_plan_timing(sdr_search_plan_t *p, int on)
{
    if (!p) return;
    p->timing = on != 0;
    p->stats.assign((size_t)K_NCLS * p->sig.size(), sdr_search_plan_t::Stat());
}

// kernel class name, optionally "@<signal group>"; sums over the matching
// (class, group) statistics
static bool plan_stat(const sdr_search_plan_t *p, const char *name, sdr_search_plan_t::Stat *out)
{
    if (!p || !name || p->stats.empty()) return false;
    std::string n(name);
    int sig = -1;
    size_t at = n.find('@');
    if (at != std::string::npos) {
        sig = atoi(n.c_str() + at + 1);
        n = n.substr(0, at);
    }
    const size_t ns = p->sig.size();
    for (int i = 0; i < K_NCLS; i++) {
        if (n != k_names[i]) continue;
        sdr_search_plan_t::Stat t;
        for (size_t g = 0; g < ns; g++) {
            if (sig >= 0 && (int)g != sig) continue;
            const auto &q = p->stats[(size_t)i * ns + g];
            t.ms += q.ms; t.cnt += q.cnt; t.elems += q.elems; t.bytes += q.bytes; t.grid += q.grid;
        }
        if (!t.cnt) return false;
        *out = t;
        return true;
    }
    return false;
}

int sdr_search_plan_kernel_bytes(const sdr_search_plan_t *p, const char *name, double *avg_bytes)
{
    sdr_search_plan_t::Stat t;
    if (!plan_stat(p, name, &t)) return 0;
    if (avg_bytes) *avg_bytes = t.bytes / t.cnt;
    return 1;
}

int sdr_search_plan_kernel_grid(const sdr_search_plan_t *p, const char *name, double *avg_threads)
{
    sdr_search_plan_t::Stat t;
    if (!plan_stat(p, name, &t)) return 0;
    if (avg_threads) *avg_threads = t.grid / t.cnt;
    return 1;
}

int sdr_search_plan_kernel_ms(const sdr_search_plan_t *p, const char *name, double *avg_ms,
                              long long *count, double *elems)
{
    sdr_search_plan_t::Stat t;
    if (!plan_stat(p, name, &t)) return 0;
    if (avg_ms) *avg_ms = t.ms / t.cnt;
    if (count) *count = t.cnt;
    if (elems) *elems = t.elems / t.cnt;
    return 1;
}

void sdr_search_plan_free(sdr_search_plan_t *p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->st);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->own_buf) {
        for (auto d : p->d_buf) pfree(d);
    }
    for (auto &s : p->sig) {
        pfree(s.cft); pfree(s.cfn); pfree(s.cfh); pfree(s.cf_nat);
        pfree(s.P); pfree(s.fds);
        pfree(s.mix_ix); pfree(s.mix_nseg); pfree(s.mix_p0);
        pfree(s.mix_p1); pfree(s.mix_s);
        pfree(s.cn0); pfree(s.ix); pfree(s.dop); pfree(s.part);
    }
    for (auto &ln : p->lanes) {
        (void)hipStreamSynchronize(ln.st);
        pfree(ln.tmp); pfree(ln.X); pfree(ln.Z); pfree(ln.iq16);
        (void)hipEventDestroy(ln.done);
        (void)hipStreamDestroy(ln.st);
    }
    for (auto &e : p->gdone) (void)hipEventDestroy(e);
    for (auto &r : p->raw) pfree(r.dlut);
    if (p->fork) (void)hipEventDestroy(p->fork);
    for (auto &e : p->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    (void)hipStreamDestroy(p->st);
    delete p;
}

}  // extern "C"
