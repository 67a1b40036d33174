// psa_anylen.hip - DFT of any length N for the drop-in entry points whose
// reference computes every length with PocketFFT (sdr_cpx_fft, sdr_corr_fft,
// sdr_search_code; src/sdr_func.c:236-264, lib/pocketfft/pocketfft_wrap.cpp).
//
// Lengths whose prime factors are all 2, 3, 5 use the four-step FFT
// directly.  Any other N uses Bluestein's chirp-z algorithm (L. I. Bluestein,
// "A linear filtering approach to the computation of the discrete Fourier
// transform", 1970): with nk = (n^2 + k^2 - (k - n)^2) / 2 and
// c_m = exp(s i pi m^2 / N) (s = -1 forward, +1 backward),
//   X_k = sum_n x_n exp(s 2 pi i n k / N) = c_k sum_n (x_n c_n) conj(c_{k-n}),
// a linear convolution of length 2 N - 1 computed as a circular one of a
// 2-3-5-smooth length L >= 2 N - 1 with the four-step FFT:
//   a = x c (zero padded to L), b_m = conj(c_m) for |m| < N (circularly),
//   X = c * IFFT(FFT(a) FFT(b) / L)[0 .. N).
// The chirp angles use m^2 mod 2N in 64-bit integers, so they are exact
// before the double-precision cos / sin; FFT(b) / L is precomputed per N.
#include <math.h>
#include <map>
#include "psa_internal.h"

namespace psa {

__global__ void __launch_bounds__(256) chirp_pre_kernel(const float2 *__restrict__ x, int N, int L,
                                                         const float2 *__restrict__ c,
                                                         float2 *__restrict__ a, long long total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long long b = t / L;
    const int n = (int)(t - b * L);
    float2 v = make_float2(0.0f, 0.0f);
    if (n < N) {
        const float2 u = x[b * N + n], w = c[n];
        v = make_float2(__fsub_rn(__fmul_rn(u.x, w.x), __fmul_rn(u.y, w.y)),
                        __fadd_rn(__fmul_rn(u.x, w.y), __fmul_rn(u.y, w.x)));
    }
    a[t] = v;
}

__global__ void __launch_bounds__(256) chirp_mul_kernel(float2 *__restrict__ a, const float2 *__restrict__ bf,
                                                         int L, long long total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const float2 u = a[t], w = bf[t % L];
    a[t] = make_float2(__fsub_rn(__fmul_rn(u.x, w.x), __fmul_rn(u.y, w.y)),
                       __fadd_rn(__fmul_rn(u.x, w.y), __fmul_rn(u.y, w.x)));
}

__global__ void __launch_bounds__(256) chirp_post_kernel(const float2 *__restrict__ a, int N, int L,
                                                          const float2 *__restrict__ c,
                                                          float2 *__restrict__ X, long long total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long long b = t / N;
    const int k = (int)(t - b * N);
    const float2 u = a[b * L + k], w = c[k];
    X[t] = make_float2(__fsub_rn(__fmul_rn(u.x, w.x), __fmul_rn(u.y, w.y)),
                       __fadd_rn(__fmul_rn(u.x, w.y), __fmul_rn(u.y, w.x)));
}

__global__ void __launch_bounds__(256) cpx16_float_kernel(const char2 *__restrict__ iq, long long n,
                                                           float2 *__restrict__ out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const char2 v = iq[t];
    out[t] = make_float2(__fmul_rn((float)v.x, PSA_CSCALE), __fmul_rn((float)v.y, PSA_CSCALE));
}

static unsigned grid_of(long long n)
{
    const long long g = (n + 255) / 256;
    if (g > 0x7FFFFFFFLL) PSA_FAIL("any-length FFT launch too large");
    return (unsigned)g;
}

void launch_cpx16_to_float(const int8_t *iq, long long n, float2 *out, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(cpx16_float_kernel, dim3(grid_of(n)), dim3(256), 0, st,
                       reinterpret_cast<const char2 *>(iq), n, out);
    PSA_CHECK(hipGetLastError());
}

static std::mutex g_any_mtx;
static std::map<int, AnyPlan *> g_any;

const AnyPlan *any_plan(int N)
{
    if (N < 1) return nullptr;
    std::lock_guard<std::mutex> lk(g_any_mtx);
    auto it = g_any.find(N);
    if (it != g_any.end()) return it->second;
    // smallest smooth L >= 2N - 1 with a four-step split
    const FFTPlan *pl = nullptr;
    long long L = 2LL * N - 1;
    for (; L <= 4096LL * 4096; L++) {
        if (!fft_supported((int)L)) continue;
        pl = fft_plan((int)L);
        if (pl) break;
    }
    if (!pl) return nullptr;
    AnyPlan *ap = new AnyPlan();
    ap->N = N;
    ap->L = (int)L;
    ap->pl = pl;
    const double PI_D = 3.14159265358979323846;
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    std::vector<float2> hc(N), hb(L);
    float2 *dtmp;
    PSA_CHECK(hipMalloc(&dtmp, sizeof(float2) * (size_t)L));
    for (int dir = 0; dir < 2; dir++) {
        const double sg = dir == 0 ? -1.0 : 1.0;
        for (int m = 0; m < N; m++) {
            const long long q = (long long)m * m % (2LL * N);        // exact m^2 mod 2N
            const double ang = sg * PI_D * (double)q / N;
            hc[m] = make_float2((float)cos(ang), (float)sin(ang));
        }
        // b_m = conj(c_m) for |m| < N, zero elsewhere; stored /L after the FFT
        std::fill(hb.begin(), hb.end(), make_float2(0.0f, 0.0f));
        for (int m = 0; m < N; m++) {
            const long long q = (long long)m * m % (2LL * N);
            const double ang = -sg * PI_D * (double)q / N;
            const float2 v = make_float2((float)(cos(ang) / L), (float)(sin(ang) / L));
            hb[m] = v;
            if (m > 0) hb[L - m] = v;
        }
        PSA_CHECK(hipMalloc(&ap->chirp[dir], sizeof(float2) * (size_t)N));
        PSA_CHECK(hipMalloc(&ap->bf[dir], sizeof(float2) * (size_t)L));
        PSA_CHECK(hipMemcpy(ap->chirp[dir], hc.data(), sizeof(float2) * (size_t)N, hipMemcpyHostToDevice));
        PSA_CHECK(hipMemcpy(ap->bf[dir], hb.data(), sizeof(float2) * (size_t)L, hipMemcpyHostToDevice));
        // FFT(b) in place (the column pass writes dtmp, the row pass back)
        fft_c2c(pl, ap->bf[dir], ap->bf[dir], dtmp, 1, 0, st);
        PSA_CHECK(hipStreamSynchronize(st));
    }
    PSA_CHECK(hipFree(dtmp));
    g_any[N] = ap;
    return ap;
}

void fft_any(const AnyPlan *ap, const float2 *in, float2 *out, float2 *tmp, int nbatch, int dir,
             hipStream_t st)
{
    if (nbatch <= 0) return;
    const int N = ap->N, L = ap->L;
    const long long tl = (long long)nbatch * L, tn = (long long)nbatch * N;
    float2 *a = tmp, *w = tmp + tl;
    hipLaunchKernelGGL(chirp_pre_kernel, dim3(grid_of(tl)), dim3(256), 0, st, in, N, L, ap->chirp[dir], a, tl);
    PSA_CHECK(hipGetLastError());
    fft_c2c(ap->pl, a, a, w, nbatch, 0, st);
    hipLaunchKernelGGL(chirp_mul_kernel, dim3(grid_of(tl)), dim3(256), 0, st, a, ap->bf[dir], L, tl);
    PSA_CHECK(hipGetLastError());
    fft_c2c(ap->pl, a, a, w, nbatch, 1, st);
    hipLaunchKernelGGL(chirp_post_kernel, dim3(grid_of(tn)), dim3(256), 0, st, a, N, L, ap->chirp[dir], out, tn);
    PSA_CHECK(hipGetLastError());
}

}  // namespace psa
