// psa_noalloc.hip - sdr_search_code (src/sdr_func.c:664-687) on a fixed
// device scratch array, for callers that cannot allocate device memory
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// The scratch is one file-scope device array of PSA_NOALLOC_BYTES per GPU,
// held from the library's load on that GPU until the process ends.  Its
// device address is looked up once per device with hipGetSymbolAddress (in
// sdr_func_init for the current device, else on the entry point's first call
// on a device) and only that address is passed to kernels and copies.  Every
// call on one GPU shares the array and leaves its contents behind: calls that
// overlap in time on one GPU race and must be serialised by the caller.
//
// The dataflow, kernels, FFT split and twiddle values are those of
// sdr_search_code, so the power surfaces are bit-identical to it; bins are
// processed in chunks sized to the scratch.
#include <math.h>
#include <string.h>
#include <algorithm>
#include "psa_internal.h"

#define PSA_NOALLOC_BYTES ((size_t)64 << 20)   // 64 MiB per GPU
#define PSA_NOALLOC_MAX_DEV 64
#define PSA_NOALLOC_MAX_CHUNK 256              // bins per pass (host parameter staging)

namespace psa {

static __device__ __align__(256) uint8_t scratch[PSA_NOALLOC_BYTES];

static std::mutex g_na_mtx;
static uint8_t *g_na_ptr[PSA_NOALLOC_MAX_DEV];   // device address of scratch, per device
static int g_na_tabM[PSA_NOALLOC_MAX_DEV];       // FFT length whose tables the array holds
static FFTPlan g_na_plan[PSA_NOALLOC_MAX_DEV];   // its geometry and table addresses
static size_t g_na_ntab[PSA_NOALLOC_MAX_DEV];
// host staging of the FFT tables: 2 (M1 + M2) + 2 (M / 1024 + 1 + 1024)
// float2, M1, M2 <= 4096
static float2 g_na_tab_host[4 * 4096 + 2 * (4096 + 1024)];

static size_t al256(size_t n) { return (n + 255) / 256 * 256; }

// current device's scratch address (looked up on the first call per device)
static uint8_t *scratch_ptr(int *dev_out)
{
    int dev = 0;
    PSA_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= PSA_NOALLOC_MAX_DEV) PSA_FAIL("device index %d out of range", dev);
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(g_na_mtx);
    if (!g_na_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            PSA_FAIL("no-allocation scratch: hipGetSymbolAddress failed (%s)", hipGetErrorName(e));
        }
        g_na_ptr[dev] = (uint8_t *)ptr;
        g_na_tabM[dev] = 0;
    }
    return g_na_ptr[dev];
}

void noalloc_setup()
{
    int dev;
    (void)scratch_ptr(&dev);
}

// bytes of the fixed part of the layout for FFT length M (tables, carrier,
// IF block, natural and digit-transposed code spectra) and per bin of a chunk
// (the four per-chunk buffers add at most 1 KiB of alignment padding)
static void layout_sizes(int M, size_t ntab, size_t *fixed, size_t *per_bin)
{
    *fixed = al256(sizeof(uint32_t) * 256) + al256(sizeof(float2) * ntab) + al256((size_t)M) +
             2 * al256(sizeof(float2) * (size_t)M);
    // mixing parameters (5 words), X work buffers (2 x M complex), P row
    *per_bin = 5 * sizeof(uint32_t) + 2 * sizeof(float2) * (size_t)M + sizeof(float) * (size_t)M;
}

}  // namespace psa

using namespace psa;

extern "C" int sdr_search_code_noalloc(const sdr_cpx_t *code_fft, double T, const sdr_buff_t *buff,
    int ix, int N, double fs, double fi, const float *fds, int len_fds, float *P)
{
    (void)T;
    if (N <= 0 || len_fds <= 0) return 1;
    if (!code_fft || !buff || !buff->data || !fds || !P || buff->N < N || ix < 0 || ix >= buff->N) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_code_noalloc: invalid arguments\n");
        return 0;
    }
    if (!init_done()) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_code_noalloc: call sdr_func_init first\n");
        return 0;
    }
    if (!fft_supported(N)) {
        fprintf(stderr, "pocketsdr_amd: sdr_search_code_noalloc: FFT length %d has prime factors "
                        "other than 2, 3, 5\n", N);
        return 0;
    }
    int dev;
    uint8_t *base = scratch_ptr(&dev);
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    fft_init_attrs();

    // FFT geometry and tables of the default split (the plan sdr_search_code
    // uses), staged into the front of the array when the length changes
    FFTPlan pl;
    size_t ntab;
    {
        std::lock_guard<std::mutex> lk(g_na_mtx);
        if (g_na_tabM[dev] != N) {
            float2 *dtab = (float2 *)(base + al256(sizeof(uint32_t) * 256));
            ntab = fft_plan_tables(N, &pl, g_na_tab_host, sizeof(g_na_tab_host) / sizeof(float2), dtab);
            if (!ntab) {
                fprintf(stderr, "pocketsdr_amd: sdr_search_code_noalloc: FFT length %d has no split\n", N);
                return 0;
            }
            size_t fixed, per_bin;
            layout_sizes(N, ntab, &fixed, &per_bin);
            if (fixed + per_bin + 1024 > PSA_NOALLOC_BYTES) {
                fprintf(stderr, "pocketsdr_amd: sdr_search_code_noalloc: FFT length %d exceeds the scratch "
                                "(%zu bytes for one bin, %zu available)\n", N, fixed + per_bin,
                        (size_t)PSA_NOALLOC_BYTES);
                return 0;
            }
            g_na_tabM[dev] = 0;
            PSA_CHECK(hipMemcpyAsync(base, host_carrier(), sizeof(uint32_t) * 256, hipMemcpyHostToDevice, st));
            PSA_CHECK(hipMemcpyAsync(dtab, g_na_tab_host, sizeof(float2) * ntab, hipMemcpyHostToDevice, st));
            PSA_CHECK(hipStreamSynchronize(st));
            g_na_tabM[dev] = N;
            g_na_plan[dev] = pl;
            g_na_ntab[dev] = ntab;
        }
        pl = g_na_plan[dev];
        ntab = g_na_ntab[dev];
    }
    size_t fixed, per_bin;
    layout_sizes(N, ntab, &fixed, &per_bin);
    const int chunk = (int)std::min<size_t>({(size_t)len_fds, (size_t)PSA_NOALLOC_MAX_CHUNK,
                                             (PSA_NOALLOC_BYTES - fixed - 1024) / per_bin});
    const uint32_t *dcarr = (const uint32_t *)base;
    size_t o = al256(sizeof(uint32_t) * 256) + al256(sizeof(float2) * ntab);
    uint8_t *din = base + o;                        o += al256((size_t)N);
    float2 *cnat = (float2 *)(base + o);            o += al256(sizeof(float2) * (size_t)N);
    float2 *cft = (float2 *)(base + o);             o += al256(sizeof(float2) * (size_t)N);
    uint32_t *dprm = (uint32_t *)(base + o);        o += al256(5 * sizeof(uint32_t) * (size_t)chunk);
    float2 *ta = (float2 *)(base + o);              o += al256(sizeof(float2) * (size_t)N * chunk);
    float2 *tb = (float2 *)(base + o);              o += al256(sizeof(float2) * (size_t)N * chunk);
    float *dP = (float *)(base + o);                o += al256(sizeof(float) * (size_t)N * chunk);
    if (o > PSA_NOALLOC_BYTES) PSA_FAIL("sdr_search_code_noalloc: scratch layout overflow");

    // IF block (the ring's two segments back to back) and the code spectrum
    const int n1 = std::min(N, buff->N - ix);
    PSA_CHECK(hipMemcpyAsync(din, buff->data + ix, n1, hipMemcpyHostToDevice, st));
    if (n1 < N) PSA_CHECK(hipMemcpyAsync(din + n1, buff->data, N - n1, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipMemcpyAsync(cnat, code_fft, sizeof(float2) * (size_t)N, hipMemcpyHostToDevice, st));
    fft_to_transposed(&pl, cnat, cft, 1, st);

    uint32_t hprm[5 * PSA_NOALLOC_MAX_CHUNK];
    for (int b0 = 0; b0 < len_fds; b0 += chunk) {
        const int nb = std::min(chunk, len_fds - b0);
        // mixing parameters per bin: phi = 0 (src/sdr_func.c:674)
        for (int b = 0; b < nb; b++) {
            uint32_t p0, p1, s;
            int nseg;
            mix_params(buff->N, ix, N, fs, fi + fds[b0 + b], 0.0, &p0, &p1, &nseg, &s);
            hprm[b] = 0;                    // start in the contiguous block
            hprm[nb + b] = (uint32_t)nseg;
            hprm[2 * nb + b] = p0;
            hprm[3 * nb + b] = p1;
            hprm[4 * nb + b] = s;
        }
        const size_t prm = sizeof(uint32_t) * (size_t)nb;
        const size_t pbytes = sizeof(float) * (size_t)N * nb;
        PSA_CHECK(hipMemcpyAsync(dprm, hprm, 5 * prm, hipMemcpyHostToDevice, st));
        PSA_CHECK(hipMemcpyAsync(dP, P + (size_t)b0 * N, pbytes, hipMemcpyHostToDevice, st));

        PassArgs a = {};
        args_fwd_col(&pl, a);
        a.carr = dcarr;
        a.buf = din; a.buffN = N; a.mix_compact = 1;
        a.mix_ix = (const int *)dprm; a.mix_nseg = (const int *)(dprm + nb);
        a.mix_p0 = dprm + 2 * nb; a.mix_p1 = dprm + 3 * nb; a.mix_s = dprm + 4 * nb;
        a.out = ta; a.out_bstride = N;
        launch_pass(LD_MIX, ST_CPX, a, nb, st);
        PassArgs r = {};
        args_fwd_row(&pl, r);
        r.carr = dcarr;
        r.in = ta; r.out = tb; r.in_bstride = r.out_bstride = N;
        launch_pass(LD_CPX, ST_CPX, r, nb, st);
        PassArgs ir = {};
        args_inv_row(&pl, ir);
        ir.carr = dcarr;
        ir.in = tb; ir.in_bstride = N; ir.cf = cft; ir.cf_bstride = 0;
        ir.prod_xdiv = 1; ir.prod_xmod = nb; ir.prod_cdiv = nb;
        set_prod(&pl, ir, 1.0f / N / N);
        ir.out = ta; ir.out_bstride = N;
        launch_pass(LD_PROD, ST_CPX, ir, nb, st);
        PassArgs ic = {};
        args_inv_col(&pl, ic);
        ic.carr = dcarr;
        ic.in = ta; ic.in_bstride = N;
        ic.P = dP; ic.p_bstride = N; ic.p_div = 1; ic.p_off1 = 0; ic.p_mul = 1; ic.p_off2 = 0;
        ic.nlag = N;
        launch_pass(LD_CPX, ST_POW, ic, nb, st);
        PSA_CHECK(hipMemcpyAsync(P + (size_t)b0 * N, dP, pbytes, hipMemcpyDeviceToHost, st));
        // the host staging of the next chunk reuses hprm
        PSA_CHECK(hipStreamSynchronize(st));
    }
    return 1;
}
