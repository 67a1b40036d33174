// psa_l6.hip - QZSS L6D / L6E FFT tracking and CSK decoding on gfx950
// (src/sdr_ch.c:627-664, 694-720).
//
// Per (channel, epoch) item the reference mixes the epoch's N samples from
// the code-offset sample i = floor(coff fs), FFT-correlates them against
// the code spectrum bank j = (int)((coff fs - i) SDR_N_CODES) (src/sdr_ch.c:
// 256-262: N-point circular spectra, the code shifted by -j / 10 sample) and
// decodes the CSK symbol from the correlation: the peak of |corr| over code
// shifts of -255 ... 255 chips, interpolated linearly between samples.
//
// The plan batches every item of every channel:
//   l6_state_kernel : open-loop state per channel (FP64, trk_open_loop),
//                     mixing parameters and spectrum index per item
//   fwd column pass : IF bytes, carrier mix (LUT semantics), * CSCALE,
//                     length-M1 FFTs, twiddle                     -> A
//   fused row pass  : length-M2 FFTs, * Cf[index] / N^2 (sdr_cpx_mul
//                     arithmetic), length-M2 inverse FFTs, twiddle -> B
//   inv column pass : length-M1 inverse FFTs; only the lags CSK reads
//                     (|lag| < win) are stored, unless every
//                     correlation is kept                         -> A
//   csk_kernel      : 511 interpolated powers, first maximum, symbol, taps
// The FFT passes are those of the batched search (psa_fft.hip) with the
// code spectrum chosen per item through PassArgs::cf_index; for row lengths
// without a specialised kernel the row pass runs as the separate forward
// and inverse passes.
//
// Closed loop (sdr_l6_plan_new_loop): the epochs run in order, each a batch
// of the channels' items (item e n_ch + c): l6_loop_kernel advances every
// channel's NCO and code offset by its loop's Doppler (psa_loop.h) and sets
// the item's mixing parameters and spectrum index, the FFT passes and
// csk_kernel run on the batch, and the next l6_loop_kernel launch feeds the
// CSK's taps to the loop update (FLL / PLL / DLL / C/N0 / loss, the P tap as
// the prompt, src/sdr_ch.c:717-763) before advancing the next epoch.
#include <math.h>
#include <string.h>
#include <string>
#include <algorithm>
#include "psa_loop.h"

namespace psa {

#define CSK_THREADS 64
#define CSK_NP 511              // code shifts -255 ... 255 (src/sdr_ch.c:647)

// Per-item state of the batch: the sdr_mix_carr parameters of the epoch's
// samples and the index of its code spectrum (channel * SDR_N_CODES + j).
struct L6Items {
    int *ix, *nseg, *cf;
    uint32_t *p0, *p1, *s;
};

// src/sdr_ch.c:705-712 after the open-loop advance: item c * n_epoch + e.
// The ring start (e N) mod buffN + i is taken mod buffN (the receiver passes
// at most buffN, where a start of buffN mixes the same samples as 0).
__global__ void __launch_bounds__(TRK_SCHUNK) l6_state_kernel(const TrkChDev *__restrict__ chs, int n_epoch,
                                                             int N, int buffN, L6Items it)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x;
    const TrkChDev ch = chs[c];
    trk_open_loop(ch, n_epoch, [&](int e, double ph, double coff, double fc) {
#pragma clang fp contract(off)
        const long long o = (long long)c * n_epoch + e;
        const double cs = coff * ch.fs;           // one rounded product, used twice
        const int i = (int)floor(cs);
        int j = (int)((cs - i) * PSA_N_CODES);
        j = min(max(j, 0), PSA_N_CODES - 1);      // (cs - i) lies in [0, 1)
        long long st = ((long long)e * N) % buffN + i;
        st %= buffN;
        if (st < 0) st += buffN;
        const double phi = ph + fc * i / ch.fs;
        uint32_t p0, p1, s;
        int nseg;
        trk_mix_params(buffN, (int)st, N, ch.fs, fc, phi, p0, p1, nseg, s);
        it.ix[o] = (int)st;
        it.nseg[o] = nseg;
        it.p0[o] = p0;
        it.p1[o] = p1;
        it.s[o] = s;
        it.cf[o] = c * PSA_N_CODES + j;
    });
}

// Closed loop, one lane per channel: the loop update of epoch e (e >= 0)
// from the CSK's symbol and taps of the batch (item c of it), results to
// the channel-major outputs, then the advance to epoch e + 1 and its mixing
// parameters (src/sdr_ch.c:694-712).  A channel lost at or before epoch e
// records idle epochs (zero symbol and taps) and keeps its last mixing
// parameters for the batch's passes, whose outputs it ignores.
__global__ void l6_loop_kernel(TrkLoopCh *__restrict__ chs, int n_ch, int e, int n_epoch, int N, int buffN,
                               L6Items it, const uint8_t *__restrict__ sym_e, const float2 *__restrict__ taps_e,
                               int n_pos, uint8_t *__restrict__ sym_out, float2 *__restrict__ taps_out,
                               TrkLoopOut *__restrict__ out)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_ch) return;
    TrkLoopCh S = chs[c];
    if (e >= 0) {
        const long long oc = (long long)c * n_epoch + e;
        if (S.idle) {
            TrkLoopOut r;
            r.fd = S.fd; r.coff = S.coff; r.phi = S.phi; r.cn0 = S.cn0;
            r.lock = S.lock; r.idle = 1; r.P = make_float2(0.0f, 0.0f);
            out[oc] = r;
            sym_out[oc] = 0;
            for (int k = 0; k < n_pos; k++) taps_out[oc * n_pos + k] = make_float2(0.0f, 0.0f);
        } else {
            const float2 *tp = taps_e + (long long)c * n_pos;
            float2 tap[6];
            for (int k = 0; k < 6; k++) tap[k] = k < n_pos ? tp[k] : make_float2(0.0f, 0.0f);
            TrkEpoch q = {};
            TrkLoopOut r;
            tl_update_p(S, tap, tap[0].x, tap[0].y, r, nullptr, nullptr, q);
            out[oc] = r;
            sym_out[oc] = sym_e[c];
            for (int k = 0; k < n_pos; k++) taps_out[oc * n_pos + k] = tp[k];
        }
    }
    if (e + 1 < n_epoch) {
        if (!S.idle) tl_nco(S, nullptr);
        const double fc = S.fi + S.fd;
        const double cs = S.coff * S.fs;
        const int i = (int)floor(cs);
        int j = (int)((cs - i) * PSA_N_CODES);
        j = min(max(j, 0), PSA_N_CODES - 1);
        long long st = ((long long)(e + 1) * N) % buffN + i;
        st %= buffN;
        if (st < 0) st += buffN;
        const double phi = S.phi + fc * i / S.fs;
        uint32_t p0, p1, sv;
        int nseg;
        trk_mix_params(buffN, (int)st, N, S.fs, fc, phi, p0, p1, nseg, sv);
        const long long o = (long long)(e + 1) * n_ch + c;
        it.ix[o] = (int)st;
        it.nseg[o] = nseg;
        it.p0[o] = p0;
        it.p1[o] = p1;
        it.s[o] = sv;
        it.cf[o] = c * PSA_N_CODES + j;
    }
    chs[c] = S;
}

// position of lag k (0 <= k < N, natural order) in a correlation stored
// whole (win = 0) or as its lags -win ... win-1 at lag + win; the plan sizes
// win to hold every lag CSK reads, the clamp only keeps a bad tap position
// inside the buffer
__device__ __forceinline__ int csk_slot(int k, int N, int win)
{
    if (win == 0) return k;
    const int w = k < win ? k + win : k - (N - win);
    return min(max(w, 0), 2 * win - 1);
}

// interp_corr (src/sdr_ch.c:627-633) on the correlation rotated so that lag
// 0 sits at index n: C[k] = corr[(k - n) mod N], the reference's copy for
// 0 <= k < 2n and its circular extension outside.  The reference build
// contracts a0 C[i] + a1 C[i+1] to fma(a0, C[i], a1 C[i+1]) in double.
__device__ __forceinline__ float2 csk_interp(const float2 *__restrict__ corr, int N, int win, int n, double x)
{
#pragma clang fp contract(off)
    const int i = (int)x;                        // truncation toward zero, as (int)
    const double a1 = x - (double)i, a0 = 1.0 - a1;
    int k0 = (int)(((long long)i - n) % N);
    if (k0 < 0) k0 += N;
    const int k1 = k0 + 1 == N ? 0 : k0 + 1;
    const float2 c0 = corr[csk_slot(k0, N, win)], c1 = corr[csk_slot(k1, N, win)];
    return make_float2((float)fma(a0, (double)c0.x, a1 * (double)c1.x),
                       (float)fma(a0, (double)c0.y, a1 * (double)c1.y));
}

// CSK (src/sdr_ch.c:636-664), one wave per item.  Lane l evaluates shifts
// -255 + l + 64 k, x = fma(i, R, n) as the reference build contracts
// n + i R, and |c| = sqrtf(fmaf(re, re, im * im)) (sdr_cpx_abs,
// src/sdr_func.c:179-182, contracted likewise), correctly rounded.  The
// reference keeps the first strict maximum over values > 0 (ix = 0 when none
// is), which the wave reduction reproduces: largest value, smallest shift
// among equal ones.  A NaN power changes the meaning of the sequential test
// (P <= P_max fails for it), so a wave that sees one replays the reference
// loop over the 511 powers on lane 0.
__global__ void __launch_bounds__(CSK_THREADS) csk_kernel(const float2 *__restrict__ corr_all, long long corr_stride,
                                                         int N, int win, int n, double R,
                                                         const double *__restrict__ pos,
                                                         int n_pos, uint8_t *__restrict__ syms,
                                                         float2 *__restrict__ taps)
{
#pragma clang fp contract(off)
    __shared__ float sP[CSK_NP + 1];
    const long long item = blockIdx.x;
    const int lane = threadIdx.x;
    const float2 *corr = corr_all + item * corr_stride;
    const double dn = (double)n;
    float bm = 0.0f;
    int bi = 0x7FFFFFFF;
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int i = -255 + lane + 64 * k;
        if (i <= 255) {
            const float2 c = csk_interp(corr, N, win, n, fma((double)i, R, dn));
            const float P = __fsqrt_rn(__fmaf_rn(c.x, c.x, __fmul_rn(c.y, c.y)));
            sP[i + 255] = P;
            nan |= P != P;
            if (P > bm) { bm = P; bi = i; }       // i rises with k: first wins
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(bm, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (om > bm || (om == bm && oi < bi)) { bm = om; bi = oi; }
    }
    int ix = bm > 0.0f ? bi : 0;
    if (__any(nan)) {
        __syncthreads();
        if (lane == 0) {
            double pm = 0.0;
            int x = 0;
            for (int i = -255; i <= 255; i++) {
                const double P = (double)sP[i + 255];
                if (P <= pm) continue;
                pm = P;
                x = i;
            }
            sP[CSK_NP] = __int_as_float(x);
        }
        __syncthreads();
        ix = __float_as_int(sP[CSK_NP]);
    }
    if (lane == 0) syms[item] = (uint8_t)(255 - ix);
    const double base = fma((double)ix, R, dn);
    for (int t = lane; t < n_pos; t += CSK_THREADS) {
        taps[item * n_pos + t] = csk_interp(corr, N, win, n, base + pos[t]);
    }
}

// R = N / (len_code / 2) samples per chip, n = (int)(280 R) (src/sdr_ch.c:638-639)
static void csk_geometry(int N, int len_code, double *R, int *n)
{
    *R = (double)N / (len_code / 2);
    *n = (int)(280 * *R);
}

// half-width of the lag window that holds every lag CSK reads: |lag| <=
// 255 R + max |pos| + 2 (0: the window would not be smaller than N)
static int csk_window(int N, int len_code, const double *pos, int n_pos)
{
    double R;
    int n;
    csk_geometry(N, len_code, &R, &n);
    double pm = 0.0;
    for (int i = 0; i < n_pos; i++) pm = std::max(pm, fabs(pos[i]));
    const double w = floor(255.0 * R + pm) + 3.0;
    return 2.0 * w < (double)N ? (int)w : 0;
}

static void launch_csk(const float2 *corr, long long stride, int win, int n_item, int N, int len_code,
                       const double *pos, int n_pos, uint8_t *syms, float2 *taps, hipStream_t st)
{
    double R;
    int n;
    csk_geometry(N, len_code, &R, &n);
    hipLaunchKernelGGL(csk_kernel, dim3((unsigned)n_item), dim3(CSK_THREADS), 0, st, corr, stride, N, win, n,
                       R, pos, n_pos, syms, taps);
    PSA_CHECK(hipGetLastError());
}

enum { L6_STATE = 0, L6_FWD_COL, L6_ROW, L6_FWD_ROW, L6_INV_ROW, L6_INV_COL, L6_CSK, L6_LOOP, L6_NCLS };
static const char *l6_names[L6_NCLS] = {"state", "fwd_col", "row", "fwd_row", "inv_row", "inv_col", "csk",
                                        "loop"};

}  // namespace psa

using namespace psa;

struct sdr_l6_plan_tag {
    int n_ch = 0, n_pos = 0, n_epoch = 0, N = 0, buffN = 0, len_code = 0, flags = 0;
    int win = 0;                    // stored lag window of the correlations (0: whole)
    bool fused = false;             // fused row pass (specialised row length)
    long long n_item = 0;
    int chunk = 0;                  // items per batch of FFT passes
    const FFTPlan *pl = nullptr;
    uint8_t *d_buf = nullptr;
    bool own_buf = true;            // false: the IF ring is the caller's device memory
    float2 *d_cf = nullptr;         // n_ch * SDR_N_CODES spectra, digit-transposed
    TrkChDev *d_ch = nullptr;
    double *d_pos = nullptr;
    L6Items it = {};
    float2 *A = nullptr, *B = nullptr, *corr = nullptr;   // corr: all items (KEEP_CORR)
    uint8_t *d_sym = nullptr;
    float2 *d_taps = nullptr;
    // closed loop: initial channel states, their working copies, per-epoch
    // results, and the CSK outputs of one epoch's batch
    bool loop = false;
    std::vector<TrkLoopCh> lch0;
    TrkLoopCh *d_lch = nullptr;
    TrkLoopOut *d_lout = nullptr;
    uint8_t *d_sym_e = nullptr;
    float2 *d_taps_e = nullptr;
    hipStream_t st = nullptr;
    hipGraphExec_t gexec = nullptr;     // the closed loop's launches, captured once
    bool timing = false;
    struct Ev { int cls; hipEvent_t a, b; };
    std::vector<Ev> evs;
    size_t ev_used = 0;
    double ms[L6_NCLS] = {}, elems[L6_NCLS] = {};
    long long cnt[L6_NCLS] = {};
};

namespace psa {

// event pair around one launch when timing is on
struct L6Tick {
    sdr_l6_plan_t *p;
    size_t slot = 0;
    L6Tick(sdr_l6_plan_t *p_, int cls) : p(p_)
    {
        if (!p->timing) return;
        if (p->ev_used == p->evs.size()) {
            sdr_l6_plan_t::Ev ev;
            PSA_CHECK(hipEventCreate(&ev.a));
            PSA_CHECK(hipEventCreate(&ev.b));
            p->evs.push_back(ev);
        }
        slot = p->ev_used++;
        p->evs[slot].cls = cls;
        PSA_CHECK(hipEventRecord(p->evs[slot].a, p->st));
    }
    ~L6Tick()
    {
        if (p->timing) PSA_CHECK(hipEventRecord(p->evs[slot].b, p->st));
    }
};

}  // namespace psa

// data: the IF ring of buffN cpx8 bytes, host memory (copied to the
// device) or, with dev, device memory read in place at every run
static sdr_l6_plan_t *l6_create(const uint8_t *data, int buffN, bool dev, int n_ch, const sdr_l6_ch_t *chs,
                                int n_pos, const double *pos, int n_epoch, int flags)
{
    if (!data || buffN <= 0 || n_ch <= 0 || n_epoch <= 0 || n_pos < 0 || n_pos > PSA_MAX_CORR || !chs ||
        (n_pos > 0 && !pos)) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: invalid arguments\n");
        return nullptr;
    }
    const int N = chs[0].N, len_code = chs[0].len_code;
    if (N <= 0 || N > buffN || len_code < 2) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: need 0 < N <= buffer length, len_code >= 2\n");
        return nullptr;
    }
    for (int c = 0; c < n_ch; c++) {
        if (chs[c].N != N || chs[c].len_code != len_code || !chs[c].code_fft) {
            fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: channels need one N, one len_code and code spectra\n");
            return nullptr;
        }
    }
    // PSA_L6_SPLIT = M1 picks another of the FFT's splits M = M1 x M2
    const FFTPlan *pl = fft_plan(N);
    if (const char *e = getenv("PSA_L6_SPLIT")) {
        int cand[8];
        const int ns = fft_splits(N, cand, 8);
        const int m1 = atoi(e);
        if (std::find(cand, cand + ns, m1) != cand + ns) pl = fft_plan(N, m1);
    }
    if (!pl) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: FFT length %d has prime factors other than 2, 3, 5\n", N);
        return nullptr;
    }
    const long long n_item = (long long)n_ch * n_epoch;
    const size_t item_bytes = sizeof(float2) * (size_t)N;
    if ((flags & SDR_L6_KEEP_CORR) && (double)n_item * item_bytes > 64.0 * (1ull << 30)) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: kept correlations exceed 64 GiB\n");
        return nullptr;
    }
    ensure_init();
    sdr_l6_plan_t *p = new sdr_l6_plan_t();
    p->n_ch = n_ch; p->n_pos = n_pos; p->n_epoch = n_epoch; p->N = N; p->buffN = buffN;
    p->len_code = len_code; p->flags = flags; p->n_item = n_item; p->pl = pl;
    // two work buffers of `chunk` transforms each, 2 GiB together at most
    const long long cap = std::max<long long>(1, (2ll << 30) / (long long)(2 * item_bytes));
    p->chunk = (int)std::min<long long>(n_item, std::min<long long>(cap, 1 << 20));
    // PSA_L6_CHUNK caps the items per batch of passes (smaller batches)
    if (const char *e = getenv("PSA_L6_CHUNK")) {
        const long long c = atoll(e);
        if (c > 0) p->chunk = (int)std::min<long long>(p->chunk, c);
    }
    PSA_CHECK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    if (dev) {
        p->d_buf = const_cast<uint8_t *>(data);
        p->own_buf = false;
    } else {
        PSA_CHECK(hipMalloc(&p->d_buf, buffN));
        PSA_CHECK(hipMemcpy(p->d_buf, data, buffN, hipMemcpyHostToDevice));
    }
    // code spectra: natural order from the caller, digit-transposed on the device
    const size_t nspec = (size_t)n_ch * PSA_N_CODES;
    PSA_CHECK(hipMalloc(&p->d_cf, nspec * item_bytes));
    float2 *stage = nullptr;
    PSA_CHECK(hipMalloc(&stage, nspec * item_bytes));
    for (int c = 0; c < n_ch; c++) {
        PSA_CHECK(hipMemcpy(stage + (size_t)c * PSA_N_CODES * N, chs[c].code_fft,
                            PSA_N_CODES * item_bytes, hipMemcpyHostToDevice));
    }
    fft_to_transposed(pl, stage, p->d_cf, (int)nspec, p->st);
    PSA_CHECK(hipStreamSynchronize(p->st));
    PSA_CHECK(hipFree(stage));
    std::vector<TrkChDev> hc(n_ch);
    for (int c = 0; c < n_ch; c++) {
        hc[c] = TrkChDev{chs[c].fs, chs[c].fi, chs[c].fc_sig, chs[c].T, chs[c].fd0,
                         chs[c].drift, chs[c].phi0, chs[c].coff0};
    }
    PSA_CHECK(hipMalloc(&p->d_ch, sizeof(TrkChDev) * n_ch));
    PSA_CHECK(hipMemcpy(p->d_ch, hc.data(), sizeof(TrkChDev) * n_ch, hipMemcpyHostToDevice));
    PSA_CHECK(hipMalloc(&p->d_pos, sizeof(double) * std::max(n_pos, 1)));
    if (n_pos > 0) PSA_CHECK(hipMemcpy(p->d_pos, pos, sizeof(double) * n_pos, hipMemcpyHostToDevice));
    const size_t ni = (size_t)n_item;
    PSA_CHECK(hipMalloc(&p->it.ix, sizeof(int) * ni));
    PSA_CHECK(hipMalloc(&p->it.nseg, sizeof(int) * ni));
    PSA_CHECK(hipMalloc(&p->it.cf, sizeof(int) * ni));
    PSA_CHECK(hipMalloc(&p->it.p0, sizeof(uint32_t) * ni));
    PSA_CHECK(hipMalloc(&p->it.p1, sizeof(uint32_t) * ni));
    PSA_CHECK(hipMalloc(&p->it.s, sizeof(uint32_t) * ni));
    PSA_CHECK(hipMalloc(&p->A, (size_t)p->chunk * item_bytes));
    PSA_CHECK(hipMalloc(&p->B, (size_t)p->chunk * item_bytes));
    if (flags & SDR_L6_KEEP_CORR) PSA_CHECK(hipMalloc(&p->corr, ni * item_bytes));
    else p->win = csk_window(N, len_code, pos, n_pos);
    p->fused = spec_length(pl->M2) && !getenv("PSA_L6_UNFUSED");
    PSA_CHECK(hipMalloc(&p->d_sym, ni));
    PSA_CHECK(hipMalloc(&p->d_taps, sizeof(float2) * ni * std::max(n_pos, 1)));
    return p;
}

extern "C" {

sdr_l6_plan_t *sdr_l6_plan_new(const sdr_buff_t *buff, int n_ch, const sdr_l6_ch_t *chs,
    int n_pos, const double *pos, int n_epoch, int flags)
{
    if (!buff) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new: invalid arguments\n");
        return nullptr;
    }
    return l6_create(buff->data, buff->N, false, n_ch, chs, n_pos, pos, n_epoch, flags);
}

sdr_l6_plan_t *sdr_l6_plan_new_dev(const uint8_t *dbuf, int buffN, int n_ch, const sdr_l6_ch_t *chs,
    int n_pos, const double *pos, int n_epoch, int flags)
{
    return l6_create(dbuf, buffN, true, n_ch, chs, n_pos, pos, n_epoch, flags);
}

sdr_l6_plan_t *sdr_l6_plan_new_loop(const sdr_buff_t *buff, int n_ch, const sdr_l6_ch_t *chs,
    const sdr_trk_loop_t *loops, int n_pos, const double *pos, int n_epoch, int flags)
{
    if (!buff || !loops || !chs || n_ch <= 0 || n_pos < TL_NTAP) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new_loop: need loop parameters and the P, E, L, N taps\n");
        return nullptr;
    }
    for (int c = 0; c < n_ch; c++) {
        if (loops[c].len_code <= 0 || !(chs[c].T > 0.0) || (int)(TL_T_CN0 / chs[c].T) <= 0) {
            fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new_loop: need len_code > 0 and 0 < T <= 0.5 s\n");
            return nullptr;
        }
        if (loops[c].len_sec_code >= 2 || (loops[c].bump_step > 0.0 && n_pos < 6)) {
            fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new_loop: L6 channels take no secondary code; "
                            "bump-jump needs the VE and VL taps (4, 5)\n");
            return nullptr;
        }
    }
    sdr_l6_plan_t *p = l6_create(buff->data, buff->N, false, n_ch, chs, n_pos, pos, n_epoch, flags);
    if (!p) return nullptr;
    if (p->chunk < n_ch) {
        fprintf(stderr, "pocketsdr_amd: sdr_l6_plan_new_loop: %d channels exceed one batch of %d\n", n_ch,
                p->chunk);
        sdr_l6_plan_free(p);
        return nullptr;
    }
    p->loop = true;
    p->lch0.assign(n_ch, TrkLoopCh{});
    for (int c = 0; c < n_ch; c++) {
        // start_track (src/sdr_ch.c:378-394) and trk_init (:364-377)
        TrkLoopCh &S = p->lch0[c];
        S.fs = chs[c].fs; S.fi = chs[c].fi; S.fc_sig = chs[c].fc_sig; S.T = chs[c].T;
        S.time = 0.0; S.fd = chs[c].fd0; S.coff = chs[c].coff0; S.phi = chs[c].phi0;
        S.cn0 = loops[c].cn0; S.pli = 1.0;
        S.len_code = loops[c].len_code; S.costas = loops[c].costas ? 1 : 0;
        S.bump_step = loops[c].bump_step > 0.0 ? loops[c].bump_step : 0.0;
    }
    const size_t ni = (size_t)p->n_item;
    PSA_CHECK(hipMalloc(&p->d_lch, sizeof(TrkLoopCh) * n_ch));
    PSA_CHECK(hipMalloc(&p->d_lout, sizeof(TrkLoopOut) * ni));
    PSA_CHECK(hipMalloc(&p->d_sym_e, (size_t)n_ch));
    PSA_CHECK(hipMalloc(&p->d_taps_e, sizeof(float2) * (size_t)n_ch * n_pos));
    return p;
}

int sdr_l6_plan_state(const sdr_l6_plan_t *p, double *state)
{
    if (!p || !state || !p->loop) return 0;
    const size_t ne = (size_t)p->n_item;
    std::vector<TrkLoopOut> r(ne);
    PSA_CHECK(hipMemcpy(r.data(), p->d_lout, sizeof(TrkLoopOut) * ne, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ne; i++) {
        double *q = state + 6 * i;
        q[0] = r[i].fd; q[1] = r[i].coff; q[2] = r[i].phi; q[3] = r[i].cn0;
        q[4] = r[i].lock; q[5] = r[i].idle;
    }
    return 1;
}

}  // extern "C"

// the FFT passes and CSK of items i0 ... i0 + nb - 1; symbols and taps to
// sym / taps (item i0 first)
static void l6_batch(sdr_l6_plan_t *p, long long i0, int nb, uint8_t *sym, float2 *taps)
{
    hipStream_t st = p->st;
    const FFTPlan *pl = p->pl;
    const int M = pl->M;
    {
        PassArgs a = {};
        args_fwd_col(pl, a);
        a.buf = p->d_buf; a.buffN = p->buffN; a.mix_compact = 0;
        a.mix_ix = p->it.ix + i0; a.mix_nseg = p->it.nseg + i0;
        a.mix_p0 = p->it.p0 + i0; a.mix_p1 = p->it.p1 + i0; a.mix_s = p->it.s + i0;
        a.out = p->A; a.out_bstride = M;
        {
            L6Tick t(p, L6_FWD_COL);
            launch_pass(LD_MIX, ST_CPX, a, nb, st);
        }
        float2 *Z, *spare;                       // inverse row output, free buffer
        if (p->fused) {
            PassArgs x = {};
            args_xcorr_row(pl, x);
            x.in = p->A; x.in_bstride = M; x.out = p->B; x.out_bstride = M;
            x.cf = p->d_cf; x.cf_bstride = M; x.cf_index = p->it.cf + i0; x.prod_cdiv = 1;
            set_prod(pl, x, 1.0f / M / M);        // sdr_corr_fft: 1.0f / N / N
            L6Tick t(p, L6_ROW);
            if (!launch_xcorr_row(x, nb, st)) PSA_FAIL("sdr_l6_plan_run: no fused row pass for M2 = %d", pl->M2);
            Z = p->B;
            spare = p->A;
        } else {
            PassArgs r = {};
            args_fwd_row(pl, r);
            r.in = p->A; r.out = p->B; r.in_bstride = r.out_bstride = M;
            {
                L6Tick t(p, L6_FWD_ROW);
                launch_pass(LD_CPX, ST_CPX, r, nb, st);
            }
            PassArgs ir = {};
            args_inv_row(pl, ir);
            ir.in = p->B; ir.in_bstride = M;
            ir.cf = p->d_cf; ir.cf_bstride = M; ir.cf_index = p->it.cf + i0;
            ir.prod_xdiv = 1; ir.prod_xmod = nb; ir.prod_cdiv = 1;
            set_prod(pl, ir, 1.0f / M / M);
            ir.out = p->A; ir.out_bstride = M;
            {
                L6Tick t(p, L6_INV_ROW);
                launch_pass(LD_PROD, ST_CPX, ir, nb, st);
            }
            Z = p->A;
            spare = p->B;
        }
        // correlations: all kept, or the lag window CSK reads
        float2 *corr = p->corr ? p->corr + i0 * M : spare;
        const long long cstride = p->win > 0 ? 2LL * p->win : M;
        PassArgs ic = {};
        args_inv_col(pl, ic);
        ic.in = Z; ic.in_bstride = M; ic.out = corr; ic.out_bstride = cstride; ic.out_win = p->win;
        {
            L6Tick t(p, L6_INV_COL);
            launch_pass(LD_CPX, ST_CPX, ic, nb, st);
        }
        {
            L6Tick t(p, L6_CSK);
            launch_csk(corr, cstride, p->win, nb, p->N, p->len_code, p->d_pos, p->n_pos, sym, taps, st);
        }
    }
}

static void l6_loop_launch(sdr_l6_plan_t *p, int e)
{
    L6Tick t(p, L6_LOOP);
    hipLaunchKernelGGL(l6_loop_kernel, dim3((p->n_ch + 63) / 64), dim3(64), 0, p->st, p->d_lch, p->n_ch, e,
                       p->n_epoch, p->N, p->buffN, p->it, p->d_sym_e, p->d_taps_e, p->n_pos, p->d_sym,
                       p->d_taps, p->d_lout);
    PSA_CHECK(hipGetLastError());
}

// the closed loop's launches: the initial loop update, then per epoch the
// FFT passes and CSK of the epoch's batch and the loop update
static void l6_loop_body(sdr_l6_plan_t *p)
{
    l6_loop_launch(p, -1);
    for (int e = 0; e < p->n_epoch; e++) {
        l6_batch(p, (long long)e * p->n_ch, p->n_ch, p->d_sym_e, p->d_taps_e);
        l6_loop_launch(p, e);
    }
}

static void l6_run(sdr_l6_plan_t *p)
{
    hipStream_t st = p->st;
    if (p->loop) {
        // every run starts from the channels' initial states
        PSA_CHECK(hipMemcpyAsync(p->d_lch, p->lch0.data(), sizeof(TrkLoopCh) * p->n_ch, hipMemcpyHostToDevice,
                                 st));
        // the 5 n_epoch + 1 launches (every argument fixed at plan creation,
        // the loop state in device memory) replayed as one hipGraph, so that
        // no host enqueue sits between the small per-epoch kernels; timed
        // runs launch them one by one (PSA_L6_NO_GRAPH=1 too)
        static const bool no_graph = getenv("PSA_L6_NO_GRAPH") != nullptr;
        if (p->timing || no_graph) {
            l6_loop_body(p);
            return;
        }
        if (!p->gexec) {
            hipGraph_t g;
            PSA_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            l6_loop_body(p);
            PSA_CHECK(hipStreamEndCapture(st, &g));
            PSA_CHECK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
            PSA_CHECK(hipGraphDestroy(g));
        }
        PSA_CHECK(hipGraphLaunch(p->gexec, st));
        return;
    }
    {
        L6Tick t(p, L6_STATE);
        hipLaunchKernelGGL(l6_state_kernel, dim3(p->n_ch), dim3(TRK_SCHUNK), 0, st, p->d_ch, p->n_epoch, p->N,
                           p->buffN, p->it);
        PSA_CHECK(hipGetLastError());
    }
    for (long long i0 = 0; i0 < p->n_item; i0 += p->chunk) {
        const int nb = (int)std::min<long long>(p->chunk, p->n_item - i0);
        l6_batch(p, i0, nb, p->d_sym + i0, p->d_taps + i0 * p->n_pos);
    }
}

extern "C" {

int sdr_l6_plan_run(sdr_l6_plan_t *p)
{
    if (!p) return 0;
    p->ev_used = 0;
    l6_run(p);
    PSA_CHECK(hipStreamSynchronize(p->st));
    if (p->timing) {
        const double M = p->pl->M;
        for (size_t i = 0; i < p->ev_used; i++) {
            float ms = 0;
            PSA_CHECK(hipEventElapsedTime(&ms, p->evs[i].a, p->evs[i].b));
            p->ms[p->evs[i].cls] += ms;
        }
        bool seen[L6_NCLS] = {};
        for (size_t i = 0; i < p->ev_used; i++) seen[p->evs[i].cls] = true;
        for (int c = 0; c < L6_NCLS; c++) {
            if (!seen[c]) continue;
            p->cnt[c]++;
            p->elems[c] += c == L6_STATE || c == L6_LOOP ? (double)p->n_item : (double)p->n_item * M;
        }
    }
    return 1;
}

int sdr_l6_plan_result(const sdr_l6_plan_t *p, uint8_t *syms, sdr_cpx_t *C)
{
    if (!p) return 0;
    const size_t ni = (size_t)p->n_item;
    if (syms) PSA_CHECK(hipMemcpy(syms, p->d_sym, ni, hipMemcpyDeviceToHost));
    if (C && p->n_pos > 0) {
        PSA_CHECK(hipMemcpy(C, p->d_taps, sizeof(float2) * ni * p->n_pos, hipMemcpyDeviceToHost));
    }
    return 1;
}

int sdr_l6_plan_corr(const sdr_l6_plan_t *p, int ch, int epoch, sdr_cpx_t *corr)
{
    if (!p || !p->corr || !corr || ch < 0 || ch >= p->n_ch || epoch < 0 || epoch >= p->n_epoch) return 0;
    const size_t item = p->loop ? (size_t)epoch * p->n_ch + ch : (size_t)ch * p->n_epoch + epoch;
    const size_t o = item * p->pl->M;
    PSA_CHECK(hipMemcpy(corr, p->corr + o, sizeof(float2) * p->N, hipMemcpyDeviceToHost));
    return 1;
}

void sdr_l6_plan_timing(sdr_l6_plan_t *p, int on)
{
    if (!p) return;
    p->timing = on != 0;
    for (int c = 0; c < L6_NCLS; c++) { p->ms[c] = 0; p->elems[c] = 0; p->cnt[c] = 0; }
}

int sdr_l6_plan_kernel_ms(const sdr_l6_plan_t *p, const char *name, double *avg_ms, long long *count,
                          double *elems)
{
    if (!p || !name) return 0;
    for (int c = 0; c < L6_NCLS; c++) {
        if (strcmp(name, l6_names[c]) || !p->cnt[c]) continue;
        if (avg_ms) *avg_ms = p->ms[c] / p->cnt[c];
        if (count) *count = p->cnt[c];
        if (elems) *elems = p->elems[c] / p->cnt[c];
        return 1;
    }
    return 0;
}

void sdr_l6_plan_free(sdr_l6_plan_t *p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->st);
    if (p->own_buf) (void)hipFree(p->d_buf);
    (void)hipFree(p->d_cf); (void)hipFree(p->d_ch); (void)hipFree(p->d_pos);
    (void)hipFree(p->it.ix); (void)hipFree(p->it.nseg); (void)hipFree(p->it.cf);
    (void)hipFree(p->it.p0); (void)hipFree(p->it.p1); (void)hipFree(p->it.s);
    (void)hipFree(p->A); (void)hipFree(p->B); (void)hipFree(p->corr);
    (void)hipFree(p->d_sym); (void)hipFree(p->d_taps);
    (void)hipFree(p->d_lch); (void)hipFree(p->d_lout); (void)hipFree(p->d_sym_e); (void)hipFree(p->d_taps_e);
    for (auto &e : p->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    (void)hipStreamDestroy(p->st);
    delete p;
}

int sdr_csk(const sdr_cpx_t *corr, int n_item, int N, int len_code, const double *pos, int n_pos,
            uint8_t *syms, sdr_cpx_t *C)
{
    if (!corr || n_item <= 0 || N <= 0 || len_code < 2 || n_pos < 0 || (n_pos > 0 && (!pos || !C)) ||
        !syms) {
        fprintf(stderr, "pocketsdr_amd: sdr_csk: invalid arguments\n");
        return 0;
    }
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    const size_t cb = sizeof(float2) * (size_t)N * n_item;
    const size_t tb = sizeof(float2) * (size_t)n_item * n_pos;
    float2 *dc = (float2 *)tc.s[SC_A].get(cb);
    uint8_t *dm = (uint8_t *)tc.s[SC_MISC].get(tb + sizeof(double) * (size_t)n_pos + (size_t)n_item + 64);
    float2 *dt = (float2 *)dm;
    double *dp = (double *)(dm + tb);
    uint8_t *ds = dm + tb + sizeof(double) * (size_t)n_pos;
    PSA_CHECK(hipMemcpyAsync(dc, corr, cb, hipMemcpyHostToDevice, st));
    if (n_pos > 0) PSA_CHECK(hipMemcpyAsync(dp, pos, sizeof(double) * n_pos, hipMemcpyHostToDevice, st));
    launch_csk(dc, N, 0, n_item, N, len_code, dp, n_pos, ds, dt, st);
    PSA_CHECK(hipMemcpyAsync(syms, ds, (size_t)n_item, hipMemcpyDeviceToHost, st));
    if (n_pos > 0) PSA_CHECK(hipMemcpyAsync(C, dt, tb, hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    return 1;
}

}  // extern "C"
