// psa_internal.h - shared definitions for the MI355X (gfx950) PocketSDR hot path.
//
// Layouts (identical to the reference ABI, src/pocket_sdr.h:105-113):
//   cpx8   1 byte / complex sample, I = low nibble, Q = high nibble (int4 each)
//   cpx16  {int8 I, int8 Q}
//   cpx    {float re, float im}
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <mutex>
#include <vector>

#include "../../include/pocket_sdr_amd.h"

#define PSA_CSCALE (1 / 11.2f)     // SDR_CSCALE, src/pocket_sdr.h:61
#define PSA_N_CODES 10             // SDR_N_CODES, src/pocket_sdr.h:60
#define PSA_MAX_CORR 107           // SDR_MAX_CORR, src/pocket_sdr.h:58

// Any HIP failure ends the process with a message, the reference's convention
// for allocation failures (src/sdr_cmn.c:43-46): no partial results.
#define PSA_CHECK(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            fprintf(stderr, "pocketsdr_amd: HIP error %s at %s:%d: %s\n",      \
                    hipGetErrorName(e_), __FILE__, __LINE__, #call);           \
            exit(-1);                                                          \
        }                                                                      \
    } while (0)

#define PSA_FAIL(...)                                                          \
    do {                                                                       \
        fprintf(stderr, "pocketsdr_amd: ");                                    \
        fprintf(stderr, __VA_ARGS__);                                          \
        fprintf(stderr, "\n");                                                 \
        exit(-1);                                                              \
    } while (0)

namespace psa {

// ---------------------------------------------------------------------------
// runtime (psa_runtime.hip)
// ---------------------------------------------------------------------------

// carrier table: round(11.2 cos/sin(-2 pi i / 256)) packed as
// (uint8 cI) | (uint8 cQ) << 8, src/sdr_func.c:122-124
void ensure_init();
bool init_done();                  // ensure_init has completed
// no-allocation entry point (psa_noalloc.hip): looks up the current
// device's scratch array once per device
void noalloc_setup();
const uint32_t *host_carrier();
const uint32_t *dev_carrier();     // the same 256 words in device memory

// Per-calling-thread HIP stream plus grow-only device scratch.  The reference
// hot path is called concurrently from one pthread per receiver channel
// (src/sdr_rcv.c:495-521), so every host thread owns its own stream and
// buffers; nothing global is locked while kernels run.
struct Scratch {
    void *ptr = nullptr;
    size_t size = 0;
    void *get(size_t bytes);
};

struct ThreadCtx {
    hipStream_t stream = nullptr;
    Scratch s[8];
    void *pinned = nullptr;
    size_t pinned_size = 0;
    void *host_pinned(size_t bytes);
};
ThreadCtx &tctx();

// scratch slot ids
enum { SC_IN = 0, SC_IQ, SC_CODE, SC_A, SC_B, SC_C, SC_P, SC_MISC };

// ---------------------------------------------------------------------------
// FFT (psa_fft.hip): four-step mixed-radix FP32 FFT, M = M1 * M2
// ---------------------------------------------------------------------------
struct SubPlan {
    int L = 0;              // sub-transform length
    int nrad = 0;
    int rad[16];            // radix sequence (product = L)
    float2 *tw_fwd = nullptr;   // exp(-2 pi i m / L), m < L (device)
    float2 *tw_bwd = nullptr;   // exp(+2 pi i m / L)
};

struct FFTPlan {
    int M = 0, M1 = 0, M2 = 0;
    SubPlan s1, s2;             // lengths M1 (column pass) and M2 (row pass)
    // inter-pass twiddle w_M^{a} = T_hi[a >> 10] * T_lo[a & 1023], a < M
    float2 *thi_fwd = nullptr, *tlo_fwd = nullptr;
    float2 *thi_bwd = nullptr, *tlo_bwd = nullptr;
};
bool fft_supported(int M);
// Geometry and tables of the default split of M without allocating: the
// tables (the values fft_plan(M) uploads) are written to host staging h
// (capacity cap float2) and p's table pointers point at the same offsets of
// the device array dev, to which the caller copies h.  Returns the float2
// count, 0 if M is unsupported or cap is too small.
size_t fft_plan_tables(int M, FFTPlan *p, float2 *h, size_t cap, float2 *dev);
// device table of scale * exp(sign 2 pi i m / L), m < L, rounded from double
float2 *make_twiddle(int L, int sign, double scale);
// cached per (M, M1), thread-safe; M1 = 0 takes the default split (rows a
// multiple of 16, then the most balanced); nullptr if unsupported
const FFTPlan *fft_plan(int M, int M1 = 0);
// column lengths of the splits of M: the default, then the other splits
// whose both lengths have specialised kernels; returns the count
int fft_splits(int M, int *m1, int nmax);
void fft_init_attrs();            // LDS limits of the pass kernels (once)

// partial peak reduction record: max value, its first row-major index within
// the PRN's surface (-1: none > 0), FP64 sum of the values
struct PeakPart {
    float m;
    int pad;
    long long i;
    double s;
};

// load / store modes of the sub-transform passes
enum LoadMode {
    LD_CPX = 0,      // float2 from global
    LD_CPX16 = 1,    // cpx16 int8 pairs -> float * CSCALE
    LD_MIX = 2,      // cpx8 IF samples, carrier-mixed through the LUT, * CSCALE
    LD_PROD = 3,     // (X * Cf) * s, the reference sdr_cpx_mul arithmetic
};
enum StoreMode {
    ST_CPX = 0,      // float2 to global (optionally times inter-pass twiddle)
    ST_POW = 1,      // P += |x|^2 (reference contraction P + fma(re, re, im*im))
};

// Arguments shared by every pass kernel.  A launch processes `nbatch`
// transforms; batch index b selects the operands through the *_bstride /
// *_bdiv fields as documented in psa_fft.hip.
struct PassArgs {
    // geometry
    int L;                  // sub-transform length
    int G;                  // sub-transforms per workgroup
    int nsub;               // sub-transforms per batch item (columns or rows)
    long long in_nstride, in_gstride;   // element strides in the input
    long long out_nstride, out_gstride;
    int dir;                // 0 forward (exp -), 1 backward (exp +)
    const float2 *tw;       // sub-transform twiddles for dir
    // inter-pass twiddle (applied on store when thi != nullptr):
    // factor w_M^{(n_idx * g_idx) mod M}
    const float2 *thi, *tlo;
    int M;
    // plain complex input / output
    const float2 *in;
    float2 *out;
    long long in_bstride, out_bstride;
    int out_win;                // ST_CPX: > 0 stores only the lags -out_win ... out_win - 1
                                // of a natural-order output, at index lag + out_win
    const float2 *tw_inv;       // fused correlator row pass: inverse sub-transform twiddles
    // LD_MIX: IF ring buffer and per-batch mixing parameters
    const uint32_t *carr;       // carrier table (set by launch_pass)
    const uint8_t *buf;
    int buffN;
    const int *mix_ix;          // [nbatch] start index in the ring
    const uint32_t *mix_p0;     // [nbatch] phase of sample 0
    const uint32_t *mix_p1;     // [nbatch] phase at the wrap split
    const uint32_t *mix_s;      // [nbatch] phase step
    const int *mix_nseg;        // [nbatch] samples before the wrap split
    int mix_compact;            // 1: buf holds the block contiguously (no ring)
    // LD_MIX on a packed raw ring (raw_bps > 0) instead of cpx8: sample i's
    // byte for the channel is buf[i raw_bps + raw_off], its cpx8 value
    // raw_lut[(raw_rtoc ? ((raw_ph0 + i) & 3) << 8 : 0) + byte]
    // (write_buff, src/sdr_rcv.c:840-914); the 1024-byte table is staged in
    // LDS after the carrier table
    int raw_bps, raw_off, raw_rtoc, raw_ph0;
    const uint8_t *raw_lut;
    // LD_CPX16: cpx16 input
    const int8_t *iq16;
    // LD_PROD: X (in, batch b -> b / prod_xdiv % prod_xmod) times Cf
    const float2 *cf;           // code spectrum, batch b -> (b / prod_cdiv)
    const int *cf_index;        // FP32 passes: per-batch spectrum index (nullptr: b / prod_cdiv)
    int prod_xdiv, prod_xmod, prod_cdiv;
    long long cf_bstride;
    float prod_scale;
    int prod_V;                 // natural indices >= prod_V use the scalar-tail form
    int prod_M1, prod_M2;       // transposed index e -> natural e / M2 + M1 (e % M2)
    // ST_POW: P accumulate; batch b -> row (b / p_div + p_off1) * p_mul
    // + p_off2 + b % p_div, row start P + row * p_bstride
    float *P;
    long long p_bstride;
    int p_div, p_off1, p_mul, p_off2;
    int nlag;               // only lags < nlag are accumulated
    int p_init;             // 1: first block, P = |x|^2 (P starts at zero)
    PeakPart *p_part;       // last block: one partial per workgroup at
                            // part[row * ngroups + group] (nullptr: none)
    int p_nbin;             // rows per PRN (row-major index (row % p_nbin) N + e)
    // LD_PROD XCD-aware order (xcd_nbin > 0, specialised lengths): items
    // b = prn * xcd_nbin + bin, prn < xcd_nprn; the PRNs of one (row group,
    // bin) tile run back to back on one XCD
    int xcd_nbin, xcd_nprn;
    int xcd_pair;               // FP32 column passes: XCD pairing of narrow column groups
    // precision form (psa_fft16.hip): a complex value is one 32-bit word of
    // two FP16 or BF16 halves, real part in the low half
    const uint32_t *hin;        // LD_CPX / LD_PROD (X) input words
    const uint32_t *hin2;       // paired power passes: the next block's input words
    uint32_t *hout;             // ST_CPX output words
    const uint32_t *hcf;        // LD_PROD code-spectrum words
    const uint4 *amat[3];       // MFMA A operands of the stage DFT matrices
    const float2 *tfull;        // ST_CPX: twiddle per output index, pre-scaled; nullptr: none
    const uint32_t *tfullh;     // the same as FP16 words (inverse row pass, FP16 form)
    float osc;                  // ST_CPX scale applied when tfull is nullptr
    float kappa;                // ST_POW: values scaled by kappa before |.|^2
    int hvar;                   // workgroup shape variant of the MFMA pass
};

// threads (workgroups x workgroup size) of the last pass launched by this
// host thread, for the plans' per-launch statistics
extern thread_local long long pass_grid;

void launch_pass(LoadMode lm, StoreMode sm, const PassArgs &a, int nbatch,
                 hipStream_t st);
// fused forward row pass + code-spectrum product + inverse row pass
// (specialised row lengths only; returns false for other lengths)
bool launch_xcorr_row(const PassArgs &a, int nbatch, hipStream_t st);
int pass_ngroups(const PassArgs &a);   // workgroups per batch item
int pass_groups(int L, int nsub);
bool spec_length(int L);          // L has a compile-time specialised kernel
// column pass: M2 transforms of length M1 at stride M2, output [k1][n2]
// row-major, inter-pass twiddle w_M^{n2 k1} on store (when M2 > 1)
void col_pass_args(const FFTPlan *pl, int dir, PassArgs &a);
// row pass: M1 contiguous transforms of length M2, output [k1][k2]
void row_pass_args(const FFTPlan *pl, int dir, PassArgs &a);

// ---------------------------------------------------------------------------
// precision form (psa_fft16.hip): FP16 / BF16 words, MFMA DFT-matrix stages
// ---------------------------------------------------------------------------
enum { HDT_F16 = 0, HDT_BF16 = 1 };
// workgroup shapes (radix split, workgroup size) per pass length
#define PSA_HVARS 4

struct HPlan {
    int M = 0, M1 = 0, M2 = 0;  // column length M1, row length M2
    int dt = HDT_F16;
    float2 *tw1[2] = {}, *tw2[2] = {};   // w_L^m of the two lengths, [dir]
    float2 *tfull[2] = {};               // inter-pass twiddle by output index, [dir]
    uint32_t *tfullh = nullptr;          // FP16 words of tfull[1] (FP16 plans)
    // DFT tables of each workgroup shape's radices, [shape][dir][stage]
    const uint4 *am1[PSA_HVARS][2][3] = {}, *am2[PSA_HVARS][2][3] = {};
    // power-of-two scales: forward column / row pass outputs, inverse row
    // pass output (folded into tfull where a pass has the twiddle)
    float osc_fc = 1, osc_fr = 1, osc_ir = 1;
};
// cached per (M, dt, M1); M1 = 0 takes the default split; nullptr when M
// has no MFMA split (or M1 is not one of its splits)
const HPlan *hfft_plan(int M, int dt, int M1 = 0);
// column lengths M1 of the MFMA splits of M, default first; returns the count
int h_splits(int M, int *m1, int nmax);
// whether M1 (column length) and M / M1 (row length) have workgroup shapes
// vc and vr
bool h_shape(int M, int M1, int vc, int vr);
// whether the power (inverse column) pass of length L, shape var, takes two
// blocks' inputs in one launch (PassArgs::hin2)
bool h_col_pairs(int dt, int L, int var);
// whether some column shape of some split of M takes paired blocks
bool h_any_pairs(int dt, int M);
// pass geometry; var selects one of the workgroup shapes of the length
void h_col_args(const HPlan *hp, int dir, PassArgs &a, int var = 0);
void h_row_args(const HPlan *hp, int dir, PassArgs &a, int var = 0);
void launch_hpass(int dt, LoadMode lm, StoreMode sm, const PassArgs &a, int nbatch,
                  hipStream_t st);
// natural-order FP32 code spectra (n x M) -> digit-transposed words * scale
void h_code_spec(const HPlan *hp, const float2 *nat, uint32_t *out, int n, float scale,
                 hipStream_t st);

// Fused inverse of whole search cells (FP16 form): one workgroup per (PRN,
// Doppler bin) runs the code-product row pass, the column pass and the power
// accumulation of every block with the intermediate and the cell's power
// surface in LDS, then writes P and the cell's peak partial once.
struct CellArgs {
    const uint32_t *X;      // forward spectra [n_blk][n_bin][M] words (digit-transposed)
    long long xblk;         // words per block of X (n_bin M)
    const uint32_t *C;      // code spectra [n_prn][M] words
    float *P;               // [n_prn][n_bin][N]
    PeakPart *part;         // [n_prn][n_bin]: one partial per cell
    int nprn, nbin, nblk, N;
    float kap2;             // power scale kappa^2
};
// the column length M1 of the split of M that has a cell kernel (0: none)
int cell_split(int M);
// launches the cell kernel of hp's split (hp: FP16, split cell_split(M));
// N <= M / 2.  prepare: only set the kernel up (plan creation, before any
// capture), no launch
void launch_cells(const HPlan *hp, const CellArgs &a, hipStream_t st, bool prepare = false);

// Full forward / backward transforms of `nbatch` contiguous length-M arrays
// (natural order in and out), used by sdr_cpx_fft and the code spectra.
void fft_c2c(const FFTPlan *pl, const float2 *in, float2 *out, float2 *tmp,
             int nbatch, int dir, hipStream_t st);
// natural-order spectrum -> the digit-transposed order used between the
// forward and inverse passes of the correlator (index k1*M2 + k2 holds
// bin k1 + M1*k2)
void fft_to_transposed(const FFTPlan *pl, const float2 *in, float2 *out,
                       int nbatch, hipStream_t st);

// ---------------------------------------------------------------------------
// pass argument helpers for the correlator
// ---------------------------------------------------------------------------
inline void set_prod(const FFTPlan *pl, PassArgs &a, float scale)
{
    a.prod_scale = scale;
    a.prod_V = pl->M >= 4 ? pl->M / 4 * 4 : 0;
    a.prod_M1 = pl->M1;
    a.prod_M2 = pl->M2;
}

inline void args_fwd_col(const FFTPlan *pl, PassArgs &a)
{
    col_pass_args(pl, 0, a);
}

inline void args_fwd_row(const FFTPlan *pl, PassArgs &a)
{
    row_pass_args(pl, 0, a);   // digit-transposed output
}

inline void args_inv_row(const FFTPlan *pl, PassArgs &a)
{
    row_pass_args(pl, 1, a);
    if (pl->M2 > 1) {          // twiddle w_M^{-n_a k1} on store
        a.thi = pl->thi_bwd;
        a.tlo = pl->tlo_bwd;
    }
}

inline void args_inv_col(const FFTPlan *pl, PassArgs &a)
{
    col_pass_args(pl, 1, a);   // natural-order output n_a + M2 n_b
    a.thi = a.tlo = nullptr;
}

// the fused row pass: forward row geometry, inverse twiddles, the inverse
// row pass's inter-pass twiddle on store
inline void args_xcorr_row(const FFTPlan *pl, PassArgs &a)
{
    row_pass_args(pl, 0, a);
    a.tw_inv = pl->s2.tw_bwd;
    if (pl->M2 > 1) {
        a.thi = pl->thi_bwd;
        a.tlo = pl->tlo_bwd;
    }
}

// ---------------------------------------------------------------------------
// misc kernels (psa_mix.hip)
// ---------------------------------------------------------------------------
// compact = 1: buf holds the N samples contiguously (second ring segment
// directly after the first); compact = 0: buf is the whole ring
void launch_mix(const uint8_t *buf, int buffN, int ix, int N, uint32_t p0,
                uint32_t p1, int nseg, uint32_t s, int8_t *IQ, hipStream_t st,
                int compact = 0);
void mix_params(int buffN, int ix, int N, double fs, double fc, double phi,
                uint32_t *p0, uint32_t *p1, int *nseg, uint32_t *s);
// c[j] = a[j] * b * s for nbatch arrays a[j], c[j] of N (stride N), one b
void launch_cpx_mul(const float2 *a, const float2 *b, int N, float s,
                    float2 *c, hipStream_t st, int nbatch = 1);

// ---------------------------------------------------------------------------
// any-length DFT (psa_anylen.hip): Bluestein's chirp-z transform on a
// 2-3-5-smooth FFT of length L >= 2 N - 1
// ---------------------------------------------------------------------------
struct AnyPlan {
    int N = 0, L = 0;
    const FFTPlan *pl = nullptr;    // length L
    float2 *chirp[2] = {};          // [dir] exp(s i pi m^2 / N), m < N, s = -1 / +1
    float2 *bf[2] = {};             // [dir] FFT of the conjugate chirp, / L
};
// cached per N, thread-safe; nullptr when no smooth L <= 4096^2 exists
const AnyPlan *any_plan(int N);
// nbatch contiguous length-N transforms (natural order); tmp holds
// 2 nbatch L complex; out may equal in
void fft_any(const AnyPlan *ap, const float2 *in, float2 *out, float2 *tmp, int nbatch, int dir,
             hipStream_t st);
// cpx16 (int8 pairs) -> float complex * CSCALE, n samples
void launch_cpx16_to_float(const int8_t *iq, long long n, float2 *out, hipStream_t st);

}  // namespace psa
