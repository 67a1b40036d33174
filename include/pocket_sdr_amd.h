/*
 * pocket_sdr_amd.h - C ABI of the MI355X (gfx950) PocketSDR search / correlate
 * hot path (libpocketsdr_amd.so).
 *
 * Part 1 re-declares, with identical signatures and semantics, the hot-path
 * functions of the reference header src/pocket_sdr.h (tomojitakasu/PocketSDR
 * v0.17).  Each declaration cites the reference prototype it replaces and the
 * definition whose behaviour it keeps.  All pointers are caller-owned host
 * memory, exactly as in the reference; the library copies to and from HBM
 * internally on a per-thread HIP stream.
 *
 * Part 2 is additive: plan-based batched entry points whose inputs stay
 * resident in HBM between runs (no reference counterpart; they batch the
 * loops of app/pocket_acq/pocket_acq.c:152-192 and src/sdr_ch.c:692-740).
 *
 * Types are ABI-identical to src/pocket_sdr.h:105-113 and 242-245; when the
 * reference header has already been included its typedefs are used.
 */
#ifndef POCKET_SDR_AMD_H
#define POCKET_SDR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef POCKET_SDR_H
#define SDR_N_CODES    10          /* src/pocket_sdr.h:60 */
#define SDR_MAX_CORR   107         /* src/pocket_sdr.h:58 */
#define SDR_FFT_FORWARD 0          /* src/pocket_sdr.h:98 */
#define SDR_FFT_BACKWARD 1         /* src/pocket_sdr.h:99 */
typedef uint8_t sdr_cpx8_t;                    /* src/pocket_sdr.h:110 */
typedef struct { int8_t I, Q; } sdr_cpx16_t;   /* src/pocket_sdr.h:111 */
typedef float sdr_cpx_t[2];                    /* src/pocket_sdr.h:113 */
typedef struct {                               /* src/pocket_sdr.h:242-245 */
    sdr_cpx8_t *data;
    int IQ, N;
} sdr_buff_t;
#endif

/* ======================================================================
 * Part 1: drop-in replacements (src/pocket_sdr.h:391-465)
 * ====================================================================== */

/* src/pocket_sdr.h:392, def. src/sdr_func.c:113-134.  Builds the carrier
 * table and uploads it to the GPU; the FFTW wisdom file is ignored. */
void sdr_func_init(const char *file);

/* src/pocket_sdr.h:393-394, def. lib/pocketfft/pocketfft_wrap.cpp:62-82 via
 * src/sdr_func.c:147-171 (32-byte aligned, exits on failure) */
sdr_cpx_t *sdr_cpx_malloc(int N);
void sdr_cpx_free(sdr_cpx_t *cpx);

/* src/pocket_sdr.h:395, def. src/sdr_func.c:182-185 */
float sdr_cpx_abs(sdr_cpx_t cpx);

/* src/pocket_sdr.h:396-397, def. src/sdr_func.c:197-233 (c = a b s) */
void sdr_cpx_mul(const sdr_cpx_t *a, const sdr_cpx_t *b, int N, float s,
    sdr_cpx_t *c);

/* src/pocket_sdr.h:398, def. src/sdr_func.c:236-264.  Unnormalised FFT of
 * any length N >= 1, dir SDR_FFT_FORWARD / SDR_FFT_BACKWARD; returns 0 for a
 * bad dir or N.  Lengths with a prime factor above 5 run Bluestein's chirp-z
 * transform on a 2-3-5-smooth length (as do sdr_corr_fft and
 * sdr_search_code). */
int sdr_cpx_fft(sdr_cpx_t *cpx1, int N, int dir, sdr_cpx_t *cpx2);

/* src/pocket_sdr.h:399-400, def. src/sdr_func.c:276-297 */
sdr_buff_t *sdr_buff_new(int N, int IQ);
void sdr_buff_free(sdr_buff_t *buff);

/* src/pocket_sdr.h:407-409, def. src/sdr_func.c:664-687.  P (len_fds x N,
 * bin-major) accumulates |corr|^2 over all N lags.  The reference's 1 ms
 * sleep every 22 bins is not reproduced. */
void sdr_search_code(const sdr_cpx_t *code_fft, double T,
    const sdr_buff_t *buff, int ix, int N, double fs, double fi,
    const float *fds, int len_fds, float *P);

/* src/pocket_sdr.h:410-412, 415, def. src/sdr_func.c:690-706, 728-744,
 * 758-769 (host-side helpers over host arrays) */
float sdr_corr_max(const float *P, int N, int Nmax, int M, double T, int *ix);
double sdr_fine_dop(const float *P, int N, const float *fds, int len_fds,
    const int *ix);
float *sdr_dop_bins(double T, float dop, float max_dop, int *len_fds);

/* src/pocket_sdr.h:416-417, def. src/sdr_func.c:981-1008 */
void sdr_corr_std(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code, int N,
    double coff, const double *pos, int n, sdr_cpx_t *corr, sdr_cpx_t *C);

/* src/pocket_sdr.h:418-420, def. src/sdr_func.c:1029-1057 */
void sdr_corr_std_cpx_code(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code,
    int N, double coff, const double *pos, int n, sdr_cpx_t *corr,
    sdr_cpx_t *C);

/* src/pocket_sdr.h:421-423, def. src/sdr_func.c:1060-1072 (Python-facing) */
void sdr_corr_std_cpx(const sdr_cpx_t *buff, int len_buff, int ix, int N,
    double fs, double fc, double phi, const float *code, const double *pos,
    int n, sdr_cpx_t *corr);

/* src/pocket_sdr.h:424-425, def. src/sdr_func.c:1088-1102 */
void sdr_corr_fft(const sdr_cpx16_t *IQ, const sdr_cpx_t *code_fft, int N,
    sdr_cpx_t *corr);

/* src/pocket_sdr.h:426-428, def. src/sdr_func.c:1105-1113 (Python-facing) */
void sdr_corr_fft_cpx(const sdr_cpx_t *buff, int len_buff, int ix, int N,
    double fs, double fc, double phi, const sdr_cpx_t *code_fft,
    sdr_cpx_t *corr);

/* src/pocket_sdr.h:429-430, def. src/sdr_func.c:772-836 */
void sdr_mix_carr(const sdr_buff_t *buff, int ix, int N, double fs, double fc,
    double phi, sdr_cpx16_t *IQ);

/* ======================================================================
 * Part 2: batched, HBM-resident plans (additive, no reference counterpart)
 * ====================================================================== */

/* --- acquisition ----------------------------------------------------------
 * One signal group = one code length: n_prn code spectra of length M
 * (conjugated, as made by sdr_gen_code_fft with Nz = N), searched over
 * len_fds Doppler bins and n_blk non-coherent blocks starting at
 * ix0 + k * blk_step in IF stream `buff`.  Mirrors search_sig of
 * app/pocket_acq/pocket_acq.c:152-192 for every PRN of the group.
 */
typedef struct {
    int buff;                  /* index into the plan's IF streams */
    int M;                     /* FFT length (2 N) */
    int N;                     /* lags searched (code period in samples) */
    double T, fs, fi;          /* code period (s), sampling rate, IF (Hz) */
    int n_prn;
    const sdr_cpx_t *code_fft; /* n_prn x M */
    int len_fds;
    const float *fds;          /* Doppler bins (Hz), as sdr_dop_bins */
    int n_blk, ix0, blk_step;
} sdr_search_sig_t;

#define SDR_SEARCH_BATCHED  0  /* shared forward FFT per (bin, block), fused */
#define SDR_SEARCH_PLAIN    1  /* reference dataflow per (PRN, block), unfused */
#define SDR_SEARCH_GRAPH    2  /* capture the run in a hipGraph and replay it */
/* Precision form of the batched search: FP16 (or BF16) complex words between
 * the FFT passes, every radix stage a DFT-matrix product on the matrix cores
 * with FP32 accumulation, FP32 twiddles and power surfaces.  Applies to FFT
 * lengths with an MFMA split (24000, 48000, 96000, 192000, 960000); other
 * groups of the plan run the FP32 batched path.  Not combined with PLAIN. */
#define SDR_SEARCH_FP16     4
#define SDR_SEARCH_BF16     8
/* Batched forms: sdr_search_plan_new times each signal group's candidate
 * FFT splits (and, in the precision form, workgroup shapes) on the group's
 * first block and keeps the fastest (adds a few runs of one block per
 * candidate to plan creation). */
#define SDR_SEARCH_TUNE     16

typedef struct sdr_search_plan_tag sdr_search_plan_t;

/* Copies the IF streams and code spectra to HBM and allocates every work
 * buffer.  Returns NULL (with a message on stderr) on invalid input. */
sdr_search_plan_t *sdr_search_plan_new(int n_buff, const sdr_buff_t *const *buffs,
    int n_sig, const sdr_search_sig_t *sigs, int flags);
/* Runs the whole search on the device (P surfaces are zeroed first), then
 * the peak search.  Blocks until done.  Returns 1 on success. */
/* The same on IF buffers in device memory (buffN bytes each, e.g. written
 * by sdr_raw_unpack_dev / sdr_arch_combine_raw_dev): read in place at every
 * run, so new samples written there between runs are searched by the next
 * run; the caller keeps them allocated while the plan lives. */
sdr_search_plan_t *sdr_search_plan_new_dev(int n_buff, const uint8_t *const *dbuf,
    const int *buffN, int n_sig, const sdr_search_sig_t *sigs, int flags);
/* The same on packed raw rings in device memory as the front end receives
 * them (no cpx8 staging pass): the mixing loaders read each sample's raw
 * byte for the stream's RF channel and convert it with the channel's table
 * as write_buff does (src/sdr_rcv.c:840-914), so the power surfaces equal
 * those of a plan on the unpacked cpx8 stream.  Batched forms only (not
 * SDR_SEARCH_PLAIN); N x bytes per sample < 2^31.  Returns NULL on invalid
 * input. */
typedef struct {
    const uint8_t *raw;        /* device memory: N samples of the format */
    int N;                     /* samples in the ring */
    int fmt;                   /* SDR_FMT_RAW8, _RAW16, _RAW16I or _RAW32
                                  (src/pocket_sdr.h:87-90) */
    int rfch;                  /* RF channel of the format */
    int rtoc;                  /* 1: fs/4 real-to-complex tables (not RAW16I) */
    const sdr_cpx8_t *lut;     /* host: the channel's table as sdr_raw_gen_lut
                                  makes it (256 entries, 1024 with rtoc) */
    int ph0;                   /* fs/4 phase of ring sample 0, (ix N) & 3 */
} sdr_raw_stream_t;
sdr_search_plan_t *sdr_search_plan_new_raw(int n_stream, const sdr_raw_stream_t *streams,
    int n_sig, const sdr_search_sig_t *sigs, int flags);
int sdr_search_plan_run(sdr_search_plan_t *plan);
/* Per-PRN results of signal group `sig` as sdr_corr_max / sdr_fine_dop
 * produce them: cn0[n_prn], ix[2 n_prn] (bin, lag), dop[n_prn]. */
int sdr_search_plan_result(const sdr_search_plan_t *plan, int sig, float *cn0,
    int *ix, double *dop);
/* Copies the power surface of one PRN (len_fds x N, bin-major). */
int sdr_search_plan_power(const sdr_search_plan_t *plan, int sig, int prn_idx,
    float *P);
/* Arithmetic form signal group `sig` runs in (-1 on a bad argument). */
#define SDR_FORM_FP32   0      /* batched FP32 passes */
#define SDR_FORM_PLAIN  1      /* per-PRN reference dataflow, FP32 */
#define SDR_FORM_FP16   2      /* batched, FP16 words, MFMA stages */
#define SDR_FORM_BF16   3      /* batched, BF16 words, MFMA stages */
int sdr_search_plan_form(const sdr_search_plan_t *plan, int sig);
/* Tiles of group `sig`: tiles[0..3] = column length M1, row length M2 of the
 * FFT split, column and row workgroup-shape variant (0 to 3; precision form
 * only).  Returns 0 on a bad argument. */
int sdr_search_plan_tiles(const sdr_search_plan_t *plan, int sig, int *tiles);
/* Workgroup shape `var` of the precision form's column (col = 1) or row
 * (col = 0) passes of length L: shape[0..4] = the radices R0, R1, R2 (1 when
 * unused), G sub-transforms and threads per workgroup.  Returns 0 when the
 * length has no such shape. */
int sdr_search_pass_shape(int col, int L, int var, int *shape);
/* Sub-transforms per workgroup of the FP32 passes' compiled kernel of length
 * L (0 when L runs on the runtime-length kernel); launches of fewer than
 * 2048 workgroups other than power passes take half as many. */
int sdr_fft_pass_groups(int L);
/* Enables per-kernel HIP-event timing of later runs (0 disables). */
void sdr_search_plan_timing(sdr_search_plan_t *plan, int on);
/* Average device time (ms) and launch count of kernel class `name`
 * ("fwd_col", "fwd_row", "inv_row", "inv_col", "mix", "mul", "pow", "peak",
 * "fill") over the timed runs; "name@g" restricts it to signal group g.
 * Returns 0 if the class never ran. */
int sdr_search_plan_kernel_ms(const sdr_search_plan_t *plan, const char *name,
    double *avg_ms, long long *count, double *elems);
/* Average algorithmic HBM bytes per launch of kernel class `name` over the
 * timed runs: the compulsory reads and writes of its operands (IF bytes once
 * per block, shared spectra once per launch, power rows read and written). */
int sdr_search_plan_kernel_bytes(const sdr_search_plan_t *plan, const char *name,
    double *avg_bytes);
/* Average launch size (workgroups x workgroup threads, the grid size a
 * profiler reports) of kernel class `name` over the timed runs. */
int sdr_search_plan_kernel_grid(const sdr_search_plan_t *plan, const char *name,
    double *avg_threads);
void sdr_search_plan_free(sdr_search_plan_t *plan);
/* Device memory (bytes) held by all live search plans of the process. */
long long sdr_search_plan_bytes(void);

/* --- tracking correlators -------------------------------------------------
 * n_ch channels on one IF stream, n_epoch consecutive epochs of N samples.
 * Each epoch advances the channel open-loop as src/sdr_ch.c:694-703 (Doppler
 * fd0 + drift t, no loop filters), mixes with sdr_mix_carr semantics and
 * correlates with sdr_corr_std semantics at the n_pos tap positions.
 */
typedef struct {
    double fs, fi, fc_sig, T;  /* sampling rate, IF, signal carrier, period */
    int N;                     /* samples per epoch (same for all channels) */
    double fd0, drift;         /* Doppler (Hz), Doppler rate (Hz/s) */
    double phi0, coff0;        /* initial carrier phase (cyc), code offset (s) */
    const sdr_cpx16_t *code;   /* resampled code bank, N x SDR_N_CODES */
} sdr_trk_ch_t;

typedef struct sdr_trk_plan_tag sdr_trk_plan_t;

sdr_trk_plan_t *sdr_trk_plan_new(const sdr_buff_t *buff, int n_ch,
    const sdr_trk_ch_t *chs, int n_pos, const double *pos, int n_epoch);
int sdr_trk_plan_run(sdr_trk_plan_t *plan);
/* corr: n_ch x n_epoch x n_pos complex; C: n_ch x n_epoch x 2 complex
 * (correlations before / after the code wrap, as sdr_corr_std's C). */
int sdr_trk_plan_result(const sdr_trk_plan_t *plan, sdr_cpx_t *corr,
    sdr_cpx_t *C);
/* P: n_ch x n_epoch full-period prompts, the trk->P history entries of
 * src/sdr_ch.c:736-740 ((C[0] + previous epoch's C[1]) / N, zero before the
 * first epoch). */
int sdr_trk_plan_prompt(const sdr_trk_plan_t *plan, sdr_cpx_t *P);
/* Closed-loop tracking (src/sdr_ch.c:453-624, 674-763: FLL for the first
 * second, then PLL, DLL, C/N0 and carrier-lock windows, the loss decision,
 * the secondary-code synchronisation and wipe-off from 1.5 s, bump-jump for
 * BOC signals; not the L6 branch); each channel starts as start_track
 * leaves it (src/sdr_ch.c:378-394) from chs[c].fd0, coff0, phi0 (drift is
 * not used) and loops[c].  Taps 0-3 must be P, E, L and the noise tap, and
 * with bump-jump taps 4 and 5 VE and VL (trk->pos, src/sdr_ch.c:228-243).
 * The prompt history (SDR_N_HIST entries, shifted with the code-offset
 * wraps as adj_coff does) lives on the GPU.  With a secondary code the
 * taps of sdr_trk_plan_result and the prompts of sdr_trk_plan_prompt are
 * the wiped ones trk->C and trk->P hold.  Every run restarts from these
 * states; atan and log10 are the GPU's (the results follow the reference to
 * their last bits). */
typedef struct {
    int len_code;              /* code length (chips), the DLL's scale */
    int costas;                /* 1: Costas discriminators (all but L6D / L6E) */
    double cn0;                /* C/N0 at acquisition (dB-Hz) */
    int len_sec_code;          /* secondary code length (0 or 1: none; at most
                                  SDR_N_HIST) */
    const int8_t *sec_code;    /* ch->sec_code (+-1), host memory */
    double bump_step;          /* bump_jump's code step (s): T / len_code, / 3
                                  for E5ABQ, / 2 for G1OCP / G2OCP
                                  (src/sdr_ch.c:552-562); 0: no bump-jump */
} sdr_trk_loop_t;
sdr_trk_plan_t *sdr_trk_plan_new_loop(const sdr_buff_t *buff, int n_ch,
    const sdr_trk_ch_t *chs, const sdr_trk_loop_t *loops, int n_pos,
    const double *pos, int n_epoch);
/* closed-loop plans: n_ch x n_epoch x {fd, coff, phi, cn0, lock, idle}
 * after each epoch (idle 1: lost at or before it, not processed) */
int sdr_trk_plan_state(const sdr_trk_plan_t *plan, double *state);
void sdr_trk_plan_timing(sdr_trk_plan_t *plan, int on);
int sdr_trk_plan_kernel_ms(const sdr_trk_plan_t *plan, const char *name,
    double *avg_ms, long long *count, double *elems);
void sdr_trk_plan_free(sdr_trk_plan_t *plan);

/* --- L6D / L6E FFT tracking and CSK ----------------------------------------
 * n_ch channels of one signal (one N and one code length) on one IF stream,
 * n_epoch consecutive epochs each.  Every epoch advances the channel
 * open-loop as the tracking plan does (src/sdr_ch.c:694-703), then follows
 * the L6D / L6E branch of track_sig (src/sdr_ch.c:705-720): i = floor(coff
 * fs), bank j = (int)((coff fs - i) SDR_N_CODES), sdr_mix_carr from ring
 * index (e N) mod buffN + i at phase phi + fc i / fs (the start taken mod
 * buffN), sdr_corr_fft against code_fft + j N, and CSK (src/sdr_ch.c:
 * 636-664): the symbol 255 - ix % 256 of the peak shift ix and the n_pos
 * correlator taps interpolated at n + ix R + pos[k].
 */
typedef struct {
    double fs, fi, fc_sig, T;  /* sampling rate, IF, signal carrier, period */
    int N;                     /* samples per epoch = correlator FFT length */
    int len_code;              /* code length (20460 for L6D / L6E, TDM) */
    double fd0, drift;         /* Doppler (Hz), Doppler rate (Hz/s) */
    double phi0, coff0;        /* initial carrier phase (cyc), code offset (s) */
    const sdr_cpx_t *code_fft; /* SDR_N_CODES x N conjugated code spectra, the
                                  layout of trk->code_fft (src/sdr_ch.c:256-262) */
} sdr_l6_ch_t;

#define SDR_L6_KEEP_CORR 1     /* keep every item's correlation for _corr */

typedef struct sdr_l6_plan_tag sdr_l6_plan_t;

sdr_l6_plan_t *sdr_l6_plan_new(const sdr_buff_t *buff, int n_ch,
    const sdr_l6_ch_t *chs, int n_pos, const double *pos, int n_epoch,
    int flags);
/* The same on an IF ring of buffN cpx8 bytes already in device memory (for
 * example the front end's output), read in place at every run; the caller
 * keeps it allocated until sdr_l6_plan_free. */
sdr_l6_plan_t *sdr_l6_plan_new_dev(const uint8_t *dbuf, int buffN, int n_ch,
    const sdr_l6_ch_t *chs, int n_pos, const double *pos, int n_epoch,
    int flags);
/* Closed-loop L6D / L6E tracking: the L6 branch of track_sig
 * (src/sdr_ch.c:705-720) inside the loop of :741-763 (FLL / PLL / DLL, C/N0,
 * loss), the CSK's P tap as the prompt.  Per channel the loop parameters
 * of sdr_trk_plan_new_loop: {ch->len_code, ch->costas (0 for L6D / L6E,
 * src/sdr_ch.c:326), cn0, no secondary code, no bump-jump}; pos starts
 * with P, E, L, N (src/sdr_ch.c:228-231).  The epochs run in order, all
 * channels of one epoch in one batch of FFT passes.  sdr_l6_plan_result
 * and sdr_l6_plan_state then give every epoch's symbol, taps and loop
 * state (zeros and idle after a loss of lock). */
sdr_l6_plan_t *sdr_l6_plan_new_loop(const sdr_buff_t *buff, int n_ch,
    const sdr_l6_ch_t *chs, const sdr_trk_loop_t *loops, int n_pos,
    const double *pos, int n_epoch, int flags);
/* state: n_ch x n_epoch x {fd, coff, phi, cn0, lock, idle} after each epoch
 * (closed-loop plans) */
int sdr_l6_plan_state(const sdr_l6_plan_t *plan, double *state);
int sdr_l6_plan_run(sdr_l6_plan_t *plan);
/* syms: n_ch x n_epoch CSK symbols; C: n_ch x n_epoch x n_pos taps (the
 * trk->C of src/sdr_ch.c:660-662). */
int sdr_l6_plan_result(const sdr_l6_plan_t *plan, uint8_t *syms,
    sdr_cpx_t *C);
/* N-sample correlation of one (channel, epoch), ch->corr of
 * src/sdr_ch.c:715 (plans made with SDR_L6_KEEP_CORR). */
int sdr_l6_plan_corr(const sdr_l6_plan_t *plan, int ch, int epoch,
    sdr_cpx_t *corr);
void sdr_l6_plan_timing(sdr_l6_plan_t *plan, int on);
/* kernel classes: "state", "fwd_col", "row" (fused row pass), or "fwd_row" and
 * "inv_row" (row lengths without a fused kernel), "inv_col", "csk", "loop"
 * (closed-loop plans) */
int sdr_l6_plan_kernel_ms(const sdr_l6_plan_t *plan, const char *name,
    double *avg_ms, long long *count, double *elems);
void sdr_l6_plan_free(sdr_l6_plan_t *plan);
/* CSK of n_item correlations of N samples each (src/sdr_ch.c:636-664, a
 * static function there): syms[n_item], C[n_item x n_pos]. */
int sdr_csk(const sdr_cpx_t *corr, int n_item, int N, int len_code,
    const double *pos, int n_pos, uint8_t *syms, sdr_cpx_t *C);

/* --- front end: raw data to cpx8 IF samples ---------------------------------
 * The receiver's write_buff (src/sdr_rcv.c:834-914, a static function there)
 * on the GPU, for the formats SDR_FMT_INT8 ... SDR_FMT_CS16
 * (src/pocket_sdr.h:85-92); the per-channel LPF and the array combine that
 * follow it are not included.  Tables as src/sdr_rcv.c:776-831 builds them:
 * raw formats take lut[k] = the cpx8 table of RF channel k (256 entries, or
 * 4 x 256 when rtoc[k]: the fs/4 real-to-complex phases); INT8, INT8X2,
 * CS8 take lut[0] = the int8 AGC table (256 entries), CS16 lut[0] / lut[1]
 * = its I / Q tables (65536 entries; lut[1] NULL: the I table).  Sample j of
 * block ix uses phase (ix N + j) & 3, as write_buff.
 */
typedef struct sdr_raw_plan sdr_raw_plan_t;

/* gen_LUTC (rtoc = 0) / gen_LUTC_rtoc (rtoc = 1: 1024 entries),
 * src/sdr_rcv.c:776-799, 817-831 */
void sdr_raw_gen_lut(int fmt, int rfch, int IQ, int bits, int rtoc,
    sdr_cpx8_t *lut);
/* gen_LUT8, src/sdr_rcv.c:802-814 (65536 entries for SDR_FMT_CS16) */
void sdr_raw_gen_lut8(int fmt, double off, double scale, int8_t *lut);
sdr_raw_plan_t *sdr_raw_plan_new(int fmt, const int *rtoc,
    const void *const *lut);
/* RF channels the format carries (1, 2, 4 or 8) */
int sdr_raw_plan_nch(const sdr_raw_plan_t *plan);
/* host pointers: N samples (N x bytes-per-sample raw bytes) of block ix to
 * out[k][0..N-1] for every channel k; returns 1 on success */
int sdr_raw_unpack(sdr_raw_plan_t *plan, const uint8_t *raw, int N,
    int64_t ix, sdr_cpx8_t *const *out);
/* the same on device memory (16-byte aligned pointers), HBM-resident */
int sdr_raw_unpack_dev(sdr_raw_plan_t *plan, const uint8_t *raw, long long N,
    int64_t ix, sdr_cpx8_t *const *out);
/* kernel time of the last unpack call (ms, HIP events) */
double sdr_raw_plan_kernel_ms(const sdr_raw_plan_t *plan);
void sdr_raw_plan_free(sdr_raw_plan_t *plan);

/* Array channels: the sample combine of sdr_arch_combine (src/sdr_array.c:
 * 553-579) for narch beams at once, on raw sample arrays instead of the
 * receiver structs.  Named apart from the reference's own sdr_arch_combine
 * (src/pocket_sdr.h:522, another signature), so that both headers can be
 * included in one translation unit and both libraries linked in one process.
 * in[j]: cpx8 samples of RF channel j (nrfch 2-8); w: Q8 weights
 * [beam][channel][re, im] as sdr_arch_t.w (src/pocket_sdr.h:297) holds them;
 * out[m]: cpx8 samples of beam m (narch 1-8).  Host pointers (returns 1) or
 * device pointers, 16-byte aligned (returns the kernel time in ms, < 0 on
 * error). */
int sdr_arch_combine_raw(int nrfch, const uint8_t *const *in, int N, int narch,
    const int16_t *w, uint8_t *const *out);
double sdr_arch_combine_raw_dev(int nrfch, const uint8_t *const *in, long long N,
    int narch, const int16_t *w, uint8_t *const *out);

/* sdr_search_code without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as sdr_search_code (bit-identical P); the work
 * buffers live in a fixed 64 MiB device array of the library (one per GPU,
 * held until the process ends), so the FFT length N must leave room for at
 * least one Doppler bin there (N up to about 1.8 million; bins are processed
 * in chunks).  Requires sdr_func_init first.  Returns 1 on success, 0 (with a
 * message on stderr, P untouched) for a length the array cannot hold, a
 * length with a prime factor above 5, or invalid arguments.  Every call on a
 * GPU shares its array: calls that overlap in time on one GPU race. */
int sdr_search_code_noalloc(const sdr_cpx_t *code_fft, double T,
    const sdr_buff_t *buff, int ix, int N, double fs, double fi,
    const float *fds, int len_fds, float *P);

/* library version string */
const char *sdr_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* POCKET_SDR_AMD_H */
