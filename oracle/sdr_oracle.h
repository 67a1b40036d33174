/*
 * sdr_oracle.h - CPU oracle for the PocketSDR search / correlate hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (tomojitakasu/PocketSDR, src/sdr_func.c and src/sdr_code.c) used
 * as the checker for the HIP library.  Only tests/, the smoke() entry and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls anything in this directory.
 *
 * Every function cites the reference file:line it restates.  Pinning status of
 * each function (which known answers / closed forms / brute-force checks pin
 * it) is listed in tests/test_oracle.py.
 *
 * Arrays use plain C layouts that are ABI-identical to the reference types:
 *   cpx8  : uint8_t, I = low nibble, Q = high nibble (src/pocket_sdr.h:105-110)
 *   cpx16 : int8_t pairs {I, Q}                       (src/pocket_sdr.h:111)
 *   cpx   : float pairs {re, im}                      (src/pocket_sdr.h:113)
 */
#ifndef SDR_ORACLE_H
#define SDR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_N_CODES 10            /* SDR_N_CODES, src/pocket_sdr.h:60 */
#define ORC_MAX_CORR 107          /* SDR_MAX_CORR, src/pocket_sdr.h:58 */

/* carrier-mixed LUT (src/sdr_func.c:113-134) */
void orc_func_init(void);
const int8_t *orc_mix_tbl(void);   /* 65536 x {I,Q} */

/* optional: route FP32 FFTs through the reference PocketFFT build
 * (oracle/_ref/libpocketfft_ref.so).  Returns 1 when loaded. */
int orc_load_ref_fft(const char *so_path);
int orc_ref_fft_loaded(void);

/* src/sdr_func.c:772-836 */
void orc_mix_carr(const uint8_t *buff, int buffN, int ix, int N, double fs,
                  double fc, double phi, int8_t *IQ);

/* src/sdr_func.c:861-912, 981-1008 */
void orc_corr_std(const int8_t *IQ, const int8_t *code, int N, double coff,
                  const double *pos, int n, float *corr, float *C);
/* src/sdr_func.c:915-956, 1029-1057 */
void orc_corr_std_cpx_code(const int8_t *IQ, const int8_t *code, int N,
                           double coff, const double *pos, int n, float *corr,
                           float *C);

/* src/sdr_func.c:197-233 */
void orc_cpx_mul(const float *a, const float *b, int N, float s, float *c);

/* src/sdr_func.c:236-264 (dir 0 forward, 1 backward, unnormalised) */
int orc_fft_f32(const float *in, int N, int dir, float *out);
/* FP64 DFT of any length (mixed radix + direct DFT for other primes) */
void orc_fft_f64(const double *in, int N, int dir, double *out);

/* src/sdr_func.c:1088-1102 */
void orc_corr_fft(const int8_t *IQ, const float *code_fft, int N, float *corr);

/* src/sdr_func.c:664-687 (the 1 ms courtesy sleep is left out) */
void orc_search_code(const float *code_fft, double T, const uint8_t *buff,
                     int buffN, int ix, int N, double fs, double fi,
                     const float *fds, int len_fds, float *P);
/* FP64 oracle: identical int8 mixing, FP64 FFT correlation, FP64 power */
void orc_search_code_f64(const double *code_fft, double T, const uint8_t *buff,
                         int buffN, int ix, int N, double fs, double fi,
                         const float *fds, int len_fds, double *P);

/* src/sdr_func.c:690-706, 728-744, 758-769 */
float orc_corr_max(const float *P, int N, int Nmax, int M, double T, int *ix);
double orc_fine_dop(const float *P, int N, const float *fds, int len_fds,
                    const int *ix);
int orc_dop_bins(double T, float dop, float max_dop, float *fds);

/* src/sdr_code.c:2556-2608 (real codes, code_Q = NULL) */
void orc_res_code(const int8_t *code, int len_code, double T, double coff,
                  double fs, int N, int Nz, int8_t *code_res);
void orc_gen_code_fft(const int8_t *code, int len_code, double T, double coff,
                      double fs, int N, int Nz, float *code_fft);
void orc_gen_code_fft_f64(const int8_t *code, int len_code, double T,
                          double coff, double fs, int N, int Nz,
                          double *code_fft);

/* GPS L1 C/A Gold code, src/sdr_code.c:792-812 (IS-GPS-200 G1/G2) */
int orc_gen_code_L1CA(int prn, int8_t *code);

/* Open-loop tracking epoch advance, src/sdr_ch.c:674-703, followed by
 * sdr_mix_carr + sdr_corr_std (src/sdr_ch.c:722-740).  One channel.
 * Outputs per epoch: corr[n_pos] (cpx), C[2] (cpx). */
typedef struct {
    double fs, fi, fc_sig, T;   /* sample rate, IF, signal carrier, code period */
    int N;                      /* samples per epoch */
    double fd0, drift;          /* Doppler (Hz) and its drift (Hz/s) */
    double phi0, coff0;         /* initial carrier phase (cyc), code offset (s) */
} orc_trk_ch_t;
void orc_trk_run(const orc_trk_ch_t *ch, const int8_t *code_bank,
                 const uint8_t *buff, int buffN, int n_epoch,
                 const double *pos, int n_pos, float *corr, float *C,
                 double *state);

/* Closed-loop tracking of one channel, src/sdr_ch.c:488-624, 694-763.
 * Per epoch: state[6] = {fd, coff, phi, cn0, lock, idle}, corr[n_pos] (cpx),
 * P (cpx). */
/* closed-loop state and the single-epoch updates of orc_trkl_run */
typedef struct {
    double T;
    int len_code, costas;
    double fd, coff, cn0, pli;
    double err_phas, phas_acc, code_int;
    double sumE, sumL, sumP, sumN, sumD;
    float C0[2];
    int lock, lost_cnt, pli_valid, idle;
    /* secondary code (sync_sec_code), bump-jump and the prompt history */
    int len_sec;                /* secondary code length (< 2: none) */
    const int8_t *sec_code;
    int sec_sync, sec_pol;
    double bump_step;           /* bump-jump code step (s); 0: no VE / VL taps */
    double sumVE, sumVL;
    float *hist;                /* ORC_N_HIST (re, im) prompts, newest last, or NULL */
} orc_tl_t;
#define ORC_N_HIST 5000         /* SDR_N_HIST, src/pocket_sdr.h:59 */
void orc_tl_fll(orc_tl_t *s, const float *C);
void orc_tl_pll(orc_tl_t *s, const float *C);
void orc_tl_dll(orc_tl_t *s, const float *C);
int orc_tl_cn0(orc_tl_t *s, const float *C, float Px, float Py);
void orc_tl_push(orc_tl_t *s, float Px, float Py);
int orc_tl_sec(orc_tl_t *s, float *C, int n_tap);
int orc_tl_update(orc_tl_t *s, float *C, float Px, float Py);
void orc_trkl_run(const orc_trk_ch_t *ch, int len_code, int costas, double cn0_0,
                  const int8_t *code_bank, const uint8_t *buff, int buffN, int n_epoch,
                  const double *pos, int n_pos, double *state, float *corr, float *P,
                  int len_sec, const int8_t *sec_code, double bump_step);

/* Prompt history of the standard correlator branch, src/sdr_ch.c:736-740:
 * P[e] = (C[e][0] + C[e-1][1]) / N, C[-1][1] = 0 (cpx arrays). */
void orc_trk_stitch(const float *C, int n_epoch, int N, float *P);

/* L6D / L6E CSK, src/sdr_ch.c:627-664: returns the chip shift ix of the
 * peak, C receives the n_pos interpolated taps (cpx). */
int orc_csk(const float *corr, int N, int len_code, const double *pos, int n_pos,
            float *C);
/* L6D / L6E FFT tracking, src/sdr_ch.c:694-720 (open loop as orc_trk_run).
 * Per epoch: syms[e], C[n_pos] (cpx), corr[N] (cpx, if corr_out), state[4]. */
void orc_trk_l6_run(const orc_trk_ch_t *ch, int len_code, const float *code_fft,
                    const uint8_t *buff, int buffN, int n_epoch, const double *pos,
                    int n_pos, uint8_t *syms, float *C, float *corr_out, double *state);

/* L6D / L6E closed-loop tracking (src/sdr_ch.c:694-720, 741-763): the FFT
 * correlator and CSK inside the FLL / PLL / DLL / C/N0 loop.  Per epoch:
 * state[6] = {fd, coff, phi, cn0, lock, idle}, syms[e], C[n_pos] (cpx). */
void orc_l6l_run(const orc_trk_ch_t *ch, int len_code, int costas, double cn0_0, const float *code_fft,
                 const uint8_t *buff, int buffN, int n_epoch, const double *pos, int n_pos,
                 double *state, uint8_t *syms, float *C);

/* Front-end raw data to cpx8 IF buffers, src/sdr_rcv.c:776-936.  fmt: the
 * SDR_FMT_* codes of src/pocket_sdr.h:85-92.  LUT generators: gen_LUTC
 * (256 cpx8), gen_LUTC_rtoc (4 x 256 cpx8, the fs/4 real-to-complex phases),
 * gen_LUT8 (256, or 65536 for CS16, int8 clipped to +-7). */
void orc_gen_lutc(int fmt, int rfch, int IQ, int bits, uint8_t *lut);
void orc_gen_lutc_rtoc(int rfch, uint8_t *lut);
void orc_gen_lut8(int fmt, double off, double scale, int8_t *lut);
/* write_buff for N samples of block index ix: lut[k] (cpx8 LUT of RF channel
 * k, or LUT8 / LUTQ for the int8 formats), rtoc[k] the real-to-complex flag
 * of channel k; out[k] receives N cpx8 samples of channel k.  Returns the
 * number of channels written. */
int orc_write_buff(int fmt, const uint8_t *raw, int N, long long ix, const int *rtoc,
                   const void *const *lut, uint8_t *const *out);
/* sdr_arch_combine, src/sdr_array.c:553-579, with build_lut
 * (src/sdr_array.c:486-498): one beam, w = Q8 weights [channel][re, im] */
void orc_arch_combine(int nrfch, const uint8_t *const *in, int N, const int16_t *w, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
