/*
 * harness.c - a C caller of libpocketsdr_amd through include/pocket_sdr_amd.h
 * only, the way the reference applications call libsdr (test infrastructure:
 * it also links the CPU oracle, oracle/liboracle.so, as the checker).
 *
 *   acq      search_sig of app/pocket_acq/pocket_acq.c:152-192: sdr_search_code
 *            over every block of the buffer with P accumulated, then
 *            sdr_corr_max and sdr_fine_dop; P against the oracle's loop
 *            (normwise <= 1e-5), the same peak (bin, lag), C/N0 within 1e-3
 *            relative, Doppler within 0.5 Hz.
 *   trk      track_sig's per-epoch sdr_mix_carr + sdr_corr_std with the
 *            prompt stitch (src/sdr_ch.c:722-740) on an open-loop NCO:
 *            mixed samples, taps and stitched prompts bit-exact.
 *   threads  NT host threads (the receiver's one pthread per channel,
 *            src/sdr_rcv.c:495-521), each NC epochs of its own channel's
 *            sdr_mix_carr + sdr_corr_std, every tenth epoch also an
 *            sdr_search_code, all at the same time: bit-exact / <= 1e-5.
 *
 * Prints one line per check and "harness ok" when all pass; exit status 0
 * on success, 1 on a mismatch.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pocket_sdr_amd.h"
#include "sdr_oracle.h"

#define FS 12e6
#define T_CODE 1e-3
#define N_SMP 12000                /* samples per code period at FS */
#define N_BUF (N_SMP * 12)         /* 12 ms of IF */
#define F_L1 1575.42e6

static uint8_t g_if[N_BUF];

static uint32_t lcg(uint32_t *s)
{
    *s = *s * 1664525u + 1013904223u;
    return *s >> 8;
}

/* cpx8 IF: PRN 5 at 1250 Hz Doppler and code phase 3711 samples, amplitude
 * 2.5 per component, uniform noise of +-2, rounded and clipped to [-7, 7] */
static void make_if(void)
{
    int8_t code[1023];
    orc_gen_code_L1CA(5, code);
    uint32_t s = 12345u;
    const double PI = 3.14159265358979323846;
    for (int i = 0; i < N_BUF; i++) {
        const double t = (double)((i + N_BUF - 3711) % N_BUF) / FS;
        const int c = code[(int)(t * 1.023e6) % 1023];
        const double ph = 2.0 * PI * 1250.0 * i / FS;
        double re = 2.5 * c * cos(ph) + ((lcg(&s) & 0xFFFF) / 65536.0 - 0.5) * 4.0;
        double im = 2.5 * c * sin(ph) + ((lcg(&s) & 0xFFFF) / 65536.0 - 0.5) * 4.0;
        int I = (int)lrint(re), Q = (int)lrint(im);
        I = I < -7 ? -7 : I > 7 ? 7 : I;
        Q = Q < -7 ? -7 : Q > 7 ? 7 : Q;
        g_if[i] = (uint8_t)(((Q & 0xF) << 4) | (I & 0xF));
    }
}

static double relnorm(const float *a, const float *b, size_t n)
{
    double d = 0.0, r = 0.0;
    for (size_t i = 0; i < n; i++) {
        d += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
        r += (double)b[i] * b[i];
    }
    return r > 0.0 ? sqrt(d / r) : sqrt(d);
}

static int check_acq(void)
{
    int8_t code[1023];
    orc_gen_code_L1CA(5, code);
    const int N = N_SMP, M = 2 * N;
    sdr_cpx_t *code_fft = sdr_cpx_malloc(M);
    orc_gen_code_fft(code, 1023, T_CODE, 0.0, FS, N, N, (float *)code_fft);
    int len_fds = 0;
    float *fds = sdr_dop_bins(T_CODE, 0.0f, 5000.0f, &len_fds);
    float *P = calloc((size_t)M * len_fds, sizeof(float));
    float *R = calloc((size_t)M * len_fds, sizeof(float));
    sdr_buff_t buff = {g_if, 2, N_BUF};
    for (int i = 0; i < N_BUF - 2 * N + 1; i += N) {
        sdr_search_code((const sdr_cpx_t *)code_fft, T_CODE, &buff, i, M, FS, 0.0, fds, len_fds, P);
        orc_search_code((const float *)code_fft, T_CODE, g_if, N_BUF, i, M, FS, 0.0, fds, len_fds, R);
    }
    int ix[2] = {0}, rx[2] = {0};
    const float cn0 = sdr_corr_max(P, M, N, len_fds, T_CODE, ix);
    const float rcn0 = orc_corr_max(R, M, N, len_fds, T_CODE, rx);
    const double dop = sdr_fine_dop(P, M, fds, len_fds, ix);
    const double rdop = orc_fine_dop(R, M, fds, len_fds, rx);
    const double e = relnorm(P, R, (size_t)M * len_fds);
    const int ok = e <= 1e-5 && ix[0] == rx[0] && ix[1] == rx[1] && fabs(cn0 - rcn0) <= 1e-3 * fabs(rcn0) &&
                   fabs(dop - rdop) <= 0.5;
    printf("acq: P rel err %.2e, peak (%d, %d) / oracle (%d, %d), C/N0 %.3f / %.3f dB-Hz, "
           "Doppler %.2f / %.2f Hz, code offset %.3f us: %s\n", e, ix[0], ix[1], rx[0], rx[1], cn0, rcn0,
           dop, rdop, ix[1] / FS * 1e6, ok ? "ok" : "MISMATCH");
    sdr_cpx_free(code_fft);
    free(fds);
    free(P);
    free(R);
    return ok;
}

/* one channel's epochs (track_sig's mix + correlate, open-loop NCO) */
typedef struct {
    int prn, n_epoch, with_search;
    double fd, phi, coff;
    int bad;
    double search_err;
} chan_t;

#define N_POS 7

static void *run_channel(void *arg)
{
    chan_t *c = (chan_t *)arg;
    const int N = N_SMP;
    int8_t code[1023];
    orc_gen_code_L1CA(c->prn, code);
    int8_t *bank = malloc((size_t)2 * N * SDR_N_CODES);
    for (int k = 0; k < SDR_N_CODES; k++) {
        orc_res_code(code, 1023, T_CODE, -k / FS / SDR_N_CODES, FS, N, 0, bank + (size_t)2 * N * k);
    }
    const double sc = T_CODE / 1023 * FS;   /* samples per chip */
    const double pos[N_POS] = {0.0, -0.5 * sc, 0.5 * sc, -120.0, -2.0 * sc, 1.5 * sc, 0.25};
    sdr_cpx16_t *IQ = malloc(sizeof(sdr_cpx16_t) * N);
    int8_t *rIQ = malloc((size_t)2 * N);
    sdr_cpx_t corr[N_POS], C1[2], C1p = {0.0f, 0.0f}, P;
    float rcorr[2 * N_POS], rC1[4], rC1p[2] = {0.0f, 0.0f}, rP[2];
    sdr_cpx_t *code_fft = NULL;
    float *fds = NULL, *SP = NULL, *SR = NULL;
    int len_fds = 0;
    if (c->with_search) {
        code_fft = sdr_cpx_malloc(2 * N);
        orc_gen_code_fft(code, 1023, T_CODE, 0.0, FS, N, N, (float *)code_fft);
        fds = sdr_dop_bins(T_CODE, 0.0f, 1000.0f, &len_fds);
        SP = malloc(sizeof(float) * 2 * N * len_fds);
        SR = malloc(sizeof(float) * 2 * N * len_fds);
    }
    sdr_buff_t buff = {g_if, 2, N_BUF};
    double phi = c->phi, coff = c->coff;
    for (int e = 0; e < c->n_epoch; e++) {
        const int ix = (int)(((long long)e * N) % N_BUF);
        const double fc = c->fd;                     /* IF 0 */
        sdr_mix_carr(&buff, ix, N, FS, fc, phi, IQ);
        orc_mix_carr(g_if, N_BUF, ix, N, FS, fc, phi, rIQ);
        if (memcmp(IQ, rIQ, (size_t)2 * N)) c->bad++;
        sdr_corr_std(IQ, (const sdr_cpx16_t *)bank, N, coff * FS, pos, N_POS, corr, C1);
        orc_corr_std(rIQ, bank, N, coff * FS, pos, N_POS, rcorr, rC1);
        if (memcmp(corr, rcorr, sizeof(rcorr)) || memcmp(C1, rC1, sizeof(rC1))) c->bad++;
        /* prompt stitch, src/sdr_ch.c:736-740 */
        for (int i = 0; i < 2; i++) {
            P[i] = (C1[0][i] + C1p[i]) / N;
            C1p[i] = C1[1][i];
            rP[i] = (rC1[i] + rC1p[i]) / N;
            rC1p[i] = rC1[2 + i];
        }
        if (memcmp(P, rP, sizeof(rP))) c->bad++;
        if (c->with_search && e % 10 == 0) {
            memset(SP, 0, sizeof(float) * 2 * N * len_fds);
            memset(SR, 0, sizeof(float) * 2 * N * len_fds);
            sdr_search_code((const sdr_cpx_t *)code_fft, T_CODE, &buff, ix, 2 * N, FS, 0.0, fds, len_fds, SP);
            orc_search_code((const float *)code_fft, T_CODE, g_if, N_BUF, ix, 2 * N, FS, 0.0, fds, len_fds, SR);
            const double er = relnorm(SP, SR, (size_t)2 * N * len_fds);
            if (er > c->search_err) c->search_err = er;
            if (er > 1e-5) c->bad++;
        }
        /* open-loop NCO and code advance, src/sdr_ch.c:694-703 */
        phi += fc * T_CODE;
        phi -= floor(phi);
        coff -= c->fd / F_L1 * T_CODE;
        if (coff < 0.0) coff += T_CODE;
    }
    free(bank);
    free(IQ);
    free(rIQ);
    if (code_fft) {
        sdr_cpx_free(code_fft);
        free(fds);
        free(SP);
        free(SR);
    }
    return NULL;
}

static int check_trk(void)
{
    chan_t c = {5, 60, 0, 1250.0, 0.3, 3711.0 / FS, 0, 0.0};
    run_channel(&c);
    printf("trk: 60 epochs of sdr_mix_carr + sdr_corr_std (%d taps) + prompt stitch: %s\n", N_POS,
           c.bad ? "MISMATCH" : "bit-exact");
    return c.bad == 0;
}

static int check_threads(int nt, int nc)
{
    chan_t *c = calloc((size_t)nt, sizeof(chan_t));
    pthread_t *th = calloc((size_t)nt, sizeof(pthread_t));
    for (int t = 0; t < nt; t++) {
        c[t].prn = 1 + t % 32;
        c[t].n_epoch = nc;
        c[t].with_search = 1;
        c[t].fd = -4000.0 + 1000.0 * t;
        c[t].phi = 0.1 * t;
        c[t].coff = (1000.0 * t + 17.0) / FS;
        pthread_create(&th[t], NULL, run_channel, &c[t]);
    }
    int bad = 0;
    double worst = 0.0;
    for (int t = 0; t < nt; t++) {
        pthread_join(th[t], NULL);
        bad += c[t].bad;
        if (c[t].search_err > worst) worst = c[t].search_err;
    }
    printf("threads: %d host threads x %d epochs of mix + corr_std (+ sdr_search_code every 10th), "
           "concurrent: %d mismatches, worst search P rel err %.2e: %s\n", nt, nc, bad, worst,
           bad ? "MISMATCH" : "ok");
    free(c);
    free(th);
    return bad == 0;
}

int main(int argc, char **argv)
{
    const char *ref = getenv("ORACLE_REF_FFT");
    if (ref && *ref) printf("oracle FFT: reference PocketFFT %s\n", orc_load_ref_fft(ref) ? "loaded" : "not loaded");
    sdr_func_init("");
    orc_func_init();
    make_if();
    int ok = 1;
    const int nt = argc > 1 ? atoi(argv[1]) : 8, nc = argc > 2 ? atoi(argv[2]) : 50;
    ok &= check_acq();
    ok &= check_trk();
    ok &= check_threads(nt, nc);
    printf("%s\n", ok ? "harness ok" : "harness FAILED");
    return ok ? 0 : 1;
}
