// psa_host.hip - host-side entry points of the drop-in ABI that do no
// sample-rate work: init, allocation helpers, and the per-PRN peak helpers
// over caller-owned host arrays (src/sdr_func.c:113-297, 690-769).
#include <math.h>
#include <string.h>
#include "psa_internal.h"

using namespace psa;

extern "C" {

// src/sdr_func.c:113-134.  The FFTW wisdom file has no meaning for the GPU
// FFT and is ignored; the carrier table goes to device constant memory.
void sdr_func_init(const char *file)
{
    (void)file;
    ensure_init();
    noalloc_setup();
}

// src/sdr_func.c:147-171 (fftwf_malloc: 32-byte aligned; exits on failure)
sdr_cpx_t *sdr_cpx_malloc(int N)
{
    void *p = nullptr;
    size_t bytes = sizeof(sdr_cpx_t) * (size_t)(N > 0 ? N : 1);
    if (posix_memalign(&p, 32, bytes) != 0 || !p) {
        fprintf(stderr, "sdr_cpx_t memory allocation error N=%d\n", N);
        exit(-1);
    }
    return (sdr_cpx_t *)p;
}

void sdr_cpx_free(sdr_cpx_t *cpx)
{
    free(cpx);
}

// src/sdr_func.c:182-185.  The reference build (-O3 -mavx2 -mfma, implicit
// contraction) fuses SQR(re) + SQR(im) into fmaf(re, re, im * im); the fused
// form is written out because this library compiles without contraction.
float sdr_cpx_abs(sdr_cpx_t cpx)
{
    return sqrtf(fmaf(cpx[0], cpx[0], cpx[1] * cpx[1]));
}

// src/sdr_func.c:276-297 (sdr_malloc = calloc, src/sdr_cmn.c:39-48)
sdr_buff_t *sdr_buff_new(int N, int IQ)
{
    sdr_buff_t *buff = (sdr_buff_t *)calloc(1, sizeof(sdr_buff_t));
    if (!buff) { fprintf(stderr, "memory allocation error\n"); exit(-1); }
    buff->data = (sdr_cpx8_t *)calloc(N > 0 ? N : 1, sizeof(sdr_cpx8_t));
    if (!buff->data) { fprintf(stderr, "memory allocation error\n"); exit(-1); }
    buff->N = N;
    buff->IQ = IQ;
    return buff;
}

void sdr_buff_free(sdr_buff_t *buff)
{
    if (!buff) return;
    free(buff->data);
    free(buff);
}

// src/sdr_func.c:690-706: argmax over bins 0..M-1 and lags 0..Nmax-1 of a
// row-major P with row stride N (strict '>': first maximum wins; P <= 0 leaves
// ix untouched), running mean in double, C/N0 in dB-Hz.
float sdr_corr_max(const float *P, int N, int Nmax, int M, double T, int *ix)
{
    double P_max = 0.0, P_ave = 0.0;
    int n = 0;
    for (int i = 0; i < M; i++) {
        for (int j = 0; j < Nmax; j++) {
            double v = P[(size_t)i * N + j];
            P_ave += (v - P_ave) / ++n;
            if (v <= P_max) continue;
            P_max = v;
            ix[0] = i;
            ix[1] = j;
        }
    }
    return (P_ave > 0.0) ? (float)(10.0 * log10((P_max - P_ave) / P_ave / T)) : 0.0f;
}

// src/sdr_func.c:709-744: quadratic through the three bins around the peak
// (the reference's least-squares fit of three points interpolates them); the
// vertex is solved in closed form.  Edge bins return the bin centre.
double sdr_fine_dop(const float *P, int N, const float *fds, int len_fds, const int *ix)
{
    if (ix[0] == 0 || ix[0] == len_fds - 1) return fds[ix[0]];
    double x[3], y[3];
    for (int i = 0; i < 3; i++) {
        x[i] = fds[ix[0] - 1 + i];
        y[i] = P[(size_t)(ix[0] - 1 + i) * N + ix[1]];
    }
    double d01 = x[0] - x[1], d02 = x[0] - x[2], d12 = x[1] - x[2];
    double den = d01 * d02 * d12;
    if (den == 0.0) return fds[ix[0]];
    double p2 = (x[2] * (y[1] - y[0]) + x[1] * (y[0] - y[2]) + x[0] * (y[2] - y[1])) / den;
    double p1 = (x[2] * x[2] * (y[0] - y[1]) + x[1] * x[1] * (y[2] - y[0]) +
                 x[0] * x[0] * (y[1] - y[2])) / den;
    if (p2 == 0.0) return fds[ix[0]];
    return -p1 / (2.0 * p2);
}

// src/sdr_func.c:758-769 (float step and bin values, caller frees with free)
float *sdr_dop_bins(double T, float dop, float max_dop, int *len_fds)
{
    float step = (float)(0.5 / T);
    *len_fds = (int)(2.0 * max_dop / step) + 1;
    float *fds = (float *)calloc(*len_fds > 0 ? *len_fds : 1, sizeof(float));
    if (!fds) { fprintf(stderr, "memory allocation error\n"); exit(-1); }
    for (int i = 0; i < *len_fds; i++) fds[i] = dop - max_dop + i * step;
    return fds;
}

}  // extern "C"
