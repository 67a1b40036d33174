// psa_loop.h - closed-loop tracking state and its per-epoch steps, shared
// by the standard-correlator loop (psa_corr.hip) and the L6D / L6E loop
// (psa_l6.hip): the FLL / PLL / DLL, C/N0 and loss decision of track_sig
// (src/sdr_ch.c:488-624, 694-763), the secondary-code sync and wipe-off
// (:453-486) and bump-jump (:551-572).  Every step runs on one lane per
// channel in FP64 with the reference build's fused multiply-adds written as
// explicit fma; atan and log10 are the GPU's, so results follow the
// reference to the last bits of those functions.  The loop constants are
// the reference's defaults (src/sdr_ch.c:29-49).
#pragma once
#include "psa_trk.h"

namespace psa {

#define TL_NTAP 4             // taps the loop reads: P, E, L, noise
#define TL_T_DLL 0.02         // T_DLL
#define TL_T_CN0 0.5          // T_CN0
#define TL_T_FPULLIN 1.0      // T_FPULLIN
#define TL_T_FPULLIN_W 0.5    // T_FPULLIN_W
#define TL_B_DLL 0.25         // B_DLL
#define TL_B_PLL 5.0          // B_PLL
#define TL_B_FLL_W 5.0        // B_FLL_W
#define TL_B_FLL_N 2.0        // B_FLL_N
#define TL_THRES_CN0_U 30.0   // THRES_CN0_U
#define TL_THRES_PLI 0.25     // THRES_PLI
#define TL_LOST_TH 4          // LOST_TH
#define TL_FILT_CN0 0.5       // FILT_CN0
#define TL_DPI (2.0 * 3.1415926535897932)

#define TL_N_HIST 5000          // SDR_N_HIST (src/pocket_sdr.h:59)
#define TL_T_NPULLIN 1.5        // T_NPULLIN
#define TL_THRES_SYNC 0.02      // THRES_SYNC
#define TL_THRES_LOST 0.002     // THRES_LOST
#define TL_THRES_SEC_RATIO 0.8  // THRES_SEC_RATIO
#define TL_BUMP_K 1.3           // BUMP_K

// one channel's tracking state (the sdr_ch_t / sdr_trk_t fields the loop uses)
struct TrkLoopCh {
    double fs, fi, fc_sig, T;
    double time, fd, coff, phi, cn0, pli;
    double err_phas, phas_acc, code_int, err_code;
    double sumE, sumL, sumP, sumN, sumD;   // DLL E / L, C/N0 windows
    double bump_step, sumVE, sumVL;        // bump-jump step (0: none) and VE / VL sums
    float C0x, C0y, C1x, C1y;              // FLL previous prompt, stitch carry
    int lock, lost_cnt, pli_valid, idle, len_code, costas, ep;
    int len_sec, sec_off, sec_sync, sec_pol;   // secondary code (length < 2: none)
    int hb;                                // prompt history: logical P[0] at hist[hb]
};

// The prompt history trk->P (SDR_N_HIST entries, newest last) as a ring in
// LDS: logical entry i at hist[(hb + i) mod TL_N_HIST].  sdr_add_buff drops
// the oldest and appends; adj_coff's memmoves shift the whole history by one
// entry either way (src/sdr_ch.c:674-689), each an index step plus the one
// entry the memmove leaves in place.
__device__ __forceinline__ float2 &tl_hist(float2 *hist, const TrkLoopCh &S, int i)
{
    int k = S.hb + i;
    if (k >= TL_N_HIST) k -= TL_N_HIST;
    return hist[k];
}
__device__ __forceinline__ void tl_hist_push(float2 *hist, TrkLoopCh &S, float2 v)
{
    S.hb = S.hb + 1 == TL_N_HIST ? 0 : S.hb + 1;
    tl_hist(hist, S, TL_N_HIST - 1) = v;
}

// per (channel, epoch): state after the epoch (idle: the channel was lost at
// or before it and the epoch was not processed)
struct TrkLoopOut {
    double fd, coff, phi, cn0;
    int lock, idle;
    float2 P;
};

// lane 0: the NCO and code-offset advance of track_sig (src/sdr_ch.c:
// 694-703) by the loop's Doppler, with adj_coff (674-689) moving the lock
// count and, for a channel with a secondary code, the prompt history
__device__ __forceinline__ void tl_nco(TrkLoopCh &S, float2 *hist)
{
#pragma clang fp contract(off)
    const double t = (double)(S.ep + 1) * S.T;
    const double tau = t - S.time;
    const double fc = S.fi + S.fd;
    S.coff = fma(-(S.fd / S.fc_sig), tau, S.coff);
    S.phi = fma(fc, tau, S.phi);
    S.phi = S.phi - floor(S.phi);
    S.time = t;
    if (S.coff >= S.T) {
        S.coff = S.coff - S.T; S.lock--;
        if (S.len_sec >= 2) {
            // memmove(P + 1, P): every entry one later, P[0] kept
            const float2 p0 = tl_hist(hist, S, 0);
            S.hb = S.hb == 0 ? TL_N_HIST - 1 : S.hb - 1;
            tl_hist(hist, S, 0) = p0;
        }
    } else if (S.coff < 0.0) {
        S.coff = S.coff + S.T; S.lock++;
        if (S.len_sec >= 2) {
            // memmove(P, P + 1): every entry one earlier, P[N_HIST - 1] kept
            const float2 pl = tl_hist(hist, S, TL_N_HIST - 1);
            S.hb = S.hb + 1 == TL_N_HIST ? 0 : S.hb + 1;
            tl_hist(hist, S, TL_N_HIST - 1) = pl;
        }
    }
}

// lane 0: tl_nco and the standard correlators' mixing parameters of the
// epoch (src/sdr_ch.c:722-734)
__device__ __forceinline__ void tl_advance(TrkLoopCh &S, int N, int buffN, TrkEpoch &q, float2 *hist)
{
#pragma clang fp contract(off)
    tl_nco(S, hist);
    q.ix = (int)(((long long)S.ep * N) % buffN);
    trk_mix_params(buffN, q.ix, N, S.fs, S.fi + S.fd, S.phi, q.p0, q.p1, q.nseg, q.s);
    q.idle = 0;
    q.neg = 0;
    q.cs = S.coff * S.fs;
}

// lane 0: sync_sec_code (src/sdr_ch.c:453-486) over the last N = len_sec
// prompts: unsynchronised, P = sum x_i c_i / N and R = sum |x_i| / N in
// float (x the in-phase prompts, c the code) declare sync at this lock count
// with the sign of P when |P| >= 0.8 R and R >= 0.02; synchronised, every N
// epochs the plain mean below 0.002 drops it.  Returns the multiplier of
// this epoch's taps and newest prompt (1 when not synchronised).  A lock
// count more than N below the sync point (the reference's index then goes
// negative) wraps into the code.
__device__ __forceinline__ int tl_sec(TrkLoopCh &S, float2 *hist, const int8_t *__restrict__ sec)
{
#pragma clang fp contract(off)
    const int N = S.len_sec;
    const float fN = (float)N;
    if (S.sec_sync == 0) {
        float P = 0.0f, R = 0.0f;
        for (int i = 0; i < N; i++) {
            const float x = tl_hist(hist, S, TL_N_HIST - N + i).x;
            P = __fadd_rn(P, __fdiv_rn(__fmul_rn(x, (float)sec[i % N]), fN));
            R = __fadd_rn(R, __fdiv_rn(fabsf(x), fN));
        }
        if ((double)fabsf(P) >= TL_THRES_SEC_RATIO * (double)R && (double)R >= TL_THRES_SYNC) {
            S.sec_sync = S.lock;
            S.sec_pol = P > 0.0f ? 1 : -1;
        }
    } else if ((S.lock - S.sec_sync) % N == 0) {
        float P = 0.0f;
        for (int i = 0; i < N; i++) P = __fadd_rn(P, __fdiv_rn(tl_hist(hist, S, TL_N_HIST - N + i).x, fN));
        if ((double)fabsf(P) < TL_THRES_LOST) S.sec_sync = S.sec_pol = 0;
    }
    if (S.sec_sync <= 0) return 1;
    int idx = (S.lock - S.sec_sync - 1 + N) % N % N;
    if (idx < 0) idx += N;
    return sec[idx] * S.sec_pol;
}

// lane 0: the loop after the correlation (src/sdr_ch.c:453-620, 740-763)
// given the epoch's prompt (Px, Py) for the history; tap 0 = P, 1 = E,
// 2 = L, 3 = the noise tap (4, 5 = VE, VL with bump-jump), wiped in place;
// q.neg records a negated epoch
__device__ __forceinline__ void tl_update_p(TrkLoopCh &S, float2 *tap, float Px, float Py, TrkLoopOut &o,
                                            float2 *hist, const int8_t *__restrict__ sec, TrkEpoch &q)
{
#pragma clang fp contract(off)
    if (S.len_sec >= 2) tl_hist_push(hist, S, make_float2(Px, Py));
    S.lock++;
    // secondary code (src/sdr_ch.c:743-746)
    if (S.len_sec >= 2 && S.lock * S.T >= TL_T_NPULLIN) {
        const int m = tl_sec(S, hist, sec);
        if (m < 0) {
            const int nt = S.bump_step > 0.0 ? 6 : TL_NTAP;
            for (int i = 0; i < nt; i++) tap[i] = make_float2(-tap[i].x, -tap[i].y);
            float2 &pn = tl_hist(hist, S, TL_N_HIST - 1);
            pn = make_float2(-pn.x, -pn.y);
            q.neg = 1;
        }
        Px = tl_hist(hist, S, TL_N_HIST - 1).x;
        Py = tl_hist(hist, S, TL_N_HIST - 1).y;
    }
    if (S.lock * S.T <= TL_T_FPULLIN) {
        // FLL (src/sdr_ch.c:488-506)
        if (S.lock >= 2) {
            const double IP1 = tap[0].x, QP1 = tap[0].y, IP2 = S.C0x, QP2 = S.C0y;
            const double dot = fma(IP1, IP2, QP1 * QP2);
            const double cross = fma(-QP1, IP2, IP1 * QP2);
            if (dot != 0.0) {
                const double B = S.lock * S.T < TL_T_FPULLIN_W ? TL_B_FLL_W : TL_B_FLL_N;
                const double err = S.costas ? atan(cross / dot) : atan2(cross, dot);
                S.fd = S.fd - B / 0.25 * err / TL_DPI;
            }
        }
        S.C0x = tap[0].x;
        S.C0y = tap[0].y;
    } else {
        // PLL (src/sdr_ch.c:509-521)
        const double IP = tap[0].x, QP = tap[0].y;
        if (IP != 0.0) {
            const double err = (S.costas ? atan(QP / IP) : atan2(QP, IP)) / TL_DPI;
            const double W = TL_B_PLL / 0.7845;
            S.phas_acc = fma(W * W * W * err, S.T, S.phas_acc);
            const double u = 1.1 * W * W * err * S.T;
            S.fd = S.fd + fma(S.T, S.phas_acc, fma(2.4 * W, err - S.err_phas, u));
            S.err_phas = err;
        }
    }
    // DLL (src/sdr_ch.c:524-549): the E and L power sums
    const int ND = max(1, (int)(TL_T_DLL / S.T));
    S.sumE = S.sumE + (double)__fmaf_rn(tap[1].x, tap[1].x, __fmul_rn(tap[1].y, tap[1].y));
    S.sumL = S.sumL + (double)__fmaf_rn(tap[2].x, tap[2].x, __fmul_rn(tap[2].y, tap[2].y));
    if (S.lock % ND == 0) {
        const double E = sqrt(S.sumE), L = sqrt(S.sumL);
        if (E + L > 0.0) {
            const double err = (E - L) / (E + L) * 0.5 * S.T / S.len_code;
            const double W = TL_B_DLL / 0.53;
            const double dt = S.T * ND;
            S.code_int = fma(W * W * err, dt, S.code_int);
            S.coff = fma(-fma(1.414 * W, err, S.code_int), dt, S.coff);
            S.err_code = err;
        }
        S.sumE = S.sumL = 0.0;
    }
    // C/N0 and carrier lock (src/sdr_ch.c:567-605)
    S.sumP = S.sumP + (double)__fmaf_rn(Px, Px, __fmul_rn(Py, Py));
    S.sumD = S.sumD + (double)__fmaf_rn(Px, Px, -__fmul_rn(Py, Py));
    S.sumN = S.sumN + (double)__fmaf_rn(tap[3].x, tap[3].x, __fmul_rn(tap[3].y, tap[3].y));
    if (S.bump_step > 0.0) {                   // npos >= 6: VE, VL (src/sdr_ch.c:586-589)
        S.sumVE = S.sumVE + (double)__fmaf_rn(tap[4].x, tap[4].x, __fmul_rn(tap[4].y, tap[4].y));
        S.sumVL = S.sumVL + (double)__fmaf_rn(tap[5].x, tap[5].x, __fmul_rn(tap[5].y, tap[5].y));
    }
    if (S.lock % (int)(TL_T_CN0 / S.T) == 0) {
        if (S.sumP > 0.0) {
            S.pli = S.sumD / S.sumP;
            S.pli_valid = 1;
        }
        if (S.sumN > 0.0) {
            const double l = log10(S.sumP / S.sumN / S.T);
            S.cn0 = fma(fma(l, 10.0, -S.cn0), TL_FILT_CN0, S.cn0);
        }
        if (S.bump_step > 0.0) {
            // bump_jump (src/sdr_ch.c:552-572)
            if (S.sumVL > TL_BUMP_K * S.sumP && S.sumP > TL_BUMP_K * S.sumVE) S.coff = S.coff + S.bump_step;
            else if (S.sumVE > TL_BUMP_K * S.sumP && S.sumP > TL_BUMP_K * S.sumVL) S.coff = S.coff - S.bump_step;
            S.sumVE = S.sumVL = 0.0;
        }
        S.sumP = S.sumN = S.sumD = 0.0;
        // loss of lock (src/sdr_ch.c:608-624)
        const bool bad = S.cn0 < TL_THRES_CN0_U || (S.costas && S.pli_valid && S.pli < TL_THRES_PLI);
        S.lost_cnt = bad ? S.lost_cnt + 1 : 0;
        if (S.lost_cnt >= TL_LOST_TH) {
            S.idle = 1;
            S.lock = 0;
        }
    }
    S.ep++;
    o.fd = S.fd; o.coff = S.coff; o.phi = S.phi; o.cn0 = S.cn0;
    o.lock = S.lock; o.idle = 0;
    o.P = make_float2(Px, Py);
}

// lane 0: tl_update_p with the standard correlators' prompt: this epoch's
// correlation before the code wrap (C[0]) joined with the previous epoch's
// after it, (C1[0] + trk->C1) / N in float (src/sdr_ch.c:736-740)
__device__ __forceinline__ void tl_update(TrkLoopCh &S, int N, float2 *tap, const float2 *C, TrkLoopOut &o,
                                          float2 *hist, const int8_t *__restrict__ sec, TrkEpoch &q)
{
#pragma clang fp contract(off)
    const float fN = (float)N;
    const float Px = __fdiv_rn(__fadd_rn(C[0].x, S.C1x), fN), Py = __fdiv_rn(__fadd_rn(C[0].y, S.C1y), fN);
    S.C1x = C[1].x;
    S.C1y = C[1].y;
    tl_update_p(S, tap, Px, Py, o, hist, sec, q);
}

}  // namespace psa
