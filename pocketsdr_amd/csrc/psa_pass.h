// psa_pass.h - device-side element loaders and storers shared by the FP32
// pass kernels (psa_fft.hip) and the FP16 / BF16 MFMA pass kernels
// (psa_fft16.hip).
#pragma once
#include "psa_internal.h"

namespace psa {

#define PASS_THREADS 256

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -a.y * b.y), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// multiply by -i (forward) or +i (backward): sgn = -1 forward, +1 backward
__device__ __forceinline__ float2 mul_i(float2 a, float sgn) { return make_float2(-sgn * a.y, sgn * a.x); }

// ---------------------------------------------------------------------------
// wave reduction of a tracked peak (value bm at index bi, -1 = none) and a
// sum bs: the larger value wins, the smaller index on ties (a commutative
// rule, so a lane and its partner reach the same record).  The whole wave
// must be active.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void peak_take(float &bm, long long &bi, float om, long long oi)
{
    if (oi >= 0 && (bi < 0 || om > bm || (om == bm && oi < bi))) { bm = om; bi = oi; }
}

// partner within a row of 16 by DPP control CTRL (quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror, row_mirror): after the four steps every lane
// of a row holds the row's record
template <int CTRL>
__device__ __forceinline__ void peak_dpp_step(float &bm, long long &bi, double &bs)
{
    auto mv = [](unsigned v) { return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true); };
    const unsigned long long ui = (unsigned long long)bi, us = (unsigned long long)__double_as_longlong(bs);
    const float om = __uint_as_float(mv(__float_as_uint(bm)));
    const long long oi = (long long)(((unsigned long long)mv((unsigned)(ui >> 32)) << 32) | mv((unsigned)ui));
    bs += __longlong_as_double((long long)(((unsigned long long)mv((unsigned)(us >> 32)) << 32) | mv((unsigned)us)));
    peak_take(bm, bi, om, oi);
}

// rows 0, 1 with rows 2, 3 (W32 = false: row 0 with 1 and 2 with 3; true:
// lanes 0-31 with 32-63) by v_permlane16_swap / v_permlane32_swap of a value
// with itself: the two outputs hold the lower and the upper group's value in
// every lane of both groups
template <bool W32>
__device__ __forceinline__ void peak_swap_step(float &bm, long long &bi, double &bs)
{
    auto sw = [](unsigned v) {
        if constexpr (W32) return __builtin_amdgcn_permlane32_swap(v, v, false, false);
        else return __builtin_amdgcn_permlane16_swap(v, v, false, false);
    };
    const unsigned long long ui = (unsigned long long)bi, us = (unsigned long long)__double_as_longlong(bs);
    const auto m = sw(__float_as_uint(bm)), ih = sw((unsigned)(ui >> 32)), il = sw((unsigned)ui);
    const auto sh = sw((unsigned)(us >> 32)), sl = sw((unsigned)us);
    bm = __uint_as_float(m[0]);
    bi = (long long)(((unsigned long long)ih[0] << 32) | il[0]);
    bs = __longlong_as_double((long long)(((unsigned long long)sh[0] << 32) | sl[0])) +
         __longlong_as_double((long long)(((unsigned long long)sh[1] << 32) | sl[1]));
    peak_take(bm, bi, __uint_as_float(m[1]), (long long)(((unsigned long long)ih[1] << 32) | il[1]));
}

// ---------------------------------------------------------------------------
// element loaders / storers of the pass kernels
// ---------------------------------------------------------------------------
// The product of sdr_cpx_mul (src/sdr_func.c:197-233) times its scale s:
// the AVX2 lanes compute re = fl(ar br) + fl(ai (-bi)), im = fl(ar bi) +
// fl(ai br) (src/sdr_func.c:203-213); the last M % 4 natural indices (tail)
// take the scalar form, contracted by the reference build
// (src/sdr_func.c:229-232).
__device__ __forceinline__ float2 prod_value(float2 x, float2 c, bool tail, float s)
{
    float re, im;
    if (!tail) {
        re = x.x * c.x + x.y * -c.y;
        im = x.x * c.y + x.y * c.x;
    } else {
        re = __builtin_fmaf(x.x, c.x, -(x.y * c.y));
        im = __builtin_fmaf(x.x, c.y, x.y * c.x);
    }
    return make_float2(re * s, im * s);
}

template <int LM>
struct Loader {
    const float2 *pin = nullptr;
    const char2 *qin = nullptr;
    const float2 *xin = nullptr, *cin = nullptr;
    const uint8_t *buf = nullptr;
    const uint32_t *carr = nullptr;     // LDS copy of the carrier table
    int ix = 0, nseg = 0, compact = 0;
    int rbps = 0, roff = 0, rrtoc = 0, rph0 = 0;     // packed raw ring (rbps > 0)
    const uint8_t *rlut = nullptr;                  // its table (LDS, after the carrier table)
    uint32_t p0 = 0, p1 = 0, s = 0;
    bool prod_tail = false;
    int prod_V = 0, prod_M1 = 1, prod_M2 = 1;
    float prod_scale = 1.0f;

    __device__ __forceinline__ void init(const PassArgs &a, int b, const uint32_t *carr_lds)
    {
        if (LM == LD_CPX) pin = a.in + (long long)b * a.in_bstride;
        if (LM == LD_CPX16) qin = reinterpret_cast<const char2 *>(a.iq16) + (long long)b * a.in_bstride;
        if (LM == LD_MIX) {
            carr = carr_lds; buf = a.buf; compact = a.mix_compact;
            ix = a.mix_ix[b]; nseg = a.mix_nseg[b];
            p0 = a.mix_p0[b]; p1 = a.mix_p1[b]; s = a.mix_s[b];
            rbps = a.raw_bps; roff = a.raw_off; rrtoc = a.raw_rtoc; rph0 = a.raw_ph0;
            rlut = reinterpret_cast<const uint8_t *>(carr_lds + 256);
        }
        if (LM == LD_PROD) {
            const long long xb = (b / a.prod_xdiv) % a.prod_xmod;
            const long long cb = a.cf_index ? a.cf_index[b] : b / a.prod_cdiv;
            xin = a.in + xb * a.in_bstride;
            cin = a.cf + cb * a.cf_bstride;
            prod_tail = a.prod_V < a.M;
            prod_V = a.prod_V; prod_M1 = a.prod_M1; prod_M2 = a.prod_M2;
            prod_scale = a.prod_scale;
        }
    }

    // LD_MIX: the carrier-mixed int8 pair of sample e, as mix_tbl holds it
    // (src/sdr_func.c:113-134, 786-806): mix_byte fetches the IF byte,
    // mix_prod forms the product (split so that callers can keep loads in
    // flight)
    __device__ __forceinline__ int mix_index(int e) const
    {
        return e < nseg ? ix + e : (compact ? e : e - nseg);
    }
    // the IF byte (raw rings: the byte holding the channel's bits)
    __device__ __forceinline__ uint32_t mix_byte(int e) const
    {
        const int i = mix_index(e);
        return rbps ? buf[i * rbps + roff] : buf[i];
    }
    __device__ __forceinline__ int2 mix_int(int e) const { return mix_prod(e, mix_byte(e)); }
    __device__ __forceinline__ int2 mix_prod(int e, uint32_t d) const
    {
        if (rbps) d = rlut[(rrtoc ? ((rph0 + mix_index(e)) & 3) << 8 : 0) + d];
        const uint32_t ph = e < nseg ? p0 + (uint32_t)e * s : p1 + (uint32_t)(e - nseg) * s;
        const int sI = (int)((int8_t)(uint8_t)(d << 4)) >> 4;
        const int sQ = (int)((int8_t)(uint8_t)d) >> 4;
        const uint32_t c = carr[ph >> 24];
        const int cI = (int8_t)(c & 0xFF), cQ = (int8_t)(c >> 8);
        return make_int2((int8_t)(sI * cI - sQ * cQ), (int8_t)(sI * cQ + sQ * cI));
    }

    // element e of the batch item (natural index for LD_MIX / LD_CPX16)
    __device__ __forceinline__ float2 load(int e) const
    {
        if (LM == LD_CPX) return pin[e];
        if (LM == LD_CPX16) {
            const char2 q = qin[e];
            return make_float2((float)q.x * PSA_CSCALE, (float)q.y * PSA_CSCALE);
        }
        if (LM == LD_MIX) {
            const int2 m = mix_int(e);
            return make_float2((float)m.x * PSA_CSCALE, (float)m.y * PSA_CSCALE);
        }
        // LD_PROD: X * Cf with the sdr_cpx_mul arithmetic
        bool tail = false;
        if (prod_tail) tail = (e / prod_M2 + prod_M1 * (e % prod_M2)) >= prod_V;
        return prod_value(xin[e], cin[e], tail, prod_scale);
    }
};

template <int SM>
struct Storer {
    float2 *pout = nullptr;
    float *prow = nullptr;
    const float2 *thi = nullptr, *tlo = nullptr;
    int nlag = 0, ogs = 0, ons = 0;
    int win = 0, wM = 0;            // ST_CPX lag window (0: every index)
    bool pinit = false, track = false;
    long long ibase = 0;            // row-major index of lag 0 of this row
    float bm = 0.0f;                // peak tracking of the values stored:
    int be = -1;                    // lag of the maximum in this row (-1: none)
    double bs = 0.0;

    __device__ __forceinline__ void init(const PassArgs &a, int b)
    {
        thi = a.thi; tlo = a.tlo;
        ogs = (int)a.out_gstride; ons = (int)a.out_nstride;
        if (SM == ST_CPX) {
            pout = a.out + (long long)b * a.out_bstride;
            win = a.out_win;
            wM = a.M;
        }
        if (SM == ST_POW) {
            const long long row = (long long)(b / a.p_div + a.p_off1) * a.p_mul + a.p_off2 + b % a.p_div;
            prow = a.P + row * a.p_bstride;
            nlag = a.nlag;
            pinit = a.p_init != 0;
            track = a.p_part != nullptr;
            if (track) ibase = (row % a.p_nbin) * (long long)a.nlag;
        }
    }

    // output k of global sub-transform gg
    __device__ __forceinline__ void store(int gg, int k, float2 v)
    {
        if (thi) {
            // w_M^{gg k}: gg < nsub and k < L with nsub * L = M, so gg k < M
            const int ti = gg * k;
            v = cmul(v, cmul(thi[ti >> 10], tlo[ti & 1023]));
        }
        const int e = gg * ogs + k * ons;
        if (SM == ST_CPX) {
            if (win == 0) {
                pout[e] = v;
            } else {
                // lags -win ... win-1 only, at e + win (mod M)
                const int w = e < win ? e + win : e - (wM - win);
                if (w >= 0 && w < 2 * win) pout[w] = v;
            }
        } else if (e < nlag) {
            // reference build: P + fmaf(re, re, im * im) (src/sdr_func.c:678-680);
            // on the first block P is zero and 0 + x = x exactly
            const float pw = __builtin_fmaf(v.x, v.x, v.y * v.y);
            const float nv = pinit ? pw : prow[e] + pw;
            prow[e] = nv;
            if (track) {
                bs += (double)nv;
                if (nv > bm || (nv == bm && be >= 0 && e < be)) {
                    if (nv > 0.0f) { bm = nv; be = e; }
                }
            }
        }
    }

    // ST_POW at lag e given the previous P value (ignored on the first
    // block); values at e >= nlag are dropped
    __device__ __forceinline__ void store_pow(int e, float2 v, float pold)
    {
        if (e >= nlag) return;
        const float pw = __builtin_fmaf(v.x, v.x, v.y * v.y);
        const float nv = pinit ? pw : pold + pw;
        prow[e] = nv;
        if (track) {
            bs += (double)nv;
            if (nv > bm || (nv == bm && be >= 0 && e < be)) {
                if (nv > 0.0f) { bm = nv; be = e; }
            }
        }
    }

    // the same without branches (bitwise conditions, selects): the store
    // goes through put(ok, e, value), which drops it when !ok, and only ok
    // values enter the peak tracking.  The sum for the mean is kept per lane
    // in FP32 (fsum, a few dozen values) and joins bs at flush_peak.
    float fsum = 0.0f;
    // The power is k2 (re^2 + im^2).
    template <class PutF>
    __device__ __forceinline__ void store_pow_sel(int e, float2 v, float k2, float pold, bool ok, const PutF &put)
    {
        ok = ok & (e < nlag);
        const float pw = k2 * __builtin_fmaf(v.x, v.x, v.y * v.y);
        const float nv = pinit ? pw : pold + pw;
        put(ok, e, nv);
        if (track) {
            fsum += ok ? nv : 0.0f;
            const bool take = ok & (nv > 0.0f) & ((nv > bm) | ((nv == bm) & (be >= 0) & (e < be)));
            bm = take ? nv : bm;
            be = take ? e : be;
        }
    }

    // workgroup reduction of the tracked peak into part[slot]; red holds one
    // record per wave
    __device__ __forceinline__ void flush_peak(PeakPart *part, long long slot, PeakPart *red)
    {
        bs += (double)fsum;
        fsum = 0.0f;
        long long bi = be >= 0 ? ibase + be : -1;
        // within rows of 16 lanes: partners by DPP; then across rows
        peak_dpp_step<0xB1>(bm, bi, bs);
        peak_dpp_step<0x4E>(bm, bi, bs);
        peak_dpp_step<0x141>(bm, bi, bs);
        peak_dpp_step<0x140>(bm, bi, bs);
        peak_swap_step<false>(bm, bi, bs);
        peak_swap_step<true>(bm, bi, bs);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { red[w].m = bm; red[w].i = bi; red[w].s = bs; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
                bs += red[i].s;
                const long long oi = red[i].i;
                const float om = red[i].m;
                if (oi >= 0 && (bi < 0 || om > bm || (om == bm && oi < bi))) { bm = om; bi = oi; }
            }
            PeakPart q;
            q.m = bm; q.pad = 0; q.i = bi; q.s = bs;
            part[slot] = q;
        }
    }
};

// Batch item b and workgroup-group grp of this workgroup.  With xcd_nbin > 0
// (inverse row passes) the order is XCD-aware: workgroups go round-robin to
// the 8 XCDs, so tile T = grp * nbin + bin keeps XCD T % 8 and its PRNs run
// back to back there, re-reading the shared X rows from that XCD's L2.  The
// grid is padded to whole rounds of 8; returns false for the padding.
// KX > 1 (column passes whose G columns span less than a 128-byte line):
// the KX workgroups of adjacent column groups, which read and write the KX
// parts of the same lines, go to one XCD in consecutive rounds, so
// that each line is fetched from HBM into one L2 only (workgroup h of a
// whole block of 8 KX runs logical item (s / KX) 8 KX + (h % 8) KX + s % KX,
// s = h / 8; the last partial block keeps the identity order).
__device__ __forceinline__ bool block_item(const PassArgs &a, bool prod, int ngroups, int *b, int *grp, int kx = 1)
{
    if (kx > 1 && !prod) {
        int h = (int)blockIdx.x;
        if (h < (int)(gridDim.x / (8u * kx)) * 8 * kx) {
            const int s = h >> 3;
            h = (s / kx) * (8 * kx) + (h & 7) * kx + s % kx;
        }
        *b = h / ngroups;
        *grp = h - *b * ngroups;
        return true;
    }
    if (prod && a.xcd_nbin > 0) {
        const int w = (int)blockIdx.x;
        const int r = w >> 3;
        const int prn = r % a.xcd_nprn;
        const int T = (r / a.xcd_nprn) * 8 + (w & 7);
        if (T >= ngroups * a.xcd_nbin) return false;
        *grp = T / a.xcd_nbin;
        const int bin = T - *grp * a.xcd_nbin;
        *b = prn * a.xcd_nbin + bin;
    } else {
        *b = (int)(blockIdx.x / (unsigned)ngroups);
        *grp = (int)blockIdx.x - *b * ngroups;
    }
    return true;
}

}  // namespace psa
