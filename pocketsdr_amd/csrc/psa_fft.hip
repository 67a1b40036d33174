// psa_fft.hip - FP32 mixed-radix FFT for gfx950, built for the PCPS
// correlator of src/sdr_func.c:1088-1102 and sdr_cpx_fft (src/sdr_func.c:236).
//
// Length M = M1 * M2 (all prime factors in {2, 3, 5}) is done as a four-step
// transform: with n = M2 n1 + n2 and k = k1 + M1 k2,
//   X[k1 + M1 k2] = sum_n2 w_M2^{n2 k2} [ w_M^{n2 k1} sum_n1 x[M2 n1 + n2] w_M1^{n1 k1} ]
// i.e. a "column" pass (M2 transforms of length M1, stride M2) that also
// applies the twiddle w_M^{n2 k1}, and a "row" pass (M1 contiguous transforms
// of length M2).  Each pass is one kernel: a workgroup loads G sub-transforms
// into LDS with coalesced accesses, runs Stockham radix-2/3/4/5/8 stages in
// LDS, and stores.  Load and store are templated so the correlator fuses
// carrier mixing, the code-spectrum product and the power accumulation into
// the first and last pass.
#include <math.h>
#include <map>
#include <cstdlib>
#include "psa_pass.h"

namespace psa {

thread_local long long pass_grid = 0;

// ---------------------------------------------------------------------------
// radix-r DFT butterflies, sgn = -1 (forward) / +1 (backward)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bfly2(float2 *a)
{
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

__device__ __forceinline__ void bfly4(float2 *a, float sgn)
{
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = mul_i(csub(a[1], a[3]), sgn);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
}

__device__ __forceinline__ void bfly8(float2 *a, float sgn)
{
    const float R = 0.70710678118654752f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    bfly4(e, sgn);
    bfly4(o, sgn);
    // o[k] *= w8^k, w8 = exp(sgn i pi / 4)
    o[1] = make_float2(R * (o[1].x - sgn * o[1].y), R * (o[1].y + sgn * o[1].x));
    o[2] = mul_i(o[2], sgn);
    o[3] = make_float2(R * (-o[3].x - sgn * o[3].y), R * (-o[3].y + sgn * o[3].x));
    for (int k = 0; k < 4; k++) {
        a[k] = cadd(e[k], o[k]);
        a[k + 4] = csub(e[k], o[k]);
    }
}

__device__ __forceinline__ void bfly3(float2 *a, float sgn)
{
    const float C = -0.5f, S = 0.86602540378443865f;
    float2 t = cadd(a[1], a[2]);
    float2 d = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x + C * t.x, a[0].y + C * t.y);
    float2 r = make_float2(-sgn * S * d.y, sgn * S * d.x);   // sgn i S d
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
}

__device__ __forceinline__ void bfly5(float2 *a, float sgn)
{
    const float C1 = 0.30901699437494742f, C2 = -0.80901699437494742f;
    const float S1 = 0.95105651629515357f, S2 = 0.58778525229247313f;
    float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    float2 m1 = make_float2(a[0].x + C1 * t1.x + C2 * t2.x, a[0].y + C1 * t1.y + C2 * t2.y);
    float2 m2 = make_float2(a[0].x + C2 * t1.x + C1 * t2.x, a[0].y + C2 * t1.y + C1 * t2.y);
    // s1 = sgn i (S1 d1 + S2 d2), s2 = sgn i (S2 d1 - S1 d2)
    float2 u1 = make_float2(S1 * d1.x + S2 * d2.x, S1 * d1.y + S2 * d2.y);
    float2 u2 = make_float2(S2 * d1.x - S1 * d2.x, S2 * d1.y - S1 * d2.y);
    float2 s1 = mul_i(u1, sgn), s2 = mul_i(u2, sgn);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, s1);
    a[4] = csub(m1, s1);
    a[2] = cadd(m2, s2);
    a[3] = csub(m2, s2);
}

// One Stockham stage of radix R over G sub-transforms of length L in LDS.
// Called with compile-time L, G, Ns on the specialised path (divisions become
// multiplies) and with runtime values on the generic path.
template <int R>
__device__ __forceinline__ void stage(const float2 *__restrict__ src, float2 *__restrict__ dst,
                                      int L, int G, int Ns, const float2 *__restrict__ tw,
                                      float sgn)
{
    const int Lr = L / R;
    const int tstep = L / (Ns * R);
    const int nb = G * Lr;
    for (int bi = threadIdx.x; bi < nb; bi += PASS_THREADS) {
        const int g = bi / Lr, j = bi - g * Lr;
        const int k = j % Ns;
        const float2 *s = src + g * L + j;
        float2 a[R];
#pragma unroll
        for (int q = 0; q < R; q++) a[q] = s[q * Lr];
        if (Ns > 1) {
#pragma unroll
            for (int q = 1; q < R; q++) a[q] = cmul(a[q], tw[q * k * tstep]);
        }
        if (R == 2) bfly2(a);
        else if (R == 3) bfly3(a, sgn);
        else if (R == 4) bfly4(a, sgn);
        else if (R == 5) bfly5(a, sgn);
        else if (R == 8) bfly8(a, sgn);
        float2 *d = dst + g * L + (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) d[q * Ns] = a[q];
    }
}

struct RadixSeq {
    int n;
    int r[16];
};

// radix plan of a sub-transform length: 8s first, then 4, 2, 5s, 3s
struct RadPlan {
    int n;
    int r[16];
};
__host__ __device__ constexpr RadPlan make_rad_plan(int L)
{
    RadPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    int n = L;
    while (n % 8 == 0 && n > 1) { p.r[p.n++] = 8; n /= 8; }
    if (n % 4 == 0) { p.r[p.n++] = 4; n /= 4; }
    if (n % 2 == 0) { p.r[p.n++] = 2; n /= 2; }
    while (n % 5 == 0 && n > 1) { p.r[p.n++] = 5; n /= 5; }
    while (n % 3 == 0 && n > 1) { p.r[p.n++] = 3; n /= 3; }
    return p;
}

__device__ float2 *run_stages_rt(float2 *src, float2 *dst, int L, int G, const RadixSeq &rs,
                                 const float2 *tw, float sgn)
{
    int Ns = 1;
    for (int p = 0; p < rs.n; p++) {
        const int r = rs.r[p];
        switch (r) {
        case 2: stage<2>(src, dst, L, G, Ns, tw, sgn); break;
        case 3: stage<3>(src, dst, L, G, Ns, tw, sgn); break;
        case 4: stage<4>(src, dst, L, G, Ns, tw, sgn); break;
        case 5: stage<5>(src, dst, L, G, Ns, tw, sgn); break;
        case 8: stage<8>(src, dst, L, G, Ns, tw, sgn); break;
        }
        __syncthreads();
        float2 *t = src; src = dst; dst = t;
        Ns *= r;
    }
    return src;
}

template <int R>
__device__ __forceinline__ void bfly(float2 *a, float sgn)
{
    if (R == 2) bfly2(a);
    else if (R == 3) bfly3(a, sgn);
    else if (R == 4) bfly4(a, sgn);
    else if (R == 5) bfly5(a, sgn);
    else if (R == 8) bfly8(a, sgn);
}

// ---------------------------------------------------------------------------
// generic pass kernel (runtime length): LDS ping-pong for load, every stage
// and store
// ---------------------------------------------------------------------------
template <int LM, int SM>
__global__ void __launch_bounds__(PASS_THREADS) pass_kernel_rt(PassArgs a, RadixSeq rs, int ngroups)
{
    extern __shared__ float2 smem[];
    __shared__ uint32_t carr[512];      // carrier table, then a raw ring's table
    const int L = a.L, G = a.G;
    float2 *buf0 = smem, *buf1 = smem + G * L;
    const int b = (int)(blockIdx.x / (unsigned)ngroups);
    const int g0 = ((int)blockIdx.x - b * ngroups) * G;
    const int gcount = min(G, a.nsub - g0);
    const float sgn = a.dir == 0 ? -1.0f : 1.0f;
    if (LM == LD_MIX) {
        for (int i = threadIdx.x; i < 256; i += PASS_THREADS) carr[i] = a.carr[i];
        if (a.raw_bps) {
            for (int i = threadIdx.x; i < 256; i += PASS_THREADS)
                carr[256 + i] = reinterpret_cast<const uint32_t *>(a.raw_lut)[i];
        }
        __syncthreads();
    }
    Loader<LM> ld;
    ld.init(a, b, carr);
    const int igs = (int)a.in_gstride, ins = (int)a.in_nstride;
    const bool in_nfast = (ins == 1);
    const int tot = G * L;
    for (int idx = threadIdx.x; idx < tot; idx += PASS_THREADS) {
        int g, n;
        if (in_nfast) { g = idx / L; n = idx - g * L; }
        else { n = idx / G; g = idx - n * G; }
        buf0[g * L + n] = g < gcount ? ld.load((g0 + g) * igs + n * ins) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 *res = run_stages_rt(buf0, buf1, L, G, rs, a.tw, sgn);
    Storer<SM> st;
    st.init(a, b);
    const bool out_nfast = (a.out_nstride == 1);
    for (int idx = threadIdx.x; idx < tot; idx += PASS_THREADS) {
        int g, k;
        if (out_nfast) { g = idx / L; k = idx - g * L; }
        else { k = idx / G; g = idx - k * G; }
        if (g < gcount) st.store(g0 + g, k, res[g * L + k]);
    }
}

// ---------------------------------------------------------------------------
// specialised pass kernel (compile-time length CL, CG sub-transforms per
// workgroup).  The first radix stage runs in registers on elements loaded
// straight from global memory; the middle stages run in place in one LDS
// buffer (rows padded by one element to spread the banks); the last stage
// reads LDS and stores straight to global memory.
// ---------------------------------------------------------------------------
template <int R, int CL, int CG, int NS>
__device__ __forceinline__ void stage_inplace(float2 *buf, const float2 *__restrict__ tw, float sgn)
{
    constexpr int LP = CL + 1;
    constexpr int Lr = CL / R;
    constexpr int tstep = CL / (NS * R);
    constexpr int NB = CG * Lr;
    constexpr int PER = (NB + PASS_THREADS - 1) / PASS_THREADS;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int bi = threadIdx.x + p * PASS_THREADS;
        if (bi < NB) {
            const int g = bi / Lr, j = bi - g * Lr;
            const int k = j % NS;
#pragma unroll
            for (int q = 0; q < R; q++) v[p][q] = buf[g * LP + j + q * Lr];
#pragma unroll
            for (int q = 1; q < R; q++) v[p][q] = cmul(v[p][q], tw[q * k * tstep]);
            bfly<R>(v[p], sgn);
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int bi = threadIdx.x + p * PASS_THREADS;
        if (bi < NB) {
            const int g = bi / Lr, j = bi - g * Lr;
            const int k = j % NS;
            float2 *d = buf + g * LP + (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) d[q * NS] = v[p][q];
        }
    }
    __syncthreads();
}

template <int CL, int CG, int IDX, int NS>
__device__ __forceinline__ void middle_stages(float2 *buf, const float2 *tw, float sgn)
{
    constexpr RadPlan P = make_rad_plan(CL);
    if constexpr (IDX < P.n - 1) {
        constexpr int R = P.r[IDX];
        stage_inplace<R, CL, CG, NS>(buf, tw, sgn);
        middle_stages<CL, CG, IDX + 1, NS * R>(buf, tw, sgn);
    }
}

// last stage (Ns = L / RL, k = j): LDS -> registers -> global
template <int SM, int CL, int CG>
__device__ __forceinline__ void last_stage(const PassArgs &a, Storer<SM> &st, int g0, int gcount,
                                           const float2 *buf, float sgn, const float2 *tw)
{
    constexpr RadPlan P = make_rad_plan(CL);
    constexpr int LP = CL + 1;
    constexpr int RL = P.r[P.n - 1];
    constexpr int LrL = CL / RL;
    const bool out_nfast = (a.out_nstride == 1);
    constexpr int NB = CG * LrL;
    constexpr int PER = (NB + PASS_THREADS - 1) / PASS_THREADS;
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int bi = threadIdx.x + p * PASS_THREADS;
        if (bi < NB) {
            int g, j;
            if (out_nfast) { g = bi / LrL; j = bi - g * LrL; }
            else { j = bi / CG; g = bi - j * CG; }
            if (SM == ST_POW) {
                // previous P values requested before the transform (clamped
                // indices, 0 on the first block), stored after it
                const bool okg = g < gcount;
                const int eb = (g0 + (okg ? g : 0)) * st.ogs + j * st.ons;
                float pold[RL];
#pragma unroll
                for (int q = 0; q < RL; q++) {
                    const int e = eb + q * LrL * st.ons;
                    pold[q] = st.pinit ? 0.0f : st.prow[e < st.nlag ? e : 0];
                }
                float2 v[RL];
#pragma unroll
                for (int q = 0; q < RL; q++) v[q] = buf[g * LP + j + q * LrL];
#pragma unroll
                for (int q = 1; q < RL; q++) v[q] = cmul(v[q], tw[q * j]);
                bfly<RL>(v, sgn);
                if (okg) {
#pragma unroll
                    for (int q = 0; q < RL; q++) {
                        if (st.thi) {     // inter-pass twiddle (not used by the power passes)
                            const int ti = (g0 + g) * (j + q * LrL);
                            v[q] = cmul(v[q], cmul(st.thi[ti >> 10], st.tlo[ti & 1023]));
                        }
                        st.store_pow(eb + q * LrL * st.ons, v[q], pold[q]);
                    }
                }
            } else if (g < gcount) {
                float2 v[RL];
#pragma unroll
                for (int q = 0; q < RL; q++) v[q] = buf[g * LP + j + q * LrL];
#pragma unroll
                for (int q = 1; q < RL; q++) v[q] = cmul(v[q], tw[q * j]);
                bfly<RL>(v, sgn);
#pragma unroll
                for (int q = 0; q < RL; q++) st.store(g0 + g, j + q * LrL, v[q]);
            }
        }
    }
}

// first stage (Ns = 1, no twiddles): global -> registers -> LDS
template <int LM, int CL, int CG>
__device__ __forceinline__ void first_stage(const PassArgs &a, const Loader<LM> &ld, int g0, int gcount,
                                            float2 *buf, float sgn)
{
    constexpr RadPlan P = make_rad_plan(CL);
    constexpr int LP = CL + 1;
    constexpr int R0 = P.r[0];
    constexpr int Lr0 = CL / R0;
    const int igs = (int)a.in_gstride, ins = (int)a.in_nstride;
    const bool in_nfast = (ins == 1);
    {
        constexpr int NB = CG * Lr0;
        constexpr int PER = (NB + PASS_THREADS - 1) / PASS_THREADS;
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int bi = threadIdx.x + p * PASS_THREADS;
            if (bi < NB) {
                int g, j;
                if (in_nfast) { g = bi / Lr0; j = bi - g * Lr0; }
                else { j = bi / CG; g = bi - j * CG; }
                float2 v[R0];
                // rows past the end of the item read row 0 and are zeroed, so
                // the loads are not serialised behind a branch
                const bool okg = g < gcount;
                const int e0 = (g0 + (okg ? g : 0)) * igs + j * ins;
#pragma unroll
                for (int q = 0; q < R0; q++) {
                    const float2 x = ld.load(e0 + q * Lr0 * ins);
                    v[q] = okg ? x : make_float2(0.f, 0.f);
                }
                bfly<R0>(v, sgn);
#pragma unroll
                for (int q = 0; q < R0; q++) buf[g * LP + j * R0 + q] = v[q];
            }
        }
        __syncthreads();
    }
}

// one item of the specialised pass: first stage from global, middle stages
// in LDS, last stage to global
// (tw: the length-CL twiddle table, staged in LDS by the caller before the
// first stage's closing barrier)
template <int LM, int SM, int CL, int CG>
__device__ __forceinline__ void pass_item(const PassArgs &a, int b, int g0, int gcount, float2 *buf,
                                          const uint32_t *carr, float sgn, PeakPart *red, long long pslot,
                                          const float2 *tw)
{
    constexpr int R0 = make_rad_plan(CL).r[0];
    Loader<LM> ld;
    ld.init(a, b, carr);
    first_stage<LM, CL, CG>(a, ld, g0, gcount, buf, sgn);
    // ---- middle stages in place ----
    middle_stages<CL, CG, 1, R0>(buf, tw, sgn);
    // ---- last stage ----
    Storer<SM> st;
    st.init(a, b);
    last_stage<SM, CL, CG>(a, st, g0, gcount, buf, sgn, tw);
    if (SM == ST_POW && a.p_part) st.flush_peak(a.p_part, pslot, red);
}

template <int LM, int SM, int CL, int CG>
__global__ void __launch_bounds__(PASS_THREADS) pass_kernel_ct(PassArgs a, int ngroups)
{
    static_assert(make_rad_plan(CL).n >= 2, "specialised lengths need at least two radix stages");
    __shared__ float2 buf[CG * (CL + 1)];
    __shared__ float2 twl[CL];          // the stages' twiddles W_CL^i
    __shared__ uint32_t carr[512];      // carrier table, then a raw ring's table
    __shared__ PeakPart red[PASS_THREADS / 64];
    int b, grp;
    // column passes: CG float2 (inverse) or CG IF bytes (mixing) per row
    // segment, kx workgroups to one 128-byte line (xcd_pair = 0: identity)
    const int segb = LM == LD_MIX ? CG : 8 * CG;
    const int kx = a.xcd_pair && a.in_nstride != 1 && segb < 128 ? 128 / segb : 1;
    if (!block_item(a, LM == LD_PROD, ngroups, &b, &grp, kx)) return;
    const int g0 = grp * CG;
    const int gcount = min(CG, a.nsub - g0);
    const float sgn = a.dir == 0 ? -1.0f : 1.0f;
    if (LM == LD_MIX) {
        for (int i = threadIdx.x; i < 256; i += PASS_THREADS) carr[i] = a.carr[i];
        if (a.raw_bps) {
            for (int i = threadIdx.x; i < 256; i += PASS_THREADS)
                carr[256 + i] = reinterpret_cast<const uint32_t *>(a.raw_lut)[i];
        }
        __syncthreads();
    }
    long long pslot = 0;
    if (SM == ST_POW && a.p_part) {
        const long long row = (long long)(b / a.p_div + a.p_off1) * a.p_mul + a.p_off2 + b % a.p_div;
        pslot = row * ngroups + grp;
    }
    for (int i = threadIdx.x; i < CL; i += PASS_THREADS) twl[i] = a.tw[i];
    pass_item<LM, SM, CL, CG>(a, b, g0, gcount, buf, carr, sgn, red, pslot, twl);
}

// Fused correlator row pass for one code spectrum per item (the FFT
// tracking of L6D / L6E): per workgroup CG rows of the forward column
// pass's output go through the forward length-M2 transforms, the product
// with the item's code spectrum (sdr_cpx_mul arithmetic and scale, spectrum
// cf_index[b], or b / prod_cdiv) and the inverse length-M2 transforms with
// the inverse inter-pass twiddle on store, all in one LDS buffer: X and the
// product never go through HBM.  The output is the inverse row pass's Z.
template <int CL, int CG>
__global__ void __launch_bounds__(PASS_THREADS) xcorr_row_kernel(PassArgs a, int ngroups)
{
    static_assert(make_rad_plan(CL).n >= 2, "specialised lengths need at least two radix stages");
    constexpr RadPlan P = make_rad_plan(CL);
    constexpr int LP = CL + 1;
    constexpr int RL = P.r[P.n - 1];
    __shared__ float2 buf[CG * LP];
    __shared__ float2 twl[CL];          // forward, then inverse twiddles W_CL^(-/+i)
    const int b = (int)(blockIdx.x / (unsigned)ngroups);
    const int grp = (int)blockIdx.x - b * ngroups;
    const int g0 = grp * CG;
    const int gcount = min(CG, a.nsub - g0);
    for (int i = threadIdx.x; i < CL; i += PASS_THREADS) twl[i] = a.tw[i];
    // forward transforms, natural order in LDS
    Loader<LD_CPX> ld;
    ld.init(a, b, nullptr);
    first_stage<LD_CPX, CL, CG>(a, ld, g0, gcount, buf, -1.0f);
    middle_stages<CL, CG, 1, P.r[0]>(buf, twl, -1.0f);
    stage_inplace<RL, CL, CG, CL / RL>(buf, twl, -1.0f);
    // (the last forward stage ends with a barrier: the inverse table
    // replaces the forward one during the product, whose closing barrier
    // precedes the inverse stages)
    for (int i = threadIdx.x; i < CL; i += PASS_THREADS) twl[i] = a.tw_inv[i];
    // X * Cf at transposed index (g0 + g) M2 + k
    {
        const long long cb = a.cf_index ? a.cf_index[b] : b / a.prod_cdiv;
        const float2 *__restrict__ cin = a.cf + cb * a.cf_bstride;
        const bool tailm = a.prod_V < a.M;
        for (int idx = threadIdx.x; idx < CG * CL; idx += PASS_THREADS) {
            const int g = idx / CL, k = idx - g * CL;
            if (g < gcount) {
                const int e = (g0 + g) * CL + k;
                bool tail = false;
                if (tailm) tail = (e / a.prod_M2 + a.prod_M1 * (e % a.prod_M2)) >= a.prod_V;
                buf[g * LP + k] = prod_value(buf[g * LP + k], cin[e], tail, a.prod_scale);
            }
        }
        __syncthreads();
    }
    // inverse transforms: stages 0 .. n-2 in LDS, the last one to global
    middle_stages<CL, CG, 0, 1>(buf, twl, 1.0f);
    Storer<ST_CPX> st;
    st.init(a, b);
    last_stage<ST_CPX, CL, CG>(a, st, g0, gcount, buf, 1.0f, twl);
}

// ---------------------------------------------------------------------------
// host side: plans and launches
// ---------------------------------------------------------------------------
static bool factor235(int n)
{
    if (n < 1) return false;
    for (int p : {2, 3, 5}) while (n % p == 0) n /= p;
    return n == 1;
}

bool fft_supported(int M) { return M >= 1 && factor235(M); }

static void radix_seq(int L, SubPlan &sp)
{
    sp.L = L;
    sp.nrad = 0;
    int n = L;
    while (n % 8 == 0) { sp.rad[sp.nrad++] = 8; n /= 8; }
    if (n % 4 == 0) { sp.rad[sp.nrad++] = 4; n /= 4; }
    if (n % 2 == 0) { sp.rad[sp.nrad++] = 2; n /= 2; }
    while (n % 5 == 0) { sp.rad[sp.nrad++] = 5; n /= 5; }
    while (n % 3 == 0) { sp.rad[sp.nrad++] = 3; n /= 3; }
}

// scale * exp(sign 2 pi i m / L), m < L, rounded from double
static void fill_twiddle(int L, int sign, double scale, float2 *h)
{
    const double PI_D = 3.14159265358979323846;
    for (int m = 0; m < L; m++) {
        double ang = sign * 2.0 * PI_D * (double)m / L;
        h[m] = make_float2((float)(scale * cos(ang)), (float)(scale * sin(ang)));
    }
}

float2 *make_twiddle(int L, int sign, double scale)
{
    std::vector<float2> h(L > 0 ? L : 1);
    fill_twiddle(L, sign, scale, h.data());
    float2 *d;
    PSA_CHECK(hipMalloc(&d, sizeof(float2) * h.size()));
    PSA_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
    return d;
}

// w_M^{a} for a = hi * 1024 + lo: T_hi[hi] = exp(s 2 pi i hi 1024 / M),
// T_lo[lo] = exp(s 2 pi i lo / M), both rounded from double; hi has
// (M + 1023) / 1024 entries, lo 1024
static void fill_split_twiddle(int M, int sign, float2 *hi, float2 *lo)
{
    const double PI_D = 3.14159265358979323846;
    int nhi = (M + 1023) / 1024;
    for (int i = 0; i < nhi; i++) {
        double ang = sign * 2.0 * PI_D * (double)((long long)i * 1024 % M) / M;
        hi[i] = make_float2((float)cos(ang), (float)sin(ang));
    }
    for (int i = 0; i < 1024; i++) {
        double ang = sign * 2.0 * PI_D * (double)i / M;
        lo[i] = make_float2((float)cos(ang), (float)sin(ang));
    }
}

static void make_split_twiddle(int M, int sign, float2 **thi, float2 **tlo)
{
    int nhi = (M + 1023) / 1024;
    std::vector<float2> hi(nhi), lo(1024);
    fill_split_twiddle(M, sign, hi.data(), lo.data());
    PSA_CHECK(hipMalloc(thi, sizeof(float2) * nhi));
    PSA_CHECK(hipMalloc(tlo, sizeof(float2) * 1024));
    PSA_CHECK(hipMemcpy(*thi, hi.data(), sizeof(float2) * nhi, hipMemcpyHostToDevice));
    PSA_CHECK(hipMemcpy(*tlo, lo.data(), sizeof(float2) * 1024, hipMemcpyHostToDevice));
}

// Split M into M1 * M2 (column length M1, row length M2), both <= 4096.
// Rows of a multiple of 16 complex (128 bytes) keep the 16-wide column
// segments of the column passes on cache-line boundaries; among those
// splits the most balanced one wins (any split when none has such rows).
static void split_len(int M, int *M1, int *M2)
{
    int best1 = M, best2 = 1;
    if (M > 4096) {
        best1 = 0;
        long long bestscore = 0;
        for (int d = 1; d <= 4096 && d <= M; d++) {
            if (M % d) continue;
            const int e = M / d;          // d = M1, e = M2
            if (e > 4096) continue;
            const long long score = (e % 16 == 0 ? 0 : (1LL << 40)) + (long long)std::abs(d - e);
            if (!best1 || score < bestscore) { best1 = d; best2 = e; bestscore = score; }
        }
    }
    *M1 = best1;
    *M2 = best2;
}

static std::mutex g_plan_mtx;
static std::map<std::pair<int, int>, FFTPlan *> g_plans;

const FFTPlan *fft_plan(int M, int M1)
{
    if (!fft_supported(M)) return nullptr;
    ensure_init();
    fft_init_attrs();
    int M2;
    if (M1 == 0) split_len(M, &M1, &M2);
    if (M1 <= 0 || M % M1 || M1 > 4096 || M / M1 > 4096) return nullptr;
    M2 = M / M1;
    std::lock_guard<std::mutex> lk(g_plan_mtx);
    auto it = g_plans.find({M, M1});
    if (it != g_plans.end()) return it->second;
    FFTPlan *p = new FFTPlan();
    p->M = M; p->M1 = M1; p->M2 = M2;
    radix_seq(M1, p->s1);
    radix_seq(M2, p->s2);
    p->s1.tw_fwd = make_twiddle(M1, -1, 1.0);
    p->s1.tw_bwd = make_twiddle(M1, +1, 1.0);
    p->s2.tw_fwd = make_twiddle(M2, -1, 1.0);
    p->s2.tw_bwd = make_twiddle(M2, +1, 1.0);
    make_split_twiddle(M, -1, &p->thi_fwd, &p->tlo_fwd);
    make_split_twiddle(M, +1, &p->thi_bwd, &p->tlo_bwd);
    g_plans[{M, M1}] = p;
    return p;
}

size_t fft_plan_tables(int M, FFTPlan *p, float2 *h, size_t cap, float2 *dev)
{
    if (!fft_supported(M)) return 0;
    int M1, M2;
    split_len(M, &M1, &M2);
    if (M1 <= 0 || M1 > 4096 || M2 > 4096) return 0;
    const int nhi = (M + 1023) / 1024;
    const size_t n = 2 * (size_t)M1 + 2 * (size_t)M2 + 2 * ((size_t)nhi + 1024);
    if (n > cap) return 0;
    *p = FFTPlan();
    p->M = M; p->M1 = M1; p->M2 = M2;
    radix_seq(M1, p->s1);
    radix_seq(M2, p->s2);
    size_t o = 0;
    auto take = [&](int len) { float2 *hp = h + o; float2 *dp = dev + o; o += len; return std::make_pair(hp, dp); };
    auto t = take(M1); fill_twiddle(M1, -1, 1.0, t.first); p->s1.tw_fwd = t.second;
    t = take(M1); fill_twiddle(M1, +1, 1.0, t.first); p->s1.tw_bwd = t.second;
    t = take(M2); fill_twiddle(M2, -1, 1.0, t.first); p->s2.tw_fwd = t.second;
    t = take(M2); fill_twiddle(M2, +1, 1.0, t.first); p->s2.tw_bwd = t.second;
    auto hi = take(nhi), lo = take(1024);
    fill_split_twiddle(M, -1, hi.first, lo.first);
    p->thi_fwd = hi.second; p->tlo_fwd = lo.second;
    hi = take(nhi); lo = take(1024);
    fill_split_twiddle(M, +1, hi.first, lo.first);
    p->thi_bwd = hi.second; p->tlo_bwd = lo.second;
    return o;
}

// (L, G) pairs with compile-time kernels: the sub-lengths of the plans for
// M = 24000 (150 x 160), 48000 (200 x 240), 96000 (300 x 320),
// 192000 (400 x 480) and 960000 (1000 x 960).  Column lengths take G = 16
// (one 128-byte segment per row; 8 for 1000 to fit LDS), row lengths keep
// about 4096 complex per workgroup.
#define PSA_SPECS(X) \
    X(150, 8) X(160, 8) X(200, 8) X(240, 8) X(300, 16) X(320, 12) \
    X(400, 16) X(480, 4) X(1000, 8) X(960, 2)

bool spec_length(int L);

static int spec_groups(int L)
{
#define PSA_G(L_, G_) if (L == L_) return G_;
    PSA_SPECS(PSA_G)
#undef PSA_G
    return 0;
}

template <int LM, int SM, int CL, int CG>
static void launch_t(const PassArgs &a, const RadixSeq &rs, int nbatch, hipStream_t st)
{
    const int ngroups = (a.nsub + a.G - 1) / a.G;
    const long long nblk = (long long)ngroups * nbatch;
    if (nblk <= 0) return;
    if (nblk > 0x7FFFFFFFLL) PSA_FAIL("FFT launch too large (%lld workgroups)", nblk);
    if ((long long)nbatch * a.M > 0x7FFFFFFFLL && (a.in_bstride == 0 || a.out_bstride == 0)) {
        PSA_FAIL("FFT launch too large");
    }
    if constexpr (CL > 0) {
        long long nb = nblk;
        if (LM == LD_PROD && a.xcd_nbin > 0) {
            // tiles padded to a multiple of 8 (one per XCD)
            const long long ntile = (long long)ngroups * a.xcd_nbin;
            nb = (ntile + 7) / 8 * 8 * a.xcd_nprn;
        }
        pass_grid = nb * PASS_THREADS;
        hipLaunchKernelGGL((pass_kernel_ct<LM, SM, CL, CG>), dim3((unsigned)nb), dim3(PASS_THREADS), 0, st,
                           a, ngroups);
    } else {
        const size_t shm = sizeof(float2) * 2 * (size_t)a.G * a.L;
        pass_grid = nblk * PASS_THREADS;
        hipLaunchKernelGGL((pass_kernel_rt<LM, SM>), dim3((unsigned)nblk), dim3(PASS_THREADS), shm, st,
                           a, rs, ngroups);
    }
    PSA_CHECK(hipGetLastError());
}

template <int LM, int SM>
static void set_attr()
{
    PSA_CHECK(hipFuncSetAttribute((const void *)pass_kernel_rt<LM, SM>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
}

// the generic kernels take up to 64 KiB of dynamic LDS (two 4096-point buffers)
void fft_init_attrs()
{
    static std::once_flag once;
    std::call_once(once, [] {
        set_attr<LD_CPX, ST_CPX>();
        set_attr<LD_CPX16, ST_CPX>();
        set_attr<LD_MIX, ST_CPX>();
        set_attr<LD_PROD, ST_CPX>();
        set_attr<LD_CPX, ST_POW>();
        set_attr<LD_PROD, ST_POW>();
    });
}

template <int CL, int CG>
static void launch_modes(LoadMode lm, StoreMode sm, const PassArgs &a, const RadixSeq &rs,
                         int nbatch, hipStream_t st)
{
#define PSA_LAUNCH(L_, S_) if (lm == L_ && sm == S_) { launch_t<L_, S_, CL, CG>(a, rs, nbatch, st); return; }
    PSA_LAUNCH(LD_CPX, ST_CPX)
    PSA_LAUNCH(LD_CPX16, ST_CPX)
    PSA_LAUNCH(LD_MIX, ST_CPX)
    PSA_LAUNCH(LD_PROD, ST_CPX)
    PSA_LAUNCH(LD_CPX, ST_POW)
    PSA_LAUNCH(LD_PROD, ST_POW)
#undef PSA_LAUNCH
    PSA_FAIL("unsupported FFT pass mode %d/%d", (int)lm, (int)sm);
}

template <int CL, int CG>
static void launch_xrow_t(const PassArgs &a, int nbatch, hipStream_t st)
{
    const int ngroups = (a.nsub + CG - 1) / CG;
    const long long nblk = (long long)ngroups * nbatch;
    if (nblk <= 0) return;
    if (nblk > 0x7FFFFFFFLL) PSA_FAIL("FFT launch too large (%lld workgroups)", nblk);
    pass_grid = nblk * PASS_THREADS;
    hipLaunchKernelGGL((xcorr_row_kernel<CL, CG>), dim3((unsigned)nblk), dim3(PASS_THREADS), 0, st, a, ngroups);
    PSA_CHECK(hipGetLastError());
}

bool launch_xcorr_row(const PassArgs &a_in, int nbatch, hipStream_t st)
{
    PassArgs a = a_in;
    if ((long long)a.M * 2 > 0x7FFFFFFFLL) PSA_FAIL("FFT length too large");
    if (a.in_nstride != 1 || a.out_nstride != 1 || a.in_gstride != a.L || a.out_gstride != a.L) {
        PSA_FAIL("fused correlator row pass: rows must be contiguous");
    }
    // small batches (the closed loop's one epoch of a few channels): a
    // half or a quarter of the rows per workgroup, so that the launch still
    // spreads over the CUs
#define PSA_X(L_, G_)                                                                          \
    if (a.L == L_) {                                                                           \
        const long long nb = (long long)((a.nsub + G_ - 1) / G_) * nbatch;                     \
        if (G_ >= 4 && nb < 1024) {                                                            \
            a.G = G_ >= 4 ? G_ / 4 : 1;                                                        \
            launch_xrow_t<L_, (G_ >= 4 ? G_ / 4 : 1)>(a, nbatch, st);                          \
        } else if (G_ >= 2 && nb < 2048) {                                                     \
            a.G = (G_ + 1) / 2;                                                                \
            launch_xrow_t<L_, (G_ + 1) / 2>(a, nbatch, st);                                    \
        } else {                                                                               \
            a.G = G_;                                                                          \
            launch_xrow_t<L_, G_>(a, nbatch, st);                                              \
        }                                                                                      \
        return true;                                                                           \
    }
    PSA_SPECS(PSA_X)
#undef PSA_X
    return false;
}

void launch_pass(LoadMode lm, StoreMode sm, const PassArgs &a_in, int nbatch, hipStream_t st)
{
    PassArgs a = a_in;
    if (!a.carr) a.carr = dev_carrier();
    // XCD pairing of narrow column groups (PSA_XCD_PAIR=0 turns it off)
    static const int pair = [] {
        const char *e = getenv("PSA_XCD_PAIR");
        return e ? atoi(e) : 1;
    }();
    a.xcd_pair = pair;
    RadixSeq rs;
    const RadPlan rp = make_rad_plan(a.L);
    rs.n = rp.n;
    for (int i = 0; i < 16; i++) rs.r[i] = rp.r[i];
    if ((long long)a.M * 2 > 0x7FFFFFFFLL) PSA_FAIL("FFT length too large");
    // small batches (the closed loop's one epoch of a few channels): half
    // the sub-transforms per workgroup, so that the launch still spreads
    // over the CUs; not for power passes, whose peak partials are laid out
    // by the plan's group count
#define PSA_L(L_, G_)                                                                          \
    if (a.L == L_ && a.G == G_) {                                                              \
        if (G_ >= 2 && sm != ST_POW && (long long)((a.nsub + G_ - 1) / G_) * nbatch < 2048) {  \
            a.G = (G_ + 1) / 2;                                                                \
            launch_modes<L_, (G_ + 1) / 2>(lm, sm, a, rs, nbatch, st);                         \
        } else {                                                                               \
            launch_modes<L_, G_>(lm, sm, a, rs, nbatch, st);                                   \
        }                                                                                      \
        return;                                                                                \
    }
    PSA_SPECS(PSA_L)
#undef PSA_L
    launch_modes<0, 0>(lm, sm, a, rs, nbatch, st);
}

bool spec_length(int L)
{
    return spec_groups(L) > 0;
}

extern "C" int sdr_fft_pass_groups(int L)
{
    return spec_groups(L);
}

int fft_splits(int M, int *m1, int nmax)
{
    int d1, d2, n = 0;
    split_len(M, &d1, &d2);
    if (d1 > 0 && n < nmax) m1[n++] = d1;
    if (M <= 4096) return n;
    for (int L = 1; L <= 4096 && n < nmax; L++) {
        if (M % L || L == d1 || !spec_length(L) || !spec_length(M / L)) continue;
        m1[n++] = L;
    }
    return n;
}

// G sub-transforms per workgroup: the specialised count where one exists,
// else as many as fit 4096 complex per LDS buffer
int pass_ngroups(const PassArgs &a)
{
    return (a.nsub + a.G - 1) / a.G;
}

int pass_groups(int L, int nsub)
{
    int G = spec_groups(L);
    if (G > 0) return G;
    G = 4096 / (L > 0 ? L : 1);
    if (G < 1) G = 1;
    if (G > nsub) G = nsub;
    return G;
}

// column pass geometry (M2 transforms of length M1, stride M2), with the
// inter-pass twiddle applied on store; output [k1][n2] row-major
void col_pass_args(const FFTPlan *pl, int dir, PassArgs &a)
{
    a.L = pl->M1;
    a.nsub = pl->M2;
    a.G = pass_groups(a.L, a.nsub);
    a.in_nstride = pl->M2; a.in_gstride = 1;
    a.out_nstride = pl->M2; a.out_gstride = 1;
    a.dir = dir;
    a.tw = dir == 0 ? pl->s1.tw_fwd : pl->s1.tw_bwd;
    a.M = pl->M;
    if (pl->M2 > 1) {
        a.thi = dir == 0 ? pl->thi_fwd : pl->thi_bwd;
        a.tlo = dir == 0 ? pl->tlo_fwd : pl->tlo_bwd;
    } else {
        a.thi = a.tlo = nullptr;
    }
}

// row pass geometry (M1 contiguous transforms of length M2)
void row_pass_args(const FFTPlan *pl, int dir, PassArgs &a)
{
    a.L = pl->M2;
    a.nsub = pl->M1;
    a.G = pass_groups(a.L, a.nsub);
    a.in_nstride = 1; a.in_gstride = pl->M2;
    a.out_nstride = 1; a.out_gstride = pl->M2;
    a.dir = dir;
    a.tw = dir == 0 ? pl->s2.tw_fwd : pl->s2.tw_bwd;
    a.M = pl->M;
    a.thi = a.tlo = nullptr;
}

void fft_c2c(const FFTPlan *pl, const float2 *in, float2 *out, float2 *tmp,
             int nbatch, int dir, hipStream_t st)
{
    PassArgs a = {};
    col_pass_args(pl, dir, a);
    a.in = in; a.out = tmp;
    a.in_bstride = a.out_bstride = pl->M;
    launch_pass(LD_CPX, ST_CPX, a, nbatch, st);
    PassArgs r = {};
    row_pass_args(pl, dir, r);
    r.in = tmp; r.out = out;
    r.in_bstride = r.out_bstride = pl->M;
    // natural order: element k2 of row k1 goes to k1 + M1 k2
    r.out_nstride = pl->M1;
    r.out_gstride = 1;
    launch_pass(LD_CPX, ST_CPX, r, nbatch, st);
}

__global__ void transpose_spec_kernel(const float2 *__restrict__ in, float2 *__restrict__ out,
                                      int M, int M1, int M2, long long total)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    long long b = t / M;
    int i = (int)(t - b * M);          // transposed index k1 * M2 + k2
    int k1 = i / M2, k2 = i - k1 * M2;
    out[t] = in[b * M + k1 + (long long)M1 * k2];
}

void fft_to_transposed(const FFTPlan *pl, const float2 *in, float2 *out,
                       int nbatch, hipStream_t st)
{
    long long total = (long long)pl->M * nbatch;
    if (total <= 0) return;
    hipLaunchKernelGGL(transpose_spec_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       in, out, pl->M, pl->M1, pl->M2, total);
    PSA_CHECK(hipGetLastError());
}

}  // namespace psa
