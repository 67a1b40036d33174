// psa_fft16.hip - precision form of the correlator passes for gfx950: complex
// values stored as one 32-bit word of two FP16 or BF16 halves, the radix
// stages of every pass done as dense DFT-matrix products on the matrix cores
// (v_mfma_f32_32x32x16_f16 / _bf16, FP32 accumulation), twiddles and all
// other arithmetic in FP32.
//
// The transform is the four-step split of psa_fft.hip (M = M1 M2, column pass
// of length M1 with the inter-pass twiddle, row pass of length M2, spectra in
// digit-transposed order between the forward and inverse passes).  Inside a
// pass of length L = R0 R1 (R2) every stage is a Stockham radix-R step: the
// G L / R sub-transform columns of the workgroup are processed 32 at a time
// as B (16 x 32 halves per K step, (re, im) pairs of the inputs), the real
// 2R x 2R form of the R-point DFT is A, and the 32 x 32 FP32 result holds the
// R outputs of each column.  The next stage's twiddle is applied to the FP32
// accumulators before they are packed into LDS.
#include <math.h>
#include <map>
#include <mutex>
#include <vector>
#include "psa_pass.h"

namespace psa {

typedef _Float16 hx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bx8 __attribute__((ext_vector_type(8)));
typedef float fx16 __attribute__((ext_vector_type(16)));

#define MP_THREADS 256
#define MP_WAVES (MP_THREADS / 64)

// complex value -> one word (re in the low half), round to nearest even
template <int DT>
__device__ __forceinline__ uint32_t pack_c(float re, float im)
{
    if constexpr (DT == HDT_F16) {
        const _Float16 x = (_Float16)re, y = (_Float16)im;
        return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
    } else {
        const __bf16 x = (__bf16)re, y = (__bf16)im;
        return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
    }
}

template <int DT>
__device__ __forceinline__ float2 unpack_c(uint32_t w)
{
    if constexpr (DT == HDT_F16) {
        return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xFFFFu)),
                           (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
    } else {
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xFFFF0000u));
    }
}

template <int DT>
__device__ __forceinline__ fx16 mfma(uint4 a, uint4 b, fx16 c)
{
    if constexpr (DT == HDT_F16) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(hx8, a), __builtin_bit_cast(hx8, b),
                                                      c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bx8, a), __builtin_bit_cast(bx8, b),
                                                       c, 0, 0, 0);
    }
}

// LDS layout of a workgroup's G sub-transforms: row g at g * RP, element n
// at phi(n).  When R0 is even one pad word follows every R0 words, so the
// first stage's stride-R0 stores spread over the banks; the row pitch RP is
// odd so that lanes walking across rows do too.
template <int L, int R0>
struct MGeom {
    static constexpr int PAD = (R0 % 2 == 0) ? R0 : 0;
    static constexpr int RP = (L + (PAD ? L / PAD : 0)) | 1;
    __device__ static __forceinline__ int phi(int n) { return PAD ? n + n / PAD : n; }
};

// One Stockham radix-R stage (NS = product of the earlier radices) over the
// G sub-transforms of the workgroup.  Column c = (g, j), j < L / R, holds the
// inputs j + t L / R (t < R) and produces outputs q = (j / NS) NS R + j % NS
// + t' NS.  NF: columns run along j (rows contiguous in memory) rather than
// along g.  loadw(g, n) returns input word n of sub-transform g;
// out(g, q, v) takes output q as FP32.
template <int DT, int L, int R, int NS, int G, bool NF, class LoadW, class Out>
__device__ __forceinline__ void mstage(const uint4 *__restrict__ amat, const LoadW &loadw, const Out &out)
{
    constexpr int LR = L / R;
    constexpr int C = G * LR;
    constexpr int NT = (C + 31) / 32;
    constexpr int KS = (2 * R + 15) / 16;       // K steps of 16 halves
    constexpr int RT = (2 * R + 31) / 32;       // 32-row output tiles
    const int lane = threadIdx.x & 63, h = lane >> 5, wave = threadIdx.x >> 6;
    // BF16: A = A_hi + A_lo (the second table follows the first), so the DFT
    // coefficients carry 16 significant bits
    constexpr int NA = DT == HDT_BF16 ? 2 : 1;
    uint4 A[NA][RT][KS];
#pragma unroll
    for (int x = 0; x < NA; x++) {
#pragma unroll
        for (int r = 0; r < RT; r++) {
#pragma unroll
            for (int k = 0; k < KS; k++) A[x][r][k] = amat[((x * RT + r) * KS + k) * 64 + lane];
        }
    }
    for (int tile = wave; tile < NT; tile += MP_WAVES) {
        const int c = tile * 32 + (lane & 31);
        const bool cv = (C % 32 == 0) || c < C;
        int g, j;
        if (NF) { g = c / LR; j = c - g * LR; }
        else { j = c / G; g = c - j * G; }
        // B: lane (column, h) supplies inputs t = 8 k + 4 h + u, u < 4
        uint4 B[KS];
#pragma unroll
        for (int k = 0; k < KS; k++) {
            uint32_t w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = 8 * k + 4 * h + u;
                w[u] = (cv && t < R) ? loadw(g, j + t * LR) : 0u;
            }
            B[k] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        fx16 acc[RT];
#pragma unroll
        for (int r = 0; r < RT; r++) {
#pragma unroll
            for (int e = 0; e < 16; e++) acc[r][e] = 0.0f;
#pragma unroll
            for (int x = 0; x < NA; x++) {
#pragma unroll
                for (int k = 0; k < KS; k++) acc[r] = mfma<DT>(A[x][r][k], B[k], acc[r]);
            }
        }
        if (cv) {
            const int qb = (j / NS) * NS * R + (j % NS);
            // accumulator 4 m + 2 w (+1) of row tile r: output t' = 16 r +
            // 4 m + 2 h + w, real (imaginary) part
#pragma unroll
            for (int r = 0; r < RT; r++) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
#pragma unroll
                    for (int w = 0; w < 2; w++) {
                        const int t = 16 * r + 4 * m + 2 * h + w;
                        if (t < R) out(g, qb + t * NS, make_float2(acc[r][4 * m + 2 * w], acc[r][4 * m + 2 * w + 1]));
                    }
                }
            }
        }
    }
}

// output q of a stage -> input of the next stage (radix RN, NSN = product of
// the radices before it): times w_L^{t k L / (NSN RN)} with t = q / (L / RN),
// k = (q mod L / RN) mod NSN, packed into LDS
template <int DT, int L, int R0, int RN, int NSN>
__device__ __forceinline__ void to_lds(uint32_t *buf, const float2 *tw, int g, int q, float2 v)
{
    constexpr int LRN = L / RN;
    constexpr int TST = L / (NSN * RN);
    const int tn = q / LRN, jn = q - tn * LRN;
    const int k = jn % NSN;
    v = cmul(v, tw[tn * k * TST]);
    buf[g * MGeom<L, R0>::RP + MGeom<L, R0>::phi(q)] = pack_c<DT>(v.x, v.y);
}

template <int LM, int SM, int DT, int L, int R0, int R1, int R2, int G, bool NF>
__global__ void __launch_bounds__(MP_THREADS) mpass_kernel(PassArgs a, int ngroups)
{
    using Geo = MGeom<L, R0>;
    constexpr bool S3 = R2 > 1;
    __shared__ uint32_t buf0[G * Geo::RP];
    __shared__ uint32_t buf1[S3 ? G * Geo::RP : 1];
    __shared__ float2 tws[L];
    __shared__ uint32_t carr[LM == LD_MIX ? 256 : 1];
    __shared__ PeakPart red[MP_THREADS / 64];
    int b, grp;
    if (!block_item(a, LM == LD_PROD, ngroups, &b, &grp)) return;
    const int g0 = grp * G;
    const int gcount = min(G, a.nsub - g0);
    for (int i = threadIdx.x; i < L; i += MP_THREADS) tws[i] = a.tw[i];
    if constexpr (LM == LD_MIX) {
        for (int i = threadIdx.x; i < 256; i += MP_THREADS) carr[i] = a.carr[i];
    }
    __syncthreads();

    // ---- stage 0: global -> MFMA -> LDS ----
    {
        const int igs = (int)a.in_gstride, ins = (int)a.in_nstride;
        const uint32_t *xin = nullptr, *cin = nullptr;
        Loader<LD_MIX> mx;
        if constexpr (LM == LD_MIX) mx.init(a, b, carr);
        if constexpr (LM == LD_CPX) xin = a.hin + (long long)b * a.in_bstride;
        if constexpr (LM == LD_PROD) {
            xin = a.hin + (long long)((b / a.prod_xdiv) % a.prod_xmod) * a.in_bstride;
            cin = a.hcf + (long long)(b / a.prod_cdiv) * a.cf_bstride;
        }
        auto ld = [&](int g, int n) -> uint32_t {
            if (g >= gcount) return 0u;
            const int e = (g0 + g) * igs + n * ins;
            if constexpr (LM == LD_MIX) {
                const int2 m = mx.mix_int(e);
                return pack_c<DT>((float)m.x, (float)m.y);
            } else if constexpr (LM == LD_CPX) {
                return xin[e];
            } else {
                const float2 p = cmul(unpack_c<DT>(xin[e]), unpack_c<DT>(cin[e]));
                return pack_c<DT>(p.x, p.y);
            }
        };
        auto st = [&](int g, int q, float2 v) { to_lds<DT, L, R0, R1, R0>(buf0, tws, g, q, v); };
        mstage<DT, L, R0, 1, G, NF>(a.amat[0], ld, st);
    }
    __syncthreads();

    // ---- last stage: LDS -> MFMA -> global ----
    Storer<ST_POW> sp;
    if constexpr (SM == ST_POW) sp.init(a, b);
    uint32_t *hout = nullptr;
    if constexpr (SM == ST_CPX) hout = a.hout + (long long)b * a.out_bstride;
    const float2 *tf = a.tfull;
    const float osc = a.osc, kap = a.kappa;
    const int ogs = (int)a.out_gstride, ons = (int)a.out_nstride;
    auto outg = [&](int g, int k, float2 v) {
        if (g >= gcount) return;
        const int gg = g0 + g;
        if constexpr (SM == ST_POW) {
            sp.store(gg, k, make_float2(v.x * kap, v.y * kap));
        } else {
            // the twiddle table is laid out like the output (index e)
            const int e = gg * ogs + k * ons;
            if (tf) v = cmul(v, tf[e]);
            else v = make_float2(v.x * osc, v.y * osc);
            hout[e] = pack_c<DT>(v.x, v.y);
        }
    };
    auto ld0 = [&](int g, int n) -> uint32_t { return buf0[g * Geo::RP + Geo::phi(n)]; };
    if constexpr (!S3) {
        mstage<DT, L, R1, R0, G, NF>(a.amat[1], ld0, outg);
    } else {
        auto st1 = [&](int g, int q, float2 v) { to_lds<DT, L, R0, R2, R0 * R1>(buf1, tws, g, q, v); };
        mstage<DT, L, R1, R0, G, NF>(a.amat[1], ld0, st1);
        __syncthreads();
        auto ld1 = [&](int g, int n) -> uint32_t { return buf1[g * Geo::RP + Geo::phi(n)]; };
        mstage<DT, L, R2, R0 * R1, G, NF>(a.amat[2], ld1, outg);
    }
    if constexpr (SM == ST_POW) {
        if (a.p_part) {
            const long long row = (long long)(b / a.p_div + a.p_off1) * a.p_mul + a.p_off2 + b % a.p_div;
            sp.flush_peak(a.p_part, row * ngroups + grp, red);
        }
    }
}

// A operand of the R-point DFT in the 32x32x16 layout: lane l, element e of
// row tile r and K step k holds A[i][kk], i = 32 r + l % 32, kk = 16 k +
// 8 (l / 32) + e; row 2t' (+1) is the real (imaginary) part of output t',
// column 2t (+1) multiplies the real (imaginary) part of input t.  lo = 1
// writes the rounding remainder (coefficient minus its rounded value).
template <int DT>
__global__ void amat_kernel(int R, int dir, int lo, uint4 *out)
{
    const int KS = (2 * R + 15) / 16;
    const int r = blockIdx.x / KS, k = blockIdx.x - r * KS;
    const int lane = threadIdx.x;
    const int i = 32 * r + (lane & 31);
    const int to = i >> 1, po = i & 1;
    uint32_t w[4];
    for (int e2 = 0; e2 < 4; e2++) {
        float v[2];
        for (int pi = 0; pi < 2; pi++) {
            const int kk = 16 * k + 8 * (lane >> 5) + 2 * e2 + pi;
            const int ti = kk >> 1;
            double x = 0.0;
            if (to < R && ti < R) {
                double s, c;
                sincospi(2.0 * (double)((to * ti) % R) / R, &s, &c);
                if (dir == 0) s = -s;
                x = po == 0 ? (pi == 0 ? c : -s) : (pi == 0 ? s : c);
            }
            v[pi] = (float)x;
            if (lo) {
                const float2 hi = unpack_c<DT>(pack_c<DT>(v[pi], 0.0f));
                v[pi] = (float)(x - (double)hi.x);
            }
        }
        w[e2] = pack_c<DT>(v[0], v[1]);
    }
    out[blockIdx.x * 64 + lane] = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void h_spec_kernel(const float2 *__restrict__ in, uint32_t *__restrict__ out, int M, int M1,
                              int M2, long long total, float scale, int dt)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    long long b = t / M;
    int i = (int)(t - b * M);          // transposed index k1 * M2 + k2
    int k1 = i / M2, k2 = i - k1 * M2;
    const float2 v = in[b * M + k1 + (long long)M1 * k2];
    out[t] = dt == HDT_F16 ? pack_c<HDT_F16>(v.x * scale, v.y * scale)
                           : pack_c<HDT_BF16>(v.x * scale, v.y * scale);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// (L, R0, R1, R2, G): column-pass lengths (G columns = 128 bytes of words per
// row segment; 16 for 960 to fit two workgroups per CU) and row-pass lengths
// (G rows, about 4-8 K points per workgroup, G L / R a multiple of 32 where
// possible).  Splits: 24000 = 150 x 160, 48000 = 200 x 240, 96000 = 300 x 320,
// 192000 = 400 x 480, 960000 = 960 x 1000.
#define PSA_HCOL(X) X(150, 10, 15, 1, 32) X(200, 10, 20, 1, 32) X(300, 15, 20, 1, 32) \
    X(400, 16, 25, 1, 32) X(960, 32, 30, 1, 16)
#define PSA_HROW(X) X(160, 16, 10, 1, 32) X(240, 16, 15, 1, 32) X(320, 16, 20, 1, 16) \
    X(480, 16, 30, 1, 16) X(1000, 10, 10, 10, 4)

struct HSpec { int L, r[3], G; };
static const HSpec h_col_specs[] = {
#define PSA_S(L_, A_, B_, C_, G_) {L_, {A_, B_, C_}, G_},
    PSA_HCOL(PSA_S)
};
static const HSpec h_row_specs[] = {
    PSA_HROW(PSA_S)
#undef PSA_S
};

static const HSpec *find_spec(const HSpec *t, int n, int L)
{
    for (int i = 0; i < n; i++) if (t[i].L == L) return &t[i];
    return nullptr;
}
static const HSpec *col_spec(int L) { return find_spec(h_col_specs, (int)(sizeof(h_col_specs) / sizeof(HSpec)), L); }
static const HSpec *row_spec(int L) { return find_spec(h_row_specs, (int)(sizeof(h_row_specs) / sizeof(HSpec)), L); }

static const uint4 *make_amat(int R, int dir, int dt)
{
    static std::mutex mtx;
    static std::map<int, uint4 *> cache;
    std::lock_guard<std::mutex> lk(mtx);
    const int key = (R * 2 + dir) * 2 + dt;
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const int KS = (2 * R + 15) / 16, RT = (2 * R + 31) / 32;
    const int n = 64 * KS * RT;         // uint4 per table
    uint4 *d;
    PSA_CHECK(hipMalloc(&d, sizeof(uint4) * n * (dt == HDT_BF16 ? 2 : 1)));
    if (dt == HDT_F16) {
        hipLaunchKernelGGL(amat_kernel<HDT_F16>, dim3(KS * RT), dim3(64), 0, 0, R, dir, 0, d);
    } else {
        hipLaunchKernelGGL(amat_kernel<HDT_BF16>, dim3(KS * RT), dim3(64), 0, 0, R, dir, 0, d);
        hipLaunchKernelGGL(amat_kernel<HDT_BF16>, dim3(KS * RT), dim3(64), 0, 0, R, dir, 1, d + n);
    }
    PSA_CHECK(hipGetLastError());
    PSA_CHECK(hipDeviceSynchronize());
    cache[key] = d;
    return d;
}

// The inter-pass twiddle in the index space of the pass outputs that take
// it: T[e] = scale w_M^{sign (e / M2)(e % M2)}.  The forward column pass
// stores X[k1][n2] at e = k1 M2 + n2 (twiddle w_M^{-n2 k1}), the inverse row
// pass Z[k1][n] at e = k1 M2 + n (twiddle w_M^{+k1 n}); both read T[e] with
// the offset of their own store, so the loads are coalesced.
static float2 *make_digit_twiddle(int M, int M2, int sign, double scale)
{
    const double PI_D = 3.14159265358979323846;
    std::vector<float2> h(M);
    for (int e = 0; e < M; e++) {
        const long long m = (long long)(e / M2) * (e % M2) % M;
        const double ang = sign * 2.0 * PI_D * (double)m / M;
        h[e] = make_float2((float)(scale * cos(ang)), (float)(scale * sin(ang)));
    }
    float2 *d;
    PSA_CHECK(hipMalloc(&d, sizeof(float2) * M));
    PSA_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
    return d;
}

static float pow2_near(double x)    // 2^round(log2 x)
{
    return (float)ldexp(1.0, (int)lround(log2(x)));
}

const HPlan *hfft_plan(int M, int dt)
{
    static std::mutex mtx;
    static std::map<int, HPlan *> plans;
    if (dt != HDT_F16 && dt != HDT_BF16) return nullptr;
    int M1 = 0, M2 = 0;
    for (const HSpec &c : h_col_specs) {
        if (M % c.L == 0 && row_spec(M / c.L)) { M1 = c.L; M2 = M / c.L; break; }
    }
    if (!M1) return nullptr;
    ensure_init();
    std::lock_guard<std::mutex> lk(mtx);
    auto it = plans.find(M * 2 + dt);
    if (it != plans.end()) return it->second;
    HPlan *p = new HPlan();
    p->M = M; p->M1 = M1; p->M2 = M2; p->dt = dt;
    // forward: the IF products are integers up to 128 in magnitude; 1/8 and
    // 1/sqrt(L) per pass keep |X| <= 23 sqrt(M) (inside FP16 up to M = 960000);
    // inverse row pass: unscaled (the code-spectrum scale keeps |Y| <= |X|)
    p->osc_fc = 0.125f / pow2_near(sqrt((double)M1));
    p->osc_fr = 1.0f / pow2_near(sqrt((double)M2));
    p->osc_ir = 1.0f;
    for (int dir = 0; dir < 2; dir++) {
        const int sg = dir == 0 ? -1 : 1;
        p->tw1[dir] = make_twiddle(M1, sg, 1.0);
        p->tw2[dir] = make_twiddle(M2, sg, 1.0);
        p->tfull[dir] = make_digit_twiddle(M, M2, sg, dir == 0 ? p->osc_fc : p->osc_ir);
        const HSpec *cs = col_spec(M1), *rs = row_spec(M2);
        for (int s = 0; s < 3; s++) {
            p->am1[dir][s] = cs->r[s] > 1 ? make_amat(cs->r[s], dir, dt) : nullptr;
            p->am2[dir][s] = rs->r[s] > 1 ? make_amat(rs->r[s], dir, dt) : nullptr;
        }
    }
    plans[M * 2 + dt] = p;
    return p;
}

void h_col_args(const HPlan *hp, int dir, PassArgs &a)
{
    const HSpec *cs = col_spec(hp->M1);
    a.L = hp->M1;
    a.nsub = hp->M2;
    a.G = cs->G;
    a.in_nstride = hp->M2; a.in_gstride = 1;
    a.out_nstride = hp->M2; a.out_gstride = 1;
    a.dir = dir;
    a.tw = hp->tw1[dir];
    a.M = hp->M;
    a.thi = a.tlo = nullptr;
    a.tfull = nullptr;
    a.osc = 1.0f;
    a.kappa = 1.0f;
    for (int s = 0; s < 3; s++) a.amat[s] = hp->am1[dir][s];
}

void h_row_args(const HPlan *hp, int dir, PassArgs &a)
{
    const HSpec *rs = row_spec(hp->M2);
    a.L = hp->M2;
    a.nsub = hp->M1;
    a.G = rs->G;
    a.in_nstride = 1; a.in_gstride = hp->M2;
    a.out_nstride = 1; a.out_gstride = hp->M2;
    a.dir = dir;
    a.tw = hp->tw2[dir];
    a.M = hp->M;
    a.thi = a.tlo = nullptr;
    a.tfull = nullptr;
    a.osc = 1.0f;
    a.kappa = 1.0f;
    for (int s = 0; s < 3; s++) a.amat[s] = hp->am2[dir][s];
}

template <int LM, int SM, int DT, int L, int R0, int R1, int R2, int G, bool NF>
static void hlaunch_t(const PassArgs &a, int nbatch, hipStream_t st)
{
    const int ngroups = (a.nsub + G - 1) / G;
    long long nb = (long long)ngroups * nbatch;
    if (LM == LD_PROD && a.xcd_nbin > 0) {
        const long long ntile = (long long)ngroups * a.xcd_nbin;
        nb = (ntile + 7) / 8 * 8 * a.xcd_nprn;
    }
    if (nb <= 0) return;
    if (nb > 0x7FFFFFFFLL) PSA_FAIL("FFT launch too large (%lld workgroups)", nb);
    hipLaunchKernelGGL((mpass_kernel<LM, SM, DT, L, R0, R1, R2, G, NF>), dim3((unsigned)nb), dim3(MP_THREADS), 0,
                       st, a, ngroups);
    PSA_CHECK(hipGetLastError());
}

template <int DT>
static void hlaunch_dt(LoadMode lm, StoreMode sm, const PassArgs &a, int nbatch, hipStream_t st)
{
    // column lengths: forward (mixing) and inverse (power) column passes
#define PSA_C(L_, A_, B_, C_, G_)                                                          \
    if (a.L == L_ && a.in_nstride != 1) {                                                  \
        if (lm == LD_MIX && sm == ST_CPX) { hlaunch_t<LD_MIX, ST_CPX, DT, L_, A_, B_, C_, G_, false>(a, nbatch, st); return; } \
        if (lm == LD_CPX && sm == ST_POW) { hlaunch_t<LD_CPX, ST_POW, DT, L_, A_, B_, C_, G_, false>(a, nbatch, st); return; } \
    }
    PSA_HCOL(PSA_C)
#undef PSA_C
    // row lengths: forward row pass and inverse (code product) row pass
#define PSA_R(L_, A_, B_, C_, G_)                                                          \
    if (a.L == L_ && a.in_nstride == 1) {                                                  \
        if (lm == LD_CPX && sm == ST_CPX) { hlaunch_t<LD_CPX, ST_CPX, DT, L_, A_, B_, C_, G_, true>(a, nbatch, st); return; } \
        if (lm == LD_PROD && sm == ST_CPX) { hlaunch_t<LD_PROD, ST_CPX, DT, L_, A_, B_, C_, G_, true>(a, nbatch, st); return; } \
    }
    PSA_HROW(PSA_R)
#undef PSA_R
    PSA_FAIL("no MFMA pass for length %d, mode %d/%d", a.L, (int)lm, (int)sm);
}

void launch_hpass(int dt, LoadMode lm, StoreMode sm, const PassArgs &a_in, int nbatch, hipStream_t st)
{
    PassArgs a = a_in;
    a.carr = dev_carrier();
    if (dt == HDT_F16) hlaunch_dt<HDT_F16>(lm, sm, a, nbatch, st);
    else hlaunch_dt<HDT_BF16>(lm, sm, a, nbatch, st);
}

void h_code_spec(const HPlan *hp, const float2 *nat, uint32_t *out, int n, float scale, hipStream_t st)
{
    long long total = (long long)hp->M * n;
    if (total <= 0) return;
    hipLaunchKernelGGL(h_spec_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, nat, out, hp->M,
                       hp->M1, hp->M2, total, scale, hp->dt);
    PSA_CHECK(hipGetLastError());
}

}  // namespace psa
