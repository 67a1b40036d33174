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
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>
#include "psa_pass.h"

namespace psa {

typedef _Float16 hx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bx8 __attribute__((ext_vector_type(8)));
typedef float fx16 __attribute__((ext_vector_type(16)));


typedef _Float16 hx2 __attribute__((ext_vector_type(2)));
typedef __bf16 bx2 __attribute__((ext_vector_type(2)));
typedef float fx2 __attribute__((ext_vector_type(2)));

// complex value -> one word (re in the low half), round to nearest even
// (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32)
template <int DT>
__device__ __forceinline__ uint32_t pack_c(float re, float im)
{
    const fx2 v = {re, im};
    if constexpr (DT == HDT_F16) return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, hx2));
    else return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bx2));
}

template <int DT>
__device__ __forceinline__ float2 unpack_c(uint32_t w)
{
    if constexpr (DT == HDT_F16) {
        const fx2 v = __builtin_convertvector(__builtin_bit_cast(hx2, w), fx2);
        return make_float2(v.x, v.y);
    } else {
        return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xFFFF0000u));
    }
}

// x * c of two packed words: FP16 in packed half arithmetic (the product of
// two words in range stays in range, |c| <= 1), BF16 through FP32
template <int DT>
__device__ __forceinline__ uint32_t cmul_w(uint32_t xw, uint32_t cw)
{
    if constexpr (DT == HDT_F16) {
        const hx2 x = __builtin_bit_cast(hx2, xw), c = __builtin_bit_cast(hx2, cw);
        const hx2 t = x * c.xx;                                    // (xr cr, xi cr)
        const hx2 u = {-c.y, c.y};
        return __builtin_bit_cast(uint32_t, __builtin_elementwise_fma(x.yx, u, t));  // (xr cr - xi ci, xi cr + xr ci)
    } else {
        const float2 p = cmul(unpack_c<DT>(xw), unpack_c<DT>(cw));
        return pack_c<DT>(p.x, p.y);
    }
}

// raw buffer access with 32-bit element offsets from a per-item base (no
// 64-bit address arithmetic per element).  The resource covers 0x7FFFFFFF
// bytes; a lane with ok = false uses byte offset 0x80000000, past that
// range, where the hardware's range check returns 0 for loads and drops
// stores (masking without branches).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ int boff(int i, uint32_t size, bool ok)
{
    return (int)(ok ? (uint32_t)i * size : 0x80000000u);
}
__device__ __forceinline__ uint32_t bload32(__amdgpu_buffer_rsrc_t r, int i, bool ok = true)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, boff(i, 4, ok), 0, 0);
}
__device__ __forceinline__ float2 bload_f2(__amdgpu_buffer_rsrc_t r, int i)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, boff(i, 8, true), 0, 0));
}
// element base + c of a word / float2 array, masked on the base: with c a
// compile-time constant the select is shared by the accesses of one base and
// c becomes the instruction's immediate offset
__device__ __forceinline__ uint32_t bload32_at(__amdgpu_buffer_rsrc_t r, int base, int c, bool ok)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)((uint32_t)boff(base, 4, ok) + (uint32_t)c * 4u), 0, 0);
}
__device__ __forceinline__ float2 bload_f2_at(__amdgpu_buffer_rsrc_t r, int base, int c, bool ok)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                          r, (int)((uint32_t)boff(base, 8, ok) + (uint32_t)c * 8u), 0, 0));
}
__device__ __forceinline__ void bstore32(__amdgpu_buffer_rsrc_t r, int i, uint32_t v, bool ok = true)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, boff(i, 4, ok), 0, 0);
}
// (the masked base passes through an empty asm statement: otherwise the
// compiler folds + c into both arms of the select, one select and one add
// per store instead of a shared base and immediate offsets)
__device__ __forceinline__ void bstore32_at(__amdgpu_buffer_rsrc_t r, int base, int c, uint32_t v, bool ok)
{
    int vb = boff(base, 4, ok);
    asm("" : "+v"(vb));
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)((uint32_t)vb + (uint32_t)c * 4u), 0, 0);
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
// at n + n / R0 when R0 is even (one pad word after every R0 words, so the
// first stage's stride-R0 stores spread over the banks), else at n.  The row
// pitch RP is odd so that lanes walking across rows do too.
template <int L, int R0>
struct MGeom {
    static constexpr int PAD = (R0 % 2 == 0) ? 1 : 0;
    static constexpr int RP = (L + (PAD ? L / R0 : 0)) | 1;
};

// Power-pass shapes that take paired blocks (measured faster): their LDS
// limits the CU to one 512-thread workgroup (960 and 1000 with 32 columns:
// up to 256 registers) or to two (400 as 16 x 25 with 32 columns: up to 128
// registers), and the kept powers (THL outputs of the last radix per lane
// and tile) fit those registers (the compiler's counts: 202, 241 and 108);
// 300 as 15 x 20 (three workgroups per CU by LDS) loses one of them to the
// kept powers and still moves enough fewer bytes to gain 11%
constexpr bool pb_shape(int L, int R0, int R1, int R2, int G, int NTH)
{
    return R2 == 1 && G == 32 && NTH == 512 &&
           ((L == 960 && R0 == 32 && R1 == 30) || (L == 1000 && R0 == 25 && R1 == 40) ||
            (L == 400 && R0 == 16 && R1 == 25) || (L == 300 && R0 == 15 && R1 == 20));
}

// One Stockham radix-R stage (NS = product of the earlier radices) over the
// G sub-transforms of the workgroup.  Column (g, j), j < L / R, reads inputs
// t < R at j + t L / R; its output t' sits at the Stockham position
// (j / NS) NS R + j % NS + t' NS.  The callers' functors split addressing
// into a per-column part, computed once per tile, and a per-element part:
//   ib = ibase(g, j)      base of the column's inputs (rows past the end of
//                         the item clamped to a valid one),
//   raw = iload(ib, tc, th, ok)      issues the loads of input t,
//   word = icook(ib, tc, th, ok, raw)  forms the packed input from them,
//   ob = obase(g, j)      (.x, .y, .z: bases and stride, .w: row valid),
//   f = omul(ob, tc, th, ok)  multiplier of output t' - complex for MULK 2,
//                         its real part for MULK 1 (.y then carries a value
//                         for oput), none for 0; MULK 3: f is passed to oput
//                         unapplied (a packed FP16 twiddle),
//   oput(ob, tc, th, v, f, ok)  stores output t' (FP32, multiplied) if ok.
// Element indices come split as t = tc + th: tc a compile-time constant of
// the unrolled loops, th = 4 h (inputs) or 2 h (outputs) from the lane half
// h, so that per-element offsets fold into immediates; ok = false marks an
// input t >= R (its value is not used) or an output that is not stored.
// NF: columns run along j (rows contiguous in memory) rather than along g.
// Outputs t' >= TH are not produced.
//
// Inputs t >= R of the B operand meet zero columns of the DFT table, and the
// results of columns past the end are discarded, so inputs are loaded on
// clamped indices and never masked.  Each wave walks its tiles
// software-pipelined (PF, stages that read global memory): the loads of the
// next tile are issued before this tile's MFMAs and only consumed (icook)
// after this tile's outputs are stored, so they stay in flight across the
// whole tile; the loop body has no branches (masked outputs select a dummy
// LDS word or an out-of-range buffer offset), so nothing pulls a wait forward.
// The A operands of one stage (R-point DFT, outputs t' < TH) in registers.
// load() issues the loads; complete() waits for them.  Callers issue the
// loads early (before a barrier or other loads) and complete them before the
// stage's tile loop: left pending into the loop, the wait the first MFMA needs
// for them would also wait, in every iteration, for the previous tile's
// output stores.
template <int DT, int R, int TH>
struct ATab {
    static constexpr int KS = (2 * R + 15) / 16;       // K steps of 16 halves
    static constexpr int RT = (2 * TH + 31) / 32;      // 32-row output tiles needed
    static constexpr int RTA = (2 * R + 31) / 32;      // row tiles of the table
    // BF16: A = A_hi + A_lo (the second table follows the first), so the DFT
    // coefficients carry 16 significant bits
    static constexpr int NA = DT == HDT_BF16 ? 2 : 1;
    uint4 v[NA][RT][KS];
    __device__ __forceinline__ void load(const uint4 *__restrict__ amat)
    {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int x = 0; x < NA; x++) {
#pragma unroll
            for (int r = 0; r < RT; r++) {
#pragma unroll
                for (int k = 0; k < KS; k++) v[x][r][k] = amat[((x * RTA + r) * KS + k) * 64 + lane];
            }
        }
    }
    __device__ __forceinline__ void complete() const
    {
#pragma unroll
        for (int x = 0; x < NA; x++) {
#pragma unroll
            for (int r = 0; r < RT; r++) {
#pragma unroll
                for (int k = 0; k < KS; k++) {
                    const uint4 &q = v[x][r][k];
                    asm volatile("" : : "v"(q.x), "v"(q.y), "v"(q.z), "v"(q.w));
                }
            }
        }
    }
};

// pre() runs after the first tile's loads are issued and before anything
// uses them or the stage's table (e.g. filling LDS tables and the barrier);
// the stage then completes the table loads.
// ALL: every tile's inputs of the wave are read before pre() and the first
// output store (stages whose inputs and outputs share one LDS buffer, pre()
// then being the barrier between them; first stages with few tiles, whose
// global loads are then all in flight together).
// UNR: the tile loop fully unrolled (stages reading LDS), ob.y of each tile
// set to its iteration index, so that the functors can keep per-tile values
// in registers (the paired-block power pass).
template <int DT, int NTH, int L, int R, int G, bool NF, int TH, int MULK, bool PF, bool ALL = false,
          bool MASKT = true, bool PFM = false, bool UNR = false, class IBase, class ILoad, class ICook, class OBase,
          class OMul, class OPut, class Pre>
__device__ __forceinline__ void mstage(const ATab<DT, R, TH> &at, const IBase &ibase, const ILoad &iload,
                                       const ICook &icook, const OBase &obase, const OMul &omul,
                                       const OPut &oput, const Pre &pre)
{
    constexpr int LR = L / R;
    constexpr int C = G * LR;
    constexpr int NT = (C + 31) / 32;
    constexpr int KS = ATab<DT, R, TH>::KS;
    constexpr int RT = ATab<DT, R, TH>::RT;
    constexpr int NA = ATab<DT, R, TH>::NA;
    constexpr int NW = NTH / 64;
    // the wave index in a scalar register: tile loop control is wave-uniform
    const int lane = threadIdx.x & 63, h = lane >> 5;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const auto &A = at.v;
    // column of this lane in a tile; past the last column the indices are
    // clamped to a valid one (the lane's results are then discarded)
    auto column = [&](int tile, int &g, int &j) -> bool {
        const int c = tile * 32 + (lane & 31);
        const bool ok = (C % 32 == 0) || c < C;
        const int cc = ok ? c : C - 1;
        if (NF) { g = cc / LR; j = cc - g * LR; }
        else { j = cc / G; g = cc - j * G; }
        return ok;
    };
    // B: lane (column, h) supplies inputs t = 8 k + 4 h + u, u < 4 (inputs
    // t >= R read t = R - 1)
    using Raw = decltype(iload(0, 0, 0, true));
    // input t = tc + 4 h is used if t < R
    auto inok = [&](int tc) -> bool { return tc + 4 < R || (tc < R && h == 0); };
    auto fetch = [&](int ib, Raw (&w)[KS][4]) {
#pragma unroll
        for (int k = 0; k < KS; k++) {
#pragma unroll
            for (int u = 0; u < 4; u++) w[k][u] = iload(ib, 8 * k + u, 4 * h, inok(8 * k + u));
        }
    };
    auto cook = [&](int ib, const Raw (&w)[KS][4], uint4 *B) {
#pragma unroll
        for (int k = 0; k < KS; k++) {
            uint32_t c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) c[u] = icook(ib, 8 * k + u, 4 * h, inok(8 * k + u), w[k][u]);
            B[k] = make_uint4(c[0], c[1], c[2], c[3]);
        }
    };
    // output t' = tc + 2 h is produced if t' < TH
    auto outok = [&](int tc) -> bool { return tc + 2 < TH || (tc < TH && h == 0); };
    // the multipliers of a tile's outputs (accumulator 4 m + 2 w (+1) of row
    // tile r is output t' = 16 r + 4 m + 2 h + w, real (imaginary) part);
    // PFM: fetched one tile ahead (ALL: with the inputs), so that global
    // multipliers (twiddle words, previous P values) are in flight across
    // the tile before - at the cost of registers
    using MulT = decltype(omul(int4{}, 0, 0, true));
    struct Muls { MulT v[RT][4][2]; };
    auto get_muls = [&](const int4 &ob, Muls &mu) {
#pragma unroll
        for (int r = 0; r < RT; r++) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
#pragma unroll
                for (int w = 0; w < 2; w++) {
                    const int tc = 16 * r + 4 * m + w;
                    if (MULK && tc < TH) mu.v[r][m][w] = omul(ob, tc, 2 * h, outok(tc));
                }
            }
        }
    };
    // one tile: MFMAs, outputs
    auto tile_body = [&](const int4 &ob, bool cv, const uint4 (&B)[KS], const Muls &mu) {
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
        // MASKT: ok marks lanes past the last column or row; without it only
        // outputs t' >= TH are masked (the caller's invalid lanes then store
        // duplicates of valid ones, or into rows nothing reads), and the
        // masks of all but the outputs next to TH fold away
        const bool ok = MASKT ? (cv && ob.w) : true;
#pragma unroll
        for (int r = 0; r < RT; r++) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
#pragma unroll
                for (int w = 0; w < 2; w++) {
                    const int tc = 16 * r + 4 * m + w;
                    if (tc >= TH) continue;      // past TH in both halves
                    float2 v = make_float2(acc[r][4 * m + 2 * w], acc[r][4 * m + 2 * w + 1]);
                    const MulT f = MULK ? mu.v[r][m][w] : MulT{};
                    if constexpr (MULK == 2) v = cmul(v, f);
                    else if constexpr (MULK == 1) v = make_float2(v.x * f.x, v.y * f.x);
                    oput(ob, tc, 2 * h, v, f, ok && outok(tc));
                }
            }
        }
    };
    // all of the wave's tiles read before the first is computed
    constexpr int MT = (NT + NW - 1) / NW;
    if constexpr (ALL) {
        static_assert(MT <= 6, "all tiles in flight: at most 6 per wave");
        Raw rw[MT][KS][4];
        Muls mus[MT];
        int ibs[MT];
        int4 obs[MT];
        bool cvs[MT];
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const int ti = wave + i * NW;
            int g, j;
            cvs[i] = column(ti < NT ? ti : NT - 1, g, j);
            ibs[i] = ibase(g, j);
            obs[i] = obase(g, j);
            fetch(ibs[i], rw[i]);
            if (PFM) get_muls(obs[i], mus[i]);
        }
        pre();
        at.complete();
#pragma unroll
        for (int i = 0; i < MT; i++) {
            if (wave + i * NW < NT) {
                uint4 Bt[KS];
                cook(ibs[i], rw[i], Bt);
                if (!PFM) get_muls(obs[i], mus[i]);
                tile_body(obs[i], cvs[i], Bt, mus[i]);
            }
        }
        return;
    }
    if constexpr (UNR) {
        static_assert(!PF, "unrolled tile loop: stages reading LDS");
        Muls mus[2];
        {
            int g = 0, j = 0;
            column(wave < NT ? wave : NT - 1, g, j);
            int4 ob = obase(g, j);
            ob.y = 0;
            pre();
            if (PFM) get_muls(ob, mus[0]);
        }
        at.complete();
#pragma unroll
        for (int it = 0; it < MT; it++) {
            const int tl = wave + it * NW;
            if (tl < NT) {
                int g = 0, j = 0;
                const bool cv = column(tl, g, j);
                int4 ob = obase(g, j);
                ob.y = it;
                if (PFM && it + 1 < MT) {
                    // the next tile's multipliers (a wave-uniform branch)
                    if (tl + NW < NT) {
                        int gn = 0, jn = 0;
                        column(tl + NW, gn, jn);
                        int4 obn = obase(gn, jn);
                        obn.y = it + 1;
                        get_muls(obn, mus[(it + 1) & 1]);
                    }
                }
                const int ib = ibase(g, j);
                Raw raw[KS][4];
                uint4 B[KS];
                fetch(ib, raw);
                cook(ib, raw, B);
                if (!PFM) get_muls(ob, mus[it & 1]);
                tile_body(ob, cv, B, mus[it & 1]);
            }
        }
        return;
    }
    // (a wave without tiles fetches the last one, unused, so that no load
    // sits in a branch)
    int tile = wave, g = 0, j = 0;
    bool cv = column(tile < NT ? tile : NT - 1, g, j);
    uint4 B[KS];
    Raw raw[KS][4];
    Muls mu;
    int4 ob = obase(g, j);
    const int ib0 = ibase(g, j);
    if (PF) fetch(ib0, raw);
    pre();
    if (PFM) get_muls(ob, mu);
    at.complete();
    if (PF) cook(ib0, raw, B);
    for (; tile < NT; tile += NW) {
        const int tn = tile + NW;
        int gn = 0, jn = 0, ibn = 0;
        bool cvn = false;
        int4 obn = ob;
        Muls mun;
        // next tile (a wave-uniform branch: its loads are not waited on
        // inside it)
        if (tn < NT && (PF || PFM)) {
            cvn = column(tn, gn, jn);
            obn = obase(gn, jn);
            if (PFM) get_muls(obn, mun);
            if (PF) {
                ibn = ibase(gn, jn);
                fetch(ibn, raw);
            }
        }
        if (!PF) {
            const int ib = ibase(g, j);
            fetch(ib, raw);
            cook(ib, raw, B);
        }
        if (!PFM) get_muls(ob, mu);
        tile_body(ob, cv, B, mu);
        if (tn < NT) {
            if (PF) cook(ibn, raw, B);
            if (PF || PFM) {
                g = gn; j = jn; cv = cvn; ob = obn;
                if (PFM) mu = mun;
            } else {
                cv = column(tn, g, j);
                ob = obase(g, j);
            }
        }
    }
}

template <int LM, int SM, int DT, int L, int R0, int R1, int R2, int G, bool NF, int NTH>
__global__ void __launch_bounds__(NTH) mpass_kernel(PassArgs a, int ngroups)
{
    using Geo = MGeom<L, R0>;
    // code-product row passes with few tiles per wave stage their inputs in
    // LDS (STG, below); their rows are 16-byte aligned for 16-byte staging
    // stores and 16 banks apart
    constexpr int NT0 = (G * (L / R0) + 31) / 32, MT0 = (NT0 + NTH / 64 - 1) / (NTH / 64);
    constexpr bool STG = LM == LD_PROD && NF && MT0 <= 4 && L % 4 == 0;
    constexpr int RP = STG ? (L + (Geo::PAD ? L / R0 : 0) + 15) / 32 * 32 + 16 : Geo::RP;
    constexpr int PD = Geo::PAD;
    constexpr bool S3 = R2 > 1;
    constexpr int RL = S3 ? R2 : R1;          // last radix
    constexpr int NSL = L / RL;               // its NS (= columns per sub-transform)
    // power passes keep lags < M / 2 only: outputs t' >= ceil(RL / 2) of the
    // last stage land at k >= L / 2 (lag >= M / 2) and are not produced
    constexpr int THL = SM == ST_POW ? (RL + 1) / 2 : RL;
    // the forward column and inverse row passes apply the inter-pass twiddle
    constexpr bool TW = SM == ST_CPX && LM != LD_CPX;
    // FP16: unit-magnitude twiddles (the stage twiddles, the inverse row
    // pass's inter-pass twiddle) are applied as packed FP16 words after the
    // result is packed (v_cvt_pk + two packed-half ops instead of four FP32
    // ops and v_cvt_pk); |v w| = |v| keeps the packing range the same.  The
    // forward column pass's inter-pass twiddle carries its scale and stays
    // FP32, as does all of BF16.
    constexpr bool PKT = DT == HDT_F16;
    constexpr bool TWP = PKT && TW && LM == LD_PROD;
    using TwT = std::conditional_t<PKT, uint32_t, float2>;
    // one dummy word per lane after the rows takes the masked LDS stores
    __shared__ __align__(16) uint32_t buf0[G * RP + 64];
    __shared__ uint32_t buf1[S3 ? G * RP + 64 : 1];
    // the LDS tables are filled without branches (entries past the end are
    // copies, unused)
    constexpr int NTW = (L + NTH - 1) / NTH;
    __shared__ TwT tws[NTW * NTH];
    // carrier table, then a raw ring's table (256 words each)
    __shared__ uint32_t carr[LM == LD_MIX ? 512 : 1];
    __shared__ PeakPart red[NTH / 64];
    int b, grp;
    // column passes: G words (inverse) or G IF bytes (mixing) per row
    // segment, KX workgroups to one 128-byte line
    constexpr int SEGB = LM == LD_MIX ? G : 4 * G;
    constexpr int KX = (!NF && SEGB < 128) ? 128 / SEGB : 1;
    if (!block_item(a, LM == LD_PROD, ngroups, &b, &grp, KX)) return;
    const int g0 = grp * G;
    const int gcount = min(G, a.nsub - g0);
    // issue order: stage-0 table, the LDS tables' words, the first tile's
    // inputs (in mstage); the LDS fill then waits only for the first two
    ATab<DT, R0, R0> at0;
    at0.load(a.amat[0]);
    // stage twiddles: two-stage lengths as [t][j] = w_L^{t j} (t < R0,
    // j < L / R0), so that the lanes of a row pass, which walk j, read
    // consecutive words at immediate offsets; three-stage lengths as w_L^m
    float2 twv[NTW];
#pragma unroll
    for (int q = 0; q < NTW; q++) {
        const int i = min((int)threadIdx.x + q * NTH, L - 1);
        twv[q] = a.tw[S3 ? i : (i / (L / R0)) * (i % (L / R0))];
    }
    constexpr int NCW = LM == LD_MIX ? (512 + NTH - 1) / NTH : 1;
    uint32_t cav[NCW] = {};
    if constexpr (LM == LD_MIX) {
#pragma unroll
        for (int q = 0; q < NCW; q++) {
            const int i = (int)threadIdx.x + q * NTH;
            if (i < 256) cav[q] = a.carr[i];
            else if (i < 512 && a.raw_bps) cav[q] = reinterpret_cast<const uint32_t *>(a.raw_lut)[i - 256];
        }
    }
    auto fill = [&]() {
#pragma unroll
        for (int q = 0; q < NTW; q++) {
            if constexpr (PKT) tws[threadIdx.x + q * NTH] = pack_c<DT>(twv[q].x, twv[q].y);
            else tws[threadIdx.x + q * NTH] = twv[q];
        }
        if constexpr (LM == LD_MIX) {
#pragma unroll
            for (int q = 0; q < NCW; q++) {
                const int i = (int)threadIdx.x + q * NTH;
                if (i < 512) carr[i] = cav[q];
            }
        }
        __syncthreads();
    };
    auto barrier = []() { __syncthreads(); };

    // LDS stages: input t of column j of a later stage is element
    // n = j + t LRs (LRs = L / R of that stage), at n + n / R0 (PD)
    auto lds_base = [&](int g, int j) -> int { return g * RP + j + (PD ? j / R0 : 0); };
    const int dummy = G * RP + (int)(threadIdx.x & 63);

    // ---- last stage (LDS -> MFMA -> global): storers and multipliers ----
    Storer<ST_POW> sp;
    if constexpr (SM == ST_POW) sp.init(a, b);
    const __amdgpu_buffer_rsrc_t hout = rsrc(a.hout + (SM == ST_CPX ? (long long)b * a.out_bstride : 0LL));
    const __amdgpu_buffer_rsrc_t tfr = rsrc(TWP ? (const void *)a.tfullh : (const void *)a.tfull);
    const float osc = a.osc, kap2 = a.kappa * a.kappa;
    const int ogs = (int)a.out_gstride, ons = NF ? 1 : (int)a.out_nstride;
    const int estep = NSL * ons;
    // output t of column j (< NSL) is element k = j + t NSL of its
    // sub-transform, at e = gg ogs + k ons.  ST_CPX multiplies by the
    // inter-pass twiddle (its table is laid out like the output) or by osc;
    // ST_POW scales the power by kappa^2 (the multiplier is 1), and the
    // previous P value rides along in .y (fetched with the tile's other
    // loads; 0 on the first block)
    const bool pacc = SM == ST_POW && !a.p_init;
    const __amdgpu_buffer_rsrc_t prr = rsrc(sp.prow);
    const int nlag = a.nlag;
    auto gbase = [&](int g, int j) -> int4 {
        const bool ok = g < gcount;
        return make_int4((g0 + (ok ? g : 0)) * ogs + j * ons, 0, 0, ok);
    };
    auto gmul = [&](const int4 &ob, int tc, int th, bool ok) {
        if constexpr (SM == ST_POW) {
            const int e = ob.x + (tc + th) * estep;
            return make_float2(1.0f, __uint_as_float(bload32(prr, e, pacc && ok && e < nlag)));
        } else if constexpr (TWP) {
            return bload32_at(tfr, ob.x + th * estep, tc * estep, ok);
        } else if constexpr (TW) {
            return bload_f2_at(tfr, ob.x + th * estep, tc * estep, ok);
        } else {
            return make_float2(osc, 0.0f);
        }
    };
    using GMulT = decltype(gmul(int4{}, 0, 0, true));
    auto gput = [&](const int4 &ob, int tc, int th, float2 v, GMulT f, bool ok) {
        if constexpr (SM == ST_POW) {
            sp.store_pow_sel(ob.x + (tc + th) * estep, v, kap2, f.y, ok, [&](bool ok2, int e2, float nv) {
                bstore32(prr, e2, __float_as_uint(nv), ok2);
            });
        } else {
            uint32_t w = pack_c<DT>(v.x, v.y);
            if constexpr (TWP) w = cmul_w<DT>(w, f);
            bstore32_at(hout, ob.x + th * estep, tc * estep, w, ok);
        }
    };
    constexpr int MULKL = TWP ? 3 : TW ? 2 : 1;
    // stage 1 input t of column j: n = j + t R0 R2
    // (unused inputs read t = R - 1 of the stage: written words)
    auto iat0 = [&](int base, int tc, int th, bool ok) -> uint32_t {
        return buf0[base + (ok ? tc + th : R1 - 1) * (R0 * R2 + PD * R2)];
    };
    auto ident = [](int, int, int, bool, uint32_t w) -> uint32_t { return w; };
    // paired blocks (PB): a power pass whose LDS already limits the CU to
    // the workgroups its registers allow with the kept powers added (the
    // shapes of pb_shape), may take two blocks' inputs (hin, hin2) and
    // write P once
    constexpr int RLT = S3 ? R2 : R1;
    constexpr int NTL = (G * (L / RLT) + 31) / 32, MTL = (NTL + NTH / 64 - 1) / (NTH / 64);
    constexpr int RTL = (2 * THL + 31) / 32;
    constexpr bool PB = DT == HDT_F16 && SM == ST_POW && pb_shape(L, R0, R1, R2, G, NTH);
    float keep[PB ? MTL : 1][PB ? RTL * 8 : 1];
    // output tc = 16 r + 4 m + w of a tile: accumulator slot 8 r + 2 m + w
    auto kidx = [](int tc) { return (tc / 16) * 8 + ((tc % 16) / 4) * 2 + (tc % 4); };
    (void)kidx;
    // MODE 0: one block; 1, 2: the first and second block of a pair
    auto run_last = [&](auto mode_c) {
        constexpr int MODE = decltype(mode_c)::value;
        if constexpr (!S3) {
            ATab<DT, R1, THL> at1;
            at1.load(a.amat[1]);
            if constexpr (MODE == 0) {
                mstage<DT, NTH, L, R1, G, NF, THL, MULKL, false, false, true, SM == ST_POW>(at1, lds_base, iat0, ident,
                                                                                          gbase, gmul, gput, barrier);
            } else if constexpr (MODE == 1) {
                // first block of a pair: its accumulated power (with the previous
                // P value) stays in registers, by tile and output
                auto gput1 = [&](const int4 &ob, int tc, int th, float2 v, float2 f, bool) {
                    const float pw = kap2 * __builtin_fmaf(v.x, v.x, v.y * v.y);
                    keep[ob.y][kidx(tc)] = sp.pinit ? pw : f.y + pw;
                };
                mstage<DT, NTH, L, R1, G, NF, THL, MULKL, false, false, true, true, true>(at1, lds_base, iat0, ident,
                                                                                        gbase, gmul, gput1, barrier);
            } else {
                // second block: the kept value is the previous P value
                auto gmul2 = [&](const int4 &ob, int tc, int, bool) { return make_float2(1.0f, keep[ob.y][kidx(tc)]); };
                sp.pinit = false;
                mstage<DT, NTH, L, R1, G, NF, THL, MULKL, false, false, true, false, true>(at1, lds_base, iat0, ident,
                                                                                         gbase, gmul2, gput, barrier);
            }
        } else {
            // middle stage (NS = R0): output t of column j at q = (j / R0) R0 R1
            // + j % R0 + t R0, next twiddle w_L^{(j / R0)(j % R0 + t R0)}
            auto obase1 = [&](int g, int j) -> int4 {
                const int jh = j / R0, jl = j - jh * R0;
                return make_int4(g * RP + jh * R0 * R1 + jl + (PD ? jh * R1 : 0), jh * jl, jh * R0, 1);
            };
            auto omul1 = [&](const int4 &ob, int tc, int th, bool ok) -> TwT {
                return tws[ob.y + (ok ? tc + th : 0) * ob.z];
            };
            auto oput1 = [&](const int4 &ob, int tc, int th, float2 v, TwT f, bool ok) {
                uint32_t w = pack_c<DT>(v.x, v.y);
                if constexpr (PKT) w = cmul_w<DT>(w, f);
                buf1[ok ? ob.x + (tc + th) * (R0 + PD) : dummy] = w;
            };
            ATab<DT, R1, R1> at1;
            at1.load(a.amat[1]);
            mstage<DT, NTH, L, R1, G, NF, R1, PKT ? 3 : 2, false, false, false>(at1, lds_base, iat0, ident, obase1,
                                                                                 omul1, oput1, barrier);
            // stage 2 input t of column j: n = j + t R0 R1
            auto iat1 = [&](int base, int tc, int th, bool ok) -> uint32_t {
                return buf1[base + (ok ? tc + th : R2 - 1) * (R0 * R1 + PD * R1)];
            };
            ATab<DT, R2, THL> at2;
            at2.load(a.amat[2]);
            mstage<DT, NTH, L, R2, G, NF, THL, MULKL, false, false, true, SM == ST_POW>(at2, lds_base, iat1, ident,
                                                                                      gbase, gmul, gput, barrier);
        }
    };

    // ---- stage 0: global -> MFMA -> LDS ----
    // output t of column j sits at q = j R0 + t (LDS j (R0 + PD) + t); the
    // next stage's twiddle is w_L^{t (j / R2) R2} (w_L^{j t}: two stages,
    // tws[t L / R0 + j])
    auto obase = [&](int g, int j) -> int4 { return make_int4(g * RP + j * (R0 + PD), 0, (j / R2) * R2, 1); };
    auto omul = [&](const int4 &ob, int tc, int th, bool ok) -> TwT {
        if constexpr (S3) return tws[(ok ? tc + th : 0) * ob.z];
        else return tws[(ok ? tc + th : 0) * (L / R0) + ob.z];
    };
    auto oput = [&](const int4 &ob, int tc, int th, float2 v, TwT f, bool ok) {
        uint32_t w = pack_c<DT>(v.x, v.y);
        if constexpr (PKT) w = cmul_w<DT>(w, f);
        buf0[ok ? ob.x + tc + th : dummy] = w;
    };
    // Code-product row passes with few tiles per wave stage their inputs:
    // the workgroup loads its rows of X and Cf with 16-byte loads (instead of
    // one 4-byte load per element and operand, in runs of L / R0 words),
    // multiplies them (the same cmul_w as the per-element loader, so the
    // results do not change) and stores the products unpadded in buf0;
    // stage 0 then reads all its inputs from LDS before the barrier that
    // precedes its first output store, so buf0 holds both.
    bool paired = false;
    if constexpr (STG) {
        const long long xo = (long long)((b / a.prod_xdiv) % a.prod_xmod) * a.in_bstride;
        const long long co = (long long)(b / a.prod_cdiv) * a.cf_bstride;
        const __amdgpu_buffer_rsrc_t xin = rsrc(a.hin + xo), cin = rsrc(a.hcf + co);
        const int igs = (int)a.in_gstride;
        constexpr int Q = G * L / 4;              // 16-byte words of the rows
        constexpr int NQ = (Q + NTH - 1) / NTH;
        uint4 xv[NQ], cv[NQ];
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int qq = min((int)threadIdx.x + u * NTH, Q - 1);
            const int g = qq / (L / 4), n = (qq - g * (L / 4)) * 4;
            const int e = (g0 + (g < gcount ? g : 0)) * igs + n;
            xv[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xin, e * 4, 0, 0));
            cv[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(cin, e * 4, 0, 0));
        }
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q0 = (int)threadIdx.x + u * NTH;
            if (q0 < Q) {
                const int g = q0 / (L / 4), n = (q0 - g * (L / 4)) * 4;
                uint4 pw;
                pw.x = cmul_w<DT>(xv[u].x, cv[u].x);
                pw.y = cmul_w<DT>(xv[u].y, cv[u].y);
                pw.z = cmul_w<DT>(xv[u].z, cv[u].z);
                pw.w = cmul_w<DT>(xv[u].w, cv[u].w);
                *reinterpret_cast<uint4 *>(&buf0[g * RP + n]) = pw;
            }
        }
        fill();
        const int tstride = L / R0;
        // inputs unpadded: element n of row g at g RP + n
        auto ibase = [&](int g, int j) -> int { return g * RP + j; };
        auto iload = [&](int base, int tc, int th, bool ok) -> uint32_t {
            return buf0[base + (ok ? tc + th : R0 - 1) * tstride];
        };
        auto icook = [](int, int, int, bool, uint32_t w) -> uint32_t { return w; };
        mstage<DT, NTH, L, R0, G, NF, R0, PKT ? 3 : 2, true, true, false>(at0, ibase, iload, icook, obase, omul,
                                                                           oput, barrier);
    } else {
        // row passes (NF) read and write rows with unit element stride (a
        // compile-time constant: per-element offsets become immediates)
        const int igs = (int)a.in_gstride, ins = NF ? 1 : (int)a.in_nstride;
        const int tstride = (L / R0) * ins;   // input t of column j at j ins + t tstride
        Loader<LD_MIX> mx;
        if constexpr (LM == LD_MIX) mx.init(a, b, carr);
        long long xo = 0, co = 0;
        if constexpr (LM == LD_CPX) xo = (long long)b * a.in_bstride;
        if constexpr (LM == LD_PROD) {
            xo = (long long)((b / a.prod_xdiv) % a.prod_xmod) * a.in_bstride;
            co = (long long)(b / a.prod_cdiv) * a.cf_bstride;
        }
        __amdgpu_buffer_rsrc_t xin = rsrc(a.hin + xo);
        const __amdgpu_buffer_rsrc_t cin = rsrc(a.hcf + co);
        // rows past the end of the batch item (last group) read row 0
        auto ibase = [&](int g, int j) -> int { return (g0 + (g < gcount ? g : 0)) * igs + j * ins; };
        // unused inputs (ok = false): the IF byte of t = R0 - 1, or a masked
        // load (returns 0)
        auto iload = [&](int base, int tc, int th, bool ok) {
            if constexpr (LM == LD_MIX) {
                return mx.mix_byte(base + (ok ? tc + th : R0 - 1) * tstride);
            } else {
                const int bh = base + th * tstride, c = tc * tstride;
                if constexpr (LM == LD_CPX) return bload32_at(xin, bh, c, ok);
                else return make_uint2(bload32_at(xin, bh, c, ok), bload32_at(cin, bh, c, ok));
            }
        };
        auto icook = [&](int base, int tc, int th, bool ok, auto raw) -> uint32_t {
            if constexpr (LM == LD_MIX) {
                const int2 m = mx.mix_prod(base + (ok ? tc + th : R0 - 1) * tstride, raw);
                return pack_c<DT>((float)m.x, (float)m.y);
            } else if constexpr (LM == LD_CPX) {
                return raw;
            } else {
                return cmul_w<DT>(raw.x, raw.y);
            }
        };
        // complex-word inputs with few tiles per wave: every tile's loads
        // are issued at once (before the LDS tables are filled), not one
        // tile ahead (the mixing and product loaders hold more registers
        // per input)
        constexpr bool ALL0 = LM == LD_CPX && MT0 <= 6;
        mstage<DT, NTH, L, R0, G, NF, R0, PKT ? 3 : 2, true, ALL0, false>(at0, ibase, iload, icook, obase, omul,
                                                                           oput, fill);
        if constexpr (PB) {
            if (a.hin2) {                   // uniform over the grid
                run_last(std::integral_constant<int, 1>{});
                // the second block's item: its stage 0 stores to buf0 after
                // the barrier that follows every wave's reads of the first
                xin = rsrc(a.hin2 + xo);
                // (loads one tile ahead: the kept powers hold registers)
                mstage<DT, NTH, L, R0, G, NF, R0, PKT ? 3 : 2, true, false, false>(at0, ibase, iload, icook, obase,
                                                                                    omul, oput, barrier);
                run_last(std::integral_constant<int, 2>{});
                paired = true;
            }
        }
    }

    if (!paired) run_last(std::integral_constant<int, 0>{});
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
// Fused inverse of search cells (FP16).  One workgroup per cell (PRN, Doppler
// bin) runs, for every block k, the code-product row pass (M1 rows of length
// M2 = LR = RR0 RR1), the column pass (M2 columns of length M1 = LC = RC0 RC1)
// and the power accumulation, with the intermediate Z (M words) and the
// cell's power surface P (N <= M / 2 floats) resident in LDS: Z never goes
// to HBM, X and the code spectrum are read once per block (from L2: the
// workgroups of one XCD hold one bin at a time), and P is written once, on
// the last block, with the cell's peak partial.  Every stage is the mstage of
// the two-pass form with the same inputs, DFT tables, twiddles and roundings,
// so P is bit-identical to mpass_kernel's.
//
// LDS (32-bit words): Z[r][c] at r ZP + c + c / RR0 (a pad word after every
// RR0 words of a row, as the row pass's first stage stores them; ZP odd), one
// dummy word per lane for masked stores, P[lag], the row and column stage
// twiddles (packed FP16), the peak reduction records.  The in-place stages
// (all but the last column stage) read every input of their tiles before a
// barrier and then write (mstage ALL).
template <int LR, int RR0, int RR1, int LC, int RC0, int RC1, int GR, int GC, int NTH>
struct CellGeom {
    static constexpr int M1 = LC, M2 = LR, M = LC * LR;
    // row pitch: a multiple of 4 words (16-byte staging stores), and 16
    // banks apart from row to row, so that the last row stage's RR0
    // (= 16) lanes per row read disjoint banks
    static constexpr int ZP = (M2 + M2 / RR0 + 15) / 32 * 32 + 16;
    static constexpr int ZW = M1 * ZP;
    static constexpr int PO = ZW + 64;                    // P
    // words up to the peak records (8-byte aligned) for P of N floats
    static constexpr int tail(int N) { return (PO + N + LR + LC + 1) & ~1; }
    // the four DFT tables (16-byte aligned) after the peak records
    static constexpr int ATW[4] = {64 * ((2 * RR0 + 31) / 32) * ((2 * RR0 + 15) / 16),
                                   64 * ((2 * RR1 + 31) / 32) * ((2 * RR1 + 15) / 16),
                                   64 * ((2 * RC0 + 31) / 32) * ((2 * RC0 + 15) / 16),
                                   64 * ((2 * RC1 + 31) / 32) * ((2 * RC1 + 15) / 16)};
    static constexpr int ato(int N)
    {
        return (tail(N) + (int)(sizeof(PeakPart) * (NTH / 64) / 4) + 3) & ~3;
    }
    static constexpr size_t bytes(int N) { return (size_t)ato(N) * 4 + 16 * (size_t)(ATW[0] + ATW[1] + ATW[2] + ATW[3]); }
};

template <int LR, int RR0, int RR1, int LC, int RC0, int RC1, int GR, int GC, int NTH>
__global__ void __launch_bounds__(NTH) cell_kernel(CellArgs a, const uint4 *am_r0, const uint4 *am_r1,
                                                   const uint4 *am_c0, const uint4 *am_c1, const float2 *tw_r,
                                                   const float2 *tw_c, const uint32_t *tfullh)
{
    constexpr int DT = HDT_F16;
    using Geo = CellGeom<LR, RR0, RR1, LC, RC0, RC1, GR, GC, NTH>;
    constexpr int M1 = Geo::M1, M2 = Geo::M2, M = Geo::M, ZP = Geo::ZP, ZW = Geo::ZW;
    constexpr int THL = (RC1 + 1) / 2;        // last stage outputs kept (lags < M / 2)
    static_assert(M2 % GC == 0, "whole column groups");
    extern __shared__ uint32_t lds[];
    uint32_t *Z = lds;
    uint32_t *twr = lds + Geo::PO + a.N, *twc = twr + LR;
    PeakPart *red = reinterpret_cast<PeakPart *>(lds + Geo::tail(a.N));
    const int dummy = ZW + (int)(threadIdx.x & 63);
    const int N = a.N;
    // cell of this workgroup: round r = w / 8 over the 8 XCDs (w % 8); an
    // XCD runs the PRNs of one bin back to back (bins x, x + 8, ...)
    const int w = (int)blockIdx.x, r = w >> 3;
    const int prn = r % a.nprn, bin = (r / a.nprn) * 8 + (w & 7);
    if (bin >= a.nbin) return;

    // the stages' DFT tables in LDS (each stage loads its A operands from
    // there, keeping registers for the tiles)
    uint4 *atl = reinterpret_cast<uint4 *>(lds + Geo::ato(N));
    uint4 *at_r0 = atl, *at_r1 = at_r0 + Geo::ATW[0], *at_c0 = at_r1 + Geo::ATW[1], *at_c1 = at_c0 + Geo::ATW[2];
    for (int i = (int)threadIdx.x; i < Geo::ATW[0]; i += NTH) at_r0[i] = am_r0[i];
    for (int i = (int)threadIdx.x; i < Geo::ATW[1]; i += NTH) at_r1[i] = am_r1[i];
    for (int i = (int)threadIdx.x; i < Geo::ATW[2]; i += NTH) at_c0[i] = am_c0[i];
    for (int i = (int)threadIdx.x; i < Geo::ATW[3]; i += NTH) at_c1[i] = am_c1[i];
    // stage-0 twiddles as [t][j] = w_L^{t j} (t < R0, j < L / R0)
    for (int i = (int)threadIdx.x; i < LR; i += NTH) {
        const float2 v = tw_r[(i / (LR / RR0)) * (i % (LR / RR0))];
        twr[i] = pack_c<DT>(v.x, v.y);
    }
    for (int i = (int)threadIdx.x; i < LC; i += NTH) {
        const float2 v = tw_c[(i / (LC / RC0)) * (i % (LC / RC0))];
        twc[i] = pack_c<DT>(v.x, v.y);
    }

    const __amdgpu_buffer_rsrc_t cin = rsrc(a.C + (long long)prn * M);
    const __amdgpu_buffer_rsrc_t tfr = rsrc(tfullh);
    const __amdgpu_buffer_rsrc_t prr = rsrc(a.P + ((long long)prn * a.nbin + bin) * N);
    auto barrier = []() { __syncthreads(); };
    auto none = []() {};
    auto ident = [](int, int, int, bool, uint32_t v) -> uint32_t { return v; };
    Storer<ST_POW> sp;
    sp.nlag = N;
    sp.ibase = (long long)bin * N;

    // the X and Cf words of a row group (16-byte loads; rows past the end
    // read row g0), fetched one group ahead: the next group's loads are
    // issued once the current group's last stage has issued its own (so
    // that waiting for those does not wait for these)
    constexpr int Q = GR * LR / 4, NQ = (Q + NTH - 1) / NTH;
    uint4 xv[NQ], cv[NQ];
    auto fetch_group = [&](int kk, int gg0) {
        const __amdgpu_buffer_rsrc_t xin = rsrc(a.X + (long long)kk * a.xblk + (long long)bin * M);
        const int gc = min(GR, M1 - gg0);
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int qq = min((int)threadIdx.x + u * NTH, Q - 1);
            const int g = qq / (LR / 4), n = (qq - g * (LR / 4)) * 4;
            const int e = (gg0 + (g < gc ? g : 0)) * M2 + n;
            xv[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xin, e * 4, 0, 0));
            cv[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(cin, e * 4, 0, 0));
        }
    };
    fetch_group(0, 0);

    for (int k = 0; k < a.nblk; k++) {
        // ---- row pass: groups of GR rows ----
        for (int g0 = 0; g0 < M1; g0 += GR) {
            const int gcount = min(GR, M1 - g0);
            // the products X Cf of the group's rows, unpadded in their Z rows
            // (rows past the end are not stored)
#pragma unroll
            for (int u = 0; u < NQ; u++) {
                const int q0 = (int)threadIdx.x + u * NTH;
                const int g = q0 / (LR / 4), n = (q0 - g * (LR / 4)) * 4;
                if (q0 < Q && g < gcount) {
                    const uint32_t xw[4] = {xv[u].x, xv[u].y, xv[u].z, xv[u].w};
                    const uint32_t cw[4] = {cv[u].x, cv[u].y, cv[u].z, cv[u].w};
                    uint4 pw;
                    pw.x = cmul_w<DT>(xw[0], cw[0]);
                    pw.y = cmul_w<DT>(xw[1], cw[1]);
                    pw.z = cmul_w<DT>(xw[2], cw[2]);
                    pw.w = cmul_w<DT>(xw[3], cw[3]);
                    *reinterpret_cast<uint4 *>(Z + (g0 + g) * ZP + n) = pw;
                }
            }
            __syncthreads();
            // stage 0 (radix RR0): input t of column j at n = j + t LR / RR0,
            // output t at q = j RR0 + t (padded: j (RR0 + 1) + t), next
            // twiddle w_LR^{t j} = twr[t LR / RR0 + j]
            auto rb0 = [&](int g, int j) -> int { return (g0 + (g < gcount ? g : 0)) * ZP + j; };
            auto rl0 = [&](int base, int tc, int th, bool ok) -> uint32_t {
                return Z[base + (ok ? tc + th : RR0 - 1) * (LR / RR0)];
            };
            // (rows past the end: the clamped row's outputs again, to its place)
            auto ro0 = [&](int g, int j) -> int4 {
                return make_int4((g0 + (g < gcount ? g : 0)) * ZP + j * (RR0 + 1), 0, j, 1);
            };
            auto rm0 = [&](const int4 &ob, int tc, int th, bool ok) -> uint32_t {
                return twr[(ok ? tc + th : 0) * (LR / RR0) + ob.z];
            };
            auto rp0 = [&](const int4 &ob, int tc, int th, float2 v, uint32_t f, bool ok) {
                Z[ok ? ob.x + tc + th : dummy] = cmul_w<DT>(pack_c<DT>(v.x, v.y), f);
            };
            ATab<DT, RR0, RR0> ar0;
            ar0.load(at_r0);
            mstage<DT, NTH, LR, RR0, GR, true, RR0, 3, true, true, false>(ar0, rb0, rl0, ident, ro0, rm0, rp0,
                                                                          barrier);
            __syncthreads();
            // stage 1 (radix RR1): input t of column j (< RR0) at element
            // j + t RR0 (padded j + t (RR0 + 1)); output t is element
            // k = j + t RR0 of the row, times the inter-pass twiddle
            // tfullh[row M2 + k], in place
            auto rb1 = [&](int g, int j) -> int { return (g0 + (g < gcount ? g : 0)) * ZP + j; };
            auto rl1 = [&](int base, int tc, int th, bool ok) -> uint32_t {
                return Z[base + (ok ? tc + th : RR1 - 1) * (RR0 + 1)];
            };
            auto ro1 = [&](int g, int j) -> int4 {
                const int gg = g0 + (g < gcount ? g : 0);
                return make_int4(gg * ZP + j, gg * M2 + j, 0, 1);
            };
            auto rm1 = [&](const int4 &ob, int tc, int th, bool ok) -> uint32_t {
                return bload32_at(tfr, ob.y + th * RR0, tc * RR0, ok);
            };
            auto rp1 = [&](const int4 &ob, int tc, int th, float2 v, uint32_t f, bool ok) {
                Z[ok ? ob.x + (tc + th) * (RR0 + 1) : dummy] = cmul_w<DT>(pack_c<DT>(v.x, v.y), f);
            };
            ATab<DT, RR1, RR1> ar1;
            ar1.load(at_r1);
            // after this stage's loads: the next group's (or the next
            // block's first group's) X and Cf words, then the barrier
            auto pre1 = [&]() {
                if (g0 + GR < M1) fetch_group(k, g0 + GR);
                else if (k + 1 < a.nblk) fetch_group(k + 1, 0);
                __syncthreads();
            };
            mstage<DT, NTH, LR, RR1, GR, true, RR1, 3, true, true, false, true>(ar1, rb1, rl1, ident, ro1, rm1,
                                                                                rp1, pre1);
        }
        __syncthreads();
        // ---- column pass: groups of GC columns ----
        sp.pinit = k == 0;
        sp.track = k == a.nblk - 1;
        const bool last = k == a.nblk - 1;
        for (int c0 = 0; c0 < M2; c0 += GC) {
            // Z column c (sub-transform g = c - c0), element (row) n
            auto zc = [&](int g) -> int { return c0 + g + (c0 + g) / RR0; };
            // stage 0 (radix RC0): input t of column j at row j + t LC / RC0,
            // output t at row j RC0 + t, next twiddle w_LC^{t j}
            auto cb0 = [&](int g, int j) -> int { return j * ZP + zc(g); };
            auto cl0 = [&](int base, int tc, int th, bool ok) -> uint32_t {
                return Z[base + (ok ? tc + th : RC0 - 1) * (LC / RC0) * ZP];
            };
            auto co0 = [&](int g, int j) -> int4 { return make_int4(j * RC0 * ZP + zc(g), 0, j, 1); };
            auto cm0 = [&](const int4 &ob, int tc, int th, bool ok) -> uint32_t {
                return twc[(ok ? tc + th : 0) * (LC / RC0) + ob.z];
            };
            auto cp0 = [&](const int4 &ob, int tc, int th, float2 v, uint32_t f, bool ok) {
                Z[ok ? ob.x + (tc + th) * ZP : dummy] = cmul_w<DT>(pack_c<DT>(v.x, v.y), f);
            };
            ATab<DT, RC0, RC0> ac0;
            ac0.load(at_c0);
            mstage<DT, NTH, LC, RC0, GC, false, RC0, 3, true, true, false>(ac0, cb0, cl0, ident, co0, cm0, cp0,
                                                                           barrier);
            __syncthreads();
            // stage 1 (radix RC1, last): input t of column j (< RC0) at row
            // j + t RC0; output t' < THL is element k = j + t' RC0 of the
            // column: lag c + M2 k, P += kappa^2 |v|^2 for lags < N (in LDS;
            // on the last block to global P, with the peak tracking)
            auto cb1 = [&](int g, int j) -> int { return j * ZP + zc(g); };
            auto cl1 = [&](int base, int tc, int th, bool ok) -> uint32_t {
                return Z[base + (ok ? tc + th : RC1 - 1) * RC0 * ZP];
            };
            auto co1 = [&](int g, int j) -> int4 { return make_int4(c0 + g + M2 * j, 0, 0, 1); };
            auto cm1 = [](const int4 &, int, int, bool) -> int { return 0; };
            ATab<DT, RC1, THL> ac1;
            ac1.load(at_c1);
            if (last) {
                // P to global memory, with the peak tracking
                auto cp1 = [&](const int4 &ob, int tc, int th, float2 v, int, bool ok) {
                    const int e = ob.x + (tc + th) * RC0 * M2;
                    const float pold = __uint_as_float(lds[Geo::PO + (e < N ? e : 0)]);
                    sp.store_pow_sel(e, v, a.kap2, pold, ok, [&](bool ok2, int e2, float nv) {
                        bstore32(prr, e2, __float_as_uint(nv), ok2);
                    });
                };
                mstage<DT, NTH, LC, RC1, GC, false, THL, 0, true, true>(ac1, cb1, cl1, ident, co1, cm1, cp1, none);
            } else {
                // P in LDS (the same arithmetic as store_pow_sel)
                auto cp1 = [&](const int4 &ob, int tc, int th, float2 v, int, bool ok) {
                    const int e = ob.x + (tc + th) * RC0 * M2;
                    const bool in = ok && e < N;
                    const float pw = a.kap2 * __builtin_fmaf(v.x, v.x, v.y * v.y);
                    const float pold = __uint_as_float(lds[Geo::PO + (in ? e : 0)]);
                    lds[in ? Geo::PO + e : dummy] = __float_as_uint(k == 0 ? pw : pold + pw);
                };
                mstage<DT, NTH, LC, RC1, GC, false, THL, 0, true, true>(ac1, cb1, cl1, ident, co1, cm1, cp1, none);
            }
        }
        __syncthreads();
    }
    sp.flush_peak(a.part, (long long)prn * a.nbin + bin, red);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// (L, R0, R1, R2, G, threads, variant): column-pass lengths (G columns =
// 128 bytes of words per row segment; the 16-column shapes of 960 and 1000
// fit two workgroups per CU, their workgroups paired on one XCD) and
// row-pass lengths (G rows, 4-8 K points per workgroup, G L / R a multiple
// of 32 where possible).  1000 runs as 25 x 40 (two stages, 20 or 40 outputs
// per lane and tile) or as 10 x 10 x 10.  Each length has two workgroup shapes; the
// plan's tuner (SDR_SEARCH_TUNE) times both on the real shapes, variant 0 is
// the default.  Splits: 24000 = 150 x 160, 48000 = 150 x 320 or 200 x 240,
// 96000 = 300 x 320 or 200 x 480, 192000 = 400 x 480, 960000 = 960 x 1000
// or 1000 x 960.
#define PSA_HCOL(X) \
    X(150, 10, 15, 1, 32, 256, 0) X(150, 10, 15, 1, 32, 512, 1) \
    X(200, 10, 20, 1, 32, 512, 0) X(200, 10, 20, 1, 32, 256, 1) \
    X(200, 8, 25, 1, 32, 256, 2) X(200, 5, 40, 1, 32, 256, 3) \
    X(300, 15, 20, 1, 32, 512, 0) X(300, 15, 20, 1, 32, 256, 1) \
    X(300, 10, 30, 1, 32, 512, 2) X(300, 12, 25, 1, 32, 512, 3) \
    X(400, 16, 25, 1, 32, 512, 0) X(400, 16, 25, 1, 32, 256, 1) \
    X(400, 10, 40, 1, 32, 512, 2) X(400, 20, 20, 1, 32, 512, 3) \
    X(960, 32, 30, 1, 32, 512, 0) X(960, 32, 30, 1, 16, 256, 1) \
    X(1000, 25, 40, 1, 32, 512, 0) X(1000, 25, 40, 1, 16, 512, 1)
#define PSA_HROW(X) \
    X(160, 16, 10, 1, 32, 256, 0) X(160, 16, 10, 1, 32, 512, 1) \
    X(160, 10, 16, 1, 32, 256, 2) X(160, 20, 8, 1, 32, 256, 3) \
    X(240, 16, 15, 1, 32, 256, 0) X(240, 16, 15, 1, 32, 512, 1) \
    X(240, 8, 30, 1, 32, 256, 2) X(240, 15, 16, 1, 32, 256, 3) \
    X(320, 16, 20, 1, 16, 256, 0) X(320, 16, 20, 1, 32, 512, 1) \
    X(320, 32, 10, 1, 16, 256, 2) X(320, 10, 32, 1, 16, 256, 3) \
    X(480, 16, 30, 1, 16, 512, 0) X(480, 16, 30, 1, 16, 256, 1) \
    X(480, 32, 15, 1, 16, 256, 2) X(480, 12, 40, 1, 16, 256, 3) \
    X(1000, 10, 10, 10, 4, 256, 0) X(1000, 25, 40, 1, 8, 512, 1) \
    X(960, 32, 30, 1, 8, 256, 0) X(960, 32, 30, 1, 16, 512, 1) \
    X(960, 30, 32, 1, 8, 256, 2) X(960, 32, 30, 1, 4, 256, 3)

struct HSpec { int L, r[3], G, nth, var; };
static const HSpec h_col_specs[] = {
#define PSA_S(L_, A_, B_, C_, G_, T_, V_) {L_, {A_, B_, C_}, G_, T_, V_},
    PSA_HCOL(PSA_S)
};
static const HSpec h_row_specs[] = {
    PSA_HROW(PSA_S)
#undef PSA_S
};

static const HSpec *find_spec(const HSpec *t, int n, int L, int var)
{
    for (int i = 0; i < n; i++) if (t[i].L == L && t[i].var == var) return &t[i];
    return nullptr;
}
static const HSpec *col_spec(int L, int var = 0)
{
    return find_spec(h_col_specs, (int)(sizeof(h_col_specs) / sizeof(HSpec)), L, var);
}
static const HSpec *row_spec(int L, int var = 0)
{
    return find_spec(h_row_specs, (int)(sizeof(h_row_specs) / sizeof(HSpec)), L, var);
}

extern "C" int sdr_search_pass_shape(int col, int L, int var, int *shape)
{
    const psa::HSpec *h = col ? psa::col_spec(L, var) : psa::row_spec(L, var);
    if (!h || !shape) return 0;
    for (int i = 0; i < 3; i++) shape[i] = h->r[i];
    shape[3] = h->G;
    shape[4] = h->nth;
    return 1;
}

bool h_col_pairs(int dt, int L, int var)
{
    // the kernel's rule (mpass_kernel PB)
    const HSpec *h = col_spec(L, var);
    return dt == HDT_F16 && h && pb_shape(L, h->r[0], h->r[1], h->r[2], h->G, h->nth);
}

bool h_any_pairs(int dt, int M)
{
    for (const HSpec &cs : h_col_specs)
        if (M % cs.L == 0 && row_spec(M / cs.L, 0) && h_col_pairs(dt, cs.L, cs.var)) return true;
    return false;
}

bool h_shape(int M, int M1, int vc, int vr)
{
    return M1 > 0 && M % M1 == 0 && col_spec(M1, vc) && row_spec(M / M1, vr);
}

int h_splits(int M, int *m1, int nmax)
{
    // the default first: the smallest largest radix (fewest padded MFMA rows
    // and registers), then the shorter column pass
    std::vector<std::pair<int, int>> c;
    for (const HSpec &cs : h_col_specs) {
        if (cs.var != 0) continue;
        const HSpec *r = M % cs.L == 0 ? row_spec(M / cs.L) : nullptr;
        if (!r) continue;
        int mr = 0;
        for (int i = 0; i < 3; i++) mr = std::max(mr, std::max(cs.r[i], r->r[i]));
        c.push_back({mr * 4096 + cs.L, cs.L});
    }
    std::sort(c.begin(), c.end());
    int n = 0;
    for (auto &x : c) if (n < nmax) m1[n++] = x.second;
    return n;
}

// A-operand table of the R-point DFT (BF16: hi part, then lo part), cached
// per (R, dir, dt); amat_radix returns the R of a cached table (0: none)
static std::mutex amat_mtx;
static std::map<int, uint4 *> amat_cache;
static int amat_radix(const uint4 *p)
{
    std::lock_guard<std::mutex> lk(amat_mtx);
    for (const auto &kv : amat_cache) if (kv.second == p) return kv.first / 4;
    return 0;
}
static const uint4 *make_amat(int R, int dir, int dt)
{
    std::lock_guard<std::mutex> lk(amat_mtx);
    const int key = (R * 2 + dir) * 2 + dt;
    auto it = amat_cache.find(key);
    if (it != amat_cache.end()) return it->second;
    const int KS = (2 * R + 15) / 16, RT = (2 * R + 31) / 32;
    const int n = 64 * KS * RT;         // uint4 per part
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
    amat_cache[key] = d;
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

// FP16 (re, im) words of a float2 table, round to nearest even
__global__ void pack_f16_kernel(const float2 *__restrict__ in, uint32_t *__restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pack_c<HDT_F16>(in[i].x, in[i].y);
}
static uint32_t *pack_table_f16(const float2 *d, int n)
{
    uint32_t *o;
    PSA_CHECK(hipMalloc(&o, sizeof(uint32_t) * n));
    hipLaunchKernelGGL(pack_f16_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d, o, n);
    PSA_CHECK(hipGetLastError());
    PSA_CHECK(hipDeviceSynchronize());
    return o;
}

static float pow2_near(double x)    // 2^round(log2 x)
{
    return (float)ldexp(1.0, (int)lround(log2(x)));
}

const HPlan *hfft_plan(int M, int dt, int M1)
{
    static std::mutex mtx;
    static std::map<std::vector<int>, HPlan *> plans;
    if (dt != HDT_F16 && dt != HDT_BF16) return nullptr;
    int cand[8];
    const int nc = h_splits(M, cand, 8);
    if (nc == 0) return nullptr;
    if (M1 == 0) M1 = cand[0];
    if (std::find(cand, cand + nc, M1) == cand + nc) return nullptr;
    const int M2 = M / M1;
    ensure_init();
    std::lock_guard<std::mutex> lk(mtx);
    const std::vector<int> key = {M, dt, M1};
    auto it = plans.find(key);
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
        if (dir == 1 && dt == HDT_F16) p->tfullh = pack_table_f16(p->tfull[dir], M);
        // the two workgroup shapes of a length may split it into different
        // radices: each shape has its own tables
        for (int v = 0; v < PSA_HVARS; v++) {
            const HSpec *cs = col_spec(M1, v), *rs = row_spec(M2, v);
            for (int s = 0; s < 3; s++) {
                if (cs) p->am1[v][dir][s] = cs->r[s] > 1 ? make_amat(cs->r[s], dir, dt) : nullptr;
                if (rs) p->am2[v][dir][s] = rs->r[s] > 1 ? make_amat(rs->r[s], dir, dt) : nullptr;
            }
        }
    }
    plans[key] = p;
    return p;
}

void h_col_args(const HPlan *hp, int dir, PassArgs &a, int var)
{
    const HSpec *cs = col_spec(hp->M1, var);
    if (!cs) PSA_FAIL("MFMA pass: no column shape %d for length %d", var, hp->M1);
    a.L = hp->M1;
    a.nsub = hp->M2;
    a.G = cs->G;
    a.hvar = var;
    a.in_nstride = hp->M2; a.in_gstride = 1;
    a.out_nstride = hp->M2; a.out_gstride = 1;
    a.dir = dir;
    a.tw = hp->tw1[dir];
    a.M = hp->M;
    a.thi = a.tlo = nullptr;
    a.tfull = nullptr;
    a.osc = 1.0f;
    a.kappa = 1.0f;
    for (int s = 0; s < 3; s++) a.amat[s] = hp->am1[var][dir][s];
}

void h_row_args(const HPlan *hp, int dir, PassArgs &a, int var)
{
    const HSpec *rs = row_spec(hp->M2, var);
    if (!rs) PSA_FAIL("MFMA pass: no row shape %d for length %d", var, hp->M2);
    a.L = hp->M2;
    a.nsub = hp->M1;
    a.G = rs->G;
    a.hvar = var;
    a.in_nstride = 1; a.in_gstride = hp->M2;
    a.out_nstride = 1; a.out_gstride = hp->M2;
    a.dir = dir;
    a.tw = hp->tw2[dir];
    a.M = hp->M;
    a.thi = a.tlo = nullptr;
    a.tfull = nullptr;
    a.osc = 1.0f;
    a.kappa = 1.0f;
    for (int s = 0; s < 3; s++) a.amat[s] = hp->am2[var][dir][s];
}

template <int LM, int SM, int DT, int L, int R0, int R1, int R2, int G, bool NF, int NTH>
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
    if (NF && (a.in_nstride != 1 || a.out_nstride != 1)) PSA_FAIL("MFMA row pass needs unit element strides");
    if (amat_radix(a.amat[0]) != R0 || (R1 > 1 && amat_radix(a.amat[1]) != R1) ||
        (R2 > 1 && amat_radix(a.amat[2]) != R2))
        PSA_FAIL("MFMA pass: DFT tables do not match length %d (%d x %d x %d)", L, R0, R1, R2);
    pass_grid = nb * NTH;
    hipLaunchKernelGGL((mpass_kernel<LM, SM, DT, L, R0, R1, R2, G, NF, NTH>), dim3((unsigned)nb), dim3(NTH), 0,
                       st, a, ngroups);
    PSA_CHECK(hipGetLastError());
}

template <int DT>
static void hlaunch_dt(LoadMode lm, StoreMode sm, const PassArgs &a, int nbatch, hipStream_t st)
{
    // column lengths: forward (mixing) and inverse (power) column passes
#define PSA_C(L_, A_, B_, C_, G_, T_, V_)                                                      \
    if (a.L == L_ && a.hvar == V_ && a.in_nstride != 1) {                                  \
        if (lm == LD_MIX && sm == ST_CPX) { hlaunch_t<LD_MIX, ST_CPX, DT, L_, A_, B_, C_, G_, false, T_>(a, nbatch, st); return; } \
        if (lm == LD_CPX && sm == ST_POW) { hlaunch_t<LD_CPX, ST_POW, DT, L_, A_, B_, C_, G_, false, T_>(a, nbatch, st); return; } \
    }
    PSA_HCOL(PSA_C)
#undef PSA_C
    // row lengths: forward row pass and inverse (code product) row pass
#define PSA_R(L_, A_, B_, C_, G_, T_, V_)                                                      \
    if (a.L == L_ && a.hvar == V_ && a.in_nstride == 1) {                                  \
        if (lm == LD_CPX && sm == ST_CPX) { hlaunch_t<LD_CPX, ST_CPX, DT, L_, A_, B_, C_, G_, true, T_>(a, nbatch, st); return; } \
        if (lm == LD_PROD && sm == ST_CPX) { hlaunch_t<LD_PROD, ST_CPX, DT, L_, A_, B_, C_, G_, true, T_>(a, nbatch, st); return; } \
    }
    PSA_HROW(PSA_R)
#undef PSA_R
    PSA_FAIL("no MFMA pass for length %d, mode %d/%d", a.L, (int)lm, (int)sm);
}

void launch_hpass(int dt, LoadMode lm, StoreMode sm, const PassArgs &a_in, int nbatch, hipStream_t st)
{
    // the forward column and inverse row passes take the inter-pass twiddle,
    // the forward row pass does not
    if (sm == ST_CPX && (lm != LD_CPX) != (a_in.tfull != nullptr)) PSA_FAIL("MFMA pass: twiddle table mismatch");
    if (dt == HDT_F16 && lm == LD_PROD && !a_in.tfullh) PSA_FAIL("MFMA pass: FP16 twiddle table missing");
    if (a_in.hin2 && (sm != ST_POW || !h_col_pairs(dt, a_in.L, a_in.hvar)))
        PSA_FAIL("MFMA pass: length %d shape %d takes no paired blocks", a_in.L, a_in.hvar);
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

namespace psa {

// cell kernels: (M1 = LC, RC0, RC1, M2 = LR, RR0, RR1, GR, GC, threads) with
// the radices of the two-pass shapes of those lengths (variant 0), so that
// both forms use the same DFT tables
#define PSA_CELL(X) X(150, 10, 15, 160, 16, 10, 50, 80, 1024)

int cell_split(int M)
{
#define PSA_CS(LC_, C0_, C1_, LR_, R0_, R1_, GR_, GC_, T_) if (M == LC_ * LR_) return LC_;
    PSA_CELL(PSA_CS)
#undef PSA_CS
    return 0;
}

// prepare: only raise the kernel's dynamic LDS limit (once; called at plan
// set-up, outside any stream capture)
template <int LR, int RR0, int RR1, int LC, int RC0, int RC1, int GR, int GC, int NTH>
static void cell_launch_t(const HPlan *hp, const CellArgs &a, hipStream_t st, bool prepare)
{
    using Geo = CellGeom<LR, RR0, RR1, LC, RC0, RC1, GR, GC, NTH>;
    auto kern = cell_kernel<LR, RR0, RR1, LC, RC0, RC1, GR, GC, NTH>;
    static std::once_flag once;
    std::call_once(once, [&]() {
        PSA_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)Geo::bytes(Geo::M / 2)));
    });
    if (prepare) return;
    const size_t lds = Geo::bytes(a.N);
    if (2 * a.N > hp->M || lds > 160 * 1024 || Geo::bytes(Geo::M / 2) > 160 * 1024)
        PSA_FAIL("cell kernel: N = %d does not fit (M = %d, %zu bytes of LDS)", a.N, hp->M, lds);
    const HSpec *cs = col_spec(LC, 0), *rs = row_spec(LR, 0);
    if (!cs || !rs || cs->r[0] != RC0 || cs->r[1] != RC1 || rs->r[0] != RR0 || rs->r[1] != RR1)
        PSA_FAIL("cell kernel: radices differ from the pass shapes of %d x %d", LC, LR);
    if (amat_radix(hp->am2[0][1][0]) != RR0 || amat_radix(hp->am2[0][1][1]) != RR1 ||
        amat_radix(hp->am1[0][1][0]) != RC0 || amat_radix(hp->am1[0][1][1]) != RC1)
        PSA_FAIL("cell kernel: DFT tables do not match %d x %d", LC, LR);
    const long long nb = (long long)(a.nbin + 7) / 8 * 8 * a.nprn;
    if (nb <= 0) return;
    if (nb > 0x7FFFFFFFLL) PSA_FAIL("cell kernel launch too large (%lld workgroups)", nb);
    pass_grid = nb * NTH;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(NTH), lds, st, a, hp->am2[0][1][0], hp->am2[0][1][1],
                       hp->am1[0][1][0], hp->am1[0][1][1], hp->tw2[1], hp->tw1[1], hp->tfullh);
    PSA_CHECK(hipGetLastError());
}

void launch_cells(const HPlan *hp, const CellArgs &a, hipStream_t st, bool prepare)
{
    if (hp->dt != HDT_F16 || !hp->tfullh) PSA_FAIL("cell kernel: FP16 plans only");
#define PSA_CL(LC_, C0_, C1_, LR_, R0_, R1_, GR_, GC_, T_)                                      \
    if (hp->M1 == LC_ && hp->M2 == LR_) {                                                    \
        cell_launch_t<LR_, R0_, R1_, LC_, C0_, C1_, GR_, GC_, T_>(hp, a, st, prepare);      \
        return;                                                                              \
    }
    PSA_CELL(PSA_CL)
#undef PSA_CL
    PSA_FAIL("no cell kernel for %d x %d", hp->M1, hp->M2);
}

}  // namespace psa
