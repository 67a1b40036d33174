// psa_corr.hip - standard correlators (src/sdr_func.c:861-1072) and the
// batched open-loop tracking plan (src/sdr_ch.c:674-740).
//
// Bit-exactness with the reference AVX2 build: each half-correlation
// (corr1 = IQ[0:j] . bank_k[N-j:N], corr2 = IQ[j:N] . bank_k[0:N-j]) is
//   float(S_vec) + t_0 + t_1 + ... (sequential float adds)
// where S_vec is the int32 sum over the first 16*floor(len/16) samples with
// _mm256_sign_epi8 semantics (x * sign(c), -(-128) stays -128) and t_i are the
// remaining samples' true int products (src/sdr_func.c:872-910).  Integer sums
// are order independent, so any GPU reduction order reproduces S_vec exactly.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <string>
#include "psa_loop.h"

namespace psa {

__device__ __forceinline__ int sgn8(int x, int c)
{
    if (c < 0) return (int)(int8_t)(uint8_t)(-x);
    return c == 0 ? 0 : x;
}

// aI += sum over NW words of x_I . c_I (4 int8 lanes each), aQ likewise;
// EXACT: _mm256_sign_epi8 products
template <bool EXACT, int NW = 4>
__device__ __forceinline__ void dot4x4(const uint32_t *xi, const uint32_t *xq, const uint32_t *ci,
                                       const uint32_t *cq, int &aI, int &aQ)
{
#pragma unroll
    for (int v = 0; v < NW; v++) {
        if (EXACT) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                aI += sgn8((int8_t)(xi[v] >> (8 * b)), (int8_t)(ci[v] >> (8 * b)));
                aQ += sgn8((int8_t)(xq[v] >> (8 * b)), (int8_t)(cq[v] >> (8 * b)));
            }
        } else {
            aI = __builtin_amdgcn_sdot4((int)xi[v], (int)ci[v], aI, false);
            aQ = __builtin_amdgcn_sdot4((int)xq[v], (int)cq[v], aQ, false);
        }
    }
}

// 16-byte load from a 4-byte aligned address
__device__ __forceinline__ uint4 load_u4(const uint32_t *p)
{
    uint4 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
}

// v + the value of lane (lane ^ 1), (lane ^ 2), 7 - lane or 15 - lane within
// a row of 16 (DPP quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror); the whole wave must be active
template <int CTRL>
__device__ __forceinline__ int dpp_add(int v)
{
    return v + __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

// Sums of four int32 values over the 64 lanes of a (fully active) wave:
// a0 .. a3's totals end in every lane of rows 0 .. 3 (lanes 16 c .. 16 c + 15).
// v_permlane32_swap trades the upper half of its first operand for the lower
// half of its second, so the two outputs' sum holds a0 over both halves in
// lanes 0-31 and a2 in lanes 32-63; v_permlane16_swap does the same with
// rows 1, 3 of the first and rows 0, 2 of the second; four DPP adds then sum
// each row.  Integer sums: any order gives the same result.
__device__ __forceinline__ int wave_sum4_rows(int a0, int a1, int a2, int a3)
{
    const auto s02 = __builtin_amdgcn_permlane32_swap((unsigned)a0, (unsigned)a2, false, false);
    const auto s13 = __builtin_amdgcn_permlane32_swap((unsigned)a1, (unsigned)a3, false, false);
    const int k0 = (int)(s02[0] + s02[1]), k1 = (int)(s13[0] + s13[1]);
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)k0, (unsigned)k1, false, false);
    int v = (int)(r[0] + r[1]);
    v = dpp_add<0xB1>(v);
    v = dpp_add<0x4E>(v);
    v = dpp_add<0x141>(v);
    return dpp_add<0x140>(v);
}

// the four totals of wave_sum4_rows as one int4 (uniform)
__device__ __forceinline__ int4 wave_sum4(int a0, int a1, int a2, int a3)
{
    const int v = wave_sum4_rows(a0, a1, a2, a3);
    return make_int4(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16),
                     __builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48));
}

// block-wide sum of 4 ints (256 threads); result valid in all threads
__device__ __forceinline__ int4 block_sum4(int4 v, int4 *red)
{
    v = wave_sum4(v.x, v.y, v.z, v.w);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    int4 r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
        r.x += red[i].x; r.y += red[i].y; r.z += red[i].z; r.w += red[i].w;
    }
    return r;
}

// Tail of one half: samples [v, len) of data d against code c, added one by
// one in float to float(S_vec) (scalar loop of src/sdr_func.c:906-909 /
// 951-954), then scaled by (1.0f * SDR_CSCALE).
template <bool CPX>
__device__ void half_epilogue(int svecA, int svecB, const char2 *d, const char2 *c, int v, int len,
                              float *outA, float *outB)
{
    float a = (float)svecA, bq = (float)svecB;
    for (int i = v; i < len; i++) {
        int I = d[i].x, Q = d[i].y, cI = c[i].x, cQ = c[i].y;
        if (CPX) {
            a += (float)(I * cI + Q * cQ);
            bq += (float)(Q * cI - I * cQ);
        } else {
            a += (float)(I * cI);
            bq += (float)(Q * cQ);
        }
    }
    const float sc = 1.0f * PSA_CSCALE;
    *outA = a * sc;
    *outB = bq * sc;
}

// One workgroup per tap: vector-part sums of both halves, then the float
// epilogue of the tap.  corr12[t] = {corr1.re, corr1.im, corr2.re, corr2.im}.
template <bool CPX>
__global__ void __launch_bounds__(256) corr_tap_kernel(const char2 *__restrict__ IQ,
                                                       const char2 *__restrict__ code, int N,
                                                       const int2 *__restrict__ jk,
                                                       float4 *__restrict__ corr12)
{
    __shared__ int4 red[4];
    const int t = blockIdx.x;
    const int j = jk[t].x, k = jk[t].y;
    const char2 *bank = code + (long long)k * N;
    const int V1 = j >= 16 ? j / 16 * 16 : 0;
    const int len2 = N - j;
    const int V2 = len2 >= 16 ? len2 / 16 * 16 : 0;
    int4 s = make_int4(0, 0, 0, 0);
    // half 1: IQ[m] with bank[N - j + m], m < V1
    for (int m = threadIdx.x; m < V1; m += blockDim.x) {
        char2 d = IQ[m], c = bank[N - j + m];
        if (CPX) {
            s.x += sgn8(d.x, c.x) + sgn8(d.y, c.y);
            s.y += sgn8(d.y, c.x) + sgn8(d.x, (int)(int8_t)(uint8_t)(-c.y));
        } else {
            s.x += sgn8(d.x, c.x);
            s.y += sgn8(d.y, c.y);
        }
    }
    // half 2: IQ[j + r] with bank[r], r < V2
    for (int r = threadIdx.x; r < V2; r += blockDim.x) {
        char2 d = IQ[j + r], c = bank[r];
        if (CPX) {
            s.z += sgn8(d.x, c.x) + sgn8(d.y, c.y);
            s.w += sgn8(d.y, c.x) + sgn8(d.x, (int)(int8_t)(uint8_t)(-c.y));
        } else {
            s.z += sgn8(d.x, c.x);
            s.w += sgn8(d.y, c.y);
        }
    }
    int4 tot = block_sum4(s, red);
    if (threadIdx.x == 0) {
        float4 o;
        half_epilogue<CPX>(tot.x, tot.y, IQ, bank + N - j, V1, j, &o.x, &o.y);
        half_epilogue<CPX>(tot.z, tot.w, IQ + j, bank, V2, len2, &o.z, &o.w);
        corr12[t] = o;
    }
}

// Polarity and output (src/sdr_func.c:995-1007): the reference build
// contracts the EPL sum to d += fmaf(c1.re, c2.re, c1.im * c2.im).
__global__ void corr_final_kernel(const float4 *__restrict__ corr12, int n, int N,
                                  float2 *__restrict__ corr, float2 *__restrict__ C)
{
    float dot = 0.0f;
    for (int i = 0; i < 3 && i < n; i++) {
        float4 c = corr12[i];
        dot = __fadd_rn(dot, __fmaf_rn(c.x, c.z, __fmul_rn(c.y, c.w)));
    }
    const float sign = dot < 0.0f ? -1.0f : 1.0f;
    const float fN = (float)N;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        float4 c = corr12[i];
        corr[i] = make_float2(__fdiv_rn(__fadd_rn(c.x, __fmul_rn(sign, c.z)), fN),
                              __fdiv_rn(__fadd_rn(c.y, __fmul_rn(sign, c.w)), fN));
    }
    if (threadIdx.x == 0 && n > 0) {
        float4 c = corr12[0];
        C[0] = make_float2(c.x, c.y);
        C[1] = make_float2(c.z, c.w);
    }
}

// tap geometry, src/sdr_func.c:988-991 (double arithmetic as the reference)
static void tap_geometry(int N, double coff, const double *pos, int n, int2 *jk)
{
    for (int i = 0; i < n; i++) {
        double p = coff + pos[i];
        int j = (int)floor(p), k = (int)((p - j) * PSA_N_CODES);
        if (j < 0) j += N; else if (j >= N) j -= N;
        if (j < 0 || j > N || k < 0 || k >= PSA_N_CODES) {
            PSA_FAIL("corr_std: tap %d position %.3f outside one code period (N=%d)", i, p, N);
        }
        jk[i] = make_int2(j, k);
    }
}

// Per-thread cache of device code banks (the per-tap path): keyed on the
// bank's address, size and a 64-entry fingerprint, as the reference's banks
// are built once per channel and never modified (src/sdr_ch.c:266-272).
static uint64_t bank_fp(const uint8_t *p, size_t bytes)
{
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 64; i++) {
        const size_t o = bytes * (size_t)i / 64;
        h = (h ^ p[o]) * 1099511628211ull;
        h = (h ^ p[o + 1 < bytes ? o + 1 : o]) * 1099511628211ull;
    }
    return h;
}

struct BankCache {
    struct Ent {
        const void *host = nullptr;
        size_t bytes = 0;
        uint64_t fp = 0;
        void *dev = nullptr;
        size_t cap = 0;
        unsigned long long used = 0;
    } e[4];
    unsigned long long tick = 0;
    ~BankCache()
    {
        for (auto &x : e) if (x.dev) (void)hipFree(x.dev);
    }
    const void *get(const void *host, size_t bytes, hipStream_t st)
    {
        tick++;
        const uint64_t fp = bank_fp((const uint8_t *)host, bytes);
        for (auto &x : e) {
            if (x.host == host && x.bytes == bytes && x.dev && x.fp == fp) {
                x.used = tick;
                return x.dev;
            }
        }
        Ent *v = &e[0];
        for (auto &x : e) if (x.used < v->used) v = &x;
        if (v->cap < bytes || !v->dev) {
            if (v->dev) PSA_CHECK(hipFree(v->dev));
            PSA_CHECK(hipMalloc(&v->dev, bytes));
            v->cap = bytes;
        }
        v->host = host;
        v->bytes = bytes;
        v->fp = fp;
        v->used = tick;
        PSA_CHECK(hipMemcpyAsync(v->dev, host, bytes, hipMemcpyHostToDevice, st));
        return v->dev;
    }
};

static BankCache &bank_cache()
{
    thread_local BankCache c;
    return c;
}

template <bool CPX>
static void corr_std_impl(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code, int N, double coff,
                          const double *pos, int n, sdr_cpx_t *corr, sdr_cpx_t *C)
{
    if (n <= 0 || N <= 0) return;
    if (n > PSA_MAX_CORR) PSA_FAIL("corr_std: n=%d exceeds SDR_MAX_CORR", n);
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    std::vector<int2> jk(n);
    tap_geometry(N, coff, pos, n, jk.data());

    const void *dcode = bank_cache().get(code, sizeof(sdr_cpx16_t) * (size_t)N * PSA_N_CODES, st);
    size_t iq_bytes = sizeof(sdr_cpx16_t) * (size_t)N;
    char *din = (char *)tc.s[SC_IQ].get(iq_bytes + 256 + sizeof(int2) * n);
    int2 *djk = (int2 *)(din + ((iq_bytes + 255) / 256) * 256);
    // one pinned staging buffer for the whole call (inputs, then outputs)
    uint8_t *hpin = (uint8_t *)tc.host_pinned(std::max(iq_bytes + sizeof(int2) * n,
                                                       sizeof(float2) * (n + 2)));
    memcpy(hpin, IQ, iq_bytes);
    memcpy(hpin + iq_bytes, jk.data(), sizeof(int2) * n);
    PSA_CHECK(hipMemcpyAsync(din, hpin, iq_bytes, hipMemcpyHostToDevice, st));
    PSA_CHECK(hipMemcpyAsync(djk, hpin + iq_bytes, sizeof(int2) * n, hipMemcpyHostToDevice, st));

    float *dout = (float *)tc.s[SC_MISC].get(sizeof(float4) * n + sizeof(float2) * (n + 2));
    float4 *c12 = (float4 *)dout;
    float2 *dcorr = (float2 *)(c12 + n);
    float2 *dC = dcorr + n;
    hipLaunchKernelGGL(corr_tap_kernel<CPX>, dim3(n), dim3(256), 0, st,
                       (const char2 *)din, (const char2 *)dcode, N, djk, c12);
    PSA_CHECK(hipGetLastError());
    hipLaunchKernelGGL(corr_final_kernel, dim3(1), dim3(128), 0, st, c12, n, N, dcorr, dC);
    PSA_CHECK(hipGetLastError());
    PSA_CHECK(hipStreamSynchronize(st));
    float *hout = (float *)hpin;
    PSA_CHECK(hipMemcpyAsync(hout, dcorr, sizeof(float2) * (n + 2), hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(corr, hout, sizeof(float2) * n);
    memcpy(C, hout + 2 * n, sizeof(float2) * 2);
}

// ===========================================================================
// batched open-loop tracking
// ===========================================================================

// Open-loop state per (channel, epoch), one 64-thread workgroup per channel
// (the recurrence of trk_open_loop), then the sdr_mix_carr parameters of
// the epoch's N samples at ring index (e N) mod buffN and cs = coff fs.
__global__ void __launch_bounds__(TRK_SCHUNK) trk_state_kernel(const TrkChDev *__restrict__ chs, int n_epoch,
                                                              int N, int buffN, TrkEpoch *__restrict__ ep)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x;
    const TrkChDev ch = chs[c];
    trk_open_loop(ch, n_epoch, [&](int e, double ph, double coff, double fc) {
#pragma clang fp contract(off)
        const int ix = (int)(((long long)e * N) % buffN);
        TrkEpoch q;
        q.ix = ix;
        trk_mix_params(buffN, ix, N, ch.fs, fc, ph, q.p0, q.p1, q.nseg, q.s);
        q.idle = 0;
        q.neg = 0;
        q.cs = coff * ch.fs;
        ep[(long long)c * n_epoch + e] = q;
    });
}

// Fused mix + correlate: one workgroup per (channel, epoch).  The epoch's N
// samples are carrier-mixed into LDS as separate int8 I and Q planes.  Each
// wave then owns up to TPW taps and walks all data words (4 samples, one
// per lane per step): per word one aligned code-word load per tap and v_dot4
// products.  The bank of each tap is read circularly shifted by j from a
// doubled copy (bank twice in a row, 4 byte-shifted copies so every word
// load is aligned), so both halves of sdr_corr_std's split (before / after the
// code wrap at sample j) read one contiguous sequence; a step's 64 words lie
// wholly in one half except at the split, so the half choice is a
// wave-uniform branch.  The one word that straddles j is left to the
// epilogue.  EXACT selects _mm256_sign_epi8 semantics for IF buffers that
// hold -8 nibbles (the only source of -128 mixed values); SAMEQ (I = Q in
// every bank, true for real codes) loads one code plane.
#define TRK_THREADS 512
#define TRK_TPW 6

// A carrier word (int8 cI in byte 0, cQ in byte 1) as one v_dot4 operand,
// the int8 lanes (cI, cQ, kI, kQ) with kI = 8 cI + 7 cQ and kQ = 8 (cI + cQ)
// modulo 256.  A sample with uI = sI + 8, uQ = sQ + 8 and vQ = 7 - sQ (all
// 0 .. 15) gives sI cI - sQ cQ as the product with (uI, vQ, -1, 0) and
// sI cQ + sQ cI as the product with (uQ, uI, 0, -1), modulo 256 (only the
// low byte, the int8 wrap, is kept).
__device__ __forceinline__ uint32_t carr_dot_entry(uint32_t c)
{
    const int cI = (int8_t)(c & 0xFF), cQ = (int8_t)(c >> 8);
    return (uint32_t)(uint8_t)cI | ((uint32_t)(uint8_t)cQ << 8) | ((uint32_t)(uint8_t)(8 * cI + 7 * cQ) << 16) |
           ((uint32_t)(uint8_t)(8 * (cI + cQ)) << 24);
}

// LDS working set of one (channel, epoch) correlation
struct TrkLds {
    uint32_t *planeI, *planeQ;   // mixed samples, Nq4 words each
    const uint32_t *carr;        // carrier table as v_dot4 operands
    float4 *c12s;                // per tap: half-1 / half-2 sums (I, Q)
    int2 *jk;                    // per tap: code offset j, bank k
    int *sgn;                    // the data-sign decision of the taps 0-2
    int4 *psum;                  // per wave: word-range partial sums (split taps)
};

// One epoch's sdr_mix_carr + sdr_corr_std by the whole workgroup: the
// epoch's state q, the channel's code banks from cbase; corr_out receives the
// n_pos taps and C_out the two halves of tap 0 (global or LDS).  Starts and
// ends with the workgroup in step (barriers inside).
// FROM16: buf holds N already-mixed cpx16 samples (int8 I, Q pairs, 8-byte
// aligned) instead of the IF ring; they are split into the planes as they are
// (sdr_corr_std's input).
template <bool EXACT, bool SAMEQ, int TPW = TRK_TPW, int NTH = TRK_THREADS, bool FROM16 = false>
__device__ __forceinline__ void trk_correlate(const TrkEpoch &q, const uint8_t *__restrict__ buf, int N,
                                              int n_pos, const double *__restrict__ pos,
                                              const uint32_t *__restrict__ cbase, long long code_copy_stride,
                                              int NP, const TrkLds &L, float2 *corr_out, float2 *C_out)
{
    const int Nq = (N + 3) / 4;
    const int Nq4 = (Nq + 255) / 256 * 256;    // planes padded to whole steps
    uint32_t *planeI = L.planeI, *planeQ = L.planeQ;
    const uint32_t *carr = L.carr;
    float4 *c12s = L.c12s;
    int2 *jk = L.jk;
    // sdr_corr_std tap geometry (src/sdr_func.c:988-991)
    for (int i = threadIdx.x; i < n_pos; i += NTH) {
        const double p = q.cs + pos[i];
        int j = (int)floor(p);
        const int k = (int)((p - j) * PSA_N_CODES);
        if (j < 0) j += N;
        else if (j >= N) j -= N;
        jk[i] = make_int2(j, k);
    }
    __syncthreads();

    // ---- mix N samples into the I / Q planes (zero padded to Nq4 words) ----
    // A word whose 4 samples lie in one ring segment takes its 4 IF bytes
    // with two aligned dword loads and a byte align (the ring is padded by a
    // dword) and steps the phase by s; other words go sample by sample.
    if constexpr (FROM16) {
        const uint2 *__restrict__ iq = reinterpret_cast<const uint2 *>(buf);
        for (int w = threadIdx.x; w < Nq4; w += NTH) {
            const int i0 = 4 * w;
            uint32_t wi = 0, wq = 0;
            if (i0 + 4 <= N) {
                // samples i0 .. i0 + 3 as (I, Q) byte pairs in one 8-byte
                // load: I in bytes 0, 2 of each dword, Q in bytes 1, 3
                const uint2 v = iq[w];
                wi = (v.x & 0xFFu) | ((v.x >> 8) & 0xFF00u) | ((v.y & 0xFFu) << 16) | ((v.y << 8) & 0xFF000000u);
                wq = ((v.x >> 8) & 0xFFu) | ((v.x >> 16) & 0xFF00u) | ((v.y << 8) & 0xFF0000u) | (v.y & 0xFF000000u);
            } else {
                const uint8_t *b8 = buf;
                for (int b = 0; b < 4; b++) {
                    const int i = i0 + b;
                    if (i < N) {
                        wi |= (uint32_t)b8[2 * i] << (8 * b);
                        wq |= (uint32_t)b8[2 * i + 1] << (8 * b);
                    }
                }
            }
            planeI[w] = wi;
            planeQ[w] = wq;
        }
        __syncthreads();
    } else {
    // phase of sample i in the first segment's terms, p0 + i s, carried
    // from word to word; the second segment's p1 + (i - nseg) s differs from
    // it by a constant (uint32 arithmetic, mod 2^32)
    const uint32_t dseg = q.p1 - q.p0 - (uint32_t)q.nseg * q.s;
    uint32_t phw = q.p0 + (uint32_t)(4 * (int)threadIdx.x) * q.s;
    const uint32_t phstep = (uint32_t)(4 * NTH) * q.s;
    const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(buf), (short)0,
                                                                         0x7FFFFFFF, 0x00020000);
    // MU words per thread at a time: their IF loads are all in flight before
    // the first is mixed (the closed loop's single workgroup per CU would
    // otherwise wait out one load latency per word).  Whether a word lies
    // below Nq4 is the same for a whole wave (Nq4 and NTH are multiples of
    // 64), so those tests are scalar branches.
    constexpr int MU = 6;
    for (int w0 = threadIdx.x; w0 < Nq4; w0 += MU * NTH) {
        uint32_t A0[MU], A1[MU];
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int w = w0 + u * NTH, i0 = 4 * w;
            if (w < Nq4) {
                const bool s1 = i0 + 4 <= N && i0 + 4 <= q.nseg, s2 = i0 + 4 <= N && i0 >= q.nseg;
                // words that go sample by sample load the ring's first dword pair
                const int ab = (s1 ? q.ix + i0 : s2 ? i0 - q.nseg : 0) & ~3;
                A0[u] = __builtin_amdgcn_raw_buffer_load_b32(brs, ab, 0, 0);
                A1[u] = __builtin_amdgcn_raw_buffer_load_b32(brs, ab, 4, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < MU; u++, phw += phstep) {
            const int w = w0 + u * NTH, i0 = 4 * w;
            if (w >= Nq4) break;
            uint32_t wi = 0, wq = 0;
            const bool s1 = i0 + 4 <= N && i0 + 4 <= q.nseg, s2 = i0 + 4 <= N && i0 >= q.nseg;
            if (s1 || s2) {
                const int src = s1 ? q.ix + i0 : i0 - q.nseg;
                uint32_t ph = s1 ? phw : phw + dseg;
                const uint32_t d4 = __builtin_amdgcn_alignbyte(A1[u], A0[u], (uint32_t)(src & 3));
                // per byte: uI = sI + 8, uQ = sQ + 8 and vQ = 7 - sQ (0 .. 15)
                const uint32_t uI = (d4 & 0x0F0F0F0Fu) ^ 0x08080808u;
                const uint32_t uQ = ((d4 >> 4) & 0x0F0F0F0Fu) ^ 0x08080808u;
                const uint32_t vQ = 0x0F0F0F0Fu - uQ;
                int vi[4], vq[4];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    // (uI_b, vQ_b, -1, 0) and (uQ_b, uI_b, 0, -1)
                    const int x1 = (int)__builtin_amdgcn_perm(vQ, uI, 0x0C0D0400u | (0x0101u * b));
                    const int x2 = (int)__builtin_amdgcn_perm(uQ, uI, 0x0D0C0004u | (0x0101u * b));
                    const int c = (int)carr[ph >> 24];
                    ph += q.s;
                    vi[b] = __builtin_amdgcn_sdot4(x1, c, 0, false);
                    vq[b] = __builtin_amdgcn_sdot4(x2, c, 0, false);
                }
                // byte 0 of each product into byte b of the word
                wi = __builtin_amdgcn_perm(__builtin_amdgcn_perm((uint32_t)vi[3], (uint32_t)vi[2], 0x0C0C0400u),
                                           __builtin_amdgcn_perm((uint32_t)vi[1], (uint32_t)vi[0], 0x0C0C0400u),
                                           0x05040100u);
                wq = __builtin_amdgcn_perm(__builtin_amdgcn_perm((uint32_t)vq[3], (uint32_t)vq[2], 0x0C0C0400u),
                                           __builtin_amdgcn_perm((uint32_t)vq[1], (uint32_t)vq[0], 0x0C0C0400u),
                                           0x05040100u);
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    int i = i0 + b;
                    if (i < N) {
                        int src;
                        uint32_t ph;
                        if (i < q.nseg) { src = q.ix + i; ph = q.p0 + (uint32_t)i * q.s; }
                        else { src = i - q.nseg; ph = q.p1 + (uint32_t)(i - q.nseg) * q.s; }
                        uint32_t d = buf[src];
                        int sI = (int)((int8_t)(uint8_t)(d << 4)) >> 4;
                        int sQ = (int)((int8_t)(uint8_t)d) >> 4;
                        const uint32_t c = carr[ph >> 24];
                        const int cI = (int8_t)(c & 0xFF), cQ = (int8_t)(c >> 8);
                        wi |= (uint32_t)(uint8_t)(int8_t)(sI * cI - sQ * cQ) << (8 * b);
                        wq |= (uint32_t)(uint8_t)(int8_t)(sI * cQ + sQ * cI) << (8 * b);
                    }
                }
            }
            planeI[w] = wi;
            planeQ[w] = wq;
        }
    }
    __syncthreads();
    }

    const uint8_t *cb = (const uint8_t *)cbase;   // [copy][bankI|bankQ][k][NP]
    const int cstride = (int)code_copy_stride;
    // the wave index and everything per tap are wave-uniform: held in scalar
    // registers, so the half choice below is a scalar branch and the code
    // loads take a scalar base
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
    const int QOFF = PSA_N_CODES * NP;       // byte offset of the Q banks
    // ---- taps: wave `wave` owns t = t0 + s nwave, s < TPW ----
    // With at most half as many taps as waves (the closed loop's four), the
    // words are split instead: wave w takes tap w % n_pos over word range
    // part w / n_pos, the parts' exact integer sums meet in LDS and the
    // wave of part 0 runs the epilogue, so the results equal the unsplit
    // ones.
    const int nsplit = n_pos > 0 && 2 * n_pos <= nwave ? nwave / n_pos : 1;
    const int part = nsplit > 1 ? wave / n_pos : 0;
    const int nstep = (Nq + 255) / 256, sps = (nstep + nsplit - 1) / nsplit;
    const int wlo = min(part * sps, nstep) * 256, whi = min((part + 1) * sps * 256, Nq);
    // (split: every wave runs one pass, waves past the last part with an
    // empty word range, so the barrier below is reached by all)
    for (int t0 = nsplit > 1 ? wave % n_pos : wave; t0 < n_pos; t0 += nwave * TPW) {
        int jt[TPW], co[TPW];
        // per tap and lane: the running sum of every word so far (tot), its
        // value through the word straddling sample j (snap) and through the
        // last word wholly before j (h1, the half-1 sum); half 2 is then
        // tot - snap (exact int32 sums)
        int tot[TPW][2], snap[TPW][2], h1[TPW][2];
#pragma unroll
        for (int s = 0; s < TPW; s++) {
            const int t = t0 + s * nwave;
            const int2 g = jk[t < n_pos ? t : 0];
            const int gx = __builtin_amdgcn_readfirstlane(g.x), gy = __builtin_amdgcn_readfirstlane(g.y);
            // code sample of data sample m: doubled bank byte N - j + m
            const int c0 = N - gx, cpy = c0 & 3;
            jt[s] = t < n_pos ? gx : N;
            co[s] = cpy * cstride + ((gy * NP) >> 2) + (c0 - cpy) / 4;
#pragma unroll
            for (int c = 0; c < 2; c++) tot[s][c] = snap[s][c] = h1[s][c] = 0;
        }
        // a step covers 256 words: lane l takes words 4 l .. 4 l + 3 of it with
        // one 16-byte load per tap (planes padded with zero words to Nq4):
        // buffer loads at the tap's scalar byte offset plus 16 l
        const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t *>(cbase), (short)0, 0x7FFFFFFF, 0x00020000);
        // one step: this lane's 4 words of the planes against each tap's
        // code words cI / cQ
        auto step = [&](const int wb, const uint4 *cI, const uint4 *cQ) {
            const int w = wb + 4 * lane;
            const uint4 xI = *reinterpret_cast<const uint4 *>(planeI + w);
            const uint4 xQ = *reinterpret_cast<const uint4 *>(planeQ + w);
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int j = jt[s];
                const uint32_t xi[4] = {xI.x, xI.y, xI.z, xI.w}, xq[4] = {xQ.x, xQ.y, xQ.z, xQ.w};
                const uint32_t ci[4] = {cI[s].x, cI[s].y, cI[s].z, cI[s].w};
                const uint32_t cq[4] = {cQ[s].x, cQ[s].y, cQ[s].z, cQ[s].w};
                // this step covers samples 4 wb .. 4 wb + 1023 (a scalar test)
                if (!(4 * wb + 1024 <= j || 4 * wb > j)) {
                    // the step holding sample j: half 1 ends with the words
                    // wholly before j (h1), half 2 is the final sum minus
                    // everything up to and including the word straddling j
                    // (snap), which the epilogue adds on its own
                    int bI = 0, bQ = 0, sI = 0, sQ = 0;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        int dI = 0, dQ = 0;
                        dot4x4<EXACT, 1>(xi + v, xq + v, ci + v, cq + v, dI, dQ);
                        const int m = 4 * (w + v);
                        if (m + 4 <= j) { bI += dI; bQ += dQ; }
                        else if (m < j) { sI += dI; sQ += dQ; }
                    }
                    h1[s][0] = tot[s][0] + bI;
                    h1[s][1] = tot[s][1] + bQ;
                    snap[s][0] = h1[s][0] + sI;
                    snap[s][1] = h1[s][1] + sQ;
                }
                // every step adds all its words to the running sum
                dot4x4<EXACT>(xi, xq, ci, cq, tot[s][0], tot[s][1]);
            }
        };
        auto load_codes = [&](const int wb, uint4 *cI, uint4 *cQ) {
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int so = 4 * (co[s] + wb);
                cI[s] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(crs, 16 * lane, so, 0));
                cQ[s] = SAMEQ ? cI[s]
                              : __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(crs, 16 * lane,
                                                                                                 so + QOFF, 0));
            }
        };
        if constexpr (TPW == 1) {
            // one tap per wave (the closed loop): the code words of KB steps
            // are in flight together, one load latency per KB steps
            constexpr int KB = 4;
            for (int wb0 = wlo; wb0 < whi; wb0 += KB * 256) {
                uint4 cI[KB], cQ[KB];
#pragma unroll
                for (int k = 0; k < KB; k++) {
                    if (wb0 + k * 256 < whi) load_codes(wb0 + k * 256, &cI[k], &cQ[k]);
                }
#pragma unroll
                for (int k = 0; k < KB; k++) {
                    if (wb0 + k * 256 < whi) step(wb0 + k * 256, &cI[k], &cQ[k]);
                }
            }
        } else {
            for (int wb = wlo; wb < whi; wb += 256) {
                uint4 cI[TPW], cQ[TPW];
                load_codes(wb, cI, cQ);
                step(wb, cI, cQ);
            }
        }
        if (nsplit > 1) {
            // a part wholly before the step holding j is all half 1
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                if (4 * whi <= jt[s]) {
#pragma unroll
                    for (int c = 0; c < 2; c++) h1[s][c] = snap[s][c] = tot[s][c];
                }
            }
            if (part > 0) {
                // this part's four sums, reduced over the lanes, to LDS
                const int4 p = wave_sum4(h1[0][0], h1[0][1], tot[0][0] - snap[0][0], tot[0][1] - snap[0][1]);
                if (lane == 0) L.psum[wave] = p;
            }
            __syncthreads();
            if (part > 0) continue;
        }
        // ---- per-tap float epilogue (the owning wave): the integer sums hold
        // the first 16 floor(len / 16) samples of each half, the rest are
        // added in float one by one (dot_IQ_code's AVX2 body and scalar
        // tail, src/sdr_func.c:861-912).  Per tap, lanes 0-15 / 16-31 fetch
        // the tail samples of half 1 / 2 and lanes 32-35 the word straddling
        // j (not in the running sums); each lane folds its integer correction
        // (+ straddle, - tail) into its partial sums, all taps' partials are
        // reduced together, and the ordered float tail is added to each.
        const uint8_t *pI = (const uint8_t *)planeI, *pQ = (const uint8_t *)planeQ;
#pragma unroll
        for (int s = 0; s < TPW; s++) {
            const int t = t0 + s * nwave;
            if (t >= n_pos) break;
            const int j = jt[s];
            const int k = __builtin_amdgcn_readfirstlane(jk[t].y);
            const int len2 = N - j;
            const int V1 = j >= 16 ? j / 16 * 16 : 0;
            const int V2 = len2 >= 16 ? len2 / 16 * 16 : 0;
            int m = -1, ci = 0, slot = 0;   // slot 0/1: tail of half 1/2; 2/3: straddle in half 1/2
            if (lane < 16) {
                const int i = V1 + lane;
                if (i < j) { m = i; ci = N - j + i; slot = 0; }
            } else if (lane < 32) {
                const int i = V2 + lane - 16;
                if (i < len2) { m = j + i; ci = i; slot = 1; }
            } else if (lane < 36) {
                const int mm = (j & ~3) + lane - 32;
                if ((j & 3) && mm < N) { m = mm; ci = mm < j ? N - j + mm : mm - j; slot = mm < j ? 2 : 3; }
            }
            int xI = 0, xQ = 0, vI = 0, vQ = 0;
            if (m >= 0) {
                xI = (int8_t)pI[m]; xQ = (int8_t)pQ[m];
                vI = (int8_t)cb[(long long)k * NP + ci];
                vQ = (int8_t)cb[QOFF + (long long)k * NP + ci];
            }
            const int eI = EXACT ? sgn8(xI, vI) : xI * vI, eQ = EXACT ? sgn8(xQ, vQ) : xQ * vQ;
            const int fI = xI * vI, fQ = xQ * vQ;
            // tail samples leave the running sums (-e) and are added in float
            // one by one below; with N < 120000 every partial sum is an
            // integer below 2^24 (vector part |x sign(c)| <= 128 per sample,
            // at most 15 tail products |x c| <= 128^2), so those float adds
            // are exact and the tail joins the integer sum instead (+f - e),
            // converted once: the same float
            const bool itail = N < 120000;
            const int tI = itail ? fI - eI : -eI, tQ = itail ? fQ - eQ : -eQ;
            const int aI = m < 0 ? 0 : (slot >= 2 ? eI : tI), aQ = m < 0 ? 0 : (slot >= 2 ? eQ : tQ);
            const bool first = slot == 0 || slot == 2;
            int p[4] = {h1[s][0] + (first ? aI : 0), h1[s][1] + (first ? aQ : 0),
                        tot[s][0] - snap[s][0] + (first ? 0 : aI),
                        tot[s][1] - snap[s][1] + (first ? 0 : aQ)};
            const int4 r = wave_sum4(p[0], p[1], p[2], p[3]);
            int4 ex = make_int4(0, 0, 0, 0);    // the other parts' sums (split)
            for (int q = 1; q < nsplit; q++) {
                const int4 v = L.psum[t + q * n_pos];
                ex.x += v.x; ex.y += v.y; ex.z += v.z; ex.w += v.w;
            }
            float f1I = (float)(r.x + ex.x);
            float f1Q = (float)(r.y + ex.y);
            float f2I = (float)(r.z + ex.z);
            float f2Q = (float)(r.w + ex.w);
            for (int i = 0; !itail && i < j - V1; i++) {
                f1I += (float)__builtin_amdgcn_readlane(fI, i);
                f1Q += (float)__builtin_amdgcn_readlane(fQ, i);
            }
            for (int i = 0; !itail && i < len2 - V2; i++) {
                f2I += (float)__builtin_amdgcn_readlane(fI, 16 + i);
                f2Q += (float)__builtin_amdgcn_readlane(fQ, 16 + i);
            }
            if (lane == 0) {
                c12s[t] = make_float4(f1I * (1.0f * PSA_CSCALE), f1Q * (1.0f * PSA_CSCALE),
                                      f2I * (1.0f * PSA_CSCALE), f2Q * (1.0f * PSA_CSCALE));
            }
        }
    }
    __syncthreads();

    if (threadIdx.x == 0) {
        float dot = 0.0f;
        for (int i = 0; i < 3 && i < n_pos; i++) {
            float4 c = c12s[i];
            dot = __fadd_rn(dot, __fmaf_rn(c.x, c.z, __fmul_rn(c.y, c.w)));
        }
        *L.sgn = dot < 0.0f ? -1 : 1;
        if (n_pos > 0) {
            C_out[0] = make_float2(c12s[0].x, c12s[0].y);
            C_out[1] = make_float2(c12s[0].z, c12s[0].w);
        }
    }
    __syncthreads();
    const float sign = (float)*L.sgn;
    const float fN = (float)N;
    for (int t = threadIdx.x; t < n_pos; t += blockDim.x) {
        float4 c = c12s[t];
        corr_out[t] = make_float2(__fdiv_rn(__fadd_rn(c.x, __fmul_rn(sign, c.z)), fN),
                                  __fdiv_rn(__fadd_rn(c.y, __fmul_rn(sign, c.w)), fN));
    }
    __syncthreads();
}

template <bool EXACT, bool SAMEQ>
__device__ __forceinline__ void trk_corr_body(
    const uint8_t *__restrict__ buf, int buffN, int N, int n_epoch, int n_pos,
    const TrkEpoch *__restrict__ ep, const double *__restrict__ pos,
    const uint32_t *__restrict__ codes, long long code_ch_stride, long long code_copy_stride,
    int NP,  // padded bank length in bytes (multiple of 4)
    const uint32_t *__restrict__ carr_g,
    float2 *__restrict__ corr_out, float2 *__restrict__ C_out, int e0, int nec)
{
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t carr[256];
    __shared__ float4 c12s[PSA_MAX_CORR];
    __shared__ int2 jk[PSA_MAX_CORR];
    __shared__ int sgn;
    __shared__ int4 psum[TRK_THREADS / 64];
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    // the grid covers epochs e0 .. e0 + nec - 1 of every channel
    const long long o = (long long)(blockIdx.x / (unsigned)nec) * n_epoch + e0 + blockIdx.x % (unsigned)nec;
    const int ch = (int)(o / n_epoch);
    const TrkEpoch q = ep[o];
    if (q.idle) {                            // uniform over the workgroup
        for (int t = threadIdx.x; t < n_pos; t += TRK_THREADS) corr_out[o * n_pos + t] = make_float2(0.0f, 0.0f);
        if (threadIdx.x < 2) C_out[2 * o + threadIdx.x] = make_float2(0.0f, 0.0f);
        return;
    }
    for (int i = threadIdx.x; i < 256; i += TRK_THREADS) carr[i] = carr_dot_entry(carr_g[i]);
    const TrkLds L{lds, lds + Nq4, carr, c12s, jk, &sgn, psum};
    trk_correlate<EXACT, SAMEQ>(q, buf, N, n_pos, pos, codes + (long long)ch * code_ch_stride,
                                code_copy_stride, NP, L, corr_out + o * n_pos, C_out + 2 * o);
    if (q.neg) {                             // a closed loop wiped its secondary code: trk->C
        __syncthreads();
        for (int t = threadIdx.x; t < n_pos; t += TRK_THREADS) {
            const float2 v = corr_out[o * n_pos + t];
            corr_out[o * n_pos + t] = make_float2(-v.x, -v.y);
        }
    }
}

#define PSA_TRK_ARGS                                                                                   \
    const uint8_t *__restrict__ buf, int buffN, int N, int n_epoch, int n_pos,                         \
        const TrkEpoch *__restrict__ ep, const double *__restrict__ pos,                               \
        const uint32_t *__restrict__ codes, long long code_ch_stride, long long code_copy_stride, int NP, \
        const uint32_t *__restrict__ carr_g, float2 *__restrict__ corr_out, float2 *__restrict__ C_out, int e0, \
        int nec
#define PSA_TRK_PASS \
    buf, buffN, N, n_epoch, n_pos, ep, pos, codes, code_ch_stride, code_copy_stride, NP, carr_g, corr_out, C_out, \
        e0, nec

template <bool EXACT, bool SAMEQ>
__global__ void __launch_bounds__(TRK_THREADS) trk_corr_kernel(PSA_TRK_ARGS)
{
    trk_corr_body<EXACT, SAMEQ>(PSA_TRK_PASS);
}

// real codes without -8 nibbles (the common form): compiled for six waves per
// SIMD (80 registers), so that three workgroups share a CU
__global__ void __launch_bounds__(TRK_THREADS) __attribute__((amdgpu_waves_per_eu(6)))
trk_corr_kernel_real(PSA_TRK_ARGS)
{
    trk_corr_body<false, true>(PSA_TRK_PASS);
}

// ===========================================================================
// closed-loop tracking with the standard correlators: track_sig
// (src/sdr_ch.c:453-624, 694-763) with the secondary-code sync and
// bump-jump, the loop steps of psa_loop.h.  One persistent workgroup per
// channel runs the channel's epochs in order: lane 0 advances the NCO and
// code offset, the workgroup mixes and correlates the epoch (trk_correlate)
// on the taps the loop reads (P, E, L and the noise tap, TL_NTAP, and VE /
// VL with bump-jump), lane 0 stitches the prompt and updates the loop.
// Every epoch's mixing parameters go to ep (idle after a loss, neg for a
// wiped chip of -1); trk_corr_kernel then correlates all taps of all epochs
// from them in one launch over the whole GPU.  The taps are computed
// independently of each other, so the loop's equal the full launch's.
// ===========================================================================
// loop workgroup size: 1024 threads where the correlator fits 128 VGPRs
// (no -8 nibbles, real codes), else 512
template <bool EXACT, bool SAMEQ>
constexpr int trk_loop_threads() { return EXACT || !SAMEQ ? 512 : 1024; }

template <bool EXACT, bool SAMEQ, int NTH>
__global__ void __launch_bounds__(NTH) trk_loop_kernel(
    const uint8_t *__restrict__ buf, int buffN, int N, int n_epoch, int n_pos,
    const double *__restrict__ pos, const uint32_t *__restrict__ codes, long long code_ch_stride,
    long long code_copy_stride, int NP, const uint32_t *__restrict__ carr_g,
    TrkLoopCh *__restrict__ chs, TrkLoopOut *__restrict__ out, TrkEpoch *__restrict__ ep,
    const int8_t *__restrict__ sec_all, int e0, int e1, float2 *__restrict__ ghist)
{
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t carr[256];
    __shared__ float4 c12s[PSA_MAX_CORR];
    __shared__ int2 jk[PSA_MAX_CORR];
    __shared__ float2 tap[PSA_MAX_CORR];
    __shared__ float2 Ch[2];
    __shared__ int sgn, stop;
    __shared__ int4 psum[NTH / 64];
    __shared__ TrkEpoch qs;
    __shared__ TrkLoopCh S;
    const int c = blockIdx.x;
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    for (int i = threadIdx.x; i < 256; i += NTH) carr[i] = carr_dot_entry(carr_g[i]);
    if (threadIdx.x == 0) {
        S = chs[c];
        stop = S.idle;
    }
    __syncthreads();
    const TrkLds L{lds, lds + Nq4, carr, c12s, jk, &sgn, psum};
    const uint32_t *cbase = codes + (long long)c * code_ch_stride;
    // the prompt history after the IF planes (channels with a secondary
    // code; trk->P starts at zero, src/sdr_ch.c:369)
    float2 *hist = reinterpret_cast<float2 *>(lds + 2 * Nq4);
    const bool use_hist = S.len_sec >= 2;
    // epochs e0 .. e1 - 1 of a run: the prompt history continues from the
    // previous launch's (kept in ghist)
    if (use_hist) {
        for (int i = threadIdx.x; i < TL_N_HIST; i += NTH)
            hist[i] = e0 == 0 ? make_float2(0.0f, 0.0f) : ghist[(long long)c * TL_N_HIST + i];
    }
    const int8_t *sec = sec_all + (use_hist ? S.sec_off : 0);
    const int nlt = S.bump_step > 0.0 ? 6 : TL_NTAP;       // the loop's taps
    __syncthreads();
    int e = e0;
    for (; e < e1; e++) {
        if (stop) break;                       // written by lane 0 before a barrier
        if (threadIdx.x == 0) tl_advance(S, N, buffN, qs, hist);
        __syncthreads();
        const TrkEpoch q = qs;
        trk_correlate<EXACT, SAMEQ, 1, NTH>(q, buf, N, nlt, pos, cbase, code_copy_stride, NP, L, tap, Ch);
        const long long o = (long long)c * n_epoch + e;
        if (threadIdx.x == 0) {
            TrkEpoch qe = q;
            TrkLoopOut r;
            tl_update(S, N, tap, Ch, r, hist, sec, qe);
            ep[o] = qe;
            out[o] = r;
            stop = S.idle;
        }
        __syncthreads();
    }
    // epochs after a loss of lock: marked idle, state as at the loss
    for (int e2 = e + threadIdx.x; e2 < e1; e2 += NTH) {
        TrkLoopOut r;
        r.fd = S.fd; r.coff = S.coff; r.phi = S.phi; r.cn0 = S.cn0;
        r.lock = S.lock; r.idle = 1;
        r.P = make_float2(0.0f, 0.0f);
        out[(long long)c * n_epoch + e2] = r;
        TrkEpoch qi = {};
        qi.idle = 1;
        ep[(long long)c * n_epoch + e2] = qi;
    }
    if (use_hist && e1 < n_epoch) {
        for (int i = threadIdx.x; i < TL_N_HIST; i += NTH) ghist[(long long)c * TL_N_HIST + i] = hist[i];
    }
    if (threadIdx.x == 0) chs[c] = S;
}

// Prompt history of track_sig (src/sdr_ch.c:736-740): the full-period prompt
// of epoch e joins this epoch's correlation before the code wrap with the
// previous epoch's after it, (C1[0] + trk->C1) / N in float, trk->C1 = 0
// before the first epoch (src/sdr_ch.c:368).
__global__ void trk_stitch_kernel(const float2 *__restrict__ C, long long n_item, int n_epoch, int N,
                                  float2 *__restrict__ P)
{
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_item) return;
    const int e = (int)(o % n_epoch);
    const float2 a = C[2 * o];
    const float2 b = e > 0 ? C[2 * o - 1] : make_float2(0.0f, 0.0f);
    const float fN = (float)N;
    P[o] = make_float2(__fdiv_rn(__fadd_rn(a.x, b.x), fN), __fdiv_rn(__fadd_rn(a.y, b.y), fN));
}

}  // namespace psa

using namespace psa;

struct sdr_trk_plan_tag {
    int n_ch, n_pos, n_epoch, N, buffN;
    bool exact, sameq;
    int NP;
    long long code_ch_stride, code_copy_stride;   // in uint32 words
    uint8_t *d_buf = nullptr;
    uint32_t *d_codes = nullptr;
    TrkChDev *d_ch = nullptr;
    double *d_pos = nullptr;
    TrkEpoch *d_ep = nullptr;
    float2 *d_corr = nullptr, *d_C = nullptr, *d_P = nullptr;
    // closed-loop mode: initial channel states, their working copies, and
    // the per-epoch state records
    bool loop = false;
    std::vector<TrkLoopCh> lch0;
    TrkLoopCh *d_lch = nullptr;
    TrkLoopOut *d_lout = nullptr;
    int8_t *d_sec = nullptr;                   // secondary codes of the loop channels
    float2 *d_hist = nullptr;                  // their prompt histories between launches
    hipStream_t st = nullptr;
    // closed loop: the all-tap pass of each chunk of epochs runs on st2
    // while the loop continues on st
    hipStream_t st2 = nullptr;
    std::vector<hipEvent_t> evc;
    bool timing = false;
    hipEvent_t ev[4];
    double ms_state = 0, ms_corr = 0;
    long long n_state = 0, n_corr = 0;
};

// ---------------------------------------------------------------------------
// sdr_corr_std on the tracking correlator's device code: one workgroup splits
// the caller's mixed samples into int8 I / Q planes in LDS and correlates all
// taps with v_dot4 words from the bank in the correlator's layout (each of the
// SDR_N_CODES banks doubled and in four byte-shifted copies).  The device
// banks are cached per calling thread, keyed on the bank's address, size and
// a 64-entry fingerprint: the reference's banks are built once per channel
// and never modified (src/sdr_ch.c:266-272), so no call compares the whole
// bank.
// ---------------------------------------------------------------------------
namespace psa {

// doubled, byte-shifted code layout of one bank set: [copy 0..3][I banks
// 0..9 | Q banks 0..9][NP bytes], byte i of a row = bank sample (i + copy)
// mod N for i < 2 N - copy, zero after
static void build_trk_bank(const sdr_cpx16_t *code, int N, int NP, uint8_t *dst)
{
    const size_t bank_bytes = (size_t)2 * PSA_N_CODES * NP;
    for (int cp = 0; cp < 4; cp++) {
        uint8_t *d = dst + (size_t)cp * bank_bytes;
        for (int comp = 0; comp < 2; comp++) {
            for (int k = 0; k < PSA_N_CODES; k++) {
                uint8_t *row = d + ((size_t)comp * PSA_N_CODES + k) * NP;
                const sdr_cpx16_t *src = code + (size_t)k * N;
                for (int i = 0; i < 2 * N - cp; i++) {
                    const sdr_cpx16_t &v = src[(i + cp) % N];
                    row[i] = (uint8_t)(comp == 0 ? v.I : v.Q);
                }
            }
        }
    }
}

template <bool EXACT, bool SAMEQ>
__global__ void __launch_bounds__(TRK_THREADS) corr_dropin_kernel(
    const uint8_t *__restrict__ iq, int N, double coff, int n_pos, const double *__restrict__ pos,
    const uint32_t *__restrict__ codes, long long code_copy_stride, int NP, float2 *__restrict__ corr_out,
    float2 *__restrict__ C_out)
{
    extern __shared__ uint32_t lds[];
    __shared__ float4 c12s[PSA_MAX_CORR];
    __shared__ int2 jk[PSA_MAX_CORR];
    __shared__ int sgn;
    __shared__ int4 psum[TRK_THREADS / 64];
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    TrkEpoch q = {};
    q.cs = coff;
    const TrkLds L{lds, lds + Nq4, nullptr, c12s, jk, &sgn, psum};
    trk_correlate<EXACT, SAMEQ, TRK_TPW, TRK_THREADS, true>(q, iq, N, n_pos, pos, codes, code_copy_stride, NP, L,
                                                            corr_out, C_out);
}

struct DropinBank {
    const void *host = nullptr;
    size_t bytes = 0;
    uint64_t fp = 0;
    uint8_t *dev = nullptr;
    int N = 0, NP = 0;
    bool sameq = false, ok = false;     // ok: values in {-1, 0, 1}
    unsigned long long used = 0;
};

struct DropinBankCache {
    DropinBank e[4];
    unsigned long long tick = 0;
    ~DropinBankCache()
    {
        for (auto &x : e) if (x.dev) (void)hipFree(x.dev);
    }
    const DropinBank &get(const sdr_cpx16_t *code, int N)
    {
        tick++;
        const size_t bytes = sizeof(sdr_cpx16_t) * (size_t)N * PSA_N_CODES;
        const uint64_t fp = bank_fp((const uint8_t *)code, bytes);
        for (auto &x : e) {
            if (x.host == code && x.bytes == bytes && x.fp == fp && x.N == N) {
                x.used = tick;
                return x;
            }
        }
        DropinBank *v = &e[0];
        for (auto &x : e) if (x.used < v->used) v = &x;
        v->host = code; v->bytes = bytes; v->fp = fp; v->N = N; v->used = tick;
        v->ok = true;
        v->sameq = true;
        for (size_t i = 0; i < (size_t)N * PSA_N_CODES; i++) {
            const int a = code[i].I, b = code[i].Q;
            if (a < -1 || a > 1 || b < -1 || b > 1) v->ok = false;
            if (a != b) v->sameq = false;
        }
        v->NP = (2 * N + 1024 + 16 + 3) / 4 * 4;
        const size_t lay = (size_t)4 * 2 * PSA_N_CODES * v->NP;
        if (v->dev) { PSA_CHECK(hipFree(v->dev)); v->dev = nullptr; }
        if (v->ok) {
            std::vector<uint8_t> h(lay, 0);
            build_trk_bank(code, N, v->NP, h.data());
            PSA_CHECK(hipMalloc(&v->dev, lay));
            PSA_CHECK(hipMemcpy(v->dev, h.data(), lay, hipMemcpyHostToDevice));
        }
        return *v;
    }
};

template <bool EXACT, bool SAMEQ>
static void dropin_launch_t(const uint8_t *iq, int N, double coff, int n, const double *pos, const DropinBank &bk,
                            float2 *corr, float2 *C, hipStream_t st)
{
    static std::once_flag once;
    std::call_once(once, [] {
        PSA_CHECK(hipFuncSetAttribute((const void *)corr_dropin_kernel<EXACT, SAMEQ>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    });
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    const long long copy_stride = (long long)(2 * PSA_N_CODES) * bk.NP / 4;
    hipLaunchKernelGGL((corr_dropin_kernel<EXACT, SAMEQ>), dim3(1), dim3(TRK_THREADS), sizeof(uint32_t) * 2 * Nq4,
                       st, iq, N, coff, n, pos, (const uint32_t *)bk.dev, copy_stride, bk.NP, corr, C);
    PSA_CHECK(hipGetLastError());
}

// sdr_corr_std through the tracking correlator; false when the bank holds
// values outside {-1, 0, 1} or N exceeds one workgroup's LDS (the caller
// then takes the per-tap kernels)
static bool corr_std_fast(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code, int N, double coff, const double *pos,
                          int n, sdr_cpx_t *corr, sdr_cpx_t *C)
{
    if (n > PSA_MAX_CORR) PSA_FAIL("corr_std: n=%d exceeds SDR_MAX_CORR", n);
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    if ((size_t)Nq4 * 8 + 4096 > 150 * 1024) return false;
    thread_local DropinBankCache cache;
    const DropinBank &bk = cache.get(code, N);
    if (!bk.ok) return false;
    std::vector<int2> jk(n);
    tap_geometry(N, coff, pos, n, jk.data());      // fails on taps outside one period
    ThreadCtx &tc = tctx();
    hipStream_t st = tc.stream;
    const size_t iq_bytes = sizeof(sdr_cpx16_t) * (size_t)N;
    const size_t po = (iq_bytes + 255) / 256 * 256;
    uint8_t *din = (uint8_t *)tc.s[SC_IQ].get(po + sizeof(double) * n);
    uint8_t *hpin = (uint8_t *)tc.host_pinned(std::max(po + sizeof(double) * n, sizeof(float2) * (n + 2)));
    memcpy(hpin, IQ, iq_bytes);
    memcpy(hpin + po, pos, sizeof(double) * n);
    // _mm256_sign_epi8 semantics only matter for samples of -128
    const bool exact = memchr(hpin, 0x80, iq_bytes) != nullptr;
    PSA_CHECK(hipMemcpyAsync(din, hpin, po + sizeof(double) * n, hipMemcpyHostToDevice, st));
    float2 *dout = (float2 *)tc.s[SC_MISC].get(sizeof(float2) * (n + 2));
    const double *dpos = (const double *)(din + po);
    if (exact) {
        if (bk.sameq) dropin_launch_t<true, true>(din, N, coff, n, dpos, bk, dout, dout + n, st);
        else dropin_launch_t<true, false>(din, N, coff, n, dpos, bk, dout, dout + n, st);
    } else {
        if (bk.sameq) dropin_launch_t<false, true>(din, N, coff, n, dpos, bk, dout, dout + n, st);
        else dropin_launch_t<false, false>(din, N, coff, n, dpos, bk, dout, dout + n, st);
    }
    PSA_CHECK(hipMemcpyAsync(hpin, dout, sizeof(float2) * (n + 2), hipMemcpyDeviceToHost, st));
    PSA_CHECK(hipStreamSynchronize(st));
    memcpy(corr, hpin, sizeof(float2) * n);
    memcpy(C, hpin + sizeof(float2) * n, sizeof(float2) * 2);
    return true;
}

}  // namespace psa

extern "C" {

void sdr_corr_std(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code, int N,
    double coff, const double *pos, int n, sdr_cpx_t *corr, sdr_cpx_t *C)
{
    if (n <= 0 || N <= 0) return;
    if (!corr_std_fast(IQ, code, N, coff, pos, n, corr, C)) corr_std_impl<false>(IQ, code, N, coff, pos, n, corr, C);
}

void sdr_corr_std_cpx_code(const sdr_cpx16_t *IQ, const sdr_cpx16_t *code,
    int N, double coff, const double *pos, int n, sdr_cpx_t *corr,
    sdr_cpx_t *C)
{
    corr_std_impl<true>(IQ, code, N, coff, pos, n, corr, C);
}

// src/sdr_func.c:1060-1072: float complex buffer -> cpx8 (int8 casts), mix,
// real code loaded into a cpx16 bank of one code (I = Q = (int8)code[i]),
// correlate with coff = 0.  The reference's mix_carr_cpx builds the cpx8
// buffer with SDR_CPX8 (src/pocket_sdr.h:106).
void sdr_corr_std_cpx(const sdr_cpx_t *buff, int len_buff, int ix, int N,
    double fs, double fc, double phi, const float *code, const double *pos,
    int n, sdr_cpx_t *corr)
{
    std::vector<uint8_t> b8(N);
    for (int i = 0, j = ix; i < N; i++, j = (j + 1) % len_buff) {
        int8_t re = (int8_t)buff[j][0], im = (int8_t)buff[j][1];
        b8[i] = (uint8_t)(((int8_t)im << 4) | (((int8_t)(re << 4) >> 4) & 0xF));
    }
    sdr_buff_t tmp;
    tmp.data = b8.data();
    tmp.IQ = 2;
    tmp.N = N;
    std::vector<sdr_cpx16_t> IQ((size_t)N * 2);
    sdr_mix_carr(&tmp, 0, N, fs, fc, phi, IQ.data());
    for (int i = 0; i < N; i++) IQ[N + i].I = IQ[N + i].Q = (int8_t)code[i];
    // the reference correlates against IQ + N as a bank; only bank k = 0 is
    // addressed because coff = 0 and integer pos give k = 0 there
    std::vector<sdr_cpx16_t> bank((size_t)N * PSA_N_CODES);
    for (int kk = 0; kk < PSA_N_CODES; kk++) {
        for (int i = 0; i < N; i++) {
            size_t src = (size_t)N + (size_t)kk * N + i;
            bank[(size_t)kk * N + i] = src < IQ.size() ? IQ[src] : sdr_cpx16_t{0, 0};
        }
    }
    sdr_cpx_t C[2];
    sdr_corr_std(IQ.data(), bank.data(), N, 0.0, pos, n, corr, C);
}

sdr_trk_plan_t *sdr_trk_plan_new(const sdr_buff_t *buff, int n_ch,
    const sdr_trk_ch_t *chs, int n_pos, const double *pos, int n_epoch)
{
    if (!buff || !buff->data || buff->N <= 0 || n_ch <= 0 || n_epoch <= 0 ||
        n_pos <= 0 || n_pos > PSA_MAX_CORR || !chs || !pos) {
        fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new: invalid arguments\n");
        return nullptr;
    }
    const int N = chs[0].N;
    if (N <= 0 || N > buff->N) {
        fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new: need 0 < N <= buffer length\n");
        return nullptr;
    }
    for (int c = 0; c < n_ch; c++) {
        if (chs[c].N != N || !chs[c].code) {
            fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new: channels need one N and a code bank\n");
            return nullptr;
        }
        for (size_t i = 0; i < (size_t)N * PSA_N_CODES; i++) {
            int a = chs[c].code[i].I, b = chs[c].code[i].Q;
            if (a < -1 || a > 1 || b < -1 || b > 1) {
                fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new: code values must be -1, 0 or 1\n");
                return nullptr;
            }
        }
    }
    // planar int8 I / Q for all tap geometries plus one word of slack
    const int Nq4 = ((N + 3) / 4 + 255) / 256 * 256;
    if ((size_t)Nq4 * 8 + 4096 > 150 * 1024) {
        fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new: N=%d too large for one workgroup\n", N);
        return nullptr;
    }
    ensure_init();
    sdr_trk_plan_t *p = new sdr_trk_plan_t();
    p->n_ch = n_ch; p->n_pos = n_pos; p->n_epoch = n_epoch; p->N = N; p->buffN = buff->N;
    // doubled bank plus the words the last (zero-data) step reads past it
    p->NP = (2 * N + 1024 + 16 + 3) / 4 * 4;
    p->sameq = true;
    for (int c = 0; c < n_ch && p->sameq; c++) {
        for (size_t i = 0; i < (size_t)N * PSA_N_CODES; i++) {
            if (chs[c].code[i].I != chs[c].code[i].Q) { p->sameq = false; break; }
        }
    }
    // exact sign semantics are needed only if some sample has a -8 nibble
    p->exact = false;
    for (int i = 0; i < buff->N && !p->exact; i++) {
        uint8_t d = buff->data[i];
        if ((d & 0xF) == 0x8 || (d >> 4) == 0x8) p->exact = true;
    }
    PSA_CHECK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    // one dword of padding (zero) past the ring for the mixer's aligned loads
    PSA_CHECK(hipMalloc(&p->d_buf, (size_t)buff->N + 8));
    PSA_CHECK(hipMemset(p->d_buf + buff->N, 0, 8));
    PSA_CHECK(hipMemcpy(p->d_buf, buff->data, buff->N, hipMemcpyHostToDevice));
    // code layout per channel: [copy 0..3][I banks 0..9 | Q banks 0..9][NP bytes],
    // each bank doubled (byte i holds bank sample i mod N, i < 2N); copy c holds
    // those bytes shifted by c, so word w of copy c = bytes 4w+c..4w+c+3
    const size_t bank_bytes = (size_t)2 * PSA_N_CODES * p->NP;
    p->code_copy_stride = (long long)(bank_bytes / 4);
    p->code_ch_stride = p->code_copy_stride * 4;
    std::vector<uint8_t> h((size_t)n_ch * 4 * bank_bytes, 0);
    for (int c = 0; c < n_ch; c++) {
        for (int cp = 0; cp < 4; cp++) {
            uint8_t *dst = h.data() + ((size_t)c * 4 + cp) * bank_bytes;
            for (int comp = 0; comp < 2; comp++) {
                for (int k = 0; k < PSA_N_CODES; k++) {
                    uint8_t *row = dst + ((size_t)comp * PSA_N_CODES + k) * p->NP;
                    const sdr_cpx16_t *src = chs[c].code + (size_t)k * N;
                    for (int i = 0; i < 2 * N - cp; i++) {
                        const sdr_cpx16_t &v = src[(i + cp) % N];
                        row[i] = (uint8_t)(comp == 0 ? v.I : v.Q);
                    }
                }
            }
        }
    }
    PSA_CHECK(hipMalloc(&p->d_codes, h.size()));
    PSA_CHECK(hipMemcpy(p->d_codes, h.data(), h.size(), hipMemcpyHostToDevice));
    std::vector<TrkChDev> hc(n_ch);
    for (int c = 0; c < n_ch; c++) {
        hc[c] = TrkChDev{chs[c].fs, chs[c].fi, chs[c].fc_sig, chs[c].T, chs[c].fd0,
                         chs[c].drift, chs[c].phi0, chs[c].coff0};
    }
    PSA_CHECK(hipMalloc(&p->d_ch, sizeof(TrkChDev) * n_ch));
    PSA_CHECK(hipMemcpy(p->d_ch, hc.data(), sizeof(TrkChDev) * n_ch, hipMemcpyHostToDevice));
    PSA_CHECK(hipMalloc(&p->d_pos, sizeof(double) * n_pos));
    PSA_CHECK(hipMemcpy(p->d_pos, pos, sizeof(double) * n_pos, hipMemcpyHostToDevice));
    const size_t ne = (size_t)n_ch * n_epoch;
    PSA_CHECK(hipMalloc(&p->d_ep, sizeof(TrkEpoch) * ne));
    PSA_CHECK(hipMalloc(&p->d_corr, sizeof(float2) * ne * n_pos));
    PSA_CHECK(hipMalloc(&p->d_C, sizeof(float2) * ne * 2));
    PSA_CHECK(hipMalloc(&p->d_P, sizeof(float2) * ne));
    for (auto &e : p->ev) PSA_CHECK(hipEventCreate(&e));
    return p;
}

sdr_trk_plan_t *sdr_trk_plan_new_loop(const sdr_buff_t *buff, int n_ch, const sdr_trk_ch_t *chs,
    const sdr_trk_loop_t *loops, int n_pos, const double *pos, int n_epoch)
{
    if (!loops || n_pos < TL_NTAP) {
        fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new_loop: need loop parameters and the P, E, L, N taps\n");
        return nullptr;
    }
    size_t n_sec = 0;
    for (int c = 0; c < n_ch; c++) {
        if (loops[c].len_code <= 0 || !(chs[c].T > 0.0) || (int)(TL_T_CN0 / chs[c].T) <= 0) {
            fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new_loop: need len_code > 0 and 0 < T <= 0.5 s\n");
            return nullptr;
        }
        if (loops[c].len_sec_code >= 2 && (!loops[c].sec_code || loops[c].len_sec_code > TL_N_HIST)) {
            fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new_loop: a secondary code needs its chips "
                            "(at most %d)\n", TL_N_HIST);
            return nullptr;
        }
        if (loops[c].bump_step > 0.0 && n_pos < 6) {
            fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new_loop: bump-jump needs the VE and VL taps (4, 5)\n");
            return nullptr;
        }
        if (loops[c].len_sec_code >= 2) n_sec += (size_t)loops[c].len_sec_code;
    }
    sdr_trk_plan_t *p = sdr_trk_plan_new(buff, n_ch, chs, n_pos, pos, n_epoch);
    if (!p) return nullptr;
    if (n_sec && (size_t)(((p->N + 3) / 4 + 255) / 256 * 256) * 8 + 4096 + sizeof(float2) * TL_N_HIST >
                     150 * 1024) {
        fprintf(stderr, "pocketsdr_amd: sdr_trk_plan_new_loop: N=%d too large for a secondary-code loop "
                        "(the IF planes and the prompt history share one workgroup's LDS)\n", p->N);
        sdr_trk_plan_free(p);
        return nullptr;
    }
    p->loop = true;
    p->lch0.assign(n_ch, TrkLoopCh{});
    for (int c = 0; c < n_ch; c++) {
        // start_track (src/sdr_ch.c:378-394) and trk_init (:364-377)
        TrkLoopCh &S = p->lch0[c];
        S.fs = chs[c].fs; S.fi = chs[c].fi; S.fc_sig = chs[c].fc_sig; S.T = chs[c].T;
        S.time = 0.0; S.fd = chs[c].fd0; S.coff = chs[c].coff0; S.phi = chs[c].phi0;
        S.cn0 = loops[c].cn0; S.pli = 1.0;
        S.len_code = loops[c].len_code; S.costas = loops[c].costas ? 1 : 0;
        S.bump_step = loops[c].bump_step > 0.0 ? loops[c].bump_step : 0.0;
    }
    if (n_sec) {
        // the channels' secondary codes, one after the other
        std::vector<int8_t> h(n_sec);
        size_t off = 0;
        for (int c = 0; c < n_ch; c++) {
            if (loops[c].len_sec_code < 2) continue;
            p->lch0[c].len_sec = loops[c].len_sec_code;
            p->lch0[c].sec_off = (int)off;
            memcpy(h.data() + off, loops[c].sec_code, (size_t)loops[c].len_sec_code);
            off += (size_t)loops[c].len_sec_code;
        }
        PSA_CHECK(hipMalloc(&p->d_sec, n_sec));
        PSA_CHECK(hipMemcpy(p->d_sec, h.data(), n_sec, hipMemcpyHostToDevice));
    }
    PSA_CHECK(hipMalloc(&p->d_lch, sizeof(TrkLoopCh) * n_ch));
    PSA_CHECK(hipMalloc(&p->d_lout, sizeof(TrkLoopOut) * (size_t)n_ch * n_epoch));
    // up to 8 chunks of at least 32 epochs (the all-tap pass of the last one
    // is what the run waits for after the loop)
    const int nchunk = std::max(1, std::min(8, n_epoch / 32));
    if (nchunk > 1) {
        if (n_sec) PSA_CHECK(hipMalloc(&p->d_hist, sizeof(float2) * TL_N_HIST * (size_t)n_ch));
    }
    PSA_CHECK(hipStreamCreateWithFlags(&p->st2, hipStreamNonBlocking));
    p->evc.resize(nchunk + 1);
    for (auto &e : p->evc) PSA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return p;
}

}  // extern "C"

template <bool EXACT, bool SAMEQ>
static void trk_launch_t(sdr_trk_plan_t *p, size_t shm, hipStream_t st, int e0, int nec)
{
    const unsigned ne = (unsigned)((size_t)p->n_ch * nec);
    const void *kern = (EXACT || !SAMEQ) ? (const void *)trk_corr_kernel<EXACT, SAMEQ>
                                         : (const void *)trk_corr_kernel_real;
    static std::once_flag once;
    std::call_once(once, [kern] {
        PSA_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    });
    if (EXACT || !SAMEQ) {
        hipLaunchKernelGGL((trk_corr_kernel<EXACT, SAMEQ>), dim3(ne), dim3(TRK_THREADS), shm, st,
                           p->d_buf, p->buffN, p->N, p->n_epoch, p->n_pos, p->d_ep, p->d_pos,
                           p->d_codes, p->code_ch_stride, p->code_copy_stride, p->NP,
                           dev_carrier(), p->d_corr, p->d_C, e0, nec);
    } else {
        hipLaunchKernelGGL(trk_corr_kernel_real, dim3(ne), dim3(TRK_THREADS), shm, st,
                           p->d_buf, p->buffN, p->N, p->n_epoch, p->n_pos, p->d_ep, p->d_pos,
                           p->d_codes, p->code_ch_stride, p->code_copy_stride, p->NP,
                           dev_carrier(), p->d_corr, p->d_C, e0, nec);
    }
    PSA_CHECK(hipGetLastError());
}

template <bool EXACT, bool SAMEQ>
static void trk_loop_launch_t(sdr_trk_plan_t *p, size_t shm, int e0, int e1)
{
    static std::once_flag once;
    constexpr int NTH = trk_loop_threads<EXACT, SAMEQ>();
    std::call_once(once, [] {
        PSA_CHECK(hipFuncSetAttribute((const void *)trk_loop_kernel<EXACT, SAMEQ, NTH>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    });
    hipLaunchKernelGGL((trk_loop_kernel<EXACT, SAMEQ, NTH>), dim3(p->n_ch), dim3(NTH), shm, p->st,
                       p->d_buf, p->buffN, p->N, p->n_epoch, p->n_pos, p->d_pos, p->d_codes,
                       p->code_ch_stride, p->code_copy_stride, p->NP, dev_carrier(), p->d_lch, p->d_lout,
                       p->d_ep, p->d_sec, e0, e1, p->d_hist);
}

static void trk_loop_launch(sdr_trk_plan_t *p, size_t shm, int e0, int e1)
{
    if (p->exact && p->sameq) trk_loop_launch_t<true, true>(p, shm, e0, e1);
    else if (p->exact) trk_loop_launch_t<true, false>(p, shm, e0, e1);
    else if (p->sameq) trk_loop_launch_t<false, true>(p, shm, e0, e1);
    else trk_loop_launch_t<false, false>(p, shm, e0, e1);
}

// the all-tap pass over epochs e0 .. e0 + nec - 1 of every channel, on st
static void trk_launch(sdr_trk_plan_t *p, size_t shm, hipStream_t st, int e0, int nec)
{
    if (p->exact && p->sameq) trk_launch_t<true, true>(p, shm, st, e0, nec);
    else if (p->exact) trk_launch_t<true, false>(p, shm, st, e0, nec);
    else if (p->sameq) trk_launch_t<false, true>(p, shm, st, e0, nec);
    else trk_launch_t<false, false>(p, shm, st, e0, nec);
}

extern "C" {

int sdr_trk_plan_run(sdr_trk_plan_t *p)
{
    if (!p) return 0;
    hipStream_t st = p->st;
    const size_t ne = (size_t)p->n_ch * p->n_epoch;
    if (p->loop) {
        // every run starts from the channels' initial states
        PSA_CHECK(hipMemcpyAsync(p->d_lch, p->lch0.data(), sizeof(TrkLoopCh) * p->n_ch, hipMemcpyHostToDevice,
                                 st));
        if (p->timing) PSA_CHECK(hipEventRecord(p->ev[0], st));
        const size_t shm = sizeof(uint32_t) * 2 * (size_t)(((p->N + 3) / 4 + 255) / 256 * 256);
        // The loop runs in chunks of epochs on st (each launch continues from
        // the channel states and prompt histories the previous one left);
        // the all-tap pass of a chunk (one workgroup per channel and epoch)
        // runs on st2 once its chunk's epoch records exist, on the CUs the
        // loop's few workgroups leave free, while the loop goes on.
        const int nchunk = (int)p->evc.size() - 1;   // (the last event joins st2 to st)
        const int csz = (p->n_epoch + nchunk - 1) / nchunk;
        for (int k = 0, e0 = 0; e0 < p->n_epoch; k++, e0 += csz) {
            const int e1 = std::min(p->n_epoch, e0 + csz);
            // (the loop kernel keeps the prompt histories after the planes)
            trk_loop_launch(p, shm + (p->d_sec ? sizeof(float2) * TL_N_HIST : 0), e0, e1);
            PSA_CHECK(hipGetLastError());
            PSA_CHECK(hipEventRecord(p->evc[k], st));
            PSA_CHECK(hipStreamWaitEvent(p->st2, p->evc[k], 0));
            trk_launch(p, shm, p->st2, e0, e1 - e0);   // all taps of the chunk's epochs
        }
        PSA_CHECK(hipEventRecord(p->evc[nchunk], p->st2));
        PSA_CHECK(hipStreamWaitEvent(st, p->evc[nchunk], 0));
        if (p->timing) PSA_CHECK(hipEventRecord(p->ev[1], st));
        PSA_CHECK(hipStreamSynchronize(st));
        if (p->timing) {
            float a = 0;
            PSA_CHECK(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
            p->ms_corr += a; p->n_corr++;
        }
        return 1;
    }
    if (p->timing) PSA_CHECK(hipEventRecord(p->ev[0], st));
    hipLaunchKernelGGL(trk_state_kernel, dim3(p->n_ch), dim3(TRK_SCHUNK), 0, st, p->d_ch, p->n_epoch, p->N,
                       p->buffN, p->d_ep);
    PSA_CHECK(hipGetLastError());
    if (p->timing) PSA_CHECK(hipEventRecord(p->ev[1], st));
    const size_t shm = sizeof(uint32_t) * 2 * (size_t)(((p->N + 3) / 4 + 255) / 256 * 256);
    trk_launch(p, shm, st, 0, p->n_epoch);
    PSA_CHECK(hipGetLastError());
    hipLaunchKernelGGL(trk_stitch_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, p->d_C,
                       (long long)ne, p->n_epoch, p->N, p->d_P);
    PSA_CHECK(hipGetLastError());
    if (p->timing) PSA_CHECK(hipEventRecord(p->ev[2], st));
    PSA_CHECK(hipStreamSynchronize(st));
    if (p->timing) {
        float a = 0, b = 0;
        PSA_CHECK(hipEventElapsedTime(&a, p->ev[0], p->ev[1]));
        PSA_CHECK(hipEventElapsedTime(&b, p->ev[1], p->ev[2]));
        p->ms_state += a; p->n_state++;
        p->ms_corr += b; p->n_corr++;
    }
    return 1;
}

int sdr_trk_plan_prompt(const sdr_trk_plan_t *p, sdr_cpx_t *P)
{
    if (!p || !P) return 0;
    const size_t ne = (size_t)p->n_ch * p->n_epoch;
    if (p->loop) {
        std::vector<TrkLoopOut> r(ne);
        PSA_CHECK(hipMemcpy(r.data(), p->d_lout, sizeof(TrkLoopOut) * ne, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < ne; i++) { P[i][0] = r[i].P.x; P[i][1] = r[i].P.y; }
        return 1;
    }
    PSA_CHECK(hipMemcpy(P, p->d_P, sizeof(float2) * ne, hipMemcpyDeviceToHost));
    return 1;
}

int sdr_trk_plan_state(const sdr_trk_plan_t *p, double *state)
{
    if (!p || !state || !p->loop) return 0;
    const size_t ne = (size_t)p->n_ch * p->n_epoch;
    std::vector<TrkLoopOut> r(ne);
    PSA_CHECK(hipMemcpy(r.data(), p->d_lout, sizeof(TrkLoopOut) * ne, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ne; i++) {
        double *q = state + 6 * i;
        q[0] = r[i].fd; q[1] = r[i].coff; q[2] = r[i].phi; q[3] = r[i].cn0;
        q[4] = r[i].lock; q[5] = r[i].idle;
    }
    return 1;
}

int sdr_trk_plan_result(const sdr_trk_plan_t *p, sdr_cpx_t *corr, sdr_cpx_t *C)
{
    if (!p) return 0;
    const size_t ne = (size_t)p->n_ch * p->n_epoch;
    if (corr) PSA_CHECK(hipMemcpy(corr, p->d_corr, sizeof(float2) * ne * p->n_pos, hipMemcpyDeviceToHost));
    if (C) PSA_CHECK(hipMemcpy(C, p->d_C, sizeof(float2) * ne * 2, hipMemcpyDeviceToHost));
    return 1;
}

void sdr_trk_plan_timing(sdr_trk_plan_t *p, int on)
{
    if (!p) return;
    p->timing = on != 0;
    p->ms_state = p->ms_corr = 0;
    p->n_state = p->n_corr = 0;
}

int sdr_trk_plan_kernel_ms(const sdr_trk_plan_t *p, const char *name, double *avg_ms,
                           long long *count, double *elems)
{
    if (!p || !name) return 0;
    std::string n(name);
    if (n == "trk_state" && p->n_state) {
        if (avg_ms) *avg_ms = p->ms_state / p->n_state;
        if (count) *count = p->n_state;
        if (elems) *elems = (double)p->n_ch * p->n_epoch;
        return 1;
    }
    if (n == "trk_corr" && p->n_corr) {
        if (avg_ms) *avg_ms = p->ms_corr / p->n_corr;
        if (count) *count = p->n_corr;
        if (elems) *elems = (double)p->n_ch * p->n_epoch * p->N;
        return 1;
    }
    return 0;
}

void sdr_trk_plan_free(sdr_trk_plan_t *p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->st);
    (void)hipFree(p->d_buf); (void)hipFree(p->d_codes); (void)hipFree(p->d_ch);
    (void)hipFree(p->d_pos); (void)hipFree(p->d_ep);
    (void)hipFree(p->d_corr); (void)hipFree(p->d_C); (void)hipFree(p->d_P);
    (void)hipFree(p->d_lch); (void)hipFree(p->d_lout); (void)hipFree(p->d_sec); (void)hipFree(p->d_hist);
    for (auto &e : p->ev) (void)hipEventDestroy(e);
    for (auto &e : p->evc) (void)hipEventDestroy(e);
    if (p->st2) (void)hipStreamDestroy(p->st2);
    (void)hipStreamDestroy(p->st);
    delete p;
}

}  // extern "C"
