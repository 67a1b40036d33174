// psa_raw.hip - front-end raw data to cpx8 IF samples: the receiver's
// write_buff (src/sdr_rcv.c:834-914) and its LUT generators
// (src/sdr_rcv.c:776-831) for the Pocket SDR / Spider SDR packed raw formats
// (RAW8 2 channels, RAW16 4, RAW16I 8, RAW32 8; 2-bit I/Q, 2- or 3-bit I and
// the fs/4 real-to-complex phases) and the int8 / int16 formats (INT8,
// INT8X2, CS8, CS16 through the AGC table), and the array combine that
// follows it in the receiver (below).  The optional per-channel LPF is not
// included.
//
// Each thread converts 16 consecutive samples of every channel: one vector
// load of 16 x (bytes per sample), table lookups in LDS (the int16 tables of
// CS16, 64 KiB each, are read through the cache), one 16-byte store per
// channel.  The work is a byte stream: bound by HBM.
#include <string.h>
#include <algorithm>
#include <vector>
#include "psa_internal.h"

namespace psa {

enum { F_INT8 = 1, F_INT8X2 = 2, F_RAW8 = 3, F_RAW16 = 4, F_RAW16I = 5, F_RAW32 = 6, F_CS8 = 7, F_CS16 = 8 };

static int raw_nch(int fmt)
{
    switch (fmt) {
    case F_RAW8: return 2;
    case F_RAW16: return 4;
    case F_RAW16I: case F_RAW32: return 8;
    case F_INT8: case F_INT8X2: case F_CS8: case F_CS16: return 1;
    default: return 0;
    }
}
static int raw_bps(int fmt)      // raw bytes per sample
{
    switch (fmt) {
    case F_INT8: case F_RAW8: return 1;
    case F_INT8X2: case F_CS8: case F_RAW16: case F_RAW16I: return 2;
    case F_RAW32: case F_CS16: return 4;
    default: return 0;
    }
}

// SDR_CPX8 (src/pocket_sdr.h:105) for components in [-8, 7]
__host__ __device__ __forceinline__ uint8_t mk_cpx8(int re, int im)
{
    return (uint8_t)((((unsigned)im & 0xF) << 4) | ((unsigned)re & 0xF));
}

struct RawArgs {
    const uint8_t *raw;
    long long N;
    int ph0;                    // phase of sample 0: (ix N) & 3
    int rtoc;                   // bit k: channel k takes the fs/4 phase tables
    const uint8_t *lutc;        // raw formats: [8][1024] cpx8 (phase p at p * 256)
    const int8_t *l8, *lq;      // int8 formats: I (and CS16 Q) tables
    uint8_t *out[8];
};

// cpx8 of channel k for raw byte `byte` at phase ph
template <int FMT>
__device__ __forceinline__ uint8_t raw_ch(const uint8_t *lut, int rtoc, int k, uint32_t byte, int ph)
{
    const int off = (FMT != F_RAW16I && ((rtoc >> k) & 1)) ? ph << 8 : 0;
    return lut[k * 1024 + off + byte];
}

template <int FMT>
__global__ void __launch_bounds__(256) raw_unpack_kernel(RawArgs a)
{
    constexpr int NCH = FMT == F_RAW8 ? 2 : FMT == F_RAW16 ? 4 : (FMT == F_RAW16I || FMT == F_RAW32) ? 8 : 1;
    constexpr int BPS = (FMT == F_INT8 || FMT == F_RAW8) ? 1
                        : (FMT == F_RAW32 || FMT == F_CS16) ? 4 : 2;
    constexpr bool RAWF = FMT >= F_RAW8 && FMT <= F_RAW32;
    constexpr int LUTB = RAWF ? 8 * 1024 : (FMT == F_CS16 ? 4 : 256);
    __shared__ __align__(16) uint8_t lut[LUTB];
    // table to LDS (CS16 reads its 64 KiB tables from global memory)
    if constexpr (RAWF) {
        for (int i = threadIdx.x; i < LUTB / 16; i += 256)
            reinterpret_cast<uint4 *>(lut)[i] = reinterpret_cast<const uint4 *>(a.lutc)[i];
    } else if constexpr (FMT != F_CS16) {
        lut[threadIdx.x] = (uint8_t)a.l8[threadIdx.x];
    }
    __syncthreads();
    const int8_t *l8s = reinterpret_cast<const int8_t *>(lut);
    const long long stride = (long long)gridDim.x * 256 * 16;
    for (long long s0 = ((long long)blockIdx.x * 256 + threadIdx.x) * 16; s0 < a.N; s0 += stride) {
        const int nq = (int)min(16LL, a.N - s0);
        // raw bytes of the 16 samples (a partial group reads byte by byte)
        uint32_t w[BPS * 4];
        if (nq == 16) {
#pragma unroll
            for (int v = 0; v < BPS; v++) {
                const uint4 q = reinterpret_cast<const uint4 *>(a.raw + s0 * BPS)[v];
                w[4 * v] = q.x; w[4 * v + 1] = q.y; w[4 * v + 2] = q.z; w[4 * v + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < BPS * 4; i++) w[i] = 0;
            for (int i = 0; i < nq * BPS; i++) w[i >> 2] |= (uint32_t)a.raw[s0 * BPS + i] << (8 * (i & 3));
        }
        auto rb = [&](int i) -> uint32_t { return (w[i >> 2] >> (8 * (i & 3))) & 0xFF; };   // raw byte i
        uint32_t o[NCH][4];
#pragma unroll
        for (int k = 0; k < NCH; k++) o[k][0] = o[k][1] = o[k][2] = o[k][3] = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int ph = (a.ph0 + (int)(s0 & 3) + q) & 3;
            uint32_t c[NCH];
            if constexpr (FMT == F_INT8) {
                const int I = l8s[rb(q)];
                const bool r = a.rtoc & 1;
                c[0] = !r || ph == 0 ? mk_cpx8(I, 0) : ph == 1 ? mk_cpx8(0, -I) : ph == 2 ? mk_cpx8(-I, 0)
                                                                                          : mk_cpx8(0, I);
            } else if constexpr (FMT == F_INT8X2) {
                c[0] = mk_cpx8(l8s[rb(2 * q)], -l8s[rb(2 * q + 1)]);
            } else if constexpr (FMT == F_CS8) {
                c[0] = mk_cpx8(l8s[rb(2 * q)], l8s[rb(2 * q + 1)]);
            } else if constexpr (FMT == F_CS16) {
                const uint32_t wi = w[q] & 0xFFFF, wq = w[q] >> 16;
                c[0] = mk_cpx8(a.l8[wi], a.lq[wq]);
            } else if constexpr (FMT == F_RAW8) {
                c[0] = raw_ch<FMT>(lut, a.rtoc, 0, rb(q), ph);
                c[1] = raw_ch<FMT>(lut, a.rtoc, 1, rb(q), ph);
            } else if constexpr (FMT == F_RAW16) {
#pragma unroll
                for (int k = 0; k < 4; k++) c[k] = raw_ch<FMT>(lut, a.rtoc, k, rb(2 * q + (k >> 1)), ph);
            } else if constexpr (FMT == F_RAW16I) {
#pragma unroll
                for (int k = 0; k < 8; k++) c[k] = raw_ch<FMT>(lut, a.rtoc, k, rb(2 * q + (k >> 2)), ph);
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++) c[k] = raw_ch<FMT>(lut, a.rtoc, k, rb(4 * q + (k >> 1)), ph);
            }
#pragma unroll
            for (int k = 0; k < NCH; k++) o[k][q >> 2] |= c[k] << (8 * (q & 3));
        }
        if (nq == 16) {
#pragma unroll
            for (int k = 0; k < NCH; k++)
                *reinterpret_cast<uint4 *>(a.out[k] + s0) = make_uint4(o[k][0], o[k][1], o[k][2], o[k][3]);
        } else {
#pragma unroll
            for (int k = 0; k < NCH; k++) {
                for (int q = 0; q < nq; q++) a.out[k][s0 + q] = (uint8_t)(o[k][q >> 2] >> (8 * (q & 3)));
            }
        }
    }
}

}  // namespace psa

using namespace psa;

struct sdr_raw_plan {
    int fmt = 0, nch = 0, rtoc = 0;
    uint8_t *lutc = nullptr;            // device [8][1024]
    int8_t *l8 = nullptr, *lq = nullptr;
    uint8_t *draw = nullptr, *dout = nullptr;   // staging of the host-pointer entry
    size_t raw_cap = 0, out_cap = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float last_ms = 0.0f;
};

namespace psa {

static void raw_launch(sdr_raw_plan_t *p, const uint8_t *raw, long long N, long long ix, uint8_t *const *out)
{
    RawArgs a = {};
    a.raw = raw; a.N = N;
    a.ph0 = (int)((ix * N) & 3);
    a.rtoc = p->rtoc;
    a.lutc = p->lutc; a.l8 = p->l8; a.lq = p->lq ? p->lq : p->l8;
    for (int k = 0; k < 8; k++) a.out[k] = k < p->nch ? out[k] : nullptr;
    const long long groups = (N + 15) / 16;
    const int blocks = (int)std::max(1LL, std::min((groups + 255) / 256, 8192LL));
    PSA_CHECK(hipEventRecord(p->e0, p->st));
    switch (p->fmt) {
#define PSA_RAW(F) case F: hipLaunchKernelGGL(raw_unpack_kernel<F>, dim3(blocks), dim3(256), 0, p->st, a); break;
    PSA_RAW(F_INT8) PSA_RAW(F_INT8X2) PSA_RAW(F_RAW8) PSA_RAW(F_RAW16) PSA_RAW(F_RAW16I) PSA_RAW(F_RAW32)
    PSA_RAW(F_CS8) PSA_RAW(F_CS16)
#undef PSA_RAW
    }
    PSA_CHECK(hipGetLastError());
    PSA_CHECK(hipEventRecord(p->e1, p->st));
}

}  // namespace psa

extern "C" {

// gen_LUTC / gen_LUTC_rtoc, src/sdr_rcv.c:776-799, 817-831
void sdr_raw_gen_lut(int fmt, int rfch, int IQ, int bits, int rtoc, sdr_cpx8_t *lut)
{
    static const int8_t vi2[] = {1, 3, -1, -3}, vq2[] = {-1, -3, 1, 3};
    static const int8_t vi3[] = {1, 3, 5, 7, -1, -3, -5, -7};
    for (int i = 0; i < 256; i++) {
        const int nib = (i >> (rfch % 2 ? 4 : 0)) & 0xF;
        if (rtoc) {
            const int I = vi3[((nib << 1) & 6) + ((nib >> 3) & 1)];
            lut[i] = mk_cpx8(I, 0);
            lut[256 + i] = mk_cpx8(0, -I);
            lut[512 + i] = mk_cpx8(-I, 0);
            lut[768 + i] = mk_cpx8(0, I);
            continue;
        }
        int I, Q = 0;
        if (fmt == F_RAW16I) I = vi2[(i >> ((rfch % 4) * 2)) & 3];
        else if (IQ == 1 && bits == 2) I = vi2[nib & 3];
        else if (IQ == 1 && bits == 3) I = vi3[((nib << 1) & 6) + ((nib >> 3) & 1)];
        else {
            I = vi2[nib & 3];
            Q = vq2[(i >> (rfch % 2 ? 6 : 2)) & 3];
        }
        lut[i] = mk_cpx8(I, Q);
    }
}

// gen_LUT8, src/sdr_rcv.c:802-814 (256 entries; 65536 for CS16)
void sdr_raw_gen_lut8(int fmt, double off, double scale, int8_t *lut)
{
    const int N = fmt == F_CS16 ? 65536 : 256;
    if (scale <= 0.0) scale = 1.0;
    for (int i = -N / 2; i < N / 2; i++) {
        const int val = (int)floor((i - off) / scale + 0.5);
        const int j = fmt == F_CS16 ? (uint16_t)i : (uint8_t)i;
        lut[j] = (int8_t)std::min(7, std::max(-7, val));
    }
}

sdr_raw_plan_t *sdr_raw_plan_new(int fmt, const int *rtoc, const void *const *lut)
{
    const int nch = raw_nch(fmt);
    if (!nch || !lut || !lut[0]) {
        fprintf(stderr, "pocketsdr_amd: sdr_raw_plan_new: bad format %d or tables\n", fmt);
        return nullptr;
    }
    ensure_init();
    sdr_raw_plan_t *p = new sdr_raw_plan_t();
    p->fmt = fmt;
    p->nch = nch;
    if (fmt >= F_RAW8 && fmt <= F_RAW32) {
        std::vector<uint8_t> h(8 * 1024, 0);
        for (int k = 0; k < nch; k++) {
            if (!lut[k]) {
                fprintf(stderr, "pocketsdr_amd: sdr_raw_plan_new: channel %d has no table\n", k);
                delete p;
                return nullptr;
            }
            const int r = fmt != F_RAW16I && rtoc && rtoc[k];
            p->rtoc |= r << k;
            memcpy(&h[k * 1024], lut[k], r ? 1024 : 256);
        }
        PSA_CHECK(hipMalloc(&p->lutc, h.size()));
        PSA_CHECK(hipMemcpy(p->lutc, h.data(), h.size(), hipMemcpyHostToDevice));
    } else {
        const size_t n = fmt == F_CS16 ? 65536 : 256;
        p->rtoc = fmt == F_INT8 && rtoc && rtoc[0] ? 1 : 0;
        PSA_CHECK(hipMalloc(&p->l8, n));
        PSA_CHECK(hipMemcpy(p->l8, lut[0], n, hipMemcpyHostToDevice));
        if (fmt == F_CS16 && lut[1]) {
            PSA_CHECK(hipMalloc(&p->lq, n));
            PSA_CHECK(hipMemcpy(p->lq, lut[1], n, hipMemcpyHostToDevice));
        }
    }
    PSA_CHECK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    PSA_CHECK(hipEventCreate(&p->e0));
    PSA_CHECK(hipEventCreate(&p->e1));
    return p;
}

int sdr_raw_plan_nch(const sdr_raw_plan_t *p) { return p ? p->nch : 0; }

int sdr_raw_unpack(sdr_raw_plan_t *p, const uint8_t *raw, int N, int64_t ix, sdr_cpx8_t *const *out)
{
    if (!p || !raw || !out || N < 0) return 0;
    if (N == 0) return 1;
    const size_t rb = (size_t)N * raw_bps(p->fmt), ob = (size_t)N * p->nch;
    if (rb > p->raw_cap) {
        (void)hipFree(p->draw);
        PSA_CHECK(hipMalloc(&p->draw, rb));
        p->raw_cap = rb;
    }
    if (ob > p->out_cap) {
        (void)hipFree(p->dout);
        PSA_CHECK(hipMalloc(&p->dout, ob));
        p->out_cap = ob;
    }
    PSA_CHECK(hipMemcpyAsync(p->draw, raw, rb, hipMemcpyHostToDevice, p->st));
    uint8_t *o[8];
    for (int k = 0; k < p->nch; k++) o[k] = p->dout + (size_t)k * N;
    raw_launch(p, p->draw, N, ix, o);
    for (int k = 0; k < p->nch; k++)
        PSA_CHECK(hipMemcpyAsync(out[k], o[k], N, hipMemcpyDeviceToHost, p->st));
    PSA_CHECK(hipStreamSynchronize(p->st));
    PSA_CHECK(hipEventElapsedTime(&p->last_ms, p->e0, p->e1));
    return 1;
}

int sdr_raw_unpack_dev(sdr_raw_plan_t *p, const uint8_t *raw, long long N, int64_t ix, sdr_cpx8_t *const *out)
{
    if (!p || !raw || !out || N < 0) return 0;
    if (((uintptr_t)raw & 15) != 0) {
        fprintf(stderr, "pocketsdr_amd: sdr_raw_unpack_dev: raw data must be 16-byte aligned\n");
        return 0;
    }
    for (int k = 0; k < p->nch; k++) {
        if (!out[k] || ((uintptr_t)out[k] & 15) != 0) {
            fprintf(stderr, "pocketsdr_amd: sdr_raw_unpack_dev: output %d must be 16-byte aligned\n", k);
            return 0;
        }
    }
    if (N == 0) return 1;
    raw_launch(p, raw, N, ix, out);
    PSA_CHECK(hipStreamSynchronize(p->st));
    PSA_CHECK(hipEventElapsedTime(&p->last_ms, p->e0, p->e1));
    return 1;
}

double sdr_raw_plan_kernel_ms(const sdr_raw_plan_t *p) { return p ? p->last_ms : 0.0; }

void sdr_raw_plan_free(sdr_raw_plan_t *p)
{
    if (!p) return;
    if (p->st) (void)hipStreamSynchronize(p->st);
    (void)hipFree(p->lutc); (void)hipFree(p->l8); (void)hipFree(p->lq);
    (void)hipFree(p->draw); (void)hipFree(p->dout);
    if (p->e0) (void)hipEventDestroy(p->e0);
    if (p->e1) (void)hipEventDestroy(p->e1);
    if (p->st) (void)hipStreamDestroy(p->st);
    delete p;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// array combine: sdr_arch_combine (src/sdr_array.c:553-579) with its weight
// table (build_lut, src/sdr_array.c:486-498) folded into the arithmetic: the
// table entry of channel j for cpx8 b is (I wr - Q wi, I wi + Q wr) with (I, Q)
// the sign-extended nibbles of b and (wr, wi) the channel's Q8 weights, so the
// products are formed directly (exact integers).  All beams of one call read
// each input sample once.
// ---------------------------------------------------------------------------
namespace psa {

struct ArchArgs {
    const uint8_t *in[8];
    long long N;
    int nrfch, narch;
    int16_t w[8][8][2];         // [beam][channel][re, im]
    uint8_t *out[8];
};

// one beam per grid row (blockIdx.y; its weights are wave-uniform), NRF RF
// channels unrolled; each thread combines 16 consecutive samples
template <int NRF>
__global__ void __launch_bounds__(256) arch_combine_kernel(ArchArgs a)
{
    const int m = blockIdx.y;
    // A sample's nibbles enter as u = nibble ^ 8 = value + 8 (0 .. 15), two
    // int16 lanes (uI, uQ) per channel: v_dot2 with (wr, -wi) gives
    // uI wr - uQ wi and with (wi, wr) uI wi + uQ wr; the offsets' share,
    // 8 sum (wr - wi) and 8 sum (wi + wr), is the accumulators' start value
    // (exact integers throughout).
    typedef short v2s __attribute__((ext_vector_type(2)));
    v2s w1[NRF], w2[NRF];
    int kI = 0, kQ = 0;
#pragma unroll
    for (int j = 0; j < NRF; j++) {
        const int wr = __builtin_amdgcn_readfirstlane((int)a.w[m][j][0]);
        const int wi = __builtin_amdgcn_readfirstlane((int)a.w[m][j][1]);
        w1[j] = v2s{(short)wr, (short)-wi};
        w2[j] = v2s{(short)wi, (short)wr};
        kI -= 8 * (wr - wi);
        kQ -= 8 * (wi + wr);
    }
    uint8_t *out = a.out[m];
    const long long stride = (long long)gridDim.x * 256 * 16;
    for (long long s0 = ((long long)blockIdx.x * 256 + threadIdx.x) * 16; s0 < a.N; s0 += stride) {
        const int nq = (int)min(16LL, a.N - s0);
        uint32_t x[NRF][4];
#pragma unroll
        for (int j = 0; j < NRF; j++) {
            if (nq == 16) {
                const uint4 v = *reinterpret_cast<const uint4 *>(a.in[j] + s0);
                x[j][0] = v.x; x[j][1] = v.y; x[j][2] = v.z; x[j][3] = v.w;
            } else {
                x[j][0] = x[j][1] = x[j][2] = x[j][3] = 0;
                for (int q = 0; q < nq; q++) x[j][q >> 2] |= (uint32_t)a.in[j][s0 + q] << (8 * (q & 3));
            }
        }
        // per byte: the I nibble ^ 8 (lo) and the Q nibble ^ 8 (hi)
        uint32_t lo[NRF][4], hi[NRF][4];
#pragma unroll
        for (int j = 0; j < NRF; j++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                lo[j][k] = (x[j][k] & 0x0F0F0F0Fu) ^ 0x08080808u;
                hi[j][k] = ((x[j][k] >> 4) & 0x0F0F0F0Fu) ^ 0x08080808u;
            }
        }
        uint32_t o[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 16; q++) {
            int32_t I = kI, Q = kQ;
            const int b = q & 3;
#pragma unroll
            for (int j = 0; j < NRF; j++) {
                // int16 lanes (uI, uQ) of sample q: byte b of lo, byte b of hi
                const uint32_t u = __builtin_amdgcn_perm(hi[j][q >> 2], lo[j][q >> 2],
                                                         0x0C040C00u + (uint32_t)b + ((uint32_t)b << 16));
                const v2s uv = __builtin_bit_cast(v2s, u);
                I = __builtin_amdgcn_sdot2(uv, w1[j], I, false);
                Q = __builtin_amdgcn_sdot2(uv, w2[j], Q, false);
            }
            // round half away from zero, divide by the Q8 scale (C division
            // truncates), clip to the cpx8 range
            I = (I + (I >= 0 ? 128 : -128)) / 256;
            Q = (Q + (Q >= 0 ? 128 : -128)) / 256;
            I = I < -8 ? -8 : (I > 7 ? 7 : I);
            Q = Q < -8 ? -8 : (Q > 7 ? 7 : Q);
            o[q >> 2] |= (uint32_t)mk_cpx8(I, Q) << (8 * (q & 3));
        }
        if (nq == 16) {
            *reinterpret_cast<uint4 *>(out + s0) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
            for (int q = 0; q < nq; q++) out[s0 + q] = (uint8_t)(o[q >> 2] >> (8 * (q & 3)));
        }
    }
}

static float arch_launch(int nrfch, const uint8_t *const *in, long long N, int narch, const int16_t *w,
                         uint8_t *const *out, hipStream_t st)
{
    ArchArgs a = {};
    for (int j = 0; j < nrfch; j++) a.in[j] = in[j];
    for (int m = 0; m < narch; m++) {
        a.out[m] = out[m];
        for (int j = 0; j < nrfch; j++) {
            a.w[m][j][0] = w[(m * nrfch + j) * 2];
            a.w[m][j][1] = w[(m * nrfch + j) * 2 + 1];
        }
    }
    a.N = N; a.nrfch = nrfch; a.narch = narch;
    const long long groups = (N + 15) / 16;
    const int blocks = (int)std::max(1LL, std::min((groups + 255) / 256, 4096LL));
    hipEvent_t e0, e1;
    PSA_CHECK(hipEventCreate(&e0));
    PSA_CHECK(hipEventCreate(&e1));
    PSA_CHECK(hipEventRecord(e0, st));
    const dim3 grid(blocks, narch);
    switch (nrfch) {
#define PSA_ARCH(R) case R: hipLaunchKernelGGL(arch_combine_kernel<R>, grid, dim3(256), 0, st, a); break;
    PSA_ARCH(2) PSA_ARCH(3) PSA_ARCH(4) PSA_ARCH(5) PSA_ARCH(6) PSA_ARCH(7) PSA_ARCH(8)
#undef PSA_ARCH
    }
    PSA_CHECK(hipGetLastError());
    PSA_CHECK(hipEventRecord(e1, st));
    PSA_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    PSA_CHECK(hipEventElapsedTime(&ms, e0, e1));
    PSA_CHECK(hipEventDestroy(e0));
    PSA_CHECK(hipEventDestroy(e1));
    return ms;
}

static bool arch_args_ok(int nrfch, int narch, const void *in, const int16_t *w, const void *out)
{
    if (nrfch < 2 || nrfch > 8 || narch < 1 || narch > 8 || !in || !w || !out) {
        fprintf(stderr, "pocketsdr_amd: sdr_arch_combine_raw: 2-8 RF channels and 1-8 beams\n");
        return false;
    }
    return true;
}

}  // namespace psa

extern "C" {

double sdr_arch_combine_raw_dev(int nrfch, const uint8_t *const *in, long long N, int narch, const int16_t *w,
                            uint8_t *const *out)
{
    if (!arch_args_ok(nrfch, narch, in, w, out) || N < 0) return -1.0;
    for (int j = 0; j < nrfch; j++) {
        if (!in[j] || ((uintptr_t)in[j] & 15)) {
            fprintf(stderr, "pocketsdr_amd: sdr_arch_combine_raw_dev: input %d must be 16-byte aligned\n", j);
            return -1.0;
        }
    }
    for (int m = 0; m < narch; m++) {
        if (!out[m] || ((uintptr_t)out[m] & 15)) {
            fprintf(stderr, "pocketsdr_amd: sdr_arch_combine_raw_dev: output %d must be 16-byte aligned\n", m);
            return -1.0;
        }
    }
    if (N == 0) return 0.0;
    ensure_init();
    return arch_launch(nrfch, in, N, narch, w, out, tctx().stream);
}

int sdr_arch_combine_raw(int nrfch, const uint8_t *const *in, int N, int narch, const int16_t *w, uint8_t *const *out)
{
    if (!arch_args_ok(nrfch, narch, in, w, out) || N < 0) return 0;
    if (N == 0) return 1;
    ensure_init();
    ThreadCtx &c = tctx();
    uint8_t *d = static_cast<uint8_t *>(c.s[SC_MISC].get((size_t)N * (nrfch + narch) + 16 * 16));
    const uint8_t *di[8];
    uint8_t *dout[8];
    size_t off = 0;
    for (int j = 0; j < nrfch; j++, off += ((size_t)N + 15) / 16 * 16) {
        di[j] = d + off;
        PSA_CHECK(hipMemcpyAsync(d + off, in[j], N, hipMemcpyHostToDevice, c.stream));
    }
    for (int m = 0; m < narch; m++, off += ((size_t)N + 15) / 16 * 16) dout[m] = d + off;
    arch_launch(nrfch, di, N, narch, w, dout, c.stream);
    for (int m = 0; m < narch; m++) PSA_CHECK(hipMemcpyAsync(out[m], dout[m], N, hipMemcpyDeviceToHost, c.stream));
    PSA_CHECK(hipStreamSynchronize(c.stream));
    return 1;
}

}  // extern "C"
