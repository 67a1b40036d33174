// psa_mix.hip - carrier wipe-off (src/sdr_func.c:772-836) and the complex
// multiply of src/sdr_func.c:197-233 as standalone kernels.
#include <math.h>
#include "psa_internal.h"

namespace psa {

// Phase setup of mix_carr (src/sdr_func.c:776-778), including the split at
// the ring-buffer end (src/sdr_func.c:831-835) where the second segment
// restarts from phi + step * n, re-derived in double (the reference build
// fuses it: fma(n, step, phi)).
void mix_params(int buffN, int ix, int N, double fs, double fc, double phi,
                uint32_t *p0, uint32_t *p1, int *nseg, uint32_t *s)
{
    const double scale = (double)(1 << 24) * 256;
    double step = fc / fs;
    *p0 = (uint32_t)((phi - floor(phi)) * scale);
    *s = (uint32_t)(int)(step * scale);
    if (ix + N <= buffN) {
        *nseg = N;
        *p1 = 0;
    } else {
        int n = buffN - ix;
        double phi2 = fma((double)n, step, phi);
        *nseg = n;
        *p1 = (uint32_t)((phi2 - floor(phi2)) * scale);
    }
}

// Mixed sample = cpx8 * carrier(phase >> 24), each product component stored
// as int8 (the reference LUT entry, src/sdr_func.c:125-130).
__device__ __forceinline__ uint32_t mix_one(uint32_t d, uint32_t c)
{
    int sI = (int)((int8_t)(uint8_t)(d << 4)) >> 4;
    int sQ = (int)((int8_t)(uint8_t)d) >> 4;
    int cI = (int8_t)(c & 0xFF), cQ = (int8_t)(c >> 8);
    uint32_t mI = (uint8_t)(int8_t)(sI * cI - sQ * cQ);
    uint32_t mQ = (uint8_t)(int8_t)(sI * cQ + sQ * cI);
    return mI | (mQ << 8);
}

// Each thread mixes 4 consecutive samples and stores them as one 8-byte word.
__global__ void __launch_bounds__(256) mix_kernel(const uint8_t *__restrict__ buf, int buffN, int ix,
                                                  int N, uint32_t p0, uint32_t p1, int nseg, uint32_t s,
                                                  int compact, const uint32_t *__restrict__ carr_g,
                                                  int8_t *__restrict__ IQ)
{
    __shared__ uint32_t carr[256];
    carr[threadIdx.x] = carr_g[threadIdx.x];
    __syncthreads();
    int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= N) return;
    uint32_t out[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int i = i0 + q;
        out[q] = 0;
        if (i < N) {
            int src;
            uint32_t ph;
            if (i < nseg) { src = ix + i; ph = p0 + (uint32_t)i * s; }
            else { src = compact ? i : i - nseg; ph = p1 + (uint32_t)(i - nseg) * s; }
            (void)buffN;
            out[q] = mix_one(buf[src], carr[ph >> 24]);
        }
    }
    uint16_t *o = reinterpret_cast<uint16_t *>(IQ);
    if (i0 + 3 < N) {
        uint2 w = make_uint2(out[0] | (out[1] << 16), out[2] | (out[3] << 16));
        if (((uintptr_t)(o + i0) & 7) == 0) {
            *reinterpret_cast<uint2 *>(o + i0) = w;
            return;
        }
    }
    for (int q = 0; q < 4 && i0 + q < N; q++) o[i0 + q] = (uint16_t)out[q];
}

void launch_mix(const uint8_t *buf, int buffN, int ix, int N, uint32_t p0,
                uint32_t p1, int nseg, uint32_t s, int8_t *IQ, hipStream_t st, int compact)
{
    if (N <= 0) return;
    int nthr = (N + 3) / 4;
    hipLaunchKernelGGL(mix_kernel, dim3((nthr + 255) / 256), dim3(256), 0, st,
                       buf, buffN, ix, N, p0, p1, nseg, s, compact, dev_carrier(), IQ);
    PSA_CHECK(hipGetLastError());
}

// sdr_cpx_mul: the AVX2 lanes (src/sdr_func.c:203-213) compute
// re = fl(ar br) + fl(ai (-bi)), im = fl(ar bi) + fl(ai br), then * s; the
// scalar tail (N % 4 elements, src/sdr_func.c:229-232) is contracted by the
// reference build to fmaf(ar, br, -(ai bi)) and fmaf(ar, bi, ai br).
__global__ void __launch_bounds__(256) cpx_mul_kernel(const float2 *__restrict__ a,
                                                      const float2 *__restrict__ b, int N, int V,
                                                      float s, float2 *__restrict__ c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    // batch blockIdx.y: a and c advance by N, b is shared
    const long long o = (long long)blockIdx.y * N;
    float2 x = a[o + i], y = b[i];
    float re, im;
    if (i < V) {
        re = __fadd_rn(__fmul_rn(x.x, y.x), __fmul_rn(x.y, -y.y));
        im = __fadd_rn(__fmul_rn(x.x, y.y), __fmul_rn(x.y, y.x));
    } else {
        re = __fmaf_rn(x.x, y.x, -__fmul_rn(x.y, y.y));
        im = __fmaf_rn(x.x, y.y, __fmul_rn(x.y, y.x));
    }
    c[o + i] = make_float2(__fmul_rn(re, s), __fmul_rn(im, s));
}

void launch_cpx_mul(const float2 *a, const float2 *b, int N, float s,
                    float2 *c, hipStream_t st, int nbatch)
{
    if (N <= 0 || nbatch <= 0) return;
    if (nbatch > 65535) PSA_FAIL("cpx_mul: batch of %d exceeds the grid", nbatch);
    int V = N >= 4 ? N / 4 * 4 : 0;
    hipLaunchKernelGGL(cpx_mul_kernel, dim3((N + 255) / 256, nbatch), dim3(256), 0, st, a, b, N, V, s, c);
    PSA_CHECK(hipGetLastError());
}

}  // namespace psa
