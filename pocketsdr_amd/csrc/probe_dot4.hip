// probe_dot4.hip - sustained issue rate of v_dot4_i32_i8 (the tracking
// correlator's MAC) and of v_fma_f32 on this GPU: 8 independent chains per
// lane, 4096 iterations, 4096 workgroups of 256 threads.  Not part of the
// library; built by `make -C pocketsdr_amd probe`.
#include <hip/hip_runtime.h>
#include <stdio.h>
__global__ void __launch_bounds__(256) k_dot4(int *out, int K, int seed)
{
    int a = threadIdx.x * 2654435761u + seed, b = blockIdx.x * 40503 + seed;
    int c0 = 0, c1 = 1, c2 = 2, c3 = 3, c4 = 4, c5 = 5, c6 = 6, c7 = 7;
    for (int i = 0; i < K; i++) {
        c0 = __builtin_amdgcn_sdot4(a, b, c0, false);
        c1 = __builtin_amdgcn_sdot4(a, b ^ 1, c1, false);
        c2 = __builtin_amdgcn_sdot4(a ^ 2, b, c2, false);
        c3 = __builtin_amdgcn_sdot4(a ^ 3, b ^ 3, c3, false);
        c4 = __builtin_amdgcn_sdot4(a ^ 4, b, c4, false);
        c5 = __builtin_amdgcn_sdot4(a, b ^ 5, c5, false);
        c6 = __builtin_amdgcn_sdot4(a ^ 6, b ^ 6, c6, false);
        c7 = __builtin_amdgcn_sdot4(a ^ 7, b ^ 7, c7, false);
        a += 0x01010101;
    }
    out[blockIdx.x * 256 + threadIdx.x] = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
}
__global__ void __launch_bounds__(256) k_fma(float *out, int K, float seed)
{
    float a = threadIdx.x * 1e-3f + seed, b = 0.999f;
    float c0 = 0, c1 = 1, c2 = 2, c3 = 3, c4 = 4, c5 = 5, c6 = 6, c7 = 7;
    for (int i = 0; i < K; i++) {
        c0 = __builtin_fmaf(a, b, c0); c1 = __builtin_fmaf(a, b, c1);
        c2 = __builtin_fmaf(a, b, c2); c3 = __builtin_fmaf(a, b, c3);
        c4 = __builtin_fmaf(a, b, c4); c5 = __builtin_fmaf(a, b, c5);
        c6 = __builtin_fmaf(a, b, c6); c7 = __builtin_fmaf(a, b, c7);
        a += 1e-7f;
    }
    out[blockIdx.x * 256 + threadIdx.x] = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
}
#define CK(x) do { if ((x) != hipSuccess) { fprintf(stderr, "probe_dot4: %s failed\n", #x); return 1; } } while (0)

int main()
{
    const int NB = 256 * 16, K = 4096;
    int *d;
    float *f;
    CK(hipMalloc(&d, sizeof(int) * NB * 256));
    CK(hipMalloc(&f, sizeof(float) * NB * 256));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const double n = (double)NB * 256 * K * 8;
    for (int rep = 0; rep < 3; rep++) {
        float ms;
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_dot4, dim3(NB), dim3(256), 0, 0, d, K, rep);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
        printf("dot4: %.3f ms, %.2f T lane-dot4/s = %.1f TOPS (4 int8 MACs per dot4)\n", ms, n / ms / 1e9,
               n * 8 / ms / 1e9);
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_fma, dim3(NB), dim3(256), 0, 0, f, K, (float)rep);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
        printf("fma : %.3f ms, %.2f T lane-fma/s = %.1f TFLOPS\n", ms, n / ms / 1e9, n * 2 / ms / 1e9);
    }
    CK(hipFree(d));
    CK(hipFree(f));
    return 0;
}
