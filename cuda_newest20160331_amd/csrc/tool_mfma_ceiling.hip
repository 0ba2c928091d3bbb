// tool_mfma_ceiling.hip -- the attainable v_mfma_f32_32x32x16_bf16 rate of this
// device, for the correlation kernel's roofline.  One wave per SIMD (one
// 256-thread workgroup per CU), sixteen independent 32x32 accumulators, three
// products per stage and a workgroup barrier every 48 MFMAs -- the issue pattern
// of the split-BF16 tile kernel without its memory traffic.  Run with operands
// that are constant and with operands that are random bf16 values of the size of
// normalised series entries: MFMA power, and so the clock the device holds,
// depends on the data.  The shader clock is read in the kernel (s_memtime against
// the 100 MHz s_memrealtime).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

static __device__ unsigned int mix(unsigned int x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <int RANDOM>
__global__ __launch_bounds__(256, 1) void ceiling_kernel(float *out, long long *clk, int iters)
{
    u32x4 f[8];
    for (int q = 0; q < 8; q++)
        for (int c = 0; c < 4; c++) {
            const unsigned int h = mix(threadIdx.x * 131u + blockIdx.x * 7919u + q * 17u + c);
            // bf16 pairs of magnitude about 2^-6..2^-5 (entries of a unit-norm series of ~1000 points)
            unsigned int lo = 0x3c80u, hi = 0x3c80u;
            if (RANDOM) {
                lo = 0x3c00u | (h & 0x80ffu);
                hi = 0x3c00u | ((h >> 16) & 0x80ffu);
            }
            f[q][c] = lo | (hi << 16);
        }
    const long long t0 = __builtin_amdgcn_s_memtime();
    const long long r0 = __builtin_amdgcn_s_memrealtime();
    f32x16_t acc[16];
    for (int q = 0; q < 16; q++)
        for (int e = 0; e < 16; e++) acc[q][e] = 0.f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 16; q++)
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8_t, f[(q >> 2) + (p == 2 ? 4 : 0)]),
                    __builtin_bit_cast(bf16x8_t, f[4 + (q & 3) - (p == 1 ? 4 : 0)]), acc[q], 0, 0, 0);
        __syncthreads();
    }
    float s = 0.f;
    for (int q = 0; q < 16; q++)
        for (int e = 0; e < 16; e++) s += acc[q][e];
    const long long t1 = __builtin_amdgcn_s_memtime();
    const long long r1 = __builtin_amdgcn_s_memrealtime();
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
    if (threadIdx.x == 0) {
        clk[2 * blockIdx.x] = t1 - t0;
        clk[2 * blockIdx.x + 1] = r1 - r0;
    }
}

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));            \
            return 1;                                                               \
        }                                                                           \
    } while (0)

template <int RANDOM>
static int run(int blocks, float *out, long long *clk)
{
    const int iters = 4000;
    CHECK(hipFuncSetAttribute((const void *)ceiling_kernel<RANDOM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              100 << 10));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int rep = 0; rep < 4; rep++) {
        CHECK(hipEventRecord(e0));
        // 100 KB of dynamic LDS keeps one workgroup per CU
        hipLaunchKernelGGL(ceiling_kernel<RANDOM>, dim3(blocks), dim3(256), 100 << 10, 0, out, clk, iters);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::vector<long long> h(2 * (size_t)blocks);
        CHECK(hipMemcpy(h.data(), clk, sizeof(long long) * 2 * blocks, hipMemcpyDeviceToHost));
        double cyc = 0, rt = 0;
        for (int b = 0; b < blocks; b++) {
            cyc += (double)h[2 * b];
            rt += (double)h[2 * b + 1];
        }
        const double flops = (double)blocks * 4 * iters * 48 * 32768.0;
        const double ghz = cyc / rt * 0.1;
        std::printf("%s operands, rep %d: %.3f ms, %.1f TFLOP/s, shader clock %.3f GHz\n",
                    RANDOM ? "random  " : "constant", rep, ms, flops / ms / 1e9, ghz);
    }
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    int ncu = 0;
    CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
    float *out = nullptr;
    long long *clk = nullptr;
    CHECK(hipMalloc(&out, sizeof(float) * 256 * ncu));
    CHECK(hipMalloc(&clk, sizeof(long long) * 2 * ncu));
    std::printf("v_mfma_f32_32x32x16_bf16, %d workgroups (one per CU), 16 accumulators per wave\n", ncu);
    if (run<0>(ncu, out, clk) || run<1>(ncu, out, clk) || run<0>(ncu, out, clk)) return 1;
    CHECK(hipFree(out));
    CHECK(hipFree(clk));
    return 0;
}
