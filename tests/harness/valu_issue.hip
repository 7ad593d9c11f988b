// Test harness (not product): issue cost of VALU instruction kinds on one
// GPU.  Every lane runs 8 independent dependency chains of one instruction
// kind for a fixed number of iterations (bounded loops, no memory access but
// one store per lane); 8 workgroups of 4 waves per CU.  Prints the time per
// (wave, chain element) per SIMD, i.e. the SIMD issue time of one wave64
// instruction of that kind when nothing else competes.
//   usage: valu_issue
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <vector>
#define ITER 4096
template <int K>
__global__ __launch_bounds__(256) void kern(float *out, float s)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    float f[8]; double d[8];
    for (int k = 0; k < 8; ++k) { f[k] = tid * 1e-7f + k; d[k] = tid * 1e-9 + k; }
    for (int it = 0; it < ITER; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if constexpr (K == 0) f[k] = f[k] * s + 0.5f;                    // v_fma_f32
            if constexpr (K == 1) f[k] = f[k] * s;                           // v_mul_f32
            if constexpr (K == 2) d[k] = d[k] + (double)s;                   // v_add_f64
            if constexpr (K == 3) d[k] = d[k] * (double)s + 0.5;             // v_fma_f64
            if constexpr (K == 4) f[k] = (float)(d[k] - (double)f[k]);       // add_f64 + cvt f64->f32 + cvt f32->f64
            if constexpr (K == 5) f[k] = __builtin_amdgcn_rsqf(f[k] + s);    // add + rsq
            if constexpr (K == 6) f[k] = (float)((double)f[k] + d[k]);       // cvt f32->f64, add_f64, cvt back
        }
    }
    float acc = 0;
    for (int k = 0; k < 8; ++k) acc += f[k] + (float)d[k];
    out[tid] = acc;
}
template <int K>
float run(float *d_out, int blocks)
{
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL(kern<K>, dim3(blocks), dim3(256), 0, 0, d_out, 0.999f);
    (void)hipEventRecord(a);
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(kern<K>, dim3(blocks), dim3(256), 0, 0, d_out, 0.999f);
    (void)hipEventRecord(b);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / 3;
}
int main()
{
    const int blocks = 256 * 8;   // 8 blocks of 4 waves per CU -> 8 waves per SIMD
    float *d_out = nullptr;
    if (hipMalloc(&d_out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
    const double waves = blocks * 4.0;
    const double instr = 8.0 * ITER;   // per wave, per loop kind (ops per chain element)
    const char *names[] = {"fma_f32", "mul_f32", "add_f64", "fma_f64", "sub_f64+cvt_f32+cvt_f64", "add_f32+rsq_f32", "cvt_f64+add_f64+cvt_f32"};
    float t[7];
    t[0] = run<0>(d_out, blocks); t[1] = run<1>(d_out, blocks); t[2] = run<2>(d_out, blocks);
    t[3] = run<3>(d_out, blocks); t[4] = run<4>(d_out, blocks); t[5] = run<5>(d_out, blocks);
    t[6] = run<6>(d_out, blocks);
    for (int k = 0; k < 7; ++k) {
        // SIMD-ns per wave-op group: time * SIMDs / (waves * ops per wave)
        const double ns_per = t[k] * 1e6 * 1024.0 / (waves * instr);
        printf("%-28s %8.3f ms  %.3f ns per (wave, chain element) per SIMD\n", names[k], t[k], ns_per);
    }
    (void)hipFree(d_out);
    return 0;
}
