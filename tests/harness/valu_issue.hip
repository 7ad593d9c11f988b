// Test harness (not product): issue cost of the VALU instruction kinds of the
// pair loop on one GPU, alone and mixed.  Every lane runs ITER iterations of
// a fixed group of 8 inline-asm instructions (8 independent register chains,
// VGPR operands only, as in the pair loop) over registers only: bounded
// loops, no memory access but one store per lane at the end; 8 workgroups
// of 4 waves per CU (8 waves per SIMD).  Prints the SIMD time per wave64
// instruction: ~1.9 ns (4 clocks at ~2.1 GHz) for an instruction that
// issues alone, less for kinds that two waves can issue in the same cycle.
//   usage: valu_issue
#include <hip/hip_runtime.h>
#include <stdio.h>

#define ITER 2048
#define FMA(a) asm volatile("v_fmac_f32_e32 %0, %1, %2" : "+v"(a) : "v"(x1), "v"(x2))
#define FMAS(a) asm volatile("v_fma_f32 %0, %1, %2, 0.5" : "+v"(a) : "s"(sx))
#define MUL(a) asm volatile("v_mul_f32_e32 %0, %0, %1" : "+v"(a) : "v"(x1))
#define CVT(a) asm volatile("v_cvt_f32_f64_e32 %0, %1" : "=v"(a) : "v"(dd))
#define ADD64(a) asm volatile("v_add_f64 %0, %1, -%0" : "+v"(a) : "v"(dd))
#define RSQ(a) asm volatile("v_rsq_f32_e32 %0, %0" : "+v"(a))
#define CMP() asm volatile("v_cmp_ngt_f32_e32 vcc, %0, %1" : : "v"(x1), "v"(x2) : "vcc")
#define MIN(a) asm volatile("v_min_f32_e32 %0, %0, %1" : "+v"(a) : "v"(x1))

template <int K>
__global__ __launch_bounds__(256) void kern(float *out, float s)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const float x1 = tid * 1e-7f + 1.0f, x2 = s, sx = s;
    float f0 = 1, f1 = 1, f2 = 1, f3 = 1, f4 = 1, f5 = 1, f6 = 1, f7 = 1;
    const double dd = tid * 1e-9 + 1.0;
    double g0 = 0, g1 = 0, g2 = 0, g3 = 0;
    for (int it = 0; it < ITER; ++it) {
        if constexpr (K == 0) { FMA(f0); FMA(f1); FMA(f2); FMA(f3); FMA(f4); FMA(f5); FMA(f6); FMA(f7); }
        if constexpr (K == 1) { FMAS(f0); FMAS(f1); FMAS(f2); FMAS(f3); FMAS(f4); FMAS(f5); FMAS(f6); FMAS(f7); }
        if constexpr (K == 2) { MUL(f0); MUL(f1); MUL(f2); MUL(f3); MUL(f4); MUL(f5); MUL(f6); MUL(f7); }
        if constexpr (K == 3) { MIN(f0); MIN(f1); MIN(f2); MIN(f3); MIN(f4); MIN(f5); MIN(f6); MIN(f7); }
        if constexpr (K == 4) { ADD64(g0); ADD64(g1); ADD64(g2); ADD64(g3); ADD64(g0); ADD64(g1); ADD64(g2); ADD64(g3); }
        if constexpr (K == 5) { CVT(f0); CVT(f1); CVT(f2); CVT(f3); CVT(f4); CVT(f5); CVT(f6); CVT(f7); }
        if constexpr (K == 6) { RSQ(f0); RSQ(f1); RSQ(f2); RSQ(f3); RSQ(f4); RSQ(f5); RSQ(f6); RSQ(f7); }
        if constexpr (K == 7) { CMP(); CMP(); CMP(); CMP(); CMP(); CMP(); CMP(); CMP(); }
        if constexpr (K == 8) { FMA(f0); ADD64(g0); FMA(f2); ADD64(g1); FMA(f4); ADD64(g2); FMA(f6); ADD64(g3); }
        if constexpr (K == 9) { FMA(f0); CVT(f1); FMA(f2); CVT(f3); FMA(f4); CVT(f5); FMA(f6); CVT(f7); }
        if constexpr (K == 10) { FMA(f0); RSQ(f1); FMA(f2); RSQ(f3); FMA(f4); RSQ(f5); FMA(f6); RSQ(f7); }
        if constexpr (K == 11) { FMA(f0); MUL(f1); FMA(f2); MUL(f3); FMA(f4); MUL(f5); FMA(f6); MUL(f7); }
    }
    out[tid] = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + (float)(g0 + g1 + g2 + g3);
}

template <int K>
static float run(float *d_out, int blocks)
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

static const char *kNames[] = {
    "v_fmac_f32 (VGPR operands)", "v_fma_f32 (SGPR operand)", "v_mul_f32", "v_min_f32",
    "v_add_f64", "v_cvt_f32_f64", "v_rsq_f32", "v_cmp_f32 -> vcc",
    "v_fmac_f32 + v_add_f64", "v_fmac_f32 + v_cvt_f32_f64", "v_fmac_f32 + v_rsq_f32",
    "v_fmac_f32 + v_mul_f32"};

template <int K>
static void one(float *d_out, int blocks)
{
    const float ms = run<K>(d_out, blocks);
    const double waves = blocks * 4.0;
    // SIMD-ns per wave instruction: time * SIMDs / (waves * instructions per wave)
    const double ns = ms * 1e6 * 1024.0 / (waves * 8.0 * ITER);
    printf("%-30s %8.3f ms  %.3f ns per wave64 instruction per SIMD\n", kNames[K], ms, ns);
}

int main()
{
    const int blocks = 256 * 8;   // 8 workgroups of 4 waves per CU
    float *d_out = nullptr;
    if (hipMalloc(&d_out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
    one<0>(d_out, blocks); one<1>(d_out, blocks); one<2>(d_out, blocks); one<3>(d_out, blocks);
    one<4>(d_out, blocks); one<5>(d_out, blocks); one<6>(d_out, blocks); one<7>(d_out, blocks);
    one<8>(d_out, blocks); one<9>(d_out, blocks); one<10>(d_out, blocks); one<11>(d_out, blocks);
    (void)hipFree(d_out);
    return 0;
}
