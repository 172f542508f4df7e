// VALU issue-rate probe for gfx950 (MI355X): the time a SIMD takes per wave64
// instruction of one kind when eight waves per SIMD issue them back to back.
// Each thread runs ITER rounds of 16 instructions (inline assembly on
// registers only: r_i = a op b, independent of each other, except the
// accumulating forms, r_i = a b + r_i, sixteen chains per wave);
// one launch of 8,192 waves is timed with HIP events.
//
//   hipcc --offload-arch=gfx950 -O3 -o issue_rate bench_issue_rate.hip && ./issue_rate
//
// The figure converts instruction counts (SQ_INSTS_VALU) into VALU issue
// time for the kernels' roofline notes.
#include <hip/hip_runtime.h>

#include <cstdio>

#define ITER 4096
#define OUTS                                                                                 \
  "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]), "+v"(r[6]),       \
      "+v"(r[7]), "+v"(r[8]), "+v"(r[9]), "+v"(r[10]), "+v"(r[11]), "+v"(r[12]), "+v"(r[13]), \
      "+v"(r[14]), "+v"(r[15])
#define I16(INS)                                                                           \
  INS " %0, %16, %17\n" INS " %1, %16, %17\n" INS " %2, %16, %17\n" INS " %3, %16, %17\n"   \
  INS " %4, %16, %17\n" INS " %5, %16, %17\n" INS " %6, %16, %17\n" INS " %7, %16, %17\n"   \
  INS " %8, %16, %17\n" INS " %9, %16, %17\n" INS " %10, %16, %17\n" INS " %11, %16, %17\n" \
  INS " %12, %16, %17\n" INS " %13, %16, %17\n" INS " %14, %16, %17\n" INS " %15, %16, %17\n"

#define F16(INS)                                                                                   \
  INS " %0, %16, %17, %17\n" INS " %1, %16, %17, %17\n" INS " %2, %16, %17, %17\n"                \
  INS " %3, %16, %17, %17\n" INS " %4, %16, %17, %17\n" INS " %5, %16, %17, %17\n"                \
  INS " %6, %16, %17, %17\n" INS " %7, %16, %17, %17\n" INS " %8, %16, %17, %17\n"                \
  INS " %9, %16, %17, %17\n" INS " %10, %16, %17, %17\n" INS " %11, %16, %17, %17\n"              \
  INS " %12, %16, %17, %17\n" INS " %13, %16, %17, %17\n" INS " %14, %16, %17, %17\n"             \
  INS " %15, %16, %17, %17\n"

constexpr int kKinds = 8;
const char* kNames[kKinds] = {"v_add_u32",     "v_mul_f32",    "v_fma_f32",    "v_fmac_f32 (acc)",
                              "v_mul_lo_u32",  "v_mul_u32_u24", "v_mul_hi_u32", "v_fma_f64 (acc)"};

template <int K>
__global__ void __launch_bounds__(256) k_issue(unsigned* sink, unsigned seed) {
  if constexpr (K < 7) {
    unsigned r[16];
    const unsigned a = threadIdx.x + seed, b = blockIdx.x | 1;
    for (int i = 0; i < 16; i++) r[i] = a + i;
    for (int it = 0; it < ITER; it++) {
      if constexpr (K == 0) asm volatile(I16("v_add_u32") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 1) asm volatile(I16("v_mul_f32") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 2) asm volatile(F16("v_fma_f32") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 3) asm volatile(I16("v_fmac_f32") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 4) asm volatile(I16("v_mul_lo_u32") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 5) asm volatile(I16("v_mul_u32_u24") : OUTS : "v"(a), "v"(b));
      if constexpr (K == 6) asm volatile(I16("v_mul_hi_u32") : OUTS : "v"(a), "v"(b));
    }
    unsigned s = 0;
    for (int i = 0; i < 16; i++) s ^= r[i];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
  } else {
    double r[16];
    const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * seed;
    for (int i = 0; i < 16; i++) r[i] = a + i;
    for (int it = 0; it < ITER; it++)
      asm volatile(
          "v_fma_f64 %0, %16, %17, %0\n v_fma_f64 %1, %16, %17, %1\n"
          "v_fma_f64 %2, %16, %17, %2\n v_fma_f64 %3, %16, %17, %3\n"
          "v_fma_f64 %4, %16, %17, %4\n v_fma_f64 %5, %16, %17, %5\n"
          "v_fma_f64 %6, %16, %17, %6\n v_fma_f64 %7, %16, %17, %7\n"
          "v_fma_f64 %8, %16, %17, %8\n v_fma_f64 %9, %16, %17, %9\n"
          "v_fma_f64 %10, %16, %17, %10\n v_fma_f64 %11, %16, %17, %11\n"
          "v_fma_f64 %12, %16, %17, %12\n v_fma_f64 %13, %16, %17, %13\n"
          "v_fma_f64 %14, %16, %17, %14\n v_fma_f64 %15, %16, %17, %15\n"
          : OUTS
          : "v"(a), "v"(b));
    double s = 0;
    for (int i = 0; i < 16; i++) s += r[i];
    sink[blockIdx.x * 256 + threadIdx.x] = (unsigned)(s > 0);
  }
}

int main() {
  // 256 CUs x 4 SIMDs x 8 waves: 8,192 waves in 2,048 blocks of 4 waves
  const int blocks = 2048, waves = blocks * 4;
  unsigned* d_sink = nullptr;
  if (hipMalloc(&d_sink, sizeof(unsigned) * blocks * 256) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const double per_simd = (double)waves * ITER * 16 / 1024.0;  // instructions per SIMD
  std::printf("%-18s %10s %s\n", "instruction", "ms", "ns per wave64 instruction per SIMD");
  for (int k = 0; k < kKinds; k++) {
    auto f = k == 0 ? k_issue<0> : k == 1 ? k_issue<1> : k == 2 ? k_issue<2> : k == 3 ? k_issue<3>
           : k == 4 ? k_issue<4> : k == 5 ? k_issue<5> : k == 6 ? k_issue<6> : k_issue<7>;
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      if (hipEventRecord(e0, 0) != hipSuccess) return 2;
      hipLaunchKernelGGL(f, dim3(blocks), dim3(256), 0, 0, d_sink, 1u + rep);
      if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 2;
      float ms = 0;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 3;
      if (ms < best) best = ms;
    }
    std::printf("%-18s %10.3f %.3f\n", kNames[k], best, best * 1e6 / per_simd);
  }
  return 0;
}
