// VALU issue-rate probe for gfx950: how many VALU wave-instructions per clock a CU issues for
// one instruction type, and for two types interleaved.  The projector kernels' issue ceiling
// depends on their instruction mix; this measures it.
//
// Each lane runs 16 independent chains (ops in inline asm, so the instruction is exactly the
// one named), 4096 iterations, 8 workgroups of 256 threads per CU (8 waves per SIMD); the
// chains of a pair alternate X, Y, X, Y...  Rate = wave-instructions / (time x CUs x clock).
// No memory is touched but the one result store per lane.
//
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 valu_issue.hip -o valu_issue   (build.py does this)
#include <hip/hip_runtime.h>
#include <cstdio>

#define CK(x)                                                                              \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);          \
      return 1;                                                                            \
    }                                                                                      \
  } while (0)

enum Op {
  FMA, ADD, SUB, MUL, FMA_CLAMP, ADD_U32, AND, ADD_SGPR,      // FP32 / integer arithmetic
  CVT_FLR, CVT_I32, CVT_F32, FLOOR, FRACT, MAX, MIN3, MED3,  // conversions, rounding, min/max
  LSHL_ADD, LSHLREV, CNDMASK,                                // shifts, select
  ADD_DPP, MOV_DPP, MAD_I24, MAD_U24, MUL_U24, RCP, NOPS
};
static const char* kName[NOPS] = {
    "fma", "add", "sub", "mul", "fma clamp", "add_u32", "and", "add (sgpr)",
    "cvt_flr_i32", "cvt_i32_f32", "cvt_f32_i32", "floor", "fract", "max", "min3", "med3",
    "lshl_add", "lshlrev", "cndmask", "add_f32_dpp", "mov_dpp", "mad_i32_i24", "mad_u32_u24",
    "mul_u32_u24", "rcp"};

template <int OP>
__device__ __forceinline__ void op(float& a, float c, float k, float ks) {
  if constexpr (OP == FMA) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == ADD) asm volatile("v_add_f32 %0, %0, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == SUB) asm volatile("v_sub_f32 %0, %0, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == MUL) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a) : "v"(c));
  if constexpr (OP == FMA_CLAMP) asm volatile("v_fma_f32 %0, %0, %1, %2 clamp" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == ADD_U32) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == AND) asm volatile("v_and_b32 %0, %0, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == ADD_SGPR) asm volatile("v_add_f32 %0, %1, %0" : "+v"(a) : "s"(ks));
  if constexpr (OP == CVT_FLR) asm volatile("v_cvt_flr_i32_f32 %0, %0" : "+v"(a));
  if constexpr (OP == CVT_I32) asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(a));
  if constexpr (OP == CVT_F32) asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(a));
  if constexpr (OP == FLOOR) asm volatile("v_floor_f32 %0, %0" : "+v"(a));
  if constexpr (OP == FRACT) asm volatile("v_fract_f32 %0, %0" : "+v"(a));
  if constexpr (OP == MAX) asm volatile("v_max_f32 %0, %0, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == MIN3) asm volatile("v_min3_f32 %0, %0, %1, %2" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == MED3) asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == LSHL_ADD) asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(a) : "v"(k));
  if constexpr (OP == LSHLREV) asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(a));
  if constexpr (OP == CNDMASK) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(a) : "v"(k));
  // DPP source: a loop-invariant register (no VALU write shortly before it)
  if constexpr (OP == ADD_DPP)
    asm volatile("v_add_f32_dpp %0, %1, %0 row_newbcast:3 row_mask:0xf bank_mask:0xf bound_ctrl:1" : "+v"(a) : "v"(k));
  if constexpr (OP == MOV_DPP)
    asm volatile("v_mov_b32_dpp %0, %1 row_newbcast:3 row_mask:0xf bank_mask:0xf bound_ctrl:1" : "=v"(a) : "v"(k));
  if constexpr (OP == MAD_I24) asm volatile("v_mad_i32_i24 %0, %0, %1, %2" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == MAD_U24) asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(a) : "v"(c), "v"(k));
  if constexpr (OP == MUL_U24) asm volatile("v_mul_u32_u24 %0, %0, %1" : "+v"(a) : "v"(c));
  if constexpr (OP == RCP) asm volatile("v_rcp_f32 %0, %0" : "+v"(a));
}

constexpr int ITERS = 4096;
constexpr int CHAINS = 16;

template <int X, int Y>
__global__ __launch_bounds__(256) void k_mix(float* out, float c, float k, float ks) {
  float a[CHAINS];
#pragma unroll
  for (int j = 0; j < CHAINS; ++j) a[j] = 1.5f + 0.01f * (float)(threadIdx.x + j);
  asm volatile("s_mov_b64 vcc, -1" ::: "vcc");
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int j = 0; j < CHAINS; j += 2) {
      op<X>(a[j], c, k, ks);
      op<Y>(a[j + 1], c, k, ks);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < CHAINS; ++j) s += a[j];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

struct Probe {
  int cus = 0, blocks = 0;
  double ghz = 0.0;
  float* out = nullptr;
  hipEvent_t e0{}, e1{};

  template <int X, int Y>
  int run() {
    auto launch = [&] { hipLaunchKernelGGL((k_mix<X, Y>), dim3(blocks), dim3(256), 0, 0, out, 0.999f, 1e-3f, 1e-3f); };
    launch();
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      CK(hipEventRecord(e0));
      launch();
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) best = ms;
    }
    const double winst = (double)blocks * 4 * ITERS * CHAINS;
    const double rate = winst / (best * 1e-3 * cus * ghz * 1e9);
    std::printf("| %s | %s | %.3f | %.3f |\n", kName[X], X == Y ? "" : kName[Y], best, rate);
    std::fflush(stdout);
    return 0;
  }
};

#define RUN(x, y) \
  if (p.run<x, y>()) return 1;

int main() {
  Probe p;
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  p.cus = prop.multiProcessorCount;
  p.ghz = prop.clockRate / 1e6;
  p.blocks = p.cus * 8;
  CK(hipMalloc(&p.out, (size_t)p.blocks * 256 * sizeof(float)));
  CK(hipEventCreate(&p.e0));
  CK(hipEventCreate(&p.e1));
  std::printf("%s: %d CUs at %.3f GHz; %d workgroups of 256, %d chains x %d iterations per lane\n\n",
              prop.gcnArchName, p.cus, p.ghz, p.blocks, CHAINS, ITERS);
  std::printf("| X | Y (alternating) | ms | VALU wave-instructions per CU-clock |\n|---|---|---|---|\n");
  // one type
  RUN(FMA, FMA) RUN(ADD, ADD) RUN(SUB, SUB) RUN(MUL, MUL) RUN(FMA_CLAMP, FMA_CLAMP) RUN(ADD_U32, ADD_U32)
  RUN(AND, AND) RUN(ADD_SGPR, ADD_SGPR) RUN(CVT_FLR, CVT_FLR) RUN(CVT_I32, CVT_I32) RUN(CVT_F32, CVT_F32)
  RUN(FLOOR, FLOOR) RUN(FRACT, FRACT) RUN(MAX, MAX) RUN(MIN3, MIN3) RUN(MED3, MED3) RUN(LSHL_ADD, LSHL_ADD)
  RUN(LSHLREV, LSHLREV) RUN(CNDMASK, CNDMASK) RUN(ADD_DPP, ADD_DPP) RUN(MOV_DPP, MOV_DPP)
  RUN(MAD_I24, MAD_I24) RUN(MAD_U24, MAD_U24) RUN(MUL_U24, MUL_U24) RUN(RCP, RCP)
  // FP32 FMA with each other type
  RUN(FMA, ADD) RUN(FMA, SUB) RUN(FMA, ADD_U32) RUN(FMA, AND) RUN(FMA, ADD_SGPR) RUN(FMA, CVT_FLR)
  RUN(FMA, CVT_I32) RUN(FMA, CVT_F32) RUN(FMA, FLOOR) RUN(FMA, FRACT) RUN(FMA, MAX) RUN(FMA, MIN3)
  RUN(FMA, MED3) RUN(FMA, LSHL_ADD) RUN(FMA, LSHLREV) RUN(FMA, CNDMASK) RUN(FMA, ADD_DPP)
  RUN(FMA, MOV_DPP) RUN(FMA, MAD_I24) RUN(FMA, MAD_U24) RUN(FMA, MUL_U24) RUN(FMA, RCP)
  // other pairs
  RUN(ADD, ADD_U32) RUN(ADD, MUL) RUN(SUB, CVT_FLR) RUN(SUB, FLOOR) RUN(SUB, MIN3) RUN(FMA_CLAMP, CVT_FLR)
  RUN(FMA_CLAMP, FLOOR) RUN(ADD_SGPR, CVT_FLR) RUN(CVT_FLR, FLOOR) RUN(CVT_FLR, LSHL_ADD) RUN(CVT_FLR, MAX)
  RUN(CVT_FLR, MIN3) RUN(ADD_U32, CVT_FLR) RUN(ADD_U32, MAD_I24) RUN(MIN3, ADD_U32) RUN(SUB, ADD_DPP)
  RUN(SUB, RCP)
  CK(hipFree(p.out));
  return 0;
}
