// probe_valu_rates.hip -- standalone measurement program (not part of libmmo_ff.so).
//
// Issue rate of single VALU instruction forms on the GPU: each kernel runs 8
// independent chains of one form per lane (256 blocks per CU x 256 threads), so the
// rate is set by the SIMD's issue of that form alone.  The pair loops of the scoring
// kernels are bound by VALU issue; these rates say which operand forms to avoid there
// (an SGPR operand, a literal, a conversion, a transcendental).
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         mmo_amd/csrc/probe_valu_rates.hip -o probe_valu_rates && ./probe_valu_rates
//
// Per-dispatch instruction and cycle counts: run it under
//   rocprofv3 --pmc SQ_INSTS_VALU GRBM_GUI_ACTIVE -- ./probe_valu_rates
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kIters = 2048;

__device__ __forceinline__ float vreg(float x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ int vregi(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

#define PROBE(NAME, INIT, BODY, FINAL)                                      \
  __global__ void __launch_bounds__(256) NAME(float* out, float b, int ib) { \
    INIT;                                                                   \
    for (int it = 0; it < kIters; ++it) {                                   \
      _Pragma("unroll") for (int i = 0; i < 8; ++i) { BODY; }               \
    }                                                                       \
    float s = 0;                                                            \
    for (int i = 0; i < 8; ++i) s += FINAL;                                 \
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;                         \
  }

#define INIT_F float a[8]; for (int i = 0; i < 8; ++i) a[i] = threadIdx.x * 1e-3f + i + 1.0f
#define INIT_I int a[8]; for (int i = 0; i < 8; ++i) a[i] = threadIdx.x + i

// v_fmac_f32 v, v, v (all VGPR)
PROBE(k_fmac_v, INIT_F; float x[8]; for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 2e-3f - i,
      a[i] = fmaf(x[i], x[(i + 1) & 7], a[i]), a[i])
// v_mul_f32 v, v, v / v, s, v
PROBE(k_mul_v, INIT_F; const float bv = vreg(b), a[i] = a[i] * bv, a[i])
PROBE(k_mul_s, INIT_F, a[i] = a[i] * b, a[i])
// v_fma_f32 ... clamp with a VGPR / an SGPR operand
PROBE(k_fmaclamp_v, INIT_F; const float bv = vreg(b),
      a[i] = __builtin_amdgcn_fmed3f(fmaf(a[i], bv, 1.0f), 0.0f, 1.0f), a[i])
PROBE(k_fmaclamp_s, INIT_F, a[i] = __builtin_amdgcn_fmed3f(fmaf(a[i], b, 1.0f), 0.0f, 1.0f), a[i])
// v_sub_u32 v, v, v / v, s, v
PROBE(k_sub_v, INIT_I; const int iv = vregi(ib), a[i] = a[i] - iv; asm volatile("" : "+v"(a[i])),
      (float)a[i])
PROBE(k_sub_s, INIT_I, a[i] = a[i] - ib; asm volatile("" : "+v"(a[i])), (float)a[i])
// v_cvt_f32_i32
PROBE(k_cvt, INIT_I, a[i] = __float_as_int((float)a[i]), (float)a[i])
// v_rsq_f32
PROBE(k_rsq, INIT_F, a[i] = __builtin_amdgcn_rsqf(a[i]), a[i])

}  // namespace

int main() {
  const int blocks = 256 * 64, threads = 256;
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * blocks * threads) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  struct K {
    const char* name;
    void (*f)(float*, float, int);
  };
  const K ks[] = {{"fmac v,v,v", k_fmac_v},     {"mul v,v,v", k_mul_v},
                  {"mul v,s,v", k_mul_s},       {"fma v,v,v,1.0 clamp", k_fmaclamp_v},
                  {"fma v,v,s,1.0 clamp", k_fmaclamp_s}, {"sub_u32 v,v,v", k_sub_v},
                  {"sub_u32 v,s,v", k_sub_s},   {"cvt_f32_i32", k_cvt},
                  {"rsq_f32", k_rsq}};
  const double winstr = (double)blocks * threads / 64 * kIters * 8;
  printf("%-22s %10s %16s\n", "form", "ms", "wave-instr/ns");
  for (const K& k : ks) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      (void)hipEventRecord(e0);
      hipLaunchKernelGGL(k.f, dim3(blocks), dim3(threads), 0, 0, d, 0.999f, 3);
      (void)hipEventRecord(e1);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) best = ms;
    }
    printf("%-22s %10.3f %16.1f\n", k.name, best, winstr / (best * 1e6));
  }
  (void)hipFree(d);
  return 0;
}
