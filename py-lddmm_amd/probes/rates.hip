// Instruction-rate probe for the roofline of the pair kernels on gfx950:
// v_fma_f32, v_exp_f32, a 2:8 exp/fma mix, v_mfma_f32_16x16x4f32 alone and
// interleaved with VALU FMAs.  Every thread runs independent chains and
// stores its result, so nothing is dead code.
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o rates py-lddmm_amd/probes/rates.hip && ./rates
#include <hip/hip_runtime.h>
#include <cstdio>
typedef float v4f __attribute__((ext_vector_type(4)));
#define N_IT 4096
__global__ void k_fma(float* out, float s) {
  float a0=threadIdx.x*1e-3f,a1=a0+1,a2=a0+2,a3=a0+3,a4=a0+4,a5=a0+5,a6=a0+6,a7=a0+7;
  for (int i=0;i<N_IT;++i){ a0=fmaf(a0,s,0.5f);a1=fmaf(a1,s,0.5f);a2=fmaf(a2,s,0.5f);a3=fmaf(a3,s,0.5f);
    a4=fmaf(a4,s,0.5f);a5=fmaf(a5,s,0.5f);a6=fmaf(a6,s,0.5f);a7=fmaf(a7,s,0.5f);}
  out[blockIdx.x*blockDim.x+threadIdx.x]=a0+a1+a2+a3+a4+a5+a6+a7;
}
__global__ void k_exp(float* out, float s) {
  float a0=threadIdx.x*1e-6f,a1=a0+0.1f,a2=a0+0.2f,a3=a0+0.3f,a4=a0+0.4f,a5=a0+0.5f,a6=a0+0.6f,a7=a0+0.7f;
  for (int i=0;i<N_IT;++i){ a0=__builtin_amdgcn_exp2f(-a0);a1=__builtin_amdgcn_exp2f(-a1);a2=__builtin_amdgcn_exp2f(-a2);a3=__builtin_amdgcn_exp2f(-a3);
    a4=__builtin_amdgcn_exp2f(-a4);a5=__builtin_amdgcn_exp2f(-a5);a6=__builtin_amdgcn_exp2f(-a6);a7=__builtin_amdgcn_exp2f(-a7);}
  out[blockIdx.x*blockDim.x+threadIdx.x]=a0+a1+a2+a3+a4+a5+a6+a7;
}
// 1 exp per 4 fma (independent)
__global__ void k_mix(float* out, float s) {
  float e0=threadIdx.x*1e-6f,e1=e0+0.1f, f0=e0,f1=e0+1,f2=e0+2,f3=e0+3,f4=e0+4,f5=e0+5,f6=e0+6,f7=e0+7;
  for (int i=0;i<N_IT;++i){ e0=__builtin_amdgcn_exp2f(-e0); f0=fmaf(f0,s,0.5f);f1=fmaf(f1,s,0.5f);f2=fmaf(f2,s,0.5f);f3=fmaf(f3,s,0.5f);
    e1=__builtin_amdgcn_exp2f(-e1); f4=fmaf(f4,s,0.5f);f5=fmaf(f5,s,0.5f);f6=fmaf(f6,s,0.5f);f7=fmaf(f7,s,0.5f);}
  out[blockIdx.x*blockDim.x+threadIdx.x]=e0+e1+f0+f1+f2+f3+f4+f5+f6+f7;
}
__global__ void k_mfma(float* out, float s) {
  v4f c0={0,0,0,0},c1=c0,c2=c0,c3=c0; float a=threadIdx.x*1e-3f, b=a+1;
  for (int i=0;i<N_IT;++i){ c0=__builtin_amdgcn_mfma_f32_16x16x4f32(a,b,c0,0,0,0); c1=__builtin_amdgcn_mfma_f32_16x16x4f32(a,b,c1,0,0,0);
    c2=__builtin_amdgcn_mfma_f32_16x16x4f32(a,b,c2,0,0,0); c3=__builtin_amdgcn_mfma_f32_16x16x4f32(a,b,c3,0,0,0);}
  v4f c=c0+c1+c2+c3; out[blockIdx.x*blockDim.x+threadIdx.x]=c[0]+c[1]+c[2]+c[3];
}
// 1 MFMA per 8 VALU fma (independent chains)
__global__ void k_mfma_valu(float* out, float s) {
  v4f c0={0,0,0,0},c1=c0; float a=threadIdx.x*1e-3f, b=a+1;
  float f0=a,f1=a+1,f2=a+2,f3=a+3,f4=a+4,f5=a+5,f6=a+6,f7=a+7;
  for (int i=0;i<N_IT;++i){ c0=__builtin_amdgcn_mfma_f32_16x16x4f32(a,b,c0,0,0,0);
    f0=fmaf(f0,s,0.5f);f1=fmaf(f1,s,0.5f);f2=fmaf(f2,s,0.5f);f3=fmaf(f3,s,0.5f);
    f4=fmaf(f4,s,0.5f);f5=fmaf(f5,s,0.5f);f6=fmaf(f6,s,0.5f);f7=fmaf(f7,s,0.5f);
    c1=__builtin_amdgcn_mfma_f32_16x16x4f32(b,a,c1,0,0,0);
    f0=fmaf(f0,s,0.25f);f1=fmaf(f1,s,0.25f);f2=fmaf(f2,s,0.25f);f3=fmaf(f3,s,0.25f);
    f4=fmaf(f4,s,0.25f);f5=fmaf(f5,s,0.25f);f6=fmaf(f6,s,0.25f);f7=fmaf(f7,s,0.25f);}
  v4f c=c0+c1; out[blockIdx.x*blockDim.x+threadIdx.x]=c[0]+c[1]+c[2]+c[3]+f0+f1+f2+f3+f4+f5+f6+f7;
}
int main() {
  int ncu; hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0);
  int clk; hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, 0);
  int blocks = ncu * 8, threads = 256;  // 8 waves/SIMD
  float* out; hipMalloc(&out, sizeof(float)*blocks*threads);
  hipEvent_t e0,e1; hipEventCreate(&e0); hipEventCreate(&e1);
  double lanes = double(blocks)*threads;
  struct { const char* name; void(*k)(float*,float); double ops; } ks[] = {
    {"fma (8/it)", k_fma, 8.0*N_IT}, {"exp (8/it)", k_exp, 8.0*N_IT}, {"mix 2exp+8fma /it (count 10)", k_mix, 10.0*N_IT},
    {"mfma16x16x4 (4/it, per-lane-instr)", k_mfma, 4.0*N_IT}, {"mfma 2 + 16 fma /it (count fma)", k_mfma_valu, 16.0*N_IT}};
  printf("CUs %d clock %d kHz\n", ncu, clk);
  for (auto& k : ks) {
    for (int rep=0; rep<3; ++rep) {
      hipEventRecord(e0); hipLaunchKernelGGL(k.k, dim3(blocks), dim3(threads), 0, 0, out, 0.999f); hipEventRecord(e1);
      hipEventSynchronize(e1); float ms; hipEventElapsedTime(&ms, e0, e1);
      if (rep==2) printf("%-40s %8.3f ms  %.3e lane-ops/s  (%.3f of %d CU*64 lanes*clk)\n", k.name, ms,
                         lanes*k.ops/(ms*1e-3), lanes*k.ops/(ms*1e-3)/(double(ncu)*64*clk*1e3), ncu);
    }
  }
  return 0;
}
