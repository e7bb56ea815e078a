// hgx_k_probe.hip -- register-only throughput probe of the library's own
// modular arithmetic (csrc/modarith.hpp): the integer-ALU ceiling that the
// transform kernels are measured against.  Every thread runs 8 independent
// chains of one operation on values in registers (no memory traffic inside
// the loop); thread 0 of workgroup 0 reads the shader cycle counter and the
// 100 MHz real-time counter around its loop, which gives the clock under
// that load.
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"

namespace {

constexpr int kChains = 8;

template <int OP>
__global__ __launch_bounds__(256) void k_probe(Mod M, u64 w, u64 wq, u64 wm, int iters, u64 *out, u64 *clk) {
  const u64 id = (u64)blockIdx.x * 256 + threadIdx.x;
  u64 x[kChains], y[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++) {
    x[c] = mod_reduce64(id * 0x9E3779B97F4A7C15ull + (u64)c * 0xBF58476D1CE4E5B9ull, M);
    y[c] = mod_reduce64(id * 0x94D049BB133111EBull + (u64)c * 0x2545F4914F6CDD1Dull, M);
  }
  const bool timer = blockIdx.x == 0 && threadIdx.x == 0;
  u64 c0 = 0, r0 = 0;
  if (timer) {
    c0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) {
      if constexpr (OP == 0) {  // Shoup product by a fixed factor
        x[c] = shoup_mul(x[c], w, wq, M);
      } else if constexpr (OP == 1) {  // Montgomery product
        x[c] = mont_mul(x[c], wm, M);
      } else if constexpr (OP == 2) {  // modular sum and difference
        const u64 u = x[c], v = y[c];
        x[c] = mod_add(u, v, M);
        y[c] = mod_sub(u, v, M);
      } else if constexpr (OP == 3) {  // forward (DIF) radix-2 butterfly, as round_regs
        const u64 u = x[c], v = y[c];
        x[c] = mod_add(u, v, M);
        y[c] = shoup_mul(u + (M.p - v), w, wq, M);
      } else {  // inverse (DIT) radix-2 butterfly
        const u64 u = x[c], v = shoup_mul(y[c], w, wq, M);
        x[c] = mod_add(u, v, M);
        y[c] = mod_sub(u, v, M);
      }
    }
  }
  if (timer) {
    clk[0] = __builtin_amdgcn_s_memtime() - c0;
    clk[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
  u64 s = 0;
#pragma unroll
  for (int c = 0; c < kChains; c++) s ^= x[c] ^ y[c];
  out[id] = s;
}

}  // namespace

int hgx_probe_chains() { return kChains; }

bool hgx_launch_probe(int op, dim3 grid, hipStream_t s, const Mod &M, u64 w, u64 wq, u64 wm, int iters, u64 *out,
                      u64 *clk) {
  switch (op) {
    case 0: hipLaunchKernelGGL(k_probe<0>, grid, dim3(256), 0, s, M, w, wq, wm, iters, out, clk); return true;
    case 1: hipLaunchKernelGGL(k_probe<1>, grid, dim3(256), 0, s, M, w, wq, wm, iters, out, clk); return true;
    case 2: hipLaunchKernelGGL(k_probe<2>, grid, dim3(256), 0, s, M, w, wq, wm, iters, out, clk); return true;
    case 3: hipLaunchKernelGGL(k_probe<3>, grid, dim3(256), 0, s, M, w, wq, wm, iters, out, clk); return true;
    case 4: hipLaunchKernelGGL(k_probe<4>, grid, dim3(256), 0, s, M, w, wq, wm, iters, out, clk); return true;
    default: return false;
  }
}
