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

// Exchange of 16 64-bit words per thread between the lanes of groups of 8
// (the transposition of an 8 x 8 block of words, twice), as a round of a row
// pass hands its values to the next: OP 5 through LDS (padded, one barrier),
// OP 6 inside the wave, by three xor stages of ds_swizzle (the LDS crossbar,
// no LDS memory) with selects on the VALU.  One add per word keeps the values
// live.  Counted in words exchanged.
__device__ __forceinline__ u64 swz_xor(u64 v, int d) {
  const int lo = (int)(u32)v, hi = (int)(u32)(v >> 32);
  int rl, rh;
  switch (d) {  // BitMode pattern: and 0x1f, or 0, xor d
    case 1: rl = __builtin_amdgcn_ds_swizzle(lo, (1 << 10) | 0x1f); rh = __builtin_amdgcn_ds_swizzle(hi, (1 << 10) | 0x1f); break;
    case 2: rl = __builtin_amdgcn_ds_swizzle(lo, (2 << 10) | 0x1f); rh = __builtin_amdgcn_ds_swizzle(hi, (2 << 10) | 0x1f); break;
    default: rl = __builtin_amdgcn_ds_swizzle(lo, (4 << 10) | 0x1f); rh = __builtin_amdgcn_ds_swizzle(hi, (4 << 10) | 0x1f); break;
  }
  return ((u64)(u32)rh << 32) | (u32)rl;
}

template <int D>
__device__ __forceinline__ void xor_stage(u64 *x, int lane) {
  const bool upper = (lane & D) != 0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j & D) continue;
    const u64 send = upper ? x[j] : x[j ^ D];
    const u64 recv = swz_xor(send, D);
    if (upper)
      x[j] = recv;
    else
      x[j ^ D] = recv;
  }
}

template <int OP>
__global__ __launch_bounds__(256) void k_exchange(Mod M, int iters, u64 *out, u64 *clk) {
  __shared__ u64 lds[OP == 5 ? 256 * 17 : 1];
  const u64 id = (u64)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 7, grp = threadIdx.x >> 3;
  u64 x[16];
#pragma unroll
  for (int c = 0; c < 16; c++) x[c] = mod_reduce64(id * 0x9E3779B97F4A7C15ull + (u64)c * 0xBF58476D1CE4E5B9ull, M);
  const bool timer = blockIdx.x == 0 && threadIdx.x == 0;
  u64 c0 = 0, r0 = 0;
  if (timer) {
    c0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (int i = 0; i < iters; i++) {
    if constexpr (OP == 5) {
      // word (lane, j) of the group's 8 x 8 block h to slot (j, lane)
#pragma unroll
      for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int j = 0; j < 8; j++) lds[HGX_PAD((grp * 2 + h) * 64 + lane * 8 + j)] = x[8 * h + j];
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int j = 0; j < 8; j++) x[8 * h + j] = lds[HGX_PAD((grp * 2 + h) * 64 + j * 8 + lane)];
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        xor_stage<4>(x + 8 * h, lane);
        xor_stage<2>(x + 8 * h, lane);
        xor_stage<1>(x + 8 * h, lane);
      }
    }
#pragma unroll
    for (int c = 0; c < 16; c++) x[c] += (u64)c;
  }
  if (timer) {
    clk[0] = __builtin_amdgcn_s_memtime() - c0;
    clk[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
  u64 sum = 0;
#pragma unroll
  for (int c = 0; c < 16; c++) sum ^= x[c];
  out[id] = sum;
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
    case 5: hipLaunchKernelGGL(k_exchange<5>, grid, dim3(256), 0, s, M, iters, out, clk); return true;
    case 6: hipLaunchKernelGGL(k_exchange<6>, grid, dim3(256), 0, s, M, iters, out, clk); return true;
    default: return false;
  }
}
