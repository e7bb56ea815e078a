// hgx_k_col.hip -- column-pass, carry and input-generation kernels.
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"

// per-run maps and each block's combined map.
// grid = jobs << bpj_bits (bpj = blocks of 1024 runs per job)
__global__ __launch_bounds__(HGX_NT) void k_carry_map(JobRef jr, int bpj_bits, u64 nruns, int rl) {
  __shared__ u32 sh[HGX_NT];
  const int tid = threadIdx.x;
  const MulJob J = load_job(jr, blockIdx.x >> bpj_bits);
  const u64 blk = blockIdx.x & ((1u << bpj_bits) - 1);
  const u64 r0 = blk * HGX_CARRY_RUNS_PER_BLOCK + (u64)tid * HGX_CARRY_RUNS_PER_THREAD;
  u32 summary = 2u;
#pragma unroll
  for (int i = 0; i < HGX_CARRY_RUNS_PER_THREAD; i++) {
    const u64 r = r0 + i;
    if (r < nruns) {
      const u32 m = run_map(J.out, J.nout, J.ovf, J.flags, r, rl);
      J.maps[r] = (uint8_t)m;
      summary = map_combine(summary, m);
    }
  }
  sh[tid] = summary;
  __syncthreads();
  for (int s = 1; s < HGX_NT; s <<= 1) {
    if ((tid & (2 * s - 1)) == 0) sh[tid] = map_combine(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  if (tid == 0) J.agg[blk] = (uint8_t)sh[0];
}

// carry into every block of runs; one workgroup of 1024 threads per job
__global__ __launch_bounds__(1024) void k_carry_blocks(JobRef jr, u64 nblocks) {
  __shared__ u32 sh[1024];
  const int tid = threadIdx.x;
  const MulJob J = load_job(jr, blockIdx.x);
  const u64 chunk = (nblocks + 1023) / 1024;
  const u64 b0 = (u64)tid * chunk;
  u64 b1 = b0 + chunk;
  if (b1 > nblocks) b1 = nblocks;
  u32 summary = 2u;
  for (u64 b = b0; b < b1; b++) summary = map_combine(summary, J.agg[b]);
  u32 state = block_exclusive_last(summary, sh);
  for (u64 b = b0; b < b1; b++) {
    J.blkin[b] = (state == 3u) ? 1 : 0;
    state = map_combine(state, J.agg[b]);
  }
}

// add ovf[r-1] and the carry bit into every run (grid as k_carry_map)
__global__ __launch_bounds__(HGX_NT) void k_carry_apply(JobRef jr, int bpj_bits, u64 nruns, int rl) {
  __shared__ u32 sh[HGX_NT];
  const int tid = threadIdx.x;
  const MulJob J = load_job(jr, blockIdx.x >> bpj_bits);
  const u64 blk = blockIdx.x & ((1u << bpj_bits) - 1);
  const u64 r0 = blk * HGX_CARRY_RUNS_PER_BLOCK + (u64)tid * HGX_CARRY_RUNS_PER_THREAD;
  u32 summary = 2u;
#pragma unroll
  for (int i = 0; i < HGX_CARRY_RUNS_PER_THREAD; i++) {
    const u64 r = r0 + i;
    if (r < nruns) summary = map_combine(summary, J.maps[r]);
  }
  const u32 ex = block_exclusive_last(summary, sh);
  u64 e = (ex == 2u) ? (u64)J.blkin[blk] : (ex == 3u ? 1ull : 0ull);
  for (int i = 0; i < HGX_CARRY_RUNS_PER_THREAD; i++) {
    const u64 r = r0 + i;
    if (r >= nruns) break;
    u64 x0 = 0, x1 = 0;
    if (r > 0) {
      x0 = J.ovf[2 * (r - 1)];
      x1 = J.ovf[2 * (r - 1) + 1];
    }
    const u64 base = r * (u64)rl;
    if (x0 | x1 | e) {
      u64 carry = e;
      for (int j = 0; j < rl; j++) {
        const u64 pos = base + j;
        if (pos >= J.nout) break;
        const u64 add = (j == 0) ? x0 : (j == 1 ? x1 : 0);
        if (j >= 2 && carry == 0) break;
        const u64 d = J.out[pos];
        const u64 s = d + add;
        u64 cy = (s < d) ? 1 : 0;
        const u64 s2 = s + carry;
        cy += (s2 < s) ? 1 : 0;
        J.out[pos] = s2;
        carry = cy;
      }
    }
    e = (J.maps[r] >> e) & 1u;
  }
}

// SplitMix64 stream: out[i] = mix(seed + (i+1) * 0x9E3779B97F4A7C15)
__global__ void k_splitmix64(u64 *out, u64 n, u64 seed, int force_top) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  if (force_top && i == n - 1 && z == 0) z = 1;
  out[i] = z;
}

void hgx_launch_fwd_col(bool first, dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr,
                        int jt_bits, u64 L, int bbits, int kbase) {
  if (first)
    hipLaunchKernelGGL((k_fwd_col<true>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
  else
    hipLaunchKernelGGL((k_fwd_col<false>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
}

void hgx_launch_inv_col(bool last, dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr,
                        int jt_bits, u64 L, int bbits, int kbase) {
  if (last)
    hipLaunchKernelGGL((k_inv_col<true>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
  else
    hipLaunchKernelGGL((k_inv_col<false>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
}

void hgx_launch_carry_map(dim3 grid, hipStream_t s, const JobRef &jr, int bpj_bits, u64 nruns, int rl) {
  hipLaunchKernelGGL(k_carry_map, grid, dim3(HGX_NT), 0, s, jr, bpj_bits, nruns, rl);
}

void hgx_launch_carry_blocks(dim3 grid, hipStream_t s, const JobRef &jr, u64 nblocks) {
  hipLaunchKernelGGL(k_carry_blocks, grid, dim3(1024), 0, s, jr, nblocks);
}

void hgx_launch_carry_apply(dim3 grid, hipStream_t s, const JobRef &jr, int bpj_bits, u64 nruns, int rl) {
  hipLaunchKernelGGL(k_carry_apply, grid, dim3(HGX_NT), 0, s, jr, bpj_bits, nruns, rl);
}

void hgx_launch_splitmix64(dim3 grid, hipStream_t s, u64 *out, u64 n, u64 seed, int force_top) {
  hipLaunchKernelGGL(k_splitmix64, grid, dim3(256), 0, s, out, n, seed, force_top);
}
