// hgx_k_arith.hip -- magnitude addition and subtraction of device-resident
// limb vectors, the additions between the products of a Newton iteration
// (mpz-+ / mpz-- on magnitudes, src/mpz-ram.lisp:185-291).
//
// A limb's carry (or borrow) out is a map of its carry in: kill (0),
// propagate (2) or generate (3) -- the same last-non-propagate scan as the
// product's carry resolution.  Three launches: the composed map of every
// block of 1024 limbs, a scan over the blocks, and the final pass that
// recomputes each limb with its incoming carry.
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"
#include "hgx_sum.h"

namespace {

constexpr int kAsPerThread = 4;
constexpr int kAsBlock = HGX_NT * kAsPerThread;  // limbs per workgroup

__device__ __forceinline__ u64 limb(const u64 *x, u64 nx, u64 i) { return i < nx ? x[i] : 0; }

// raw limb result and its carry map
template <bool SUB>
__device__ __forceinline__ u64 limb_op(u64 x, u64 y, u32 &m) {
  if (SUB) {
    const u64 d = x - y;
    m = (x < y) ? 3u : (d == 0 ? 2u : 0u);
    return d;
  }
  const u64 s = x + y;
  m = (s < x) ? 3u : (s == ~0ull ? 2u : 0u);
  return s;
}

template <bool SUB>
__global__ __launch_bounds__(HGX_NT) void k_as_blocks(const u64 *a, u64 na, const u64 *b, u64 nb, u64 n, uint8_t *agg) {
  __shared__ u32 sh[HGX_NT];
  const u64 i0 = (u64)blockIdx.x * kAsBlock + (u64)threadIdx.x * kAsPerThread;
  u32 summary = 2u;
#pragma unroll
  for (int j = 0; j < kAsPerThread; j++) {
    const u64 i = i0 + j;
    if (i < n) {
      u32 m;
      (void)limb_op<SUB>(limb(a, na, i), limb(b, nb, i), m);
      summary = map_combine(summary, m);
    }
  }
  sh[threadIdx.x] = summary;
  __syncthreads();
  for (int s = 1; s < HGX_NT; s <<= 1) {
    if ((threadIdx.x & (2 * s - 1)) == 0) sh[threadIdx.x] = map_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) agg[blockIdx.x] = (uint8_t)sh[0];
}

// p[i] = v, i < n (graph-safe zeroing: no memset node)
__global__ __launch_bounds__(256) void k_fill(u64 *p, u64 n, u64 v) {
  const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// carry into every block; one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_as_scan(const uint8_t *agg, uint8_t *blkin, u64 nblocks) {
  __shared__ u32 sh[1024];
  const u64 chunk = (nblocks + 1023) / 1024;
  const u64 b0 = (u64)threadIdx.x * chunk;
  u64 b1 = b0 + chunk;
  if (b1 > nblocks) b1 = nblocks;
  u32 summary = 2u;
  for (u64 b = b0; b < b1; b++) summary = map_combine(summary, agg[b]);
  u32 state = block_exclusive_last(summary, sh);
  for (u64 b = b0; b < b1; b++) {
    blkin[b] = (state == 3u) ? 1 : 0;
    state = map_combine(state, agg[b]);
  }
}

// out[i] = a[i] +- b[i] +- carry in, i < n; an add writes its final carry to
// *carry_out when that is not null
template <bool SUB>
__global__ __launch_bounds__(HGX_NT) void k_as_apply(const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out, u64 n,
                                                      const uint8_t *blkin, u64 *carry_out) {
  __shared__ u32 sh[HGX_NT];
  const u64 i0 = (u64)blockIdx.x * kAsBlock + (u64)threadIdx.x * kAsPerThread;
  u64 r[kAsPerThread];
  u32 m[kAsPerThread];
  u32 summary = 2u;
#pragma unroll
  for (int j = 0; j < kAsPerThread; j++) {
    const u64 i = i0 + j;
    r[j] = 0;
    m[j] = 2u;
    if (i < n) {
      r[j] = limb_op<SUB>(limb(a, na, i), limb(b, nb, i), m[j]);
      summary = map_combine(summary, m[j]);
    }
  }
  const u32 ex = block_exclusive_last(summary, sh);
  u64 c = (ex == 3u) ? 1u : (ex == 2u ? (u64)blkin[blockIdx.x] : 0u);
#pragma unroll
  for (int j = 0; j < kAsPerThread; j++) {
    const u64 i = i0 + j;
    if (i < n) {
      out[i] = SUB ? r[j] - c : r[j] + c;
      c = (m[j] == 2u) ? c : (m[j] == 3u ? 1u : 0u);
      if (!SUB && i == n - 1 && carry_out) *carry_out = c;
    }
  }
}

// ---- scans of device limb vectors (sizes and comparisons read back) ----
constexpr unsigned kScanThreads = 256;
constexpr unsigned kScanMaxBlocks = 16384;

unsigned scan_blocks(u64 n) {
  const u64 b = (n + kScanThreads - 1) / kScanThreads;
  return (unsigned)(b < kScanMaxBlocks ? (b ? b : 1) : kScanMaxBlocks);
}

// the largest of the workgroup's values V, by thread 0
__device__ __forceinline__ u64 block_max(u64 v, u64 *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (unsigned s = kScanThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s && sh[threadIdx.x + s] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// *res = max(*res, 1 + the largest i < n with a[i] != b[i]) (b zero-extended
// beyond nb; b = null: compares with zero); *res starts at 0
__global__ __launch_bounds__(kScanThreads) void k_top_diff(const u64 *a, u64 na, const u64 *b, u64 nb, u64 n,
                                                           unsigned long long *res) {
  __shared__ u64 sh[kScanThreads];
  u64 best = 0;
  const u64 stride = (u64)gridDim.x * kScanThreads;
  for (u64 i = (u64)blockIdx.x * kScanThreads + threadIdx.x; i < n; i += stride) {
    const u64 x = i < na ? a[i] : 0, y = (b && i < nb) ? b[i] : 0;
    if (x != y) best = i + 1;
  }
  const u64 m = block_max(best, sh);
  if (threadIdx.x == 0 && m) atomicMax(res, (unsigned long long)m);
}

// res[1] = sign of a - b (magnitudes) from res[0] = 1 + the top differing limb
__global__ void k_cmp_sign(const u64 *a, u64 na, const u64 *b, u64 nb, unsigned long long *res) {
  const u64 t = res[0];
  if (t == 0) {
    res[1] = 0;
    return;
  }
  const u64 x = t - 1 < na ? a[t - 1] : 0, y = t - 1 < nb ? b[t - 1] : 0;
  res[1] = x > y ? 1ull : ~0ull;
}

// out[i] = bits [64 i - s, 64 i - s + 64) of a, i <= n: a * 2^s in n + 1 limbs
__global__ __launch_bounds__(256) void k_shl(const u64 *a, u64 n, int s, u64 *out) {
  const u64 stride = (u64)gridDim.x * 256;
  for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i <= n; i += stride) {
    const u64 lo = i ? a[i - 1] : 0, hi = i < n ? a[i] : 0;
    out[i] = s ? (hi << s) | (lo >> (64 - s)) : (i < n ? hi : 0);
  }
}

// ---- batched signed sums (HgxSumJob) ----
// The results' blocks of kAsBlock limbs are numbered across the jobs: job j
// owns blocks blk0 .. blk0 + ceil(n / kAsBlock) - 1.  Per job, RES holds
// four words: 1 + the top limb where |X| and |Y| differ (subtractions), 1 +
// the top non-zero limb of the sum, its sign (0, 1, or 2 for -1), and
// whether a subtraction takes |Y| - |X| (1) rather than |X| - |Y| (0).
constexpr int kSumWords = 4;

// the job of block GB: the last job whose first block is at most GB (found
// by thread 0, shared through SH[0])
__device__ __forceinline__ u64 sum_job(const HgxSumJob *jobs, u64 njobs, u64 gb, u64 *sh) {
  if (threadIdx.x == 0) {
    u64 lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const u64 mid = (lo + hi + 1) / 2;
      if (jobs[mid].blk0 <= gb)
        lo = mid;
      else
        hi = mid - 1;
    }
    sh[0] = lo;
  }
  __syncthreads();
  const u64 j = sh[0];
  __syncthreads();
  return j;
}

__device__ __forceinline__ bool sum_is_sub(const HgxSumJob &J) { return J.sx * J.sy < 0; }

// top differing limb of |X| and |Y| for the subtractions
__global__ __launch_bounds__(HGX_NT) void k_sums_top(const HgxSumJob *jobs, u64 njobs, unsigned long long *res) {
  __shared__ u64 sh[kScanThreads];
  const u64 j = sum_job(jobs, njobs, blockIdx.x, sh);
  const HgxSumJob J = jobs[j];
  if (!sum_is_sub(J)) return;  // the whole workgroup
  const u64 i0 = ((u64)blockIdx.x - J.blk0) * kAsBlock + (u64)threadIdx.x * kAsPerThread;
  u64 best = 0;
#pragma unroll
  for (int k = 0; k < kAsPerThread; k++) {
    const u64 i = i0 + k;
    if (i < J.n && limb(J.x, J.nx, i) != limb(J.y, J.ny, i)) best = i + 1;
  }
  const u64 m = block_max(best, sh);
  if (threadIdx.x == 0 && m) atomicMax(&res[kSumWords * j], (unsigned long long)m);
}

// the operands in the order of the operation: a subtraction takes the larger
// magnitude first (SWAP: |Y| > |X|)
struct SumOps {
  const u64 *a, *b;
  u64 na, nb;
  bool sub;
};
__device__ __forceinline__ SumOps sum_ops(const HgxSumJob &J, bool swap) {
  const u64 nx = J.sx ? J.nx : 0;
  SumOps o{J.x, J.y, nx, J.ny, sum_is_sub(J)};
  if (o.sub && swap) {
    o.a = J.y;
    o.b = J.x;
    o.na = J.ny;
    o.nb = nx;
  }
  return o;
}

// each block's carry map; the job's first block records the order
__global__ __launch_bounds__(HGX_NT) void k_sums_blocks(const HgxSumJob *jobs, u64 njobs, unsigned long long *res,
                                                        uint8_t *agg) {
  __shared__ u64 shj[kScanThreads];
  const u64 j = sum_job(jobs, njobs, blockIdx.x, shj);
  const HgxSumJob J = jobs[j];
  bool swap = false;
  if (sum_is_sub(J)) {
    const u64 t = res[kSumWords * j];  // 0: equal magnitudes
    if (t) swap = limb(J.x, J.nx, t - 1) < limb(J.y, J.ny, t - 1);
    if (threadIdx.x == 0 && (u64)blockIdx.x == J.blk0) res[kSumWords * j + 3] = swap ? 1 : 0;
  }
  const SumOps o = sum_ops(J, swap);
  u32 *sh = (u32 *)shj;
  const u64 i0 = ((u64)blockIdx.x - J.blk0) * kAsBlock + (u64)threadIdx.x * kAsPerThread;
  u32 summary = 2u;
#pragma unroll
  for (int k = 0; k < kAsPerThread; k++) {
    const u64 i = i0 + k;
    if (i < J.n) {
      u32 m;
      if (o.sub)
        (void)limb_op<true>(limb(o.a, o.na, i), limb(o.b, o.nb, i), m);
      else
        (void)limb_op<false>(limb(o.a, o.na, i), limb(o.b, o.nb, i), m);
      summary = map_combine(summary, m);
    }
  }
  sh[threadIdx.x] = summary;
  __syncthreads();
  for (int st = 1; st < HGX_NT; st <<= 1) {
    if ((threadIdx.x & (2 * st - 1)) == 0) sh[threadIdx.x] = map_combine(sh[threadIdx.x], sh[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) agg[blockIdx.x] = (uint8_t)sh[0];
}

// carry into every block of one job per workgroup
__global__ __launch_bounds__(HGX_NT) void k_sums_scan(const HgxSumJob *jobs, const uint8_t *agg, uint8_t *blkin) {
  __shared__ u32 sh[HGX_NT];
  const HgxSumJob J = jobs[blockIdx.x];
  const u64 nb = (J.n + kAsBlock - 1) / kAsBlock;
  const u64 chunk = (nb + HGX_NT - 1) / HGX_NT;
  const u64 b0 = (u64)threadIdx.x * chunk;
  const u64 b1 = b0 + chunk < nb ? b0 + chunk : nb;
  u32 summary = 2u;
  for (u64 b = b0; b < b1; b++) summary = map_combine(summary, agg[J.blk0 + b]);
  u32 state = block_exclusive_last(summary, sh);
  for (u64 b = b0; b < b1; b++) {
    blkin[J.blk0 + b] = (state == 3u) ? 1 : 0;
    state = map_combine(state, agg[J.blk0 + b]);
  }
}

// every limb with its incoming carry, written over X; the size and sign
__global__ __launch_bounds__(HGX_NT) void k_sums_apply(const HgxSumJob *jobs, u64 njobs, unsigned long long *res,
                                                       const uint8_t *blkin) {
  __shared__ u64 sh[kScanThreads];
  const u64 j = sum_job(jobs, njobs, blockIdx.x, sh);
  const HgxSumJob J = jobs[j];
  const bool swap = sum_is_sub(J) && res[kSumWords * j + 3] != 0;
  const SumOps o = sum_ops(J, swap);
  const u64 i0 = ((u64)blockIdx.x - J.blk0) * kAsBlock + (u64)threadIdx.x * kAsPerThread;
  u64 r[kAsPerThread];
  u32 m[kAsPerThread];
  u32 summary = 2u;
#pragma unroll
  for (int k = 0; k < kAsPerThread; k++) {
    const u64 i = i0 + k;
    r[k] = 0;
    m[k] = 2u;
    if (i < J.n) {
      if (o.sub)
        r[k] = limb_op<true>(limb(o.a, o.na, i), limb(o.b, o.nb, i), m[k]);
      else
        r[k] = limb_op<false>(limb(o.a, o.na, i), limb(o.b, o.nb, i), m[k]);
      summary = map_combine(summary, m[k]);
    }
  }
  const u32 ex = block_exclusive_last(summary, (u32 *)sh);
  u64 c = (ex == 3u) ? 1u : (ex == 2u ? (u64)blkin[blockIdx.x] : 0u);
  u64 best = 0;
#pragma unroll
  for (int k = 0; k < kAsPerThread; k++) {
    const u64 i = i0 + k;
    if (i < J.n) {
      const u64 v = o.sub ? r[k] - c : r[k] + c;
      J.x[i] = v;
      if (v) best = i + 1;
      c = (m[k] == 2u) ? c : (m[k] == 3u ? 1u : 0u);
    }
  }
  const u64 top = block_max(best, sh);
  if (threadIdx.x == 0) {
    if (top) atomicMax(&res[kSumWords * j + 1], (unsigned long long)top);
    if ((u64)blockIdx.x == J.blk0) {
      int sg = J.sy;  // a sum of magnitudes takes Y's sign (X's, when X is not zero)
      if (o.sub) sg = res[kSumWords * j] == 0 ? 0 : (swap ? J.sy : J.sx);
      res[kSumWords * j + 2] = sg > 0 ? 1u : (sg < 0 ? 2u : 0u);
    }
  }
}

}  // namespace

void hgx_launch_sums(hipStream_t s, const HgxSumJob *jobs, u64 njobs, u64 nblk, uint8_t *agg, uint8_t *blkin,
                     u64 *res) {
  const u64 nres = (u64)kSumWords * njobs;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((nres + 255) / 256)), dim3(256), 0, s, res, nres, (u64)0);
  unsigned long long *r = (unsigned long long *)res;
  hipLaunchKernelGGL(k_sums_top, dim3((unsigned)nblk), dim3(HGX_NT), 0, s, jobs, njobs, r);
  hipLaunchKernelGGL(k_sums_blocks, dim3((unsigned)nblk), dim3(HGX_NT), 0, s, jobs, njobs, r, agg);
  hipLaunchKernelGGL(k_sums_scan, dim3((unsigned)njobs), dim3(HGX_NT), 0, s, jobs, (const uint8_t *)agg, blkin);
  hipLaunchKernelGGL(k_sums_apply, dim3((unsigned)nblk), dim3(HGX_NT), 0, s, jobs, njobs, r,
                     (const uint8_t *)blkin);
}

void hgx_launch_top_diff(hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *res) {
  const u64 n = na > nb ? na : nb;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, res, (u64)2, (u64)0);
  if (n)
    hipLaunchKernelGGL(k_top_diff, dim3(scan_blocks(n)), dim3(kScanThreads), 0, s, a, na, b, nb, n,
                       (unsigned long long *)res);
}

void hgx_launch_cmp_sign(hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *res) {
  hipLaunchKernelGGL(k_cmp_sign, dim3(1), dim3(1), 0, s, a, na, b, nb, (unsigned long long *)res);
}

void hgx_launch_shl(hipStream_t s, const u64 *a, u64 n, int bits, u64 *out) {
  hipLaunchKernelGGL(k_shl, dim3(scan_blocks(n + 1)), dim3(256), 0, s, a, n, bits, out);
}

void hgx_launch_fill(hipStream_t s, u64 *p, u64 n, u64 v) {
  if (n) hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

void hgx_launch_addsub(bool sub, hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out, u64 n,
                       uint8_t *agg, uint8_t *blkin, u64 *carry_out) {
  const unsigned nblk = (unsigned)((n + kAsBlock - 1) / kAsBlock);
  if (sub) {
    hipLaunchKernelGGL(k_as_blocks<true>, dim3(nblk), dim3(HGX_NT), 0, s, a, na, b, nb, n, agg);
    hipLaunchKernelGGL(k_as_scan, dim3(1), dim3(1024), 0, s, agg, blkin, (u64)nblk);
    hipLaunchKernelGGL(k_as_apply<true>, dim3(nblk), dim3(HGX_NT), 0, s, a, na, b, nb, out, n, blkin, carry_out);
  } else {
    hipLaunchKernelGGL(k_as_blocks<false>, dim3(nblk), dim3(HGX_NT), 0, s, a, na, b, nb, n, agg);
    hipLaunchKernelGGL(k_as_scan, dim3(1), dim3(1024), 0, s, agg, blkin, (u64)nblk);
    hipLaunchKernelGGL(k_as_apply<false>, dim3(nblk), dim3(HGX_NT), 0, s, a, na, b, nb, out, n, blkin, carry_out);
  }
}

u64 hgx_addsub_blocks(u64 n) { return (n + kAsBlock - 1) / kAsBlock; }
