// hgx_k_fold.hip -- the leaf folds of binary splitting on the device: one
// thread per leaf (hgx_fold.hpp), then one pass that packs each leaf's
// P, Q, B, R from the strided fold buffers into consecutive limbs.
//
// A fold is a chain of 64 x 64-bit products and carries over values of a
// few dozen limbs: VALU work with strided, wave-coalesced loads and stores.
// Workgroups of one wave spread the leaves over every compute unit.
#include "hgx_fold.hpp"
#include "hgx_launch.hpp"

namespace {

constexpr int kFoldThreads = 64;

__global__ __launch_bounds__(kFoldThreads) void k_fold(FoldSeries S, const u64 *bounds, u64 count, FoldSlots F,
                                                       int *codes) {
  const u64 t = (u64)blockIdx.x * kFoldThreads + threadIdx.x;
  if (t >= count) return;
  int c[4];
  if (!fold_leaf(S, bounds[2 * t], bounds[2 * t + 1], t, F, c)) {
    c[0] = kFoldFailed;
    c[1] = c[2] = c[3] = 0;
  }
  int4 *out = (int4 *)codes;
  out[t] = make_int4(c[0], c[1], c[2], c[3]);
}

__global__ __launch_bounds__(kFoldThreads) void k_fold_pack(FoldSlots F, u64 count, const int *codes, const u64 *off,
                                                            u64 *dst) {
  const u64 t = (u64)blockIdx.x * kFoldThreads + threadIdx.x;
  if (t >= count) return;
  const int4 c4 = ((const int4 *)codes)[t];
  const int c[4] = {c4.x, c4.y, c4.z, c4.w};
  fold_pack_leaf(F, t, c, dst + off[t]);
}

}  // namespace

void hgx_launch_fold(hipStream_t s, const FoldSeries &S, const u64 *bounds, u64 count, const FoldSlots &F, int *codes) {
  const unsigned grid = (unsigned)((count + kFoldThreads - 1) / kFoldThreads);
  hipLaunchKernelGGL(k_fold, dim3(grid), dim3(kFoldThreads), 0, s, S, bounds, count, F, codes);
}

void hgx_launch_fold_pack(hipStream_t s, const FoldSlots &F, u64 count, const int *codes, const u64 *off, u64 *dst) {
  const unsigned grid = (unsigned)((count + kFoldThreads - 1) / kFoldThreads);
  hipLaunchKernelGGL(k_fold_pack, dim3(grid), dim3(kFoldThreads), 0, s, F, count, codes, off, dst);
}
