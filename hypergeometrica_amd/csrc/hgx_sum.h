// hgx_sum.h -- the runtime's internal entry points for the binary-splitting
// engine (hgx_series.cpp): batched signed sums of device limb vectors (also
// shared with the kernels), the device memory left, and gathered uploads.
// The sum R = B_r Q_r R_l + B_l P_l R_r of COMBINE
// (src/binary-splitting.lisp:112-121) for every node of a tree height whose
// values stay in device memory: mpz-+ on signed values
// (src/mpz-ram.lisp:185-291), magnitudes added, or the smaller subtracted
// from the larger.
#pragma once
#include <cstddef>
#include <cstdint>

struct HgxSumJob {
  uint64_t *x;        // |X|, NX limbs, in a slot of at least max(NX, NY) + 1 limbs: receives the sum
  const uint64_t *y;  // |Y|, NY limbs (a different slot)
  uint64_t nx, ny;
  int sx, sy;     // signs (-1, 0, 1); SY != 0; SX = 0 reads X as zero
  uint64_t n;     // set by the runtime: limbs written, max(NX, NY) + 1
  uint64_t blk0;  // set by the runtime: the job's first block of the launch
};

// Every job's X slot receives SX|X| + SY|Y| (N limbs); SIGNS[j] its sign and
// SIZES[j] its limb count up to the top non-zero limb.  Queued on the
// library's stream after the work before it; returns when the sums are
// done.
int hgx_internal_sums_device(HgxSumJob *jobs, size_t count, int *signs, uint64_t *sizes);

// device memory the library may still allocate under its HBM budget
size_t hgx_internal_device_room();

// D_DST[...] = the COUNT host pieces SRC[k] (NLIMBS[k] limbs each, pageable
// memory) one after the other: gathered by copy threads into the library's
// pinned staging buffers, each moved by one DMA copy while the next fills.
// Returns when the copies are done.
int hgx_internal_upload_gather(uint64_t *d_dst, const uint64_t *const *src, const size_t *nlimbs, size_t count);

// The binary-splitting leaves folded on the device (hgx_fold.hpp): leaf t
// covers [BOUNDS[2t], BOUNDS[2t+1]), CAP holds the fold buffers'
// capacities.  On success CODES[4t + w] = sign * limbs of leaf t's P, Q, B,
// R (w = 0..3), and *D_VALS is a block of hgx_device_alloc memory (freed by
// hgx_device_free) that holds them leaf after leaf, P, Q, B, R each
// |CODES[...]| limbs, *NLIMBS in all.  HGX_ETOOBIG when a leaf outgrew its
// capacities, HGX_ENOMEM when device memory is short: then nothing is left
// allocated.  Returns when the values are in place.
struct FoldSeries;
int hgx_internal_fold_device(const FoldSeries *S, const uint64_t *bounds, size_t count, const uint32_t *cap,
                             int32_t *codes, uint64_t **d_vals, size_t *nlimbs);

// D_DST[0..N) = D_SRC[0..N), device to device on the library's stream;
// returns when the copy is done
int hgx_internal_copy_device(uint64_t *d_dst, const uint64_t *d_src, size_t n);
