// hgx_launch.hpp -- host launchers of the kernels (one per kernel family).
// The kernels are compiled in separate translation units (hgx_k_*.hip);
// the runtime calls these functions and holds no device code itself.
#pragma once
#include <hip/hip_runtime.h>

#include "hgx_types.hpp"

// column passes; WIDTH 256 or 512 picks the tile's thread count (512: 8
// rows per thread, blocks of at most 2^28 words), 0 the default rule
void hgx_launch_fwd_col(bool first, dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr,
                        int jt_bits, u64 L, int bbits, int kbase, int width = 0);
void hgx_launch_inv_col(dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr, int jt_bits,
                        u64 L, int bbits, int kbase, int width = 0);
// blocks of 2^cbits limbs (cbits 9 or 10)
void hgx_launch_crt(dim3 grid, hipStream_t s, const PrimeSet &ps, const JobRef &jr, int bpj_bits, int cbits, u64 L);
// row passes; each returns false when SBITS is out of range
bool hgx_launch_row_mul(bool single, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                        const JobRef &jr, int jt_bits);
// multiply row pass of a product (two) or square on 2^tb-word tiles (tb 10
// or 11), rows of at most 2^tb words
bool hgx_launch_row_half(bool two, int tb, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps,
                         const RowParams &rp, const JobRef &jr, int jt_bits);
bool hgx_launch_row_sqr(bool single, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                        const JobRef &jr, int jt_bits);
// whole small transforms (5 <= SBITS <= 10), 2^JBITS jobs per tile
// (SBITS + JBITS <= 12), NJOBS jobs in all: products and squares
bool hgx_launch_row_pack_mul(int sbits, int jbits, unsigned njobs, dim3 grid, hipStream_t s, const PrimeSet &ps,
                             const RowParams &rp, const JobRef &jr);
bool hgx_launch_row_pack_sqr(int sbits, int jbits, unsigned njobs, dim3 grid, hipStream_t s, const PrimeSet &ps,
                             const RowParams &rp, const JobRef &jr);
bool hgx_launch_row_tr(int mode, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                       const JobRef &jr, int jt_bits);
bool hgx_launch_row_inv_scaled(int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                               const JobRef &jr);
void hgx_launch_carry_map(dim3 grid, hipStream_t s, const JobRef &jr, int bpj_bits, u64 nruns, int rl);
void hgx_launch_carry_blocks(dim3 grid, hipStream_t s, const JobRef &jr, u64 nblocks);
void hgx_launch_carry_apply(dim3 grid, hipStream_t s, const JobRef &jr, int bpj_bits, u64 nruns, int rl);
// the three carry steps in one launch when a job has at most 1024 runs
void hgx_launch_carry_small(dim3 grid, hipStream_t s, const JobRef &jr, u64 nruns, int rl);
void hgx_launch_splitmix64(dim3 grid, hipStream_t s, u64 *out, u64 n, u64 seed, int force_top);
void hgx_launch_addsub(bool sub, hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *out, u64 n,
                       uint8_t *agg, uint8_t *blkin, u64 *carry_out);
void hgx_launch_fill(hipStream_t s, u64 *p, u64 n, u64 v);
// res[0] (zeroed first) = 1 + the top index where a and b differ (b null or
// shorter: zero-extended); 0 when equal.  cmp_sign then sets res[1] to the
// sign of a - b (1, 0 or 2^64 - 1).
void hgx_launch_top_diff(hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *res);
void hgx_launch_cmp_sign(hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, u64 *res);
// out[0..n] = a * 2^bits, 0 <= bits < 64 (out must not overlap a)
void hgx_launch_shl(hipStream_t s, const u64 *a, u64 n, int bits, u64 *out);
u64 hgx_addsub_blocks(u64 n);
// the signed sums of NJOBS jobs (hgx_sum.h) whose blk0 and n are set, NBLK
// blocks in all; AGG and BLKIN hold NBLK bytes each, RES 4 words per job
struct HgxSumJob;
void hgx_launch_sums(hipStream_t s, const HgxSumJob *jobs, u64 njobs, u64 nblk, uint8_t *agg, uint8_t *blkin,
                     u64 *res);
// the leaf folds of binary splitting (hgx_fold.hpp): BOUNDS holds each
// leaf's lo and up, CODES 4 ints per leaf; then each leaf's P, Q, B, R
// packed at DST + OFF[t]
struct FoldSeries;
struct FoldSlots;
void hgx_launch_fold(hipStream_t s, const FoldSeries &S, const u64 *bounds, u64 count, const FoldSlots &F, int *codes);
void hgx_launch_fold_pack(hipStream_t s, const FoldSlots &F, u64 count, const int *codes, const u64 *off, u64 *dst);
// register-only probe of one modular operation (0 Shoup product, 1
// Montgomery product, 2 sum and difference, 3 forward butterfly, 4 inverse
// butterfly), hgx_probe_chains() chains per thread, ITERS steps each;
// CLK[0..1] = shader cycles and 100 MHz ticks of workgroup 0's loop
bool hgx_launch_probe(int op, dim3 grid, hipStream_t s, const Mod &M, u64 w, u64 wq, u64 wm, int iters, u64 *out,
                      u64 *clk);
int hgx_probe_chains();
// a ROWS x COLS table of (SCALE R^(i j), companion) pairs; RM, SM, ONE in
// Montgomery form, PINV = p^-1 mod 2^64
void hgx_launch_pow_table(hipStream_t s, u64 *tab, u32 rows, u64 cols, u64 rm, u64 sm, u64 one, u64 pinv,
                          const Mod &M);
void hgx_launch_direct_mont(hipStream_t s, u64 *DM, const u64 *TS, const u64 *TT, int tbits, const Mod &M, u64 rm,
                            u64 rmq);
