// hgx_types.hpp -- constants and the structs shared by the kernels and the
// host runtime (job descriptors, prime set, root-table views).
#pragma once
#include "modarith.hpp"


#define HGX_TILE 4096
#define HGX_NT 256
#define HGX_EPT 16  // elements of a full tile per thread
#define HGX_PAD(x) ((x) + ((x) >> 4))
#define HGX_TILE_LDS (HGX_TILE + HGX_TILE / 16)
#define HGX_HALF_TILE_BITS 11  // tiles of the low-latency multiply row pass (rows of at most 2^11 words)
#define HGX_RUN 4  // limbs per carry run: one run per thread of a CRT workgroup
#define HGX_CRT_BLOCK_BITS 10
#define HGX_CRT_BLOCK (1 << HGX_CRT_BLOCK_BITS)  // limbs per CRT workgroup (the unit of the global carry scan)


enum { MODE_MUL = 0, MODE_SQR = 1, MODE_FWD = 2, MODE_INV = 3 };

// HBM arrays as global-address-space pointers: accesses through them compile
// to global loads and stores (a uniform base in SGPRs plus a 32-bit VGPR
// offset where the offset fits) instead of flat accesses, which take 64-bit
// VGPR addresses and also count against the LDS wait counter.
// gat(base, i): word i of a global array whose offsets in bytes stay below
// 2^32 (i < 2^29), addressed as base + a 32-bit byte offset.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) u64 gu64;
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gu64 *gmem(u64 *p) { return (gu64 *)p; }
__device__ __forceinline__ const gu64 *gmem(const u64 *p) { return (const gu64 *)p; }
__device__ __forceinline__ gu64 &gat(gu64 *base, u32 i) { return *(gu64 *)((gchar *)base + (i << 3)); }
__device__ __forceinline__ const gu64 &gat(const gu64 *base, u32 i) {
  return *(const gu64 *)((const gchar *)base + (i << 3));
}
// pair i (16 bytes) of a global table of pairs
__device__ __forceinline__ ulonglong2 gat2(const gu64 *base, u32 i) {
  return *(const __attribute__((address_space(1))) ulonglong2 *)((const gchar *)base + (i << 4));
}
#else
typedef u64 gu64;
HGX_HD u64 *gmem(u64 *p) { return p; }
HGX_HD const u64 *gmem(const u64 *p) { return p; }
HGX_HD u64 &gat(u64 *base, u32 i) { return base[i]; }
HGX_HD const u64 &gat(const u64 *base, u32 i) { return base[i]; }
HGX_HD ulonglong2 gat2(const u64 *base, u32 i) { return *(const ulonglong2 *)(base + 2 * (size_t)i); }
#endif

// Moduli and in-tile root tables.  tw_fwd[k] holds the pairs
// (w_4096^i, floor(w_4096^i * 2^64 / p_k)) for i < 2048 -- the root and its
// Shoup companion, as the reference's roots table holds (w, w~)
// (src/moduli.lisp:321-332) -- and tw_inv[k] the same for w_4096^-i, with
// w_4096 = g^((p-1)/4096) (primitive-root-from-generator, src/moduli.lisp:182-187).
struct PrimeSet {
  u64 p[3];
  u64 np[3];  // 2^64 - p
  u32 c[3];
  const u64 *tw_fwd[3];
  const u64 *tw_inv[3];
  // Garner constants p0^-1 mod p1, p0^-1 mod p2, p1^-1 mod p2 with their
  // Shoup companions
  u64 g01, g02, g12;
  u64 g01q, g02q, g12q;
};

// Roots of one column pass on blocks of length B = 2^b (T = B/256 columns).
// Direct form (B <= 2^20): D[k] holds (w_B^(t q), Shoup companion) pairs at
// index q*T + t, read coalesced along t, one Shoup product per element.
// Direct Montgomery form (large blocks, 8 bytes per word of the block):
// DM[k] holds w_B^(t q) (times the pass's scale) in Montgomery form at
// index q*T + t, one REDC product per element.
// Tile-factored form: w_B^(t q) = w_B^(t0 q) * w_B^(tt q) for the tile's
// first column t0 = t - tt (tt < 16): TS[k] holds (w_B^(tt q), companion)
// pairs at q*16 + tt (times the pass's scale), TT[k] holds (w_B^(t0 q),
// companion) pairs at (t0/16)*256 + q; two Shoup products.
enum { TW_FACTORED = 0, TW_DIRECT = 1, TW_DIRECT_MONT = 2 };
struct ColTw {
  const u64 *D[3];
  const u64 *DM[3];
  const u64 *TS[3];
  const u64 *TT[3];
  int tbits;  // log2 T
};

__host__ __device__ inline Mod mod_of(const PrimeSet &ps, int k) { return Mod{ps.p[k], ps.np[k], ps.c[k]}; }

// One multiplication as the kernels see it: operands, product and work
// arrays in device memory.  A launch covers one job (passed by value) or a
// table of jobs of the same transform length (a batch); blockIdx.x selects
// job = blockIdx.x >> jt_bits and its tile = blockIdx.x & (2^jt_bits - 1).
struct MulJob {
  const u64 *a;
  const u64 *b;
  u64 na, nb;
  u64 *out;       // product limbs
  u64 nout;       // limbs to store (na + nb); positions at or above are zero
  u64 *res0;      // 3 L residues of operand a (prime k at + k L); the product spectrum lands here
  u64 *res1;      // 3 L residues of operand b
  u64 *ovf;       // 2 words per run: what the run's sum carries past its 16 limbs
  u32 *flags;     // per run: limbs 2..15 of the run are all ones
  uint8_t *maps;  // per run: carry map (kill / propagate / generate)
  uint8_t *agg;   // per block of 1024 runs: combined map
  uint8_t *blkin; // per block of 1024 runs: carry into the block
};

struct JobRef {
  MulJob one;
  const MulJob *table;  // null: every block works on ONE
};

// The job table does not change during a launch: read through the constant
// address space, the uniform job lands in SGPRs (scalar loads), and so do the
// array bases derived from it.
__device__ __forceinline__ MulJob load_job(const JobRef &jr, u32 j) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (jr.table) return ((const __attribute__((address_space(4))) MulJob *)jr.table)[j];
  return jr.one;
#else
  return jr.table ? jr.table[j] : jr.one;
#endif
}

// Parameters of a row pass (see k_row).
struct RowParams {
  u64 L;
  int tw_bits;
  int kbase;
  u64 scale[3];  // Montgomery-form factor applied to the pointwise product when SINGLE
};

#define HGX_CARRY_RUNS_PER_THREAD 4
#define HGX_CARRY_RUNS_PER_BLOCK (HGX_NT * HGX_CARRY_RUNS_PER_THREAD)
