// hgx_kernels.hpp -- gfx950 kernels of the exact NTT multiply.
//
// Transform structure (equivalent to ntt-forward / ntt-reverse,
// src/number-theoretic-transform.lisp:40-143, same residues out):
//
//   A length-L decimation-in-frequency transform with bit-reversed output is
//   split into passes over HBM.  A "column pass" on blocks of length B treats
//   each block as a 256 x (B/256) matrix (row m, column t; element
//   t + (B/256) m), runs a 256-point DIF down every column (output row m'
//   holds frequency q = bitrev8(m')), and multiplies that element by
//   w_B^(t q).  Each row of the block is then a contiguous sub-block of
//   length B/256 whose own bit-reversed DIF finishes the transform, so the
//   next pass recurses on it; no transposes are needed.  The final "row
//   pass" transforms contiguous blocks of at most 4096 words in LDS.  The
//   inverse runs the same passes in reverse with w^-1 (decimation in time,
//   twiddles applied before each column transform).
//
//   Every pass moves one 4096-word tile (32 KiB) per workgroup through LDS:
//   a coalesced load, radix-16/8/4/2 rounds in registers with LDS exchanges,
//   and a coalesced store.  Column tiles are 16 adjacent columns x 256 rows
//   (16 consecutive words = 128 B per row segment); row tiles are 4096
//   consecutive words.
//
// Multiply pipeline (one L, three primes):
//   k_fwd_col<true>   limbs -> residues mod p_k for all k, first column pass
//   k_fwd_col         further column passes (in place)
//   k_row<MUL|SQR>    last forward pass of both operands, pointwise product,
//                     first inverse pass, per prime
//   k_inv_col         inverse column passes (in place); the outermost one
//                     folds in the 1/L scale
//   k_crt             CRT (Garner form) of every coefficient and the exact
//                     sum of every 16-limb run, with each run's overflow
//   k_carry_*         resolves the run-to-run carries (block scan)
//   For L <= 4096 one k_row<..., SINGLE> does all the transform work.
//   Launches too small to fill the GPU (fewer than 512 of these workgroups)
//   run k_fwd_col8 / k_inv_col8 (512 threads per column tile) and k_row_half
//   (2048- or 1024-word row tiles) instead, with the same results.
#pragma once
#include "hgx_types.hpp"

__device__ __forceinline__ u32 bitrev8(u32 m) { return __builtin_bitreverse32(m) >> 24; }

// x * w_B^(t q) for column t, frequency q: one Shoup product from the direct
// table, one REDC product from the direct Montgomery table, or two Shoup
// products from the tile-factored tables.
// The tables are indexed with 32-bit byte offsets: the direct tables hold
// 2^bbits pairs with bbits <= 22, the direct Montgomery tables 2^bbits words
// with bbits <= 28, the tile tables 2^(bbits-4) pairs with bbits <= 32.
template <int FORM>
__device__ __forceinline__ u64 apply_twiddle(u64 x, const ColTw &tw, int k, u32 t, u32 q, const Mod &M) {
  if constexpr (FORM == TW_DIRECT) {
    const ulonglong2 w = gat2(gmem(tw.D[k]), (q << tw.tbits) + t);
    return shoup_mul(x, w.x, w.y, M);
  } else if constexpr (FORM == TW_DIRECT_MONT) {
    return mont_mul(x, gat(gmem(tw.DM[k]), (q << tw.tbits) + t), M);
  } else {
    const ulonglong2 s = gat2(gmem(tw.TS[k]), (q << 4) + (t & 15));
    const ulonglong2 f = gat2(gmem(tw.TT[k]), ((t >> 4) << 8) + q);
    return shoup_mul(shoup_mul(x, s.x, s.y, M), f.x, f.y, M);
  }
}

// Tile-factored twiddle with the tile's 256 w_B^(t0 q) pairs staged in LDS
// (tts[q]); t0 is the same for every column of a tile.
__device__ __forceinline__ u64 apply_twiddle_tt(u64 x, const ColTw &tw, int k, u32 t, u32 q, const ulonglong2 *tts,
                                                const Mod &M) {
  const ulonglong2 s = gat2(gmem(tw.TS[k]), (q << 4) + (t & 15));
  const ulonglong2 f = tts[q];
  return shoup_mul(shoup_mul(x, s.x, s.y, M), f.x, f.y, M);
}

// ---------------------------------------------------------------------------
// In-tile rounds.  A round does R radix-2 stages (A .. A+R-1 of S_BITS) on
// groups of 2^R elements spaced D = S >> (A+R) rows apart, one group per
// thread iteration, values in registers.
// ---------------------------------------------------------------------------
// RMAX: the largest radix exponent of a round (4: radix-16 rounds; 3:
// radix-8 rounds, for tiles of 2048 words on 256 threads).
template <int S_BITS, int RMAX = 4>
struct RoundPlan {
  static constexpr int NR = (S_BITS + RMAX - 1) / RMAX;
  static constexpr int bits(int i) { return S_BITS / NR + (i < S_BITS % NR ? 1 : 0); }
  static constexpr int start(int i) {
    int a = 0;
    for (int k = 0; k < i; k++) a += bits(k);
    return a;
  }
};

template <int V>
struct Log2 { static constexpr int value = 1 + Log2<V / 2>::value; };
template <>
struct Log2<1> { static constexpr int value = 0; };

// A scheduling fence after every radix-2 stage keeps the twiddle loads of
// one stage (at most 8 pairs) in registers at a time.
#if defined(HGX_NO_STAGE_FENCE)
#define HGX_STAGE_FENCE() ((void)0)
#else
#define HGX_STAGE_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif

// Butterflies of one round on the 2^R values of one group, in registers:
// x[j] is element base + j*D of the group, and goff = base mod D selects the
// twiddles (w_2h^e from the w_4096 table, e = goff + i*D).
// ZHI (forward only): the upper half x[RR/2 ..] is zero and the lower half
// holds unreduced 64-bit words, so the first stage is x[j + RR/2] = x[j] w
// (Shoup's product takes any 64-bit multiplicand) and x[j] = x[j] mod p.
template <int S_BITS, int A, int R, bool INV, bool ZHI = false>
__device__ __forceinline__ void round_regs(u64 *x, int goff, const u64 *__restrict__ tw, const Mod &M) {
  constexpr int S = 1 << S_BITS;
  constexpr int D = S >> (A + R);
  constexpr int RR = 1 << R;
  static_assert(!(ZHI && INV), "zero upper half is a forward-only case");
  if (!INV) {
#pragma unroll
    for (int l = 0; l < R; l++) {
      const int half = 1 << (R - 1 - l);
      const int h = half * D;
      const int tstride = 2048 / h;
      if (ZHI && l == 0) {
#pragma unroll
        for (int j = 0; j < half; j++) {
          const u64 u = x[j];
          if (D == 1 && j == 0) {
            x[j + half] = mod_reduce64(u, M);  // twiddle w^0 = 1
          } else {
            const int e = goff + j * D;
            const ulonglong2 w = *(const ulonglong2 *)(tw + 2 * (e * tstride));
            x[j + half] = shoup_mul(u, w.x, w.y, M);
          }
          x[j] = mod_reduce64(u, M);
        }
        HGX_STAGE_FENCE();
        continue;
      }
#pragma unroll
      for (int j = 0; j < RR; j++) {
        if (j & half) continue;
        const u64 u = x[j], v = x[j + half];
        x[j] = mod_add(u, v, M);
        if (D == 1 && (j & (half - 1)) == 0) {
          x[j + half] = mod_sub(u, v, M);  // twiddle w^0 = 1
        } else {
          const int e = goff + (j & (half - 1)) * D;
          const ulonglong2 w = *(const ulonglong2 *)(tw + 2 * (e * tstride));
          x[j + half] = shoup_mul(u + (M.p - v), w.x, w.y, M);
        }
      }
      HGX_STAGE_FENCE();
    }
  } else {
    // An output that the next stage of this round only multiplies by a
    // nontrivial twiddle may stay in [0, 2p): Shoup's product takes any
    // 64-bit multiplicand and returns a canonical residue.  Outputs that the
    // next stage adds or subtracts, and every output of the round's last
    // stage, are canonical.  The roles are fixed at compile time.
    auto lazy_ok = [](int l, int q) {
      const int nh = 1 << (R - l);  // the next stage's half
      return l > 0 && (q & nh) != 0 && !(D == 1 && (q & (nh - 1)) == 0);
    };
#pragma unroll
    for (int l = R - 1; l >= 0; l--) {
      const int half = 1 << (R - 1 - l);
      const int h = half * D;
      const int tstride = 2048 / h;
#pragma unroll
      for (int j = 0; j < RR; j++) {
        if (j & half) continue;
        const u64 u = x[j];
        u64 v;
        if (D == 1 && (j & (half - 1)) == 0) {
          v = x[j + half];  // twiddle w^0 = 1
        } else {
          const int e = goff + (j & (half - 1)) * D;
          const ulonglong2 w = *(const ulonglong2 *)(tw + 2 * (e * tstride));
          v = shoup_mul(x[j + half], w.x, w.y, M);
        }
        x[j] = lazy_ok(l, j) ? u + v : mod_add(u, v, M);
        x[j + half] = lazy_ok(l, j + half) ? u + (M.p - v) : mod_sub(u, v, M);
      }
      HGX_STAGE_FENCE();
    }
  }
}

// COLFAST: column tile (16 columns, LDS index m*16 + t, lanes run along t).
// Otherwise: row tile (LDS index t*S + m, lanes run along m).
template <int S_BITS, int A, int R, bool INV, bool COLFAST>
__device__ __forceinline__ void tile_round(u64 *lds, int wbits, const u64 *__restrict__ tw, const Mod &M) {
  constexpr int S = 1 << S_BITS;
  constexpr int D = S >> (A + R);
  constexpr int LD = Log2<D>::value;
  constexpr int RR = 1 << R;
  constexpr int GPC_BITS = S_BITS - R;
  const int ngroups = 1 << (wbits + GPC_BITS);
  for (int gid = threadIdx.x; gid < ngroups; gid += HGX_NT) {
    int t, g;
    if (COLFAST) {
      t = gid & ((1 << wbits) - 1);
      g = gid >> wbits;
    } else {
      g = gid & ((1 << GPC_BITS) - 1);
      t = gid >> GPC_BITS;
    }
    const int goff = g & (D - 1);
    const int base = ((g >> LD) << (LD + R)) + goff;
    u64 x[RR];
#pragma unroll
    for (int j = 0; j < RR; j++) {
      const int m = base + j * D;
      const int li = COLFAST ? (m * 16 + t) : (t * S + m);
      x[j] = lds[HGX_PAD(li)];
    }
    round_regs<S_BITS, A, R, INV>(x, goff, tw, M);
#pragma unroll
    for (int j = 0; j < RR; j++) {
      const int m = base + j * D;
      const int li = COLFAST ? (m * 16 + t) : (t * S + m);
      lds[HGX_PAD(li)] = x[j];
    }
  }
}

template <int S_BITS, int I, bool COLFAST>
struct FwdRounds {
  static __device__ __forceinline__ void run(u64 *lds, int wbits, const u64 *tw, const Mod &M) {
    if constexpr (I < RoundPlan<S_BITS>::NR) {
      tile_round<S_BITS, RoundPlan<S_BITS>::start(I), RoundPlan<S_BITS>::bits(I), false, COLFAST>(lds, wbits, tw, M);
      __syncthreads();
      FwdRounds<S_BITS, I + 1, COLFAST>::run(lds, wbits, tw, M);
    }
  }
};

template <int S_BITS, int I, bool COLFAST>
struct InvRounds {  // I counts down from NR-1
  static __device__ __forceinline__ void run(u64 *lds, int wbits, const u64 *tw, const Mod &M) {
    if constexpr (I >= 0) {
      tile_round<S_BITS, RoundPlan<S_BITS>::start(I), RoundPlan<S_BITS>::bits(I), true, COLFAST>(lds, wbits, tw, M);
      __syncthreads();
      InvRounds<S_BITS, I - 1, COLFAST>::run(lds, wbits, tw, M);
    }
  }
};

template <int S_BITS, bool COLFAST>
__device__ __forceinline__ void tile_forward(u64 *lds, int wbits, const u64 *tw, const Mod &M) {
  FwdRounds<S_BITS, 0, COLFAST>::run(lds, wbits, tw, M);
}
template <int S_BITS, bool COLFAST>
__device__ __forceinline__ void tile_inverse(u64 *lds, int wbits, const u64 *tw, const Mod &M) {
  InvRounds<S_BITS, RoundPlan<S_BITS>::NR - 1, COLFAST>::run(lds, wbits, tw, M);
}

// ---------------------------------------------------------------------------
// CRT of one coefficient (Garner form) and the carry sum of one run.
// c = r0 + p0 (v1 + p1 v2) is the unique value below p0 p1 p2 with the given
// residues, i.e. the same number as the reference's
// (mod (sum f_k r_k) composite), src/ntt-multiply.lisp:242-252.
// ---------------------------------------------------------------------------
struct U192 { u64 w0, w1, w2; };

__device__ __forceinline__ U192 garner3(u64 r0, u64 r1, u64 r2, const PrimeSet &ps) {
  const u64 p0 = ps.p[0], p1 = ps.p[1], p2 = ps.p[2];
  const Mod M1 = mod_of(ps, 1), M2 = mod_of(ps, 2);
  // v1 = (r1 - r0) p0^-1 mod p1
  const u64 r0m1 = (r0 >= p1) ? r0 - p1 : r0;
  const u64 v1 = shoup_mul(mod_sub(r1, r0m1, M1), ps.g01, ps.g01q, M1);
  // v2 = ((r2 - r0) p0^-1 - v1) p1^-1 mod p2
  const u64 r0m2 = (r0 >= p2) ? r0 - p2 : r0;
  const u64 v1m2 = (v1 >= p2) ? v1 - p2 : v1;
  u64 t = shoup_mul(mod_sub(r2, r0m2, M2), ps.g02, ps.g02q, M2);
  const u64 v2 = shoup_mul(mod_sub(t, v1m2, M2), ps.g12, ps.g12q, M2);
  // u = v1 + p1 * v2  (< p1 p2 < 2^126)
  u64 ulo, uhi;
  mul64x64(p1, v2, ulo, uhi);
  ulo += v1;
  uhi += (ulo < v1) ? 1 : 0;
  // c = r0 + p0 * u  (< p0 p1 p2 < 2^188)
  u64 a0, a1, b0, b1;
  mul64x64(p0, ulo, a0, a1);
  mul64x64(p0, uhi, b0, b1);
  U192 c;
  c.w0 = a0 + r0;
  u64 cy = (c.w0 < a0) ? 1 : 0;
  u64 m1 = a1 + b0;
  u64 cy2 = (m1 < a1) ? 1 : 0;
  c.w1 = m1 + cy;
  cy2 += (c.w1 < m1) ? 1 : 0;
  c.w2 = b1 + cy2;
  return c;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Column pass, forward.  Tile = 16 adjacent columns of one block of length
// 2^bbits.  FIRST: the input is the operand's limbs (zero above its length),
// reduced mod prime y (as %make-ntt-work, src/number-theoretic-transform.lisp:7-22)
// into residue array y.  Otherwise: in place on the residue array of prime
// kbase + y.
// grid = (jobs << jt_bits, nprimes, noperands); FIRST: grid =
// (3 * (jobs << jt_bits), 1, noperands) with the prime in blockIdx.x mod 3.
// Tile-factored twiddles: the tile's 256 pairs w_B^(t0 q) are staged in LDS
// by one coalesced load at the start.
// BIG: blocks above 2^28 words (tile-factored twiddles only), whose in-block
// offsets need 64 bits; otherwise every access is a uniform base plus a
// 32-bit per-thread offset.
template <bool FIRST, int FORM, bool BIG = false>
__global__ __launch_bounds__(HGX_NT) void k_fwd_col(PrimeSet ps, ColTw tw, JobRef jr, int jt_bits, u64 L, int bbits,
                                                      int kbase) {
  // The 256-point column DFT is two rounds of 16: round 1 (stages 0-3) takes
  // rows g + 16 j of column t straight from HBM, round 2 (stages 4-7) rows
  // 16 g + j, which go straight back to HBM with their twiddles; one LDS
  // exchange between the rounds.  Every HBM access is a 128-byte row segment.
  __shared__ u64 lds[HGX_TILE_LDS];
  __shared__ ulonglong2 tts[FORM == TW_FACTORED ? 256 : 1];
  const int tid = threadIdx.x;
  const int t = tid & 15, g = tid >> 4;
  // FIRST: the three primes of a tile are consecutive blocks (grid.y = 1), so
  // the limbs they share are read from HBM close together in time
  const u32 bx = FIRST ? blockIdx.x / 3 : blockIdx.x;
  const int ky = FIRST ? (int)(blockIdx.x - 3 * bx) : (int)blockIdx.y;
  const MulJob J = load_job(jr, bx >> jt_bits);
  const u64 tile = bx & ((1u << jt_bits) - 1);
  const int tpb_bits = bbits - 12;  // tiles per block
  const u64 blk = tile >> tpb_bits;
  const u64 t0 = (tile & ((1ull << tpb_bits) - 1)) << 4;
  const u64 rs = 1ull << (bbits - 8);
  const u64 cbase = (blk << bbits) + t0;  // the tile's first column in the block
  if constexpr (FORM == TW_FACTORED) tts[tid] = gat2(gmem(tw.TT[kbase + ky]), (u32)((t0 >> 4) << 8) + tid);
  // LDS slots (padded, HGX_PAD) of rows g + 16 j and 16 g + j of column t
  const int l1 = 17 * g + t, l2 = 272 * g + t;
  // word offsets of rows g + 16 j (o1 + 16 rs j) and 16 g + j (o2 + rs j) of
  // column t from the tile's first column
  const u32 o1 = t + (u32)rs * g;
  const u64 o2 = t + rs * 16 * g;
  const int op = blockIdx.z;
  u64 *res = op ? J.res1 : J.res0;
  u64 x[16];

  const int k = kbase + ky;
  const Mod M = mod_of(ps, k);
  gu64 *arr = gmem(res + (u64)ky * L + cbase);
  if (FIRST) {
    // limbs -> residues mod p_k (the limbs are read once per prime)
    const u64 *limbs = op ? J.b : J.a;
    const u64 n = op ? J.nb : J.na;
    const gu64 *lb = gmem(limbs + cbase);
    // rows 128.. of the block (j >= 8 here) lie at or above the operand's
    // length: the first stage has a zero upper half
    if ((blk << bbits) + (rs << 7) >= n) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const u64 pos = cbase + 16 * rs * j + o1;
        x[j] = (pos < n) ? gat(lb + 16 * rs * j, o1) : 0;
      }
      round_regs<8, 0, 4, false, true>(x, g, ps.tw_fwd[k], M);
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const u64 pos = cbase + 16 * rs * j + o1;
        x[j] = mod_reduce64((pos < n) ? gat(lb + 16 * rs * j, o1) : 0, M);
      }
      round_regs<8, 0, 4, false>(x, g, ps.tw_fwd[k], M);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = gat(arr + 16 * rs * j, o1);
    round_regs<8, 0, 4, false>(x, g, ps.tw_fwd[k], M);
  }
#pragma unroll
  for (int j = 0; j < 16; j++) lds[l1 + 272 * j] = x[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = lds[l2 + 17 * j];
  round_regs<8, 4, 4, false>(x, 0, ps.tw_fwd[k], M);
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const u32 m = 16 * g + j;
    u64 v;
    if constexpr (FORM == TW_FACTORED)
      v = apply_twiddle_tt(x[j], tw, k, (u32)t0 + t, bitrev8(m), tts, M);
    else
      v = apply_twiddle<FORM>(x[j], tw, k, (u32)t0 + t, bitrev8(m), M);
    if (BIG)
      arr[rs * j + o2] = v;
    else
      gat(arr + rs * j, (u32)o2) = v;
  }
}

// Round I of a row transform (rows of 2^S_BITS words, row t of the tile at
// index t << S_BITS): group gid is row t, elements base + j*D.
template <int S_BITS, int I, int RMAX = 4>
struct RowRound {
  static constexpr int A = RoundPlan<S_BITS, RMAX>::start(I);
  static constexpr int R = RoundPlan<S_BITS, RMAX>::bits(I);
  static constexpr int RR = 1 << R;
  static constexpr int D = (1 << S_BITS) >> (A + R);
  static constexpr int LD = Log2<D>::value;
  static constexpr int GPC_BITS = S_BITS - R;
  static __device__ __forceinline__ void where(int gid, int &row, int &goff, int &base) {
    const int g = gid & ((1 << GPC_BITS) - 1);
    row = (gid >> GPC_BITS) << S_BITS;
    goff = g & (D - 1);
    base = ((g >> LD) << (LD + R)) + goff;
  }
};

// One forward round on the same groups of two row tiles: both tiles' values
// are read from LDS before either is transformed, so the second read's
// latency hides behind the first tile's butterflies.
template <int S_BITS, int I, int RMAX>
__device__ __forceinline__ void row_round_two(u64 *ldsA, u64 *ldsB, int wbits, const u64 *__restrict__ tw,
                                              const Mod &M) {
  using RI = RowRound<S_BITS, I, RMAX>;
#pragma unroll 1
  for (int gid = threadIdx.x; gid < (1 << (wbits + RI::GPC_BITS)); gid += HGX_NT) {
    int row, goff, base;
    RI::where(gid, row, goff, base);
    u64 a[RI::RR], b[RI::RR];
#pragma unroll
    for (int j = 0; j < RI::RR; j++) a[j] = ldsA[HGX_PAD(row + base + j * RI::D)];
#pragma unroll
    for (int j = 0; j < RI::RR; j++) b[j] = ldsB[HGX_PAD(row + base + j * RI::D)];
    round_regs<S_BITS, RI::A, RI::R, false>(a, goff, tw, M);
#pragma unroll
    for (int j = 0; j < RI::RR; j++) ldsA[HGX_PAD(row + base + j * RI::D)] = a[j];
    round_regs<S_BITS, RI::A, RI::R, false>(b, goff, tw, M);
#pragma unroll
    for (int j = 0; j < RI::RR; j++) ldsB[HGX_PAD(row + base + j * RI::D)] = b[j];
  }
}

// Forward rounds I .. I1 in LDS on one or two tiles, a barrier after each.
template <int S_BITS, int I, int I1, bool TWO, int RMAX = 4>
struct FwdRange {
  static __device__ __forceinline__ void run(u64 *ldsA, u64 *ldsB, int wbits, const u64 *tw, const Mod &M) {
    using RP = RoundPlan<S_BITS, RMAX>;
    if constexpr (I <= I1) {
      if constexpr (TWO)
        row_round_two<S_BITS, I, RMAX>(ldsA, ldsB, wbits, tw, M);
      else
        tile_round<S_BITS, RP::start(I), RP::bits(I), false, false>(ldsA, wbits, tw, M);
      __syncthreads();
      FwdRange<S_BITS, I + 1, I1, TWO, RMAX>::run(ldsA, ldsB, wbits, tw, M);
    }
  }
};
// Inverse rounds I down to I0 in LDS, a barrier after each.
template <int S_BITS, int I, int I0, int RMAX = 4>
struct InvRange {
  static __device__ __forceinline__ void run(u64 *lds, int wbits, const u64 *tw, const Mod &M) {
    using RP = RoundPlan<S_BITS, RMAX>;
    if constexpr (I >= I0) {
      tile_round<S_BITS, RP::start(I), RP::bits(I), true, false>(lds, wbits, tw, M);
      __syncthreads();
      InvRange<S_BITS, I - 1, I0, RMAX>::run(lds, wbits, tw, M);
    }
  }
};

// Last forward pass of both operands (or the one operand of a square), the
// pointwise product a*b/R (REDC) and the first inverse pass, on one tile of
// rows.  Round 0 reads HBM and the last inverse round writes it; the last
// forward round, the product and the first inverse round run per group in
// registers, so the operand tiles meet in LDS only between rounds.
template <int S_BITS, bool TWO, int RMAX = 4>
__device__ __forceinline__ void row_mul_tile(u64 *A, const u64 *Bp, u64 *ldsA, u64 *ldsB, int wbits, const u64 *twf,
                                             const u64 *twi, const Mod &M) {
  constexpr int NR = RoundPlan<S_BITS, RMAX>::NR;
  using R0 = RowRound<S_BITS, 0, RMAX>;
  using RL = RowRound<S_BITS, NR - 1, RMAX>;
  const int tid = threadIdx.x;
  if constexpr (NR == 1) {
#pragma unroll 1
    for (int gid = tid; gid < (1 << (wbits + R0::GPC_BITS)); gid += HGX_NT) {
      int row, goff, base;
      R0::where(gid, row, goff, base);
      u64 a[R0::RR], b[R0::RR];
#pragma unroll
      for (int j = 0; j < R0::RR; j++) a[j] = A[row + base + j * R0::D];
      round_regs<S_BITS, 0, R0::R, false>(a, goff, twf, M);
      if (TWO) {
#pragma unroll
        for (int j = 0; j < R0::RR; j++) b[j] = Bp[row + base + j * R0::D];
        round_regs<S_BITS, 0, R0::R, false>(b, goff, twf, M);
      }
#pragma unroll
      for (int j = 0; j < R0::RR; j++) a[j] = mont_mul(a[j], TWO ? b[j] : a[j], M);
      round_regs<S_BITS, 0, R0::R, true>(a, goff, twi, M);
#pragma unroll
      for (int j = 0; j < R0::RR; j++) A[row + base + j * R0::D] = a[j];
    }
  } else {
    // forward round 0 from HBM; both operands' loads are issued before the
    // first butterfly, so the second operand's HBM latency is hidden
#pragma unroll 1
    for (int gid = tid; gid < (1 << (wbits + R0::GPC_BITS)); gid += HGX_NT) {
      int row, goff, base;
      R0::where(gid, row, goff, base);
      u64 x[R0::RR], y[R0::RR];
#pragma unroll
      for (int j = 0; j < R0::RR; j++) x[j] = A[row + base + j * R0::D];
      if (TWO) {
#pragma unroll
        for (int j = 0; j < R0::RR; j++) y[j] = Bp[row + base + j * R0::D];
      }
      round_regs<S_BITS, 0, R0::R, false>(x, goff, twf, M);
#pragma unroll
      for (int j = 0; j < R0::RR; j++) ldsA[HGX_PAD(row + base + j * R0::D)] = x[j];
      if (TWO) {
        round_regs<S_BITS, 0, R0::R, false>(y, goff, twf, M);
#pragma unroll
        for (int j = 0; j < R0::RR; j++) ldsB[HGX_PAD(row + base + j * R0::D)] = y[j];
      }
    }
    __syncthreads();
    if constexpr (NR > 2) FwdRange<S_BITS, 1, NR - 2, TWO, RMAX>::run(ldsA, ldsB, wbits, twf, M);
    // last forward round, product, first inverse round
#pragma unroll 1
    for (int gid = tid; gid < (1 << (wbits + RL::GPC_BITS)); gid += HGX_NT) {
      int row, goff, base;
      RL::where(gid, row, goff, base);
      u64 a[RL::RR], b[RL::RR];
#pragma unroll
      for (int j = 0; j < RL::RR; j++) a[j] = ldsA[HGX_PAD(row + base + j * RL::D)];
      if (TWO) {
#pragma unroll
        for (int j = 0; j < RL::RR; j++) b[j] = ldsB[HGX_PAD(row + base + j * RL::D)];
      }
      round_regs<S_BITS, RL::A, RL::R, false>(a, goff, twf, M);
      if (TWO) round_regs<S_BITS, RL::A, RL::R, false>(b, goff, twf, M);
#pragma unroll
      for (int j = 0; j < RL::RR; j++) a[j] = mont_mul(a[j], TWO ? b[j] : a[j], M);
      round_regs<S_BITS, RL::A, RL::R, true>(a, goff, twi, M);
#pragma unroll
      for (int j = 0; j < RL::RR; j++) ldsA[HGX_PAD(row + base + j * RL::D)] = a[j];
    }
    __syncthreads();
    if constexpr (NR > 2) InvRange<S_BITS, NR - 2, 1, RMAX>::run(ldsA, wbits, twi, M);
    // inverse round 0 to HBM
#pragma unroll 1
    for (int gid = tid; gid < (1 << (wbits + R0::GPC_BITS)); gid += HGX_NT) {
      int row, goff, base;
      R0::where(gid, row, goff, base);
      u64 x[R0::RR];
#pragma unroll
      for (int j = 0; j < R0::RR; j++) x[j] = ldsA[HGX_PAD(row + base + j * R0::D)];
      round_regs<S_BITS, 0, R0::R, true>(x, goff, twi, M);
#pragma unroll
      for (int j = 0; j < R0::RR; j++) A[row + base + j * R0::D] = x[j];
    }
  }
}

// Row pass: contiguous blocks of 2^S_BITS words; a tile is
// min(L, 4096) consecutive words (2^tw_bits).
//   MODE_MUL / MODE_SQR, !SINGLE: last forward pass of both operands (or the
//     one operand), pointwise product (a*b/R by REDC), first inverse pass.
//   MODE_FWD / MODE_INV: transform-only row pass on res0 (prime kbase + y).
//   SINGLE (L <= 4096): the limbs come in, one workgroup per prime
//     (grid = (jobs, 3)), and the three residue arrays of the product go out
//     (k_crt finishes).

template <int S_BITS, int MODE, bool SINGLE>
__global__ __launch_bounds__(HGX_NT) void k_row(PrimeSet ps, RowParams rp, JobRef jr, int jt_bits) {
  constexpr bool TWO = (MODE == MODE_MUL);
  __shared__ u64 ldsA[HGX_TILE_LDS];
  __shared__ u64 ldsB[TWO ? HGX_TILE_LDS : 1];
  const int tid = threadIdx.x;
  const int TW = 1 << rp.tw_bits;
  const int wbits = rp.tw_bits - S_BITS;
  const MulJob J = load_job(jr, blockIdx.x >> jt_bits);
  const u64 tile = blockIdx.x & ((1u << jt_bits) - 1);
  const u64 gbase = tile << rp.tw_bits;

  if (!SINGLE) {
    const int k = rp.kbase + blockIdx.y;
    const Mod M = mod_of(ps, k);
    u64 *A = J.res0 + (u64)blockIdx.y * rp.L + gbase;
    if constexpr (MODE == MODE_MUL || MODE == MODE_SQR) {
      const u64 *Bp = TWO ? (J.res1 + (u64)blockIdx.y * rp.L + gbase) : nullptr;
      row_mul_tile<S_BITS, TWO>(A, Bp, ldsA, ldsB, wbits, ps.tw_fwd[k], ps.tw_inv[k], M);
    } else {
      for (int x = tid; x < TW; x += HGX_NT) ldsA[HGX_PAD(x)] = A[x];
      __syncthreads();
      if (MODE == MODE_FWD)
        tile_forward<S_BITS, false>(ldsA, wbits, ps.tw_fwd[k], M);
      else
        tile_inverse<S_BITS, false>(ldsA, wbits, ps.tw_inv[k], M);
      for (int x = tid; x < TW; x += HGX_NT) A[x] = ldsA[HGX_PAD(x)];
    }
  } else {
    // whole transform in one tile (L <= 4096): limbs in, one prime per
    // workgroup (blockIdx.y), residues of the product out (k_crt finishes)
    const int k = blockIdx.y;
    const Mod M = mod_of(ps, k);
    for (int x = tid; x < TW; x += HGX_NT) {
      const u64 pos = gbase + x;
      ldsA[HGX_PAD(x)] = (pos < J.na) ? mod_reduce64(J.a[pos], M) : 0;
      if (TWO) ldsB[HGX_PAD(x)] = (pos < J.nb) ? mod_reduce64(J.b[pos], M) : 0;
    }
    __syncthreads();
    tile_forward<S_BITS, false>(ldsA, wbits, ps.tw_fwd[k], M);
    if (TWO) tile_forward<S_BITS, false>(ldsB, wbits, ps.tw_fwd[k], M);
    for (int x = tid; x < TW; x += HGX_NT) {
      const u64 a = ldsA[HGX_PAD(x)];
      const u64 b = TWO ? ldsB[HGX_PAD(x)] : a;
      ldsA[HGX_PAD(x)] = mont_mul(mont_mul(a, b, M), rp.scale[k], M);
    }
    __syncthreads();
    tile_inverse<S_BITS, false>(ldsA, wbits, ps.tw_inv[k], M);
    u64 *dst = J.res0 + (u64)k * rp.L + gbase;
    for (int x = tid; x < TW; x += HGX_NT) dst[x] = ldsA[HGX_PAD(x)];
  }
}

// Whole transforms of 2^S_BITS words (S_BITS <= 10) for a batch of small
// products of one length: 2^jbits jobs per tile of 2^(S_BITS + jbits) <=
// 4096 words, one row per job, so that a workgroup's tile is full where
// k_row<SINGLE> would hold one job's 2^S_BITS words.  One workgroup per tile
// and prime (grid = (ceil(njobs / 2^jbits), 3)); limbs in, the product's
// residues out (k_crt finishes).  The same residues as k_row<SINGLE>.
template <int S_BITS, int MODE>
__global__ __launch_bounds__(HGX_NT) void k_row_pack(PrimeSet ps, RowParams rp, JobRef jr, int jbits, u32 njobs) {
  constexpr bool TWO = (MODE == MODE_MUL);
  constexpr int S = 1 << S_BITS;
  constexpr int NJ = HGX_TILE >> S_BITS;  // the most jobs of a tile
  __shared__ u64 ldsA[HGX_TILE_LDS];
  __shared__ u64 ldsB[TWO ? HGX_TILE_LDS : 1];
  __shared__ const u64 *sa[NJ];
  __shared__ const u64 *sb[NJ];
  __shared__ u64 *sres[NJ];
  __shared__ u64 sna[NJ], snb[NJ];
  const int tid = threadIdx.x;
  const int k = blockIdx.y;
  const Mod M = mod_of(ps, k);
  const int nj = 1 << jbits;
  const int TW = nj << S_BITS;
  const u32 j0 = blockIdx.x << jbits;
  for (int r = tid; r < nj; r += HGX_NT) {
    const u32 j = j0 + r;
    if (j < njobs) {
      const MulJob J = load_job(jr, j);
      sa[r] = J.a;
      sb[r] = TWO ? J.b : nullptr;
      sna[r] = J.na;
      snb[r] = TWO ? J.nb : 0;
      sres[r] = J.res0;
    } else {  // past the last job: zero rows, nothing stored
      sa[r] = nullptr;
      sb[r] = nullptr;
      sna[r] = 0;
      snb[r] = 0;
      sres[r] = nullptr;
    }
  }
  __syncthreads();
  for (int x = tid; x < TW; x += HGX_NT) {
    const int r = x >> S_BITS, pos = x & (S - 1);
    ldsA[HGX_PAD(x)] = ((u64)pos < sna[r]) ? mod_reduce64(sa[r][pos], M) : 0;
    if (TWO) ldsB[HGX_PAD(x)] = ((u64)pos < snb[r]) ? mod_reduce64(sb[r][pos], M) : 0;
  }
  __syncthreads();
  tile_forward<S_BITS, false>(ldsA, jbits, ps.tw_fwd[k], M);
  if (TWO) tile_forward<S_BITS, false>(ldsB, jbits, ps.tw_fwd[k], M);
  for (int x = tid; x < TW; x += HGX_NT) {
    const u64 a = ldsA[HGX_PAD(x)];
    const u64 b = TWO ? ldsB[HGX_PAD(x)] : a;
    ldsA[HGX_PAD(x)] = mont_mul(mont_mul(a, b, M), rp.scale[k], M);
  }
  __syncthreads();
  tile_inverse<S_BITS, false>(ldsA, jbits, ps.tw_inv[k], M);
  for (int x = tid; x < TW; x += HGX_NT) {
    u64 *dst = sres[x >> S_BITS];
    if (dst) dst[(u64)k * rp.L + (x & (S - 1))] = ldsA[HGX_PAD(x)];
  }
}

// Row pass of a product (TWO) or square on tiles of 2^TB words (rows of
// 2^S_BITS <= 2^TB words): 2048-word tiles with radix-8 rounds or 1024-word
// tiles with radix-4 rounds, two or four times the workgroups of k_row's
// 4096-word tiles for the same rows, which lowers the latency of products
// too small to fill the GPU (the runtime uses them when k_row would launch
// fewer than two workgroups per CU).  Same residues as k_row.
template <int S_BITS, bool TWO, int TB>
__global__ __launch_bounds__(HGX_NT) void k_row_half(PrimeSet ps, RowParams rp, JobRef jr, int jt_bits) {
  static_assert(S_BITS <= TB && (TB == 10 || TB == 11), "whole rows per tile, one group per thread per round");
  constexpr int NLDS = (1 << TB) + (1 << TB) / 16;
  __shared__ u64 ldsA[NLDS];
  __shared__ u64 ldsB[TWO ? NLDS : 1];
  const MulJob J = load_job(jr, blockIdx.x >> jt_bits);
  const u64 tile = blockIdx.x & ((1u << jt_bits) - 1);
  const u64 gbase = tile << TB;
  const int k = rp.kbase + blockIdx.y;
  const Mod M = mod_of(ps, k);
  u64 *A = J.res0 + (u64)blockIdx.y * rp.L + gbase;
  const u64 *Bp = TWO ? (J.res1 + (u64)blockIdx.y * rp.L + gbase) : nullptr;
  row_mul_tile<S_BITS, TWO, TB - 8>(A, Bp, ldsA, ldsB, TB - S_BITS, ps.tw_fwd[k], ps.tw_inv[k], M);
}

// Transform-only inverse when L <= 4096 (ntt-reverse on one tile, with the
// 1/L factor applied at the end).
template <int S_BITS>
__global__ __launch_bounds__(HGX_NT) void k_row_inv_scaled(PrimeSet ps, RowParams rp, JobRef jr) {
  __shared__ u64 ldsA[HGX_TILE_LDS];
  const int tid = threadIdx.x;
  const int TW = 1 << rp.tw_bits;
  const int wbits = rp.tw_bits - S_BITS;
  const int k = rp.kbase;
  const Mod M = mod_of(ps, k);
  u64 *A = jr.one.res0 + ((u64)blockIdx.x << rp.tw_bits);
  for (int x = tid; x < TW; x += HGX_NT) ldsA[HGX_PAD(x)] = A[x];
  __syncthreads();
  tile_inverse<S_BITS, false>(ldsA, wbits, ps.tw_inv[k], M);
  for (int x = tid; x < TW; x += HGX_NT) A[x] = mont_mul(ldsA[HGX_PAD(x)], rp.scale[k], M);
}

// ---------------------------------------------------------------------------
// Run-to-run carry resolution.  Run r holds digits D (16 limbs, written by the
// CRT pass) and receives X = ovf[r-1] plus a carry bit e from run r-1.  Its
// carry out is 1 iff limbs 2..15 are all ones and D_lo2 + X + e >= 2^128, so
// each run is a map e -> e' that kills (0), propagates (2) or generates (3).
// The carry into a run is the value of the nearest non-propagating map before
// it (0 if none): a last-non-P scan, done per 1024 runs and then across blocks.
// ---------------------------------------------------------------------------

__device__ __forceinline__ u32 run_map(const u64 *out, u64 n_out, const u64 *ovf, const u32 *flags, u64 r,
                                       int rl) {
  u64 x0 = 0, x1 = 0;
  if (r > 0) {
    x0 = ovf[2 * (r - 1)];
    x1 = ovf[2 * (r - 1) + 1];
  }
  const u64 base = r * (u64)rl;
  const u64 d0 = (base < n_out) ? out[base] : 0;
  const u64 d1 = (rl > 1 && base + 1 < n_out) ? out[base + 1] : 0;
  const u32 hi_ones = flags[r];
  u64 s0 = d0 + x0;
  u64 cy = (s0 < d0) ? 1 : 0;
  u64 s1 = d1 + x1;
  u64 c1 = (s1 < d1) ? 1 : 0;
  u64 s1b = s1 + cy;
  c1 += (s1b < s1) ? 1 : 0;
  const u32 f0 = c1 ? 1u : 0u;
  const u32 f1 = (c1 || (s0 == ~0ull && s1b == ~0ull)) ? 1u : 0u;
  if (!hi_ones) return 0u;
  return f0 | (f1 << 1);
}

// combine(earlier, later): later unless it propagates
__device__ __forceinline__ u32 map_combine(u32 a, u32 b) { return (b == 2u) ? a : b; }

__device__ inline u32 block_exclusive_last(u32 v, u32 *sh) {
  // inclusive Hillis-Steele scan of map_combine over the block, then shift
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    u32 other = (tid >= off) ? sh[tid - off] : 2u;
    __syncthreads();
    sh[tid] = map_combine(other, sh[tid]);
    __syncthreads();
  }
  u32 ex = (tid > 0) ? sh[tid - 1] : 2u;
  __syncthreads();
  return ex;
}

