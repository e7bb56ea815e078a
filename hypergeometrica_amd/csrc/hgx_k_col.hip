// hgx_k_col.hip -- column-pass, carry and input-generation kernels.
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"

// Column pass, inverse.  Twiddles w_B^-(t q) (times the final scale on the
// outermost pass) are applied as rows 16 g + j of column t come in from
// HBM; the 256-point DIT then runs as two rounds of 16 (stages 7-4 in
// registers, one LDS exchange, stages 3-0) and rows g + 16 j go straight
// back, in place on the residue array of prime kbase + y.
// grid = (jobs << jt_bits, nprimes)
template <int FORM, bool BIG = false>
__global__ __launch_bounds__(HGX_NT) void k_inv_col(PrimeSet ps, ColTw tw, JobRef jr, int jt_bits, u64 L, int bbits,
                                                      int kbase) {
  __shared__ u64 lds[HGX_TILE_LDS];
  const int tid = threadIdx.x;
  const int t = tid & 15, g = tid >> 4;
  const MulJob J = load_job(jr, blockIdx.x >> jt_bits);
  const u64 tile = blockIdx.x & ((1u << jt_bits) - 1);
  const int tpb_bits = bbits - 12;
  const u64 blk = tile >> tpb_bits;
  const u64 t0 = (tile & ((1ull << tpb_bits) - 1)) << 4;
  const u64 rs = 1ull << (bbits - 8);
  // LDS slots (padded, HGX_PAD) of rows g + 16 j and 16 g + j of column t
  const int l1 = 17 * g + t, l2 = 272 * g + t;
  // word offsets of rows g + 16 j (o1 + 16 rs j) and 16 g + j (o2 + rs j)
  // from the tile's first column; BIG: blocks above 2^28 words
  const u32 o1 = t + (u32)rs * g;
  const u64 o2 = t + rs * 16 * g;
  const int k = kbase + blockIdx.y;
  const Mod M = mod_of(ps, k);
  gu64 *arr = gmem(J.res0 + (u64)blockIdx.y * L + (blk << bbits) + t0);
  u64 x[16];
  if constexpr (FORM == TW_FACTORED) {
    __shared__ ulonglong2 tts[256];
    tts[tid] = gat2(gmem(tw.TT[k]), (u32)((t0 >> 4) << 8) + tid);
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = BIG ? (u64)arr[rs * j + o2] : (u64)gat(arr + rs * j, (u32)o2);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = apply_twiddle_tt(x[j], tw, k, (u32)t0 + t, bitrev8(16 * g + j), tts, M);
  } else {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const u32 m = 16 * g + j;
      const u64 v = BIG ? (u64)arr[rs * j + o2] : (u64)gat(arr + rs * j, (u32)o2);
      x[j] = apply_twiddle<FORM>(v, tw, k, (u32)t0 + t, bitrev8(m), M);
    }
  }
  round_regs<8, 4, 4, true>(x, 0, ps.tw_inv[k], M);
#pragma unroll
  for (int j = 0; j < 16; j++) lds[l2 + 17 * j] = x[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; j++) x[j] = lds[l1 + 272 * j];
  round_regs<8, 0, 4, true>(x, g, ps.tw_inv[k], M);
#pragma unroll
  for (int j = 0; j < 16; j++) gat(arr + 16 * rs * j, o1) = x[j];
}

// CRT and carry sums.  A workgroup takes a block of min(L, 2^cbits)
// consecutive coefficients (1024, or 512 when blocks of 1024 would be too few
// to fill the GPU): the threads reconstruct them from the three residue
// arrays (coalesced loads, Garner form) into LDS; then thread t sums its run
// of 4 coefficients exactly in registers (4 digits + a 2-word overflow); the
// overflows are carried from run to run inside the block (kill / propagate /
// generate scan over the block's 256 or 128 runs), so the block leaves exact
// digits plus one 2-word overflow and an all-ones flag for limbs 2.. of the
// block.
// add-big-digit (src/ntt-multiply.lisp:10-51) adds the same values at the
// same positions.  grid = jobs << bpj_bits
static_assert(HGX_CRT_BLOCK == HGX_NT * HGX_RUN, "one run per thread");
__global__ __launch_bounds__(HGX_NT) void k_crt(PrimeSet ps, JobRef jr, int bpj_bits, int cbits, u64 L) {
  constexpr int PADN = HGX_CRT_BLOCK + HGX_CRT_BLOCK / 16;
  __shared__ u64 w0[PADN];
  __shared__ u64 w1[PADN];
  __shared__ u64 w2[PADN];
  __shared__ u64 ovs[2 * HGX_NT];
  __shared__ u32 scan[HGX_NT];
  const int tid = threadIdx.x;
  const MulJob J = load_job(jr, blockIdx.x >> bpj_bits);
  const u64 blk = blockIdx.x & ((1u << bpj_bits) - 1);
  const int P = L < (1ull << cbits) ? (int)L : (1 << cbits);  // cbits <= HGX_CRT_BLOCK_BITS
  const int rl = P < HGX_RUN ? P : HGX_RUN;
  const int nruns = P / rl;
  const bool act = tid < nruns;
  const u64 base = blk << cbits;
  const gu64 *r0 = gmem(J.res0 + base), *r1 = gmem(J.res0 + L + base), *r2 = gmem(J.res0 + 2 * L + base);
  for (int x = tid; x < P; x += HGX_NT) {
    const U192 cf = garner3(gat(r0, x), gat(r1, x), gat(r2, x), ps);
    w0[HGX_PAD(x)] = cf.w0;
    w1[HGX_PAD(x)] = cf.w1;
    w2[HGX_PAD(x)] = cf.w2;
  }
  __syncthreads();
  // exact sum of this thread's run: digits d[], overflow (a0, a1)
  u64 d[HGX_RUN];
  u64 a0 = 0, a1 = 0, andhi = ~0ull;
#pragma unroll
  for (int j = 0; j < HGX_RUN; j++) {
    d[j] = 0;
    if (act && j < rl) {
      const int x = tid * rl + j;
      const u64 c0 = w0[HGX_PAD(x)], c1 = w1[HGX_PAD(x)], c2 = w2[HGX_PAD(x)];
      const u64 s0 = a0 + c0;
      const u64 cy = (s0 < a0) ? 1 : 0;
      const u64 s1 = a1 + c1;
      u64 cy1 = (s1 < a1) ? 1 : 0;
      const u64 s1b = s1 + cy;
      cy1 += (s1b < s1) ? 1 : 0;
      d[j] = s0;
      if (j >= 2) andhi &= s0;
      a0 = s1b;
      a1 = c2 + cy1;
    }
  }
  if (act) {
    ovs[2 * tid] = a0;
    ovs[2 * tid + 1] = a1;
  }
  const bool hi_ones = andhi == ~0ull;  // limbs 2.. of the run (vacuous for runs of 2)
  __syncthreads();
  // carry map of the run for the overflow of the previous run
  u32 m = 2u;
  u64 x0 = 0, x1 = 0;
  if (act && tid > 0) {
    x0 = ovs[2 * (tid - 1)];
    x1 = ovs[2 * (tid - 1) + 1];
    const u64 s0 = d[0] + x0;
    const u64 cy = (s0 < d[0]) ? 1 : 0;
    const u64 s1 = d[1] + x1;
    u64 c1 = (s1 < d[1]) ? 1 : 0;
    const u64 s1b = s1 + cy;
    c1 += (s1b < s1) ? 1 : 0;
    const u32 f0 = c1 ? 1u : 0u;
    const u32 f1 = (c1 || (s0 == ~0ull && s1b == ~0ull)) ? 1u : 0u;
    m = hi_ones ? (f0 | (f1 << 1)) : 0u;
  } else if (tid == 0) {
    m = 0u;  // the block's first run: incoming carries are added by k_carry_apply
  }
  const u32 ex = block_exclusive_last(m, scan);
  const u64 e = (ex == 3u) ? 1u : 0u;  // carry bit into this run from the run before
  if (act && tid > 0) {
    u64 carry = e;
#pragma unroll
    for (int j = 0; j < HGX_RUN; j++) {
      if (j < rl) {
        const u64 add = (j == 0) ? x0 : (j == 1 ? x1 : 0);
        const u64 s = d[j] + add;
        u64 cy = (s < d[j]) ? 1 : 0;
        const u64 s2 = s + carry;
        cy += (s2 < s) ? 1 : 0;
        d[j] = s2;
        carry = cy;
      }
    }
  }
  // the block's overflow: the last run's overflow plus the carry out of it
  if (tid == nruns - 1) {
    const u64 eo = (m >> e) & 1u;
    const u64 n0 = a0 + eo;
    J.ovf[2 * blk] = n0;
    J.ovf[2 * blk + 1] = a1 + ((n0 < a0) ? 1 : 0);
  }
  // limbs 2.. of the block all ones?
  u64 andall = ~0ull;
#pragma unroll
  for (int j = 0; j < HGX_RUN; j++)
    if (act && j < rl && tid * rl + j >= 2) andall &= d[j];
  // digits out through LDS, coalesced
#pragma unroll
  for (int j = 0; j < HGX_RUN; j++)
    if (act && j < rl) w0[HGX_PAD(tid * rl + j)] = d[j];
  const int allones = __syncthreads_and(andall == ~0ull);
  gu64 *outb = gmem(J.out + base);
  const u64 nleft = J.nout > base ? J.nout - base : 0;  // product limbs from this block on
  for (int x = tid; x < P; x += HGX_NT)
    if ((u64)x < nleft) gat(outb, x) = w0[HGX_PAD(x)];
  if (tid == 0) J.flags[blk] = allones ? 1u : 0u;
}

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

// Run r receives the previous run's overflow (two words added at its limbs 0
// and 1) and the carry bit E; the carry ripples until it stops.
__device__ __forceinline__ void carry_into_run(const MulJob &J, u64 r, int rl, u64 e) {
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
    carry_into_run(J, r, rl, e);
    e = (J.maps[r] >> e) & 1u;
  }
}

// All three steps in one workgroup per job when a job has at most 1024 runs
// (L <= 2^20): thread r takes run r's map, one scan over the workgroup gives
// every run its incoming carry, and the thread adds it.  grid = jobs,
// blockDim = the runs rounded up to a power of two, at least 64.
__global__ __launch_bounds__(1024) void k_carry_small(JobRef jr, u64 nruns, int rl) {
  __shared__ u32 sh[1024];
  const u64 r = threadIdx.x;
  const MulJob J = load_job(jr, blockIdx.x);
  const u32 m = (r < nruns) ? run_map(J.out, J.nout, J.ovf, J.flags, r, rl) : 2u;
  const u32 ex = block_exclusive_last(m, sh);
  if (r < nruns) carry_into_run(J, r, rl, (ex == 3u) ? 1ull : 0ull);
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

// Column passes for launches too small to fill the GPU: the same 4096-word
// tile (16 columns x 256 rows) and the same residues as k_fwd_col and
// k_inv_col, on 512 threads that hold 8 rows each, so twice the waves share
// a tile's work.  The 256-point transform runs as rounds of 8, 8 and 4:
// stages 0-2 on rows g + 32 j (straight from or to HBM), 3-5 on rows
// 32 (g / 4) + g % 4 + 4 j and 6-7 on rows 4 g' + j for g' = g, g + 32 (with
// the twiddles w_B^(t q) and HBM on the other side), two LDS exchanges.
// LDS slot of row r, column t: 17 r + t (padded as HGX_PAD).
constexpr int kColNT8 = 512;

template <bool FIRST, int FORM>
__global__ __launch_bounds__(kColNT8) void k_fwd_col8(PrimeSet ps, ColTw tw, JobRef jr, int jt_bits, u64 L, int bbits,
                                                     int kbase) {
  __shared__ u64 lds[HGX_TILE_LDS];
  __shared__ ulonglong2 tts[FORM == TW_FACTORED ? 256 : 1];
  const int tid = threadIdx.x;
  const int t = tid & 15, g = tid >> 4;
  const u32 bx = FIRST ? blockIdx.x / 3 : blockIdx.x;
  const int ky = FIRST ? (int)(blockIdx.x - 3 * bx) : (int)blockIdx.y;
  const MulJob J = load_job(jr, bx >> jt_bits);
  const u64 tile = bx & ((1u << jt_bits) - 1);
  const int tpb_bits = bbits - 12;
  const u64 blk = tile >> tpb_bits;
  const u64 t0 = (tile & ((1ull << tpb_bits) - 1)) << 4;
  const u64 rs = 1ull << (bbits - 8);
  const u64 cbase = (blk << bbits) + t0;
  if constexpr (FORM == TW_FACTORED) {
    if (tid < 256) tts[tid] = gat2(gmem(tw.TT[kbase + ky]), (u32)((t0 >> 4) << 8) + tid);
  }
  const int op = blockIdx.z;
  u64 *res = op ? J.res1 : J.res0;
  const int k = kbase + ky;
  const Mod M = mod_of(ps, k);
  gu64 *arr = gmem(res + (u64)ky * L + cbase);
  const u32 o1 = t + (u32)rs * g;  // row g of column t; row g + 32 j at o1 + 32 rs j
  u64 x[8];
  if (FIRST) {
    const u64 *limbs = op ? J.b : J.a;
    const u64 n = op ? J.nb : J.na;
    const gu64 *lb = gmem(limbs + cbase);
    if ((blk << bbits) + (rs << 7) >= n) {
      // rows 128.. (j >= 4) lie at or above the operand's length
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const u64 pos = cbase + 32 * rs * j + o1;
        x[j] = (pos < n) ? gat(lb + 32 * rs * j, o1) : 0;
      }
      round_regs<8, 0, 3, false, true>(x, g, ps.tw_fwd[k], M);
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const u64 pos = cbase + 32 * rs * j + o1;
        x[j] = mod_reduce64((pos < n) ? gat(lb + 32 * rs * j, o1) : 0, M);
      }
      round_regs<8, 0, 3, false>(x, g, ps.tw_fwd[k], M);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = gat(arr + 32 * rs * j, o1);
    round_regs<8, 0, 3, false>(x, g, ps.tw_fwd[k], M);
  }
#pragma unroll
  for (int j = 0; j < 8; j++) lds[17 * (g + 32 * j) + t] = x[j];
  __syncthreads();
  const int b2 = ((g >> 2) << 5) + (g & 3);
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = lds[17 * (b2 + 4 * j) + t];
  round_regs<8, 3, 3, false>(x, g & 3, ps.tw_fwd[k], M);
#pragma unroll
  for (int j = 0; j < 8; j++) lds[17 * (b2 + 4 * j) + t] = x[j];
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int gg = g + 32 * h;
    u64 y[4];
#pragma unroll
    for (int j = 0; j < 4; j++) y[j] = lds[17 * (4 * gg + j) + t];
    round_regs<8, 6, 2, false>(y, 0, ps.tw_fwd[k], M);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const u32 m = 4 * gg + j;
      u64 v;
      if constexpr (FORM == TW_FACTORED)
        v = apply_twiddle_tt(y[j], tw, k, (u32)t0 + t, bitrev8(m), tts, M);
      else
        v = apply_twiddle<FORM>(y[j], tw, k, (u32)t0 + t, bitrev8(m), M);
      gat(arr, (u32)rs * m + t) = v;
    }
  }
}

template <int FORM>
__global__ __launch_bounds__(kColNT8) void k_inv_col8(PrimeSet ps, ColTw tw, JobRef jr, int jt_bits, u64 L, int bbits,
                                                     int kbase) {
  __shared__ u64 lds[HGX_TILE_LDS];
  __shared__ ulonglong2 tts[FORM == TW_FACTORED ? 256 : 1];
  const int tid = threadIdx.x;
  const int t = tid & 15, g = tid >> 4;
  const MulJob J = load_job(jr, blockIdx.x >> jt_bits);
  const u64 tile = blockIdx.x & ((1u << jt_bits) - 1);
  const int tpb_bits = bbits - 12;
  const u64 blk = tile >> tpb_bits;
  const u64 t0 = (tile & ((1ull << tpb_bits) - 1)) << 4;
  const u64 rs = 1ull << (bbits - 8);
  const int k = kbase + blockIdx.y;
  const Mod M = mod_of(ps, k);
  gu64 *arr = gmem(J.res0 + (u64)blockIdx.y * L + (blk << bbits) + t0);
  if constexpr (FORM == TW_FACTORED) {
    if (tid < 256) tts[tid] = gat2(gmem(tw.TT[k]), (u32)((t0 >> 4) << 8) + tid);
    __syncthreads();
  }
  // stages 7-6 on rows 4 g' + j, twiddles w_B^-(t q) applied as they come in
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int gg = g + 32 * h;
    u64 y[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const u32 m = 4 * gg + j;
      const u64 v = gat(arr, (u32)rs * m + t);
      if constexpr (FORM == TW_FACTORED)
        y[j] = apply_twiddle_tt(v, tw, k, (u32)t0 + t, bitrev8(m), tts, M);
      else
        y[j] = apply_twiddle<FORM>(v, tw, k, (u32)t0 + t, bitrev8(m), M);
    }
    round_regs<8, 6, 2, true>(y, 0, ps.tw_inv[k], M);
#pragma unroll
    for (int j = 0; j < 4; j++) lds[17 * (4 * gg + j) + t] = y[j];
  }
  __syncthreads();
  const int b2 = ((g >> 2) << 5) + (g & 3);
  u64 x[8];
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = lds[17 * (b2 + 4 * j) + t];
  round_regs<8, 3, 3, true>(x, g & 3, ps.tw_inv[k], M);
#pragma unroll
  for (int j = 0; j < 8; j++) lds[17 * (b2 + 4 * j) + t] = x[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 8; j++) x[j] = lds[17 * (g + 32 * j) + t];
  round_regs<8, 0, 3, true>(x, g, ps.tw_inv[k], M);
  const u32 o1 = t + (u32)rs * g;
#pragma unroll
  for (int j = 0; j < 8; j++) gat(arr + 32 * rs * j, o1) = x[j];
}

// a column launch with fewer workgroups than this (two per CU) runs the
// 512-thread kernels above
constexpr unsigned kColSmallBelow = 512;

// the 512-thread tile (8 rows per thread): forced by WIDTH 512, refused by
// 256, else for launches too small to fill the GPU; blocks above 2^28 words
// need the 64-bit offsets of the 256-thread kernel
static bool col_small(dim3 grid, int bbits, int width) {
  if (bbits > 28 || width == 256) return false;
  return width == 512 || grid.x * grid.y * grid.z < kColSmallBelow;
}

static int tw_form(const ColTw &tw) { return tw.D[0] ? TW_DIRECT : (tw.DM[0] ? TW_DIRECT_MONT : TW_FACTORED); }

void hgx_launch_fwd_col(bool first, dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr,
                        int jt_bits, u64 L, int bbits, int kbase, int width) {
  if (col_small(grid, bbits, width)) {
    const dim3 b(kColNT8);
    switch (tw_form(tw) + (first ? 3 : 0)) {
      case TW_FACTORED: hipLaunchKernelGGL((k_fwd_col8<false, TW_FACTORED>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      case TW_DIRECT: hipLaunchKernelGGL((k_fwd_col8<false, TW_DIRECT>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      case TW_DIRECT_MONT: hipLaunchKernelGGL((k_fwd_col8<false, TW_DIRECT_MONT>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      case 3 + TW_FACTORED: hipLaunchKernelGGL((k_fwd_col8<true, TW_FACTORED>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      case 3 + TW_DIRECT: hipLaunchKernelGGL((k_fwd_col8<true, TW_DIRECT>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      default: hipLaunchKernelGGL((k_fwd_col8<true, TW_DIRECT_MONT>), grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
    }
    return;
  }
  const bool big = bbits > 28;  // tile-factored twiddles only (direct forms stop at 2^28)
  switch (tw_form(tw) + (first ? 3 : 0)) {
    case TW_FACTORED:
      if (big)
        hipLaunchKernelGGL((k_fwd_col<false, TW_FACTORED, true>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                           kbase);
      else
        hipLaunchKernelGGL((k_fwd_col<false, TW_FACTORED>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                           kbase);
      break;
    case TW_DIRECT:
      hipLaunchKernelGGL((k_fwd_col<false, TW_DIRECT>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      break;
    case TW_DIRECT_MONT:
      hipLaunchKernelGGL((k_fwd_col<false, TW_DIRECT_MONT>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                         kbase);
      break;
    case 3 + TW_FACTORED:
      if (big)
        hipLaunchKernelGGL((k_fwd_col<true, TW_FACTORED, true>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                           kbase);
      else
        hipLaunchKernelGGL((k_fwd_col<true, TW_FACTORED>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                           kbase);
      break;
    case 3 + TW_DIRECT:
      hipLaunchKernelGGL((k_fwd_col<true, TW_DIRECT>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      break;
    default:
      hipLaunchKernelGGL((k_fwd_col<true, TW_DIRECT_MONT>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits,
                         kbase);
      break;
  }
}

void hgx_launch_inv_col(dim3 grid, hipStream_t s, const PrimeSet &ps, const ColTw &tw, const JobRef &jr, int jt_bits,
                        u64 L, int bbits, int kbase, int width) {
  if (col_small(grid, bbits, width)) {
    const dim3 b(kColNT8);
    switch (tw_form(tw)) {
      case TW_FACTORED: hipLaunchKernelGGL(k_inv_col8<TW_FACTORED>, grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      case TW_DIRECT: hipLaunchKernelGGL(k_inv_col8<TW_DIRECT>, grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
      default: hipLaunchKernelGGL(k_inv_col8<TW_DIRECT_MONT>, grid, b, 0, s, ps, tw, jr, jt_bits, L, bbits, kbase); break;
    }
    return;
  }
  switch (tw_form(tw)) {
    case TW_FACTORED:
      if (bbits > 28)
        hipLaunchKernelGGL((k_inv_col<TW_FACTORED, true>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      else
        hipLaunchKernelGGL((k_inv_col<TW_FACTORED>), grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      break;
    case TW_DIRECT:
      hipLaunchKernelGGL(k_inv_col<TW_DIRECT>, grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      break;
    default:
      hipLaunchKernelGGL(k_inv_col<TW_DIRECT_MONT>, grid, dim3(HGX_NT), 0, s, ps, tw, jr, jt_bits, L, bbits, kbase);
      break;
  }
}

// direct Montgomery table of one pass from its tile-factored pair:
// DM[q*T + t] = TT[t/16][q] * TS[q][t mod 16] * 2^64 mod p (both factors are
// plain values with Shoup companions; rm = 2^64 mod p with companion rmq
// turns the product into Montgomery form)
__global__ __launch_bounds__(256) void k_direct_mont(u64 *DM, const u64 *TS, const u64 *TT, int tbits, Mod M, u64 rm,
                                                     u64 rmq) {
  const u64 total = 256ull << tbits;
  for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < total; i += (u64)gridDim.x * 256) {
    const u32 q = (u32)(i >> tbits);
    const u64 t = i & ((1ull << tbits) - 1);
    const ulonglong2 w = *(const ulonglong2 *)(TS + 2 * ((q << 4) + (u32)(t & 15)));
    const u64 f = TT[2 * (((t >> 4) << 8) + q)];
    DM[i] = shoup_mul(shoup_mul(f, w.x, w.y, M), rm, rmq, M);
  }
}

// Twiddle tables on the device: entry (i, j) of a ROWS x COLS table, row
// major as (value, Shoup companion) pairs, is SCALE * R^(i j) mod p
// (pow_table_entry; RM, SM and ONE in Montgomery form, PINV = p^-1 mod 2^64).
__global__ __launch_bounds__(256) void k_pow_table(u64 *tab, u32 rows, u64 cols, u64 rm, u64 sm, u64 one, u64 pinv,
                                                   Mod M) {
  const u64 total = (u64)rows * cols;
  for (u64 id = (u64)blockIdx.x * 256 + threadIdx.x; id < total; id += (u64)gridDim.x * 256) {
    const u64 i = id / cols, j = id - i * cols;
    u64 v, q;
    pow_table_entry(i * j, rm, sm, one, pinv, M, v, q);
    tab[2 * id] = v;
    tab[2 * id + 1] = q;
  }
}

void hgx_launch_pow_table(hipStream_t s, u64 *tab, u32 rows, u64 cols, u64 rm, u64 sm, u64 one, u64 pinv,
                          const Mod &M) {
  const u64 total = (u64)rows * cols;
  const u64 blocks = (total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536;
  hipLaunchKernelGGL(k_pow_table, dim3((unsigned)blocks), dim3(256), 0, s, tab, rows, cols, rm, sm, one, pinv, M);
}

void hgx_launch_direct_mont(hipStream_t s, u64 *DM, const u64 *TS, const u64 *TT, int tbits, const Mod &M, u64 rm,
                            u64 rmq) {
  hipLaunchKernelGGL(k_direct_mont, dim3(4096), dim3(256), 0, s, DM, TS, TT, tbits, M, rm, rmq);
}

void hgx_launch_crt(dim3 grid, hipStream_t s, const PrimeSet &ps, const JobRef &jr, int bpj_bits, int cbits, u64 L) {
  hipLaunchKernelGGL(k_crt, grid, dim3(HGX_NT), 0, s, ps, jr, bpj_bits, cbits, L);
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

void hgx_launch_carry_small(dim3 grid, hipStream_t s, const JobRef &jr, u64 nruns, int rl) {
  // one thread per run (nruns <= 1024), whole waves: fewer scan steps for
  // fewer runs
  unsigned nt = 64;
  while (nt < nruns) nt <<= 1;
  hipLaunchKernelGGL(k_carry_small, grid, dim3(nt), 0, s, jr, nruns, rl);
}

void hgx_launch_splitmix64(dim3 grid, hipStream_t s, u64 *out, u64 n, u64 seed, int force_top) {
  hipLaunchKernelGGL(k_splitmix64, grid, dim3(256), 0, s, out, n, seed, force_top);
}
