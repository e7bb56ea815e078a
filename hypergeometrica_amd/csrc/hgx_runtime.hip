// hgx_runtime.hip -- host side of the C ABI declared in include/hgx.h.
//
// One device context per process: the prime set and its root tables are
// built once at hgx_init and stay resident (the reference builds **scheme**
// once at load, src/ntt-multiply.lisp:8); per-transform-length plans (the
// two-level root tables of the column passes) are built on first use and
// cached; work arrays grow on demand under the HBM budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hgx.h"
#include "hgx_fold.hpp"
#include "hgx_launch.hpp"
#include "hgx_sum.h"

typedef unsigned __int128 u128;

namespace {

constexpr int kMaxBits = 32;   // largest transform: 2^32 words per residue array
constexpr int kRowMaxBits = 12;
// a group of whole-tile products of 2^4 .. 2^10 words packs several per tile
constexpr int kPackMinBits = 4, kPackMaxBits = 10;
// a multiply row pass with fewer 4096-word workgroups than this (two per CU
// of the 256) runs on 2048-word tiles instead, and on 1024-word tiles when
// the 2048-word ones are still fewer
constexpr u64 kHalfTileBelow = 512;
// transforms up to this length can run the CRT on blocks of 512 limbs (twice
// the workgroups) when blocks of 1024 would be fewer than kHalfTileBelow; their
// carry arrays are sized for 512-limb blocks
constexpr int kCrtHalfMaxBits = 18;

// log2 of the CRT block (and carry run) for NJOBS products of length 2^n
int crt_block_bits(int n, unsigned njobs) {
  if (n <= kCrtHalfMaxBits && n > HGX_CRT_BLOCK_BITS &&
      ((u64)njobs << (n - HGX_CRT_BLOCK_BITS)) < kHalfTileBelow)
    return HGX_CRT_BLOCK_BITS - 1;
  return HGX_CRT_BLOCK_BITS;
}

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(HGX_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

// ---- host modular arithmetic (table construction only) ----
u64 mulmod(u64 a, u64 b, u64 m) { return (u64)(((u128)a * b) % m); }
u64 powmod(u64 a, u64 e, u64 m) {
  u64 r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = mulmod(r, a, m);
    a = mulmod(a, a, m);
    e >>= 1;
  }
  return r;
}
u64 invmod(u64 a, u64 m) { return powmod(a, m - 2, m); }  // m prime
u64 to_mont(u64 a, u64 m) { return (u64)((((u128)(a % m)) << 64) % m); }
u64 shoup_companion(u64 w, u64 m) { return (u64)((((u128)w) << 64) / m); }  // calc-b-inv
// m^-1 mod 2^64 for odd m (Newton's iteration doubles the correct low bits)
u64 inv_pow2_64(u64 m) {
  u64 x = m;
  for (int i = 0; i < 6; i++) x *= 2 - m * x;
  return x;
}

bool is_prime_small(u64 n) {
  if (n < 2) return false;
  for (u64 d = 2; d * d <= n; d++)
    if (n % d == 0) return false;
  return true;
}

// smallest prime generator of Z/pZ* (find-finite-field-generator,
// src/moduli.lisp:173-180); p - 1 = c * 2^55 with small c.
u64 find_generator(u64 p, u64 c) {
  std::vector<u64> f{2};
  u64 r = c;
  for (u64 d = 2; d * d <= r; d++)
    if (r % d == 0) {
      if (d != 2) f.push_back(d);
      while (r % d == 0) r /= d;
    }
  if (r > 1 && r != 2) f.push_back(r);
  for (u64 g = 2;; g++) {
    if (!is_prime_small(g)) continue;
    bool ok = true;
    for (u64 q : f)
      if (powmod(g, (p - 1) / q, p) == 1) {
        ok = false;
        break;
      }
    if (ok) return g;
  }
}

// The moduli of default-moduli (src/moduli.lisp:257-264): the 63-bit primes
// i * 2^55 + 1, descending.
const u64 kModuli[3] = {0x6280000000000001ull, 0x5700000000000001ull, 0x4180000000000001ull};

struct ColPassTables {
  int bbits = 0;
  bool direct = false;  // B <= 2^direct_max: direct (q, t) tables
  u64 *Df[3], *Di[3], *Di_mul[3], *Di_tr[3];
  // otherwise tile-factored: small (q, tt) tables with the scale, tile tables
  u64 *TSf[3], *TSi[3], *TSi_mul[3], *TSi_tr[3], *TTf[3], *TTi[3];
  // and, when chosen, direct Montgomery tables of the forward and of the
  // product's inverse (the transform-only inverse keeps the factored pair)
  u64 *DMf[3] = {}, *DMi[3] = {};
};

// Column passes on blocks of at most 2^kDirectMaxBits words use the direct
// twiddle table (16 bytes per word of the block, per prime and direction).
#ifndef HGX_DIRECT_MAX_BITS
#define HGX_DIRECT_MAX_BITS 20
#endif
constexpr int kDirectMaxBits = HGX_DIRECT_MAX_BITS;
// whether column passes on blocks above 2^kDirectMaxBits (and at most
// 2^28 words) use direct Montgomery tables by default
constexpr int kMontDirectDefault = 0;

struct Plan {
  int n = 0, ncol = 0, srow = 0;
  ColPassTables col[8];
  u64 *dev = nullptr;  // one allocation holding the host-built tables
  size_t bytes = 0;
  std::vector<std::pair<u64 *, size_t>> big;  // device-built direct Montgomery tables
};

// A length's pass structure: NCOL column passes (row pass of n - 8 NCOL
// bits), the largest block with the direct table (2^DM words), and whether
// larger blocks (up to 2^kMontDirectMaxBits words) use direct Montgomery tables.
// COLW and ROWT choose kernels, not tables: the column tile's thread count
// (256 or 512; 0 the default rule: 512 for launches of fewer than 512
// workgroups) and the multiply row pass's tile (2^ROWT words, 12, 11 or 10;
// 0 the default rule: smaller tiles for launches too small to fill the GPU).
struct Choice {
  int ncol, dm, md;
  int colw = 0, rowt = 0;
  bool operator==(const Choice &o) const {
    return ncol == o.ncol && dm == o.dm && md == o.md && colw == o.colw && rowt == o.rowt;
  }
  bool operator!=(const Choice &o) const { return !(*this == o); }
  bool same_tables(const Choice &o) const { return ncol == o.ncol && dm == o.dm && md == o.md; }
};
constexpr int kMontDirectMaxBits = 28;

struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct Context {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ws_done = nullptr;  // last use of the shared workspace
  size_t budget = 0, allocated = 0;
  PrimeSet ps{};
  u64 gen[3]{};
  u64 *tables = nullptr;
  std::map<int, std::unique_ptr<Plan>> plans;  // by plan_key(n, choice)
  std::map<int, Choice> choice;                 // tuned or set pass structure per length
  Workspace ws, stage, jobs_dev, pack, scan;
  MulJob *jobs_host = nullptr;  // pinned job-table staging
  size_t jobs_host_cap = 0;
  hipEvent_t jobs_copied = nullptr;
  bool split_on = true;  // unbalanced products as batched chunk products
  // captured launch sequences of batched chunks (hipGraph), keyed by shape
  bool graphs_on = true;
  // cap on the device memory of one host-memory product before it is staged
  // in chunks (0: the HBM budget alone)
  std::atomic<size_t> stage_limit{0};
  std::map<std::vector<u64>, hipGraphExec_t> graphs;
  std::recursive_mutex mu;
};

Context g_ctx;

// ---- per-kernel timing (benchmarks) ----
struct TimedLaunch {
  const char *name;
  hipEvent_t a, b;
};
bool g_timing = false;
std::vector<TimedLaunch> g_pending;
std::map<std::string, std::pair<uint64_t, double>> g_kstats;
std::vector<hipEvent_t> g_timing_events;  // drained events, reused (no creation on the launch path)

hipEvent_t timing_event() {
  if (!g_timing_events.empty()) {
    hipEvent_t e = g_timing_events.back();
    g_timing_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct KTimer {
  hipStream_t s;
  bool on;
  KTimer(const char *name, hipStream_t st) : s(st), on(g_timing) {
    if (!on) return;
    TimedLaunch t{name, timing_event(), timing_event()};
    (void)hipEventRecord(t.a, s);
    g_pending.push_back(t);
  }
  ~KTimer() {
    if (on) (void)hipEventRecord(g_pending.back().b, s);
  }
};

void timing_drain() {
  for (auto &t : g_pending) {
    (void)hipEventSynchronize(t.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    auto &st = g_kstats[t.name];
    st.first += 1;
    st.second += ms;
    g_timing_events.push_back(t.a);
    g_timing_events.push_back(t.b);
  }
  g_pending.clear();
}

#define HGX_TIMED(name, s) KTimer kt_##__LINE__(name, s)

bool slots_trim();

int dev_alloc(void **p, size_t bytes) {
  if (g_ctx.allocated + bytes > g_ctx.budget && slots_trim()) {
    // idle call slots gave their memory back; try again
  }
  if (g_ctx.allocated + bytes > g_ctx.budget)
    return fail(HGX_ENOMEM, "HBM budget exceeded: need " + std::to_string(bytes) + " more bytes, " +
                                std::to_string(g_ctx.budget - g_ctx.allocated) + " left");
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(HGX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  g_ctx.allocated += bytes;
  return HGX_OK;
}

void dev_free(void *p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  g_ctx.allocated -= bytes;
}

int ensure(Workspace &w, size_t bytes) {
  if (w.bytes >= bytes) return HGX_OK;
  HIPCHK(hipEventSynchronize(g_ctx.ws_done));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  dev_free(w.ptr, w.bytes);
  w.ptr = nullptr;
  w.bytes = 0;
  size_t want = std::max(bytes, (size_t)1 << 20);
  int rc = dev_alloc(&w.ptr, want);
  if (rc) return rc;
  w.bytes = want;
  return HGX_OK;
}

// ---- host transfers through pinned staging ----
// A caller's limbs live in pageable memory (calloc or mmap, src/ram-vec.lisp:52-59,
// src/disk-vec.lisp:34-35).  Large transfers go through kPinBufs pinned buffers
// of kPinChunk bytes of a PinStage: worker threads copy chunk k + 1 between the
// caller's memory and a staging buffer while the DMA engine moves chunk k.
// Downloads are deferred: a buffer's bytes reach the caller's memory when the
// buffer is needed again or at pin_flush().  A PinStage has one user at a time:
// the library's own (g_pin) is used under g_ctx.mu, the request server's
// stager thread owns another.
constexpr size_t kPinChunk = (size_t)32 << 20;
constexpr int kPinBufs = 4;
constexpr size_t kPinMin = (size_t)4 << 20;  // smaller transfers: one plain copy

// memcpy split over a few worker threads; the workers are never joined (the
// pool is leaked), so process exit does not wait for them.  Copies from
// different threads take turns (USE).
struct CopyPool {
  std::mutex use;
  std::mutex mu;
  std::condition_variable cv, done;
  char *dst = nullptr;
  const char *src = nullptr;
  size_t bytes = 0;
  int parts = 0, next = 0, left = 0;
  int nthreads = 0;
  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [this] { return next < parts; });
      const int p = next++;
      const size_t per = (bytes + parts - 1) / parts;
      const size_t off = std::min(bytes, (size_t)p * per), len = std::min(per, bytes - off);
      char *d = dst;
      const char *s = src;
      lk.unlock();
      if (len) std::memcpy(d + off, s + off, len);
      lk.lock();
      if (--left == 0) done.notify_all();
    }
  }
  void copy(void *d, const void *s, size_t n) {
    if (nthreads == 0 || n < kPinMin) {
      std::memcpy(d, s, n);
      return;
    }
    std::lock_guard<std::mutex> turn(use);
    std::unique_lock<std::mutex> lk(mu);
    dst = (char *)d;
    src = (const char *)s;
    bytes = n;
    parts = nthreads;
    next = 0;
    left = nthreads;
    cv.notify_all();
    done.wait(lk, [this] { return left == 0; });
    parts = 0;
  }
};

CopyPool *copy_pool() {
  static CopyPool *pool = [] {
    CopyPool *p = new CopyPool;
    const unsigned hw = std::thread::hardware_concurrency();
    const int n = (int)std::min(8u, std::max(1u, hw / 2));
    for (int i = 0; i < n; i++) std::thread([p] { p->loop(); }).detach();
    p->nthreads = n;
    return p;
  }();
  return pool;
}

struct PinStage {
  char *buf[kPinBufs] = {};
  hipEvent_t ev[kPinBufs] = {};
  int state[kPinBufs] = {};  // 0 free, 1 upload in flight, 2 download to copy out
  char *out_dst[kPinBufs] = {};
  size_t out_len[kPinBufs] = {};
  int next = 0;
  bool ready = false;
};
PinStage g_pin;

int pin_init(PinStage &P) {
  if (P.ready) return HGX_OK;
  for (int i = 0; i < kPinBufs; i++) {
    HIPCHK(hipHostMalloc((void **)&P.buf[i], kPinChunk, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&P.ev[i], hipEventDisableTiming));
  }
  (void)copy_pool();
  P.ready = true;
  return HGX_OK;
}

// wait for buffer i's transfer; a finished download is copied to its caller
int pin_settle(PinStage &P, int i) {
  if (P.state[i] == 0) return HGX_OK;
  const hipError_t e = hipEventSynchronize(P.ev[i]);
  const int was = P.state[i];
  P.state[i] = 0;
  if (e != hipSuccess) return fail(HGX_EDEVICE, std::string("staging transfer failed: ") + hipGetErrorString(e));
  if (was == 2) copy_pool()->copy(P.out_dst[i], P.buf[i], P.out_len[i]);
  return HGX_OK;
}

// every deferred download has reached the caller's memory
int pin_flush(PinStage &P) {
  if (!P.ready) return HGX_OK;
  for (int k = 0; k < kPinBufs; k++) {
    const int i = (P.next + k) % kPinBufs;  // oldest first
    int rc = pin_settle(P, i);
    if (rc) return rc;
  }
  return HGX_OK;
}
int pin_flush() { return pin_flush(g_pin); }

// host memory the HIP runtime already has pinned (hipHostMalloc): the DMA
// engine reads and writes it directly
bool host_pinned(const void *p) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is not an error here
    return false;
  }
  return attr.type == hipMemoryTypeHost;
}

// host -> device on stream s.  Staged and small pageable transfers have read
// the host memory on return; a pinned source is read by the DMA engine when
// the stream reaches the copy, so the caller keeps it until the stream
// synchronizes (every caller here does before it returns).
int h2d(PinStage &P, void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kPinMin || host_pinned(src)) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return HGX_OK;
  }
  int rc = pin_init(P);
  if (rc) return rc;
  for (size_t off = 0; off < bytes; off += kPinChunk) {
    const int i = P.next;
    P.next = (i + 1) % kPinBufs;
    rc = pin_settle(P, i);
    if (rc) return rc;
    const size_t len = std::min(kPinChunk, bytes - off);
    copy_pool()->copy(P.buf[i], (const char *)src + off, len);
    HIPCHK(hipMemcpyAsync((char *)dst + off, P.buf[i], len, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(P.ev[i], s));
    P.state[i] = 1;
  }
  return HGX_OK;
}
int h2d(void *dst, const void *src, size_t bytes, hipStream_t s) { return h2d(g_pin, dst, src, bytes, s); }

// device -> host on stream s, after the work queued before it; the bytes are
// in the host memory after pin_flush() (or a later transfer reusing the
// buffers)
int d2h(PinStage &P, void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kPinMin || host_pinned(dst)) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    return HGX_OK;
  }
  int rc = pin_init(P);
  if (rc) return rc;
  for (size_t off = 0; off < bytes; off += kPinChunk) {
    const int i = P.next;
    P.next = (i + 1) % kPinBufs;
    rc = pin_settle(P, i);
    if (rc) return rc;
    const size_t len = std::min(kPinChunk, bytes - off);
    HIPCHK(hipMemcpyAsync(P.buf[i], (const char *)src + off, len, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(P.ev[i], s));
    P.state[i] = 2;
    P.out_dst[i] = (char *)dst + off;
    P.out_len[i] = len;
  }
  return HGX_OK;
}
int d2h(void *dst, const void *src, size_t bytes, hipStream_t s) { return d2h(g_pin, dst, src, bytes, s); }

// An API call that queues downloads flushes them before it returns success;
// on any other way out the scope drops what is still pending, so no later
// call writes into memory its caller may have released.
struct PinScope {
  PinStage &P;
  explicit PinScope(PinStage &p = g_pin) : P(p) {}
  ~PinScope() {
    if (!P.ready) return;
    for (int i = 0; i < kPinBufs; i++) {
      if (P.state[i] == 0) continue;
      (void)hipEventSynchronize(P.ev[i]);
      P.state[i] = 0;
    }
  }
};

void pin_free(PinStage &P) {
  if (!P.ready) return;
  for (int i = 0; i < kPinBufs; i++) {
    (void)hipEventSynchronize(P.ev[i]);
    (void)hipEventDestroy(P.ev[i]);
    (void)hipHostFree(P.buf[i]);
  }
  P = PinStage{};
}

// ---- host packs of batched products from host memory ----
// hgx_mul_batch gathers the operands of many products into one contiguous
// pack, moves it with one DMA, and scatters the products back.  The packs
// live in pinned memory (grown on demand up to kPackHostMax bytes per
// direction, used under g_ctx.mu), so the DMA engines read and write them
// directly, and the gather and scatter run on several threads.
constexpr size_t kPackHostMax = (size_t)1 << 30;
struct HostBuf {
  void *p = nullptr;
  size_t bytes = 0;
};
HostBuf g_pack_in, g_pack_out;

// pinned memory of at least BYTES in B, or null (beyond kPackHostMax, or
// the allocation failed: the caller then uses pageable memory)
u64 *host_buf(HostBuf &b, size_t bytes) {
  if (bytes > kPackHostMax) return nullptr;
  if (b.bytes >= bytes) return (u64 *)b.p;
  if (b.p) (void)hipHostFree(b.p);
  b = HostBuf{};
  const size_t want = std::min(kPackHostMax, std::max(bytes, (size_t)64 << 20));
  if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return nullptr;
  }
  b.bytes = want;
  return (u64 *)b.p;
}

void host_buf_free(HostBuf &b) {
  if (b.p) (void)hipHostFree(b.p);
  b = HostBuf{};
}

struct Piece {
  void *dst;
  const void *src;
  size_t bytes;
};

// the copies of PIECES, their concatenated bytes split evenly over up to 8
// threads (at least 4 MiB each)
void copy_pieces(const std::vector<Piece> &pieces) {
  size_t total = 0;
  for (const Piece &p : pieces) total += p.bytes;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, std::max(1u, hw / 2)), std::max<size_t>(1, total >> 22));
  auto work = [&](size_t t) {
    const size_t lo = total * t / nt, hi = total * (t + 1) / nt;
    size_t at = 0;
    for (const Piece &p : pieces) {
      const size_t a = std::max(lo, at), b = std::min(hi, at + p.bytes);
      if (a < b) std::memcpy((char *)p.dst + (a - at), (const char *)p.src + (a - at), b - a);
      at += p.bytes;
      if (at >= hi) break;
    }
  };
  std::vector<std::thread> th;
  size_t t = 1;
  try {
    for (; t < nt; t++) th.emplace_back(work, t);
  } catch (const std::exception &) {
    // no thread for part t and above: this thread copies them
  }
  for (size_t u = t; u < nt; u++) work(u);
  work(0);
  for (auto &x : th) x.join();
}

// ---- plans ----
// A length-2^n transform is NCOL column passes of 8 bits and a row pass of
// n - 8 NCOL bits; a column pass needs blocks of at least one 4096-word tile,
// so the row pass keeps 4..12 bits.  Column passes on blocks of at most
// 2^DIRECT_MAX words use the direct twiddle table.  The default is the fewest
// passes and DIRECT_MAX = kDirectMaxBits; hgx_plan_set and hgx_autotune
// choose others per length.
int min_ncol(int n) { return n <= kRowMaxBits ? 0 : (n - kRowMaxBits + 7) / 8; }
int max_ncol(int n) { return n <= kRowMaxBits ? 0 : (n - 4) / 8; }
Choice default_choice(int n) { return Choice{min_ncol(n), kDirectMaxBits, kMontDirectDefault}; }
Choice choice_for(int n) {
  auto it = g_ctx.choice.find(n);
  return it != g_ctx.choice.end() ? it->second : default_choice(n);
}
int plan_key(int n, const Choice &c) { return (n << 13) | (c.md << 12) | (c.ncol << 6) | c.dm; }

int build_plan_variant(int n, const Choice &c, Plan **out);
int build_plan(int n, Plan **out) { return build_plan_variant(n, choice_for(n), out); }

int build_plan_variant(int n, const Choice &choice, Plan **out) {
  const int ncol = choice.ncol, dm = choice.dm;
  const int key = plan_key(n, choice);
  auto it = g_ctx.plans.find(key);
  if (it != g_ctx.plans.end()) {
    *out = it->second.get();
    return HGX_OK;
  }
  auto plan = std::make_unique<Plan>();
  plan->n = n;
  plan->ncol = ncol;
  plan->srow = n - 8 * ncol;
  const u64 L = 1ull << n;
  // the tables are filled on the device (k_pow_table): here each gets its
  // offset in the plan's block, in words
  struct Fill {
    size_t at;
    u32 rows;
    u64 cols, root, scale;
    int k;
  };
  std::vector<Fill> fills;
  size_t words = 0;
  struct Off { size_t Df, Di, Dm, Dt, Sf, Si, Sm, St, Tf, Ti; };
  std::vector<Off> offs(plan->ncol * 3);
  for (int i = 0; i < plan->ncol; i++) {
    const int b = n - 8 * i;
    const bool direct = b <= dm;
    const size_t T = (size_t)1 << (b - 8);
    plan->col[i].bbits = b;
    plan->col[i].direct = direct;
    for (int k = 0; k < 3; k++) {
      const u64 p = kModuli[k];
      const u64 w = powmod(g_ctx.gen[k], (p - 1) >> b, p);
      const u64 wi = invmod(w, p);
      const u64 Linv = invmod(L % p, p);
      const u64 Rm = (u64)((((u128)1) << 64) % p);
      const u64 kmul = mulmod(Rm, Linv, p);  // R / L: undoes the pointwise REDC's 1/R and scales by 1/L
      const u64 ktr = Linv;                   // 1 / L
      Off &o = offs[i * 3 + k];
      o = Off{};
      // ROWS x COLS pairs (SCALE * R^(i j), Shoup companion floor(value * 2^64 / p))
      auto table = [&](u32 rows, u64 cols, u64 r, u64 scale) {
        const size_t at = words;
        fills.push_back(Fill{at, rows, cols, r, scale, k});
        words += 2 * (size_t)rows * cols;
        return at;
      };
      // direct table: row q holds SCALE * (root^q)^t for t < T
      auto push_direct = [&](u64 root, u64 scale) { return table(256, T, root, scale); };
      // small table: SCALE * root^(tt q) at q*16 + tt
      auto push_small = [&](u64 root, u64 scale) { return table(256, 16, root, scale); };
      // tile table: root^(16 c q) at c*256 + q, c < T/16
      auto push_tile = [&](u64 root) { return table((u32)(T / 16), 256, powmod(root, 16, p), 1); };
      if (direct) {
        o.Df = push_direct(w, 1);
        if (i == 0) {
          o.Dm = push_direct(wi, kmul);
          o.Dt = push_direct(wi, ktr);
        } else {
          o.Di = push_direct(wi, 1);
        }
      } else {
        o.Sf = push_small(w, 1);
        o.Tf = push_tile(w);
        o.Ti = push_tile(wi);
        if (i == 0) {
          o.Sm = push_small(wi, kmul);
          o.St = push_small(wi, ktr);
        } else {
          o.Si = push_small(wi, 1);
        }
      }
    }
  }
  if (words) {
    plan->bytes = words * sizeof(u64);
    int rc = dev_alloc((void **)&plan->dev, plan->bytes);
    if (rc) return rc;
    for (const Fill &f : fills) {
      const u64 p = kModuli[f.k];
      const Mod M{p, 0 - p, (u32)((p - 1) >> 55)};
      hgx_launch_pow_table(g_ctx.stream, plan->dev + f.at, f.rows, f.cols, to_mont(f.root, p), to_mont(f.scale, p),
                           to_mont(1, p), inv_pow2_64(p), M);
    }
    const hipError_t e = hipGetLastError();
    const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(g_ctx.stream) : e;
    if (e2 != hipSuccess) {
      dev_free(plan->dev, plan->bytes);
      return fail(HGX_EDEVICE, std::string("root tables: ") + hipGetErrorString(e2));
    }
    for (int i = 0; i < plan->ncol; i++)
      for (int k = 0; k < 3; k++) {
        const Off &o = offs[i * 3 + k];
        ColPassTables &c = plan->col[i];
        if (c.direct) {
          c.Df[k] = plan->dev + o.Df;
          c.Di[k] = (i == 0) ? nullptr : plan->dev + o.Di;
          c.Di_mul[k] = (i == 0) ? plan->dev + o.Dm : nullptr;
          c.Di_tr[k] = (i == 0) ? plan->dev + o.Dt : nullptr;
        } else {
          c.TSf[k] = plan->dev + o.Sf;
          c.TSi[k] = (i == 0) ? nullptr : plan->dev + o.Si;
          c.TSi_mul[k] = (i == 0) ? plan->dev + o.Sm : nullptr;
          c.TSi_tr[k] = (i == 0) ? plan->dev + o.St : nullptr;
          c.TTf[k] = plan->dev + o.Tf;
          c.TTi[k] = plan->dev + o.Ti;
        }
      }
  }
  // direct Montgomery tables, built on the device from the factored pairs
  if (choice.md) {
    u64 rmont[3];  // 2^64 mod p_k: turns a plain value into Montgomery form
    for (int k = 0; k < 3; k++) rmont[k] = to_mont(1, kModuli[k]);
    for (int i = 0; i < plan->ncol; i++) {
      ColPassTables &c = plan->col[i];
      if (c.direct || c.bbits > kMontDirectMaxBits) continue;
      const size_t bytes = (size_t)sizeof(u64) << c.bbits;
      for (int k = 0; k < 3; k++) {
        const Mod M{kModuli[k], 0 - kModuli[k], (u32)((kModuli[k] - 1) >> 55)};
        for (int dir = 0; dir < 2; dir++) {
          u64 *dmt = nullptr;
          int rc = dev_alloc((void **)&dmt, bytes);
          if (rc) {
            for (auto &b : plan->big) dev_free(b.first, b.second);
            dev_free(plan->dev, plan->bytes);
            return rc;
          }
          plan->big.push_back({dmt, bytes});
          const u64 *TS = dir == 0 ? c.TSf[k] : (i == 0 ? c.TSi_mul[k] : c.TSi[k]);
          const u64 *TT = dir == 0 ? c.TTf[k] : c.TTi[k];
          hgx_launch_direct_mont(g_ctx.stream, dmt, TS, TT, c.bbits - 8, M, rmont[k], shoup_companion(rmont[k], M.p));
          (dir == 0 ? c.DMf : c.DMi)[k] = dmt;
        }
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g_ctx.stream));
  }
  *out = plan.get();
  g_ctx.plans[key] = std::move(plan);
  return HGX_OK;
}

void free_plan(Plan &p) {
  dev_free(p.dev, p.bytes);
  for (auto &b : p.big) dev_free(b.first, b.second);
  p.big.clear();
}

void drop_plan(int key) {
  auto it = g_ctx.plans.find(key);
  if (it == g_ctx.plans.end()) return;
  free_plan(*it->second);
  g_ctx.plans.erase(it);
}

ColTw coltw(const ColPassTables &c, int which) {
  // which: 0 forward, 1 inverse, 2 inverse with multiply scale, 3 inverse with 1/L
  ColTw t{};
  t.tbits = c.bbits - 8;
  for (int k = 0; k < 3; k++) {
    if (c.direct) {
      t.D[k] = (which == 0) ? c.Df[k] : (which == 1 ? c.Di[k] : (which == 2 ? c.Di_mul[k] : c.Di_tr[k]));
    } else if ((which == 0 && c.DMf[k]) || ((which == 1 || which == 2) && c.DMi[k])) {
      t.D[k] = nullptr;
      t.DM[k] = (which == 0) ? c.DMf[k] : c.DMi[k];
    } else {
      t.D[k] = nullptr;
      t.TS[k] = (which == 0) ? c.TSf[k] : (which == 1 ? c.TSi[k] : (which == 2 ? c.TSi_mul[k] : c.TSi_tr[k]));
      t.TT[k] = (which == 0) ? c.TTf[k] : c.TTi[k];
    }
  }
  return t;
}

// ---- row-kernel dispatch ----
template <int MODE, bool SINGLE>
int launch_row(int sbits, dim3 grid, hipStream_t s, const RowParams &rp, const JobRef &jr, int jt_bits) {
  static const char *names[4] = {SINGLE ? "row_single_mul" : "row_mul", SINGLE ? "row_single_sqr" : "row_sqr",
                                 "row_fwd", "row_inv"};
  HGX_TIMED(names[MODE], s);
  bool ok = false;
  if (MODE == MODE_MUL)
    ok = hgx_launch_row_mul(SINGLE, sbits, grid, s, g_ctx.ps, rp, jr, jt_bits);
  else if (MODE == MODE_SQR)
    ok = hgx_launch_row_sqr(SINGLE, sbits, grid, s, g_ctx.ps, rp, jr, jt_bits);
  else
    ok = hgx_launch_row_tr(MODE, sbits, grid, s, g_ctx.ps, rp, jr, jt_bits);
  return ok ? HGX_OK : fail(HGX_EINVAL, "row pass size out of range");
}

int launch_row_inv_scaled(int sbits, dim3 grid, hipStream_t s, const RowParams &rp, const JobRef &jr) {
  HGX_TIMED("row_inv_scaled", s);
  return hgx_launch_row_inv_scaled(sbits, grid, s, g_ctx.ps, rp, jr) ? HGX_OK
                                                                       : fail(HGX_EINVAL, "row pass size out of range");
}

constexpr size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int log2_ceil(u64 x) {
  int n = 0;
  while ((1ull << n) < x) n++;
  return n;
}

// Work memory of one multiply of transform length 2^n.
struct WorkLayout {
  size_t res0, res1, ovf, flags, maps, agg, blkin, total;
};
constexpr WorkLayout work_layout(int n, bool sqr) {
  const u64 L = 1ull << n;
  // carry runs of the smallest CRT block the length may use
  const u64 rl = std::min<u64>(L, 1ull << (n <= kCrtHalfMaxBits ? HGX_CRT_BLOCK_BITS - 1 : HGX_CRT_BLOCK_BITS));
  const u64 nruns = L / rl;
  const u64 nblk = (nruns + HGX_CARRY_RUNS_PER_BLOCK - 1) / HGX_CARRY_RUNS_PER_BLOCK;
  WorkLayout w{};
  size_t at = 0;
  const bool single = n <= kRowMaxBits;
  w.res0 = at;
  at += align_up(3 * L * sizeof(u64));
  w.res1 = at;
  at += (single || sqr) ? 0 : align_up(3 * L * sizeof(u64));
  w.ovf = at;
  at += align_up(2 * nruns * sizeof(u64));
  w.flags = at;
  at += align_up(nruns * sizeof(u32));
  w.maps = at;
  at += align_up(nruns);
  w.agg = at;
  at += align_up(nblk);
  w.blkin = at;
  at += align_up(nblk);
  w.total = at;
  return w;
}

// Products without allocation (hgx_mul_device_noalloc) compute in a scratch
// array the library holds on every device it runs on: the work memory of one
// product of transform length at most 2^kNoAllocMaxBits (the largest layout
// of those lengths is the general product's at the top length).
constexpr int kNoAllocMaxBits = 20;
constexpr size_t noalloc_bytes() {
  size_t m = 0;
  for (int n = 1; n <= kNoAllocMaxBits; n++)
    for (int sq = 0; sq < 2; sq++) m = std::max(m, work_layout(n, sq != 0).total);
  return m;
}
constexpr size_t kNoAllocBytes = noalloc_bytes();
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};  // the array's device address, per device (hipGetSymbolAddress)

}  // namespace

// Work memory of hgx_mul_device_noalloc: the 6 * 2^20 residue words and the
// carry arrays of one product of transform length 2^20 (48.0 MiB), held by
// every device the library runs on until the process ends.  Host code never
// uses this name as an address: hgx_init looks up each device's copy with
// hipGetSymbolAddress.
static __device__ __align__(256) uint64_t scratch[kNoAllocBytes / 8];

namespace {

MulJob make_job(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, bool sqr, const WorkLayout &wl,
                char *base) {
  MulJob J{};
  J.a = a;
  J.na = na;
  J.b = sqr ? a : b;
  J.nb = sqr ? na : nb;
  J.out = out;
  J.nout = na + (sqr ? na : nb);
  J.res0 = (u64 *)(base + wl.res0);
  J.res1 = (u64 *)(base + wl.res1);
  J.ovf = (u64 *)(base + wl.ovf);
  J.flags = (u32 *)(base + wl.flags);
  J.maps = (uint8_t *)(base + wl.maps);
  J.agg = (uint8_t *)(base + wl.agg);
  J.blkin = (uint8_t *)(base + wl.blkin);
  return J;
}
MulJob make_job(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, int n, bool sqr, char *base) {
  return make_job(a, na, b, nb, out, sqr, work_layout(n, sqr), base);
}

// Enqueue NJOBS products of transform length 2^n (all squares or all
// general products) on stream s.  JR holds the job by value (one job) or a
// device table of NJOBS jobs.
int enqueue_group(const JobRef &jr, unsigned njobs, int n, bool sqr, hipStream_t s) {
  if (n > kMaxBits) return fail(HGX_ETOOBIG, "transform length 2^" + std::to_string(n) + " above 2^32");
  Plan *plan = nullptr;
  int rc = build_plan(n, &plan);
  if (rc) return rc;
  const u64 L = 1ull << n;
  const int cbits = crt_block_bits(n, njobs);
  const u64 rl = std::min<u64>(L, 1ull << cbits);  // carry runs are whole CRT blocks
  const u64 nruns = L / rl;
  const int nops = sqr ? 1 : 2;

  if (n <= kRowMaxBits) {
    RowParams rp{};
    rp.L = L;
    rp.tw_bits = n;
    rp.kbase = 0;
    for (int k = 0; k < 3; k++) {
      const u64 p = kModuli[k];
      // pointwise value a*b/R times this (Montgomery) factor = a*b/L
      const u64 Rm = (u64)((((u128)1) << 64) % p);
      rp.scale[k] = to_mont(mulmod(Rm, invmod(L % p, p), p), p);
    }
    // several small products per tile when the group has them
    const int jb = (n >= kPackMinBits && n <= kPackMaxBits) ? std::min(kRowMaxBits - n, log2_ceil(njobs)) : 0;
    if (jb > 0) {
      HGX_TIMED(sqr ? "row_pack_sqr" : "row_pack_mul", s);
      const dim3 grid((njobs + (1u << jb) - 1) >> jb, 3);
      const bool ok = sqr ? hgx_launch_row_pack_sqr(n, jb, njobs, grid, s, g_ctx.ps, rp, jr)
                          : hgx_launch_row_pack_mul(n, jb, njobs, grid, s, g_ctx.ps, rp, jr);
      if (!ok) return fail(HGX_EINVAL, "packed row pass size out of range");
    } else {
      rc = sqr ? launch_row<MODE_SQR, true>(n, dim3(njobs, 3), s, rp, jr, 0)
               : launch_row<MODE_MUL, true>(n, dim3(njobs, 3), s, rp, jr, 0);
      if (rc) return rc;
    }
  } else {
    const int jt = n - 12;  // tiles per job
    const unsigned grid = njobs << jt;
    const Choice ch = choice_for(n);
    {
      HGX_TIMED("fwd_col_first", s);
      hgx_launch_fwd_col(true, dim3(grid * 3, 1, nops), s, g_ctx.ps, coltw(plan->col[0], 0), jr, jt, L,
                         plan->col[0].bbits, 0, ch.colw);
    }
    for (int i = 1; i < plan->ncol; i++) {
      HGX_TIMED("fwd_col", s);
      hgx_launch_fwd_col(false, dim3(grid, 3, nops), s, g_ctx.ps, coltw(plan->col[i], 0), jr, jt, L,
                         plan->col[i].bbits, 0, ch.colw);
    }
    RowParams rp{};
    rp.L = L;
    rp.tw_bits = 12;
    rp.kbase = 0;
    // the row tile: as chosen for the length, else 2048-word tiles (twice
    // the workgroups; 1024-word tiles, four times, when that is still few)
    // when 4096-word tiles are too few to fill the GPU, for a lower latency
    int tb = 0;
    if (ch.rowt == HGX_HALF_TILE_BITS || ch.rowt == HGX_HALF_TILE_BITS - 1) {
      if (plan->srow <= ch.rowt) tb = ch.rowt;
    } else if (ch.rowt == 0 && plan->srow <= HGX_HALF_TILE_BITS && 3ull * grid < kHalfTileBelow) {
      tb = (plan->srow < HGX_HALF_TILE_BITS && 6ull * grid < kHalfTileBelow) ? HGX_HALF_TILE_BITS - 1
                                                                             : HGX_HALF_TILE_BITS;
    }
    if (tb) {
      HGX_TIMED(sqr ? "row_sqr_half" : "row_mul_half", s);
      const int jh = n - tb;
      rp.tw_bits = tb;
      if (!hgx_launch_row_half(!sqr, tb, plan->srow, dim3(njobs << jh, 3), s, g_ctx.ps, rp, jr, jh))
        return fail(HGX_EINVAL, "row pass size out of range");
    } else {
      rc = sqr ? launch_row<MODE_SQR, false>(plan->srow, dim3(grid, 3), s, rp, jr, jt)
               : launch_row<MODE_MUL, false>(plan->srow, dim3(grid, 3), s, rp, jr, jt);
      if (rc) return rc;
    }
    for (int i = plan->ncol - 1; i >= 0; i--) {
      HGX_TIMED("inv_col", s);
      hgx_launch_inv_col(dim3(grid, 3), s, g_ctx.ps, coltw(plan->col[i], i == 0 ? 2 : 1), jr, jt, L, plan->col[i].bbits,
                         0, ch.colw);
    }
  }
  {
    const int cb = n > cbits ? n - cbits : 0;  // CRT blocks per job (log2)
    HGX_TIMED("crt", s);
    hgx_launch_crt(dim3(njobs << cb), s, g_ctx.ps, jr, cb, cbits, L);
  }
  // run-to-run carries between the CRT blocks
  if (nruns > 1 && nruns <= 1024) {
    HGX_TIMED("carry_small", s);
    hgx_launch_carry_small(dim3(njobs), s, jr, nruns, (int)rl);
  } else if (nruns > 1) {
    const u64 nblk = (nruns + HGX_CARRY_RUNS_PER_BLOCK - 1) / HGX_CARRY_RUNS_PER_BLOCK;
    const int bpj = log2_ceil(nblk);
    {
      HGX_TIMED("carry_map", s);
      hgx_launch_carry_map(dim3(njobs << bpj), s, jr, bpj, nruns, (int)rl);
    }
    {
      HGX_TIMED("carry_blocks", s);
      hgx_launch_carry_blocks(dim3(njobs), s, jr, nblk);
    }
    {
      HGX_TIMED("carry_apply", s);
      hgx_launch_carry_apply(dim3(njobs << bpj), s, jr, bpj, nruns, (int)rl);
    }
  }
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

// Forward or inverse transform of one residue array in place (device).
int enqueue_transform(u64 *d, size_t nwords, int mod_num, bool inverse, hipStream_t s) {
  const int n = log2_ceil(nwords);
  if ((1ull << n) != nwords) return fail(HGX_EINVAL, "transform length must be a power of two");
  if (n > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32");
  if (n == 0) return HGX_OK;  // length 1: both transforms are the identity
  Plan *plan = nullptr;
  int rc = build_plan(n, &plan);
  if (rc) return rc;
  const u64 L = nwords;
  const int jt = n >= 12 ? n - 12 : 0;
  const unsigned tiles = 1u << jt;
  JobRef jr{};
  jr.one.res0 = d;
  RowParams rp{};
  rp.L = L;
  rp.tw_bits = n >= 12 ? 12 : n;
  rp.kbase = mod_num;
  if (!inverse) {
    for (int i = 0; i < plan->ncol; i++) {
      HGX_TIMED("fwd_col", s);
      hgx_launch_fwd_col(false, dim3(tiles, 1, 1), s, g_ctx.ps, coltw(plan->col[i], 0), jr, jt, L, plan->col[i].bbits,
                         mod_num);
    }
    rc = launch_row<MODE_FWD, false>(plan->srow, dim3(tiles, 1), s, rp, jr, jt);
    if (rc) return rc;
  } else {
    if (plan->ncol == 0) {
      const u64 p = kModuli[mod_num];
      rp.scale[mod_num] = to_mont(invmod(L % p, p), p);  // REDC(x, (1/L) R) = x / L
      rc = launch_row_inv_scaled(plan->srow, dim3(tiles), s, rp, jr);
      if (rc) return rc;
    } else {
      rc = launch_row<MODE_INV, false>(plan->srow, dim3(tiles, 1), s, rp, jr, jt);
      if (rc) return rc;
      for (int i = plan->ncol - 1; i >= 0; i--) {
        HGX_TIMED("inv_col", s);
        hgx_launch_inv_col(dim3(tiles, 1), s, g_ctx.ps, coltw(plan->col[i], i == 0 ? 3 : 1), jr, jt, L,
                           plan->col[i].bbits, mod_num);
      }
    }
  }
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

int check_ready() {
  if (!g_ctx.ready) return fail(HGX_ENOTINIT, "hgx_init has not been called");
  return HGX_OK;
}

// Order work on a caller stream after the previous use of the shared
// workspace, and record this use.
int ws_acquire(hipStream_t s) {
  HIPCHK(hipStreamWaitEvent(s, g_ctx.ws_done, 0));
  return HGX_OK;
}
int ws_release(hipStream_t s) {
  HIPCHK(hipEventRecord(g_ctx.ws_done, s));
  return HGX_OK;
}

// ---- unbalanced products ----
// A product whose longer operand (NA limbs) is much longer than the shorter
// (NB) is computed as K products of consecutive M-limb chunks of the longer
// operand with the shorter one, all of transform length 2^NC (one batched
// group): chunk i's product belongs at limb i*M.  The products of even and of
// odd chunks go to two zeroed images, where they do not overlap (M >= NB), and
// one addition with carries adds the images: the same integer as the one
// transform of length >= NA + NB, for a fraction of its work.
struct Split {
  int nc = 0;
  u64 m = 0, k = 0;
};

// Splitting costs a few launches, a job-table copy and an addition per
// product: only products of transform length >= 2^kSplitMinBits are split.
constexpr int kSplitMinBits = 19;

bool choose_split(u64 na, u64 nb, Split &sp) {
  if (na < nb) std::swap(na, nb);
  const int n = log2_ceil(na + nb);
  if (n < kSplitMinBits || nb == 0) return false;
  const double full = 3.0 * std::ldexp(1.0, n) * n;
  double best = 0.8 * full;  // split only for a clear gain
  bool found = false;
  for (int nc = std::max(log2_ceil(2 * nb), 1); nc < n; nc++) {
    const u64 Lc = 1ull << nc, m = Lc - nb;
    if (m < nb) continue;
    const u64 k = (na + m - 1) / m;
    // transform work, the zeroing and the addition, and a per-job overhead
    const double c = 3.0 * (double)k * (double)Lc * nc + 8.0 * (double)(na + nb) + 2048.0 * (double)k;
    if (c < best) {
      best = c;
      sp = Split{nc, m, k};
      found = true;
    }
  }
  return found;
}

int ensure_jobs_host(size_t count) {
  HIPCHK(hipEventSynchronize(g_ctx.jobs_copied));
  if (g_ctx.jobs_host_cap < count) {
    if (g_ctx.jobs_host) (void)hipHostFree(g_ctx.jobs_host);
    g_ctx.jobs_host = nullptr;
    g_ctx.jobs_host_cap = std::max<size_t>(count, 1024);
    HIPCHK(hipHostMalloc((void **)&g_ctx.jobs_host, g_ctx.jobs_host_cap * sizeof(MulJob), hipHostMallocDefault));
  }
  return HGX_OK;
}

// Device memory of a split product: the two images, the job table and the
// chunks' work memory (MAIN), and the carry-scan scratch of the image sum.
struct SplitBytes {
  size_t main, scan;
};
SplitBytes split_bytes(size_t na, size_t nb, const Split &sp) {
  const u64 n = na + nb;
  return {2 * align_up(n * sizeof(u64)) + align_up(sp.k * sizeof(MulJob)) + sp.k * work_layout(sp.nc, false).total,
          align_up(2 * hgx_addsub_blocks(n))};
}

// The split product into D_OUT on stream S, in the device memory at BASE
// (split_bytes().main bytes) with scan scratch AGG.  JOBS (sp.k entries of
// host memory) must stay unchanged until the stream has copied it.
int split_product_at(const u64 *d_a, size_t na, const u64 *d_b, size_t nb, u64 *d_out, const Split &sp,
                     hipStream_t s, char *base, uint8_t *agg, MulJob *jobs) {
  if (na < nb) {
    std::swap(d_a, d_b);
    std::swap(na, nb);
  }
  const u64 n = na + nb;
  const WorkLayout wl = work_layout(sp.nc, false);
  const size_t img = align_up(n * sizeof(u64));
  const size_t tbl = align_up(sp.k * sizeof(MulJob));
  const u64 nblk = hgx_addsub_blocks(n);
  u64 *E = (u64 *)base, *O = (u64 *)(base + img);
  MulJob *table = (MulJob *)(base + 2 * img);
  char *work = base + 2 * img + tbl;
  hgx_launch_fill(s, E, 2 * img / sizeof(u64), 0);
  for (u64 i = 0; i < sp.k; i++) {
    const u64 off = i * sp.m;
    const u64 mi = std::min<u64>(sp.m, na - off);
    jobs[i] = make_job(d_a + off, mi, d_b, nb, ((i & 1) ? O : E) + off, sp.nc, false, work + i * wl.total);
  }
  HIPCHK(hipMemcpyAsync(table, jobs, sp.k * sizeof(MulJob), hipMemcpyHostToDevice, s));
  JobRef jr{};
  jr.table = table;
  int rc = enqueue_group(jr, (unsigned)sp.k, sp.nc, false, s);
  if (rc) return rc;
  hgx_launch_addsub(false, s, E, n, O, n, d_out, n, agg, agg + nblk, nullptr);
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

// The split product into D_OUT on stream S in the library workspace; the
// caller holds the workspace.
int split_product_locked(const u64 *d_a, size_t na, const u64 *d_b, size_t nb, u64 *d_out, const Split &sp,
                         hipStream_t s) {
  const SplitBytes sb = split_bytes(na, nb, sp);
  if (g_ctx.ws.bytes < sb.main) HIPCHK(hipStreamSynchronize(s));  // this call's earlier work may use it
  int rc = ensure(g_ctx.ws, sb.main);
  if (rc) return rc;
  if (g_ctx.scan.bytes < sb.scan) HIPCHK(hipStreamSynchronize(s));
  rc = ensure(g_ctx.scan, sb.scan);
  if (rc) return rc;
  rc = ensure_jobs_host(sp.k);
  if (rc) return rc;
  rc = split_product_at(d_a, na, d_b, nb, d_out, sp, s, (char *)g_ctx.ws.ptr, (uint8_t *)g_ctx.scan.ptr,
                        g_ctx.jobs_host);
  if (rc) return rc;
  HIPCHK(hipEventRecord(g_ctx.jobs_copied, s));
  return HGX_OK;
}

int mul_device_locked(const u64 *d_a, size_t na, const u64 *d_b, size_t nb, u64 *d_out, bool sqr, hipStream_t s) {
  const int n = log2_ceil((u64)hgx_product_length(na, nb));
  if (n > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32");
  Split sp;
  if (!sqr && g_ctx.split_on && choose_split(na, nb, sp)) {
    int rc = ws_acquire(s);
    if (rc) return rc;
    rc = split_product_locked(d_a, na, d_b, nb, d_out, sp, s);
    if (rc) return rc;
    return ws_release(s);
  }
  const WorkLayout wl = work_layout(n, sqr);
  int rc = ws_acquire(s);
  if (rc) return rc;
  rc = ensure(g_ctx.ws, wl.total);
  if (rc) return rc;
  JobRef jr{};
  jr.one = make_job(d_a, na, d_b, nb, d_out, n, sqr, (char *)g_ctx.ws.ptr);
  rc = enqueue_group(jr, 1, n, sqr, s);
  if (rc) return rc;
  return ws_release(s);
}

// ---- products beyond the memory available to them (host-staged) ----
// A host-memory product whose device working set (both operands, the
// product, one transform's work arrays and its root tables) exceeds the
// memory it may use -- the HBM budget left, capped by hgx_staging_limit -- is
// computed from M-limb chunks a_i, b_j of the operands.  The product is
// produced diagonal by diagonal: c_s = sum over i + j = s of a_i * b_j, each
// chunk product on the GPU (transform length 2M) and summed there by the
// carry-scan addition into a window W of 2M + 2 limbs that starts at limb
// s*M.  After diagonal s the low M limbs of W are final and go to the host;
// the next diagonal's first addition reads W from limb M on into the other
// window.  With D pairs per diagonal, W < 2D * 2^(128M) by induction
// (W_s = c_s + floor(W_(s-1) / 2^(64M)), c_s < D * 2^(128M)), so 2M + 1 limbs
// hold it.  Only two chunks, one chunk product, the windows and one
// transform's work arrays are in HBM at a time; operands and product stay in
// host memory.  This takes the place of the reference's disk-backed vectors
// for numbers beyond memory (src/ram-vec.lisp:31-97, src/disk-vec.lisp:78-91):
// the result is the same exact product.
constexpr u64 kStageMinChunk = 1ull << 12;
constexpr size_t kStageSlack = (size_t)8 << 20;  // allocation granules, job tables

// device bytes of the root tables of length 2^n still to be built
size_t plan_bytes_needed(int n) {
  const Choice c = choice_for(n);
  if (g_ctx.plans.count(plan_key(n, c))) return 0;
  size_t words = 0, big = 0;
  for (int i = 0; i < c.ncol; i++) {
    const int b = n - 8 * i;
    const size_t B = (size_t)1 << b, tables = (i == 0) ? 3 : 2;
    if (b <= c.dm) {
      words += 3 * tables * 2 * B;
    } else {
      words += 3 * (tables * 2 * 4096 + 4 * (B >> 4));
      if (c.md && b <= kMontDirectMaxBits) big += 3 * 2 * B * sizeof(u64);
    }
  }
  return align_up(words * sizeof(u64)) + big;
}

// device bytes of the work memory of one product (its transform's, or for an
// unbalanced product that mul_device_locked splits, the chunk products' two
// images, job table and work memory and the scan scratch of the image sum),
// root tables still to be built included
size_t work_bytes(size_t na, size_t nb, bool sqr) {
  Split sp;
  if (!sqr && g_ctx.split_on && choose_split(na, nb, sp)) {
    const SplitBytes sb = split_bytes(na, nb, sp);
    return sb.main + sb.scan + plan_bytes_needed(sp.nc);
  }
  const int n = log2_ceil((u64)hgx_product_length(na, nb));
  return work_layout(n, sqr).total + plan_bytes_needed(n);
}

// device bytes of one product computed whole from host memory: operands,
// product and work memory
size_t inhbm_bytes(size_t na, size_t nb, bool sqr) {
  return align_up(na * 8) + (sqr ? 0 : align_up(nb * 8)) + align_up((na + nb) * 8) + work_bytes(na, nb, sqr);
}

// Device memory of a staged product with M-limb chunks: two operand chunks,
// the chunk product, two result windows, the work memory of one transform of
// length 2M and the scan scratch of the window additions.
struct StagedLayout {
  int n;
  size_t ca, cb, P, W0, W1, work, scan, total;
};
StagedLayout staged_layout(u64 M) {
  StagedLayout l{};
  l.n = log2_ceil(2 * M);
  size_t at = 0;
  l.ca = at;
  at += align_up(M * 8);
  l.cb = at;
  at += align_up(M * 8);
  l.P = at;
  at += align_up(2 * M * 8);
  l.W0 = at;
  at += align_up((2 * M + 2) * 8);
  l.W1 = at;
  at += align_up((2 * M + 2) * 8);
  l.work = at;
  at += work_layout(l.n, false).total;
  l.scan = at;
  at += align_up(2 * hgx_addsub_blocks(2 * M + 2));
  l.total = at;
  return l;
}

// device bytes of the staged product with M-limb chunks, root tables included
size_t staged_bytes(u64 M) { return staged_layout(M).total + plan_bytes_needed(log2_ceil(2 * M)); }

// device memory a host-memory product may use: the budget not held by other
// allocations, capped by the staging limit.  With REUSE the library's
// workspaces, which the product reuses, count as free.
size_t memory_for_product(bool reuse = true) {
  const size_t reusable = reuse ? g_ctx.ws.bytes + g_ctx.stage.bytes + g_ctx.scan.bytes : 0;
  const size_t held = g_ctx.allocated - reusable;
  size_t avail = g_ctx.budget > held ? g_ctx.budget - held : 0;
  const size_t limit = g_ctx.stage_limit.load();
  if (limit) avail = std::min(avail, limit);
  return avail;
}

bool needs_staging(size_t na, size_t nb, bool sqr) {
  return inhbm_bytes(na, nb, sqr) + kStageSlack > memory_for_product();
}

// the largest power-of-two chunk (at most the longer operand rounded up)
// whose staged working set fits AVAIL bytes; 0 when none does
u64 staged_chunk(size_t na, size_t avail) {
  u64 M = 1ull << log2_ceil(na);
  while (M > kStageMinChunk && staged_bytes(M) + kStageSlack > avail) M >>= 1;
  return staged_bytes(M) + kStageSlack > avail ? 0 : M;
}

// The staged product with M-limb chunks on stream S, through the pinned
// staging PIN, in the device memory at BASE (staged_layout(M).total bytes).
// g_ctx.mu is taken only around the kernel launches of each chunk product, so
// a caller that does not hold it (the request server's stager thread) leaves
// the library free between them.  Returns once the product is in OUT.
int mul_staged_at(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, bool sqr, u64 M, hipStream_t s,
                  PinStage &pin, char *base) {
  if (na < nb) {
    std::swap(a, b);
    std::swap(na, nb);
  }
  const StagedLayout lay = staged_layout(M);
  const u64 ka = (na + M - 1) / M, kb = (nb + M - 1) / M, nout = na + nb, wlen = 2 * M + 2;
  u64 *ca = (u64 *)(base + lay.ca), *cbf = (u64 *)(base + lay.cb), *P = (u64 *)(base + lay.P);
  u64 *W[2] = {(u64 *)(base + lay.W0), (u64 *)(base + lay.W1)};
  uint8_t *agg = (uint8_t *)(base + lay.scan), *blkin = agg + hgx_addsub_blocks(wlen);
  char *work = base + lay.work;
  u64 in_a = ~0ull, in_b = ~0ull;  // chunks now in the device buffers
  int cur = 0, rc = HGX_OK;
  PinScope scope(pin);
  for (u64 d = 0; d + 1 < ka + kb; d++) {
    u64 *Wc = W[cur], *Wp = W[cur ^ 1];
    bool first = true;
    const u64 ilo = d >= kb - 1 ? d - (kb - 1) : 0, ihi = std::min<u64>(d, ka - 1);
    for (u64 i = ilo; i <= ihi; i++) {
      const u64 j = d - i;
      if (sqr && j < i) break;  // a square takes each unordered pair once
      const u64 mi = std::min<u64>(M, na - i * M), mj = std::min<u64>(M, nb - j * M);
      const bool sq = sqr && i == j;
      if (in_a != i) {
        rc = h2d(pin, ca, a + i * M, mi * 8, s);
        if (rc) return rc;
        in_a = i;
      }
      if (!sq && in_b != j) {
        rc = h2d(pin, cbf, b + j * M, mj * 8, s);
        if (rc) return rc;
        in_b = j;
      }
      std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
      // the chunk product (balanced: never split) in this product's own work memory
      const u64 mo = sq ? mi : mj;
      const int n = log2_ceil(mi + mo);
      JobRef jr{};
      jr.one = make_job(ca, mi, sq ? ca : cbf, mo, P, n, sq, work);
      rc = enqueue_group(jr, 1, n, sq, s);
      if (rc) return rc;
      for (int t = 0; t < ((sqr && !sq) ? 2 : 1); t++) {
        HGX_TIMED("stage_add", s);
        if (first) {  // W = (previous window from limb M) + P
          hgx_launch_addsub(false, s, d ? Wp + M : nullptr, d ? M + 2 : 0, P, mi + mo, Wc, wlen, agg, blkin, nullptr);
          first = false;
        } else {
          hgx_launch_addsub(false, s, Wc, wlen, P, mi + mo, Wc, wlen, agg, blkin, nullptr);
        }
      }
      HIPCHK(hipGetLastError());
    }
    const u64 at = d * M;
    const u64 cnt = (d + 2 == ka + kb) ? nout - at : M;
    rc = d2h(pin, out + at, Wc, cnt * 8, s);
    if (rc) return rc;
    cur ^= 1;
  }
  rc = pin_flush(pin);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return HGX_OK;
}

// A staged product for the library's synchronous calls (the caller holds
// g_ctx.mu): the library stream, its pinned staging and its staging workspace.
int mul_staged_locked(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, bool sqr) {
  const size_t avail = memory_for_product();
  const u64 M = staged_chunk(std::max(na, nb), avail);
  if (!M)
    return fail(HGX_ENOMEM, "HBM budget too small even for a host-staged product: " + std::to_string(avail) +
                                " bytes available, " + std::to_string(staged_bytes(kStageMinChunk) + kStageSlack) +
                                " needed");
  int rc = ensure(g_ctx.stage, staged_layout(M).total);
  if (rc) return rc;
  return mul_staged_at(a, na, b, nb, out, sqr, M, g_ctx.stream, g_pin, (char *)g_ctx.stage.ptr);
}

// Replay the captured launch sequence for KEY on stream S, capturing it on
// first use (PREPARE runs first, outside the capture).  The job table the
// kernels read is copied before every replay, so a graph depends only on the
// shape of the chunk and on the table and work-memory addresses in KEY.
template <class Prepare, class Enqueue>
int launch_graph(const std::vector<u64> &key, hipStream_t s, Prepare prepare, Enqueue enqueue) {
  auto it = g_ctx.graphs.find(key);
  if (it == g_ctx.graphs.end()) {
    int rc = prepare();
    if (rc) return rc;
    if (g_ctx.graphs.size() >= 64) {  // bounded cache
      for (auto &kv : g_ctx.graphs) (void)hipGraphExecDestroy(kv.second);
      g_ctx.graphs.clear();
    }
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueue();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(HGX_EDEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return fail(HGX_EDEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(e2));
    it = g_ctx.graphs.emplace(key, exec).first;
  }
  HIPCHK(hipGraphLaunch(it->second, s));
  return HGX_OK;
}

// Batched products of device operands: grouped by (transform length,
// square), one launch sequence per group over a device job table.  Groups
// whose work memory exceeds the budget are split.
int batch_device_locked(const hgx_mul_desc *descs, size_t count, hipStream_t s) {
  struct Item { size_t idx; int n; bool sqr; size_t bytes; };
  // one work layout per (length, square) key: key = 2 n + sqr
  constexpr int kKeys = 2 * (kMaxBits + 1);
  WorkLayout layout[kKeys];
  bool have[kKeys] = {};
  size_t per_key[kKeys + 1] = {};
  std::vector<Item> items;
  std::vector<std::pair<size_t, Split>> splits;  // unbalanced products, done first
  items.reserve(count);
  for (size_t i = 0; i < count; i++) {
    const hgx_mul_desc &d = descs[i];
    if (!d.a || !d.b || !d.out || d.na == 0 || d.nb == 0) return fail(HGX_EINVAL, "bad descriptor " + std::to_string(i));
    const bool sqr = (d.a == d.b && d.na == d.nb);
    const int n = log2_ceil((u64)hgx_product_length(d.na, d.nb));
    if (n > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32 in descriptor " + std::to_string(i));
    Split sp;
    if (!sqr && g_ctx.split_on && choose_split(d.na, d.nb, sp)) {
      splits.push_back({i, sp});
      continue;
    }
    const int k = 2 * n + (sqr ? 1 : 0);
    if (!have[k]) {
      layout[k] = work_layout(n, sqr);
      have[k] = true;
    }
    per_key[k + 1]++;
    items.push_back({i, n, sqr, layout[k].total});
  }
  {
    // items in order of (length, square), stable: a counting sort over the
    // keys (linear in the batch, which can hold ~10^6 small products)
    for (int k = 0; k < kKeys; k++) per_key[k + 1] += per_key[k];
    std::vector<Item> sorted(items.size());
    for (const Item &it : items) sorted[per_key[2 * it.n + (it.sqr ? 1 : 0)]++] = it;
    items.swap(sorted);
  }
  int rc = ws_acquire(s);
  if (rc) return rc;
  for (auto &sv : splits) {
    const hgx_mul_desc &d = descs[sv.first];
    rc = split_product_locked(d.a, d.na, d.b, d.nb, d.out, sv.second, s);
    if (rc) return rc;
  }
  size_t pos = 0;
  while (pos < items.size()) {
    // a chunk: consecutive items that fit the budget together
    const size_t budget_left = g_ctx.budget - (g_ctx.allocated - g_ctx.ws.bytes - g_ctx.jobs_dev.bytes);
    size_t end = pos, bytes = 0;
    while (end < items.size()) {
      const size_t add = items[end].bytes;
      const size_t tbl = align_up((end - pos + 1) * sizeof(MulJob));
      if (end > pos && bytes + add + tbl + (2u << 20) > budget_left) break;
      bytes += add;
      end++;
    }
    if (g_ctx.ws.bytes < bytes || g_ctx.jobs_dev.bytes < align_up((end - pos) * sizeof(MulJob)))
      HIPCHK(hipStreamSynchronize(s));  // earlier work of this call may still use them
    rc = ensure(g_ctx.ws, bytes);
    if (rc) return rc;
    rc = ensure(g_ctx.jobs_dev, align_up((end - pos) * sizeof(MulJob)));
    if (rc) return rc;
    // fill the job table (pinned) and copy it to the device
    rc = ensure_jobs_host(end - pos);
    if (rc) return rc;
    char *base = (char *)g_ctx.ws.ptr;
    for (size_t i = pos; i < end; i++) {
      const hgx_mul_desc &d = descs[items[i].idx];
      g_ctx.jobs_host[i - pos] =
          make_job(d.a, d.na, d.b, d.nb, d.out, items[i].sqr, layout[2 * items[i].n + (items[i].sqr ? 1 : 0)], base);
      base += items[i].bytes;
    }
    MulJob *table = (MulJob *)g_ctx.jobs_dev.ptr;
    HIPCHK(hipMemcpyAsync(table, g_ctx.jobs_host, (end - pos) * sizeof(MulJob), hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(g_ctx.jobs_copied, s));
    // the chunk's groups: (length, square, count) runs of the sorted items
    std::vector<u64> key{(u64)(uintptr_t)table, (u64)(uintptr_t)g_ctx.ws.ptr};
    for (size_t g = pos; g < end;) {
      size_t h = g;
      while (h < end && items[h].n == items[g].n && items[h].sqr == items[g].sqr) h++;
      key.push_back(((u64)items[g].n << 1) | (items[g].sqr ? 1u : 0u));
      key.push_back(h - g);
      g = h;
    }
    auto enqueue_chunk = [&]() -> int {
      size_t g = pos;
      while (g < end) {
        size_t h = g;
        while (h < end && items[h].n == items[g].n && items[h].sqr == items[g].sqr) h++;
        JobRef jr{};
        jr.table = table + (g - pos);
        int r = enqueue_group(jr, (unsigned)(h - g), items[g].n, items[g].sqr, s);
        if (r) return r;
        g = h;
      }
      return (int)HGX_OK;
    };
    if (g_ctx.graphs_on && !g_timing) {
      rc = launch_graph(key, s, [&]() -> int {
        // root tables first: building them is host work and copies
        for (size_t i = pos; i < end; i++) {
          Plan *pl = nullptr;
          int r = build_plan(items[i].n, &pl);
          if (r) return r;
        }
        return (int)HGX_OK;
      }, enqueue_chunk);
    } else {
      rc = enqueue_chunk();
    }
    if (rc) return rc;
    pos = end;
    if (pos < items.size()) {
      // the next chunk reuses the work memory
      HIPCHK(hipStreamSynchronize(s));
    }
  }
  return ws_release(s);
}

// ---- request server (continuous batching) ----
//
// hgx_submit queues a request and returns.  A server thread admits queued
// requests in arrival order while their memory (operands, product, work
// arrays) fits in a device arena carved from the HBM budget (a round takes at
// most a 1/lanes share of the arena), groups each admission round by
// transform length into batched launches, and puts each round on the lane
// with the least work in flight, so rounds overlap on the device.  A lane is
// the stream of a round's kernels; the copies go to upload and download
// streams shared by the lanes (or, in the lane layout, run on the lane's
// stream in order with its kernels).  A request that arrives
// while others run joins the next round as soon as memory frees up.
// Requests larger than the arena go to a stager thread, which computes them
// host-staged on a stream of its own and holds the library only while it
// launches each chunk product, so rounds keep being admitted meanwhile.
// Completion is observed from events; the arena block is then returned and
// the ticket completes.  A ticket is released by hgx_wait, or by the hgx_poll
// call that reports its completion.
struct SrvReq {
  hgx_mul_desc d{};
  bool sqr = false;
  bool split = false;  // unbalanced: chunk products in the request's own arena block
  Split sp{};
  int n = 0;
  size_t bytes = 0, off = 0;
  size_t table_off = (size_t)-1, table_bytes = 0;  // job table of its group (last request of a group owns it)
  // host copy of that table (or of a split product's chunk table), alive
  // until the request completes: pinned, from the server's table pool, so
  // its upload is asynchronous (a copy from pageable memory would hold the
  // server thread until the stream reached it); table_spill when the pool
  // is exhausted
  MulJob *table_host = nullptr;
  size_t host_off = (size_t)-1, host_bytes = 0;
  std::vector<MulJob> table_spill;
  hipEvent_t done_ev = nullptr;
  bool ev_ok = true;  // done_ev was recorded after the request's last work
  int lane = -1;      // lane of its round
  int state = 0;      // 0 queued, 1 in flight, 2 done, <0 failed
  // copy streams: the end of its round's uploads (held by the round's first
  // request), the end of its kernels, and the download stream once the
  // download is issued (-1 before)
  hipEvent_t up_ev = nullptr, kern_ev = nullptr;
  bool kern_ok = true;     // kern_ev was recorded after its kernels
  bool lane_done = false;  // its kernels finished: no longer counted on the lane
  int up = -1, dl = -1;
  std::string err;
};

// Stream layouts.  By default the copies run on streams shared by the
// lanes: all uploads of a round go to an upload stream, the round's kernels
// wait for them on a compute lane, and each product's download is issued on
// the less loaded of two download streams once its kernels are seen to be
// done.  The copy engines move this mix at 48 GB/s each way with one or two
// streams per direction, against 38 GB/s when six streams each run uploads
// and downloads in turn.  Served with 32 requests outstanding, a saturated
// mixed-size stream runs at 4.0-4.3e9 limbs/s in this layout against
// 3.8e9 with every copy of a round on its lane's stream, in order with the
// round's kernels (HGX_SERVE_COPY_STREAMS=0).  Downloads queued at launch
// behind the kernels' event (instead of issued once the kernels are seen
// done) served 2.2e9: a download then waits in its stream for kernels that
// end later than the next ones.  HGX_SERVE_SPLIT_D2H=1 (lane layout) gives
// each lane a second stream for its downloads; HGX_SERVE_UPLOADS /
// HGX_SERVE_DOWNLOADS (1..4) set the copy stream counts and HGX_SERVE_LANES
// (1..16) the compute lanes.
constexpr int kMaxLanes = 16;
constexpr int kMaxCopyStreams = 4;
constexpr int kDefaultLanes = 6;
constexpr int kDefaultSplitD2H = 0;
constexpr int kDefaultCopyStreams = 1;
constexpr int kDefaultUploads = 1;
constexpr int kDefaultDownloads = 2;
constexpr size_t kServeTableBytes = 4 << 20;  // pinned pool of the server's host job tables

struct Lane {
  hipStream_t s = nullptr;  // uploads and kernels
  hipStream_t d = nullptr;  // downloads (the same stream unless split)
  hipEvent_t k = nullptr;   // end of the latest round's kernels on s, for split downloads
  size_t pending = 0;       // arena bytes of the requests in flight on the lane
};

struct Server {
  bool started = false, stop = false;
  std::thread th, stager;
  std::mutex mu;
  std::condition_variable cv;        // request state changes and arrivals
  std::condition_variable stage_cv;  // requests for the stager
  std::deque<std::shared_ptr<SrvReq>> queue, staged;
  std::vector<std::shared_ptr<SrvReq>> inflight;
  std::unordered_map<hgx_ticket, std::shared_ptr<SrvReq>> tickets;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  std::map<size_t, size_t> free_list;  // offset -> size
  char *tables = nullptr;               // pinned host pool of job tables
  std::map<size_t, size_t> table_free;  // offset -> size in it
  int nlanes = kDefaultLanes;
  bool split_d2h = kDefaultSplitD2H;
  Lane lanes[kMaxLanes];
  bool copy_streams = kDefaultCopyStreams;
  int nup = kDefaultUploads, ndown = kDefaultDownloads;
  hipStream_t up[kMaxCopyStreams] = {}, down[kMaxCopyStreams] = {};
  size_t up_pending[kMaxCopyStreams] = {}, down_pending[kMaxCopyStreams] = {};  // bytes not yet seen done
  hipStream_t stage_stream = nullptr;  // the stager's
  PinStage stage_pin;                  // the stager's pinned staging
  std::vector<hipEvent_t> event_pool;
  // statistics since the server started
  std::chrono::steady_clock::time_point t0;
  uint64_t completed = 0, failed = 0, rounds = 0, staged_done = 0;
  double launch_s = 0;  // server thread time spent issuing rounds
  double retire_s = 0;  // ... observing completions
  double wait_s = 0;    // ... waiting for arrivals or completions
  // failure injection for tests: the FAIL_COPY-th copy the server issues is
  // reported failed without being issued (HGX_SERVE_FAIL_COPY; 0: none)
  uint64_t fail_copy = 0, copies = 0;
};

Server g_srv;

// First-fit ranges of a pool (offset -> size of each free range).
size_t range_alloc(std::map<size_t, size_t> &fl, size_t bytes) {
  bytes = align_up(bytes);
  for (auto it = fl.begin(); it != fl.end(); ++it) {
    if (it->second >= bytes) {
      const size_t off = it->first, sz = it->second;
      fl.erase(it);
      if (sz > bytes) fl[off + bytes] = sz - bytes;
      return off;
    }
  }
  return (size_t)-1;
}

void range_free(std::map<size_t, size_t> &fl, size_t off, size_t bytes) {
  bytes = align_up(bytes);
  auto it = fl.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != fl.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    fl.erase(nx);
  }
  if (it != fl.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      fl.erase(it);
    }
  }
}

size_t arena_alloc(size_t bytes) { return range_alloc(g_srv.free_list, bytes); }
void arena_free(size_t off, size_t bytes) { range_free(g_srv.free_list, off, bytes); }

// Host storage of COUNT jobs for request R, kept until it completes (caller
// holds g_srv.mu).
MulJob *host_table(SrvReq &r, size_t count) {
  const size_t bytes = count * sizeof(MulJob);
  const size_t off = g_srv.tables ? range_alloc(g_srv.table_free, bytes) : (size_t)-1;
  if (off != (size_t)-1) {
    r.host_off = off;
    r.host_bytes = bytes;
    r.table_host = (MulJob *)(g_srv.tables + off);
  } else {
    r.table_spill.resize(count);
    r.table_host = r.table_spill.data();
  }
  return r.table_host;
}

size_t request_bytes(const SrvReq &r) {
  const size_t in = align_up(r.d.na * 8) + (r.sqr ? 0 : align_up(r.d.nb * 8));
  if (r.split) {
    const SplitBytes sb = split_bytes(r.d.na, r.d.nb, r.sp);
    return in + align_up((r.d.na + r.d.nb) * 8) + sb.main + sb.scan;
  }
  return in + align_up((r.d.na + r.d.nb) * 8) + work_layout(r.n, r.sqr).total;
}

// a completion event from the pool (null when none can be created)
hipEvent_t take_event() {
  if (!g_srv.event_pool.empty()) {
    hipEvent_t e = g_srv.event_pool.back();
    g_srv.event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

// A copy of the server thread; on failure R carries the error and false is
// returned.
bool srv_copy(SrvReq &r, void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  const uint64_t k = ++g_srv.copies;
  const hipError_t e = (k == g_srv.fail_copy) ? hipErrorInvalidValue : hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e == hipSuccess) return true;
  (void)hipGetLastError();
  if (r.err.empty())
    r.err = std::string(kind == hipMemcpyHostToDevice ? "upload" : "download") + " of a served request failed: " +
            hipGetErrorString(e);
  return false;
}

// Launch one admission round on lane LI, its uploads on upload stream UI
// with copy streams (called without g_srv.mu held).  Every failure lands in
// the error of the requests it concerns; a request with an error is not
// downloaded, and completes as failed.  With copy streams the products'
// downloads are issued later (issue_download), once their kernels are done.
void launch_round(std::vector<std::shared_ptr<SrvReq>> &batch, int li, int ui) {
  Lane &ln = g_srv.lanes[li];
  const hipStream_t s = ln.s;
  const bool cstreams = g_srv.copy_streams;
  const hipStream_t cs = cstreams ? g_srv.up[ui] : s;  // copies in
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  // unbalanced products (split into chunk products) after the grouped ones
  std::vector<std::shared_ptr<SrvReq>> splits;
  {
    std::vector<std::shared_ptr<SrvReq>> grouped;
    for (auto &r : batch) (r->split ? splits : grouped).push_back(r);
    batch = grouped;
  }
  std::stable_sort(batch.begin(), batch.end(), [](const std::shared_ptr<SrvReq> &x, const std::shared_ptr<SrvReq> &y) {
    return x->n != y->n ? x->n < y->n : x->sqr < y->sqr;
  });
  // operands in
  std::vector<MulJob> jobs(batch.size());
  for (size_t i = 0; i < batch.size(); i++) {
    SrvReq &r = *batch[i];
    char *base = g_srv.arena + r.off;
    u64 *da = (u64 *)base;
    u64 *db = r.sqr ? da : (u64 *)(base + align_up(r.d.na * 8));
    u64 *dout = (u64 *)(base + align_up(r.d.na * 8) + (r.sqr ? 0 : align_up(r.d.nb * 8)));
    char *work = (char *)dout + align_up((r.d.na + r.d.nb) * 8);
    if (srv_copy(r, da, r.d.a, r.d.na * 8, hipMemcpyHostToDevice, cs) && !r.sqr)
      srv_copy(r, db, r.d.b, r.d.nb * 8, hipMemcpyHostToDevice, cs);
    jobs[i] = make_job(da, r.d.na, db, r.d.nb, dout, r.n, r.sqr, work);
  }
  // the split products' operands in
  for (auto &rp : splits) {
    SrvReq &r = *rp;
    u64 *da = (u64 *)(g_srv.arena + r.off);
    u64 *db = (u64 *)(g_srv.arena + r.off + align_up(r.d.na * 8));
    if (srv_copy(r, da, r.d.a, r.d.na * 8, hipMemcpyHostToDevice, cs))
      srv_copy(r, db, r.d.b, r.d.nb * 8, hipMemcpyHostToDevice, cs);
  }
  // one launch sequence per (length, square) group, its job table (several
  // jobs) in the arena, uploaded with the operands; every job's memory lies
  // in the arena, so a request whose upload failed only computes garbage
  struct Group {
    size_t g, h;
    MulJob *dtab;  // null: one launch sequence per job
    bool table_failed;
    std::string err;
  };
  std::vector<Group> groups;
  for (size_t g = 0; g < batch.size();) {
    size_t h = g;
    while (h < batch.size() && batch[h]->n == batch[g]->n && batch[h]->sqr == batch[g]->sqr) h++;
    Group gr{g, h, nullptr, false, std::string()};
    size_t toff = (size_t)-1;
    SrvReq &owner = *batch[h - 1];
    if (h - g > 1) {
      std::lock_guard<std::mutex> lk2(g_srv.mu);
      toff = arena_alloc((h - g) * sizeof(MulJob));
      if (toff != (size_t)-1) host_table(owner, h - g);
    }
    if (toff != (size_t)-1) {
      owner.table_off = toff;
      owner.table_bytes = (h - g) * sizeof(MulJob);
      std::copy(jobs.begin() + g, jobs.begin() + h, owner.table_host);
      gr.dtab = (MulJob *)(g_srv.arena + toff);
      SrvReq probe;
      if (!srv_copy(probe, gr.dtab, owner.table_host, owner.table_bytes, hipMemcpyHostToDevice, cs)) {
        gr.table_failed = true;
        gr.err = probe.err;
      }
    }
    groups.push_back(gr);
    g = h;
  }
  // with copy streams the lane's kernels wait for the uploads; if the wait
  // cannot be queued, the server waits for the uploads itself
  if (cstreams) {
    hipEvent_t ue = nullptr;  // held by one request of the round
    for (auto *v : {&batch, &splits})
      for (auto &r : *v)
        if (!ue && r->up_ev) ue = r->up_ev;
    hipError_t e = ue ? hipEventRecord(ue, cs) : hipErrorInvalidHandle;
    if (e == hipSuccess) e = hipStreamWaitEvent(s, ue, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (hipStreamSynchronize(cs) != hipSuccess) {
        (void)hipGetLastError();
        for (auto *v : {&batch, &splits})
          for (auto &r : *v)
            if (r->err.empty()) r->err = "the uploads of a served round could not be awaited";
      }
    }
  }
  for (const Group &gr : groups) {
    if (gr.table_failed) {
      for (size_t i = gr.g; i < gr.h; i++)
        if (batch[i]->err.empty()) batch[i]->err = gr.err;
      continue;
    }
    int rc = HGX_OK;
    if (gr.dtab) {
      JobRef jr{};
      jr.table = gr.dtab;
      rc = enqueue_group(jr, (unsigned)(gr.h - gr.g), batch[gr.g]->n, batch[gr.g]->sqr, s);
    } else {
      for (size_t i = gr.g; i < gr.h && rc == HGX_OK; i++) {
        JobRef jr{};
        jr.one = jobs[i];
        rc = enqueue_group(jr, 1, batch[i]->n, batch[i]->sqr, s);
      }
    }
    if (rc) {
      for (size_t i = gr.g; i < gr.h; i++)
        if (batch[i]->err.empty()) batch[i]->err = g_err;
    }
  }
  // the split products: chunk products and the image sum in the request's
  // own arena block, so split products on different lanes overlap
  for (size_t i = 0; i < splits.size(); i++) {
    SrvReq &r = *splits[i];
    char *base = g_srv.arena + r.off;
    u64 *da = (u64 *)base;
    u64 *db = (u64 *)(base + align_up(r.d.na * 8));
    u64 *dout = (u64 *)(base + align_up(r.d.na * 8) + align_up(r.d.nb * 8));
    char *work = (char *)dout + align_up((r.d.na + r.d.nb) * 8);
    const SplitBytes sb = split_bytes(r.d.na, r.d.nb, r.sp);
    MulJob *tab;
    {
      std::lock_guard<std::mutex> lk2(g_srv.mu);
      tab = host_table(r, r.sp.k);
    }
    int rc = split_product_at(da, r.d.na, db, r.d.nb, dout, r.sp, s, work, (uint8_t *)(work + sb.main), tab);
    if (rc && r.err.empty()) r.err = g_err;
    batch.push_back(splits[i]);
  }
  if (cstreams) {
    // the end of every request's kernels; its download comes later
    for (auto &rp : batch) {
      SrvReq &r = *rp;
      const hipError_t e = hipEventRecord(r.kern_ev, s);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        r.kern_ok = false;
        if (r.err.empty()) r.err = std::string("recording the kernels' end failed: ") + hipGetErrorString(e);
      }
    }
    return;
  }
  // products out and completion events (the product follows the operands in
  // every request's arena block), on the download stream after the kernels
  hipStream_t ds = s;
  if (ln.d != ln.s) {
    hipError_t e = hipEventRecord(ln.k, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(ln.d, ln.k, 0);
    if (e == hipSuccess)
      ds = ln.d;
    else
      (void)hipGetLastError();  // downloads stay on the upload stream, in order
  }
  for (size_t i = 0; i < batch.size(); i++) {
    SrvReq &r = *batch[i];
    const u64 *dout = (const u64 *)(g_srv.arena + r.off + align_up(r.d.na * 8) + (r.sqr ? 0 : align_up(r.d.nb * 8)));
    if (r.err.empty()) srv_copy(r, r.d.out, dout, (r.d.na + r.d.nb) * 8, hipMemcpyDeviceToHost, ds);
    const hipError_t e = hipEventRecord(r.done_ev, ds);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      r.ev_ok = false;
      if (r.err.empty()) r.err = std::string("recording the completion event failed: ") + hipGetErrorString(e);
    }
  }
}

// Return a finished request's memory and complete its ticket (g_srv.mu held).
size_t product_bytes(const SrvReq &r) { return (r.d.na + r.d.nb) * 8; }

// the request's kernels are done (or failed): it no longer counts on its
// lane or its upload stream (g_srv.mu held)
void release_lane_locked(SrvReq &r) {
  if (r.lane_done) return;
  r.lane_done = true;
  g_srv.lanes[r.lane].pending -= r.bytes;
  if (r.up >= 0) g_srv.up_pending[r.up] -= r.bytes;
}

void finish_locked(SrvReq &r) {
  arena_free(r.off, r.bytes);
  if (r.table_off != (size_t)-1) arena_free(r.table_off, r.table_bytes);
  if (r.host_off != (size_t)-1) range_free(g_srv.table_free, r.host_off, r.host_bytes);
  release_lane_locked(r);
  if (r.dl >= 0) g_srv.down_pending[r.dl] -= product_bytes(r);
  for (hipEvent_t *e : {&r.done_ev, &r.kern_ev, &r.up_ev}) {
    if (*e) g_srv.event_pool.push_back(*e);
    *e = nullptr;
  }
  r.state = r.err.empty() ? 2 : HGX_EDEVICE;
  (r.state == 2 ? g_srv.completed : g_srv.failed)++;
}

// Wait for every stream a request may still use (after a failed event).
void sync_request_streams(const SrvReq &r) {
  const Lane &ln = g_srv.lanes[r.lane];
  (void)hipStreamSynchronize(ln.s);
  if (ln.d != ln.s) (void)hipStreamSynchronize(ln.d);
  if (r.up >= 0) (void)hipStreamSynchronize(g_srv.up[r.up]);
  if (r.dl >= 0) (void)hipStreamSynchronize(g_srv.down[r.dl]);
  (void)hipGetLastError();
}

// Retire the requests whose work has finished (g_srv.mu held).  A request
// whose completion event was not recorded or reports an error waits for the
// streams it used before its memory is returned, and fails.  With copy
// streams, a request whose kernels are seen done gets a download stream (the
// one with fewer bytes in flight) and is returned, for the caller to issue
// its download (issue_download) without the server's lock.
std::vector<std::shared_ptr<SrvReq>> retire_locked() {
  std::vector<std::shared_ptr<SrvReq>> downloads;
  bool any = false;
  for (size_t i = 0; i < g_srv.inflight.size();) {
    SrvReq &r = *g_srv.inflight[i];
    const bool kernels = g_srv.copy_streams && r.dl < 0;  // waiting for its kernels
    const hipError_t q = kernels ? (r.kern_ok ? hipEventQuery(r.kern_ev) : hipErrorInvalidHandle)
                                 : (r.ev_ok ? hipEventQuery(r.done_ev) : hipErrorInvalidHandle);
    if (q == hipErrorNotReady) {
      i++;
      continue;
    }
    if (q == hipSuccess && kernels) {
      release_lane_locked(r);
      int d = 0;
      for (int k = 1; k < g_srv.ndown; k++)
        if (g_srv.down_pending[k] < g_srv.down_pending[d]) d = k;
      r.dl = d;
      g_srv.down_pending[d] += product_bytes(r);
      downloads.push_back(g_srv.inflight[i]);
      i++;
      continue;
    }
    if (q != hipSuccess) {
      (void)hipGetLastError();
      if (r.err.empty())
        r.err = std::string(kernels ? "the kernels' end event failed: " : "completion event failed: ") +
                hipGetErrorString(q);
      sync_request_streams(r);
    }
    finish_locked(r);
    g_srv.inflight[i] = g_srv.inflight.back();
    g_srv.inflight.pop_back();
    any = true;
  }
  if (any) g_srv.cv.notify_all();
  return downloads;
}

// The download of a request whose kernels are done, on its download stream,
// and its completion event (server thread, without g_srv.mu: a download into
// pageable memory may take the copy's time).  A request with an error is not
// downloaded.
void issue_download(SrvReq &r) {
  const hipStream_t d = g_srv.down[r.dl];
  const u64 *dout = (const u64 *)(g_srv.arena + r.off + align_up(r.d.na * 8) + (r.sqr ? 0 : align_up(r.d.nb * 8)));
  if (r.err.empty()) srv_copy(r, r.d.out, dout, product_bytes(r), hipMemcpyDeviceToHost, d);
  const hipError_t e = hipEventRecord(r.done_ev, d);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    r.ev_ok = false;
    if (r.err.empty()) r.err = std::string("recording the completion event failed: ") + hipGetErrorString(e);
  }
}

void server_loop() {
  std::unique_lock<std::mutex> lk(g_srv.mu);
  using clock = std::chrono::steady_clock;
  auto since = [](clock::time_point t) { return std::chrono::duration<double>(clock::now() - t).count(); };
  while (!g_srv.stop) {
    const auto tr = clock::now();
    auto downloads = retire_locked();
    g_srv.retire_s += since(tr);
    if (!downloads.empty()) {
      lk.unlock();
      const auto t = std::chrono::steady_clock::now();
      for (auto &r : downloads) issue_download(*r);
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
      lk.lock();
      g_srv.launch_s += dt;
    }
    // admit in arrival order while memory lasts, at most a 1/lanes share of
    // the arena per round: under load several rounds are in flight on the
    // lanes, so one round's uploads and downloads overlap another's kernels
    std::vector<std::shared_ptr<SrvReq>> round;
    size_t round_bytes = 0;
    const size_t share = g_srv.arena_bytes / g_srv.nlanes;
    while (!g_srv.queue.empty()) {
      auto r = g_srv.queue.front();
      const size_t limit = g_ctx.stage_limit.load();
      if (r->bytes > g_srv.arena_bytes || (limit && r->bytes > limit)) {
        // larger than the arena: host-staged by the stager thread while the
        // rounds go on
        g_srv.queue.pop_front();
        r->state = 1;
        g_srv.staged.push_back(r);
        g_srv.stage_cv.notify_one();
        continue;
      }
      if (!round.empty() && round_bytes + r->bytes > share) break;
      const size_t off = arena_alloc(r->bytes);
      if (off == (size_t)-1) break;
      // completion event; with copy streams also the kernels' end and, for
      // the round's first request, the end of the round's uploads
      hipEvent_t ev = take_event();
      hipEvent_t kev = (ev && g_srv.copy_streams) ? take_event() : nullptr;
      hipEvent_t uev = (kev && round.empty()) ? take_event() : nullptr;
      g_srv.queue.pop_front();
      const bool ok = ev && (!g_srv.copy_streams || (kev && (!round.empty() || uev)));
      if (!ok) {
        for (hipEvent_t e : {ev, kev, uev})
          if (e) g_srv.event_pool.push_back(e);
        arena_free(off, r->bytes);
        r->err = "no completion event could be created";
        r->state = HGX_EDEVICE;
        g_srv.failed++;
        g_srv.cv.notify_all();
        continue;
      }
      r->off = off;
      r->state = 1;
      r->done_ev = ev;
      r->kern_ev = kev;
      r->up_ev = uev;
      round.push_back(r);
      round_bytes += r->bytes;
    }
    if (!round.empty()) {
      // the lane with the least work in flight (request sizes span four
      // orders of magnitude: rotating through the lanes would queue small
      // requests behind a large one while other lanes idle), and so for the
      // upload stream
      int li = 0, ui = -1;
      for (int i = 1; i < g_srv.nlanes; i++)
        if (g_srv.lanes[i].pending < g_srv.lanes[li].pending) li = i;
      if (g_srv.copy_streams) {
        ui = 0;
        for (int i = 1; i < g_srv.nup; i++)
          if (g_srv.up_pending[i] < g_srv.up_pending[ui]) ui = i;
      }
      for (auto &r : round) {
        r->lane = li;
        r->up = ui;
        g_srv.lanes[li].pending += r->bytes;
        if (ui >= 0) g_srv.up_pending[ui] += r->bytes;
        g_srv.inflight.push_back(r);
      }
      lk.unlock();
      const auto t = std::chrono::steady_clock::now();
      launch_round(round, li, ui);
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
      lk.lock();
      g_srv.launch_s += dt;
      g_srv.rounds++;
      continue;
    }
    const auto tw = clock::now();
    if (g_srv.inflight.empty())
      g_srv.cv.wait(lk, [] { return g_srv.stop || !g_srv.queue.empty(); });
    else
      g_srv.cv.wait_for(lk, std::chrono::microseconds(20));
    g_srv.wait_s += since(tw);
  }
}

// A request larger than the arena, host-staged on the stager's stream in
// device memory of its own (allocated for the request, freed after it).
int run_staged_request(const SrvReq &r) {
  char *base = nullptr;
  size_t bytes = 0;
  u64 M = 0;
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    const size_t avail = memory_for_product(false);
    M = staged_chunk(std::max(r.d.na, r.d.nb), avail);
    if (!M)
      return fail(HGX_ENOMEM, "HBM budget too small for a host-staged served product: " + std::to_string(avail) +
                                  " bytes available");
    Plan *plan = nullptr;
    int rc = build_plan(staged_layout(M).n, &plan);  // root tables before the work memory
    if (rc) return rc;
    bytes = staged_layout(M).total;
    rc = dev_alloc((void **)&base, bytes);
    if (rc) return rc;
  }
  const int rc = mul_staged_at(r.d.a, r.d.na, r.d.b, r.d.nb, r.d.out, r.sqr, M, g_srv.stage_stream, g_srv.stage_pin,
                               base);
  (void)hipStreamSynchronize(g_srv.stage_stream);  // also after a failure: nothing may still use BASE
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  dev_free(base, bytes);
  return rc;
}

void stager_loop() {
  std::unique_lock<std::mutex> lk(g_srv.mu);
  for (;;) {
    g_srv.stage_cv.wait(lk, [] { return g_srv.stop || !g_srv.staged.empty(); });
    if (g_srv.stop) return;
    auto r = g_srv.staged.front();
    g_srv.staged.pop_front();
    lk.unlock();
    const int rc = run_staged_request(*r);
    const std::string err = rc ? g_err : std::string();
    lk.lock();
    r->err = err;
    r->state = rc ? rc : 2;
    (rc ? g_srv.failed : g_srv.staged_done)++;
    if (!rc) g_srv.completed++;
    g_srv.cv.notify_all();
  }
}

void server_stop(bool at_exit = false);
// At process exit the HIP runtime (and any tool loaded into the process) may
// already be finalized, so only the server threads are stopped and joined;
// the device memory, streams and events go with the process.
void server_stop_at_exit() { server_stop(true); }

// streams, events and memory of the server (HIP calls; not at process exit)
void server_release() {
  for (int i = 0; i < kMaxLanes; i++) {
    Lane &ln = g_srv.lanes[i];
    if (ln.d && ln.d != ln.s) (void)hipStreamDestroy(ln.d);
    if (ln.s) (void)hipStreamDestroy(ln.s);
    if (ln.k) (void)hipEventDestroy(ln.k);
    ln = Lane{};
  }
  for (int i = 0; i < kMaxCopyStreams; i++) {
    if (g_srv.up[i]) (void)hipStreamDestroy(g_srv.up[i]);
    if (g_srv.down[i]) (void)hipStreamDestroy(g_srv.down[i]);
    g_srv.up[i] = g_srv.down[i] = nullptr;
    g_srv.up_pending[i] = g_srv.down_pending[i] = 0;
  }
  if (g_srv.stage_stream) (void)hipStreamDestroy(g_srv.stage_stream);
  g_srv.stage_stream = nullptr;
  pin_free(g_srv.stage_pin);
  for (auto e : g_srv.event_pool) (void)hipEventDestroy(e);
  g_srv.event_pool.clear();
  if (g_srv.arena) dev_free(g_srv.arena, g_srv.arena_bytes);
  g_srv.arena = nullptr;
  g_srv.arena_bytes = 0;
  g_srv.free_list.clear();
  if (g_srv.tables) (void)hipHostFree(g_srv.tables);
  g_srv.tables = nullptr;
  g_srv.table_free.clear();
}

long env_long(const char *name, long dflt) {
  const char *e = getenv(name);
  return (e && *e) ? strtol(e, nullptr, 10) : dflt;
}

int server_start_locked() {
  if (g_srv.started) return HGX_OK;
  static bool registered = false;
  if (!registered) {
    std::atexit(server_stop_at_exit);  // join the server threads before static destruction
    registered = true;
  }
  const long lanes = env_long("HGX_SERVE_LANES", kDefaultLanes);
  if (lanes < 1 || lanes > kMaxLanes) return fail(HGX_EINVAL, "HGX_SERVE_LANES must be 1..16");
  g_srv.nlanes = (int)lanes;
  g_srv.split_d2h = env_long("HGX_SERVE_SPLIT_D2H", kDefaultSplitD2H) != 0;
  g_srv.copy_streams = env_long("HGX_SERVE_COPY_STREAMS", kDefaultCopyStreams) != 0;
  const long nup = env_long("HGX_SERVE_UPLOADS", kDefaultUploads);
  const long ndown = env_long("HGX_SERVE_DOWNLOADS", kDefaultDownloads);
  if (nup < 1 || nup > kMaxCopyStreams || ndown < 1 || ndown > kMaxCopyStreams)
    return fail(HGX_EINVAL, "HGX_SERVE_UPLOADS and HGX_SERVE_DOWNLOADS must be 1..4");
  g_srv.nup = (int)nup;
  g_srv.ndown = (int)ndown;
  if (g_srv.copy_streams) g_srv.split_d2h = false;  // downloads are on the download streams
  g_srv.fail_copy = (uint64_t)std::max(0L, env_long("HGX_SERVE_FAIL_COPY", 0));
  size_t arena = 0;
  if (const char *e = getenv("HGX_SERVE_ARENA_BYTES")) arena = strtoull(e, nullptr, 10);
  if (!arena) arena = (g_ctx.budget - g_ctx.allocated) / 2;
  arena &= ~(size_t)255;
  int rc = dev_alloc((void **)&g_srv.arena, arena);
  if (rc) return rc;
  g_srv.arena_bytes = arena;
  g_srv.free_list.clear();
  g_srv.free_list[0] = arena;
  g_srv.table_free.clear();
  if (hipHostMalloc((void **)&g_srv.tables, kServeTableBytes, hipHostMallocDefault) == hipSuccess) {
    g_srv.table_free[0] = kServeTableBytes;
  } else {
    g_srv.tables = nullptr;  // job tables then come from pageable memory
    (void)hipGetLastError();
  }
  hipError_t e = hipSuccess;
  for (int i = 0; i < g_srv.nlanes && e == hipSuccess; i++) {
    Lane &ln = g_srv.lanes[i];
    e = hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking);
    ln.d = ln.s;
    if (e == hipSuccess && g_srv.split_d2h) {
      e = hipStreamCreateWithFlags(&ln.d, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ln.k, hipEventDisableTiming);
      if (e != hipSuccess) ln.d = ln.s;
    }
  }
  for (int i = 0; g_srv.copy_streams && i < g_srv.nup && e == hipSuccess; i++)
    e = hipStreamCreateWithFlags(&g_srv.up[i], hipStreamNonBlocking);
  for (int i = 0; g_srv.copy_streams && i < g_srv.ndown && e == hipSuccess; i++)
    e = hipStreamCreateWithFlags(&g_srv.down[i], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g_srv.stage_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    server_release();
    return fail(HGX_EDEVICE, std::string("creating the server's streams failed: ") + hipGetErrorString(e));
  }
  {
    std::lock_guard<std::mutex> lk(g_srv.mu);
    g_srv.stop = false;
    g_srv.completed = g_srv.failed = g_srv.rounds = g_srv.staged_done = g_srv.copies = 0;
    g_srv.launch_s = g_srv.retire_s = g_srv.wait_s = 0;
    g_srv.t0 = std::chrono::steady_clock::now();
    g_srv.started = true;
  }
  g_srv.th = std::thread(server_loop);
  g_srv.stager = std::thread(stager_loop);
  return HGX_OK;
}

// Stop the server.  Rounds already launched finish and their requests
// complete normally; requests not yet launched fail with HGX_ENOTINIT, so no
// hgx_wait waits forever.
void server_stop(bool at_exit) {
  {
    std::lock_guard<std::mutex> lk(g_srv.mu);
    if (!g_srv.started) return;
    g_srv.stop = true;
  }
  g_srv.cv.notify_all();
  g_srv.stage_cv.notify_all();
  if (g_srv.th.joinable()) g_srv.th.join();
  if (g_srv.stager.joinable()) g_srv.stager.join();
  if (at_exit) {
    std::lock_guard<std::mutex> lk(g_srv.mu);
    g_srv.started = false;
    return;
  }
  // launched rounds finish: kernels end, the downloads of products whose
  // kernels are done are issued, downloads end (at most three passes)
  for (int pass = 0; pass < 3; pass++) {
    for (int i = 0; i < g_srv.nlanes; i++) {
      (void)hipStreamSynchronize(g_srv.lanes[i].s);
      if (g_srv.lanes[i].d != g_srv.lanes[i].s) (void)hipStreamSynchronize(g_srv.lanes[i].d);
    }
    for (int i = 0; i < kMaxCopyStreams; i++) {
      if (g_srv.up[i]) (void)hipStreamSynchronize(g_srv.up[i]);
      if (g_srv.down[i]) (void)hipStreamSynchronize(g_srv.down[i]);
    }
    (void)hipGetLastError();
    std::vector<std::shared_ptr<SrvReq>> downloads;
    {
      std::lock_guard<std::mutex> lk(g_srv.mu);
      downloads = retire_locked();
    }
    if (downloads.empty()) break;
    for (auto &r : downloads) issue_download(*r);
  }
  {
    std::lock_guard<std::mutex> lk(g_srv.mu);
    for (auto &r : g_srv.inflight) {  // an event that never completes
      if (r->err.empty()) r->err = "the request's completion was never observed";
      finish_locked(*r);
    }
    g_srv.inflight.clear();
    for (auto *q : {&g_srv.queue, &g_srv.staged}) {
      for (auto &r : *q) {
        r->err = "the request server stopped (hgx_shutdown) before the request ran";
        r->state = HGX_ENOTINIT;
        g_srv.failed++;
      }
      q->clear();
    }
    g_srv.started = false;
    g_srv.cv.notify_all();
  }
  server_release();
}

// ---- pass-structure choices and the autotuner ----
void clear_graphs() {
  for (auto &kv : g_ctx.graphs) (void)hipGraphExecDestroy(kv.second);
  g_ctx.graphs.clear();
}

int set_choice_locked(int bits, const Choice &c) {
  const int ncol = c.ncol, dm = c.dm;
  if (bits < 1 || bits > kMaxBits) return fail(HGX_EINVAL, "transform bits out of range");
  if (ncol < min_ncol(bits) || ncol > max_ncol(bits))
    return fail(HGX_EINVAL, "2^" + std::to_string(bits) + " takes " + std::to_string(min_ncol(bits)) + ".." +
                                std::to_string(max_ncol(bits)) + " column passes");
  if (dm < 12 || dm > 24) return fail(HGX_EINVAL, "direct-table limit must be 12..24 bits");
  if (c.md != 0 && c.md != 1) return fail(HGX_EINVAL, "direct Montgomery flag must be 0 or 1");
  if (c.colw != 0 && c.colw != 256 && c.colw != 512) return fail(HGX_EINVAL, "column tile threads must be 0, 256 or 512");
  if (c.rowt != 0 && (c.rowt < 10 || c.rowt > 12)) return fail(HGX_EINVAL, "row tile must be 0 or 10..12 bits");
  // products of this length may still run on any stream with the old tables
  HIPCHK(hipDeviceSynchronize());
  const Choice old = choice_for(bits);
  g_ctx.choice[bits] = c;
  if (old != c) {
    clear_graphs();  // captured launches hold the old kernels
    if (!old.same_tables(c)) drop_plan(plan_key(bits, old));
  }
  return HGX_OK;
}

std::string device_arch() {
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, g_ctx.device) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  return a.substr(0, a.find(':'));
}

// tuned choices persist in the file named by HGX_TUNE_CACHE: "arch bits ncol dm md" lines
void tune_cache_load() {
  const char *path = getenv("HGX_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char line[256];
  const std::string mine = device_arch();
  while (fgets(line, sizeof line, f)) {
    char arch[64];
    int bits, ncol, dm, md, colw = 0, rowt = 0;
    const int got = sscanf(line, "%63s %d %d %d %d %d %d", arch, &bits, &ncol, &dm, &md, &colw, &rowt);
    if ((got == 5 || got == 7) && mine == arch && bits >= 1 && bits <= kMaxBits && ncol >= min_ncol(bits) &&
        ncol <= max_ncol(bits) && dm >= 12 && dm <= 24 && (md == 0 || md == 1) &&
        (colw == 0 || colw == 256 || colw == 512) && (rowt == 0 || (rowt >= 10 && rowt <= 12))) {
      Choice c{ncol, dm, md};
      c.colw = colw;
      c.rowt = rowt;
      g_ctx.choice[bits] = c;
    }
  }
  fclose(f);
}

void tune_cache_save() {
  const char *path = getenv("HGX_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "w");
  if (!f) return;
  const std::string arch = device_arch();
  for (auto &kv : g_ctx.choice)
    fprintf(f, "%s %d %d %d %d %d %d\n", arch.c_str(), kv.first, kv.second.ncol, kv.second.dm, kv.second.md,
            kv.second.colw, kv.second.rowt);
  fclose(f);
}

// ---- device limb vectors for callers without an allocator of their own ----
// A caching allocator: a freed block is kept by size class and handed out
// again once the stream work queued before its release has finished (an
// event recorded at the release).  Blocks count against the HBM budget;
// cached ones are returned to the device when an allocation would exceed it.
struct DevBlock {
  void *p;
  size_t bytes;
  hipEvent_t ev;
};
struct DevPool {
  std::multimap<size_t, DevBlock> cached;  // class size -> released block
  std::unordered_map<void *, size_t> live;  // handed-out block -> class size
  u64 *res = nullptr;                        // device words of the scans' results
  u64 *hres = nullptr;                       // pinned host copy of them
};
DevPool g_dev;

// size classes: powers of two up to 1 MiB, then eighths of a power of two
size_t dev_class(size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (bytes <= ((size_t)1 << 20)) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  const int e = 63 - __builtin_clzll((unsigned long long)bytes);
  const size_t g = (size_t)1 << (e - 3);
  return (bytes + g - 1) / g * g;
}

// return cached blocks to the device: those whose work has finished, or all
// of them after waiting (WAIT)
void dev_trim(bool wait) {
  for (auto it = g_dev.cached.begin(); it != g_dev.cached.end();) {
    DevBlock &b = it->second;
    if (wait) (void)hipEventSynchronize(b.ev);
    if (hipEventQuery(b.ev) == hipSuccess) {
      (void)hipEventDestroy(b.ev);
      dev_free(b.p, b.bytes);
      it = g_dev.cached.erase(it);
    } else {
      (void)hipGetLastError();
      ++it;
    }
  }
}

int dev_block_alloc(size_t bytes, void **out) {
  const size_t c = dev_class(bytes);
  auto range = g_dev.cached.equal_range(c);
  for (auto it = range.first; it != range.second; ++it) {
    if (hipEventQuery(it->second.ev) == hipSuccess) {
      *out = it->second.p;
      (void)hipEventDestroy(it->second.ev);
      g_dev.cached.erase(it);
      g_dev.live[*out] = c;
      return HGX_OK;
    }
    (void)hipGetLastError();
  }
  int rc = dev_alloc(out, c);
  if (rc == HGX_ENOMEM && !g_dev.cached.empty()) {
    dev_trim(false);
    rc = dev_alloc(out, c);
    if (rc == HGX_ENOMEM) {
      dev_trim(true);
      rc = dev_alloc(out, c);
    }
  }
  if (rc) return rc;
  g_dev.live[*out] = c;
  return HGX_OK;
}

int dev_block_release(void *p, hipStream_t s) {
  auto it = g_dev.live.find(p);
  if (it == g_dev.live.end()) return fail(HGX_EINVAL, "not a live block of hgx_device_alloc");
  hipEvent_t ev = nullptr;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const hipError_t e = hipEventRecord(ev, s);
  if (e != hipSuccess) {
    (void)hipEventDestroy(ev);
    return fail(HGX_EDEVICE, std::string("hipEventRecord failed: ") + hipGetErrorString(e));
  }
  g_dev.cached.emplace(it->second, DevBlock{p, it->second, ev});
  g_dev.live.erase(it);
  return HGX_OK;
}

// every block, cached or handed out (hgx_shutdown)
void dev_pool_free() {
  dev_trim(true);
  for (auto &kv : g_dev.live) dev_free(kv.first, kv.second);
  g_dev.live.clear();
  if (g_dev.res) dev_free(g_dev.res, 256);
  g_dev.res = nullptr;
  if (g_dev.hres) (void)hipHostFree(g_dev.hres);
  g_dev.hres = nullptr;
}

// the result words of the scans, on first use
int dev_res_ready() {
  if (g_dev.res) return HGX_OK;
  int rc = dev_alloc((void **)&g_dev.res, 256);
  if (rc) return rc;
  const hipError_t e = hipHostMalloc((void **)&g_dev.hres, 256, hipHostMallocDefault);
  if (e != hipSuccess) {
    dev_free(g_dev.res, 256);
    g_dev.res = nullptr;
    return fail(HGX_ENOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
  }
  return HGX_OK;
}

// run the scan kernels on S and read their two result words back
int scan_read(hipStream_t s, const u64 *a, u64 na, const u64 *b, u64 nb, bool sign, u64 out[2]) {
  int rc = dev_res_ready();
  if (rc) return rc;
  hgx_launch_top_diff(s, a, na, b, nb, g_dev.res);
  if (sign) hgx_launch_cmp_sign(s, a, na, b, nb, g_dev.res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g_dev.hres, g_dev.res, 16, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  out[0] = g_dev.hres[0];
  out[1] = g_dev.hres[1];
  return HGX_OK;
}

// ---- concurrent host-memory products (call slots) ----
// hgx_mul / hgx_sqr calls from several threads run on call slots: a slot
// has a stream, device memory for its product's operands, product and work
// arrays, pinned staging and a pinned chunk table of its own.  g_ctx.mu is
// held only while a call sizes its slot and builds root tables, and while it
// queues its kernels; its transfers and the wait for its product run
// without it, so calls on different slots overlap, one's copies with
// another's kernels.  Products whose device working set exceeds 1/8 of the
// budget (or that must be host-staged) take the library for their whole
// duration, as a single product of that size fills the GPU by itself.
constexpr int kCallSlots = 4;
struct CallSlot {
  hipStream_t s = nullptr;
  Workspace mem, scan;  // operands, product and work memory; a split product's carry scratch
  PinStage pin;
  MulJob *jobs = nullptr;  // pinned chunk table of a split product
  size_t jobs_cap = 0;
  bool busy = false;
};
struct CallSlots {
  std::mutex mu;
  std::condition_variable cv;
  CallSlot slot[kCallSlots];
  int active = 0;        // calls holding a slot
  bool closing = false;  // hgx_shutdown waits for the active calls
};
CallSlots g_slots;

// a free slot (waiting for one), or -1 while the library shuts down
int slot_acquire() {
  std::unique_lock<std::mutex> lk(g_slots.mu);
  for (;;) {
    if (g_slots.closing) return -1;
    for (int i = 0; i < kCallSlots; i++)
      if (!g_slots.slot[i].busy) {
        g_slots.slot[i].busy = true;
        g_slots.active++;
        return i;
      }
    g_slots.cv.wait(lk);
  }
}

void slot_release(int i) {
  std::lock_guard<std::mutex> lk(g_slots.mu);
  g_slots.slot[i].busy = false;
  g_slots.active--;
  g_slots.cv.notify_all();
}

// device memory of a slot (g_ctx.mu held; the slot's stream is idle)
int slot_ensure(Workspace &w, size_t bytes) {
  if (w.bytes >= bytes) return HGX_OK;
  dev_free(w.ptr, w.bytes);
  w = Workspace{};
  const size_t want = std::max(bytes, (size_t)1 << 20);
  int rc = dev_alloc(&w.ptr, want);
  if (rc) return rc;
  w.bytes = want;
  return HGX_OK;
}

// release the device memory of every slot not in use (g_ctx.mu held), so
// that a larger product or the server may have it; whether any was held
bool slots_trim() {
  std::lock_guard<std::mutex> lk(g_slots.mu);
  bool any = false;
  for (CallSlot &S : g_slots.slot) {
    if (S.busy) continue;
    any = any || S.mem.ptr || S.scan.ptr;
    dev_free(S.mem.ptr, S.mem.bytes);
    dev_free(S.scan.ptr, S.scan.bytes);
    S.mem = Workspace{};
    S.scan = Workspace{};
  }
  return any;
}

// every slot's resources (g_ctx.mu held, no slot active)
void slots_free() {
  for (CallSlot &S : g_slots.slot) {
    if (S.s) (void)hipStreamSynchronize(S.s);
    dev_free(S.mem.ptr, S.mem.bytes);
    dev_free(S.scan.ptr, S.scan.bytes);
    pin_free(S.pin);
    if (S.jobs) (void)hipHostFree(S.jobs);
    if (S.s) (void)hipStreamDestroy(S.s);
    S = CallSlot{};
  }
}

// whether a host-memory product runs on a call slot (g_ctx.mu held);
// HGX_CALL_SLOTS=0 in the environment turns the slots off
bool slot_product(size_t na, size_t nb, bool sqr) {
  static const bool on = [] {
    const char *e = getenv("HGX_CALL_SLOTS");
    return !(e && *e == '0');
  }();
  return on && !needs_staging(na, nb, sqr) && inhbm_bytes(na, nb, sqr) <= g_ctx.budget / 8;
}

// The product of host-memory operands on a call slot (g_ctx.mu not held).
int mul_host_slot(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, bool sqr) {
  const int si = slot_acquire();
  if (si < 0) return fail(HGX_ENOTINIT, "the library is shutting down");
  CallSlot &S = g_slots.slot[si];
  struct Release {
    CallSlot &S;
    int i;
    ~Release() {
      // nothing may still run on the slot's memory once it is free (a
      // failed call can return with work queued)
      if (S.s && hipStreamSynchronize(S.s) != hipSuccess) (void)hipGetLastError();
      slot_release(i);
    }
  } release{S, si};
  const size_t nout = na + nb;
  const size_t ba = align_up(na * 8), bb = sqr ? 0 : align_up(nb * 8), bo = align_up(nout * 8);
  const int n = log2_ceil((u64)hgx_product_length(na, nb));
  Split sp;
  bool split = false;
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    int rc = check_ready();
    if (rc) return rc;
    if (!S.s) HIPCHK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
    split = !sqr && g_ctx.split_on && choose_split(na, nb, sp);
    size_t work = 0;
    if (split) {
      const SplitBytes sb = split_bytes(na, nb, sp);
      work = sb.main;
      rc = slot_ensure(S.scan, sb.scan);
      if (!rc && S.jobs_cap < sp.k) {
        if (S.jobs) (void)hipHostFree(S.jobs);
        S.jobs = nullptr;
        S.jobs_cap = 0;
        HIPCHK(hipHostMalloc((void **)&S.jobs, std::max<size_t>(sp.k, 64) * sizeof(MulJob), hipHostMallocDefault));
        S.jobs_cap = std::max<size_t>(sp.k, 64);
      }
    } else {
      work = work_layout(n, sqr).total;
    }
    if (!rc) rc = slot_ensure(S.mem, ba + bb + bo + work);
    Plan *plan = nullptr;
    if (!rc) rc = build_plan(split ? sp.nc : n, &plan);  // root tables, built and copied before this returns
    if (rc) return rc;
  }
  char *base = (char *)S.mem.ptr;
  u64 *d_a = (u64 *)base, *d_b = sqr ? d_a : (u64 *)(base + ba), *d_o = (u64 *)(base + ba + bb);
  char *work = base + ba + bb + bo;
  PinScope pin(S.pin);
  int rc = h2d(S.pin, d_a, a, na * 8, S.s);
  if (!rc && !sqr) rc = h2d(S.pin, d_b, b, nb * 8, S.s);
  if (rc) return rc;
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    if (split) {
      rc = split_product_at(d_a, na, d_b, nb, d_o, sp, S.s, work, (uint8_t *)S.scan.ptr, S.jobs);
    } else {
      JobRef jr{};
      jr.one = make_job(d_a, na, d_b, sqr ? na : nb, d_o, n, sqr, work);
      rc = enqueue_group(jr, 1, n, sqr, S.s);
    }
    if (rc) return rc;
  }
  rc = d2h(S.pin, out, d_o, nout * 8, S.s);
  if (!rc) rc = pin_flush(S.pin);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S.s));
  return HGX_OK;
}

}  // namespace

// for the library's other host translation units (hgx_series.cpp): record
// an error for hgx_last_error on the calling thread, and whether hgx_init
// has succeeded
int hgx_internal_fail(int code, const char *msg) { return fail(code, msg ? msg : ""); }
bool hgx_internal_ready() { return g_ctx.ready; }

size_t hgx_internal_device_room() {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  return g_ctx.ready ? memory_for_product(true) : 0;
}

int hgx_internal_upload_gather(uint64_t *d_dst, const uint64_t *const *src, const size_t *nlimbs, size_t count) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (count && (!d_dst || !src || !nlimbs)) return fail(HGX_EINVAL, "null gather upload");
  const hipStream_t s = g_ctx.stream;
  PinStage &P = g_pin;
  rc = pin_init(P);
  if (rc) return rc;
  PinScope pin;
  size_t k = 0, in = 0;  // the next piece and its limbs already staged
  u64 *dst = d_dst;
  while (k < count) {
    const int i = P.next;
    P.next = (i + 1) % kPinBufs;
    rc = pin_settle(P, i);
    if (rc) return rc;
    // fill staging buffer i from the pieces, on the copy threads
    std::vector<Piece> pieces;
    size_t fill = 0;
    while (k < count && fill < kPinChunk) {
      const size_t take = std::min((nlimbs[k] - in) * 8, kPinChunk - fill);
      if (take) pieces.push_back(Piece{P.buf[i] + fill, src[k] + in, take});
      fill += take;
      in += take / 8;
      if (in == nlimbs[k]) {
        k++;
        in = 0;
      }
    }
    if (fill == 0) break;
    copy_pieces(pieces);
    HIPCHK(hipMemcpyAsync(dst, P.buf[i], fill, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(P.ev[i], s));
    P.state[i] = 1;
    dst += fill / 8;
  }
  HIPCHK(hipStreamSynchronize(s));
  return HGX_OK;
}

int hgx_internal_sums_device(HgxSumJob *jobs, size_t count, int *signs, uint64_t *sizes) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (count == 0) return HGX_OK;
  if (!jobs || !signs || !sizes) return fail(HGX_EINVAL, "null sum jobs or results");
  u64 nblk = 0;
  for (size_t j = 0; j < count; j++) {
    HgxSumJob &J = jobs[j];
    if (!J.x || !J.y || J.ny == 0 || J.sy == 0) return fail(HGX_EINVAL, "a signed sum needs a non-zero Y");
    if (J.sx == 0) J.nx = 0;
    J.n = std::max(J.nx, J.ny) + 1;
    J.blk0 = nblk;
    nblk += hgx_addsub_blocks(J.n);
  }
  if (nblk >= ((u64)1 << 31) || count >= ((size_t)1 << 31)) return fail(HGX_ETOOBIG, "too many signed sums");
  hipStream_t s = g_ctx.stream;
  rc = ws_acquire(s);
  if (rc) return rc;
  const size_t jb = align_up(count * sizeof(HgxSumJob)), rb = align_up(count * 4 * sizeof(u64));
  rc = ensure(g_ctx.scan, jb + rb + 2 * nblk);
  if (rc) return rc;
  char *base = (char *)g_ctx.scan.ptr;
  HgxSumJob *djobs = (HgxSumJob *)base;
  u64 *dres = (u64 *)(base + jb);
  uint8_t *agg = (uint8_t *)(base + jb + rb);
  std::vector<u64> hres(4 * count);
  {
    PinScope pin;
    rc = h2d(djobs, jobs, count * sizeof(HgxSumJob), s);
    if (rc) return rc;
    hgx_launch_sums(s, djobs, count, nblk, agg, agg + nblk, dres);
    HIPCHK(hipGetLastError());
    rc = d2h(hres.data(), dres, hres.size() * sizeof(u64), s);
    if (!rc) rc = pin_flush();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
  }
  rc = ws_release(s);
  if (rc) return rc;
  for (size_t j = 0; j < count; j++) {
    sizes[j] = hres[4 * j + 1];
    signs[j] = hres[4 * j + 2] == 1 ? 1 : (hres[4 * j + 2] == 2 ? -1 : 0);
  }
  return HGX_OK;
}

// The leaf folds on the device, in a scratch block SCR laid out as: the
// fold buffers (stride ST, CAP[b] limbs each), the leaves' bounds, their
// codes, their offsets.  The packed values go to a new block *OUT.
static int fold_device_in(void *scr, const FoldSeries &S, const uint64_t *bounds, size_t count, const uint32_t *cap,
                          size_t st, size_t per, int32_t *codes, uint64_t **out, size_t *nlimbs) {
  const hipStream_t s = g_ctx.stream;
  u64 *bufs = (u64 *)scr;
  u64 *dbounds = bufs + st * per;
  int *dcodes = (int *)(dbounds + 2 * count);
  u64 *doff = (u64 *)(dcodes + 4 * count);
  FoldSlots F;
  size_t at = 0;
  for (int b = 0; b < kFoldBufs; b++) {
    F.buf[b] = bufs + at * st;
    F.cap[b] = cap[b];
    at += cap[b];
  }
  F.stride = st;
  {
    PinScope pin;
    int rc = h2d(dbounds, bounds, 16 * count, s);
    if (rc) return rc;
    hgx_launch_fold(s, S, dbounds, count, F, dcodes);
    HIPCHK(hipGetLastError());
    rc = d2h(codes, dcodes, 16 * count, s);
    if (!rc) rc = pin_flush();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
  }
  std::vector<u64> off(count);
  size_t total = 0;
  for (size_t t = 0; t < count; t++) {
    if (codes[4 * t] == kFoldFailed) return fail(HGX_ETOOBIG, "a leaf outgrew the device fold's buffers");
    off[t] = total;
    for (int w = 0; w < 4; w++) total += (size_t)std::abs((long long)codes[4 * t + w]);
  }
  void *o = nullptr;
  int rc = dev_block_alloc(std::max<size_t>(total, 1) * 8, &o);
  if (rc) return rc;
  {
    PinScope pin;
    rc = h2d(doff, off.data(), 8 * count, s);
    if (!rc) {
      hgx_launch_fold_pack(s, F, count, dcodes, doff, (u64 *)o);
      const hipError_t e = hipGetLastError();
      if (e == hipSuccess) {
        const hipError_t e2 = hipStreamSynchronize(s);
        if (e2 != hipSuccess)
          rc = fail(HGX_EDEVICE, std::string("leaf pack failed: ") + hipGetErrorString(e2));
      } else {
        rc = fail(HGX_EDEVICE, std::string("leaf pack launch failed: ") + hipGetErrorString(e));
      }
    }
  }
  if (rc) {
    (void)dev_block_release(o, s);
    return rc;
  }
  *out = (uint64_t *)o;
  *nlimbs = total;
  return HGX_OK;
}

int hgx_internal_copy_device(uint64_t *d_dst, const uint64_t *d_src, size_t n) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!n) return HGX_OK;
  if (!d_dst || !d_src) return fail(HGX_EINVAL, "null device copy");
  HIPCHK(hipMemcpyAsync(d_dst, d_src, 8 * n, hipMemcpyDeviceToDevice, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return HGX_OK;
}

int hgx_internal_fold_device(const FoldSeries *S, const uint64_t *bounds, size_t count, const uint32_t *cap,
                             int32_t *codes, uint64_t **d_vals, size_t *nlimbs) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!S || !bounds || !cap || !codes || !d_vals || !nlimbs || count == 0)
    return fail(HGX_EINVAL, "bad leaf fold arguments");
  *d_vals = nullptr;
  *nlimbs = 0;
  if (count >= ((size_t)1 << 28)) return fail(HGX_ETOOBIG, "too many leaves for the device fold");
  size_t per = 0;  // limbs of one leaf's buffers
  for (int b = 0; b < kFoldBufs; b++) {
    if (cap[b] < (u32)kFoldTV2 + 1 || cap[b] >= (1u << 30)) return fail(HGX_EINVAL, "bad leaf fold capacity");
    per += cap[b];
  }
  const size_t st = (count + 63) & ~(size_t)63;  // the stride, whole waves of leaves
  if (per > ((size_t)1 << 40) / st) return fail(HGX_ETOOBIG, "leaf fold buffers above 8 TiB");
  const size_t bytes = 8 * st * per + 16 * count + 16 * count + 8 * count;
  void *scr = nullptr;
  rc = dev_block_alloc(bytes, &scr);
  if (rc) return rc;
  rc = fold_device_in(scr, *S, bounds, count, cap, st, per, codes, d_vals, nlimbs);
  const int rc2 = dev_block_release(scr, g_ctx.stream);
  if (rc) return rc;
  if (rc2) {
    (void)dev_block_release(*d_vals, g_ctx.stream);
    *d_vals = nullptr;
    return rc2;
  }
  return HGX_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *hgx_last_error(void) { return g_err.c_str(); }

int hgx_init(int device, size_t hbm_budget_bytes) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  if (g_ctx.ready) return HGX_OK;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(HGX_EINVAL, "no such device");
  HIPCHK(hipSetDevice(device));
  g_ctx.device = device;
  if (device >= kMaxDevices) return fail(HGX_EINVAL, "device index above 63");
  if (!g_scratch[device]) {
    void *sp = nullptr;
    const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp)
      return fail(HGX_EDEVICE, std::string("looking up the no-allocation scratch array failed: ") +
                                   hipGetErrorString(e));
    g_scratch[device] = sp;
  }
  size_t freeb = 0, totalb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totalb));
  g_ctx.budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(0.9 * (double)freeb);
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&g_ctx.ws_done, hipEventDisableTiming));
  HIPCHK(hipEventRecord(g_ctx.ws_done, g_ctx.stream));
  HIPCHK(hipEventCreateWithFlags(&g_ctx.jobs_copied, hipEventDisableTiming));
  HIPCHK(hipEventRecord(g_ctx.jobs_copied, g_ctx.stream));
  // prime set: moduli, Proth factors, generators, in-tile root tables
  std::vector<u64> host(3 * 2 * 2048 * 2);
  PrimeSet ps{};
  for (int k = 0; k < 3; k++) {
    const u64 p = kModuli[k];
    ps.p[k] = p;
    ps.np[k] = 0 - p;
    ps.c[k] = (u32)((p - 1) >> 55);
    g_ctx.gen[k] = find_generator(p, ps.c[k]);
    const u64 w = powmod(g_ctx.gen[k], (p - 1) >> 12, p);
    const u64 wi = invmod(w, p);
    u64 x = 1, y = 1;
    for (int i = 0; i < 2048; i++) {
      host[((k * 2 + 0) * 2048 + i) * 2] = x;
      host[((k * 2 + 0) * 2048 + i) * 2 + 1] = shoup_companion(x, p);
      host[((k * 2 + 1) * 2048 + i) * 2] = y;
      host[((k * 2 + 1) * 2048 + i) * 2 + 1] = shoup_companion(y, p);
      x = mulmod(x, w, p);
      y = mulmod(y, wi, p);
    }
  }
  ps.g01 = invmod(kModuli[0] % kModuli[1], kModuli[1]);
  ps.g02 = invmod(kModuli[0] % kModuli[2], kModuli[2]);
  ps.g12 = invmod(kModuli[1] % kModuli[2], kModuli[2]);
  ps.g01q = shoup_companion(ps.g01, kModuli[1]);
  ps.g02q = shoup_companion(ps.g02, kModuli[2]);
  ps.g12q = shoup_companion(ps.g12, kModuli[2]);
  int rc = dev_alloc((void **)&g_ctx.tables, host.size() * sizeof(u64));
  if (rc) return rc;
  HIPCHK(hipMemcpy(g_ctx.tables, host.data(), host.size() * sizeof(u64), hipMemcpyHostToDevice));
  for (int k = 0; k < 3; k++) {
    ps.tw_fwd[k] = g_ctx.tables + (k * 2 + 0) * 2048 * 2;
    ps.tw_inv[k] = g_ctx.tables + (k * 2 + 1) * 2048 * 2;
  }
  g_ctx.ps = ps;
  g_ctx.ready = true;
  tune_cache_load();
  return HGX_OK;
}

void hgx_shutdown(void) {
  server_stop();
  {
    // calls on call slots finish first (they take g_ctx.mu to queue work)
    std::unique_lock<std::mutex> sl(g_slots.mu);
    g_slots.closing = true;
    g_slots.cv.notify_all();
    g_slots.cv.wait(sl, [] { return g_slots.active == 0; });
  }
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  slots_free();
  {
    std::lock_guard<std::mutex> sl(g_slots.mu);
    g_slots.closing = false;
  }
  if (!g_ctx.ready) return;
  (void)hipStreamSynchronize(g_ctx.stream);
  timing_drain();
  g_timing = false;
  g_kstats.clear();
  for (hipEvent_t e : g_timing_events) (void)hipEventDestroy(e);
  g_timing_events.clear();
  dev_pool_free();
  for (auto &kv : g_ctx.plans) free_plan(*kv.second);
  g_ctx.plans.clear();
  for (auto &kv : g_ctx.graphs) (void)hipGraphExecDestroy(kv.second);
  g_ctx.graphs.clear();
  dev_free(g_ctx.ws.ptr, g_ctx.ws.bytes);
  g_ctx.ws = Workspace{};
  dev_free(g_ctx.stage.ptr, g_ctx.stage.bytes);
  g_ctx.stage = Workspace{};
  dev_free(g_ctx.jobs_dev.ptr, g_ctx.jobs_dev.bytes);
  g_ctx.jobs_dev = Workspace{};
  dev_free(g_ctx.pack.ptr, g_ctx.pack.bytes);
  g_ctx.pack = Workspace{};
  dev_free(g_ctx.scan.ptr, g_ctx.scan.bytes);
  g_ctx.scan = Workspace{};
  if (g_ctx.jobs_host) (void)hipHostFree(g_ctx.jobs_host);
  g_ctx.jobs_host = nullptr;
  g_ctx.jobs_host_cap = 0;
  pin_free(g_pin);
  host_buf_free(g_pack_in);
  host_buf_free(g_pack_out);
  (void)hipEventDestroy(g_ctx.jobs_copied);
  dev_free(g_ctx.tables, 3 * 2 * 2048 * 2 * sizeof(u64));
  g_ctx.tables = nullptr;
  (void)hipEventDestroy(g_ctx.ws_done);
  (void)hipStreamDestroy(g_ctx.stream);
  g_ctx.ready = false;
}

int hgx_num_moduli(void) { return 3; }
uint64_t hgx_modulus(int k) { return (k >= 0 && k < 3) ? kModuli[k] : 0; }
int hgx_max_transform_bits(void) { return kMaxBits; }

size_t hgx_product_length(size_t na, size_t nb) {
  // least-power-of-two->=, src/math-utilities.lisp:24-28
  size_t n = na + nb;
  size_t L = 1;
  while (L < n) L <<= 1;
  return L;
}

int hgx_mul_device(const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_a || !d_b || !d_out || na == 0 || nb == 0) return fail(HGX_EINVAL, "null operand or zero length");
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  const bool sqr = (d_a == d_b && na == nb);
  return mul_device_locked(d_a, na, d_b, nb, d_out, sqr, s);
}

int hgx_sqr_device(const uint64_t *d_a, size_t na, uint64_t *d_out, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_a || !d_out || na == 0) return fail(HGX_EINVAL, "null operand or zero length");
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  return mul_device_locked(d_a, na, d_a, na, d_out, true, s);
}

int hgx_mul_device_noalloc(const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out,
                           void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_a || !d_b || !d_out || na == 0 || nb == 0) return fail(HGX_EINVAL, "null operand or zero length");
  const int n = log2_ceil((u64)hgx_product_length(na, nb));
  if (n > kNoAllocMaxBits)
    return fail(HGX_ETOOBIG, "hgx_mul_device_noalloc takes products of transform length at most 2^" +
                                 std::to_string(kNoAllocMaxBits) + " (na + nb <= " +
                                 std::to_string(1ull << kNoAllocMaxBits) + "); this one needs 2^" + std::to_string(n));
  if (!g_ctx.plans.count(plan_key(n, choice_for(n))))
    return fail(HGX_ENOTINIT, "the root tables of transform length 2^" + std::to_string(n) +
                                  " are not resident: run one hgx_mul_device product of that length first");
  void *sp = g_scratch[g_ctx.device];
  if (!sp) return fail(HGX_ENOTINIT, "no scratch array on this device");
  const bool sqr = (d_a == d_b && na == nb);
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  JobRef jr{};
  jr.one = make_job(d_a, na, d_b, nb, d_out, n, sqr, (char *)sp);
  // no per-kernel timing events here: creating one may allocate
  const bool timing = g_timing;
  g_timing = false;
  rc = enqueue_group(jr, 1, n, sqr, s);
  g_timing = timing;
  return rc;
}

// host-memory product: stage operands into device memory, multiply, copy back
static int mul_host_locked(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out, bool sqr) {
  if (log2_ceil((u64)hgx_product_length(na, nb)) > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32");
  if (needs_staging(na, nb, sqr)) return mul_staged_locked(a, na, b, nb, out, sqr);
  hipStream_t s = g_ctx.stream;
  const size_t nout = na + nb;
  const size_t bytes_a = align_up(na * 8), bytes_b = sqr ? 0 : align_up(nb * 8), bytes_o = align_up(nout * 8);
  int rc = ws_acquire(s);
  if (rc) return rc;
  rc = ensure(g_ctx.stage, bytes_a + bytes_b + bytes_o);
  if (rc) return rc;
  char *base = (char *)g_ctx.stage.ptr;
  u64 *d_a = (u64 *)base, *d_b = sqr ? d_a : (u64 *)(base + bytes_a), *d_o = (u64 *)(base + bytes_a + bytes_b);
  PinScope pin;
  rc = h2d(d_a, a, na * 8, s);
  if (!rc && !sqr) rc = h2d(d_b, b, nb * 8, s);
  if (!rc) rc = mul_device_locked(d_a, na, d_b, sqr ? na : nb, d_o, sqr, s);
  if (!rc) rc = d2h(out, d_o, nout * 8, s);
  if (!rc) rc = pin_flush();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return HGX_OK;
}

// limb ranges [x, x + nx) and [y, y + ny) share memory
static bool overlaps(const uint64_t *x, size_t nx, const uint64_t *y, size_t ny) {
  const uintptr_t x0 = (uintptr_t)x, x1 = x0 + 8 * nx, y0 = (uintptr_t)y, y1 = y0 + 8 * ny;
  return x0 < y1 && y0 < x1;
}

// A host-memory product on a call slot, or holding the library (large or
// host-staged products).
static int mul_host(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out, bool sqr) {
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    int rc = check_ready();
    if (rc) return rc;
    if (log2_ceil((u64)hgx_product_length(na, nb)) > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32");
    if (!slot_product(na, nb, sqr)) {
      slots_trim();  // the product may use what idle slots hold
      return mul_host_locked(a, na, b, nb, out, sqr);
    }
  }
  return mul_host_slot(a, na, b, nb, out, sqr);
}

int hgx_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
  if (!a || !b || !out || na == 0 || nb == 0) return fail(HGX_EINVAL, "null operand or zero length");
  if (overlaps(out, na + nb, a, na) || overlaps(out, na + nb, b, nb))
    return fail(HGX_EINVAL, "the product storage overlaps an operand");
  return mul_host(a, na, b, nb, out, a == b && na == nb);
}

int hgx_sqr(const uint64_t *a, size_t na, uint64_t *out) {
  if (!a || !out || na == 0) return fail(HGX_EINVAL, "null operand or zero length");
  if (overlaps(out, 2 * na, a, na)) return fail(HGX_EINVAL, "the square's storage overlaps the operand");
  return mul_host(a, na, a, na, out, true);
}

int hgx_mul_batch(const hgx_mul_desc *descs, size_t count) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (count == 0) return HGX_OK;
  if (!descs) return fail(HGX_EINVAL, "null descriptor array");
  for (size_t i = 0; i < count; i++) {
    const hgx_mul_desc &d = descs[i];
    if (!d.a || !d.b || !d.out || d.na == 0 || d.nb == 0) return fail(HGX_EINVAL, "bad descriptor " + std::to_string(i));
    if (log2_ceil((u64)hgx_product_length(d.na, d.nb)) > kMaxBits)
      return fail(HGX_ETOOBIG, "transform length above 2^32 in descriptor " + std::to_string(i));
  }
  // products too large for the device memory are staged one by one; the
  // others are packed into device buffers, each pack small enough that its
  // largest product's work memory still fits beside it, and each pack runs as
  // a device batch (which chunks the work memory by the budget itself)
  std::vector<size_t> rest;
  for (size_t i = 0; i < count; i++) {
    const hgx_mul_desc &d = descs[i];
    const bool sqr = (d.a == d.b && d.na == d.nb);
    if (needs_staging(d.na, d.nb, sqr)) {
      rc = mul_staged_locked(d.a, d.na, d.b, d.nb, d.out, sqr);
      if (rc) return rc;
      continue;
    }
    rest.push_back(i);
  }
  if (rest.empty()) return HGX_OK;
  hipStream_t s = g_ctx.stream;
  rc = ws_acquire(s);
  if (rc) return rc;
  for (size_t k0 = 0; k0 < rest.size();) {
    const size_t reusable = g_ctx.ws.bytes + g_ctx.pack.bytes + g_ctx.jobs_dev.bytes + g_ctx.scan.bytes;
    const size_t held = g_ctx.allocated - reusable;
    const size_t avail = g_ctx.budget > held ? g_ctx.budget - held : 0;
    size_t k1 = k0, in_words = 0, out_words = 0, max_work = 0;
    while (k1 < rest.size()) {
      const hgx_mul_desc &d = descs[rest[k1]];
      const bool sqr = (d.a == d.b && d.na == d.nb);
      const size_t iw = d.na + (sqr ? 0 : d.nb), ow = d.na + d.nb;
      const size_t w = std::max(max_work, work_bytes(d.na, d.nb, sqr));
      const size_t need = align_up((in_words + iw + out_words + ow) * 8) + w +
                          align_up((k1 - k0 + 1) * sizeof(MulJob)) + kStageSlack;
      if (k1 > k0 && need > avail) break;
      // a pack that fits the pinned buffers goes no further
      if (k1 > k0 && 8 * std::max(in_words + iw, out_words + ow) > kPackHostMax &&
          8 * std::max(in_words, out_words) <= kPackHostMax)
        break;
      in_words += iw;
      out_words += ow;
      max_work = w;
      k1++;
    }
    rc = ensure(g_ctx.pack, align_up((in_words + out_words) * 8));
    if (rc) return rc;
    // the packs: pinned when they fit, pageable (staged) otherwise
    std::unique_ptr<u64[]> own_in, own_out;
    u64 *hin = host_buf(g_pack_in, in_words * 8);
    u64 *hout = hin ? host_buf(g_pack_out, out_words * 8) : nullptr;
    if (!hin || !hout) {
      own_in.reset(new (std::nothrow) u64[std::max<size_t>(in_words, 1)]);
      own_out.reset(new (std::nothrow) u64[std::max<size_t>(out_words, 1)]);
      if (!own_in || !own_out) return fail(HGX_ENOMEM, "out of host memory for a batch's operand pack");
      hin = own_in.get();
      hout = own_out.get();
    }
    std::vector<hgx_mul_desc> dd(k1 - k0);
    std::vector<Piece> pieces;
    pieces.reserve(2 * (k1 - k0));
    u64 *din = (u64 *)g_ctx.pack.ptr, *dout = din + in_words;
    size_t at = 0, ot = 0;
    for (size_t k = k0; k < k1; k++) {
      const hgx_mul_desc &d = descs[rest[k]];
      const bool sqr = (d.a == d.b && d.na == d.nb);
      hgx_mul_desc &e = dd[k - k0];
      pieces.push_back(Piece{hin + at, d.a, d.na * 8});
      e.a = din + at;
      e.na = d.na;
      at += d.na;
      if (sqr) {
        e.b = e.a;
      } else {
        pieces.push_back(Piece{hin + at, d.b, d.nb * 8});
        e.b = din + at;
        at += d.nb;
      }
      e.nb = d.nb;
      e.out = dout + ot;
      ot += d.na + d.nb;
    }
    copy_pieces(pieces);
    PinScope pin;
    rc = h2d(din, hin, in_words * 8, s);
    if (rc) return rc;
    rc = batch_device_locked(dd.data(), dd.size(), s);
    if (rc) return rc;
    rc = d2h(hout, dout, out_words * 8, s);
    if (!rc) rc = pin_flush();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    pieces.clear();
    ot = 0;
    for (size_t k = k0; k < k1; k++) {
      const hgx_mul_desc &d = descs[rest[k]];
      pieces.push_back(Piece{d.out, hout + ot, (d.na + d.nb) * 8});
      ot += d.na + d.nb;
    }
    copy_pieces(pieces);
    k0 = k1;
  }
  return HGX_OK;
}

int hgx_mul_batch_device(const hgx_mul_desc *descs, size_t count, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (count == 0) return HGX_OK;
  if (!descs) return fail(HGX_EINVAL, "null descriptor array");
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  return batch_device_locked(descs, count, s);
}

int hgx_ntt_forward(uint64_t *data, size_t n, int mod_num) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!data || n == 0 || mod_num < 0 || mod_num > 2) return fail(HGX_EINVAL, "bad transform arguments");
  hipStream_t s = g_ctx.stream;
  rc = ws_acquire(s);
  if (rc) return rc;
  rc = ensure(g_ctx.stage, align_up(n * 8));
  if (rc) return rc;
  u64 *d = (u64 *)g_ctx.stage.ptr;
  PinScope pin;
  rc = h2d(d, data, n * 8, s);
  if (!rc) rc = enqueue_transform(d, n, mod_num, false, s);
  if (!rc) rc = d2h(data, d, n * 8, s);
  if (!rc) rc = pin_flush();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return ws_release(s);
}

int hgx_ntt_reverse(uint64_t *data, size_t n, int mod_num) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!data || n == 0 || mod_num < 0 || mod_num > 2) return fail(HGX_EINVAL, "bad transform arguments");
  hipStream_t s = g_ctx.stream;
  rc = ws_acquire(s);
  if (rc) return rc;
  rc = ensure(g_ctx.stage, align_up(n * 8));
  if (rc) return rc;
  u64 *d = (u64 *)g_ctx.stage.ptr;
  PinScope pin;
  rc = h2d(d, data, n * 8, s);
  if (!rc) rc = enqueue_transform(d, n, mod_num, true, s);
  if (!rc) rc = d2h(data, d, n * 8, s);
  if (!rc) rc = pin_flush();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return ws_release(s);
}

// magnitude sum / difference of device limb vectors (mpz-+ / mpz-- on the
// magnitudes, src/mpz-ram.lisp:185-291)
static int addsub_device(bool sub, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out,
                         void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_out || (na && !d_a) || (nb && !d_b)) return fail(HGX_EINVAL, "null operand or output");
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  const u64 n = std::max(na, nb);
  if (n == 0) return fail(HGX_EINVAL, "both operands empty");
  const u64 nblk = hgx_addsub_blocks(n);
  rc = ws_acquire(s);
  if (rc) return rc;
  rc = ensure(g_ctx.scan, 2 * nblk);
  if (rc) return rc;
  uint8_t *agg = (uint8_t *)g_ctx.scan.ptr;
  hgx_launch_addsub(sub, s, d_a, na, d_b, nb, d_out, n, agg, agg + nblk, sub ? nullptr : d_out + n);
  HIPCHK(hipGetLastError());
  return ws_release(s);
}

int hgx_add_device(const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out, void *stream) {
  return addsub_device(false, d_a, na, d_b, nb, d_out, stream);
}

int hgx_sub_device(const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out, void *stream) {
  return addsub_device(true, d_a, na, d_b, nb, d_out, stream);
}

int hgx_fill_splitmix64(uint64_t *d_out, size_t n, uint64_t seed, int force_top, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_out) return fail(HGX_EINVAL, "null output");
  if (n == 0) return HGX_OK;
  hipStream_t s = stream ? (hipStream_t)stream : g_ctx.stream;
  hgx_launch_splitmix64(dim3((unsigned)((n + 255) / 256)), s, d_out, (u64)n, seed, force_top);
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

int hgx_submit(const hgx_mul_desc *desc, hgx_ticket *ticket) {
  if (!desc || !ticket) return fail(HGX_EINVAL, "null descriptor or ticket");
  const hgx_mul_desc &d = *desc;
  if (!d.a || !d.b || !d.out || d.na == 0 || d.nb == 0) return fail(HGX_EINVAL, "null operand or zero length");
  auto r = std::make_shared<SrvReq>();
  r->d = d;
  r->sqr = (d.a == d.b && d.na == d.nb);
  r->n = log2_ceil((u64)hgx_product_length(d.na, d.nb));
  if (r->n > kMaxBits) return fail(HGX_ETOOBIG, "transform length above 2^32");
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    int rc = check_ready();
    if (rc) return rc;
    rc = server_start_locked();
    if (rc) return rc;
    r->split = !r->sqr && g_ctx.split_on && choose_split(d.na, d.nb, r->sp);
  }
  r->bytes = request_bytes(*r);
  static std::atomic<hgx_ticket> next_ticket{1};
  const hgx_ticket t = next_ticket++;
  {
    std::lock_guard<std::mutex> lk(g_srv.mu);
    if (!g_srv.started || g_srv.stop) return fail(HGX_ENOTINIT, "the request server is stopping (hgx_shutdown)");
    g_srv.tickets[t] = r;
    g_srv.queue.push_back(r);
  }
  g_srv.cv.notify_all();
  *ticket = t;
  return HGX_OK;
}

// 1 (done) or the request's error releases the ticket
int hgx_poll(hgx_ticket ticket) {
  std::lock_guard<std::mutex> lk(g_srv.mu);
  auto it = g_srv.tickets.find(ticket);
  if (it == g_srv.tickets.end()) return fail(HGX_EINVAL, "unknown ticket");
  const int st = it->second->state;
  if (st == 2) {
    g_srv.tickets.erase(it);
    return 1;
  }
  if (st < 0) {
    const std::string err = it->second->err;
    g_srv.tickets.erase(it);
    return fail(st, err);
  }
  return 0;
}

int hgx_wait(hgx_ticket ticket) {
  std::unique_lock<std::mutex> lk(g_srv.mu);
  auto it = g_srv.tickets.find(ticket);
  if (it == g_srv.tickets.end()) return fail(HGX_EINVAL, "unknown ticket");
  auto r = it->second;
  g_srv.cv.wait(lk, [&] { return r->state == 2 || r->state < 0; });
  g_srv.tickets.erase(ticket);
  if (r->state < 0) return fail(r->state, r->err);
  return HGX_OK;
}

int hgx_server_stats(hgx_server_counters *out) {
  if (!out) return fail(HGX_EINVAL, "null output");
  std::lock_guard<std::mutex> lk(g_srv.mu);
  std::memset(out, 0, sizeof *out);
  if (!g_srv.started) return HGX_OK;
  out->completed = g_srv.completed;
  out->failed = g_srv.failed;
  out->rounds = g_srv.rounds;
  out->staged = g_srv.staged_done;
  out->queued = g_srv.queue.size() + g_srv.staged.size();
  out->in_flight = g_srv.inflight.size();
  out->uptime_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - g_srv.t0).count();
  out->launch_s = g_srv.launch_s;
  out->retire_s = g_srv.retire_s;
  out->wait_s = g_srv.wait_s;
  out->lanes = g_srv.nlanes;
  out->split_downloads = g_srv.split_d2h ? 1 : 0;
  out->uploads = g_srv.copy_streams ? g_srv.nup : 0;
  out->downloads = g_srv.copy_streams ? g_srv.ndown : 0;
  return HGX_OK;
}

// ---- pass-structure choices and the autotuner ----
int hgx_plan_set(int bits, int ncol, int direct_max_bits, int mont_direct) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  Choice c = choice_for(bits);
  c.ncol = ncol;
  c.dm = direct_max_bits;
  c.md = mont_direct;
  return set_choice_locked(bits, c);
}

int hgx_kernel_set(int bits, int col_threads, int row_tile_bits) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (bits < 1 || bits > kMaxBits) return fail(HGX_EINVAL, "transform bits out of range");
  Choice c = choice_for(bits);
  c.colw = col_threads;
  c.rowt = row_tile_bits;
  return set_choice_locked(bits, c);
}

int hgx_kernel_get(int bits, int *col_threads, int *row_tile_bits) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  if (bits < 1 || bits > kMaxBits || !col_threads || !row_tile_bits) return fail(HGX_EINVAL, "bad arguments");
  const Choice c = choice_for(bits);
  *col_threads = c.colw;
  *row_tile_bits = c.rowt;
  return HGX_OK;
}

int hgx_plan_get(int bits, int *ncol, int *direct_max_bits, int *mont_direct) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  if (bits < 1 || bits > kMaxBits || !ncol || !direct_max_bits || !mont_direct)
    return fail(HGX_EINVAL, "bad arguments");
  const Choice c = choice_for(bits);
  *ncol = c.ncol;
  *direct_max_bits = c.dm;
  *mont_direct = c.md;
  return HGX_OK;
}

// For every length 2^bits, min_bits <= bits <= max_bits, time a product of two
// 2^(bits-1)-limb operands (SplitMix64 limbs, device memory) under every
// pass structure (column-pass counts; direct-table limits 12, 16, 20, 22 where
// they change a pass) and keep the fastest: best of REPS runs after one warm-up.
// Lengths from which one product fills the GPU (its launches have at least
// kHalfTileBelow workgroups): only there are the kernels tuned on a single
// product; below it the default rule adapts them to the launch size, which
// for batches is larger than one product's.
constexpr int kTuneKernelsFrom = 20;

// For every length 2^bits, min_bits <= bits <= max_bits, time a product of
// two 2^(bits-1)-limb operands (SplitMix64 limbs, device memory) and keep the
// fastest choice: first the pass structure (column-pass counts; direct-table
// limits 12, 16, 20, 22 where they change a pass; direct Montgomery tables),
// then, from 2^kTuneKernelsFrom on, the kernels on it (column tiles of 256 or
// 512 threads, row tiles of 2^12, 2^11 or 2^10 words), each the best of REPS
// runs after a warm-up.
int hgx_autotune(int min_bits, int max_bits, int reps) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (min_bits < 1 || max_bits > kMaxBits || min_bits > max_bits || reps < 1)
    return fail(HGX_EINVAL, "bad autotune range");
  hipStream_t s = g_ctx.stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int n = std::max(min_bits, kRowMaxBits + 1); n <= max_bits && !rc; n++) {
    const u64 L = 1ull << n, half = L / 2;
    // candidates: every column-pass count; the direct-table limits that
    // change which passes use the direct table; direct Montgomery tables
    // where some pass is above the direct limit and at most 2^28 words
    std::vector<Choice> cand;
    for (int ncol = min_ncol(n); ncol <= max_ncol(n); ncol++) {
      std::vector<int> seen;
      for (int dm : {12, 16, 20, 22}) {
        int ndirect = 0, nmont = 0;
        for (int i = 0; i < ncol; i++) {
          ndirect |= (n - 8 * i <= dm) << i;
          nmont |= (n - 8 * i > dm && n - 8 * i <= kMontDirectMaxBits) << i;
        }
        if (std::find(seen.begin(), seen.end(), ndirect) != seen.end()) continue;
        seen.push_back(ndirect);
        cand.push_back(Choice{ncol, dm, 0});
        if (nmont) cand.push_back(Choice{ncol, dm, 1});
      }
    }
    const bool tune_kernels = n >= kTuneKernelsFrom;
    if (cand.size() < 2 && !tune_kernels) continue;
    rc = ensure(g_ctx.stage, 16 * L);
    if (rc) break;
    u64 *a = (u64 *)g_ctx.stage.ptr, *b = a + half, *out = b + half;
    hgx_launch_splitmix64(dim3((unsigned)((half + 255) / 256)), s, a, half, 0x48475800 + n, 1);
    hgx_launch_splitmix64(dim3((unsigned)((half + 255) / 256)), s, b, half, 0x48475900 + n, 1);
    // best of REPS timed runs of one candidate (run 0 builds the tables)
    auto time_choice = [&](const Choice &c, float &t) -> int {
      int r2 = set_choice_locked(n, c);
      if (r2) return r2;
      t = 1e30f;
      for (int r = 0; r <= reps; r++) {
        HIPCHK(hipEventRecord(e0, s));
        r2 = mul_device_locked(a, half, b, half, out, false, s);
        if (r2) return r2;
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) t = std::min(t, ms);
      }
      return HGX_OK;
    };
    float best = 1e30f;
    Choice pick = choice_for(n);
    pick.colw = pick.rowt = 0;
    if (cand.size() >= 2) {
      for (auto &c : cand) {  // setting the next candidate drops this one's tables
        float t;
        rc = time_choice(c, t);
        if (rc) break;
        if (t < best) {
          best = t;
          pick = c;
        }
      }
      if (rc) break;
    } else {
      rc = time_choice(pick, best);
      if (rc) break;
    }
    if (tune_kernels) {
      // kernels on the chosen tables; the default rule stays unless an
      // explicit choice is clearly faster
      const int srow = n - 8 * pick.ncol;
      const bool colw_any = pick.ncol > 0 && n - 8 * (pick.ncol - 1) <= 28;
      Choice kpick = pick;
      float kbest = best;
      for (int colw : {0, 256, 512}) {
        if (colw && !colw_any) continue;
        for (int rowt : {0, 12, 11, 10}) {
          if (rowt && srow > rowt) continue;
          if (!colw && !rowt) continue;  // the default, timed above
          Choice c = pick;
          c.colw = colw;
          c.rowt = rowt;
          float t;
          rc = time_choice(c, t);
          if (rc) break;
          if (t < 0.98f * kbest) {
            kbest = t;
            kpick = c;
          }
        }
        if (rc) break;
      }
      if (rc) break;
      pick = kpick;
    }
    rc = set_choice_locked(n, pick);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  tune_cache_save();
  return HGX_OK;
}

int hgx_staging_limit(size_t bytes) {
  g_ctx.stage_limit.store(bytes);
  return HGX_OK;
}

int hgx_split_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  g_ctx.split_on = on != 0;
  return HGX_OK;
}

int hgx_graphs_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  g_ctx.graphs_on = on != 0;
  return HGX_OK;
}

int hgx_timing_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  timing_drain();
  g_kstats.clear();
  g_timing = on != 0;
  // events for the launches to come, created here rather than between them
  if (g_timing)
    while (g_timing_events.size() < 256) {
      hipEvent_t e = nullptr;
      HIPCHK(hipEventCreate(&e));
      g_timing_events.push_back(e);
    }
  return HGX_OK;
}

int hgx_timing_read(hgx_kernel_time *out, int max_entries) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  timing_drain();
  int n = 0;
  for (auto &kv : g_kstats) {
    if (n >= max_entries || !out) break;
    std::memset(out[n].name, 0, sizeof out[n].name);
    std::strncpy(out[n].name, kv.first.c_str(), sizeof out[n].name - 1);
    out[n].launches = kv.second.first;
    out[n].total_ms = kv.second.second;
    n++;
  }
  return n;
}


// ---- device limb vectors (callers without an allocator of their own) ----
int hgx_device_alloc(size_t nlimbs, uint64_t **d_out) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!d_out) return fail(HGX_EINVAL, "null output");
  *d_out = nullptr;
  void *p = nullptr;
  rc = dev_block_alloc(std::max<size_t>(nlimbs, 1) * 8, &p);
  if (rc) return rc;
  *d_out = (uint64_t *)p;
  return HGX_OK;
}

int hgx_device_free(uint64_t *d, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  if (!d) return HGX_OK;
  int rc = check_ready();
  if (rc) return rc;
  return dev_block_release(d, stream ? (hipStream_t)stream : g_ctx.stream);
}

// Pinned host memory moves by one DMA copy, queued under the library lock
// and awaited outside it (callers on different streams overlap); pageable
// memory goes through the library's pinned staging, under the lock.
int hgx_upload(uint64_t *d_dst, const uint64_t *src, size_t nlimbs, void *stream) {
  hipStream_t s;
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    int rc = check_ready();
    if (rc) return rc;
    if (!nlimbs) return HGX_OK;
    if (!d_dst || !src) return fail(HGX_EINVAL, "null pointer");
    s = stream ? (hipStream_t)stream : g_ctx.stream;
    if (!host_pinned(src)) {
      PinScope pin;
      rc = h2d(d_dst, src, nlimbs * 8, s);
      if (rc) return rc;
      HIPCHK(hipStreamSynchronize(s));  // SRC may be reused once this returns
      return HGX_OK;
    }
    HIPCHK(hipMemcpyAsync(d_dst, src, nlimbs * 8, hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return HGX_OK;
}

int hgx_download(uint64_t *dst, const uint64_t *d_src, size_t nlimbs, void *stream) {
  hipStream_t s;
  {
    std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
    int rc = check_ready();
    if (rc) return rc;
    if (!nlimbs) return HGX_OK;
    if (!dst || !d_src) return fail(HGX_EINVAL, "null pointer");
    s = stream ? (hipStream_t)stream : g_ctx.stream;
    if (!host_pinned(dst)) {
      PinScope pin;
      rc = d2h(dst, d_src, nlimbs * 8, s);
      if (!rc) rc = pin_flush();
      if (rc) return rc;
      HIPCHK(hipStreamSynchronize(s));
      return HGX_OK;
    }
    HIPCHK(hipMemcpyAsync(dst, d_src, nlimbs * 8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return HGX_OK;
}

int hgx_stream_create(void **stream) {
  if (!stream) return fail(HGX_EINVAL, "null output");
  *stream = nullptr;
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  hipStream_t st = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *stream = st;
  return HGX_OK;
}

int hgx_stream_destroy(void *stream) {
  if (!stream) return HGX_OK;
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  HIPCHK(hipStreamDestroy((hipStream_t)stream));
  return HGX_OK;
}

int hgx_device_copy(uint64_t *d_dst, const uint64_t *d_src, size_t nlimbs, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!nlimbs) return HGX_OK;
  if (!d_dst || !d_src) return fail(HGX_EINVAL, "null pointer");
  if (overlaps(d_dst, nlimbs, d_src, nlimbs) && d_dst != d_src) return fail(HGX_EINVAL, "overlapping copy");
  HIPCHK(hipMemcpyAsync(d_dst, d_src, nlimbs * 8, hipMemcpyDeviceToDevice,
                        stream ? (hipStream_t)stream : g_ctx.stream));
  return HGX_OK;
}

int hgx_device_fill(uint64_t *d, size_t nlimbs, uint64_t value, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!nlimbs) return HGX_OK;
  if (!d) return fail(HGX_EINVAL, "null pointer");
  hgx_launch_fill(stream ? (hipStream_t)stream : g_ctx.stream, d, nlimbs, value);
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

int hgx_device_significant(const uint64_t *d, size_t nlimbs, size_t *out, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!out || (nlimbs && !d)) return fail(HGX_EINVAL, "null pointer");
  if (!nlimbs) {
    *out = 0;
    return HGX_OK;
  }
  u64 r[2];
  rc = scan_read(stream ? (hipStream_t)stream : g_ctx.stream, d, nlimbs, nullptr, 0, false, r);
  if (rc) return rc;
  *out = (size_t)r[0];
  return HGX_OK;
}

int hgx_device_compare(const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, int *out, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!out || (na && !d_a) || (nb && !d_b)) return fail(HGX_EINVAL, "null pointer");
  if (!na && !nb) {
    *out = 0;
    return HGX_OK;
  }
  u64 r[2];
  rc = scan_read(stream ? (hipStream_t)stream : g_ctx.stream, d_a, na, d_b, nb, true, r);
  if (rc) return rc;
  *out = r[1] == 0 ? 0 : (r[1] == 1 ? 1 : -1);
  return HGX_OK;
}

int hgx_shl_device(const uint64_t *d_a, size_t n, int bits, uint64_t *d_out, void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (bits < 0 || bits > 63) return fail(HGX_EINVAL, "shift must be 0..63 bits");
  if (!d_out || (n && !d_a)) return fail(HGX_EINVAL, "null pointer");
  if (n && overlaps(d_out, n + 1, d_a, n)) return fail(HGX_EINVAL, "the shift's output overlaps its input");
  hgx_launch_shl(stream ? (hipStream_t)stream : g_ctx.stream, d_a, n, bits, d_out);
  HIPCHK(hipGetLastError());
  return HGX_OK;
}

int hgx_host_alloc(size_t bytes, void **p) {
  if (!p) return fail(HGX_EINVAL, "null output");
  *p = nullptr;
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  const hipError_t e = hipHostMalloc(p, std::max<size_t>(bytes, 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(HGX_ENOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
  }
  return HGX_OK;
}

int hgx_host_free(void *p) {
  if (!p) return HGX_OK;
  HIPCHK(hipHostFree(p));
  return HGX_OK;
}

int hgx_synchronize(void *stream) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(stream ? (hipStream_t)stream : g_ctx.stream));
  return HGX_OK;
}


// ---- throughput probe of the modular arithmetic (integer-ALU ceiling) ----
int hgx_modmul_probe(int op, int iters, double *out) {
  std::lock_guard<std::recursive_mutex> lk(g_ctx.mu);
  int rc = check_ready();
  if (rc) return rc;
  if (!out || op < 0 || op > 6 || iters < 1) return fail(HGX_EINVAL, "bad probe arguments");
  hipDeviceProp_t prop{};
  HIPCHK(hipGetDeviceProperties(&prop, g_ctx.device));
  // eight workgroups of 256 threads per CU: every SIMD holds 8 waves
  const unsigned blocks = (unsigned)prop.multiProcessorCount * 8;
  const u64 threads = (u64)blocks * 256;
  void *mem = nullptr;
  rc = dev_block_alloc(threads * 8 + 256, &mem);
  if (rc) return rc;
  u64 *outd = (u64 *)mem, *clk = outd + threads;
  const Mod M = mod_of(g_ctx.ps, 0);
  const u64 w = powmod(g_ctx.gen[0], (M.p - 1) >> 20, M.p);  // a 2^20-th root of unity
  hipStream_t s = g_ctx.stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  (void)hgx_launch_probe(op, dim3(blocks), s, M, w, shoup_companion(w, M.p), to_mont(w, M.p), 16, outd, clk);  // warm-up
  (void)hipEventRecord(e0, s);
  (void)hgx_launch_probe(op, dim3(blocks), s, M, w, shoup_companion(w, M.p), to_mont(w, M.p), iters, outd, clk);
  (void)hipEventRecord(e1, s);
  u64 cyc[2] = {0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(cyc, clk, 16, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)dev_block_release(mem, s);
  if (e != hipSuccess) return fail(HGX_EDEVICE, std::string("probe failed: ") + hipGetErrorString(e));
  const double ops = (double)threads * (op >= 5 ? 16 : hgx_probe_chains()) * iters;
  out[0] = ops / (ms * 1e-3);
  out[1] = cyc[1] ? (double)cyc[0] / ((double)cyc[1] / 100e6) / 1e9 : 0.0;
  out[2] = ms * 1e-3;
  return HGX_OK;
}

}  // extern "C"
