// hgx_series.cpp -- binary splitting of a hypergeometric series, one tree
// height at a time, with the products of each height batched on the GPU.
//
// The reference evaluates the partial sum of a series over [lower, upper)
// depth first (binary-split, src/binary-splitting.lisp:123-133): split at
// floor((lower + upper) / 2), recurse, COMBINE (:112-121) the two partials
// with seven products and one sum, and at one term the base case (:100-110).
// This engine builds the same tree, folds every subtree of at most
// kFoldTerms terms (the partial of an interval is the same four integers for
// any split, so a left fold of the base cases gives the recursion's
// numbers), and then evaluates the tree bottom up, one height at a time.  A height's products go out in two rounds (five products that need
// only the children, then the two that need round-1 results).  While the
// values are in host memory, a product whose smaller operand has more than
// *ntt-multiply-threshold* limbs (src/multiply.lisp:9) joins one batched GPU
// call (hgx_mul_batch), the others are multiplied on the host by schoolbook
// while the GPU works, as the reference's mpz-* multiplies them
// (src/multiply.lisp:21-42).  A split whose heights hold enough limbs moves
// its values to device memory and keeps them there: each round is one
// batched device call (hgx_mul_batch_device), the sums R run on the device,
// and the result stays there (hgx_partial_get downloads it,
// hgx_partial_get_device copies it device to device).  Such a split folds its
// leaves on the device too, one thread per leaf (hgx_fold.hpp); other splits
// fold them on host threads (and a split that moves later uploads them).
// All arithmetic is exact, so every partial equals the reference's.
//
// Series terms are polynomials in k with 64-bit signed coefficients and an
// optional separate value at k = 0, which covers the reference's series
// (make-e-series, make-exp-series, make-ramanujan-series,
// make-chudnovsky-series, make-catalan-series, make-apery-series).
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <utility>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hgx.h"
#include "hgx_fold.hpp"
#include "hgx_sum.h"

typedef uint64_t u64;
typedef unsigned __int128 u128;

// Limb vectors whose resize leaves new limbs uninitialised: products are
// written whole by the GPU or by mag_mul, so zeroing them first would only
// add a pass over memory.
template <class T>
struct NoInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <class U>
  NoInit(const NoInit<U> &) noexcept {}
  template <class U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
typedef std::vector<u64, NoInit<u64>> Limbs;

// defined in hgx_runtime.hip
int hgx_internal_fail(int code, const char *msg);
bool hgx_internal_ready();

namespace {

constexpr u64 kFoldTerms = 32;
constexpr size_t kLeafChunk = 512;  // leaves folded by one task
// a height moves to the GPU when its round-1 products have at least this
// many limbs, and the device has room for kDeviceRoomFactor times the live
// values (a height's products take about twice as many limbs as its
// children, the children's regions live until the height is done, and the
// largest product's transforms take about four times the live values)
constexpr size_t kDeviceMinLimbs = (size_t)1 << 20;
constexpr size_t kDeviceRoomFactor = 10;
constexpr size_t kHostMaxLimbs = 100;  // *ntt-multiply-threshold*, src/multiply.lisp:9
constexpr int kMaxDeg = 7;
constexpr int kSmallLimbs = 16;  // term values and products of two of them

// A signed integer of at most kSmallLimbs limbs (a term value).
struct Small {
  int s = 0;  // -1, 0, 1
  int n = 0;  // limbs in use (trimmed)
  u64 m[kSmallLimbs] = {};
};

// A signed integer of any size, grown in place (a leaf's fold).
struct Acc {
  int s = 0;
  Limbs m;
};

void trim(Limbs &m) {
  while (!m.empty() && m.back() == 0) m.pop_back();
}

void small_trim(Small &x) {
  while (x.n > 0 && x.m[x.n - 1] == 0) x.n--;
  if (x.n == 0) x.s = 0;
}

int mag_cmp(const u64 *a, size_t na, const u64 *b, size_t nb) {
  if (na != nb) return na < nb ? -1 : 1;
  for (size_t i = na; i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

// r[0 .. na+nb) = a * b by schoolbook, the outer loop over the shorter
// operand B (%multiply-storage/schoolboy, src/multiply.lisp:37-42).  The
// first row is written, the later ones accumulate: row j writes r[j + na]
// and adds into r[j .. j + na).
void mag_mul(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *r) {
  if (nb == 0) {
    std::fill(r, r + na, 0);
    return;
  }
  {
    const u64 b0 = b[0];
    u64 carry = 0;
    for (size_t i = 0; i < na; i++) {
      const u128 t = (u128)a[i] * b0 + carry;
      r[i] = (u64)t;
      carry = (u64)(t >> 64);
    }
    r[na] = carry;
  }
  for (size_t j = 1; j < nb; j++) {
    const u64 bj = b[j];
    u64 carry = 0;
    u64 *rj = r + j;
    if (bj)
      for (size_t i = 0; i < na; i++) {
        const u128 t = (u128)a[i] * bj + rj[i] + carry;
        rj[i] = (u64)t;
        carry = (u64)(t >> 64);
      }
    rj[na] = carry;
  }
}

// x += y (magnitudes); x grows as needed
void mag_add_to(Limbs &x, const u64 *y, size_t ny) {
  if (x.size() < ny) x.resize(ny, 0);
  u64 c = 0;
  size_t i = 0;
  for (; i < ny; i++) {
    const u128 t = (u128)x[i] + y[i] + c;
    x[i] = (u64)t;
    c = (u64)(t >> 64);
  }
  for (; c && i < x.size(); i++) {
    x[i] += 1;
    c = (x[i] == 0);
  }
  if (c) x.push_back(1);
}

// x = |x - y| for |x| >= |y| (magnitudes)
void mag_sub_from(Limbs &x, const u64 *y, size_t ny) {
  u64 b = 0;
  size_t i = 0;
  for (; i < ny; i++) {
    const u64 xi = x[i], t = xi - y[i];
    const u64 b2 = (xi < y[i]) || (t < b);
    x[i] = t - b;
    b = b2;
  }
  for (; b && i < x.size(); i++) {
    b = (x[i] == 0);
    x[i] -= 1;
  }
  trim(x);
}

// x = x + y for signed x, y (mpz-+)
void big_add(Acc &x, const Acc &y) {
  if (y.s == 0) return;
  if (x.s == 0) {
    x = y;
    return;
  }
  if (x.s == y.s) {
    mag_add_to(x.m, y.m.data(), y.m.size());
    return;
  }
  const int c = mag_cmp(x.m.data(), x.m.size(), y.m.data(), y.m.size());
  if (c == 0) {
    x.s = 0;
    x.m.clear();
  } else if (c > 0) {
    mag_sub_from(x.m, y.m.data(), y.m.size());
  } else {
    Limbs t = y.m;
    mag_sub_from(t, x.m.data(), x.m.size());
    x.m.swap(t);
    x.s = y.s;
  }
}

// x = x * v for a small v; TMP is scratch
void big_mul_small(Acc &x, const Small &v, Limbs &tmp) {
  if (x.s == 0) return;
  if (v.s == 0) {
    x.s = 0;
    x.m.clear();
    return;
  }
  tmp.resize(x.m.size() + v.n);
  mag_mul(x.m.data(), x.m.size(), v.m, v.n, tmp.data());
  trim(tmp);
  x.m.swap(tmp);
  x.s *= v.s;
}

Acc big_of(const Small &v) {
  Acc b;
  b.s = v.s;
  b.m.assign(v.m, v.m + v.n);
  return b;
}

// z = x * y for small x, y (the product fits: each factor is a term value)
int small_mul(const Small &x, const Small &y, Small &z) {
  z = Small();
  if (x.s == 0 || y.s == 0) return HGX_OK;
  if (x.n + y.n > kSmallLimbs) return hgx_internal_fail(HGX_EINVAL, "series term values exceed 1024 bits");
  u64 r[2 * kSmallLimbs];
  mag_mul(x.m, x.n, y.m, y.n, r);
  z.n = x.n + y.n;
  std::copy(r, r + z.n, z.m);
  z.s = x.s * y.s;
  small_trim(z);
  return HGX_OK;
}

// v = v * k + c for k >= 0 and a signed 64-bit c (one Horner step)
int small_horner(Small &v, u64 k, int64_t c) {
  // v * k
  u64 carry = 0;
  for (int i = 0; i < v.n; i++) {
    const u128 t = (u128)v.m[i] * k + carry;
    v.m[i] = (u64)t;
    carry = (u64)(t >> 64);
  }
  if (carry) {
    if (v.n == kSmallLimbs) return hgx_internal_fail(HGX_EINVAL, "series term values exceed 1024 bits");
    v.m[v.n++] = carry;
  }
  small_trim(v);
  if (c == 0) return HGX_OK;
  const int cs = c < 0 ? -1 : 1;
  const u64 cm = c < 0 ? (u64)0 - (u64)c : (u64)c;
  if (v.s == 0) {
    v.s = cs;
    v.n = 1;
    v.m[0] = cm;
    return HGX_OK;
  }
  if (v.s == cs) {  // magnitudes add
    u64 cy = cm;
    for (int i = 0; cy && i < v.n; i++) {
      v.m[i] += cy;
      cy = (v.m[i] < cy);
    }
    if (cy) {
      if (v.n == kSmallLimbs) return hgx_internal_fail(HGX_EINVAL, "series term values exceed 1024 bits");
      v.m[v.n++] = cy;
    }
    return HGX_OK;
  }
  // opposite signs: |v| - cm or cm - |v|
  if (v.n == 1 && v.m[0] < cm) {
    v.m[0] = cm - v.m[0];
    v.s = cs;
    return HGX_OK;
  }
  u64 b = cm;
  for (int i = 0; b && i < v.n; i++) {
    const u64 old = v.m[i];
    v.m[i] = old - b;
    b = (old < b);
  }
  small_trim(v);
  return HGX_OK;
}

// One of the series' four term functions: a polynomial in k, or v0 at k = 0.
struct Poly {
  int deg = 0;
  int64_t c[kMaxDeg + 1] = {};
  bool at_zero = false;
  int64_t v0 = 0;
};

int poly_eval(const Poly &f, u64 k, Small &out) {
  out = Small();
  if (k == 0 && f.at_zero) return small_horner(out, 0, f.v0);
  for (int i = f.deg; i >= 0; i--) {
    const int rc = small_horner(out, k, f.c[i]);
    if (rc) return rc;
  }
  return HGX_OK;
}

struct Series {
  Poly a, b, p, q;
};

struct AccPart {
  Acc P, Q, B, R;
};

// The partial over [lo, up): the base case at LO (binary-split-base-case=1,
// src/binary-splitting.lisp:100-110) combined with the base cases lo+1, ...
// one at a time on the right: COMBINE with (p_k, q_k, b_k, p_k a_k) gives
// R <- b_k q_k R + B P p_k a_k, P <- P p_k, Q <- Q q_k, B <- B b_k.
int fold(const Series &s, u64 lo, u64 up, AccPart &out) {
  Small pk, qk, bk, ak, t1, t2;
  int rc = poly_eval(s.p, lo, pk);
  if (!rc) rc = poly_eval(s.q, lo, qk);
  if (!rc) rc = poly_eval(s.b, lo, bk);
  if (!rc) rc = poly_eval(s.a, lo, ak);
  if (!rc) rc = small_mul(pk, ak, t1);
  if (rc) return rc;
  out.P = big_of(pk);
  out.Q = big_of(qk);
  out.B = big_of(bk);
  out.R = big_of(t1);
  Acc BP = out.B;  // B * P of the terms so far
  Acc x;           // BP p_k a_k (its capacity kept from term to term)
  Limbs tmp;
  big_mul_small(BP, pk, tmp);
  for (u64 k = lo + 1; k < up; k++) {
    rc = poly_eval(s.p, k, pk);
    if (!rc) rc = poly_eval(s.q, k, qk);
    if (!rc) rc = poly_eval(s.b, k, bk);
    if (!rc) rc = poly_eval(s.a, k, ak);
    if (!rc) rc = small_mul(bk, qk, t1);
    if (!rc) rc = small_mul(pk, ak, t2);
    if (rc) return rc;
    big_mul_small(out.R, t1, tmp);
    x = BP;
    big_mul_small(x, t2, tmp);
    big_add(out.R, x);
    rc = small_mul(bk, pk, t1);
    if (rc) return rc;
    big_mul_small(BP, t1, tmp);
    big_mul_small(out.P, pk, tmp);
    big_mul_small(out.Q, qk, tmp);
    big_mul_small(out.B, bk, tmp);
  }
  return HGX_OK;
}

unsigned worker_count() {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return std::min(hw, 16u);
}

// fn(i) for i < n on up to worker_count() threads; the first failure wins
int parallel_for(size_t n, const std::function<int(size_t)> &fn) {
  if (n == 0) return HGX_OK;
  const unsigned nt = (unsigned)std::min<size_t>(worker_count(), n);
  std::atomic<size_t> next{0};
  std::atomic<int> err{HGX_OK};
  std::string msg;
  std::atomic_flag msg_taken = ATOMIC_FLAG_INIT;
  auto work = [&]() {
    for (;;) {
      if (err.load() != HGX_OK) return;
      const size_t i = next.fetch_add(1);
      if (i >= n) return;
      int rc;
      try {
        rc = fn(i);
      } catch (const std::bad_alloc &) {
        rc = hgx_internal_fail(HGX_ENOMEM, "out of host memory in binary splitting");
      } catch (const std::exception &e) {
        rc = hgx_internal_fail(HGX_EINVAL, e.what());
      }
      if (rc != HGX_OK) {
        int expect = HGX_OK;
        if (err.compare_exchange_strong(expect, rc) && !msg_taken.test_and_set()) msg = hgx_last_error();
        return;
      }
    }
  };
  if (nt == 1) {
    work();
  } else {
    // a thread that cannot be created leaves its share to the others (the
    // work is taken from one counter), so creation failures cost only speed
    std::vector<std::thread> th;
    try {
      th.reserve(nt - 1);
      for (unsigned t = 1; t < nt; t++) th.emplace_back(work);
    } catch (const std::exception &) {
    }
    work();
    for (auto &t : th) t.join();
  }
  if (err.load() != HGX_OK) return hgx_internal_fail(err.load(), msg.c_str());
  return HGX_OK;
}

// ---- storage of the tree's values ----
// The values of the tree live in regions: one block of limbs holds every
// product of one round of one height, or the partials of a chunk of folded
// leaves.  A value keeps its region alive; a region that no value refers to
// goes back to the split's cache, and a later height takes a cached block of
// about the size it needs.  Each height's products take about the same
// number of limbs in all, so after the first heights the split reuses
// memory it has written before: no allocation or free per value, and no
// fresh pages for the kernel to fault in and zero.
class Cache {
 public:
  ~Cache() {
    for (int i = 0; i < n_; i++) std::free(base_[i]);
  }
  // the smallest cached block of WANT..2*WANT limbs, or null
  u64 *take(size_t want, size_t &got) {
    std::lock_guard<std::mutex> g(mu_);
    int best = -1;
    for (int i = 0; i < n_; i++)
      if (cap_[i] >= want && cap_[i] / 2 <= want && (best < 0 || cap_[i] < cap_[best])) best = i;
    if (best < 0) return nullptr;
    u64 *p = base_[best];
    got = cap_[best];
    n_--;
    base_[best] = base_[n_];
    cap_[best] = cap_[n_];
    return p;
  }
  // keeps P for reuse (up to kMax blocks), or frees it
  void give(u64 *p, size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (n_ < kMax) {
        base_[n_] = p;
        cap_[n_] = cap;
        n_++;
        return;
      }
    }
    std::free(p);
  }

 private:
  static constexpr int kMax = 64;
  std::mutex mu_;
  u64 *base_[kMax];
  size_t cap_[kMax];
  int n_ = 0;
};

struct Region {
  u64 *base = nullptr;
  size_t cap = 0;
  std::weak_ptr<Cache> cache;  // where the block goes back to, while the split runs
  ~Region() {
    if (!base) return;
    if (auto c = cache.lock())
      c->give(base, cap);
    else
      std::free(base);
  }
};
typedef std::shared_ptr<Region> RegionRef;

// a region of at least LIMBS limbs (throws std::bad_alloc)
RegionRef new_region(const std::shared_ptr<Cache> &cache, size_t limbs) {
  RegionRef r = std::make_shared<Region>();
  limbs = std::max<size_t>(limbs, 1);
  size_t got = 0;
  u64 *p = cache ? cache->take(limbs, got) : nullptr;
  if (!p) {
    p = static_cast<u64 *>(std::malloc(8 * limbs));
    if (!p) throw std::bad_alloc();
    got = limbs;
  }
  r->base = p;
  r->cap = got;
  r->cache = cache;
  return r;
}

// A signed integer in a slot of CAP limbs of a region of host memory, or of
// device memory once the split has moved to the GPU (DEV).  Host values are
// trimmed; device values may keep a zero top limb from a product.
struct Num {
  int s = 0;
  size_t n = 0;  // limbs in use
  size_t cap = 0;
  u64 *p = nullptr;
  bool dev = false;
  std::shared_ptr<void> own;  // the region
};

// one round's products in device memory (hgx_device_alloc), released after
// the work queued before the release
struct DevRegion {
  u64 *base = nullptr;
  size_t cap = 0;  // limbs
  ~DevRegion() {
    if (base) (void)hgx_device_free(base, nullptr);
  }
};

int new_dev_region(size_t limbs, std::shared_ptr<DevRegion> &out) {
  out = std::make_shared<DevRegion>();
  out->cap = std::max<size_t>(limbs, 1);
  return hgx_device_alloc(out->cap, &out->base);
}

struct Part {
  Num P, Q, B, R;
};

void num_trim(Num &x) {
  while (x.n > 0 && x.p[x.n - 1] == 0) x.n--;
  if (x.n == 0) x.s = 0;
}

// x = x + y for signed x, y (mpz-+), in x's slot, which holds at least
// max(|x|, |y|) + 1 limbs
int num_add(Num &x, const Num &y) {
  if (y.s == 0) return HGX_OK;
  const size_t xn = x.s ? x.n : 0;
  if (x.cap < std::max(xn, y.n) + 1) return hgx_internal_fail(HGX_EINVAL, "binary splitting: sum slot too small");
  if (xn == 0) {
    std::copy(y.p, y.p + y.n, x.p);
    x.n = y.n;
    x.s = y.s;
    return HGX_OK;
  }
  if (x.s == y.s) {  // |x| += |y|
    if (x.n < y.n) {
      std::fill(x.p + x.n, x.p + y.n, 0);
      x.n = y.n;
    }
    u64 c = 0;
    size_t i = 0;
    for (; i < y.n; i++) {
      const u128 t = (u128)x.p[i] + y.p[i] + c;
      x.p[i] = (u64)t;
      c = (u64)(t >> 64);
    }
    for (; c && i < x.n; i++) {
      x.p[i] += 1;
      c = (x.p[i] == 0);
    }
    if (c) x.p[x.n++] = 1;
    return HGX_OK;
  }
  const int c = mag_cmp(x.p, x.n, y.p, y.n);
  if (c == 0) {
    x.s = 0;
    x.n = 0;
    return HGX_OK;
  }
  if (c > 0) {  // |x| -= |y|
    u64 b = 0;
    size_t i = 0;
    for (; i < y.n; i++) {
      const u64 xi = x.p[i], t = xi - y.p[i];
      const u64 b2 = (xi < y.p[i]) || (t < b);
      x.p[i] = t - b;
      b = b2;
    }
    for (; b && i < x.n; i++) {
      b = (x.p[i] == 0);
      x.p[i] -= 1;
    }
  } else {  // |x| = |y| - |x|, with |x| < |y|
    std::fill(x.p + x.n, x.p + y.n, 0);
    u64 b = 0;
    for (size_t i = 0; i < y.n; i++) {
      const u64 yi = y.p[i], t = yi - x.p[i];
      const u64 b2 = (yi < x.p[i]) || (t < b);
      x.p[i] = t - b;
      b = b2;
    }
    x.n = y.n;
    x.s = y.s;
  }
  num_trim(x);
  return HGX_OK;
}

// The products OUT[i] = X[i] * Y[i] in one new region, slot i of ROOM[i] >=
// |X[i]| + |Y[i]| limbs: on the GPU in one batched call when the smaller
// magnitude has more than kHostMaxLimbs limbs, by schoolbook on the host
// threads otherwise (while the GPU call runs).
int mul_many(const std::shared_ptr<Cache> &cache, const std::vector<const Num *> &xs,
             const std::vector<const Num *> &ys, const std::vector<size_t> &room, std::vector<Num> &out) {
  const size_t n = xs.size();
  size_t total = 0;
  for (size_t i = 0; i < n; i++) total += room[i];
  const RegionRef reg = new_region(cache, total);
  out.assign(n, Num());
  std::vector<size_t> gpu, host;
  size_t at = 0;
  for (size_t i = 0; i < n; i++) {
    const Num &x = *xs[i], &y = *ys[i];
    Num &o = out[i];
    o.p = reg->base + at;
    o.cap = room[i];
    o.own = reg;
    at += room[i];
    if (x.s == 0 || y.s == 0) continue;
    if (o.cap < x.n + y.n) return hgx_internal_fail(HGX_EINVAL, "binary splitting: product slot too small");
    o.s = x.s * y.s;
    o.n = x.n + y.n;
    (std::min(x.n, y.n) > kHostMaxLimbs ? gpu : host).push_back(i);
  }
  int rc_gpu = HGX_OK;
  std::string gpu_err;
  std::thread gpu_thread;
  auto gpu_call = [&]() {
    try {
      std::vector<hgx_mul_desc> d(gpu.size());
      for (size_t k = 0; k < gpu.size(); k++) {
        const size_t i = gpu[k];
        const Num &x = *xs[i], &y = *ys[i];
        d[k] = hgx_mul_desc{x.p, x.n, y.p, y.n, out[i].p};
      }
      rc_gpu = hgx_mul_batch(d.data(), d.size());
      if (rc_gpu) gpu_err = hgx_last_error();
    } catch (const std::bad_alloc &) {
      rc_gpu = HGX_ENOMEM;
      gpu_err = "out of host memory in binary splitting";
    } catch (const std::exception &e) {
      rc_gpu = HGX_EINVAL;
      gpu_err = e.what();
    }
  };
  if (!gpu.empty()) {
    if (!hgx_internal_ready())
      return hgx_internal_fail(HGX_ENOTINIT, "hgx_init has not been called (a series product needs the GPU)");
    try {
      gpu_thread = std::thread(gpu_call);
    } catch (const std::exception &) {
      gpu_call();  // no thread: the batch runs before the host products
    }
  }
  const int rc_host = parallel_for(host.size(), [&](size_t k) {
    const size_t i = host[k];
    const Num &x = *xs[i], &y = *ys[i];
    if (x.n >= y.n)
      mag_mul(x.p, x.n, y.p, y.n, out[i].p);
    else
      mag_mul(y.p, y.n, x.p, x.n, out[i].p);
    return (int)HGX_OK;
  });
  if (gpu_thread.joinable()) gpu_thread.join();
  if (rc_gpu) return hgx_internal_fail(rc_gpu, gpu_err.c_str());
  if (rc_host) return rc_host;
  for (size_t i = 0; i < n; i++)
    if (out[i].s) num_trim(out[i]);
  return HGX_OK;
}

// The products of a height whose values live in device memory: one device
// region for all, every product in one batched call (hgx_mul_batch_device,
// queued on the library's stream).  A product keeps all |X| + |Y| limbs.
int mul_many_dev(const std::vector<const Num *> &xs, const std::vector<const Num *> &ys,
                 const std::vector<size_t> &room, std::vector<Num> &out) {
  const size_t n = xs.size();
  size_t total = 0;
  for (size_t i = 0; i < n; i++) total += room[i];
  std::shared_ptr<DevRegion> reg;
  int rc = new_dev_region(total, reg);
  if (rc) return rc;
  out.assign(n, Num());
  std::vector<hgx_mul_desc> d;
  d.reserve(n);
  size_t at = 0;
  for (size_t i = 0; i < n; i++) {
    const Num &x = *xs[i], &y = *ys[i];
    Num &o = out[i];
    o.p = reg->base + at;
    o.cap = room[i];
    o.dev = true;
    o.own = reg;
    at += room[i];
    if (x.s == 0 || y.s == 0) continue;
    if (!x.dev || !y.dev || o.cap < x.n + y.n)
      return hgx_internal_fail(HGX_EINVAL, "binary splitting: bad device product slot");
    o.s = x.s * y.s;
    o.n = x.n + y.n;
    d.push_back(hgx_mul_desc{x.p, x.n, y.p, y.n, o.p});
  }
  return d.empty() ? HGX_OK : hgx_mul_batch_device(d.data(), d.size(), nullptr);
}

// R = B_r Q_r R_l + B_l P_l R_r for every node of a device height: O2 holds
// the pairs (B_r Q_r R_l, B_l P_l R_r); the sums go over the first of each
// pair (hgx_internal_sums_device), whose sign and trimmed size come back.
int sums_dev(std::vector<Num> &o2, size_t cnt) {
  std::vector<HgxSumJob> jobs;
  std::vector<size_t> idx;
  jobs.reserve(cnt);
  idx.reserve(cnt);
  for (size_t i = 0; i < cnt; i++) {
    Num &x = o2[2 * i];
    const Num &y = o2[2 * i + 1];
    if (y.s == 0) continue;  // R is the first product as it stands
    const size_t nx = x.s ? x.n : 0;
    if (x.cap < std::max(nx, y.n) + 1) return hgx_internal_fail(HGX_EINVAL, "binary splitting: sum slot too small");
    jobs.push_back(HgxSumJob{x.p, y.p, nx, y.n, x.s, y.s, 0, 0});
    idx.push_back(i);
  }
  if (jobs.empty()) return HGX_OK;
  std::vector<int> sg(jobs.size());
  std::vector<uint64_t> sz(jobs.size());
  const int rc = hgx_internal_sums_device(jobs.data(), jobs.size(), sg.data(), sz.data());
  if (rc) return rc;
  for (size_t k = 0; k < jobs.size(); k++) {
    Num &x = o2[2 * idx[k]];
    x.s = sz[k] ? sg[k] : 0;
    x.n = x.s ? sz[k] : 0;
  }
  return HGX_OK;
}

// Moves every live value to device memory: one region, filled by one
// gathered upload through the library's pinned staging (no pinned memory of
// the split's own to fault in).
int to_device(std::vector<std::pair<u64, Part>> &live) {
  std::vector<Num *> vals;
  vals.reserve(4 * live.size());
  std::vector<const u64 *> src;
  std::vector<size_t> len;
  src.reserve(4 * live.size());
  len.reserve(4 * live.size());
  size_t total = 0;
  for (auto &e : live)
    for (Num *v : {&e.second.P, &e.second.Q, &e.second.B, &e.second.R}) {
      vals.push_back(v);
      const size_t n = v->s ? v->n : 0;
      src.push_back(v->p);
      len.push_back(n);
      total += n;
    }
  std::shared_ptr<DevRegion> reg;
  int rc = new_dev_region(total, reg);
  if (rc) return rc;
  rc = hgx_internal_upload_gather(reg->base, src.data(), len.data(), src.size());
  if (rc) return rc;
  size_t at = 0;
  for (size_t k = 0; k < vals.size(); k++) {
    Num &v = *vals[k];
    v.p = reg->base + at;
    at += len[k];
    v.cap = len[k];
    v.dev = true;
    v.own = reg;
  }
  return HGX_OK;
}

// HGX_SERIES_DEVICE: '0' keeps a split on the host (0), '1' sends any split
// to the device (2), unset or other lets the sizes decide (1)
int device_mode() {
  const char *env = std::getenv("HGX_SERIES_DEVICE");
  return env && env[0] == '0' ? 0 : (env && env[0] == '1' ? 2 : 1);
}

// The leaves folded on the device (hgx_internal_fold_device), for a split
// whose heights will live there: the leaves' values, bounded from the
// series' coefficients, hold at least kDeviceMinLimbs limbs and fit the
// device several times over (or HGX_SERIES_DEVICE=1).  DONE is set when
// every leaf's partial is in one device region; otherwise the split folds
// on the host -- also when a term value exceeds kFoldTV limbs, a leaf
// outgrows its buffers or device memory is short.  TIGHT (a test hook)
// gives the fold buffers their smallest capacities, so that leaves of more
// than a few limbs fail on the device.
int leaves_on_device(const Series &s, const std::vector<std::pair<u64, u64>> &blocks, u64 lower, u64 upper,
                     bool tight, std::vector<std::pair<u64, Part>> &leaf, bool &done) {
  done = false;
  if (!hgx_internal_ready() || blocks.empty()) return HGX_OK;
  const int mode = device_mode();
  if (mode == 0) return HGX_OK;
  FoldSeries fs;
  std::memset(&fs, 0, sizeof fs);
  const Poly *polys[4] = {&s.a, &s.b, &s.p, &s.q};
  for (int i = 0; i < 4; i++) {
    FoldPoly &f = fs.f[i];
    f.deg = polys[i]->deg;
    f.at_zero = polys[i]->at_zero ? 1 : 0;
    f.v0 = polys[i]->v0;
    for (int j = 0; j <= f.deg; j++) f.c[j] = polys[i]->c[j];
  }
  u64 terms = 0;
  for (const auto &b : blocks) terms = std::max(terms, b.second - b.first);
  u32 cap[kFoldBufs];
  u64 leaf_limbs = 0;
  if (!fold_caps(fs, lower, upper, terms, cap, leaf_limbs)) return HGX_OK;
  const size_t count = blocks.size();
  u64 per = 0;
  for (int b = 0; b < kFoldBufs; b++) per += cap[b];
  // (divisions, not products, so that no size can overflow)
  if (mode == 1 && leaf_limbs < (kDeviceMinLimbs + count - 1) / count) return HGX_OK;
  const size_t room = hgx_internal_device_room();
  if (leaf_limbs > room / (8 * kDeviceRoomFactor) / count || per > room / 16 / count) return HGX_OK;
  if (tight)
    for (int b = 0; b < kFoldBufs; b++) cap[b] = kFoldTV2 + 1;
  std::vector<u64> bounds(2 * count);
  for (size_t t = 0; t < count; t++) {
    bounds[2 * t] = blocks[t].first;
    bounds[2 * t + 1] = blocks[t].second;
  }
  std::vector<int32_t> codes(4 * count);
  u64 *dv = nullptr;
  size_t total = 0;
  const int rc = hgx_internal_fold_device(&fs, bounds.data(), count, cap, codes.data(), &dv, &total);
  if (rc == HGX_ETOOBIG || rc == HGX_ENOMEM) return HGX_OK;  // the host folds
  if (rc) return rc;
  std::shared_ptr<DevRegion> reg = std::make_shared<DevRegion>();
  reg->base = dv;
  reg->cap = std::max<size_t>(total, 1);
  size_t at = 0;
  for (size_t t = 0; t < count; t++) {
    Part &pt = leaf[t].second;
    Num *dst[4] = {&pt.P, &pt.Q, &pt.B, &pt.R};
    for (int w = 0; w < 4; w++) {
      const int32_t c = codes[4 * t + w];
      Num &d = *dst[w];
      d.s = c > 0 ? 1 : (c < 0 ? -1 : 0);
      d.n = (size_t)(c < 0 ? -(int64_t)c : c);
      d.cap = d.n;
      d.p = dv + at;
      d.dev = true;
      d.own = reg;
      at += d.n;
    }
  }
  done = true;
  return HGX_OK;
}

// Whether the split moves to the GPU before the height whose round-1
// products are XS[i] * YS[i]: once a height's products hold at least
// kDeviceMinLimbs limbs, the values stay in device memory from there on --
// every product of a height in one batched device call, small ones too, and
// no host packing, copies or host sums per height -- when the device has
// room for several times the live values.  (The products are exact either
// way; a height's products take about the same number of limbs at every
// height, so a large split moves at its first height.)  RC receives a
// failure of the move.  HGX_SERIES_DEVICE=0 keeps every height on the host,
// =1 moves any split at its first height.
bool move_to_device(std::vector<std::pair<u64, Part>> &live, const std::vector<const Num *> &xs,
                    const std::vector<const Num *> &ys, int &rc) {
  rc = HGX_OK;
  if (!hgx_internal_ready()) return false;
  const int mode = device_mode();
  if (mode == 0) return false;
  size_t all = 0;
  for (size_t i = 0; i < xs.size(); i++)
    if (xs[i]->s && ys[i]->s) all += xs[i]->n + ys[i]->n;
  if (mode == 1 && all < kDeviceMinLimbs) return false;
  size_t held = 0;
  for (const auto &e : live) held += e.second.P.n + e.second.Q.n + e.second.B.n + e.second.R.n;
  if (held > hgx_internal_device_room() / (8 * kDeviceRoomFactor)) return false;
  rc = to_device(live);
  if (rc == HGX_ENOMEM) {  // no device memory for the move: the split stays on the host
    rc = HGX_OK;
    return false;
  }
  return rc == HGX_OK;
}

// Moves every live value back to host memory (one download per device
// region that holds live values), trimmed, for a split whose device memory
// ran out.
int to_host(std::vector<std::pair<u64, Part>> &live, const std::shared_ptr<Cache> &cache) {
  std::vector<std::pair<DevRegion *, RegionRef>> moved;
  for (auto &e : live)
    for (Num *v : {&e.second.P, &e.second.Q, &e.second.B, &e.second.R}) {
      if (!v->dev) continue;
      DevRegion *d = static_cast<DevRegion *>(v->own.get());
      RegionRef h;
      for (auto &m : moved)
        if (m.first == d) h = m.second;
      if (!h) {
        h = new_region(cache, d->cap);
        const int rc = hgx_download(h->base, d->base, d->cap, nullptr);
        if (rc) return rc;
        moved.emplace_back(d, h);
      }
      v->p = h->base + (v->p - d->base);
      v->dev = false;
      v->own = h;
      if (v->s) num_trim(*v);
    }
  return HGX_OK;
}

// The split tree down to subtrees of at most kFoldTerms terms (BLOCKS);
// every inner node as (height, lo, mid, up); returns the height of [lo, up).
int build_tree(u64 lo, u64 up, std::vector<std::array<u64, 4>> &nodes, std::vector<std::pair<u64, u64>> &blocks) {
  if (up - lo <= kFoldTerms) {
    blocks.push_back({lo, up});
    return 0;
  }
  const u64 m = lo + (up - lo) / 2;  // floor((lo + up) / 2) without overflow
  const int h = 1 + std::max(build_tree(lo, m, nodes, blocks), build_tree(m, up, nodes, blocks));
  nodes.push_back({(u64)h, lo, m, up});
  return h;
}

int parse_poly(const hgx_series &s, int which, Poly &f) {
  f.deg = s.degree[which];
  if (f.deg < 0 || f.deg > kMaxDeg) return hgx_internal_fail(HGX_EINVAL, "series polynomial degree must be 0..7");
  for (int i = 0; i <= f.deg; i++) f.c[i] = s.coef[which][i];
  f.at_zero = s.at_zero[which] != 0;
  f.v0 = s.value0[which];
  return HGX_OK;
}

}  // namespace

// P, Q, B, R in one region of their own
struct hgx_partial {
  Num v[4];
};

extern "C" {

static int series_split(const hgx_series *desc, uint64_t lower, uint64_t upper, hgx_partial **out);

int hgx_series_split(const hgx_series *desc, uint64_t lower, uint64_t upper, hgx_partial **out) {
  try {
    return series_split(desc, lower, upper, out);
  } catch (const std::bad_alloc &) {
    return hgx_internal_fail(HGX_ENOMEM, "out of host memory in binary splitting");
  } catch (const std::exception &e) {
    return hgx_internal_fail(HGX_EINVAL, e.what());
  }
}

static int series_split(const hgx_series *desc, uint64_t lower, uint64_t upper, hgx_partial **out) {
  if (!desc || !out) return hgx_internal_fail(HGX_EINVAL, "null series or result pointer");
  *out = nullptr;
  if (upper <= lower) return hgx_internal_fail(HGX_EINVAL, "the interval [lower, upper) is empty");
  if (upper > (1ull << 40)) return hgx_internal_fail(HGX_EINVAL, "series terms are limited to k < 2^40");
  Series s;
  int rc = parse_poly(*desc, 0, s.a);
  if (!rc) rc = parse_poly(*desc, 1, s.b);
  if (!rc) rc = parse_poly(*desc, 2, s.p);
  if (!rc) rc = parse_poly(*desc, 3, s.q);
  if (rc) return rc;
  // declared first, so that every region goes back to it before it frees
  // its blocks
  const std::shared_ptr<Cache> cache = std::make_shared<Cache>();
  std::vector<std::array<u64, 4>> nodes;
  std::vector<std::pair<u64, u64>> blocks;
  build_tree(lower, upper, nodes, blocks);
  // test hook: device work at this height reports a failure (height 0: the
  // leaves' folds get buffers too small for them; others: HGX_ENOMEM)
  const char *fail_env = std::getenv("HGX_SERIES_FAIL_HEIGHT");
  const long fail_height = fail_env ? std::atol(fail_env) : -1;
  // the live partials by their interval's lower end (children of a node
  // start at lo and mid, and no two live partials share a lower end),
  // first the leaves: folded on the device for a split whose heights go
  // there, else on the host in chunks -- one thread folds a chunk's leaves,
  // copies their partials into one region and frees its own fold vectors
  std::vector<std::pair<u64, Part>> live(blocks.size());
  for (size_t i = 0; i < blocks.size(); i++) live[i].first = blocks[i].first;
  bool leaves_dev = false;
  rc = leaves_on_device(s, blocks, lower, upper, fail_height == 0, live, leaves_dev);
  if (rc) return rc;
  const size_t nchunk = leaves_dev ? 0 : (blocks.size() + kLeafChunk - 1) / kLeafChunk;
  rc = parallel_for(nchunk, [&](size_t c) {
    const size_t i0 = c * kLeafChunk, i1 = std::min(blocks.size(), i0 + kLeafChunk);
    std::vector<AccPart> acc(i1 - i0);
    size_t total = 0;
    for (size_t i = i0; i < i1; i++) {
      AccPart &a = acc[i - i0];
      const int r = fold(s, blocks[i].first, blocks[i].second, a);
      if (r) return r;
      total += a.P.m.size() + a.Q.m.size() + a.B.m.size() + a.R.m.size();
    }
    const RegionRef reg = new_region(cache, total);
    size_t at = 0;
    for (size_t i = i0; i < i1; i++) {
      AccPart &a = acc[i - i0];
      const Acc *src[4] = {&a.P, &a.Q, &a.B, &a.R};
      Part &pt = live[i].second;
      Num *dst[4] = {&pt.P, &pt.Q, &pt.B, &pt.R};
      for (int k = 0; k < 4; k++) {
        Num &d = *dst[k];
        d.s = src[k]->s;
        d.n = d.s ? src[k]->m.size() : 0;
        d.cap = d.n;
        d.p = reg->base + at;
        d.own = reg;
        std::copy(src[k]->m.begin(), src[k]->m.begin() + d.n, d.p);
        at += d.n;
      }
    }
    return (int)HGX_OK;
  });
  if (rc) return rc;
  // build_tree lists the leaves left to right
  auto by_lower = [](const auto &x, const auto &y) { return x.first < y.first; };
  if (!std::is_sorted(live.begin(), live.end(), by_lower)) std::sort(live.begin(), live.end(), by_lower);
  auto find = [&](u64 lo) -> Part & {
    auto it = std::lower_bound(live.begin(), live.end(), lo, [](const auto &x, u64 v) { return x.first < v; });
    return it->second;
  };
  auto prod_limbs = [](const Num &x, const Num &y) -> size_t { return x.s && y.s ? x.n + y.n : 0; };
  bool on_dev = leaves_dev, dev_off = false;
  std::stable_sort(nodes.begin(), nodes.end(), [](const auto &x, const auto &y) { return x[0] < y[0]; });
  for (size_t h0 = 0; h0 < nodes.size();) {
    size_t h1 = h0;
    while (h1 < nodes.size() && nodes[h1][0] == nodes[h0][0]) h1++;
    const size_t cnt = h1 - h0;
    std::vector<Part *> L(cnt), R(cnt);
    std::vector<const Num *> xs, ys;
    std::vector<size_t> room;
    // round 1: P_l P_r, Q_l Q_r, B_l B_r, Q_r R_l, P_l R_r (combine, :112-121)
    auto round1 = [&]() {
      xs.clear();
      ys.clear();
      room.clear();
      xs.reserve(5 * cnt);
      ys.reserve(5 * cnt);
      room.reserve(5 * cnt);
      // the children by a cursor over the live partials: a height's nodes
      // come left to right (post order, then a stable sort by height), and
      // a node's right child directly follows its left one
      size_t cur = 0;
      for (size_t i = 0; i < cnt; i++) {
        const u64 lo = nodes[h0 + i][1], mid = nodes[h0 + i][2];
        while (cur < live.size() && live[cur].first < lo) cur++;
        if (cur + 1 < live.size() && live[cur].first == lo && live[cur + 1].first == mid) {
          L[i] = &live[cur].second;
          R[i] = &live[cur + 1].second;
        } else {
          L[i] = &find(lo);
          R[i] = &find(mid);
        }
        const Part &l = *L[i], &r = *R[i];
        for (auto pr : {std::make_pair(&l.P, &r.P), std::make_pair(&l.Q, &r.Q), std::make_pair(&l.B, &r.B),
                        std::make_pair(&r.Q, &l.R), std::make_pair(&l.P, &r.R)}) {
          xs.push_back(pr.first);
          ys.push_back(pr.second);
          room.push_back(prod_limbs(*pr.first, *pr.second));
        }
      }
    };
    // the height's products and sums; the nodes' partials replace the
    // children only once all of them succeed
    auto height = [&](bool dev) -> int {
      std::vector<Num> o1;
      int r = dev ? mul_many_dev(xs, ys, room, o1) : mul_many(cache, xs, ys, room, o1);
      if (r) return r;
      if (dev && (long)nodes[h0][0] == fail_height)
        return hgx_internal_fail(HGX_ENOMEM, "binary splitting: device memory exhausted (test hook)");
      // round 2: B_r (Q_r R_l), B_l (P_l R_r); the first slot of each node
      // also holds the sum R
      std::vector<const Num *> xs2, ys2;
      std::vector<size_t> room2;
      for (size_t i = 0; i < cnt; i++) {
        const Num &br = R[i]->B, &bl = L[i]->B;
        xs2.push_back(&br);
        ys2.push_back(&o1[5 * i + 3]);
        xs2.push_back(&bl);
        ys2.push_back(&o1[5 * i + 4]);
        const size_t a = prod_limbs(br, o1[5 * i + 3]), b = prod_limbs(bl, o1[5 * i + 4]);
        room2.push_back(std::max(a, b) + 1);
        room2.push_back(b);
      }
      std::vector<Num> o2;
      r = dev ? mul_many_dev(xs2, ys2, room2, o2) : mul_many(cache, xs2, ys2, room2, o2);
      if (r) return r;
      // R = B_r Q_r R_l + B_l P_l R_r; the node's partial replaces its left
      // child's slot (same lower end), the right child's slot is emptied
      auto finish = [&](size_t i) {
        Part &l = *L[i], &rt = *R[i];
        l.P = std::move(o1[5 * i]);
        l.Q = std::move(o1[5 * i + 1]);
        l.B = std::move(o1[5 * i + 2]);
        l.R = std::move(o2[2 * i]);
        rt = Part();
      };
      if (dev) {
        r = sums_dev(o2, cnt);
        if (r) return r;
        // one thread: the values of a round share their region's reference
        // count, which threads would contend for
        for (size_t i = 0; i < cnt; i++) finish(i);
        return HGX_OK;
      }
      // host sums on host threads: the sums first, the partials replace the
      // children once every sum has succeeded
      r = parallel_for(cnt, [&](size_t i) { return num_add(o2[2 * i], o2[2 * i + 1]); });
      if (r) return r;
      for (size_t i = 0; i < cnt; i++) finish(i);
      return HGX_OK;
    };
    round1();
    if (!on_dev && !dev_off) {
      on_dev = move_to_device(live, xs, ys, rc);
      if (rc) return rc;
    }
    rc = height(on_dev);
    if (on_dev && rc == HGX_ENOMEM) {
      // device memory ran out: the children are intact; every live value
      // goes back to host memory and the split goes on there
      rc = to_host(live, cache);
      if (rc) return rc;
      on_dev = false;
      dev_off = true;
      round1();  // the trimmed host values
      rc = height(false);
    }
    if (rc) return rc;
    // drop the right children's slots
    std::vector<u64> gone;
    gone.reserve(cnt);
    for (size_t i = 0; i < cnt; i++) gone.push_back(nodes[h0 + i][2]);
    std::sort(gone.begin(), gone.end());
    {
      // both sorted: one merge pass keeps the others in order
      size_t keep = 0, g = 0;
      for (size_t k = 0; k < live.size(); k++) {
        while (g < gone.size() && gone[g] < live[k].first) g++;
        if (g < gone.size() && gone[g] == live[k].first) continue;
        if (keep != k) live[keep] = std::move(live[k]);
        keep++;
      }
      live.erase(live.begin() + keep, live.end());
    }
    h0 = h1;
  }
  // the result gets a region of its own (the last height's regions also
  // hold that height's intermediate products), in device memory for a split
  // that ran there
  const Part &top = find(lower);
  std::unique_ptr<hgx_partial> res(new (std::nothrow) hgx_partial);
  if (!res) return hgx_internal_fail(HGX_ENOMEM, "out of host memory for the series result");
  const Num *src[4] = {&top.P, &top.Q, &top.B, &top.R};
  size_t total = 0;
  for (int k = 0; k < 4; k++) total += src[k]->n;
  if (on_dev) {
    std::shared_ptr<DevRegion> dreg;
    rc = new_dev_region(total, dreg);
    if (rc) return rc;
    size_t at = 0;
    for (int k = 0; k < 4; k++) {
      Num &d = res->v[k];
      d.n = src[k]->n;
      d.cap = d.n;
      d.p = dreg->base + at;
      d.dev = true;
      d.own = dreg;
      at += d.n;
      rc = hgx_internal_copy_device(d.p, src[k]->p, d.n);
      if (rc) return rc;
      // device products keep their operands' full lengths, so a top value
      // can carry many zero limbs (~10^5 at 2^28 bits): one device search
      // for the top non-zero limb trims it
      size_t sig = 0;
      if (d.n) rc = hgx_device_significant(d.p, d.n, &sig, nullptr);
      if (rc) return rc;
      d.n = sig;
      d.s = d.n ? src[k]->s : 0;
    }
    *out = res.release();
    return HGX_OK;
  }
  const RegionRef reg = new_region(nullptr, total);
  size_t at = 0;
  for (int k = 0; k < 4; k++) {
    Num &d = res->v[k];
    d.s = src[k]->s;
    d.n = src[k]->n;
    d.cap = d.n;
    d.p = reg->base + at;
    d.own = reg;
    at += d.n;
  }
  rc = parallel_for(4, [&](size_t k) {
    std::copy(src[k]->p, src[k]->p + src[k]->n, res->v[k].p);
    return (int)HGX_OK;
  });
  if (rc) return rc;
  *out = res.release();
  return HGX_OK;
}

int hgx_partial_get(const hgx_partial *x, int which, int *sign, size_t *nlimbs, uint64_t *limbs) {
  if (!x || which < 0 || which > 3) return hgx_internal_fail(HGX_EINVAL, "bad partial or component index");
  const Num &v = x->v[which];
  if (sign) *sign = v.s;
  if (nlimbs) *nlimbs = v.n;
  if (limbs && v.n) {
    if (v.dev) return hgx_download(limbs, v.p, v.n, nullptr);
    std::memcpy(limbs, v.p, 8 * v.n);
  }
  return HGX_OK;
}

int hgx_partial_get_device(const hgx_partial *x, int which, int *sign, size_t *nlimbs, uint64_t *d_limbs,
                           void *stream) {
  if (!x || which < 0 || which > 3) return hgx_internal_fail(HGX_EINVAL, "bad partial or component index");
  const Num &v = x->v[which];
  if (sign) *sign = v.s;
  if (nlimbs) *nlimbs = v.n;
  if (d_limbs && v.n) return v.dev ? hgx_internal_copy_device(d_limbs, v.p, v.n) : hgx_upload(d_limbs, v.p, v.n, stream);
  return HGX_OK;
}

void hgx_partial_free(hgx_partial *x) { delete x; }

}  // extern "C"
