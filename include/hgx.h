//
// hgx.h -- C ABI of the MI355X exact NTT big-integer multiply.
//
// This is the drop-in boundary for Hypergeometrica's NTT multiply.  The
// reference (Common Lisp) dispatches every product of two operands of more
// than *ntt-multiply-threshold* limbs through the generic MPZ-* method
// (src/multiply.lisp:12-49) to MPZ-SQUARE for (eq x y) and to MPZ-*/NTT
// otherwise.  An sb-alien shim redefines those two functions to call
// hgx_sqr / hgx_mul; no caller changes.
//
// Data layout is the reference's storage layout: little-endian uint64 limbs
// in caller-owned memory (src/vec.lisp:17-65, src/ram-vec.lisp:52-59).
// Every entry point returns HGX_OK (0) or a negative HGX_E* code; the
// message of the last failure on the calling thread is hgx_last_error().
// All entry points are thread-safe.
#ifndef HGX_H
#define HGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HGX_OK = 0,
  HGX_EINVAL = -1,  // bad argument (null pointer, zero length, bad index)
  HGX_ETOOBIG = -2,  // transform length above the supported maximum
  HGX_ENOMEM = -3,  // device memory or the HBM budget exhausted
  HGX_EDEVICE = -4,  // HIP runtime or kernel failure
  HGX_ENOTINIT = -5  // hgx_init has not succeeded
};

// One multiplication: OUT receives NA + NB limbs of A * B.  A request with
// A == B and NA == NB is computed as a square (as src/multiply.lisp:43-44
// routes (eq x y)).
typedef struct {
  const uint64_t *a;
  size_t na;
  const uint64_t *b;
  size_t nb;
  uint64_t *out;
} hgx_mul_desc;

typedef uint64_t hgx_ticket;

// Library set-up: select DEVICE, build the resident prime / root tables
// (the scheme of src/moduli.lisp:298-353, src/ntt-multiply.lisp:8) and cap
// device work memory at HBM_BUDGET_BYTES (0 = 90% of free memory).
// Replaces the load-time construction of **scheme**.
int hgx_init(int device, size_t hbm_budget_bytes);
void hgx_shutdown(void);
const char *hgx_last_error(void);

// Scheme introspection: the moduli of default-moduli, src/moduli.lisp:257-264
// (descending), and the transform length chosen for a product,
// least-power-of-two->= (na + nb), src/ntt-multiply.lisp:176.
int hgx_num_moduli(void);
uint64_t hgx_modulus(int k);
size_t hgx_product_length(size_t na, size_t nb);
int hgx_max_transform_bits(void);

// Exact product of two non-negative integers in host memory.
// Replaces MPZ-*/NTT, src/ntt-multiply.lisp:174-254 (magnitude part; the
// sign is sign(x)*sign(y), applied by the caller as at :254).  Writes exactly
// NA + NB limbs to OUT (the reference's L-limb result trimmed by
// optimize-mpz, src/multiply.lisp:19, holds the same value).  OUT is fresh
// storage, as the reference's result is (:184): it must not overlap A or B
// (HGX_EINVAL otherwise).
int hgx_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
            uint64_t *out);

// Exact square.  Replaces MPZ-SQUARE, src/ntt-multiply.lisp:101-172.
// Writes exactly 2 * NA limbs to fresh storage OUT (:108), which must not
// overlap A.
int hgx_sqr(const uint64_t *a, size_t na, uint64_t *out);

// Many independent products in one call, grouped on the device by
// transform length (the uneven multiplies of one binary-splitting level,
// COMBINE at src/binary-splitting.lisp:112-121).  Equivalent to COUNT calls
// of hgx_mul / hgx_sqr with identical output.
int hgx_mul_batch(const hgx_mul_desc *descs, size_t count);

// Same products, with every pointer in device memory; runs on STREAM
// (a hipStream_t, or NULL for the library's stream) and returns once the
// work is queued.  Inputs resident in HBM; used by benchmarks and by
// device-resident callers.
int hgx_mul_device(const uint64_t *d_a, size_t na, const uint64_t *d_b,
                   size_t nb, uint64_t *d_out, void *stream);
int hgx_sqr_device(const uint64_t *d_a, size_t na, uint64_t *d_out,
                   void *stream);
int hgx_mul_batch_device(const hgx_mul_desc *descs, size_t count,
                         void *stream);

// hgx_mul_device without any allocation: no hipMalloc, hipFree or other
// allocating call, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).  No counterpart in
// the reference.  The work memory is a scratch array of 48 MiB that the
// library holds on every device it runs on from hgx_init until the process
// ends; products of transform length at most 2^20 (NA + NB <= 2^20 limbs)
// fit it, larger ones fail with HGX_ETOOBIG.  The root tables of the
// product's transform length must already be resident (any earlier
// hgx_mul_device product of that length builds them), else HGX_ENOTINIT.
// Never split (one transform of length >= NA + NB).  Every call on a device
// shares that device's scratch array, which keeps the last call's contents
// (nothing in it needs to start at zero): calls that overlap on one device --
// on different streams, or from different threads -- race; calls ordered on
// one stream do not.  Same output as hgx_mul_device.
int hgx_mul_device_noalloc(const uint64_t *d_a, size_t na, const uint64_t *d_b,
                           size_t nb, uint64_t *d_out, void *stream);

// Forward transform of N canonical residues (each < modulus MOD_NUM) in
// place, output in bit-reversed order.  Same result as NTT-FORWARD,
// src/number-theoretic-transform.lisp:40-90, with **scheme** and MOD_NUM.
int hgx_ntt_forward(uint64_t *data, size_t n, int mod_num);

// Inverse transform (input in bit-reversed order, output natural order,
// scaled by 1/N).  Same result as NTT-REVERSE,
// src/number-theoretic-transform.lisp:93-143.
int hgx_ntt_reverse(uint64_t *data, size_t n, int mod_num);

// Continuous-batching request server (no reference counterpart: the
// reference calls MPZ-* synchronously).  hgx_submit queues one product and
// returns at once; requests that arrive while others run are admitted as
// device memory (the HBM budget) frees up, and their work overlaps on the
// device.  Host buffers must stay valid until the ticket completes.
// hgx_poll returns 1 when done, 0 while pending, <0 on failure;
// hgx_wait blocks and returns HGX_OK or the request's error.  A ticket is
// released by hgx_wait or by the hgx_poll call that reports its completion
// (1 or an error); later calls with it return HGX_EINVAL.  hgx_shutdown lets
// the work already on the device finish and fails the requests that have not
// started with HGX_ENOTINIT.  A failed upload, download or kernel launch
// fails only the requests it concerns.
int hgx_submit(const hgx_mul_desc *desc, hgx_ticket *ticket);
int hgx_poll(hgx_ticket ticket);
int hgx_wait(hgx_ticket ticket);

// Counters of the request server since it started (all zero before the
// first hgx_submit): requests completed and failed, admission rounds
// launched, requests computed host-staged (larger than the server's arena),
// requests queued and on the device now, seconds since start, seconds the
// server thread spent issuing rounds (copies and kernel launches) -- the
// host-side cost per request that a request loop resident on the device
// would remove --, observing completions and waiting for work, and its
// streams: compute lanes, whether each lane has a
// second stream for downloads, and the upload and download streams shared by
// the lanes (0 when every copy runs on its round's lane).  No counterpart in
// the reference.
typedef struct {
  uint64_t completed, failed, rounds, staged, queued, in_flight;
  double uptime_s, launch_s, retire_s, wait_s;
  int lanes, split_downloads, uploads, downloads;
} hgx_server_counters;

int hgx_server_stats(hgx_server_counters *out);

// Magnitude sum and difference of device limb vectors, the additions
// between the products of a Newton iteration kept in device memory
// (mpz-+ and mpz-- on the magnitudes of mpz/ram, src/mpz-ram.lisp:185-291).
// N = max(NA, NB) > 0; the shorter operand reads as zero-extended.
// hgx_add_device writes N + 1 limbs (the last is the carry out);
// hgx_sub_device writes N limbs of A - B and requires A >= B.  D_OUT may be
// D_A or D_B.  Ordered on STREAM (null: the library's stream) with the
// library's other calls.
int hgx_add_device(const uint64_t *d_a, size_t na, const uint64_t *d_b,
                   size_t nb, uint64_t *d_out, void *stream);
int hgx_sub_device(const uint64_t *d_a, size_t na, const uint64_t *d_b,
                   size_t nb, uint64_t *d_out, void *stream);

// Device limb vectors for hosts without a device allocator of their own (an
// SBCL host through sb-alien): enough to keep the mantissas of a Newton
// iteration (mpd-reciprocal, src/mpd-reciprocal.lisp:27-73; mpd-inv-sqrt,
// src/mpd-sqrt.lisp:40-87) in HBM between the products and sums above, as
// the reference keeps them in its storage vectors (make-storage,
// src/mpz-ram.lisp:12-17).  hgx_device_alloc hands out NLIMBS limbs of device
// memory (counted against the HBM budget; the contents are undefined);
// hgx_device_free releases a block after the work queued on STREAM before it
// (the library keeps released blocks for reuse, and returns them to the
// device when the budget needs them).  hgx_shutdown frees every block.
// hgx_upload / hgx_download copy limbs between host memory (pageable or
// pinned) and device memory, ordered on STREAM, and return when the copy is
// complete.  hgx_device_copy (non-overlapping) and hgx_device_fill are
// queued on STREAM.  hgx_device_significant writes the limb count up to the
// top non-zero limb (mpz-size after optimize-storage, src/mpz-ram.lisp:77-92)
// and hgx_device_compare the sign (-1, 0, 1) of |A| - |B| (both read as
// zero-extended), each after the work queued before it.  hgx_shl_device
// writes N + 1 limbs of A * 2^BITS (0 <= BITS < 64; the output must not
// overlap A): mpd-half! (src/mpd.lisp:77-81) is the shift by 63 with the
// exponent lowered by one.  hgx_host_alloc / hgx_host_free give pinned host
// memory, which the DMA engines read and write directly.  STREAM null: the
// library's stream.  No counterpart in the reference beyond the storage
// layout.
int hgx_device_alloc(size_t nlimbs, uint64_t **d_out);
int hgx_device_free(uint64_t *d, void *stream);
int hgx_upload(uint64_t *d_dst, const uint64_t *src, size_t nlimbs,
               void *stream);
int hgx_download(uint64_t *dst, const uint64_t *d_src, size_t nlimbs,
                 void *stream);
int hgx_device_copy(uint64_t *d_dst, const uint64_t *d_src, size_t nlimbs,
                    void *stream);
int hgx_device_fill(uint64_t *d, size_t nlimbs, uint64_t value, void *stream);
int hgx_device_significant(const uint64_t *d, size_t nlimbs, size_t *out,
                           void *stream);
int hgx_device_compare(const uint64_t *d_a, size_t na, const uint64_t *d_b,
                       size_t nb, int *out, void *stream);
int hgx_shl_device(const uint64_t *d_a, size_t n, int bits, uint64_t *d_out,
                   void *stream);
int hgx_host_alloc(size_t bytes, void **p);
int hgx_host_free(void *p);
int hgx_synchronize(void *stream);

// Streams for such hosts (default priority, non-blocking): work on different
// streams may overlap on the device, and copies of pinned host memory on
// them overlap each other.  hgx_stream_destroy waits for the stream first.
int hgx_stream_create(void **stream);
int hgx_stream_destroy(void *stream);

// Device SplitMix64 fill for benchmarks and tests:
// d_out[i] = mix(seed + (i + 1) * 0x9E3779B97F4A7C15); when FORCE_TOP is
// non-zero the last limb is forced non-zero.
int hgx_fill_splitmix64(uint64_t *d_out, size_t n, uint64_t seed,
                        int force_top, void *stream);

// Host-staged products, for operands and products beyond the device memory:
// a product from host memory (hgx_mul, hgx_sqr, hgx_mul_batch, hgx_submit)
// whose device working set -- operands, product, one transform's work
// arrays and root tables -- exceeds the memory available to it is computed
// from chunks of the operands, diagonal by diagonal of the chunk products,
// with two chunks, one chunk product and a window of the result in HBM at a
// time.  Same exact product.  The successor of the reference's disk-backed
// vectors for numbers beyond memory (make-storage, src/mpz-ram.lisp:12-17;
// disk-vec, src/disk-vec.lisp:78-91).  BYTES caps the device memory of one
// such product below the HBM budget; 0 (the default) leaves the budget as
// the only cap.
int hgx_staging_limit(size_t bytes);

// Unbalanced products: when one operand is much longer than the other, a
// product (single, batched or served) is computed as products of chunks of
// the longer operand with the shorter one, all of one shorter transform
// length, batched, and an addition of their sums -- the same integer for a
// fraction of the work of one transform of length >= NA + NB.  No counterpart
// in the reference (its mpz-*/ntt always transforms at that length).  ON = 0
// turns it off.  Default: on.
int hgx_split_enable(int on);

// Batched calls (hgx_mul_batch, hgx_mul_batch_device) capture the launch
// sequence of each chunk in a HIP graph, keyed by its shape (the transform
// lengths and counts of its groups) and by the library's job-table and
// work-memory addresses, and replay it when the same shape comes again.
// No counterpart in the reference (batching infrastructure).  ON = 0 launches
// every kernel directly; the outputs are the same.  Default: on.
int hgx_graphs_enable(int on);

// Pass structure per transform length 2^BITS: NCOL column passes of 8 bits
// and a row pass of BITS - 8 NCOL bits (4..12); the largest column-pass
// block (2^DIRECT_MAX_BITS words, 12..24) that uses the direct twiddle table;
// and MONT_DIRECT (0/1): whether larger blocks of at most 2^28 words use
// direct tables of Montgomery-form roots (8 bytes per word of the block, per
// prime and direction, built on the device) instead of the tile-factored
// pair.  No counterpart in the reference (its transforms are one radix-2
// loop).  Every choice gives the same results; the default is the fewest
// passes, DIRECT_MAX_BITS = 20 and MONT_DIRECT as built.
int hgx_plan_set(int bits, int ncol, int direct_max_bits, int mont_direct);
int hgx_plan_get(int bits, int *ncol, int *direct_max_bits, int *mont_direct);

// Kernels per transform length 2^BITS, on whatever pass structure it has:
// COL_THREADS 256 or 512 threads per column tile (512: 8 rows per thread and
// two LDS exchanges, blocks of at most 2^28 words; 256 otherwise) and
// ROW_TILE_BITS 12, 11 or 10: the multiply row pass on tiles of 2^12 words
// (radix-16 rounds) or of 2^11 / 2^10 words (radix-8 / radix-4 rounds; only
// for rows that fit the tile).  0 for either: the default rule, which takes
// the 512-thread column tile and the smaller row tiles for launches too small
// to fill the GPU.  No counterpart in the reference.  Same results for every
// choice.
int hgx_kernel_set(int bits, int col_threads, int row_tile_bits);
int hgx_kernel_get(int bits, int *col_threads, int *row_tile_bits);

// Autotuner: for every length 2^bits with MIN_BITS <= bits <= MAX_BITS that
// offers a choice, time a product of two 2^(bits-1)-limb operands under each
// pass structure, then (from 2^20 on, where one product fills the GPU) under
// each kernel choice on the fastest structure (best of REPS runs after a
// warm-up), and keep the fastest.
// With HGX_TUNE_CACHE set in the environment the choices are saved to that
// file and read back by later hgx_init calls on the same GPU architecture.
int hgx_autotune(int min_bits, int max_bits, int reps);

// Per-kernel device timing of later calls, for benchmarks: HIP events are
// recorded around every kernel on the stream it is launched on.  Plays the
// role of the reference's *verbose* phase timers ("Forward...", "Reverse...",
// "CRT...", src/ntt-multiply.lisp:185-252).  hgx_timing_enable(1) clears the
// totals and starts recording; hgx_timing_read waits for the recorded work
// and returns the number of kernel kinds written to OUT.
typedef struct {
  char name[32];
  uint64_t launches;
  double total_ms;
} hgx_kernel_time;

int hgx_timing_enable(int on);
int hgx_timing_read(hgx_kernel_time *out, int max_entries);

// Throughput of the library's own modular arithmetic on this GPU, the
// integer-ALU ceiling the transforms are measured against: OP 0 Shoup
// product by a fixed factor (the reference's m*/fast2,
// src/modular-arithmetic.lisp:206-231), 1 Montgomery product, 2 modular sum
// and difference (m+ / m-, :25-48), 3 forward radix-2 butterfly
// (src/number-theoretic-transform.lisp:66-81), 4 inverse butterfly
// (:126-141); 5 and 6 move 16 words per thread between the lanes of groups
// of 8, as a row pass's rounds hand values on: 5 through LDS with a barrier,
// 6 inside the wave (ds_swizzle xor stages and VALU selects), counted in
// words moved.  Register-only, 8 independent chains per thread, 8
// workgroups of 256 threads per CU, ITERS steps per chain.  OUT[0] =
// operations per second, OUT[1] = shader clock in GHz under that load (cycle
// counter over the 100 MHz real-time counter), OUT[2] = seconds.
int hgx_modmul_probe(int op, int iters, double *out);

// Binary splitting of a hypergeometric series
//     sum over k >= 0 of a(k)/b(k) * p(0)...p(k) / (q(0)...q(k))
// (the series struct, src/binary-splitting.lisp:19-34) over [LOWER, UPPER):
// the partial P, Q, B, R of binary-split (:123-133), built from COMBINE
// (:112-121) and the base case (:100-110), evaluated one tree height at a
// time.  Each height's products whose smaller operand has more than 100
// limbs (*ntt-multiply-threshold*, src/multiply.lisp:9) go to the GPU in one
// hgx_mul_batch call; smaller ones are multiplied on the host, as mpz-*
// does (src/multiply.lisp:37-42).  A split whose heights' products hold at
// least 2^20 limbs instead keeps its values in device memory throughout (its
// subtrees of at most 32 terms folded by one thread each on the device,
// every product of a round in one hgx_mul_batch_device call, the sums R on
// the device), with the same result; HGX_SERIES_DEVICE=0 in the
// environment keeps every split on the host, =1 moves every split.  An
// interval whose products are all small needs no hgx_init.  Replaces a
// Lisp-side binary-split whose term functions are polynomials: a(k), b(k),
// p(k), q(k) are polynomials in k (index 0 a,
// 1 b, 2 p, 3 q; coefficients constant term first, degree 0..7) or, where
// AT_ZERO[i] is set, VALUE0[i] at k = 0 (the reference's series give 1 there:
// make-exp-series, make-e-series, make-ramanujan-series, :140-184;
// make-chudnovsky-series, src/pi.lisp:24-43; make-catalan-series,
// make-apery-series, :203-231).  UPPER <= 2^40; term values up to 1024 bits.
// *OUT receives a result that the caller reads with hgx_partial_get (WHICH 0
// P, 1 Q, 2 B, 3 R: SIGN in {-1, 0, 1}, NLIMBS of the trimmed magnitude,
// copied to LIMBS when it is not null) and releases with hgx_partial_free.
// A split that ran in device memory keeps its result there until
// hgx_partial_free (so that result needs the library initialised until
// then): hgx_partial_get downloads it into LIMBS, and hgx_partial_get_device
// copies it device to device into D_LIMBS (device memory of at least
// *NLIMBS limbs), uploading instead a result held in host memory; both
// return when the limbs are in place.  hgx_partial_get_device has no
// counterpart in the reference: it serves device-resident callers such as
// the Newton steps of mpd-pi (src/pi.lisp:53-74), which take B, Q and R
// without a round trip through host memory.
typedef struct {
  int degree[4];
  int64_t coef[4][8];
  int at_zero[4];
  int64_t value0[4];
} hgx_series;

typedef struct hgx_partial hgx_partial;

int hgx_series_split(const hgx_series *series, uint64_t lower, uint64_t upper,
                     hgx_partial **out);
int hgx_partial_get(const hgx_partial *x, int which, int *sign, size_t *nlimbs,
                    uint64_t *limbs);
int hgx_partial_get_device(const hgx_partial *x, int which, int *sign,
                           size_t *nlimbs, uint64_t *d_limbs, void *stream);
void hgx_partial_free(hgx_partial *x);

#ifdef __cplusplus
}
#endif
#endif

