/*
 * fastfib_amd: exact big-integer multiplication and F(n) on AMD Instinct
 * MI355X (gfx950) - the C-ABI of libfastfib_amd.so.
 *
 * The first group keeps the reference's C surface (shwwwa/fastfib, a unity
 * build whose functions are all static: common.h:8-11, fastfib.c:16-21), so
 * a caller of big_mul / big_sqr / big_mul_ntt / fibonacci can switch by
 * name: same Big layout, same squaring signal (a == b pointer identity,
 * mul_basic.c:177, mul_ntt.c:36,125), same ownership (results are new and
 * freed with ff_big_free), and the same failure behaviour (a message on
 * stderr and exit(1), as alloc.c:142-146). All arithmetic runs on the GPU.
 *
 * The second group is new: device-resident numbers, batched products and
 * status codes, for callers that keep operands in HBM between products.
 */
#ifndef FASTFIB_AMD_H
#define FASTFIB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define FF_API __attribute__((visibility("default")))
#else
#define FF_API
#endif

/* value = sum d[i] * 2^(64 i); normalized: n >= 1, no leading zero limb,
   zero is {0} with n = 1. Layout-identical to the reference's Big
   (src/common.h:29-33). A caller that already has that Big in scope (the
   reference's unity build, which typedefs it as an anonymous struct that C
   cannot redeclare) defines FF_CALLER_BIG before including this header, and
   every declaration below then takes the caller's own Big. */
#ifndef FF_CALLER_BIG
typedef struct {
    uint64_t *d;
    size_t n;
} Big;
#endif

/* ------------------------------------------------------------------------
 * Drop-in surface. Host pointers in and out; failures abort the process.
 * --------------------------------------------------------------------- */

/* Replaces ntt_init (src/ntt_tables.c:72): prepares the device and the
   root tables for products of up to max_limbs limbs. Idempotent and
   grow-only; every other entry point also initializes on first use. */
FF_API void ff_ntt_init(size_t max_limbs);

/* Replaces big_mul (src/mul_basic.c:64): a * b as a new Big; a == b squares. */
FF_API Big ff_big_mul(const Big *a, const Big *b);

/* Replaces big_sqr (src/mul_basic.c:109): a * a as a new Big. */
FF_API Big ff_big_sqr(const Big *a);

/* Replaces big_mul_ntt (src/mul_ntt.c:97): a * b by the three-prime NTT
   with the reference's width and length rule (mul_ntt.c:103-118); a == b
   squares with one forward transform. */
FF_API Big ff_big_mul_ntt(const Big *a, const Big *b);

/* Replaces fibonacci (src/fib.c:211): F(n) as a new Big, by fast doubling
   with every product and linear step on the GPU. */
FF_API Big ff_fibonacci(uint64_t n);

/* Replaces big_to_string (src/decimal.c:229): the decimal digits of a,
   without leading zeros ("0" for zero), as a malloc'ed NUL-terminated
   string the caller frees with free(). Divide and conquer by powers
   10^(19 * 2^i) with Barrett division, every level on the GPU. */
FF_API char *ff_big_to_string(const Big *a);

/* Replaces big_recip (src/decimal.c:39): floor(2^(2s) / P) minus at most
   4, for P of exactly s bits, by the reference's Newton recursion with
   every product on the GPU (exported for its parity tests). */
FF_API Big ff_big_recip(const Big *P, size_t s);

/* The schedule the product of a bits_a-bit and a bits_b-bit number gets
   (square: one operand), as big_mul_ntt chooses it (src/mul_ntt.c:99-118):
   coefficient width 32..35 while the CRT bound holds, the shortest length
   2^k or 3 * 2^k (at least 16) not below conv = ca + cb - 1 coefficients,
   the narrowest width among equal lengths; past 2^max_lg the split multiply
   of src/mul_split.c:146-177 (widest width the bound allows, pieces of
   length/2 coefficients transformed at `length` = 2^lg by the reference's
   cost model). A zero operand gives zero = 1 and no transform. Host-only:
   touches no device. */
typedef struct {
    unsigned width;
    uint64_t length;
    uint64_t conv;
    int radix3;
    int split;
    uint64_t pieces_a, pieces_b;
    int zero;
} ff_plan_info;
FF_API int ff_product_plan(uint64_t bits_a, uint64_t bits_b, int square, ff_plan_info *out);

/* The exact reciprocal floor(2^(2 s) / P_i) of P_i = 10^(19 * 2^i) (s its
   bit length), i in 5..40, that ff_big_to_string's Barrett levels use
   (where the reference calls big_recip, decimal.c:252-259, and tolerates
   its 4-low window): *n gets its limb count; limbs, when not NULL, gets the
   limbs (room for cap). No counterpart in the reference; for tests. */
FF_API int ff_dec_reciprocal(int i, uint64_t *limbs, size_t cap, size_t *n);

/* Replaces big_free (src/bigint.c:26): frees x's limbs, sets d = NULL, n = 0. */
FF_API void ff_big_free(Big *x);

/* Replaces fib_step_linear (src/fib.c:81): from s1 = F(k)^2 and
   s2 = F(k-1)^2 builds (lo, hi) = bit ? (F(2k), F(2k+1)) : (F(2k-1), F(2k)),
   i.e. A = s1 + s2, B = 4 s1 - s2 + 2(-1)^k (k_odd gives the sign),
   C = B - A, keeping (C, B) or (A, C). Requires s2 <= s1. */
FF_API void ff_fib_step_linear(const Big *s1, const Big *s2, int k_odd, int bit, Big *lo, Big *hi);

/* Replaces last_factors (src/fib.c:147): x = 2u + v and, if y is not NULL,
   y = 2u - v (requires 2u >= v). */
FF_API void ff_last_factors(const Big *u, const Big *v, Big *x, Big *y);

/* ------------------------------------------------------------------------
 * Device-resident API (no counterpart in the reference). Returns FF_OK or a
 * negative status; ff_last_error() describes the last failure.
 * --------------------------------------------------------------------- */

#define FF_OK 0
#define FF_ERR_DEVICE (-1)      /* a HIP call failed */
#define FF_ERR_NOMEM (-2)       /* device memory exhausted */
#define FF_ERR_ARG (-3)         /* invalid argument */
#define FF_ERR_UNSUPPORTED (-4) /* size beyond what this build handles */

/* A number in device memory: d has room for cap limbs, n is its
   normalized length and bits its bit length (0 for zero). */
typedef struct {
    uint64_t *d;
    size_t n;
    size_t cap;
    uint64_t bits;
} ff_dbig;

/* Selects the device (-1: the current one), creates the library's stream
   and uploads the root tables. */
FF_API int ff_init(int device);

/* The HIP stream (hipStream_t) all device work of the library is queued on. */
FF_API void *ff_stream(void);

/* Waits for all queued device work. */
FF_API int ff_sync(void);

FF_API const char *ff_last_error(void);

FF_API int ff_dbig_alloc(ff_dbig *x, size_t cap);
FF_API void ff_dbig_release(ff_dbig *x);
/* Copies n host limbs in (normalizes n and computes bits), into x's own
   memory when x is allocated with room for them, else into a new block. */
FF_API int ff_dbig_upload(ff_dbig *x, const uint64_t *limbs, size_t n);
/* Copies x's n limbs out to host memory with room for cap limbs. */
FF_API int ff_dbig_download(const ff_dbig *x, uint64_t *limbs, size_t cap);

/* out = a * b on the device (b == a or b == NULL squares); out is
   (re)allocated as needed. Same schedule rule as ff_big_mul_ntt. */
FF_API int ff_dmul(ff_dbig *out, const ff_dbig *a, const ff_dbig *b);

/* No counterpart in the reference: the product for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc). out = a * b as ff_dmul computes
   it (b == a or b == NULL squares; out may be a or b), but out->d is the
   caller's (out->cap >= a->n + b->n + 1 limbs, else FF_ERR_ARG) and every
   scratch buffer lives in one fixed device array of
   FF_NOALLOC_SCRATCH_BYTES that each GPU holds from ff_init until the
   process ends; nothing is allocated or freed. Products whose transform is
   longer than FF_NOALLOC_MAX_LEN words (operands of up to about 2^24 bits
   each) return FF_ERR_UNSUPPORTED with nothing launched; the reference rule
   picks the schedule. The library must be initialized (ff_init) first,
   else FF_ERR_ARG. Waits for the product, then sets out->n and out->bits.
   Every call on a GPU shares that GPU's array: calls that overlapped on one
   GPU would race; the library's lock and its single stream keep its own
   calls from overlapping. */
#define FF_NOALLOC_MAX_LEN (1u << 20)
#define FF_NOALLOC_SCRATCH_BYTES (256u + 4u * 16384u + 2u * 4u * 3u * FF_NOALLOC_MAX_LEN + 256u)
FF_API int ff_dmul_noalloc(ff_dbig *out, const ff_dbig *a, const ff_dbig *b);

/* out = F(n), computed and left in device memory. */
FF_API int ff_dfibonacci(uint64_t n, ff_dbig *out);

/* The decimal digits of a device number (as ff_big_to_string): *out gets
   a malloc'ed string, *len its length. */
FF_API int ff_dbig_to_string(const ff_dbig *a, char **out, size_t *len);

/* The same conversion with the digits left in device memory: *dev_digits
   points at the first of *len digits (no terminator), in a buffer the
   library owns and reuses at the next conversion. */
FF_API int ff_dbig_to_decimal(const ff_dbig *a, const char **dev_digits, size_t *len);

/* ------------------------------------------------------------------------
 * Batched products (no counterpart in the reference, whose callers loop
 * over big_mul). A batch is built once for a fixed list of device operands
 * and can be run many times.
 * --------------------------------------------------------------------- */

/* One product of a batch. Pointers are device memory; a and b hold na and
   nb normalized limbs; out has room for na + nb + 1 limbs. b == NULL or
   b == a squares. bits_a / bits_b may be 0 (the library reads them). nout
   receives the product's normalized length (ff_batch_finish). Each run
   re-reads the operands: their values may change between runs as long as
   na, nb stay and the bit lengths do not exceed bits_a, bits_b. */
typedef struct {
    const uint64_t *a;
    size_t na;
    const uint64_t *b;
    size_t nb;
    uint64_t *out;
    size_t nout;
    uint64_t bits_a, bits_b;
} ff_prod;

/* How a batch runs (flags for ff_batch_create). */
#define FF_BATCH_PLAIN 0   /* one product at a time, one launch per pass each */
#define FF_BATCH_GROUPED 1 /* one launch per pass kind for the whole batch (work lists) */
#define FF_BATCH_GRAPH 2   /* grouped, captured once in a HIP graph and replayed */
#define FF_BATCH_TUNED 4   /* schedules from the autotuner table (ff_tune_*) */
#define FF_BATCH_LEARNED 8 /* schedules from the learned dispatcher (ff_mlp_*) */
#define FF_BATCH_QUEUE 16  /* one persistent launch per run: workgroups claim the batch's pass blocks from a
                              task table in dependency order (with FF_BATCH_GRAPH: that launch captured) */

typedef struct ff_batch ff_batch;

FF_API int ff_batch_create(ff_batch **out, const ff_prod *items, size_t count, int flags);
/* Queues one run of the whole batch on the library stream. */
FF_API int ff_batch_run(ff_batch *b);
/* Waits for the last run and writes every item's nout. */
FF_API int ff_batch_finish(ff_batch *b, ff_prod *items);
FF_API void ff_batch_destroy(ff_batch *b);

/* ------------------------------------------------------------------------
 * Schedule autotuner and learned dispatcher (no counterpart in the
 * reference, which always takes the shortest length and fixed block sizes:
 * mul_ntt.c:103-118, ntt_conv.c:34-62, and its cost model's piece length
 * for split products, mul_split.c:160-170).
 *
 * Candidate k (0 .. ff_tune_candidates()-1) is a schedule: for k < 16 the
 * transform length (the shortest, or the next supported one up), the
 * LDS-resident inner block (log2) and the outer-pass tile (LDS words); for
 * k >= 16, products past the longest transform only, the split multiply's
 * piece transform 2^(max_lg - 4 + k - 16) (ff_tune_candidate reports it as
 * a negative inner_log: -lg). A batch created
 * with FF_BATCH_TUNED takes each product's candidate from the autotuner
 * table, with FF_BATCH_LEARNED from the MLP; when neither was set, both are
 * read from dispatch.bin beside the library on first use. Without them a
 * batch uses the reference rule (ff_tune_default_candidate()).
 * --------------------------------------------------------------------- */

FF_API int ff_tune_candidates(void);
FF_API int ff_tune_default_candidate(void);
FF_API int ff_tune_candidate(int k, int *len_up, int *inner_log, int *tile_words);

/* The dispatcher's 8 features of a product in a batch of count products
   (log2 bits of a and b, square, log2 count, free HBM fraction, log2 and
   radix-3 flag of the shortest length, width - 32) and the mask of
   candidates that apply to it. */
FF_API int ff_tune_features(uint64_t bits_a, uint64_t bits_b, int square, size_t count, float *x, uint32_t *mask);

/* Times every candidate for one product shape: per-product device
   microseconds of a grouped batch of `copies` such products with random
   operands, averaged over at least reps runs (more, up to 32, while the
   runs take under a millisecond) after one warm run (NaN: candidate does
   not apply). us has room for ff_tune_candidates() values. */
FF_API int ff_tune_measure(uint64_t bits_a, uint64_t bits_b, int square, size_t copies, int reps, uint64_t seed,
                           float *us);

/* The autotuner table: one candidate per size class; ff_tune_class gives a
   product's class (index into the table of ff_tune_table_size() entries). */
FF_API int ff_tune_table_size(void);
/* The candidate the reference rule amounts to for a product: the default
   single-transform schedule, or for a product past the longest transform
   the piece length of the reference's cost model (mul_split.c:160-170). */
FF_API int ff_tune_rule_candidate(uint64_t bits_a, uint64_t bits_b, int square);
FF_API int ff_tune_class(uint64_t bits_a, uint64_t bits_b, int square, size_t count);
FF_API int ff_tune_set_table(const int32_t *table, size_t n);

/* Loads a dispatcher file (table and MLP parameters). */
FF_API int ff_dispatch_load(const char *path);

/* The MLP 8 -> 64 -> 64 -> K (FP32, ReLU): parameters in the order
   mu[8] sd[8] (input normalization) W1[64][8] b1[64] W2[64][64] b2[64]
   W3[K][64] b3[K]; outputs are predicted relative log-times. */
FF_API size_t ff_mlp_param_count(int K);
FF_API int ff_mlp_set(int K, const float *params, size_t count);
FF_API int ff_mlp_get(int *K, float *params, size_t count);
/* Mean-squared-error loss over n rows (x: n x 8, t: n x K) and its
   gradient with respect to every parameter (mu, sd get 0), on the GPU. */
FF_API int ff_mlp_grad(const float *x, const float *t, size_t n, float *loss, float *grad);
/* Adam training on the GPU from the current parameters: epochs over the n
   rows in shuffled minibatches of `batch`; loss_out gets the last epoch's
   mean loss. */
FF_API int ff_mlp_train(const float *x, const float *t, size_t n, int epochs, int batch, float lr, uint64_t seed,
                        float *loss_out);
/* Inference on the GPU: outputs (y, n x K) and argmin over mask bits
   (choice; mask NULL allows all), either may be NULL. */
FF_API int ff_mlp_predict(const float *x, size_t n, const uint32_t *mask, float *y, int *choice);

/* Schedule knobs (the reference's compile-time tunables, Appendix of its
   README: -DNAME=value, here set at run time). Names:
     "inner_log"   log2 of the LDS-resident inner block (4..12, default 12)
     "tile_words"  LDS words of an outer-pass tile (1024..16384, default 4096)
     "force_len"   transform length to use instead of the shortest (0: off)
     "max_lg"      longest single transform 2^max_lg (12..25, default 25);
                   longer products split into transformed-once pieces (the
                   reference's table bound NTT_MAX_LG, ntt_tables.c)
     "quad_max_lg" F(n): two bits of n per step (F(4k + d) from F(k),
                   F(k-1) through weighted pointwise products of one
                   transform of each) while that step's transform is at
                   most 2^quad_max_lg words (0: off, else 13..25, default
                   18)
     "pair_steps"  F(n), bit 0: doubling steps the same way (one transform
                   of F(k) and of F(k-1), the new pair combined pointwise)
                   instead of two squarings and the linear step after the
                   CRT (fib.c:36-62); bit 1: the last product likewise
                   (F(2k) = a^2 + 2ab, F(2k+1) = 2a^2 + 2ab + b^2 formed in
                   the product's middle round, no factor pass). Default 2:
                   the doubling steps measure faster as squarings
     "quad_inner_log" inner block (log2, 8..12) of quadrupling transforms
                   of at most 2^17 words (larger ones use "inner_log");
                   0: "inner_log" for all. Default 10: a quadrupling's
                   inner workgroup transforms both operands and two
                   outputs, and those short steps have few blocks, so
                   smaller blocks spread that serial work over more
                   workgroups
     "host_bits"   F(n): the top bits of n (10..13, default 13) done on the
                   host (F(k0), k0 < 2^host_bits, at most 89 limbs) before
                   the device chain; each bit replaces one one-workgroup
                   step (about 10 us) by microseconds of host squaring
     "dispatch"    schedules of single products and F(n): 0 the reference
                   rule (default), FF_BATCH_TUNED or FF_BATCH_LEARNED
     "school_limbs" products whose operands are both shorter than this many
                   limbs (0..200) are multiplied by the schoolbook method on
                   the GPU (one wavefront or workgroup per product, no
                   transforms), as the reference does below its NTT threshold
                   of 200 limbs (mul_basic.c:33-35); default 32, where the
                   one-block NTT measured as fast from there up
     "queue_waves" register budget of the work-queue kernel (FF_BATCH_QUEUE),
                   in waves per SIMD: 1 the compiler's own allocation, no
                   spills, two workgroups per CU; 3 or 4 (default, measured
                   fastest) more workgroups per CU at the price of spilled
                   registers
     "queue_parts" claim counters of the work-queue kernel (1..8, default 8):
                   workgroup w claims from part w mod parts
     "queue_chunk" tasks a workgroup takes per claim (1..64, default 8)
     "dec_queue"   1: the decimal conversion runs each of its level batches
                   as one work-queue launch (FF_BATCH_QUEUE) instead of one
                   launch per pass kind; default 0
     "test_fail_alloc" testing hook: after this many more device or pinned
                   allocations, every further one fails as if memory were
                   exhausted, so the drop-in's "out of memory (N bytes)" and
                   exit(1) (alloc.c:142-146) can be exercised; -1 (default)
                   turns it off
   Returns FF_ERR_ARG for an unknown name or value. */
FF_API int ff_set_option(const char *name, long long value);
/* The current value of an option above (FF_ERR_ARG for an unknown name). */
FF_API int ff_get_option(const char *name, long long *value);

/* Live kernel timer (no counterpart in the reference; its profiling was a
   sampling profiler, tests/sprof.c). While enabled, each kernel launch is
   bracketed by HIP events on the library stream. ff_prof_read returns, for
   kernel kind k (0 .. until it returns FF_ERR_ARG), the kind's name, its
   launches, total milliseconds and total algorithmic bytes moved. */
FF_API int ff_prof_enable(int on);
FF_API void ff_prof_reset(void);
FF_API int ff_prof_read(int k, const char **name, long *launches, double *ms, double *bytes);
/* The same kind's total algorithmic 32-bit integer multiplies, booked by the
   launchers of the transform kernels that count them (the inner blocks: 3 per
   twiddled butterfly, a Montgomery product by a root with its precomputed
   companion; 4 per pointwise Montgomery product), 0 for the others. No
   counterpart in the reference. */
FF_API int ff_prof_read_mults(int k, double *mults);

/* Issue-rate probe (no counterpart in the reference): the GPU's sustained
   rate of 32-bit integer multiplies (v_mul_lo_u32 and v_mul_hi_u32, 8
   independent chains per lane on every SIMD), the peak the transforms'
   INT32-multiply roofline is quoted against; *seconds gets the timed
   launch's duration. */
FF_API int ff_probe_int32_mul(double *mults_per_s, double *seconds);

/* Bandwidth probe (no counterpart in the reference): a plain device copy of
   `bytes` bytes (16 bytes per lane, grid-stride, 8 workgroups per CU),
   best of three timed runs; *bytes_per_s counts bytes read plus bytes
   written. The measured HBM peak beside the 8 TB/s specification, and the
   dispatch of known size (k_copy_probe) that scales rocprofv3's FETCH_SIZE
   and WRITE_SIZE counters to bytes. */
FF_API int ff_probe_copy(size_t bytes, double *bytes_per_s, double *seconds);

/* Allocation counters (no counterpart in the reference, whose arena
   reuses warm pages, alloc.c:19-23,74-140): the device and pinned-host
   allocations the library has made since it started, and the bytes its
   device pool holds for reuse. Device memory is reused from the pool in
   stream order, so repeating the same work (F(n), a batch run, a decimal
   conversion) leaves both counts unchanged. */
FF_API int ff_mem_stats(uint64_t *device_allocs, uint64_t *host_allocs, size_t *cached_bytes);

/* Testing hook (no counterpart in the reference): marks the carry look-back
   time-out word as set, as a bounded wait that ran out would, without any
   kernel waiting. The next length read-back then fails loudly: a message on
   stderr and exit(1), as after a real time-out. For failure-path tests only. */
FF_API int ff_test_set_lb_fault(void);

#ifdef __cplusplus
}
#endif

#endif
