// Runtime services shared by the host side of fastfib_amd: the library
// stream, the stream-ordered device memory pool, errors and options.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

extern hipStream_t g_st;
extern std::recursive_mutex g_mu;
extern int g_inner_log;
extern int g_tile_words;
// outer tile of F(n)'s own steps (doubling squares under the reference rule,
// quadruplings, pair products; 0: tile_words): their small passes are
// latency, where smaller tiles mean more workgroups
extern int g_fib_tile_words;
// Levels per round of the outer passes: 0 = by the pass's row count 2^rho
// (three where bit rho of g_outer_rl3 is set, else four), 2..4 = that many
// in every outer pass.
extern int g_outer_rl;
extern int g_outer_rl3;
// Levels per round of the pair kernel: 0 = two for blocks of at most 2^10
// words, else four; 2..4 = that many.
extern int g_pair_rl;
// Output limbs above which a single CRT launch and a fused doubling step
// take four limbs per thread (default 1280 workgroups of 256: one wave of
// residency).
extern long long g_crt_wave_limbs;
extern uint64_t g_force_len;
extern int g_max_lg;    // longest single transform 2^max_lg; longer products split
extern int g_pair_steps;   // bit 0: doubling steps as weighted pairs; bit 1: the last product as one
extern int g_quad_max_lg;     // longest quadrupling transform 2^quad_max_lg (0: no quadrupling)
extern int g_quad_inner_log;  // inner block of quadrupling transforms (0: inner_log)
extern int g_host_bits;       // top bits of n done on the host before the device chain (10..13)
extern int g_school_limbs;    // products with both operands shorter than this run schoolbook (0: none)
extern int g_queue_waves;     // register budget of the batch work-queue kernel (waves per SIMD: 1 = unbounded, 3, 4)
extern int g_queue_parts;     // claim counters of the work-queue kernel (1..8)
extern int g_queue_chunk;     // tasks per claim (1..64)
extern int g_dec_queue;       // the decimal conversion's level batches run as work queues (FF_BATCH_QUEUE)
extern int g_dispatch;      // schedule source of the drop-in products and F(n): 0, FF_BATCH_TUNED or FF_BATCH_LEARNED
extern int g_dispatch_gen;  // bumped whenever the table or the network changes

void ff_set_err(const char *fmt, ...);
void ff_ensure_init(void);
int ff_is_inited(void);
int ff_device(void);  // the device the library runs on
// The allocation-free product's scratch array on `device` (ff_noalloc.hip).
int ff_noalloc_setup(int device);

// Device memory: blocks are cached by size class and reused in stream
// order (all device work runs on g_st), hipMalloc only on a miss.
void *ff_dalloc(size_t bytes);
void ff_dfree(void *p);
// The library's only allocation calls (counted; see ff_mem_stats).
hipError_t ff_hip_malloc(void **p, size_t bytes);
hipError_t ff_host_malloc(void **p, size_t bytes);
// Bytes the pool holds for reuse (free for the library, not for the driver).
size_t ff_pool_cached(void);

// Kernel timing is suspended while a graph is being captured (event
// records and their readback do not belong in a replayed graph).
void ff_prof_pause(int pause);

// Host <-> device copy ordered after everything queued on the library
// stream and complete on return. The stream is non-blocking, so a plain
// hipMemcpy (null stream) is not ordered after its kernels: a recycled pool
// block could be written while a queued kernel still reads it.
void ff_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);

// Pinned host buffer for reading small results back.
uint64_t *ff_pinned_results(size_t words);
// The same buffer's address for kernels (valid after ff_pinned_results).
uint64_t *ff_pinned_results_dev(void);

// Workspace of the single-pass carry kernels: zeroed regions (left zeroed
// by the kernels), handed out from the start again by ff_lb_begin; the word
// a look-back wait that times out sets (read back with the length keys).
void ff_lb_begin(void);
uint32_t *ff_lb_take(size_t words);
uint32_t *ff_lb_fault(void);
// What a set time-out word means for the caller: results are wrong and the
// workspace is no longer zeroed, so the process stops (a message on stderr
// and exit(1), like every other device failure of the drop-in surface).
[[noreturn]] void ff_lb_timeout_fail(void);
// Workgroups of the kernel (at this block size) the GPU holds at once.
uint64_t ff_resident_blocks(const void *kernel, int threads);
