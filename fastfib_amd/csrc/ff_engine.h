// Product schedules shared by the single-product engine (ff_engine.hip) and
// the batch engine (ff_batch.hip).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/fastfib_amd.h"
#include "ff_internal.h"

// A product's transform: width bw, length L = 2^lg or 3 * 2^lg (r3), conv
// coefficients; split = 1 when no single transform fits (L > 2^25).
struct ProdShape {
    unsigned bw;
    uint64_t L;
    int lg, r3;
    uint64_t conv;
    int split;
};

// A schedule choice beside the reference rule: the transform length (0: the
// shortest, the reference's choice; 1: the next supported length up), the
// inner block (log2, 0: default) and the outer tile (LDS words, 0: default).
struct FFSched {
    int len_up;
    int inner_log;
    int tile_words;
    int split_lg;  // split products: piece transform 2^split_lg (0: the reference's cost model)
};

ProdShape ff_product_shape(uint64_t bits_a, uint64_t bits_b);
// The CRT bound (mul_ntt.c:44-53): min(ca, cb) * 2^(2w) < p1 p2 p3.
int ff_width_ok(uint64_t bits_a, uint64_t bits_b, unsigned w);
// The shape with the schedule's length choice applied.
ProdShape ff_shape_with(const ProdShape &s, const FFSched &sc);

// One product of a launch group (device pointers).
struct MulItem {
    const uint64_t *a;
    uint64_t na, bits_a;
    const uint64_t *b;  // nullptr: square
    uint64_t nb, bits_b;
    uint64_t *out;      // room for na + nb + 1 limbs
    uint64_t conv;      // product coefficients at the group's width
    unsigned long long *lenkey;  // where the CRT books the normalized length
};

// One product past the longest transform (split into transformed-once pieces).
void ff_mul_split(const MulItem &it, int split_lg = 0);
// Both operands (normalized limb counts) below the schoolbook bound.
bool ff_school_ok(uint64_t na, uint64_t nb);
// A split product's coefficient width and piece transform length (log2).
// split_lg > 0 forces the piece length (kept within [max(10, max_lg - 4), max_lg]).
void ff_split_plan(uint64_t bits_a, uint64_t bits_b, bool sq, unsigned *bw, int *lgp, int split_lg = 0);

// Runs up to FF_MAXB products sharing one shape and schedule, one launch
// per pass (transform arrays come from the memory pool, stream-ordered).
// With hand: no CRT; hand[i] gets product i's residues, whose buffers are
// appended to keep for the caller to free after it has queued their use.
void ff_mul_group(const MulItem *it, int cnt, const ProdShape &s, const FFSched &sc, FFCrtJob *hand = nullptr,
                  std::vector<void *> *keep = nullptr);

// Effective schedule knobs (0 in sc: the global options).
int ff_sched_inner_log(const FFSched &sc);
int ff_sched_tile_words(const FFSched &sc);

// The schedules of n products (shortest shapes s0, bit lengths, square
// flags) in a batch of count products: the reference rule, the autotuner
// table (FF_BATCH_TUNED) or the learned dispatcher (FF_BATCH_LEARNED, one
// inference launch for all n).
void ff_pick_scheds(const ProdShape *s0, const uint64_t *bits_a, const uint64_t *bits_b, const uint8_t *square,
                    size_t n, size_t count, int flags, FFSched *out);

// ff_batch_create with the schedules given (force[i] for item i) instead of picked.
int ff_batch_create_sched(ff_batch **out, const ff_prod *items, size_t count, int flags, const FFSched *force);
