// Single-pass carry resolution across the workgroups of a launch, by
// decoupled look-back (Merrill and Garland, "Single-pass Parallel Prefix
// Scan with Decoupled Look-back", NVIDIA technical report NVR-2016-002) on
// binary carry chains: one launch instead of a local pass, a carry scan and
// a carry-apply pass.
//
// A job's workgroups (256 threads, one limb each) take tickets in start
// order (atomicAdd), so a workgroup only waits for workgroups that are
// already running and the walk always ends. Per round, a workgroup forms its
// aggregate for each chain (generate G: carry-out for a zero carry-in;
// propagate P: carry-out = carry-in, every limb all ones) from wavefront
// ballots, publishes it, then walks back over its predecessors' status words
// until each chain meets a predecessor that generates, kills (G = P = 0) or
// has published its inclusive carry-out (for non-adversarial limbs the walk
// stops at the first predecessor), and publishes its own inclusive
// carry-out. The carry into each limb then follows from the same ballot
// arithmetic as inside one workgroup: A = g, B = g | p, C = (A + B + c) ^ A ^ B.
//
// Workspace of a job (uint32 words, zero on entry): [0] ticket counter,
// [1] finished-group counter, then status words, then one word through
// which the second-to-last workgroup passes the bit lengths of its last two
// limbs to the last (for the length keys), then the finished-workgroup
// counters of groups of LB_GROUP workgroups. The last workgroup to finish
// zeroes it again, so it can be reused by the next launch on the stream
// (and by graph replays) without a clearing pass.
//
// Every cross-workgroup word is accessed by relaxed device-scope atomics:
// each carries its whole message, so no fence (no L2 write-back or
// invalidation across the XCDs) is needed; outputs are only read after the
// launch.
//
// Status word: bits 0-1 kind (LB_AGG, LB_INC); aggregate: G_x at bit 2 + 2x,
// P_x at bit 3 + 2x; inclusive: carry-out of chain x at bit 2 + x (x < 3).
#pragma once
#include <stdint.h>

#define LB_T 256
#define LB_NW (LB_T / 64)
#define LB_AGG 1u
#define LB_INC 2u
#define LB_SPIN_MAX (1u << 18)

struct LbShared {
    uint32_t wg[3][LB_NW], wp[3][LB_NW];
    uint32_t cin;   // the block's carry-ins, bit x for chain x
    uint32_t word;  // broadcast of thread 0's ticket / last flag
};

__device__ __forceinline__ uint32_t lb_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void lb_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A nonzero word, waited for with a bound: a wait that runs out sets *fault
// (read by the host, which then fails loudly) and returns 0.
__device__ __noinline__ uint32_t lb_wait(const uint32_t *p, uint32_t *fault) {
    for (uint32_t it = 0; it < LB_SPIN_MAX; it++) {
        const uint32_t v = lb_load(p);
        if (v) return v;
        __builtin_amdgcn_s_sleep(2);
    }
    atomicOr(fault, 1u);
    return 0;
}

// The workgroup's position among the job's workgroups, in start order. A
// grid that fits the GPU at once (every workgroup resident together) uses
// its grid index and skips the ticket atomic, which all workgroups would
// contend for at the same moment.
__device__ __forceinline__ uint32_t lb_ticket(LbShared &S, uint32_t *ws, bool resident, uint32_t bx) {
    if (resident) return bx;
    if (threadIdx.x == 0) S.word = atomicAdd(ws, 1u);
    __syncthreads();
    return S.word;
}

// One round over NC (1 to 3) chains: limb (generate, propagate) per thread
// in g[x], p[x]; st: the round's status words; init: bit x = carry into
// block 0 of chain x. Returns bit x = carry into this thread's limb.
template <int NC>
__device__ __forceinline__ uint32_t lb_round(LbShared &S, uint32_t *st, uint32_t b, const bool *g, const bool *p,
                                             uint32_t init, uint32_t *fault) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t G[NC], B[NC];
#pragma unroll
    for (int x = 0; x < NC; x++) {
        G[x] = __ballot(g[x]);
        B[x] = G[x] | __ballot(p[x]);
        if (lane == 0) {
            S.wg[x][w] = (G[x] + B[x]) < G[x];
            S.wp[x][w] = (B[x] == ~0ull && G[x] == 0);
        }
    }
    __syncthreads();
    uint64_t WG[NC], WB[NC];
#pragma unroll
    for (int x = 0; x < NC; x++) {
        WG[x] = __ballot(lane < LB_NW && S.wg[x][lane]);
        WB[x] = WG[x] | __ballot(lane < LB_NW && S.wp[x][lane]);
    }
    if (threadIdx.x == 0) {
        uint32_t agg = 0;
#pragma unroll
        for (int x = 0; x < NC; x++) {
            const uint32_t bg = (uint32_t)(((WG[x] + WB[x]) >> LB_NW) & 1);
            const uint32_t bp = (WB[x] == (1ull << LB_NW) - 1 && WG[x] == 0) ? 1u : 0u;
            agg |= (bg | (bp << 1)) << (2 * x);
        }
        uint32_t cin = 0;
        if (b == 0) {
            cin = init;
        } else {
            lb_store(st + b, LB_AGG | (agg << 2));
            uint32_t need = (1u << NC) - 1;
            for (int64_t j = (int64_t)b - 1; need; j--) {
                if (j < 0) {
                    cin |= init & need;
                    break;
                }
                const uint32_t v = lb_wait(st + j, fault);
#pragma unroll
                for (int x = 0; x < NC; x++) {
                    if (!(need >> x & 1)) continue;
                    if (v & LB_INC) {
                        cin |= ((v >> (2 + x)) & 1) << x;
                        need &= ~(1u << x);
                    } else {
                        const uint32_t gg = (v >> (2 + 2 * x)) & 1, pp = (v >> (3 + 2 * x)) & 1;
                        if (gg) cin |= 1u << x;
                        if (gg || !pp) need &= ~(1u << x);  // else the predecessor propagates: look further
                    }
                }
            }
        }
        uint32_t inc = 0;
#pragma unroll
        for (int x = 0; x < NC; x++) {
            const uint32_t bg = (agg >> (2 * x)) & 1, bp = (agg >> (2 * x + 1)) & 1;
            inc |= (bg | (bp & (cin >> x))) << x;
        }
        lb_store(st + b, LB_INC | (inc << 2));
        S.cin = cin;
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int x = 0; x < NC; x++) {
        const uint64_t WC = (WG[x] + WB[x] + ((S.cin >> x) & 1)) ^ WG[x] ^ WB[x];
        const uint64_t C = (G[x] + B[x] + ((WC >> w) & 1)) ^ G[x] ^ B[x];
        c |= (uint32_t)((C >> lane) & 1) << x;
    }
    return c;
}

// Finished-workgroup counting is two-level: a counter per group of
// LB_GROUP workgroups (after the status words, at gc), and ws[1] counting
// the complete groups, so no single word takes every workgroup's atomic.
#define LB_GROUP 32
__host__ __device__ __forceinline__ uint64_t lb_groups(uint64_t nblk) { return (nblk + LB_GROUP - 1) / LB_GROUP; }

// True in the job's last workgroup to finish (every other workgroup is past
// its last read of the workspace); b: the workgroup's position.
__device__ __forceinline__ bool lb_last(LbShared &S, uint32_t *ws, uint32_t *gc, uint32_t b, uint32_t nblk) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = b / LB_GROUP;
        const uint32_t gsize = nblk - g * LB_GROUP < LB_GROUP ? nblk - g * LB_GROUP : LB_GROUP;
        bool last = false;
        if (atomicAdd(gc + g, 1u) == gsize - 1) last = atomicAdd(ws + 1, 1u) == (uint32_t)lb_groups(nblk) - 1;
        S.word = last ? 1u : 0u;
    }
    __syncthreads();
    return S.word != 0;
}

// Zeroes a job's workspace of n words (the last workgroup, all threads).
__device__ __forceinline__ void lb_clear(uint32_t *ws, uint64_t n) {
    for (uint64_t k = threadIdx.x; k < n; k += LB_T) ws[k] = 0;
}

__device__ __forceinline__ uint32_t lb_blen(uint64_t v) { return v ? 64u - (uint32_t)__clzll((long long)v) : 0u; }

typedef unsigned __int128 lb_u128;

// A thread's LPT consecutive limbs as one link of a chain: from limb sums
// u[q] < 2^65 (u[q] = 2^64 - 1 past the result, so a carry passes through),
// the group's generate (carry-out for carry-in 0) and propagate (carry-out
// equals carry-in).
template <int LPT>
__device__ __forceinline__ void lb_group_gp(const lb_u128 *u, bool *g, bool *p) {
    uint32_t c0 = 0, c1 = 1;
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        c0 = (uint32_t)((u[q] + c0) >> 64);
        c1 = (uint32_t)((u[q] + c1) >> 64);
    }
    *g = c0 != 0;
    *p = c1 != 0 && c0 == 0;
}

// The group's limbs for carry-in c.
template <int LPT>
__device__ __forceinline__ void lb_group_apply(const lb_u128 *u, uint32_t c, uint64_t *v) {
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const lb_u128 s = u[q] + c;
        v[q] = (uint64_t)s;
        c = (uint32_t)(s >> 64);
    }
}

// Length keys of up to two len-limb outputs, thread t holding limbs
// i0 + t LPT + q (v[x * LPT + q]): the top nonzero limb among the last
// three, as (limbs << 8) | its bit count, 0 for zero. Workgroup b of nblk:
// the second-to-last publishes the bit lengths of its last two limbs in
// *kw, the last combines them with its own and stores the keys. Every
// thread of the workgroup calls it.
template <int NO, int LPT>
__device__ __forceinline__ void lb_keys(uint32_t *kw, uint32_t b, uint32_t nblk, uint64_t i0, uint64_t len,
                                        const uint64_t *v, unsigned long long *const *key, uint32_t *fault) {
    __shared__ unsigned long long best[NO];
    __shared__ uint32_t tail[NO][2];
    const int t = threadIdx.x;
    const uint64_t ib = i0 + (uint64_t)t * LPT;  // this thread's first limb
    if (nblk >= 2 && b == nblk - 2) {
        // the workgroup's last two limbs: (thread, q) = (T - 1, LPT - 2..LPT - 1), or threads T - 2, T - 1
#pragma unroll
        for (int q = 0; q < LPT; q++) {
            const int r = (t * LPT + q) - (LB_T * LPT - 2);
            if (r >= 0) {
#pragma unroll
                for (int x = 0; x < NO; x++) tail[x][r] = lb_blen(v[x * LPT + q]);
            }
        }
        __syncthreads();
        if (t == 0) {
            uint32_t w = 1u << 31;
#pragma unroll
            for (int x = 0; x < NO; x++) w |= (tail[x][0] | (tail[x][1] << 7)) << (14 * x);
            lb_store(kw, w);
        }
    }
    if (b != nblk - 1) return;  // uniform over the workgroup
    if (t < NO) best[t] = 0;
    __syncthreads();
    const uint64_t lo = len >= 3 ? len - 3 : 0;
#pragma unroll
    for (int q = 0; q < LPT; q++) {
        const uint64_t i = ib + q;
        if (i >= lo && i < len) {
#pragma unroll
            for (int x = 0; x < NO; x++)
                if (v[x * LPT + q]) atomicMax(&best[x], ((unsigned long long)(i + 1) << 8) | lb_blen(v[x * LPT + q]));
        }
    }
    __syncthreads();
    if (t == 0) {
        uint32_t w = 0;
#pragma unroll
        for (int x = 0; x < NO; x++) {
            unsigned long long k = best[x];
            // the candidates below this workgroup: limbs i0 - 1 and i0 - 2
            for (uint64_t j = i0; k == 0 && j > lo && j + 2 > i0; j--) {
                if (!w) w = lb_wait(kw, fault);
                const uint32_t bl = (w >> (14 * x + (j == i0 ? 7 : 0))) & 127;
                if (bl) k = ((unsigned long long)j << 8) | bl;
            }
            *key[x] = k;
        }
    }
}
