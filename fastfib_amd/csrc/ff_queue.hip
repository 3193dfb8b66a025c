// The batch work-queue kernel (k_queue): one persistent launch that runs a
// whole batch's transform passes, CRT jobs and schoolbook products as tasks
// claimed in table order, with per-product dependency counters.
#include <map>

#include "ff_crt_body.h"
#include "ff_internal.h"
#include "ff_lookback.h"
#include "ff_ntt_body.h"
#include "ff_runtime.h"
#include "ff_school_dev.h"

// ------------------------------------------------------------------ batch work queue
//
// A whole batch in one launch: workgroups (as many as the GPU holds at
// once) claim the tasks of the table one at a time, in table order, with one
// atomic. The table lists every block of every list launch of the grouped
// form (forward passes, shared one-block spectra, inner blocks, inverse
// passes, CRT jobs, schoolbook products, zero products) in that order, so
// each task's dependencies - the earlier passes of its own product, and the
// transform of a b operand it shares - are earlier tasks, already claimed
// by running workgroups: waits always end, whatever the residency. A
// product's next pass starts as soon as its own previous pass ends, not the
// whole batch's, and the launches' drain and fill between pass kinds are
// gone. CRT jobs take look-back tickets in claim order (never by grid index).
//
// Visibility: the transform arrays that pass between tasks are written and
// read only by agent-scope (coherent) accesses (ff_ld / ff_st; the L2s of
// the XCDs are not coherent with each other), every wave's stores complete
// before a finished task's counter is bumped, and a waiting task's loads are
// issued after its counters are seen. Agent-scope fences would do the same
// for ordinary accesses by writing back and invalidating whole L2s at every
// task, which measured several times slower than the grouped launches.
#define FQ_SPIN_MAX (1u << 20)

__device__ __noinline__ void fq_wait(const uint32_t *c, uint32_t need, uint32_t *fault) {
    for (uint32_t it = 0;; it++) {
        if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) return;
        // a wait that runs out marks the fault word (the host then fails
        // loudly); once it is marked, later waits give up at once
        if ((it & 255) == 255 && lb_load(fault)) return;
        if (it >= FQ_SPIN_MAX) {
            atomicOr(fault, 1u);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Claims: the table is dealt to nq parts (task t to part t mod nq), each
// with its own claim counter (on its own cache line: one contended address
// serializes every claim of the launch); workgroup w claims from part
// w mod nq, chunk tasks of its part per atomic, and runs them in order.
// Every part keeps at least one workgroup (the grid is at least nq), and a
// part's claims follow the table order, so the lowest unfinished task is
// always held by a workgroup whose earlier tasks are done: waits end.
template <int MINW>
__global__ __launch_bounds__(FF_THREADS, MINW) void k_queue(const FFQTask *tasks, uint32_t ntasks, const char *desc,
                                                      uint32_t *ctr, uint32_t nctr, uint32_t nq, uint32_t chunk,
                                                      const FFTables *T, FFPrimes PR, FFGarner G, uint32_t *fault) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_task, s_last;
    uint32_t *const cnt = ctr + FQ_HEADER;
    const uint32_t part = blockIdx.x % nq;
    uint32_t next = 0, left = 0;  // the next task of the held chunk, tasks left in it
    for (;;) {
        if (!left) {
            if (threadIdx.x == 0) s_task = atomicAdd(ctr + FQ_LINE * part, 1u);
            __syncthreads();
            // uniform across the workgroup: the task's fields and its item's
            // load as scalars, as in the list kernels
            const uint32_t c = __builtin_amdgcn_readfirstlane(s_task);
            next = part + nq * c * chunk;
            left = chunk;
        }
        const uint32_t t = next;
        if (t >= ntasks) break;
        next += nq;
        left--;
        const FFQTask task = tasks[t];
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < 2; k++)
                if (task.dep[k] != FQ_NONE) fq_wait(cnt + task.dep[k], task.need[k], fault);
        }
        // the data other tasks wrote is read by coherent loads, issued after
        // this barrier (so after the counters were seen)
        __syncthreads();
        const char *item = desc + task.off;
        const uint32_t sub = task.sub;
        if (task.kind < FQ_INNER) {
            const FFOuterItem &it = *(const FFOuterItem *)item;
            const uint32_t ntile = it.ps.nblocks * it.ps.tpb;
            const int pi = (int)(sub / ntile);
            const uint32_t tile = sub - (uint32_t)pi * ntile;
            switch (task.kind) {
            case FQ_OUTER | 2 | 4: outer_body<false, true, true, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 2: outer_body<false, true, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER: outer_body<false, false, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 1 | 2 | 4: outer_body<true, true, true, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            case FQ_OUTER | 1 | 2: outer_body<true, true, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            default: outer_body<true, false, false, true>(it.ps, it.x, pi, tile, T, PR, lds); break;
            }
        } else if (task.kind <= FQ_INNER_SMALL_R3) {
            const FFInnerItem &it = *(const FFInnerItem *)item;
            if (task.kind == FQ_INNER) {
                const uint32_t nblk = it.in.L / it.in.S;
                const int pi = (int)(sub / nblk);
                inner_body<false, false, true>(it.in, it.c, pi, sub - (uint32_t)pi * nblk, T, PR, lds);
            } else if (task.kind == FQ_INNER_SMALL) {
                inner_body<true, false, true>(it.in, it.c, (int)sub, 0, T, PR, lds);
            } else {
                inner_body<true, true, true>(it.in, it.c, (int)sub, 0, T, PR, lds);
            }
        } else if (task.kind == FQ_CRT) {
            const FFCrtJob &J = *(const FFCrtJob *)item;
            CrtShared &sh = *(CrtShared *)lds;
            if (crt_lpt_list(J.bw, J.nout) == 4) crt_body<4, true>(J, sub, G, fault, sh, false);
            else crt_body<1, true>(J, sub, G, fault, sh, false);
        } else if (task.kind == FQ_SCHOOL) {
            school_body(*(const FFSchoolItem *)item, *(FFSchoolShared *)lds);
        } else if (task.kind == FQ_SCHOOL_WAVE) {
            FFSchoolWaveShared &sh = *(FFSchoolWaveShared *)lds;
            const int w = threadIdx.x >> 6;
            const bool live = (uint32_t)w < sub;
            const FFSchoolItem *its = (const FFSchoolItem *)item;
            school_wave(its[live ? w : 0], live, sh.A[w], sh.B[w], sh.lo[w], sh.mid[w], sh.hi[w]);
        } else {
            if (threadIdx.x == 0) **(uint64_t *const *)item = 0;
        }
        // this task's coherent stores complete (every wave) before its counter moves
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0 && task.done != FQ_NONE) atomicAdd(cnt + task.done, 1u);
    }
    // every task is claimed; the last workgroup to get here (all others have
    // finished their last task) clears the counters for the next launch
    if (threadIdx.x == 0) s_last = atomicAdd(ctr + FQ_LINE * FQ_MAXQ, 1u) == gridDim.x - 1;
    __syncthreads();
    if (s_last) {
        for (uint32_t i = threadIdx.x; i < FQ_HEADER + nctr; i += blockDim.x)
            __hip_atomic_store(ctr + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}


// One workgroup per task up to what the GPU holds at once with this LDS
// size (more would only queue behind them).
hipError_t ff_launch_queue(const FFQTask *tasks, uint32_t ntasks, const char *desc, uint32_t *ctr, uint32_t nctr,
                           size_t lds, double bytes, hipStream_t st) {
    if (!ntasks) return hipSuccess;
    // register budget: waves per SIMD the compiler must allow (1: its own
    // choice, no spills; 3 or 4: more workgroups resident, some spills)
    const int w = g_queue_waves;
    const void *k = w >= 4 ? (const void *)k_queue<4> : w == 3 ? (const void *)k_queue<3> : (const void *)k_queue<1>;
    static std::map<std::pair<const void *, size_t>, uint32_t> resident;
    auto r = resident.find({k, lds});
    if (r == resident.end()) {
        int per_cu = 0, cus = 0;
        FF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, FF_THREADS, lds));
        FF_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ff_device()));
        r = resident.emplace(std::make_pair(k, lds), (uint32_t)(per_cu > 0 ? per_cu : 1) * (uint32_t)cus).first;
    }
    const uint32_t chunk = (uint32_t)g_queue_chunk;
    const uint32_t chunks = (ntasks + chunk - 1) / chunk;
    const uint32_t grid = chunks < r->second ? chunks : r->second;
    const uint32_t nq = (uint32_t)g_queue_parts < grid ? (uint32_t)g_queue_parts : grid;
    FFGarner G;
    ff_garner_get(&G);
    ff_prof_begin(FK_QUEUE, st);
    if (w >= 4)
        hipLaunchKernelGGL(k_queue<4>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           ff_tables_dev(),
                           ff_primes(), G, ff_lb_fault());
    else if (w == 3)
        hipLaunchKernelGGL(k_queue<3>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           ff_tables_dev(),
                           ff_primes(), G, ff_lb_fault());
    else
        hipLaunchKernelGGL(k_queue<1>, dim3(grid), dim3(FF_THREADS), lds, st, tasks, ntasks, desc, ctr, nctr, nq, chunk,
                           ff_tables_dev(),
                           ff_primes(), G, ff_lb_fault());
    ff_prof_end(FK_QUEUE, bytes, st);
    return hipGetLastError();
}
