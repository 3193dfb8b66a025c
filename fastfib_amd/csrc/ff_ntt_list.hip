// Work-list forms of the transform kernels: one launch runs the same pass
// kind of many heterogeneous products (a batch, a decimal level), the flat
// grid cut among them by a start array.
#include "ff_internal.h"
#include "ff_ntt_body.h"
#include "ff_runtime.h"

// Work-list form: item k owns blocks [start[k], start[k+1]) of the flat
// grid, 3 * nblocks * tpb of them (the prime is the major index).
template <bool INV, bool TOP, bool R3, int SH = 0>
__global__ __launch_bounds__(FF_THREADS) void k_outer_list(const FFOuterItem *items, const uint32_t *start, int n,
                                                           const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFOuterItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t ntile = it.ps.nblocks * it.ps.tpb;
    const int pi = (int)(local / ntile);
    outer_body<INV, TOP, R3, false, SH>(it.ps, it.x, pi, local - (uint32_t)pi * ntile, T, PR, lds);
}

// Work-list form: item k owns blocks [start[k], start[k+1]), 3 * nblk of them.
template <bool SMALL, bool R3, bool FIX12 = false>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner_list(const FFInnerItem *items, const uint32_t *start, int n,
                                                              const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFInnerItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t nblk = SMALL ? 1 : it.in.L / it.in.S;
    const int pi = (int)(local / nblk);
    inner_body<SMALL, R3, false, FIX12>(it.in, it.c, pi, local - (uint32_t)pi * nblk, T, PR, lds);
}

// Inner blocks of a transform on their own, in place: the forward levels
// (INV false) or the inverse levels (INV true), no pointwise product (the
// split multiply forms its pointwise sums between the two).
template <bool INV>
__global__ __launch_bounds__(FF_THREADS, 4) void k_inner_part_list(const FFInnerItem *items, const uint32_t *start,
                                                                   int n, const FFTables *T, FFPrimes PR) {
    extern __shared__ uint32_t lds[];
    const int k = ff_find_item(start, n, blockIdx.x);
    const FFInnerItem &it = items[k];
    const uint32_t local = blockIdx.x - start[k];
    const uint32_t nblk = it.in.L / it.in.S;
    const int pi = (int)(local / nblk);
    const uint32_t blkx = local - (uint32_t)pi * nblk;
    const FFPrime P = PR.p[pi];
    FFTile ta{lds, 1, 1, 0, it.in.s, 1, 1};
    uint32_t *d = it.c.fa + (size_t)pi * it.in.L + (size_t)blkx * it.in.S;
    auto ld = [&](int pos, int c) -> uint32_t {
        (void)c;
        return d[pos];
    };
    auto st = [&](int pos, int c, uint32_t v) {
        (void)c;
        d[pos] = v;
    };
    if (!INV) ff_forward_io(ta, T->rt[pi], T->rtp[pi], P.p, P.pinv, true, true, ld, st);
    else ff_inverse_io(ta, T->irt[pi], T->irtp[pi], P.p, P.pinv, true, true, ld, st);
}

// A list of one pass kind, by the items' shape (sh > 0: every item has
// that fixed-geometry shape).
template <bool INV, bool TOP>
static void launch_outer_list(bool r3, int sh, const FFOuterItem *items, const uint32_t *start, int n, dim3 g,
                              size_t lds, hipStream_t st) {
    const FFTables *T = ff_tables_dev();
    const FFPrimes PR = ff_primes();
    const dim3 b(FF_THREADS);
#define FF_SHCASE(R3, SH)                                                                                      \
    case SH: hipLaunchKernelGGL((k_outer_list<INV, TOP, R3, SH>), g, b, lds, st, items, start, n, T, PR); \
        return;
    if constexpr (TOP) {
        if (r3) {
            switch (sh) {
                FF_SHCASE(true, 9)
                FF_SHCASE(true, 10)
                FF_SHCASE(true, 11)
                FF_SHCASE(true, 13)
                FF_SHCASE(true, 14)
            }
            hipLaunchKernelGGL((k_outer_list<INV, TOP, true>), g, b, lds, st, items, start, n, T, PR);
            return;
        }
        switch (sh) {
            FF_SHCASE(false, 1)
            FF_SHCASE(false, 2)
            FF_SHCASE(false, 3)
            FF_SHCASE(false, 4)
            FF_SHCASE(false, 5)
            FF_SHCASE(false, 6)
            FF_SHCASE(false, 7)
            FF_SHCASE(false, 8)
            FF_SHCASE(false, 12)
        }
        hipLaunchKernelGGL((k_outer_list<INV, TOP, false>), g, b, lds, st, items, start, n, T, PR);
    } else {
        switch (sh) {
            FF_SHCASE(false, 6)
            FF_SHCASE(false, 7)
        }
        hipLaunchKernelGGL((k_outer_list<INV, TOP, false>), g, b, lds, st, items, start, n, T, PR);
    }
#undef FF_SHCASE
}

extern "C" hipError_t ff_launch_outer_list(bool inv, bool top, bool r3, int sh, const FFOuterItem *items,
                                           const uint32_t *start, int n, uint32_t blocks, size_t lds, double bytes,
                                           hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = inv ? (top ? FK_INV_TOP : FK_INV) : (top ? FK_FWD_TOP : FK_FWD);
    const dim3 g(blocks);
    ff_prof_begin(kind, st);
    if (!inv && top) launch_outer_list<false, true>(r3, sh, items, start, n, g, lds, st);
    else if (!inv) launch_outer_list<false, false>(r3, sh, items, start, n, g, lds, st);
    else if (top) launch_outer_list<true, true>(r3, sh, items, start, n, g, lds, st);
    else launch_outer_list<true, false>(r3, sh, items, start, n, g, lds, st);
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inner_list(bool small, bool r3, bool fix12, const FFInnerItem *items,
                                           const uint32_t *start, int n, uint32_t blocks, size_t lds, double bytes,
                                           hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = small ? FK_SMALL : FK_INNER;
    // one-block lists of short transforms: n / 4 threads (what radix-4
    // rounds keep busy), whole wavefronts, so that more items run per CU.
    // The list's LDS size bounds its longest item: lds = 8 (n + n/16 + 16).
    uint32_t threads = FF_THREADS;
    if (small) {
        const size_t words = lds / 8;
        const size_t nmax = words > 16 ? (words - 16) * 16 / 17 + 1 : 1;
        const size_t want = (nmax / 4 + 63) / 64 * 64;
        threads = want < 64 ? 64 : (want < FF_THREADS ? (uint32_t)want : FF_THREADS);
    }
    const dim3 g(blocks), b(threads);
    ff_prof_begin(kind, st);
    if (!small && fix12) hipLaunchKernelGGL((k_inner_list<false, false, true>), g, b, lds, st, items, start, n, ff_tables_dev(), ff_primes());
    else if (!small) hipLaunchKernelGGL((k_inner_list<false, false>), g, b, lds, st, items, start, n, ff_tables_dev(), ff_primes());
    else if (r3) hipLaunchKernelGGL((k_inner_list<true, true>), g, b, lds, st, items, start, n, ff_tables_dev(), ff_primes());
    else hipLaunchKernelGGL((k_inner_list<true, false>), g, b, lds, st, items, start, n, ff_tables_dev(), ff_primes());
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_inner_part_list(bool inv, const FFInnerItem *items, const uint32_t *start, int n,
                                                uint32_t blocks, size_t lds, double bytes, hipStream_t st) {
    if (!blocks) return hipSuccess;
    const int kind = inv ? FK_INV : FK_FWD;
    ff_prof_begin(kind, st);
    if (!inv) hipLaunchKernelGGL((k_inner_part_list<false>), dim3(blocks), dim3(FF_THREADS), lds, st, items, start, n,
                                 ff_tables_dev(), ff_primes());
    else hipLaunchKernelGGL((k_inner_part_list<true>), dim3(blocks), dim3(FF_THREADS), lds, st, items, start, n,
                            ff_tables_dev(), ff_primes());
    ff_prof_end(kind, bytes, st);
    return hipGetLastError();
}
