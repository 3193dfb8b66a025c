// Device building blocks of the transforms: radix-2 DIF/DIT levels on
// groups of up to 16 elements held in registers, run as "rounds" over an LDS
// tile, and the radix-3 layer.
//
// Tile layout: nsub stacked sub-transforms of M = 2^rho rows each, C columns;
// element (sub s, row v, column c) lives at lds[(s * M + v) * ld + c]. A
// round covers the row bits [b_lo, b_lo + NL): each group is the 2^NL rows
// that differ only in those bits, for one (s, c) and one value of the other
// row bits. Forward rounds run the DIF levels b = rho-1 .. 0 (Gentleman-Sande,
// natural order in, bit-reversed order out); inverse rounds run DIT levels
// b = 0 .. rho-1 with inverse roots (bit-reversed in, natural out), so that
// inverse(forward(x)) = M * x. Twiddles of level b are rt[2^b + j], j = row
// mod 2^b, the table layout of ntt_arith.c:37 and ntt_tables.c:38-65 (reference).
#pragma once
#include "ff_arith.h"

// Twiddles of one group and level: the butterflies of level b = b_lo + lb
// use rt[2^b + (low | k << b_lo)] = w^low * w_(2^(lb+1))^k, k < 2^lb, with
// w = w_(2^(b+1)). Only the base w^low is read from the table; the others
// are products with the 16th roots w_16^m = rt[8 + m] (companions at
// rtp[8 + m]), the same for every group of the prime: one gathered load
// per level instead of 2^lb, at the price of cheap multiplies.
// LDS: the tables were copied into LDS (rt points there).
// LOW0: the round starts at row bit 0 (low = 0, h = 2^lb), where every
// twiddle rt[h + k] is one of the first 16 table entries, read as it is (no
// products) and the same for every group.
// Tables in HBM (not LDS), b_lo > 0: every twiddle rt[h + low + (k << b_lo)]
// and its companion are read from the tables by buffer loads whose
// per-twiddle offset k << b_lo is uniform (a scalar register), so the loads
// cost the vector unit one address per level, and the products that would
// derive them are left to the memory pipeline, idle in rounds between two
// LDS passes.
template <int LB, bool LDS, bool LOW0 = false>
__device__ __forceinline__ void ff_group_twiddles(uint32_t tw[1 << LB], uint32_t twp[1 << LB], int h, int low,
                                                  const uint32_t *rt, const uint32_t *rtp, uint32_t p,
                                                  uint32_t pinv, int b_lo = 0) {
    // the 16th roots sit at the same address for every lane: from HBM they
    // are read through the constant address space, so they arrive by scalar
    // loads
    typedef const __attribute__((address_space(4))) uint32_t *cptr;
    if constexpr (LOW0) {
#pragma unroll
        for (int k = 0; k < (1 << LB); k++) {
            tw[k] = LDS ? rt[(1 << LB) + k] : ((cptr)rt)[(1 << LB) + k];
            twp[k] = LDS ? rtp[(1 << LB) + k] : ((cptr)rtp)[(1 << LB) + k];
        }
        return;
    }
    if constexpr (!LDS) {
        const __amdgpu_buffer_rsrc_t r0 = __builtin_amdgcn_make_buffer_rsrc((void *)rt, (short)0, FF_TW * 4, 0x00020000);
        const __amdgpu_buffer_rsrc_t r1 =
            __builtin_amdgcn_make_buffer_rsrc((void *)rtp, (short)0, FF_TW * 4, 0x00020000);
        const uint32_t vo = (uint32_t)(h + low) << 2;
#pragma unroll
        for (int k = 0; k < (1 << LB); k++) {
            const int so = (k << b_lo) << 2;
            tw[k] = __builtin_amdgcn_raw_buffer_load_b32(r0, vo, so, 0);
            twp[k] = __builtin_amdgcn_raw_buffer_load_b32(r1, vo, so, 0);
        }
        return;
    }
    tw[0] = rt[h + low];
    twp[0] = rtp[h + low];
#pragma unroll
    for (int k = 1; k < (1 << LB); k++) {
        const int m = k << (3 - LB);  // w_(2^(LB+1))^k = w_16^(k 2^(3-LB))
        const uint32_t r = LDS ? rt[8 + m] : ((cptr)rt)[8 + m];
        const uint32_t rp = LDS ? rtp[8 + m] : ((cptr)rtp)[8 + m];
        tw[k] = ff_mont_w(tw[0], r, rp, p);
        twp[k] = tw[k] * pinv;
    }
}

// LOW0: b_lo = 0 known at compile time (twiddles from the first table
// entries, the product by w^0 = 1 left out).
template <int NL, bool LDS = false, bool LOW0 = false>
__device__ __forceinline__ void ff_dif_group(uint32_t *v, int low, int b_lo, const uint32_t *rt, const uint32_t *rtp,
                                             uint32_t p, uint32_t pinv) {
#pragma unroll
    for (int lb = NL - 1; lb >= 0; lb--) {
        const int b = LOW0 ? lb : b_lo + lb;
        const int h = 1 << b;
        if (b == 0) {
#pragma unroll
            for (int j = 0; j < (1 << NL); j++) {
                if (j & (1 << lb)) continue;
                const int k = j | (1 << lb);
                const uint32_t x = v[j], y = v[k];
                v[j] = ff_add(x, y, p);
                v[k] = ff_sub(x, y, p);
            }
            continue;
        }
        uint32_t tw[8], twp[8];
        if (lb == 0) ff_group_twiddles<0, LDS, LOW0>(tw, twp, h, low, rt, rtp, p, pinv, b_lo);
        else if (lb == 1) ff_group_twiddles<1, LDS, LOW0>(tw, twp, h, low, rt, rtp, p, pinv, b_lo);
        else if (lb == 2) ff_group_twiddles<2, LDS, LOW0>(tw, twp, h, low, rt, rtp, p, pinv, b_lo);
        else ff_group_twiddles<3, LDS, LOW0>(tw, twp, h, low, rt, rtp, p, pinv, b_lo);
#pragma unroll
        for (int j = 0; j < (1 << NL); j++) {
            if (j & (1 << lb)) continue;
            const int k = j | (1 << lb);
            const uint32_t x = v[j], y = v[k];
            const int q = j & ((1 << lb) - 1);
            v[j] = ff_add(x, y, p);
            v[k] = (LOW0 && q == 0) ? ff_sub(x, y, p) : ff_mont_w_diff(x, y, tw[q], twp[q], p);
        }
    }
}

template <int NL, bool LDS = false, bool LOW0 = false>
__device__ __forceinline__ void ff_dit_group(uint32_t *v, int low, int b_lo, const uint32_t *irt, const uint32_t *irtp,
                                             uint32_t p, uint32_t pinv) {
#pragma unroll
    for (int lb = 0; lb < NL; lb++) {
        const int b = LOW0 ? lb : b_lo + lb;
        const int h = 1 << b;
        uint32_t tw[8], twp[8];
        if (b != 0) {
            if (lb == 0) ff_group_twiddles<0, LDS, LOW0>(tw, twp, h, low, irt, irtp, p, pinv, b_lo);
            else if (lb == 1) ff_group_twiddles<1, LDS, LOW0>(tw, twp, h, low, irt, irtp, p, pinv, b_lo);
            else if (lb == 2) ff_group_twiddles<2, LDS, LOW0>(tw, twp, h, low, irt, irtp, p, pinv, b_lo);
            else ff_group_twiddles<3, LDS, LOW0>(tw, twp, h, low, irt, irtp, p, pinv, b_lo);
        }
#pragma unroll
        for (int j = 0; j < (1 << NL); j++) {
            if (j & (1 << lb)) continue;
            const int k = j | (1 << lb);
            uint32_t x = v[j], y = v[k];
            const int q = j & ((1 << lb) - 1);
            if (b != 0 && !(LOW0 && q == 0)) y = ff_mont_w(y, tw[q], twp[q], p);
            v[j] = ff_add(x, y, p);
            v[k] = ff_sub(x, y, p);
        }
    }
}

// The three primes' constants, as a kernel argument.
struct FFPrimes {
    FFPrime p[3];
};
FFPrimes ff_primes(void);                // ff_ntt.hip

struct FFTile {
    uint32_t *s;  // LDS base
    int ld;       // row stride in words
    int C;        // columns (power of 2)
    int logC;
    int rho;      // rows per sub-transform = 2^rho
    int nsub;     // stacked sub-transforms (1, or 3 under a radix-3 layer, or several operands)
    int pad;      // 1: one spare word after every 16 rows (single-column tiles)
    int tid0 = 0; // the threads [tid0, tid0 + nthr) work on the tile;
    int nthr = 0; // nthr = 0: the whole workgroup
};

__device__ __forceinline__ int ff_tile_tid(const FFTile &t) { return (int)threadIdx.x - t.tid0; }
__device__ __forceinline__ int ff_tile_nthr(const FFTile &t) { return t.nthr ? t.nthr : (int)blockDim.x; }

// LDS index of row position pos (= sub * 2^rho + row), column c.
__device__ __forceinline__ int ff_pos_addr(const FFTile &t, int pos, int c) {
    return pos * t.ld + c + (t.pad ? (pos >> 4) : 0);
}

__device__ __forceinline__ int ff_tile_addr(const FFTile &t, int sub, int row, int c) {
    return ff_pos_addr(t, (sub << t.rho) + row, c);
}

// A round whose groups may come from a loader instead of LDS and go to a
// storer instead of LDS: ld(pos, c) -> value, st(pos, c, value). The first
// round of a transform reads HBM straight into registers and the last
// writes HBM straight from them, so a tile makes one fewer LDS round trip
// at each end. MID fuses the last DIF round, the pointwise product with
// the other operand's spectrum (LDS array B, or a square) and the first DIT
// round, which cover the same row bits and therefore the same groups.
enum { FF_R_DIF = 0, FF_R_DIT = 1, FF_R_MID = 2 };

// Loaders and storers of the rounds are callables ld(pos, c) -> value and
// st(pos, c, value). A row accessor (ff_rows_io: a member `rows`) also
// gives the byte offset of element (pos, c), row0(pos, c), and the byte
// stride of one row, rstep(), with ldb / stb at byte offsets: a group's
// elements then cost one addition each (offsets advance by uniform
// strides) instead of an address computation.
template <class T, class = void>
struct ff_rows_io : std::false_type {};
template <class T>
struct ff_rows_io<T, std::void_t<decltype(T::rows)>> : std::true_type {};

// Twiddles of the outer four-step passes, applied in the round that touches
// HBM (the forward's last, the inverse's first). Those rounds cover the
// lowest row bits, so a group holds 2^NL consecutive rows, whose twiddles
// are base * g^rev(j): the provider computes them from two table roots per
// group instead of one per element. FFNoTw: no twiddles (inner blocks).
// mid(x, y): the pointwise step of the middle round, x of the array being
// transformed, y of the other spectrum (or x itself for a square).
struct FFNoTw {
    static constexpr bool on = false;
    static constexpr bool lds_tables = false;  // the root tables read from HBM
    template <int NL>
    __device__ void group(int, int, uint32_t *) const {}
    __device__ uint32_t one(int, int) const { return 0; }
    __device__ uint32_t mid(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const { return ff_mont(x, y, p, pinv); }
};

// w v mod p for a weight w of 0, 1 or 2.
__device__ __forceinline__ uint32_t ff_w012(uint32_t v, int w, uint32_t p) {
    const uint32_t v2 = ff_add(v, v, p);
    return w == 0 ? 0u : (w == 1 ? v : v2);
}

// A weighted pair with one output: u = w0 x^2 + w1 xy + w2 y^2 (weights 0..2)
// as x (w0 x + w1 y) + w2 y^2: one product when w2 = 0 (F(2k) = a (a + 2b)),
// two otherwise.
struct FFPairMid : FFNoTw {
    int w0, w1, w2;
    __device__ uint32_t mid(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const {
        const uint32_t u = ff_mont(x, ff_add(ff_w012(x, w0, p), ff_w012(y, w1, p), p), p, pinv);
        return w2 ? ff_add(u, ff_w012(ff_mont(y, y, p, pinv), w2, p), p) : u;
    }
};

// nl-bit reversal of j, folded at compile time in unrolled loops.
__host__ __device__ constexpr int ff_rev_c(int j, int nl) {
    return nl == 0 ? 0 : ((j & 1) << (nl - 1)) | ff_rev_c(j >> 1, nl - 1);
}

// T[k] = base * g^k for k < 2^NL (Montgomery form), in a tree of depth NL.
template <int NL>
__device__ __forceinline__ void ff_geo(uint32_t base, uint32_t g, uint32_t *T, uint32_t p, uint32_t pinv) {
    T[0] = base;
    uint32_t gk = g;  // g^(2^i)
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const int h = 1 << i;
        const uint32_t gp = gk * pinv;
#pragma unroll
        for (int k = 0; k < h; k++) T[h + k] = ff_mont_w(T[k], gk, gp, p);
        if (i + 1 < NL) gk = ff_mont(gk, gk, p, pinv);
    }
}

template <int NL, int KIND, bool FROM_G, bool TO_G, class Ld, class St, class Gt = FFNoTw>
__device__ __forceinline__ void ff_round_io(const FFTile &t, int b_lo, const uint32_t *rt, const uint32_t *rtp,
                                            const uint32_t *irt, const uint32_t *irtp, const uint32_t *B,
                                            uint32_t p, uint32_t pinv, Ld ld, St st, const Gt &gt = Gt()) {
    const int G = 1 << NL;
    const int per_sub = (1 << (t.rho - NL)) << t.logC;
    const int ngroups = per_sub * t.nsub;
    // The group's rows differ from its first only in row bits [b_lo, b_lo +
    // NL), disjoint from every other bit of the position, so the padded LDS
    // address of row j splits into the first row's address plus an offset
    // that is the same for every group of the round (uniform: kept in
    // scalar registers, one addition per element instead of the address).
    int off[G];
#pragma unroll
    for (int j = 0; j < G; j++) {
        const int J = j << b_lo;
        off[j] = J * t.ld + (t.pad ? (J >> 4) : 0);
    }
    for (int gi = ff_tile_tid(t); gi < ngroups; gi += ff_tile_nthr(t)) {
        const int sub = gi / per_sub;
        const int g = gi - sub * per_sub;
        const int c = g & (t.C - 1);
        const int rest = g >> t.logC;
        const int low = rest & ((1 << b_lo) - 1);
        const int base = low | ((rest >> b_lo) << (b_lo + NL));
        const int pos0 = sub << t.rho;
        const int a0 = ff_pos_addr(t, pos0 + base, c);
        uint32_t v[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
            // base has no bits in [b_lo, b_lo + NL): the sum is the OR
            const int pos = pos0 + base + (j << b_lo);
            if constexpr (FROM_G && ff_rows_io<Ld>::value) {
                // row j at the group's first byte offset plus j uniform strides
                v[j] = ld.ldb(ld.row0(pos0 + base, c) + (uint32_t)j * (ld.rstep() << b_lo));
            } else {
                v[j] = FROM_G ? ld(pos, c) : t.s[a0 + off[j]];
            }
        }
        if constexpr (Gt::on && KIND == FF_R_DIT && FROM_G) {
            // b_lo = 0 here: rows pos0 + base + j
            uint32_t tw[G];
            gt.template group<NL>(pos0 + base, c, tw);
#pragma unroll
            for (int j = 0; j < G; j++) v[j] = ff_mont(v[j], tw[j], p, pinv);
        }
        // the rounds at row bit 0 (the middle round, the outer passes' HBM
        // rounds) take the constant-twiddle form
        if (KIND == FF_R_DIF) {
            if (b_lo == 0) ff_dif_group<NL, Gt::lds_tables, true>(v, 0, 0, rt, rtp, p, pinv);
            else ff_dif_group<NL, Gt::lds_tables>(v, low, b_lo, rt, rtp, p, pinv);
        } else if (KIND == FF_R_DIT) {
            if (b_lo == 0) ff_dit_group<NL, Gt::lds_tables, true>(v, 0, 0, irt, irtp, p, pinv);
            else ff_dit_group<NL, Gt::lds_tables>(v, low, b_lo, irt, irtp, p, pinv);
        } else {
            // FF_R_MID is only called at b_lo = 0
            ff_dif_group<NL, Gt::lds_tables, true>(v, 0, 0, rt, rtp, p, pinv);
#pragma unroll
            for (int j = 0; j < G; j++) v[j] = gt.mid(v[j], B ? B[a0 + off[j]] : v[j], p, pinv);
            ff_dit_group<NL, Gt::lds_tables, true>(v, 0, 0, irt, irtp, p, pinv);
        }
        if constexpr (Gt::on && KIND == FF_R_DIF && TO_G) {
            uint32_t tw[G];
            gt.template group<NL>(pos0 + base, c, tw);
#pragma unroll
            for (int j = 0; j < G; j++) v[j] = ff_mont(v[j], tw[j], p, pinv);
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
            const int pos = pos0 + base + (j << b_lo);
            if constexpr (TO_G && ff_rows_io<St>::value) {
                st.stb(st.row0(pos0 + base, c) + (uint32_t)j * (st.rstep() << b_lo), v[j]);
            } else {
                if (TO_G) st(pos, c, v[j]);
                else t.s[a0 + off[j]] = v[j];
            }
        }
    }
}

// Dispatch of one round by its level count and ends.
template <int KIND, class Ld, class St, class Gt = FFNoTw>
__device__ void ff_round_any(int nl, bool from_g, bool to_g, const FFTile &t, int b_lo, const uint32_t *rt,
                             const uint32_t *rtp, const uint32_t *irt, const uint32_t *irtp, const uint32_t *B,
                             uint32_t p, uint32_t pinv, Ld ld, St st, const Gt &gt = Gt()) {
#define FF_RCASE(N)                                                                                                \
    case N:                                                                                                        \
        if (from_g && to_g) ff_round_io<N, KIND, true, true>(t, b_lo, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt); \
        else if (from_g) ff_round_io<N, KIND, true, false>(t, b_lo, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt);   \
        else if (to_g) ff_round_io<N, KIND, false, true>(t, b_lo, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt);     \
        else ff_round_io<N, KIND, false, false>(t, b_lo, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt);              \
        break;
    switch (nl) {
        FF_RCASE(4)
        FF_RCASE(3)
        FF_RCASE(2)
        FF_RCASE(1)
    }
#undef FF_RCASE
}

// Round sizes: rounds of rl levels (4 by default) and one of the remainder
// (rho % rl, or rl). DIF rounds cover the row bits from the top and DIT
// rounds from the bottom, so the remainder takes the top bits in both: the
// round at the lowest bits (the forward's last, the inverse's first) is the
// one that meets HBM in the outer passes and applies their twiddles, two
// table roots and a geometric sequence per group, which 16-element groups
// share far better than 2-element ones.

// Forward (DIF) over the tile; the first round reads through ld when
// from_g, the last writes through st when to_g.
// A tile of one-row sub-transforms (rho = 0, under a radix-3 layer) has no
// levels: its elements only pass through.
template <class Ld, class St, class Gt = FFNoTw>
__device__ void ff_pass_through(const FFTile &t, bool from_g, bool to_g, Ld ld, St st, const Gt &gt = Gt(),
                                uint32_t p = 0, uint32_t pinv = 0) {
    const int n = t.nsub << t.logC;
    for (int i = ff_tile_tid(t); i < n; i += ff_tile_nthr(t)) {
        const int pos = i >> t.logC, c = i & (t.C - 1);
        uint32_t v = from_g ? ld(pos, c) : t.s[ff_pos_addr(t, pos, c)];
        if constexpr (Gt::on) {
            if (from_g || to_g) v = ff_mont(v, gt.one(pos, c), p, pinv);
        }
        if (to_g) st(pos, c, v);
        else t.s[ff_pos_addr(t, pos, c)] = v;
    }
}

// rl: levels per round (1..4); short tiles take fewer, so that more groups
// (2^(rho - rl)) keep the threads busy.
template <class Ld, class St, class Gt = FFNoTw>
__device__ void ff_forward_io(const FFTile &t, const uint32_t *rt, const uint32_t *rtp, uint32_t p, uint32_t pinv,
                              bool from_g, bool to_g, Ld ld, St st, const Gt &gt = Gt(), int rl = 4) {
    if (t.rho == 0) {
        ff_pass_through(t, from_g, to_g, ld, st, gt, p, pinv);
        return;
    }
    const int R = (t.rho + rl - 1) / rl;
    int b = t.rho;
    for (int r = 0; r < R; r++) {
        const int nl = r == 0 ? (t.rho % rl == 0 ? rl : t.rho % rl) : rl;
        b -= nl;
        if (r) __syncthreads();
        ff_round_any<FF_R_DIF>(nl, from_g && r == 0, to_g && r == R - 1, t, b, rt, rtp, nullptr, nullptr, nullptr,
                               p, pinv, ld, st, gt);
    }
}

// Inverse (DIT) over the tile: the rounds of ff_forward_io in reverse.
template <class Ld, class St, class Gt = FFNoTw>
__device__ void ff_inverse_io(const FFTile &t, const uint32_t *irt, const uint32_t *irtp, uint32_t p, uint32_t pinv,
                              bool from_g, bool to_g, Ld ld, St st, const Gt &gt = Gt(), int rl = 4) {
    if (t.rho == 0) {
        ff_pass_through(t, from_g, to_g, ld, st, gt, p, pinv);
        return;
    }
    const int R = (t.rho + rl - 1) / rl;
    int b = 0;
    for (int r = 0; r < R; r++) {
        const int nl = r == R - 1 ? (t.rho % rl == 0 ? rl : t.rho % rl) : rl;
        if (r) __syncthreads();
        ff_round_any<FF_R_DIT>(nl, from_g && r == 0, to_g && r == R - 1, t, b, nullptr, nullptr, irt, irtp, nullptr,
                               p, pinv, ld, st, gt);
        b += nl;
    }
}

// Whole convolution of one LDS-resident block: forward of A (first round
// through ld when from_g), pointwise with B (LDS, already transformed) or a
// square, inverse, the last round through st when to_g.
template <class Ld, class St, class Gt = FFNoTw>
// rl: levels per round (1..4). Fewer levels per round means more, shorter
// rounds with more threads busy: lower latency for small tiles in one
// workgroup; four is the throughput choice.
__device__ void ff_conv_io(const FFTile &t, const uint32_t *rt, const uint32_t *rtp, const uint32_t *irt,
                           const uint32_t *irtp, const uint32_t *B, uint32_t p, uint32_t pinv, bool from_g, bool to_g,
                           Ld ld, St st, const Gt &gt = Gt(), int rl = 4) {
    const int R = (t.rho + rl - 1) / rl;
    int b = t.rho;
    for (int r = 0; r < R - 1; r++) {
        b -= rl;
        if (r) __syncthreads();
        ff_round_any<FF_R_DIF>(rl, from_g && r == 0, false, t, b, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, st,
                               gt);
    }
    const int nl0 = t.rho % rl == 0 ? rl : t.rho % rl;
    if (R > 1) __syncthreads();
    ff_round_any<FF_R_MID>(nl0, from_g && R == 1, to_g && R == 1, t, 0, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt);
    b = nl0;
    for (int r = 1; r < R; r++) {
        __syncthreads();
        ff_round_any<FF_R_DIT>(rl, false, to_g && r == R - 1, t, b, nullptr, nullptr, irt, irtp, nullptr, p, pinv, ld,
                               st, gt);
        b += rl;
    }
}

// The same for a block of 2^12 words (a one-column tile, rho = 12, built by
// the caller from literals): rounds of four levels at row bits 8, 4, 0, 4, 8
// written out, so that with everything inlined each round's LDS offsets and
// twiddle offsets are compile-time constants (immediate offsets of the LDS
// and buffer instructions). The first round reads through ld, the last
// writes through st.
template <class Ld, class St, class Gt = FFNoTw>
__device__ __forceinline__ void ff_conv12_io(const FFTile &t, const uint32_t *rt, const uint32_t *rtp,
                                             const uint32_t *irt, const uint32_t *irtp, const uint32_t *B,
                                             uint32_t p, uint32_t pinv, Ld ld, St st, const Gt &gt = Gt()) {
    ff_round_io<4, FF_R_DIF, true, false>(t, 8, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, st, gt);
    __syncthreads();
    ff_round_io<4, FF_R_DIF, false, false>(t, 4, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, st, gt);
    __syncthreads();
    ff_round_io<4, FF_R_MID, false, false>(t, 0, rt, rtp, irt, irtp, B, p, pinv, ld, st, gt);
    __syncthreads();
    ff_round_io<4, FF_R_DIT, false, false>(t, 4, nullptr, nullptr, irt, irtp, nullptr, p, pinv, ld, st, gt);
    __syncthreads();
    ff_round_io<4, FF_R_DIT, false, true>(t, 8, nullptr, nullptr, irt, irtp, nullptr, p, pinv, ld, st, gt);
}

// ff_forward_io / ff_inverse_io with the row count 2^RHO and the round size
// RL as template parameters: round R covers literal row bits, so that with
// a tile built from literals (and everything inlined) each round's LDS and
// twiddle offsets are compile-time constants.
template <int RHO, int RL, int R, bool FROM_G, bool TO_G, class Ld, class St, class Gt>
__device__ __forceinline__ void ff_forward_fixed(const FFTile &t, const uint32_t *rt, const uint32_t *rtp, uint32_t p,
                                                 uint32_t pinv, Ld ld, St st, const Gt &gt) {
    constexpr int NR = (RHO + RL - 1) / RL;
    constexpr int NL0 = RHO % RL ? RHO % RL : RL;  // the remainder takes the top bits
    constexpr int NL = R == 0 ? NL0 : RL;
    constexpr int B = RHO - NL0 - R * RL;
    if constexpr (R > 0) __syncthreads();
    ff_round_io<NL, FF_R_DIF, FROM_G && R == 0, TO_G && R == NR - 1>(t, B, rt, rtp, nullptr, nullptr, nullptr, p,
                                                                      pinv, ld, st, gt);
    if constexpr (R + 1 < NR) ff_forward_fixed<RHO, RL, R + 1, FROM_G, TO_G>(t, rt, rtp, p, pinv, ld, st, gt);
}

template <int RHO, int RL, int R, bool FROM_G, bool TO_G, class Ld, class St, class Gt>
__device__ __forceinline__ void ff_inverse_fixed(const FFTile &t, const uint32_t *irt, const uint32_t *irtp,
                                                 uint32_t p, uint32_t pinv, Ld ld, St st, const Gt &gt) {
    constexpr int NR = (RHO + RL - 1) / RL;
    constexpr int NL0 = RHO % RL ? RHO % RL : RL;  // the last round takes the top bits
    constexpr int NL = R == NR - 1 ? NL0 : RL;
    constexpr int B = R * RL;
    if constexpr (R > 0) __syncthreads();
    ff_round_io<NL, FF_R_DIT, FROM_G && R == 0, TO_G && R == NR - 1>(t, B, nullptr, nullptr, irt, irtp, nullptr, p,
                                                                      pinv, ld, st, gt);
    if constexpr (R + 1 < NR) ff_inverse_fixed<RHO, RL, R + 1, FROM_G, TO_G>(t, irt, irtp, p, pinv, ld, st, gt);
}

// Forward transform of a 2^12-word block into its LDS tile (as ff_conv12_io).
template <class Ld>
__device__ __forceinline__ void ff_forward12_io(const FFTile &t, const uint32_t *rt, const uint32_t *rtp, uint32_t p,
                                                uint32_t pinv, Ld ld) {
    auto none = [](int, int, uint32_t) {};
    ff_round_io<4, FF_R_DIF, true, false>(t, 8, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, none);
    __syncthreads();
    ff_round_io<4, FF_R_DIF, false, false>(t, 4, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, none);
    __syncthreads();
    ff_round_io<4, FF_R_DIF, false, false>(t, 0, rt, rtp, nullptr, nullptr, nullptr, p, pinv, ld, none);
}

// Radix-3 layer over a tile of nsub = 3 stacked thirds (ntt_radix3.c:8-22,
// reference): forward, for n1 < M with a_r = (third r, row n1):
//   y0 = a0 + a1 + a2, d = w3 (a1 - a2),
//   y1 = (a0 - a2 + d) wR^n1, y2 = (a0 - a1 - d) wR^2n1
// Inverse (roots inverted): t1 = z1 wR^-n1, t2 = z2 wR^-2n1,
//   x0 = z0 + t1 + t2, d = w3^-1 (t1 - t2), x1 = z0 - t2 + d, x2 = z0 - t1 - d.
// rstep = FF_NBASE / R, so wR^e = w_NBASE^(e * rstep).
template <bool INV>
__device__ void ff_radix3_layer(const FFTile &t, const FFTables *T, int pi, const FFPrime &P, uint32_t rstep) {
    const uint32_t p = P.p;
    const int M = 1 << t.rho;
    const int n = M << t.logC;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int c = i & (t.C - 1);
        const int n1 = i >> t.logC;
        const int a0i = ff_tile_addr(t, 0, n1, c), a1i = ff_tile_addr(t, 1, n1, c), a2i = ff_tile_addr(t, 2, n1, c);
        uint32_t e1 = (uint32_t)n1 * rstep;
        uint32_t e2 = 2u * (uint32_t)n1 * rstep;
        if (INV) {
            e1 = e1 ? FF_NBASE - e1 : 0;
            e2 = e2 ? FF_NBASE - e2 : 0;
        }
        const uint32_t tw1 = ff_root(T, pi, e1, p, P.pinv), tw2 = ff_root(T, pi, e2, p, P.pinv);
        const uint32_t a0 = t.s[a0i], a1 = t.s[a1i], a2 = t.s[a2i];
        if (!INV) {
            const uint32_t y0 = ff_add(ff_add(a0, a1, p), a2, p);
            const uint32_t d = ff_mont_w_diff(a1, a2, P.w3, P.w3p, p);
            const uint32_t y1 = ff_add(ff_sub(a0, a2, p), d, p);
            const uint32_t y2 = ff_sub(ff_sub(a0, a1, p), d, p);
            t.s[a0i] = y0;
            t.s[a1i] = ff_mont(y1, tw1, p, P.pinv);
            t.s[a2i] = ff_mont(y2, tw2, p, P.pinv);
        } else {
            const uint32_t t1 = ff_mont(a1, tw1, p, P.pinv), t2 = ff_mont(a2, tw2, p, P.pinv);
            const uint32_t d = ff_mont_w_diff(t1, t2, P.iw3, P.iw3p, p);
            t.s[a0i] = ff_add(ff_add(a0, t1, p), t2, p);
            t.s[a1i] = ff_add(ff_sub(a0, t2, p), d, p);
            t.s[a2i] = ff_sub(ff_sub(a0, t1, p), d, p);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ int ff_bitrev(int x, int bits) {
    return bits ? (int)(__brev((unsigned)x) >> (32 - bits)) : 0;
}

// Frequency held by tile row position pos after the forward row transform:
// rev(v) for a power-of-2 R; 3 rev(v) + r for R = 3M (pos = r M + v).
__device__ __forceinline__ uint32_t ff_row_freq(int pos, int rho, int r3) {
    if (!r3) return (uint32_t)ff_bitrev(pos, rho);
    const int r = pos >> rho, v = pos & ((1 << rho) - 1);
    return 3u * (uint32_t)ff_bitrev(v, rho) + (uint32_t)r;
}

// b-bit coefficient k of an operand stored as nw 32-bit words (the limbs,
// little-endian), reduced mod p; zero past the words (src_load8,
// ntt_kernels.c:40-74, reference).
template <class Ptr>
__device__ __forceinline__ uint32_t ff_load_coef(Ptr w, uint32_t nw, uint64_t k, int bw, const FFPrime &P) {
    const uint64_t bit = k * (uint64_t)bw;
    const uint64_t wi = bit >> 5;
    const int sh = (int)(bit & 31);
    if (wi >= nw) return 0;
    uint64_t lo = w[wi];
    if (wi + 1 < nw) lo |= (uint64_t)w[wi + 1] << 32;
    uint64_t v = lo >> sh;
    if (sh + bw > 64 && wi + 2 < nw) v |= (uint64_t)w[wi + 2] << (64 - sh);
    v &= (bw >= 64 ? ~0ull : ((1ull << bw) - 1));
    const uint32_t lo32 = ff_red32((uint32_t)v, P.p);
    const uint32_t hi = (uint32_t)(v >> 32);
    if (!hi) return lo32;
    return ff_add(lo32, ff_mont(hi, P.r2, P.p, P.pinv), P.p);
}
