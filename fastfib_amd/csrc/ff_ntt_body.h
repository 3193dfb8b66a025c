// Device bodies of the transform passes (outer passes, inner blocks, the
// pair kernel), shared by the list and grid kernels (ff_ntt.hip) and the
// batch work-queue kernel (ff_queue.hip).
#pragma once

#include "ff_internal.h"
#include "ff_ntt_dev.h"


#define FF_THREADS 256

// Four-step twiddles of an outer pass: element (row position pos, tile
// column c), column col = c0 + c, is scaled by w^(col * f(pos) * rootstep),
// f the row's frequency (ff_row_freq), w the 3 * 2^25-th root; INV: by the
// inverse root, times the CRT scale on the top pass (scm); forward top pass
// with scm != 0: times scm * 2^-32 (2^32 mod p for scm = R^2 mod p, undoing
// the 2^-32 of coefficients unpacked by FFCoefSrc in Montgomery-reduced
// form; the rounds between are linear). group() serves the 2^NL
// consecutive rows of one butterfly group: their frequencies are
// f0 + step * rev(j), so the twiddles are base * g^rev(j) from two table
// roots (base = w^(col f0 rootstep), g = w^(col step rootstep)).
template <bool INV>
struct FFOuterTw {
    static constexpr bool on = true;
    static constexpr bool lds_tables = false;
    // outer passes have no middle round
    __device__ uint32_t mid(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const { return ff_mont(x, y, p, pinv); }
    const FFTables *T;
    int pi, rho, r3;
    bool top;
    uint32_t c0, rootstep, scm, p, pinv;

    __device__ __forceinline__ uint32_t root(uint32_t e) const {
        if (INV) e = e ? FF_NBASE - e : 0;
        return ff_root(T, pi, e, p, pinv);
    }
    __device__ __forceinline__ uint32_t one(int pos, int c) const {
        uint32_t tw = root((c0 + c) * ff_row_freq(pos, rho, r3) * rootstep);
        if (top && (INV || scm)) tw = ff_mont(tw, scm, p, pinv);
        return tw;
    }
    template <int NL>
    __device__ __forceinline__ void group(int pos, int c, uint32_t *tw) const {
        const uint32_t cs = (c0 + c) * rootstep;
        const uint32_t step = (r3 ? 3u : 1u) << (rho - NL);
        uint32_t base = root(cs * ff_row_freq(pos, rho, r3));
        if (top && (INV || scm)) base = ff_mont(base, scm, p, pinv);
        uint32_t pw[1 << NL];
        ff_geo<NL>(base, root(cs * step), pw, p, pinv);
#pragma unroll
        for (int j = 0; j < (1 << NL); j++) tw[j] = pw[ff_rev_c(j, NL)];
    }
};

// Row access to a block of n words of a transform array in HBM: tile
// element (pos, c) is word pos * Sb + c0 + c (< n). Reads and writes go
// through a buffer descriptor of the block (scalar registers: a lane
// supplies only its 32-bit byte offset), or with COH through the coherent
// accesses of ff_ld / ff_st. A row accessor for the rounds (ff_rows_io).
template <bool COH>
struct FFRows {
    static constexpr bool rows = true;
    ff_gptr<uint32_t> p;
    __amdgpu_buffer_rsrc_t rs;
    uint32_t Sb, c0;
    __device__ __forceinline__ FFRows(ff_gptr<uint32_t> p_, uint32_t n, uint32_t Sb_, uint32_t c0_)
        : p(p_), rs(__builtin_amdgcn_make_buffer_rsrc((void *)(uint32_t *)p_, (short)0, (int)(n * 4u), 0x00020000)),
          Sb(Sb_), c0(c0_) {}
    __device__ __forceinline__ uint32_t row0(int pos, int c) const { return ((uint32_t)pos * Sb + c0 + (uint32_t)c) << 2; }
    __device__ __forceinline__ uint32_t rstep() const { return Sb << 2; }
    __device__ __forceinline__ uint32_t ldb(uint32_t ob) const {
        if (COH) return ff_ld<COH>(p + (ob >> 2));
        return __builtin_amdgcn_raw_buffer_load_b32(rs, ob, 0, 0);
    }
    __device__ __forceinline__ void stb(uint32_t ob, uint32_t v) const {
        if (COH) ff_st<COH>(p + (ob >> 2), v);
        else __builtin_amdgcn_raw_buffer_store_b32(v, rs, ob, 0, 0);
    }
    __device__ __forceinline__ uint32_t operator()(int pos, int c) const { return ldb(row0(pos, c)); }
    __device__ __forceinline__ void operator()(int pos, int c, uint32_t v) const { stb(row0(pos, c), v); }
};

// The bw-bit coefficients of an operand of nw 32-bit words at src, reduced
// mod p (src_load8, ntt_kernels.c:40-74, reference): element (row pos,
// column c) of a forward top pass is coefficient pos * Sb + c0 + c.
// Coefficient positions fit 32 bits (L <= 3 * 2^25 coefficients of at most
// 35 bits). The three words holding a coefficient are read by buffer loads,
// which return zero past the operand's last word, so coefficients past the
// operand need no bounds test. A coefficient wider than 32 bits, hi * 2^32 +
// lo (hi < 8), is taken in Montgomery-reduced form instead, its value times
// 2^-32: with m = lo * p^-1 mod 2^32, hi * 2^32 + lo - m p is divisible by
// 2^32 with quotient hi - hi(m p) in (-p, 8), one correction from [0, p).
// The forward top pass's twiddles then carry the factor 2^32 back
// (FFOuterTw with scm = R^2 mod p).
struct FFCoefSrc {
    __amdgpu_buffer_rsrc_t rs;
    uint32_t nw, Sb, c0, lomask, himask, p, pinv;
    int bw;
    __device__ __forceinline__ FFCoefSrc(const uint32_t *src, uint32_t nw_, uint32_t Sb_, uint32_t c0_, int bw_,
                                         const FFPrime &P)
        : rs(__builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, (int)(nw_ * 4u), 0x00020000)),
          nw(nw_), Sb(Sb_), c0(c0_), lomask(bw_ >= 32 ? ~0u : ((1u << bw_) - 1)),
          himask(bw_ > 32 ? ((1u << (bw_ - 32)) - 1) : 0u), p(P.p), pinv(P.pinv), bw(bw_) {}
    __device__ __forceinline__ uint32_t bit(int pos, int c) const {
        return ((uint32_t)pos * Sb + c0 + (uint32_t)c) * (uint32_t)bw;
    }
    __device__ __forceinline__ void words(uint32_t bit, uint32_t *w) const {
        const uint32_t off = (bit >> 5) << 2;
        w[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
        w[1] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4, 0, 0);
        w[2] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 8, 0, 0);
    }
    // WIDE = (bw > 32)
    template <bool WIDE>
    __device__ __forceinline__ uint32_t value(const uint32_t *w, uint32_t bit) const {
        const uint32_t sh = bit & 31;
        const uint32_t lo = __builtin_amdgcn_alignbit(w[1], w[0], sh) & lomask;
        if (WIDE) {
            // (hi * 2^32 + lo) * 2^-32 mod p
            const uint32_t hi = __builtin_amdgcn_alignbit(w[2], w[1], sh) & himask;
            return ff_mont_fix(hi, __umulhi(lo * pinv, p), p);
        }
        uint32_t r = lo, q = r - p;
        r = q < r ? q : r;
        q = r - p;
        return q < r ? q : r;  // lo < 2^32 < 3p
    }
};

// The tile of a forward top pass from the operand's limbs (FFCoefSrc): each
// thread issues the loads of UB elements before it uses the first of them.
// Batches of 4 skip the zero rows at a finer grain; the 2^10-row tiles keep
// 8 loads in flight (measured: 4 is faster for tiles of at most 2^9 rows,
// 8 for 2^10).
#define FF_UNPACK_B 8
template <bool WIDE, int UB = FF_UNPACK_B>
__device__ __forceinline__ void ff_unpack_tile_w(const FFTile &t, const FFCoefSrc &cs) {
    const int n = (t.nsub << t.rho) << t.logC;
    const int tid = ff_tile_tid(t), nthr = ff_tile_nthr(t);
    if (t.C <= nthr && !t.pad) {
        // nthr is a multiple of C (both powers of 2): a thread keeps its
        // column and steps down rs rows per element, so the bit position and
        // the LDS address advance by uniform strides
        const int rs = nthr >> t.logC, rows = n >> t.logC;
        const uint32_t dbit = (uint32_t)rs * cs.Sb * (uint32_t)cs.bw;
        const int dadr = rs * t.ld;
        int row = tid >> t.logC;
        const int col = tid & (t.C - 1);
        uint32_t bit = cs.bit(row, col);
        int adr = row * t.ld + col;
        // rows from rows_nz on hold coefficients pos * Sb + c0 + c >= ncoef,
        // past the operand's last bit: zero (a product's operands fill at
        // most half its transform, so the tile's upper half is usually
        // zero). A batch whose first row (k0 rs, that of thread 0) is past
        // them, uniform across the workgroup, stores zeros with no loads.
        const uint32_t ncoef = (uint32_t)(((uint64_t)cs.nw * 32u + (uint32_t)cs.bw - 1) / (uint32_t)cs.bw);
        const uint32_t rows_nz = (ncoef + cs.Sb - 1) / cs.Sb;
        for (int k0 = 0; k0 * nthr < n; k0 += UB) {
            if ((uint32_t)(k0 * rs) >= rows_nz) {
#pragma unroll
                for (int e = 0; e < UB; e++)
                    if (row + e * rs < rows) t.s[adr + e * dadr] = 0;
                bit += UB * dbit;
                adr += UB * dadr;
                row += UB * rs;
                continue;
            }
            uint32_t w[UB][3];
#pragma unroll
            for (int e = 0; e < UB; e++) cs.words(bit + (uint32_t)e * dbit, w[e]);
#pragma unroll
            for (int e = 0; e < UB; e++) {
                const uint32_t r = cs.value<WIDE>(w[e], bit + (uint32_t)e * dbit);
                if (row + e * rs < rows) t.s[adr + e * dadr] = r;
            }
            bit += UB * dbit;
            adr += UB * dadr;
            row += UB * rs;
        }
        return;
    }
    for (int k0 = 0; k0 * nthr < n; k0 += UB) {
        uint32_t w[UB][3], b[UB];
#pragma unroll
        for (int e = 0; e < UB; e++) {
            const int i = tid + (k0 + e) * nthr;
            b[e] = cs.bit(i >> t.logC, i & (t.C - 1));
            // past the tile (i >= n) the loads are harmless and their values unused
            cs.words(b[e], w[e]);
        }
#pragma unroll
        for (int e = 0; e < UB; e++) {
            const int i = tid + (k0 + e) * nthr;
            const uint32_t r = cs.value<WIDE>(w[e], b[e]);
            if (i < n) t.s[ff_pos_addr(t, i >> t.logC, i & (t.C - 1))] = r;
        }
    }
}

template <int UB = FF_UNPACK_B>
__device__ __forceinline__ void ff_unpack_tile(const FFTile &t, const FFCoefSrc &cs) {
    if (cs.bw > 32) ff_unpack_tile_w<true, UB>(t, cs);
    else ff_unpack_tile_w<false, UB>(t, cs);
}

// Tile shapes of the outer passes with a fixed-geometry path (rows 2^rho,
// times 3 with r3, 2^logC columns, row stride ld words, rounds of rl
// levels), as the plans of products of 2^17 .. 3 * 2^24 words per prime
// give them: top passes of 2^17 .. 2^21 words (shapes 1-5, fast doubling's
// tiles), of 2^22 .. 2^25 (6-8) and of 3 * 2^21 .. 3 * 2^23 (9, 10, with
// the radix-3 layer); second passes of 3 * 2^21 .. 3 * 2^24 words (6, 7);
// top passes of batch items with 4096-word tiles: 3 * 2^17 .. 3 * 2^20
// words with the radix-3 layer (11, 13, 14), 2^17 .. 2^19 words (12).
// Shape 0: the general path.
struct FFShape {
    int rho, logC, ld, rl, r3;
};
#define FF_NSHAPES 14
__host__ __device__ constexpr FFShape ff_shape(int sh) {
    return sh == 1    ? FFShape{5, 6, 64, 3, 0}
           : sh == 2  ? FFShape{6, 5, 33, 3, 0}
           : sh == 3  ? FFShape{7, 4, 17, 4, 0}
           : sh == 4  ? FFShape{8, 4, 17, 4, 0}
           : sh == 5  ? FFShape{9, 4, 17, 3, 0}
           : sh == 6  ? FFShape{6, 6, 64, 3, 0}
           : sh == 7  ? FFShape{5, 7, 128, 3, 0}
           : sh == 8  ? FFShape{10, 4, 16, 4, 0}
           : sh == 9  ? FFShape{5, 5, 33, 3, 1}
           : sh == 10 ? FFShape{4, 6, 64, 4, 1}
           : sh == 11 ? FFShape{6, 4, 17, 3, 1}
           : sh == 12 ? FFShape{7, 5, 33, 4, 0}
           : sh == 13 ? FFShape{7, 4, 17, 4, 1}
           : sh == 14 ? FFShape{8, 4, 17, 4, 1}
                      : FFShape{0, 0, 0, 0, 0};
}

// Outer pass over one tile of R rows x C columns (tile index `tile` among
// the pass's nblocks * tpb) of array X, prime pi. COH: the array is read and
// written by coherent accesses (ff_ld / ff_st). SH > 0: the pass has tile
// shape ff_shape(SH), its rounds at literal positions.
template <bool INV, bool TOP, bool R3, bool COH = false, int SH = 0>
__device__ __forceinline__ void outer_body(const FFPass &ps, const FFXform &X, int pi, uint32_t tile,
                                           const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const uint32_t blk = tile / ps.tpb, ct = tile - blk * ps.tpb;
    const uint32_t c0 = ct << ps.logC;
    const ff_gptr<uint32_t> data = ff_g(X.data) + (size_t)pi * ps.L + (size_t)blk * ps.Lb;
    FFTile t{lds, ps.ld, ps.C, ps.logC, ps.rho, R3 ? 3 : 1, 0};
    const int n = (int)ps.R << ps.logC;

    const uint32_t pinv = P.pinv;
    // inverse top pass: the CRT scale; forward top pass at widths above 32
    // bits: R^2 mod p, for the coefficients' factor 2^-32 (FFCoefSrc)
    const uint32_t scm = TOP ? (INV ? ps.sc[pi] : (ps.bw > 32 ? P.r2 : 0u)) : 0u;
    const bool top = TOP;
    const int bw = ps.bw;
    const uint32_t Sb = ps.Sb, rootstep = ps.rootstep;
    const int rho = ps.rho, r3 = R3;
    const uint32_t nw = X.nw;
    // element (row position pos, tile column c) <-> data[pos * Sb + c0 + c]
    // (a top pass reads the limbs instead, through FFCoefSrc)
    // (offsets inside the block fit 32 bits: Lb <= 3 * 2^25); the twiddles
    // are applied by the rounds (tw), not by these accessors
    const FFRows<COH> rows(data, ps.Lb, Sb, c0);
    const FFRows<COH> &load_fwd = rows, &store_fwd = rows, &load_inv = rows, &store_inv = rows;
    const FFOuterTw<INV> tw{T, pi, rho, r3, top, c0, rootstep, scm, p, pinv};
    auto lds_load = [&](int pos, int c) -> uint32_t { return lds[ff_pos_addr(t, pos, c)]; };
    auto lds_store = [&](int pos, int c, uint32_t v) { lds[ff_pos_addr(t, pos, c)] = v; };

    if constexpr (SH > 0) {
        constexpr FFShape S = ff_shape(SH);
        static_assert((S.r3 != 0) == R3 && (!R3 || TOP), "shape and kernel agree");
        const FFTile tf{lds, S.ld, 1 << S.logC, S.logC, S.rho, R3 ? 3 : 1, 0};
        const FFOuterTw<INV> twf{T, pi, S.rho, R3, top, c0, rootstep, scm, p, pinv};
        if constexpr (!INV) {
            if constexpr (TOP) {
                ff_unpack_tile<S.rho >= 10 ? 8 : 4>(tf, FFCoefSrc(X.src, nw, Sb, c0, bw, P));
                if constexpr (R3) ff_radix3_layer<false>(tf, T, pi, P, ps.rstep3);
                else __syncthreads();
                ff_forward_fixed<S.rho, S.rl, 0, false, true>(tf, T->rt[pi], T->rtp[pi], p, pinv, lds_load, store_fwd,
                                                             twf);
            } else {
                ff_forward_fixed<S.rho, S.rl, 0, true, true>(tf, T->rt[pi], T->rtp[pi], p, pinv, load_fwd, store_fwd,
                                                            twf);
            }
        } else if constexpr (!R3) {
            ff_inverse_fixed<S.rho, S.rl, 0, true, true>(tf, T->irt[pi], T->irtp[pi], p, pinv, load_inv, store_inv,
                                                        twf);
        } else {
            ff_inverse_fixed<S.rho, S.rl, 0, true, false>(tf, T->irt[pi], T->irtp[pi], p, pinv, load_inv, lds_store,
                                                         twf);
            ff_radix3_layer<true>(tf, T, pi, P, ps.rstep3);
            constexpr int NT = (3 << S.rho) << S.logC;
            for (int i = threadIdx.x; i < NT; i += blockDim.x)
                store_inv(i >> S.logC, i & ((1 << S.logC) - 1), tf.s[ff_pos_addr(tf, i >> S.logC, i & ((1 << S.logC) - 1))]);
        }
        return;
    }
    if (!INV) {
        if (!R3 && !TOP) {
            ff_forward_io(t, T->rt[pi], T->rtp[pi], p, pinv, true, true, load_fwd, store_fwd, tw, ps.rl);
        } else {
            // the coefficients are unpacked into the LDS tile (the radix-3
            // layer works there). A first round reading them straight into
            // registers through FFCoefSrc measured 12% slower.
            ff_unpack_tile(t, FFCoefSrc(X.src, nw, Sb, c0, bw, P));
            if (R3) ff_radix3_layer<false>(t, T, pi, P, ps.rstep3);
            else __syncthreads();
            ff_forward_io(t, T->rt[pi], T->rtp[pi], p, pinv, false, true, lds_load, store_fwd, tw, ps.rl);
        }
    } else {
        if (!R3) {
            ff_inverse_io(t, T->irt[pi], T->irtp[pi], p, pinv, true, true, load_inv, store_inv, tw, ps.rl);
        } else {
            ff_inverse_io(t, T->irt[pi], T->irtp[pi], p, pinv, true, false, load_inv, lds_store, tw, ps.rl);
            ff_radix3_layer<true>(t, T, pi, P, ps.rstep3);
            for (int i = threadIdx.x; i < n; i += blockDim.x) store_inv(i >> ps.logC, i & (ps.C - 1),
                                                                        lds_load(i >> ps.logC, i & (ps.C - 1)));
        }
    }
}

// The grid position of a work-list workgroup: item k owns blocks
// [start[k], start[k+1]).
__device__ __forceinline__ int ff_find_item(const uint32_t *start, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Inner blocks: forward transform of a (and b), pointwise product, inverse,
// all in LDS. small = 1: the whole transform, read from the limbs and
// scaled for the CRT at the end.
// FIX12: every block is 2^12 words (the fixed path only, in a kernel of
// its own, whose registers are the fixed path's alone).
template <bool SMALL, bool R3, bool COH = false, bool FIX12 = false>
__device__ __forceinline__ void inner_body(const FFInner &in, const FFConv &C, int pi, uint32_t blkx,
                                           const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p;
    const bool square = SMALL ? (C.bsrc == nullptr) : (C.fb == nullptr);
    const uint32_t n = SMALL ? in.L : in.S;
    const int words = (int)(n + (n >> 4) + 16);
    uint32_t *A = lds, *B = lds + words;
    FFTile ta{A, 1, 1, 0, in.s, R3 ? 3 : 1, 1};
    FFTile tb = ta;
    tb.s = B;
    const size_t off = (size_t)pi * in.L + (SMALL ? 0 : (size_t)blkx * in.S);

    const uint32_t pinv = P.pinv;
    const bool small = SMALL;
    const int bw = in.bw;
    const uint32_t sc = in.sc[pi];
    const ff_gptr<uint32_t> fa = ff_g(C.fa), dst = fa + off;
    const ff_gptr<const uint32_t> fb = ff_g(C.fb), asrc = ff_g(C.asrc), bsrc = ff_g(C.bsrc);
    const ff_gptr<const uint32_t> fbo = C.fb ? fb + off : fb;
    // element i of the block <-> limbs' coefficient i (one-block form) or
    // the forward transform's word off + i: a uniform base plus a 32-bit
    // lane offset
    auto load_a = [&](int pos, int c) -> uint32_t {
        (void)c;
        return small ? ff_load_coef(asrc, C.anw, (uint64_t)pos, bw, P) : ff_ld<COH>(dst + (uint32_t)pos);
    };
    auto load_b = [&](int pos, int c) -> uint32_t {
        (void)c;
        return small ? ff_load_coef(bsrc, C.bnw, (uint64_t)pos, bw, P) : ff_ld<COH>(fbo + (uint32_t)pos);
    };
    auto store_a = [&](int pos, int c, uint32_t v) {
        (void)c;
        ff_st<COH>(dst + (uint32_t)pos, small ? ff_mont(v, sc, p, pinv) : v);
    };
    auto lds_a = [&](int pos, int c) -> uint32_t { return A[ff_pos_addr(ta, pos, c)]; };
    auto lds_a_st = [&](int pos, int c, uint32_t v) { A[ff_pos_addr(ta, pos, c)] = v; };
    auto lds_b_st = [&](int pos, int c, uint32_t v) { B[ff_pos_addr(tb, pos, c)] = v; };

    if (!SMALL && (FIX12 || in.s == 12)) {
        // the common block of 2^12 words: its geometry as literals
        const FFTile a12{A, 1, 1, 0, 12, 1, 1}, b12{B, 1, 1, 0, 12, 1, 1};
        const FFRows<COH> ra(dst, 4096, 1, 0);
        if (!square) {
            const FFRows<COH> rb(ff_gptr<uint32_t>(const_cast<uint32_t *>((const uint32_t *)fbo)), 4096, 1, 0);
            ff_forward12_io(b12, T->rt[pi], T->rtp[pi], p, pinv, rb);
            __syncthreads();
        }
        if (C.pair == 1) {
            FFPairMid pm;
            pm.w0 = C.cu[0], pm.w1 = C.cu[1], pm.w2 = C.cu[2];
            ff_conv12_io(a12, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], B, p, pinv, ra, ra, pm);
        } else {
            ff_conv12_io(a12, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, ra, ra);
        }
        return;
    }
    if constexpr (FIX12) return;
    if (!SMALL) {
        // the block's words as rows of one word (row accessors for the rounds)
        const FFRows<COH> ra(dst, n, 1, 0);
        if (!square) {
            const FFRows<COH> rb(ff_gptr<uint32_t>(const_cast<uint32_t *>((const uint32_t *)fbo)), n, 1, 0);
            ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, true, false, rb, lds_b_st);
            __syncthreads();
        }
        if (C.pair == 1) {
            // a weighted pair with one output: the middle round forms
            // cu0 a^2 + cu1 ab + cu2 b^2 from a's group and B
            FFPairMid pm;
            pm.w0 = C.cu[0], pm.w1 = C.cu[1], pm.w2 = C.cu[2];
            ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], B, p, pinv, true, true, ra, ra, pm);
        } else {
            ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, true,
                       true, ra, ra);
        }
        return;
    }
    const bool bgiven = C.bspec == 1;  // B's spectrum read, not computed
    // rounds of two levels up to 2^10 points: a radix-16 round of a short
    // transform keeps only n / 16 of the threads busy (16 at n = 256)
    const int small_rl = ta.rho <= 10 ? 2 : 4;
    if (C.bspec == 2) {
        // a's spectrum alone, stored by tile position for the products sharing it
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds_a_st((int)i, 0, load_a((int)i, 0));
        if (R3) ff_radix3_layer<false>(ta, T, pi, P, in.rstep3);
        else __syncthreads();
        ff_forward_io(ta, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_a_st);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) ff_st<COH>(dst + i, lds_a((int)i, 0));
        return;
    }
    if (!R3) {
        // one block read from limbs: unpack element by element into LDS
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            lds_a_st((int)i, 0, load_a((int)i, 0));
            if (bgiven) lds_b_st((int)i, 0, ff_ld<COH>(fbo + i));
            else if (!square) lds_b_st((int)i, 0, load_b((int)i, 0));
        }
        __syncthreads();
        if (!square && !bgiven) {
            ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_b_st);
            __syncthreads();
        }
        ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, false, true,
                   lds_a, store_a, FFNoTw{}, small_rl);
        return;
    }
    // 3 * 2^k in one block: the radix-3 layers work on the LDS tiles
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        lds_a_st((int)i, 0, load_a((int)i, 0));
        if (bgiven) lds_b_st((int)i, 0, ff_ld<COH>(fbo + i));
        else if (!square) lds_b_st((int)i, 0, load_b((int)i, 0));
    }
    ff_radix3_layer<false>(ta, T, pi, P, in.rstep3);
    if (!square && !bgiven) {
        ff_radix3_layer<false>(tb, T, pi, P, in.rstep3);
        ff_forward_io(tb, T->rt[pi], T->rtp[pi], p, pinv, false, false, lds_a, lds_b_st);
        __syncthreads();
    }
    ff_conv_io(ta, T->rt[pi], T->rtp[pi], T->irt[pi], T->irtp[pi], square ? nullptr : B, p, pinv, false, false,
               lds_a, lds_a_st, FFNoTw{}, small_rl);
    ff_radix3_layer<true>(ta, T, pi, P, in.rstep3);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) store_a((int)i, 0, lds_a((int)i, 0));
}

// Pair mode (multi-pass plans): 2 * FF_THREADS threads, the first half on
// the block of a, the second on the block of b, each through the forward
// levels into its LDS tile (tile A, tile B). The middle round then fuses the
// pointwise combination with the first inverse levels: each thread of the
// first half forms u = cu0 X + cu1 Y + cu2 Z of its group (X = a^2, Y = ab,
// Z = b^2 pointwise), each of the second half v likewise, both from tiles A
// and B; a barrier between those reads and the writes back keeps the two
// halves off each other's operands. The remaining inverse rounds end in
// HBM. With u alone the second half runs the rounds on an empty tile
// (nsub = 0), keeping the barrier counts equal.
//
// Degree 4 (pair & 4): the weights apply to X = x^2, Y = xy, Z = y^2 for
// x = a^2 + 2ab, y = a^2 + b^2 (F(2k), F(2k-1) from F(k), F(k-1)), and one
// product by R^3 restores the single R^-1 of a plain pointwise product that
// the inverse pass's CRT scale expects.

// w x mod p for a weight w < 8.
__device__ __forceinline__ uint32_t ff_wmul(uint32_t x, int w, uint32_t p) {
    const uint32_t x2 = ff_add(x, x, p), x4 = ff_add(x2, x2, p);
    uint32_t r = (w & 1) ? x : 0u;
    if (w & 2) r = ff_add(r, x2, p);
    if (w & 4) r = ff_add(r, x4, p);
    return r;
}

struct FFPairPoint {
    int w0, w1, w2;
    int deg4;
    uint32_t r3;  // R^3 mod p
    __device__ uint32_t operator()(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) const {
        if (deg4) {
            // x' = a (a + 2b), y' = a^2 + b^2
            const uint32_t aa = ff_mont(x, x, p, pinv), bb = ff_mont(y, y, p, pinv);
            const uint32_t y2 = ff_add(y, y, p);
            x = ff_mont(x, ff_add(x, y2, p), p, pinv);
            y = ff_add(aa, bb, p);
        }
        // w0 x^2 + w1 xy + w2 y^2 = x (w0 x + w1 y) + w2 y^2
        uint32_t u = ff_mont(x, ff_add(ff_wmul(x, w0, p), ff_wmul(y, w1, p), p), p, pinv);
        if (w2) u = ff_add(u, ff_wmul(ff_mont(y, y, p, pinv), w2, p), p);
        if (deg4) u = ff_mont(u, r3, p, pinv);
        return u;
    }
};

// The middle round on row bits [0, NL) of a one-column tile: this thread's
// output from tiles A and B, the first NL inverse levels, then (after a
// barrier) its own tile, or HBM through st when it is the last round. Every
// thread of the workgroup runs the same number of iterations.
template <int NL, class St>
__device__ __forceinline__ void pair_mid_round(const FFTile &t, bool active, const uint32_t *A, const uint32_t *B,
                                               const uint32_t *irt, const uint32_t *irtp, uint32_t p, uint32_t pinv,
                                               const FFPairPoint &pt, bool to_g, St st) {
    const int G = 1 << NL;
    const int ngroups = 1 << (t.rho - NL);
    const int nthr = ff_tile_nthr(t);
    const int iters = (ngroups + nthr - 1) / nthr;
    for (int it = 0; it < iters; it++) {
        const int g = ff_tile_tid(t) + it * nthr;
        const bool act = active && g < ngroups;
        uint32_t v[G];
        if (act) {
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int ad = ff_pos_addr(t, (g << NL) | j, 0);
                v[j] = pt(A[ad], B[ad], p, pinv);
            }
            ff_dit_group<NL, false, true>(v, 0, 0, irt, irtp, p, pinv);
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int j = 0; j < G; j++) {
                const int pos = (g << NL) | j;
                if (to_g) st(pos, 0, v[j]);
                else t.s[ff_pos_addr(t, pos, 0)] = v[j];
            }
        }
    }
}

// The inverse rounds of a pair block above its middle round: rounds of RL
// levels from row bit B up to RHO, the last one to HBM, at literal positions.
template <int RHO, int RL, int B, class St>
__device__ __forceinline__ void pair_dit_fixed(const FFTile &t, const uint32_t *irt, const uint32_t *irtp, uint32_t p,
                                               uint32_t pinv, St st) {
    if constexpr (B < RHO) {
        __syncthreads();
        ff_round_io<RL, FF_R_DIT, false, B + RL >= RHO>(t, B, nullptr, nullptr, irt, irtp, nullptr, p, pinv, st, st);
        pair_dit_fixed<RHO, RL, B + RL>(t, irt, irtp, p, pinv, st);
    }
}

// inner_pair_body for blocks of 2^RHO words in rounds of RL levels, the
// tile geometry and every round position literal; each half addresses its
// operand's block through a buffer descriptor (the half is the same for a
// whole wavefront, so the block's address is read from the first lane).
template <int RHO, int RL>
__device__ __forceinline__ void inner_pair_fixed(const FFConv &C, int pi, size_t off, const FFTables *T,
                                                 const FFPrimes &PR, uint32_t *lds) {
    constexpr int N = 1 << RHO, WORDS = N + (N >> 4) + 16;
    constexpr int NL0 = RHO % RL ? RHO % RL : RL;
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p, pinv = P.pinv;
    const int half = threadIdx.x >= FF_THREADS;
    const FFTile tf{lds + half * WORDS, 1, 1, 0, RHO, 1, 1, half * FF_THREADS, FF_THREADS};
    // (readfirstlane returns an int: each half goes back through uint32_t,
    // or a low half with bit 31 set would sign-extend over the high one)
    const uint64_t bp = (uint64_t)((half ? const_cast<uint32_t *>(C.fb) : C.fa) + off);
    const uint32_t bhi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(bp >> 32));
    const uint32_t blo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bp);
    const uint64_t bu = ((uint64_t)bhi << 32) | blo;
    const FFRows<false> rows(ff_g((uint32_t *)bu), N, 1, 0);
    ff_forward_fixed<RHO, RL, 0, true, false>(tf, T->rt[pi], T->rtp[pi], p, pinv, rows, rows, FFNoTw{});
    __syncthreads();
    const int two = (C.pair & 3) == 2;
    const FFPairPoint pt{half ? C.cv[0] : C.cu[0], half ? C.cv[1] : C.cu[1], half ? C.cv[2] : C.cu[2],
                         (C.pair & 4) != 0, ff_mont(P.r2, P.r2, p, pinv)};
    const bool active = !half || two;
    pair_mid_round<NL0>(tf, active, lds, lds + WORDS, T->irt[pi], T->irtp[pi], p, pinv, pt, false, rows);
    FFTile td = tf;
    if (!active) td.nsub = 0;
    pair_dit_fixed<RHO, RL, NL0>(td, T->irt[pi], T->irtp[pi], p, pinv, rows);
}

// PM: 0 the general body, 1 blocks of 2^12 words in rounds of four levels,
// 2 blocks of 2^10 words in rounds of two (ff_pair_mode).
__host__ __device__ inline int ff_pair_mode(const FFInner &in) {
    return in.s == 12 && in.rl == 4 ? 1 : in.s == 10 && in.rl == 2 ? 2 : 0;
}

template <int PM>
__device__ __forceinline__ void inner_pair_body(const FFInner &in, const FFConv &C, int pi, uint32_t blkx,
                                                const FFTables *T, const FFPrimes &PR, uint32_t *lds) {
    if constexpr (PM == 1) {
        inner_pair_fixed<12, 4>(C, pi, (size_t)pi * in.L + (size_t)blkx * in.S, T, PR, lds);
        return;
    }
    if constexpr (PM == 2) {
        inner_pair_fixed<10, 2>(C, pi, (size_t)pi * in.L + (size_t)blkx * in.S, T, PR, lds);
        return;
    }
    const FFPrime P = PR.p[pi];
    const uint32_t p = P.p, pinv = P.pinv;
    const uint32_t n = in.S;
    const int words = (int)(n + (n >> 4) + 16);
    const int half = threadIdx.x >= FF_THREADS;
    FFTile th{lds + half * words, 1, 1, 0, in.s, 1, 1, half * FF_THREADS, FF_THREADS};
    const size_t off = (size_t)pi * in.L + (size_t)blkx * in.S;
    const uint32_t *src = (half ? C.fb : C.fa) + off;
    uint32_t *dst = (half ? const_cast<uint32_t *>(C.fb) : C.fa) + off;
    auto load_g = [&](int pos, int c) -> uint32_t {
        (void)c;
        return src[pos];
    };
    auto lds_ld = [&](int pos, int c) -> uint32_t { return th.s[ff_pos_addr(th, pos, c)]; };
    auto lds_st = [&](int pos, int c, uint32_t v) { th.s[ff_pos_addr(th, pos, c)] = v; };
    auto store_g = [&](int pos, int c, uint32_t v) {
        (void)c;
        dst[pos] = v;
    };
    const int rl = in.rl;
    ff_forward_io(th, T->rt[pi], T->rtp[pi], p, pinv, true, false, load_g, lds_st, FFNoTw{}, rl);
    __syncthreads();
    const int two = (C.pair & 3) == 2;
    const FFPairPoint pt{half ? C.cv[0] : C.cu[0], half ? C.cv[1] : C.cu[1], half ? C.cv[2] : C.cu[2],
                         (C.pair & 4) != 0, ff_mont(P.r2, P.r2, p, pinv)};
    const bool active = !half || two;
    const int R = (in.s + rl - 1) / rl;
    const int nl0 = in.s % rl == 0 ? rl : in.s % rl;
    const uint32_t *A = lds, *B = lds + words;
    switch (nl0) {
    case 4: pair_mid_round<4>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    case 3: pair_mid_round<3>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    case 2: pair_mid_round<2>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    default: pair_mid_round<1>(th, active, A, B, T->irt[pi], T->irtp[pi], p, pinv, pt, R == 1, store_g); break;
    }
    if (!active) th.nsub = 0;
    int b = nl0;
    for (int r = 1; r < R; r++) {
        __syncthreads();
        ff_round_any<FF_R_DIT>(rl, false, r == R - 1, th, b, nullptr, nullptr, T->irt[pi], T->irtp[pi], nullptr, p,
                               pinv, lds_ld, store_g);
        b += rl;
    }
}

