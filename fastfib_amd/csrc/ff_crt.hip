// CRT recombination and carry propagation on gfx950.
//
// The three residue arrays y_p[i] (exact convolution coefficient i mod p,
// natural order, as the inverse top pass leaves them) are combined by
// Garner's method into X_i < p1 p2 p3 < 2^93 and added into the output at
// bit i * bw (crt.c:180-229, reference). Parallel form:
//   pack   one thread per chunk of 64 coefficients, which fill exactly bw
//          limbs; it writes them with a zero carry-in and records the
//          chunk's carry-out H (bits above its bw limbs) and whether limbs
//          2 .. bw-1 are all ones (then a carry could run through it);
//   scan   the carry into chunk k+1 is H_k + e_{k+1}, e_{k+1} in {0, 1} the
//          overflow of chunk k's low limbs plus its own carry-in: a
//          generate/propagate scan over chunks. Only chunks with that
//          all-ones flag can overflow, so when no chunk has it (the flag the
//          pack kernel raises stays clear) e = 0 throughout, the scan exits
//          at once and the carries are just the H's;
//   apply  adds the carry into each chunk (ripple confined to the chunk),
//          writes the final carry above the last chunk, and finds the
//          result's top nonzero limb.
// This replaces the reference's 8 sequential chunks plus carry fix
// (crt.c:326-336) with a segmented parallel carry scan.
#include <stdint.h>

#include "ff_internal.h"

typedef unsigned __int128 u128;

struct FFGarner {
    uint32_t p1, p2, p3;
    uint32_t pinv2, pinv3;
    uint32_t i12, i12p, i13, i13p, i23, i23p;  // pi^-1 mod pj, Montgomery form, and * pinv
    uint64_t p12;
};

static FFGarner g_garner;
static int g_garner_ready = 0;

static uint32_t h_pow(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

static void garner_init(void) {
    if (g_garner_ready) return;
    FFPrime P[3];
    ff_prime_consts(P);
    FFGarner &G = g_garner;
    G.p1 = P[0].p, G.p2 = P[1].p, G.p3 = P[2].p;
    G.pinv2 = P[1].pinv, G.pinv3 = P[2].pinv;
    const uint64_t R2 = (1ull << 32) % G.p2, R3 = (1ull << 32) % G.p3;
    G.i12 = (uint32_t)(h_pow(G.p1 % G.p2, G.p2 - 2, G.p2) * R2 % G.p2);
    G.i13 = (uint32_t)(h_pow(G.p1 % G.p3, G.p3 - 2, G.p3) * R3 % G.p3);
    G.i23 = (uint32_t)(h_pow(G.p2 % G.p3, G.p3 - 2, G.p3) * R3 % G.p3);
    G.i12p = G.i12 * G.pinv2;
    G.i13p = G.i13 * G.pinv3;
    G.i23p = G.i23 * G.pinv3;
    G.p12 = (uint64_t)G.p1 * G.p2;
    g_garner_ready = 1;
}

// X = x1 + x2 p1 + x3 p1 p2 from canonical residues (crt.c:180-206).
__device__ __forceinline__ u128 ff_garner(uint32_t r1, uint32_t r2, uint32_t r3, const FFGarner &G) {
    const uint32_t x1 = r1;
    const uint32_t x1m2 = x1 >= G.p2 ? x1 - G.p2 : x1;  // p1 < 2 p2
    const uint32_t x2 = ff_mont_w(r2 - x1m2 + G.p2, G.i12, G.i12p, G.p2);
    const uint32_t x1m3 = x1 >= G.p3 ? x1 - G.p3 : x1;  // p1 < 2 p3
    const uint32_t t3 = ff_mont_w(r3 - x1m3 + G.p3, G.i13, G.i13p, G.p3);
    const uint32_t x2m3 = x2 >= G.p3 ? x2 - G.p3 : x2;  // p2 < 2 p3
    const uint32_t x3 = ff_mont_w(t3 - x2m3 + G.p3, G.i23, G.i23p, G.p3);
    const uint64_t lo = (uint64_t)x2 * G.p1 + x1;  // < 2^62
    return (u128)x3 * G.p12 + lo;
}

#define PACK_THREADS 64

// One thread per chunk of 64 coefficients, read 16 at a time as 128-bit
// loads (all issued before use).
__global__ __launch_bounds__(PACK_THREADS) void k_crt_pack(FFCrtSet js, FFGarner G) {
    const FFCrtJob &J = js.j[blockIdx.y];
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * (uint64_t)blockDim.x >= nchunks) return;
    const int bw = J.bw;
    const u128 mask = ((u128)1 << bw) - 1;
    const uint32_t *__restrict__ y1 = J.y;
    const uint32_t *__restrict__ y2 = J.y + J.L;
    const uint32_t *__restrict__ y3 = J.y + 2 * (size_t)J.L;
    uint64_t *__restrict__ w = J.out + k * (uint64_t)bw;
    u128 buf = 0, c = 0;
    int nbits = 0, q = 0;
    uint32_t ones = 1;
    const uint64_t i0 = 64 * k;
    // the chunk's coefficients past conv (and past L) are zero
    const uint64_t lim = J.conv < J.L ? J.conv : J.L;
    const bool live = k < nchunks;
    for (int part = 0; part < 4 && live; part++) {
        const uint64_t b0 = i0 + 16 * part;
        uint32_t r1[16], r2[16], r3[16];
        if (b0 + 16 <= lim) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const uint4 a1 = *(const uint4 *)(y1 + b0 + 4 * v);
                const uint4 a2 = *(const uint4 *)(y2 + b0 + 4 * v);
                const uint4 a3 = *(const uint4 *)(y3 + b0 + 4 * v);
                r1[4 * v] = a1.x, r1[4 * v + 1] = a1.y, r1[4 * v + 2] = a1.z, r1[4 * v + 3] = a1.w;
                r2[4 * v] = a2.x, r2[4 * v + 1] = a2.y, r2[4 * v + 2] = a2.z, r2[4 * v + 3] = a2.w;
                r3[4 * v] = a3.x, r3[4 * v + 1] = a3.y, r3[4 * v + 2] = a3.z, r3[4 * v + 3] = a3.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const bool in = b0 + j < lim;
                r1[j] = in ? y1[b0 + j] : 0;
                r2[j] = in ? y2[b0 + j] : 0;
                r3[j] = in ? y3[b0 + j] : 0;
            }
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const u128 X = ff_garner(r1[j], r2[j], r3[j], G);
            const u128 acc = X + c;
            buf |= (acc & mask) << nbits;
            c = acc >> bw;
            nbits += bw;
            if (nbits >= 64) {
                const uint64_t limb = (uint64_t)buf;
                w[q] = limb;
                if (q >= 2 && limb != ~0ull) ones = 0;
                q++;
                buf >>= 64;
                nbits -= 64;
            }
        }
    }
    if (live) {
        FFChunk r;
        r.h_lo = (uint64_t)c;
        r.h_hi = (uint64_t)(c >> 64);
        r.ones = ones;
        r.pad_ = 0;
        J.chunks[k] = r;
    }
    // one plain store per workgroup: did any of its chunks raise the flag
    const int any = __syncthreads_or(live && ones);
    if (threadIdx.x == 0) J.wgflag[blockIdx.x] = (uint32_t)any;
}

// Exact carry-lookahead over chunks, one workgroup (1024 threads) per
// product, only when some chunk could propagate a carry.
// g_k: chunk k overflows with e_k = 0; p_k: it overflows with e_k = 1.
#define SCAN_THREADS 1024

__device__ __forceinline__ u128 chunk_h(const FFCrtJob &J, uint64_t k) {
    const FFChunk pc = J.chunks[k];
    return ((u128)pc.h_hi << 64) | pc.h_lo;
}

__device__ __forceinline__ void chunk_gp(const FFCrtJob &J, uint64_t k, uint32_t &g, uint32_t &p) {
    const FFChunk ch = J.chunks[k];
    if (!ch.ones) {
        g = p = 0;
        return;
    }
    const u128 hprev = k > 0 ? chunk_h(J, k - 1) : 0;
    const uint64_t *w = J.out + k * (uint64_t)J.bw;
    const u128 low = ((u128)w[1] << 64) | w[0];
    const u128 s = low + hprev;
    const uint32_t ov0 = s < low;  // carry out of 128 bits
    g = ov0;
    p = ov0 | (s == ~(u128)0);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_crt_scan(FFCrtSet js) {
    const FFCrtJob &J = js.j[blockIdx.x];
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t nwg = (nchunks + PACK_THREADS - 1) / PACK_THREADS;
    int mine = 0;
    for (uint64_t i = threadIdx.x; i < nwg; i += SCAN_THREADS) mine |= J.wgflag[i] != 0;
    const int any = __syncthreads_or(mine);
    if (threadIdx.x == 0) {
        *J.mode = (uint32_t)any;
        *J.lenkey = 0;  // k_crt_apply books the top limb by atomicMax
    }
    if (!any) return;  // no chunk can propagate: the H's are the carries
    const uint64_t per = (nchunks + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint64_t k0 = threadIdx.x * per, k1 = k0 + per < nchunks ? k0 + per : nchunks;
    uint32_t Gs = 0, Ps = 1;
    for (uint64_t k = k0; k < k1; k++) {
        uint32_t g, p;
        chunk_gp(J, k, g, p);
        Gs = g | (p & Gs);
        Ps = p & Ps;
    }
    __shared__ uint32_t sg[SCAN_THREADS], sp[SCAN_THREADS];
    sg[threadIdx.x] = Gs;
    sp[threadIdx.x] = Ps;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const uint32_t g2 = sg[threadIdx.x], p2 = sp[threadIdx.x];
        uint32_t g1 = 0, p1 = 1;
        if ((int)threadIdx.x >= off) g1 = sg[threadIdx.x - off], p1 = sp[threadIdx.x - off];
        __syncthreads();
        sg[threadIdx.x] = g2 | (p2 & g1);
        sp[threadIdx.x] = p2 & p1;
        __syncthreads();
    }
    uint32_t e = threadIdx.x ? sg[threadIdx.x - 1] : 0;
    for (uint64_t k = k0; k < k1; k++) {
        J.cin[k] = (k > 0 ? chunk_h(J, k - 1) : 0) + e;
        uint32_t g, p;
        chunk_gp(J, k, g, p);
        e = g | (p & e);
    }
    if (k1 == nchunks && k0 < k1) J.cin[nchunks] = chunk_h(J, nchunks - 1) + e;
}

// Top nonzero limb of d[i0..i1) booked into the product's length key.
__device__ __forceinline__ void book_top(const FFCrtJob &J, const uint64_t *d, uint64_t i0, uint64_t i1) {
    for (uint64_t i = i1; i-- > i0;) {
        const uint64_t v = d[i];
        if (v) {
            const unsigned long long key = ((unsigned long long)(i + 1) << 8) | (unsigned)(64 - __clzll((long long)v));
            atomicMax(J.lenkey, key);
            return;
        }
    }
}

// Adds the incoming carry into each chunk; the final carry goes above. The
// chunks that hold limbs from lo_top up (where the top nonzero limb must
// be) book it.
__global__ __launch_bounds__(256) void k_crt_apply(FFCrtSet js, uint64_t lo_top0, uint64_t lo_top1,
                                                   uint64_t lo_top2, uint64_t lo_top3) {
    const FFCrtJob &J = js.j[blockIdx.y];
    const uint64_t lo_top = blockIdx.y == 0 ? lo_top0 : blockIdx.y == 1 ? lo_top1 : blockIdx.y == 2 ? lo_top2 : lo_top3;
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nchunks) return;
    const bool exact = *J.mode != 0;
    const uint64_t base = k * (uint64_t)J.bw;
    uint64_t *w = J.out + base;
    const u128 c = exact ? J.cin[k] : (k > 0 ? chunk_h(J, k - 1) : (u128)0);
    if (k == nchunks) {
        // limbs above the last chunk: the final carry, then zeros up to nout
        w[0] = (uint64_t)c;
        w[1] = (uint64_t)(c >> 64);
        for (uint64_t i = base + 2; i < J.nout; i++) J.out[i] = 0;
        const uint64_t top = J.nout > base + 2 ? J.nout : base + 2;
        if (top > lo_top) book_top(J, J.out, base > lo_top ? base : lo_top, top);
        return;
    }
    if (c != 0) {
        const u128 low = ((u128)w[1] << 64) | w[0];
        const u128 s = low + c;
        w[0] = (uint64_t)s;
        w[1] = (uint64_t)(s >> 64);
        if (s < low) {
            for (int q = 2; q < J.bw; q++) {
                if (++w[q] != 0) break;
            }
        }
    }
    const uint64_t end = base + J.bw;
    if (end > lo_top) book_top(J, J.out, base > lo_top ? base : lo_top, end);
}

// Algorithmic bytes: pack reads 3 residues per coefficient and writes the
// limbs and one 24-byte record per chunk; scan (exact path only) reads the
// records and two limbs per chunk; apply reads one record per chunk.
extern "C" hipError_t ff_launch_crt(const FFCrtSet *js, int njobs, uint64_t max_chunks, double conv_total,
                                    double out_limbs_total, hipStream_t st) {
    garner_init();
    const double chunks = conv_total / 64.0;
    dim3 g1((unsigned)((max_chunks + PACK_THREADS - 1) / PACK_THREADS), njobs);
    ff_prof_begin(FK_CRT_PACK, st);
    hipLaunchKernelGGL(k_crt_pack, g1, dim3(PACK_THREADS), 0, st, *js, g_garner);
    ff_prof_end(FK_CRT_PACK, 12.0 * conv_total + 8.0 * out_limbs_total + 24.0 * chunks, st);
    ff_prof_begin(FK_CRT_SCAN, st);
    hipLaunchKernelGGL(k_crt_scan, dim3(njobs), dim3(SCAN_THREADS), 0, st, *js);
    ff_prof_end(FK_CRT_SCAN, 0.0, st);
    uint64_t lo[4] = {0, 0, 0, 0};
    for (int i = 0; i < njobs && i < 4; i++) {
        // the product of an na-limb and an nb-limb number has its top limb
        // at index >= na + nb - 2 = nout - 3 (nout = na + nb + 1)
        lo[i] = js->j[i].nout >= 3 ? js->j[i].nout - 3 : 0;
    }
    dim3 g3((unsigned)((max_chunks + 1 + 255) / 256), njobs);
    ff_prof_begin(FK_CRT_APPLY, st);
    hipLaunchKernelGGL(k_crt_apply, g3, dim3(256), 0, st, *js, lo[0], lo[1], lo[2], lo[3]);
    ff_prof_end(FK_CRT_APPLY, 24.0 * chunks, st);
    return hipGetLastError();
}
