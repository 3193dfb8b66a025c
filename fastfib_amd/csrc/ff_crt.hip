// CRT recombination and carry propagation on gfx950.
//
// The three residue arrays y_p[i] (exact convolution coefficient i mod p,
// natural order, as the inverse top pass leaves them) are combined by
// Garner's method into X_i < p1 p2 p3 < 2^93 and added into the output at
// bit i * bw (crt.c:180-229, reference). Parallel form:
//   pack   one thread per chunk of 64 coefficients, which fill exactly bw
//          limbs; it writes them with a zero carry-in and records the
//          chunk's carry-out H (bits above its bw limbs) and whether limbs
//          2 .. bw-1 are all ones;
//   scan   the carry into chunk k+1 is H_k plus e_{k+1} in {0, 1}, where
//          e_{k+1} = 1 iff adding C_k = H_{k-1} + e_k to chunk k's low limbs
//          overflows its bw limbs: a generate/propagate carry-lookahead scan
//          over chunks (one workgroup per product);
//   apply  adds C_k into chunk k (ripple bounded by the chunk) and writes the
//          final carry above the last chunk.
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

// One thread per chunk of 64 coefficients.
__global__ __launch_bounds__(256) void k_crt_pack(const FFCrtJob *jobs, FFGarner G) {
    const FFCrtJob J = jobs[blockIdx.y];
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nchunks) return;
    const int bw = J.bw;
    const u128 mask = ((u128)1 << bw) - 1;
    const uint32_t *y1 = J.y, *y2 = J.y + J.L, *y3 = J.y + 2 * (size_t)J.L;
    uint64_t *w = J.out + k * (uint64_t)bw;
    u128 buf = 0, c = 0;
    int nbits = 0, q = 0;
    uint32_t ones = 1;
    const uint64_t i0 = 64 * k;
    for (int j = 0; j < 64; j++) {
        const uint64_t i = i0 + j;
        u128 X = 0;
        if (i < J.conv && i < J.L) X = ff_garner(y1[i], y2[i], y3[i], G);
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
    FFChunk r;
    r.h_lo = (uint64_t)c;
    r.h_hi = (uint64_t)(c >> 64);
    r.ones = ones;
    r.pad_ = 0;
    J.chunks[k] = r;
}

// Carry-lookahead over chunks, one workgroup (1024 threads) per product.
// g_k: chunk k overflows with e_k = 0; p_k: it overflows with e_k = 1.
#define SCAN_THREADS 1024

__device__ __forceinline__ void chunk_gp(const FFCrtJob &J, uint64_t k, uint32_t &g, uint32_t &p) {
    const FFChunk ch = J.chunks[k];
    if (!ch.ones) {
        g = p = 0;
        return;
    }
    u128 hprev = 0;
    if (k > 0) {
        const FFChunk pc = J.chunks[k - 1];
        hprev = ((u128)pc.h_hi << 64) | pc.h_lo;
    }
    const uint64_t *w = J.out + k * (uint64_t)J.bw;
    const u128 low = ((u128)w[1] << 64) | w[0];
    const u128 s = low + hprev;
    const uint32_t ov0 = s < low;  // carry out of 128 bits
    g = ov0;
    p = ov0 | (s == ~(u128)0);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_crt_scan(const FFCrtJob *jobs) {
    const FFCrtJob J = jobs[blockIdx.x];
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t per = (nchunks + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint64_t k0 = threadIdx.x * per, k1 = k0 + per < nchunks ? k0 + per : nchunks;
    // segment summary: the composition of the chunks' (g, p) maps
    uint32_t G = 0, Pp = 1;
    for (uint64_t k = k0; k < k1; k++) {
        uint32_t g, p;
        chunk_gp(J, k, g, p);
        G = g | (p & G);
        Pp = p & Pp;
    }
    __shared__ uint32_t sg[SCAN_THREADS], sp[SCAN_THREADS];
    sg[threadIdx.x] = G;
    sp[threadIdx.x] = Pp;
    __syncthreads();
    // inclusive Hillis-Steele scan of the maps (later segment applied last)
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint32_t g2 = sg[threadIdx.x], p2 = sp[threadIdx.x];
        uint32_t g1 = 0, p1 = 1;
        if ((int)threadIdx.x >= off) g1 = sg[threadIdx.x - off], p1 = sp[threadIdx.x - off];
        __syncthreads();
        sg[threadIdx.x] = g2 | (p2 & g1);
        sp[threadIdx.x] = p2 & p1;
        __syncthreads();
    }
    // carry bit entering this segment (e_0 = 0)
    uint32_t e = threadIdx.x ? sg[threadIdx.x - 1] : 0;
    for (uint64_t k = k0; k < k1; k++) {
        u128 hprev = 0;
        if (k > 0) {
            const FFChunk pc = J.chunks[k - 1];
            hprev = ((u128)pc.h_hi << 64) | pc.h_lo;
        }
        J.cin[k] = hprev + e;
        uint32_t g, p;
        chunk_gp(J, k, g, p);
        e = g | (p & e);
    }
    if (k1 == nchunks && k0 < k1) {
        const FFChunk pc = J.chunks[nchunks - 1];
        J.cin[nchunks] = (((u128)pc.h_hi << 64) | pc.h_lo) + e;
    }
}

// Adds the incoming carry into each chunk; the final carry goes above.
__global__ __launch_bounds__(256) void k_crt_apply(const FFCrtJob *jobs) {
    const FFCrtJob J = jobs[blockIdx.y];
    const uint64_t nchunks = (J.conv + 63) / 64;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nchunks) return;
    uint64_t *w = J.out + k * (uint64_t)J.bw;
    const u128 c = J.cin[k];
    if (k == nchunks) {
        // limbs above the last chunk: the final carry, then zeros up to nout
        const uint64_t base = k * (uint64_t)J.bw;
        if (base < J.nout + 2) {
            w[0] = (uint64_t)c;
            w[1] = (uint64_t)(c >> 64);
        }
        for (uint64_t i = base + 2; i < J.nout; i++) J.out[i] = 0;
        return;
    }
    if (c == 0) return;
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

extern "C" hipError_t ff_launch_crt(const FFCrtJob *jobs_dev, int njobs, uint64_t max_chunks, hipStream_t st) {
    garner_init();
    dim3 g1((unsigned)((max_chunks + 255) / 256), njobs);
    hipLaunchKernelGGL(k_crt_pack, g1, dim3(256), 0, st, jobs_dev, g_garner);
    hipLaunchKernelGGL(k_crt_scan, dim3(njobs), dim3(SCAN_THREADS), 0, st, jobs_dev);
    dim3 g3((unsigned)((max_chunks + 1 + 255) / 256), njobs);
    hipLaunchKernelGGL(k_crt_apply, g3, dim3(256), 0, st, jobs_dev);
    return hipGetLastError();
}

// Normalized length of d[0..n): index of the top nonzero limb + 1 (>= 1).
__global__ __launch_bounds__(256) void k_normlen(const uint64_t *d, uint64_t n, uint64_t *len_out) {
    __shared__ uint64_t best;
    if (threadIdx.x == 0) best = 0;
    __syncthreads();
    for (uint64_t top = n; top > 0;) {
        const uint64_t lo = top > 256 ? top - 256 : 0;
        const uint64_t i = lo + threadIdx.x;
        if (i < top && d[i] != 0) atomicMax((unsigned long long *)&best, (unsigned long long)(i + 1));
        __syncthreads();
        if (best) break;
        top = lo;
        __syncthreads();
    }
    if (threadIdx.x == 0) *len_out = best ? best : 1;
}

extern "C" hipError_t ff_launch_normlen(const uint64_t *d, uint64_t n, uint64_t *len_out, hipStream_t st) {
    hipLaunchKernelGGL(k_normlen, dim3(1), dim3(256), 0, st, d, n, len_out);
    return hipGetLastError();
}
