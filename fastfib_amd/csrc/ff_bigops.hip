// Limb arithmetic on device numbers (the reference's bigint.c helpers on
// the GPU): batched add / subtract with carry lookahead, shifts, selection,
// and normalized single-number operations built on them.
#include <stdio.h>
#include <string.h>

#include "ff_bigops.h"
#include "ff_engine.h"
#include "ff_internal.h"
#include "ff_limbs.h"
#include "ff_runtime.h"

#define ADD_T 256  // limbs per block, one per thread (4 wavefronts)

// ------------------------------------------------------------------ add / sub

// One block of 256 limbs of one node: the limb sums u = a + b' (b' = b, ~b
// or ~2b), then carries inside the block by wavefront ballots. Lane k of a
// wave generates (g: a + b' overflows) or propagates (p: u = 2^64 - 1); with
// G, P the wave's ballots, bit k of (G + (G|P) + c) ^ G ^ (G|P) is the carry
// into lane k for wave carry-in c (lanes past the node's end propagate). The
// block's known carry-in (1 for block
// 0 of a subtraction, the two's complement +1, else 0) is applied here; the
// block summary (carry-out, all-propagate) goes to the scan.
__global__ __launch_bounds__(ADD_T) void k_add_local(FFAddNodes J) {
    __shared__ uint32_t wg[4], wp[4], wc[4];
    const uint32_t node = blockIdx.x / J.nblk, blk = blockIdx.x - node * J.nblk;
    const uint64_t i = (uint64_t)blk * ADD_T + threadIdx.x;
    const bool live = i < J.n;
    uint64_t av = 0, bv = 0;
    if (live) {
        if (i < J.na) av = J.a[node * J.sa + i];
        if (J.b) {
            const uint64_t *b = J.b + node * J.sb;
            if (J.b_mode == FF_SUB2X) {
                const uint64_t lo = i < J.nb ? b[i] : 0, pr = (i >= 1 && i - 1 < J.nb) ? b[i - 1] : 0;
                bv = (lo << 1) | (pr >> 63);
            } else if (i < J.nb) {
                bv = b[i];
            }
        }
        if (J.b_mode != FF_ADD) bv = ~bv;
    }
    const uint64_t u = av + bv;
    // lanes past the end propagate, so the block's carry-out is the carry
    // out of the node's top limb
    const bool g = live && u < av;
    const bool p = !live || u == ~0ull;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t G = __ballot(g), B = G | __ballot(p);
    if (lane == 0) {
        wg[w] = (G + B) < G;  // carry out of the wave for carry-in 0
        wp[w] = B == ~0ull && G == 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t base = (J.b_mode != FF_ADD && blk == 0) ? 1u : 0u;
        uint32_t c = base, allp = 1;
        for (int k = 0; k < 4; k++) {
            wc[k] = c;
            c = wg[k] | (wp[k] & c);
            allp &= wp[k];
        }
        const uint32_t P = blk == 0 ? 0u : allp;
        J.bgp[blockIdx.x] = c | (P << 1);
        if (J.nblk == 1 && J.cout) J.cout[node] = (uint8_t)c;
    }
    __syncthreads();
    const uint64_t cw = wc[w];
    const uint64_t C = (G + B + cw) ^ G ^ B;
    if (live) J.out[node * J.so + i] = u + ((C >> lane) & 1);
}

// Exact carry into every block of a node: the node's blocks are cut into
// 256 segments, each composed (generate, propagate), the segments scanned
// in order, then each segment walked with its carry-in.
__global__ __launch_bounds__(256) void k_add_scan(FFAddNodes J) {
    __shared__ uint8_t sg[256], sp[256], sc[256];
    const uint32_t node = blockIdx.x;
    const uint32_t nb = J.nblk, seg = (nb + 255) / 256;
    const uint32_t b0 = threadIdx.x * seg, b1 = b0 + seg < nb ? b0 + seg : nb;
    const uint32_t *bgp = J.bgp + (size_t)node * nb;
    uint32_t G = 0, P = 1;
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t v = bgp[b];
        G = (v & 1) | ((v >> 1) & G);
        P &= (v >> 1) & 1;
    }
    sg[threadIdx.x] = (uint8_t)G;
    sp[threadIdx.x] = (uint8_t)(b0 < b1 ? P : 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int t = 0; t < 256; t++) {
            sc[t] = (uint8_t)c;
            c = sg[t] | (sp[t] & c);
        }
        if (J.cout) J.cout[node] = (uint8_t)c;
    }
    __syncthreads();
    uint32_t c = sc[threadIdx.x];
    uint8_t *bcin = J.bcin + (size_t)node * nb;
    for (uint32_t b = b0; b < b1; b++) {
        bcin[b] = (uint8_t)c;
        const uint32_t v = bgp[b];
        c = (v & 1) | ((v >> 1) & c);
    }
}

// Adds each block's carry-in, one thread per limb: limb k of the block
// takes the carry when the limbs below it in the block are all ones (a
// subtraction of nearly equal numbers ripples through whole blocks, so a
// sequential ripple would be slow); wavefront ballots find the first limb
// that stops it.
__global__ __launch_bounds__(ADD_T) void k_add_apply(FFAddNodes J) {
    const uint32_t b = blockIdx.x;
    if (!J.bcin[b]) return;  // uniform over the block
    __shared__ uint32_t all1[ADD_T / 64];
    const uint32_t node = b / J.nblk, blk = b - node * J.nblk;
    uint64_t *o = J.out + node * J.so;
    const uint64_t i = (uint64_t)blk * ADD_T + threadIdx.x;
    const bool live = i < J.n;
    const uint64_t v = live ? o[i] : 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t ones = __ballot(live && v == ~0ull);
    if (lane == 0) all1[w] = ones == ~0ull;
    __syncthreads();
    bool cw = true;  // the carry reaches this wave
    for (int k = 0; k < w; k++) cw = cw && all1[k];
    // lanes up to the first that is not all ones take the carry
    const uint64_t below = lane ? (ones & ((1ull << lane) - 1)) : 0;
    const bool take = cw && below == ((lane ? (1ull << lane) - 1 : 0));
    if (live && take) o[i] = v + 1;
}

size_t ff_add_nodes_scratch(uint32_t N, uint64_t n) {
    const size_t blocks = (size_t)N * ((n + ADD_T - 1) / ADD_T);
    return (blocks * 4 + 255) / 256 * 256 + blocks + 256;
}

hipError_t ff_add_nodes(FFAddNodes J, void *scratch, hipStream_t st) {
    if (J.N == 0 || J.n == 0) return hipSuccess;
    J.nblk = (uint32_t)((J.n + ADD_T - 1) / ADD_T);
    const size_t blocks = (size_t)J.N * J.nblk;
    J.bgp = (uint32_t *)scratch;
    J.bcin = (uint8_t *)scratch + (blocks * 4 + 255) / 256 * 256;
    ff_prof_begin(FK_LINEAR, st);
    hipLaunchKernelGGL(k_add_local, dim3((unsigned)blocks), dim3(ADD_T), 0, st, J);
    if (J.nblk > 1) {
        hipLaunchKernelGGL(k_add_scan, dim3(J.N), dim3(256), 0, st, J);
        hipLaunchKernelGGL(k_add_apply, dim3((unsigned)blocks), dim3(ADD_T), 0, st, J);
    }
    ff_prof_end(FK_LINEAR, 8.0 * J.N * (3.0 * J.n), st);
    return hipGetLastError();
}

// ------------------------------------------------------------------ node kernels

__global__ void k_select3(const uint64_t *r, const uint64_t *d1, const uint64_t *d2, const uint8_t *c1,
                          const uint8_t *c2, uint64_t sr, uint64_t n, uint64_t total, uint64_t *out, uint64_t so) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t j = t / n, i = t - j * n;
    const uint64_t o = j * sr + i;
    out[j * so + i] = c2[j] ? d2[o] : (c1[j] ? d1[o] : r[o]);
}

hipError_t ff_select3(const uint64_t *r, const uint64_t *d1, const uint64_t *d2, const uint8_t *c1,
                      const uint8_t *c2, uint64_t sr, uint64_t n, uint32_t N, uint64_t *out, uint64_t so,
                      hipStream_t st) {
    const uint64_t total = (uint64_t)N * n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_select3, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, r, d1, d2, c1, c2, sr,
                       n, total, out, so);
    return hipGetLastError();
}

__global__ void k_copy_nodes(const uint64_t *q, uint64_t sq, uint64_t nq, uint64_t total, uint64_t *out, uint64_t so,
                             uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t j = t / n, i = t - j * n;
    out[j * so + i] = i < nq ? q[j * sq + i] : 0;
}

__global__ void k_add_small_nodes(const uint8_t *c1, const uint8_t *c2, uint32_t N, uint64_t *out, uint64_t so,
                                  uint64_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    uint64_t c = (uint64_t)c1[j] + (uint64_t)c2[j];
    uint64_t *o = out + j * so;
    for (uint64_t i = 0; i < n && c; i++) {
        const uint64_t v = o[i] + c;
        c = v < c ? 1 : 0;
        o[i] = v;
    }
}

hipError_t ff_copy_add_small(const uint64_t *q, uint64_t sq, uint64_t nq, const uint8_t *c1, const uint8_t *c2,
                             uint32_t N, uint64_t *out, uint64_t so, uint64_t n, hipStream_t st) {
    const uint64_t total = (uint64_t)N * n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_copy_nodes, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, q, sq, nq, total, out,
                       so, n);
    hipLaunchKernelGGL(k_add_small_nodes, dim3((N + 255) / 256), dim3(256), 0, st, c1, c2, N, out, so, n);
    return hipGetLastError();
}

__global__ void k_shr_nodes(const uint64_t *a, uint64_t sa, uint64_t na, uint64_t s, unsigned b, uint64_t total,
                            uint64_t *out, uint64_t so, uint64_t nout) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint64_t j = t / nout, i = t - j * nout;
    const uint64_t *x = a + j * sa;
    const uint64_t lo = i + s < na ? x[i + s] : 0;
    const uint64_t hi = i + s + 1 < na ? x[i + s + 1] : 0;
    out[j * so + i] = b ? (lo >> b) | (hi << (64 - b)) : lo;
}

hipError_t ff_shr_nodes(const uint64_t *a, uint64_t sa, uint64_t na, uint64_t k, uint32_t N, uint64_t *out,
                        uint64_t so, uint64_t nout, hipStream_t st) {
    const uint64_t total = (uint64_t)N * nout;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_shr_nodes, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, sa, na, k / 64,
                       (unsigned)(k % 64), total, out, so, nout);
    return hipGetLastError();
}

__global__ void k_shl(const uint64_t *a, uint64_t na, uint64_t s, unsigned b, uint64_t *out, uint64_t nout) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const uint64_t lo = (i >= s && i - s < na) ? a[i - s] : 0;
    const uint64_t pr = (b && i >= s + 1 && i - s - 1 < na) ? a[i - s - 1] : 0;
    out[i] = b ? (lo << b) | (pr >> (64 - b)) : lo;
}

// key = (top nonzero limb + 1) << 8 | its bit length, by atomicMax.
__global__ void k_norm_key(const uint64_t *d, uint64_t n, unsigned long long *key) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t v = d[i];
    if (v) atomicMax(key, ((unsigned long long)(i + 1) << 8) | (unsigned)(64 - __clzll((long long)v)));
}

// ------------------------------------------------------------------ numbers

FFNum ff_num_alloc(uint64_t cap) {
    cap = cap ? cap : 1;
    return FFNum{(uint64_t *)ff_dalloc(cap * 8), 1, 0, cap};
}

void ff_num_free(FFNum &x) {
    if (x.d) ff_dfree(x.d);
    x.d = nullptr;
    x.n = x.bits = x.cap = 0;
}

FFNum ff_num_u64(uint64_t v) {
    FFNum x = ff_num_alloc(1);
    FF_CHECK(ff_launch_fill(x.d, 1, v, g_st));
    x.bits = v ? 64 - __builtin_clzll(v) : 0;
    return x;
}

FFNum ff_num_copy(const FFNum &x) {
    FFNum y = ff_num_alloc(x.n);
    FF_CHECK(hipMemcpyAsync(y.d, x.d, x.n * 8, hipMemcpyDeviceToDevice, g_st));
    y.n = x.n, y.bits = x.bits;
    return y;
}

FFNum ff_num_pow2(uint64_t k) {
    const uint64_t n = k / 64 + 1;
    FFNum x = ff_num_alloc(n);
    FF_CHECK(ff_launch_fill(x.d, n, 0, g_st));
    FF_CHECK(ff_launch_fill(x.d + k / 64, 1, 1ull << (k % 64), g_st));
    x.n = n, x.bits = k + 1;
    return x;
}

static uint64_t limbs_of(uint64_t bits) { return bits ? (bits + 63) / 64 : 1; }

FFNum ff_num_shr(const FFNum &x, uint64_t k) {
    if (x.bits <= k) return ff_num_u64(0);
    const uint64_t bits = x.bits - k, n = limbs_of(bits);
    FFNum y = ff_num_alloc(n);
    FF_CHECK(ff_shr_nodes(x.d, 0, x.n, k, 1, y.d, 0, n, g_st));
    y.n = n, y.bits = bits;
    return y;
}

FFNum ff_num_shl(const FFNum &x, uint64_t k) {
    if (x.bits == 0) return ff_num_u64(0);
    const uint64_t bits = x.bits + k, n = limbs_of(bits);
    FFNum y = ff_num_alloc(n);
    hipLaunchKernelGGL(k_shl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g_st, x.d, x.n, k / 64,
                       (unsigned)(k % 64), y.d, n);
    FF_CHECK(hipGetLastError());
    y.n = n, y.bits = bits;
    return y;
}

void ff_num_normalize(FFNum &x) {
    unsigned long long *key = (unsigned long long *)ff_dalloc(8);
    FF_CHECK(ff_launch_fill((uint64_t *)key, 1, 0, g_st));
    hipLaunchKernelGGL(k_norm_key, dim3((unsigned)((x.n + 255) / 256)), dim3(256), 0, g_st, x.d, x.n, key);
    FF_CHECK(hipGetLastError());
    uint64_t *h = ff_pinned_results(1);
    FF_CHECK(hipMemcpyAsync(h, key, 8, hipMemcpyDeviceToHost, g_st));
    FF_CHECK(hipStreamSynchronize(g_st));
    const uint64_t k = h[0];
    ff_dfree(key);
    if (!k) {
        x.n = 1, x.bits = 0;
    } else {
        x.n = k >> 8;
        x.bits = 64 * (x.n - 1) + (k & 0xff);
    }
}

// out = a + b' on n limbs (one node); returns the carry-out, read back
// only if want_cout.
static FFNum add_one(const FFNum &a, const uint64_t *b, uint64_t nb, int mode, uint64_t n, int *cout_out) {
    FFNum r = ff_num_alloc(n);
    void *scr = ff_dalloc(ff_add_nodes_scratch(1, n) + 256);
    uint8_t *cout = (uint8_t *)scr + ff_add_nodes_scratch(1, n);
    FFAddNodes J{};
    J.a = a.d, J.na = a.n < n ? a.n : n;
    J.b = b, J.nb = nb < n ? nb : n;
    J.b_mode = mode;
    J.out = r.d, J.n = n, J.N = 1;
    J.cout = cout;
    FF_CHECK(ff_add_nodes(J, scr, g_st));
    if (cout_out) {
        uint64_t *h = ff_pinned_results(1);
        FF_CHECK(hipMemcpyAsync(h, cout, 1, hipMemcpyDeviceToHost, g_st));
        FF_CHECK(hipStreamSynchronize(g_st));
        *cout_out = (int)(h[0] & 1);
    }
    ff_dfree(scr);
    r.n = n;
    return r;
}

FFNum ff_num_add(const FFNum &a, const FFNum &b) {
    const uint64_t n = (a.n > b.n ? a.n : b.n) + 1;
    FFNum r = add_one(a, b.d, b.n, FF_ADD, n, nullptr);
    ff_num_normalize(r);
    return r;
}

FFNum ff_num_sub(const FFNum &a, const FFNum &b) {
    FFNum r = add_one(a, b.d, b.n, FF_SUB, a.n, nullptr);
    ff_num_normalize(r);
    return r;
}

int ff_num_cmp(const FFNum &a, const FFNum &b) {
    if (a.bits != b.bits) return a.bits < b.bits ? -1 : 1;
    if (a.bits == 0) return 0;
    int ge = 0;
    FFNum r = add_one(a, b.d, b.n, FF_SUB, a.n, &ge);
    int res;
    if (!ge) {
        res = -1;
    } else {
        ff_num_normalize(r);
        res = r.bits ? 1 : 0;
    }
    ff_num_free(r);
    return res;
}

FFNum ff_num_add_small(const FFNum &a, int64_t v) {
    const uint64_t mag = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
    uint64_t *dv = (uint64_t *)ff_dalloc(8);
    FF_CHECK(ff_launch_fill(dv, 1, mag, g_st));
    FFNum r = add_one(a, dv, 1, v < 0 ? FF_SUB : FF_ADD, a.n + (v < 0 ? 0 : 1), nullptr);
    ff_dfree(dv);
    ff_num_normalize(r);
    return r;
}
