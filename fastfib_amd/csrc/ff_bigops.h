// Limb arithmetic on device numbers beside the multiply: shifts, additions
// and subtractions with parallel carry lookahead, comparison, powers of two
// (the reference's bigint.c helpers, bigint.c:20-341, on the GPU).
//
// Two layers. FFAddNodes is the batched primitive: N numbers ("nodes") of
// uniform size laid out with strides, out = a + b, a - b or a - 2b modulo
// 2^(64 n), each node's carry-out reported; the decimal conversion runs its
// levels through it without waiting for the device. FFNum operations are
// single normalized numbers with exact bit lengths; those whose length
// depends on the data read it back (they wait for the device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// b_mode of FFAddNodes
#define FF_ADD 0     // a + b
#define FF_SUB 1     // a - b (two's complement: a + ~b + 1)
#define FF_SUB2X 2   // a - 2b

struct FFAddNodes {
    const uint64_t *a;  // node j: a + j * sa, na limbs (zero above)
    uint64_t sa, na;
    const uint64_t *b;  // node j: b + j * sb, nb limbs (zero above); sb = 0 shares one b
    uint64_t sb, nb;
    int b_mode;
    uint64_t *out;  // node j: out + j * so, n limbs
    uint64_t so, n;
    uint32_t N;
    uint32_t nblk;   // blocks of 256 limbs per node: ceil(n / 256)
    uint32_t *bgp;   // N * nblk block summaries: generate | propagate << 1
    uint8_t *bcin;   // N * nblk exact carries into blocks
    uint8_t *cout;   // N carry-outs (may be null)
};

// Scratch bytes FFAddNodes needs for bgp and bcin (one allocation).
size_t ff_add_nodes_scratch(uint32_t N, uint64_t n);
// Fills the derived fields (nblk, bgp, bcin from scratch) and launches.
hipError_t ff_add_nodes(FFAddNodes J, void *scratch, hipStream_t st);

// out_j = c_j2 ? d2_j : c_j1 ? d1_j : r_j (n limbs per node, strides so).
hipError_t ff_select3(const uint64_t *r, const uint64_t *d1, const uint64_t *d2, const uint8_t *c1,
                      const uint8_t *c2, uint64_t sr, uint64_t n, uint32_t N, uint64_t *out, uint64_t so,
                      hipStream_t st);
// out_j = (q_j + c1_j + c2_j), q_j of nq limbs copied into n >= nq limbs.
hipError_t ff_copy_add_small(const uint64_t *q, uint64_t sq, uint64_t nq, const uint8_t *c1, const uint8_t *c2,
                             uint32_t N, uint64_t *out, uint64_t so, uint64_t n, hipStream_t st);
// out_j = a_j >> k (bits), nout limbs per node (strides sa, so).
hipError_t ff_shr_nodes(const uint64_t *a, uint64_t sa, uint64_t na, uint64_t k, uint32_t N, uint64_t *out,
                        uint64_t so, uint64_t nout, hipStream_t st);

// A normalized device number: n limbs (top limb nonzero unless zero, then
// n = 1), bits its exact bit length, cap the limbs allocated.
struct FFNum {
    uint64_t *d;
    uint64_t n, bits, cap;
};

FFNum ff_num_alloc(uint64_t cap);
void ff_num_free(FFNum &x);
FFNum ff_num_u64(uint64_t v);
FFNum ff_num_copy(const FFNum &x);
FFNum ff_num_pow2(uint64_t k);
FFNum ff_num_shr(const FFNum &x, uint64_t k);
FFNum ff_num_shl(const FFNum &x, uint64_t k);
FFNum ff_num_add(const FFNum &a, const FFNum &b);
FFNum ff_num_sub(const FFNum &a, const FFNum &b);  // requires a >= b
int ff_num_cmp(const FFNum &a, const FFNum &b);     // -1, 0, 1
FFNum ff_num_add_small(const FFNum &a, int64_t v);  // requires a + v >= 0
// a * b (b null: square), from the multiply engine.
FFNum ff_num_mul(const FFNum &a, const FFNum *b);
// Normalizes x's length and bits from the device (waits).
void ff_num_normalize(FFNum &x);
