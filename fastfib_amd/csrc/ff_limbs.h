// Linear limb passes (fast-doubling step, last factors, carries): shared
// structs and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FF_CH 16          // limbs per chunk of the linear passes
#define FF_LIN_THREADS 256

struct FFCarrySum {
    uint64_t limb0;
    int8_t e;      // signed carry out of the chunk (zero carry-in)
    uint8_t ones;  // limbs 1.. of the chunk are all ones
    uint8_t zero;  // limbs 1.. of the chunk are all zero
    uint8_t pad_;
    uint32_t pad2_;
};

// A = s1 + s2, B = 4 s1 - s2 + cb0, C = B - A on len limbs; A or B is NULL
// (not kept). sums: 2 streams x nch chunks (stream 0: the kept one of A/B).
struct FFStepArgs {
    const uint64_t *s1, *s2;
    uint64_t n1, n2, len;
    uint64_t *A, *B, *C;
    int64_t cb0;
    FFCarrySum *sums;
    uint32_t *wgflag;  // per workgroup: a chunk (not the last) could pass on a carry
};

// X = 2u + v, Y = 2u - v (Y may be NULL) on len limbs.
struct FFLastArgs {
    const uint64_t *u, *v;
    uint64_t nu, nv, len;
    uint64_t *X, *Y;
    FFCarrySum *sums;
    uint32_t *wgflag;
};

// The carry resolution of one or two outputs of a linear pass.
struct FFCarryStreams {
    uint64_t *d[2];     // the arrays being fixed
    uint64_t len;       // limbs of each
    uint64_t nch;       // chunks per stream
    int nstreams;
    const FFCarrySum *sums;  // nstreams x nch
    int8_t *cin;             // nstreams x nch (exact path)
    const uint32_t *wgflag;  // from the local pass, one per workgroup
    uint32_t *mode;          // written by the scan: 1 = exact carries in cin
    unsigned long long *lenkey[2];  // (top nonzero limb + 1) << 8 | its bits
    uint64_t lo_top[2];      // the top nonzero limb is at or above this index
};

extern "C" {
hipError_t ff_launch_step(const FFStepArgs *a, hipStream_t st);
hipError_t ff_launch_last(const FFLastArgs *a, hipStream_t st);
hipError_t ff_launch_carry(const FFCarryStreams *cs, hipStream_t st);
hipError_t ff_launch_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey, uint64_t lo_top,
                               hipStream_t st);
}

extern "C" hipError_t ff_launch_fill(uint64_t *d, uint64_t n, uint64_t v, hipStream_t st);
