// Linear limb passes (last factors, small additions, fills): shared structs
// and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// X = 2u + v, Y = 2u - v (Y may be NULL) on len limbs, with length keys.
struct FFLastArgs {
    const uint64_t *u, *v;
    uint64_t nu, nv, len;
    uint64_t *X, *Y;
    unsigned long long *keyX, *keyY;
    uint32_t *lb;  // look-back workspace, ff_last_lb_words(len) zeroed words
};

size_t ff_last_lb_words(uint64_t len);

extern "C" {
hipError_t ff_launch_last(const FFLastArgs *a, hipStream_t st);
// d[0..len) += v, then the top limb (among the last three) rebooked.
hipError_t ff_launch_add_small(uint64_t *d, uint64_t len, int64_t v, unsigned long long *lenkey, hipStream_t st);
hipError_t ff_launch_fill(uint64_t *d, uint64_t n, uint64_t v, hipStream_t st);
}
