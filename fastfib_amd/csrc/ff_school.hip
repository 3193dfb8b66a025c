// Products below the NTT threshold by the schoolbook method: the list
// kernels over the bodies of ff_school_dev.h (one workgroup per product, or
// one wavefront per product up to 128 output limbs).
#include <stdint.h>

#include "ff_internal.h"
#include "ff_runtime.h"
#include "ff_school_dev.h"

namespace {

__global__ __launch_bounds__(FF_SCHOOL_T) void k_school_wave_list(const FFSchoolItem *items, int n) {
    __shared__ FFSchoolWaveShared sh;
    const int w = threadIdx.x >> 6;
    const int k = blockIdx.x * 4 + w;
    const bool live = k < n;
    school_wave(items[live ? k : 0], live, sh.A[w], sh.B[w], sh.lo[w], sh.mid[w], sh.hi[w]);
}

__global__ __launch_bounds__(FF_SCHOOL_T) void k_school_list(const FFSchoolItem *items, int n) {
    (void)n;
    __shared__ FFSchoolShared sh;
    school_body(items[blockIdx.x], sh);
}

__global__ __launch_bounds__(FF_SCHOOL_T) void k_school_set(FFSchoolSet set) {
    __shared__ FFSchoolShared sh;
    school_body(set.j[blockIdx.x], sh);
}

}  // namespace

extern "C" hipError_t ff_launch_school_wave_list(const FFSchoolItem *items, int n, double bytes, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    ff_prof_begin(FK_SCHOOL, st);
    hipLaunchKernelGGL(k_school_wave_list, dim3((unsigned)((n + 3) / 4)), dim3(FF_SCHOOL_T), 0, st, items, n);
    ff_prof_end(FK_SCHOOL, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_school_set(const FFSchoolSet *set, int n, double bytes, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    ff_prof_begin(FK_SCHOOL, st);
    hipLaunchKernelGGL(k_school_set, dim3((unsigned)n), dim3(FF_SCHOOL_T), 0, st, *set);
    ff_prof_end(FK_SCHOOL, bytes, st);
    return hipGetLastError();
}

extern "C" hipError_t ff_launch_school_list(const FFSchoolItem *items, int n, double bytes, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    ff_prof_begin(FK_SCHOOL, st);
    hipLaunchKernelGGL(k_school_list, dim3((unsigned)n), dim3(FF_SCHOOL_T), 0, st, items, n);
    ff_prof_end(FK_SCHOOL, bytes, st);
    return hipGetLastError();
}
