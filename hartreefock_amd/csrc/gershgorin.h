// Block-wide Gershgorin bound shared by the SP2 kernels of dense.hip and scf.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace hfgpu {

// Gershgorin interval [lo, hi] of a symmetric n x n row-major matrix, by one block of 1024 threads:
// 64 consecutive columns per pass (coalesced rows), 16 row groups per column, partial row sums
// combined in a fixed order.  part: at least 1024 doubles of LDS (free again on return).  Every
// thread returns the block's lo and hi.
__device__ __forceinline__ void gershgorin_block(const double* __restrict__ F, int n, double* part, double& lo_out,
                                                 double& hi_out) {
    const int t = threadIdx.x, c = t & 63, g = t >> 6;
    double lo = 1e300, hi = -1e300;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + c;
        double r = 0;
        if (j < n)
            for (int i = g; i < n; i += 16)
                if (i != j) r += fabs(F[(size_t)i * n + j]);
        part[t] = r;
        __syncthreads();
        if (g == 0 && j < n) {
            double s = 0;
            for (int k = 0; k < 16; ++k) s += part[k * 64 + c];
            const double d = F[(size_t)j * n + j];
            lo = fmin(lo, d - s);
            hi = fmax(hi, d + s);
        }
        __syncthreads();
    }
    if (g == 0) {                      // the first wavefront holds every column's bounds
        for (int m = 32; m >= 1; m >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, m));
            hi = fmax(hi, __shfl_xor(hi, m));
        }
        if (c == 0) {
            part[0] = lo;
            part[1] = hi;
        }
    }
    __syncthreads();
    lo_out = part[0];
    hi_out = part[1];
    __syncthreads();
}

}  // namespace hfgpu
