// Device helpers shared by the SW and LW kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace rrtmg {

__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Index of the 10000-interval lookup tables for optical depth x:
// int(tblint * x/(bpade + x) + 0.5)  (rrtmg_lw_rtrnmr.f90:597-598,
// rrtmg_sw_reftra.f90:182-183)
template <typename F>
__device__ __forceinline__ int tbl_index(F x)
{
    const F bpade = (F)(1.0 / 0.278);
    int i = (int)((F)1.e4 * (x / (bpade + x)) + (F)0.5);
    return i < 0 ? 0 : (i > 10000 ? 10000 : i);  // keeps NaN/huge inputs inside the table
}

// Sum of v over the ng adjacent lanes of each column group of a one-wave
// block; valid on the group's first lane.  Deterministic (fixed order); FP32
// sums are compensated (Kahan).  `red` is a 64-entry LDS buffer.
template <typename Real>
__device__ __forceinline__ Real seg_sum(Real v, int lane, int ng, Real *red)
{
    red[lane] = v;
    __syncthreads();
    Real s = (Real)0;
    if (lane % ng == 0 && lane + ng <= 64) {
        if (sizeof(Real) == 4) {
            Real comp = (Real)0;
            for (int j = 0; j < ng; ++j) {
                Real y = red[lane + j] - comp;
                Real t = s + y;
                comp = (t - s) - y;
                s = t;
            }
        } else {
            for (int j = 0; j < ng; ++j) s += red[lane + j];
        }
    }
    __syncthreads();
    return s;
}

}  // namespace rrtmg
