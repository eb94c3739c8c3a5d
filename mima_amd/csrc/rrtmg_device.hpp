// Device helpers shared by the SW and LW kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace rrtmg {

__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Index of the 10000-interval lookup tables for optical depth x:
// int(tblint * x/(bpade + x) + 0.5)  (rrtmg_lw_rtrnmr.f90:597-598,
// rrtmg_sw_reftra.f90:182-183)
// a / b in the arithmetic of the path: FP64 exactly (the reference's
// division); FP32 as a times the hardware reciprocal (1 ulp), the FP32-mixed
// path's division (an IEEE FP32 division is a ten-instruction sequence);
// anything else (dual numbers) with its own operator.
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename A, typename B>
__device__ __forceinline__ auto fdiv(const A &a, const B &b) -> decltype(a / b)
{
    return a / b;
}

template <typename F>
__device__ __forceinline__ int tbl_index(F x)
{
    const F bpade = (F)(1.0 / 0.278);
    int i = (int)((F)1.e4 * fdiv(x, bpade + x) + (F)0.5);
    return i < 0 ? 0 : (i > 10000 ? 10000 : i);  // keeps NaN/huge inputs inside the table
}

// Stores the records of a block's threads, staged in LDS (thread t's record
// of R elements at srec[t * R]), to global memory, where the records of the
// block's first nrec threads are contiguous from dst: consecutive threads
// store consecutive 16-byte vectors.  From each thread directly, the stores
// would be 16-byte pieces at the record's stride, one L2 request each.  Every
// thread of the block calls it, after writing its own record.
template <typename Real, int R>
__device__ __forceinline__ void store_records(Real *__restrict__ dst, const Real *srec, int nrec)
{
    static_assert(R * sizeof(Real) % 16 == 0, "records are whole 16-byte vectors");
    __syncthreads();
    const int nvec = nrec * (int)(R * sizeof(Real) / 16);
    const uint4 *s = (const uint4 *)srec;
    uint4 *d = (uint4 *)dst;
    for (int k = (int)threadIdx.x; k < nvec; k += (int)blockDim.x) d[k] = s[k];
}

// XCD-aware block numbering.  The 8 XCDs of the chip each have their own L2
// and receive consecutive workgroups round-robin (workgroup i runs on XCD
// i mod 8).  Renumbering i -> (i mod 8) * (grid/8) + i/8 hands each XCD a
// contiguous range of logical blocks, so the wave units of one column tile,
// which share that tile's coefficient fields, run behind the same L2.  The
// grid is padded to a multiple of 8 (xcd_grid); logical blocks >= nblocks
// have no work.
__host__ __device__ __forceinline__ int xcd_grid(int nblocks) { return (nblocks + 7) / 8 * 8; }
__device__ __forceinline__ int xcd_block()
{
    const int per = (int)gridDim.x / 8;
    return ((int)blockIdx.x % 8) * per + (int)blockIdx.x / 8;
}

// Element i of an array whose uniform base sits in SGPRs, addressed by a
// 32-bit unsigned byte offset (one global access with a VGPR offset and no
// per-lane 64-bit address arithmetic).  Valid for arrays below 4 GiB, which
// the C ABI checks for the per-field arrays it passes to the band kernels.
template <typename T>
__device__ __forceinline__ T &at32(T *base, unsigned i)
{
    return *(T *)((char *)base + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &at32(const T *base, unsigned i)
{
    return *(const T *)((const char *)base + i * (unsigned)sizeof(T));
}

// Zeroes the work-item counter of a persistent band kernel (a kernel, not a
// memset, so that the sequence can be captured in a graph).
// Pointers typed by address space (global memory, LDS).  A called function
// knows of a generic pointer argument only that it is flat: its accesses
// compile to flat_* instructions, which wait on both the vector-memory and
// the LDS counters.  Arguments typed by address space keep global_* and ds_*.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
using lptr = __attribute__((address_space(3))) T *;

[[maybe_unused]] static __global__ void reset_items_kernel(int *items) { *items = 0; }

// Next work item of a persistent workgroup, claimed from a global counter
// (dynamic balance: items differ in cost, e.g. night columns in SW); every
// wave of the workgroup gets the same item.  The leading barrier also
// guarantees that all waves are done with the previous item.
__device__ __forceinline__ int next_item(int *items, int *s_item)
{
    __syncthreads();
    if (threadIdx.x == 0) *s_item = atomicAdd(items, 1);
    __syncthreads();
    return *s_item;
}

// Sum of v over the ng adjacent lanes of each column group of a one-wave
// block (g = the lane's position in its group); valid on the group's first
// lane.  Deterministic (fixed order).  FP64: sequential through LDS, the
// reference's order over g-points; FP32 (the g-point flux sums): a
// compensated pairwise tree of cross-lane shuffles, log2 ng steps, each
// addition's rounding error recovered exactly (Knuth's TwoSum) and the
// errors summed beside it and added once at the end, so the sum is good to
// about one rounding instead of log2(ng).  `red` is a 64-entry LDS buffer
// (FP64 only).  Called by all 64 lanes.
template <typename Real>
__device__ __forceinline__ Real seg_sum(Real v, int lane, int g, int ng, Real *red)
{
    if constexpr (sizeof(Real) == 4) {
        Real e = (Real)0;
        for (int s = 1; s < ng; s <<= 1) {
            const Real o = __shfl_down(v, s);
            const Real oe = __shfl_down(e, s);
            if (g + s < ng) {
                const Real t = v + o;
                const Real bp = t - v;
                const Real err = (v - (t - bp)) + (o - bp);
                v = t;
                e = e + (oe + err);
            }
        }
        return g == 0 ? v + e : (Real)0;
    } else {
        red[lane] = v;
        __syncthreads();
        Real s = (Real)0;
        if (g == 0 && lane + ng <= 64)
            for (int j = 0; j < ng; ++j) s += red[lane + j];
        __syncthreads();
        return s;
    }
}

// seg_sum with the lane arithmetic done once per work unit (the forward band
// kernels call it for every level): the byte address of each step's
// partner lane for ds_bpermute and whether the lane adds at that step, and
// selects instead of a branch.  Bitwise the same sums as seg_sum.  A lane
// whose partner lies past lane 63 never adds (its group would extend past
// the wave), so the wrapped address only feeds lanes whose sum is unused.
template <typename Real, int NG>
struct GSum {
    static constexpr int STEPS = NG <= 2 ? 1 : NG <= 4 ? 2 : NG <= 8 ? 3 : 4;
    int addr[STEPS];
    bool take[STEPS];
    int lane, g;
    __device__ GSum(int lane_, int g_) : lane(lane_), g(g_)
    {
#pragma unroll
        for (int k = 0; k < STEPS; ++k) {
            addr[k] = ((lane_ + (1 << k)) & 63) << 2;
            take[k] = g_ + (1 << k) < NG;
        }
    }
    __device__ __forceinline__ Real sum(Real v, Real *red) const
    {
        if constexpr (sizeof(Real) == 4) {
            Real e = (Real)0;
#pragma unroll
            for (int k = 0; k < STEPS; ++k) {
                const Real o = __int_as_float(__builtin_amdgcn_ds_bpermute(addr[k], __float_as_int(v)));
                const Real oe = __int_as_float(__builtin_amdgcn_ds_bpermute(addr[k], __float_as_int(e)));
                const Real t = v + o;
                const Real bp = t - v;
                const Real err = (v - (t - bp)) + (o - bp);
                e = take[k] ? e + (oe + err) : e;
                v = take[k] ? t : v;
            }
            return g == 0 ? v + e : (Real)0;
        } else {
            return seg_sum(v, lane, g, NG, red);
        }
    }
};

// seg_sum of N values at once; v[i] is valid on the group's first lane.
// `red` holds N x 64 entries (FP64 only).
template <int N, typename Real>
__device__ __forceinline__ void seg_sum_n(Real (&v)[N], int lane, int g, int ng, Real *red)
{
    if constexpr (sizeof(Real) == 4) {
        for (int s = 1; s < ng; s <<= 1)
            for (int i = 0; i < N; ++i) {
                const Real o = __shfl_down(v[i], s);
                if (g + s < ng) v[i] += o;
            }
    } else {
        for (int i = 0; i < N; ++i) red[i * 64 + lane] = v[i];
        __syncthreads();
        if (g == 0 && lane + ng <= 64) {
            for (int i = 0; i < N; ++i) {
                Real s = (Real)0;
                for (int j = 0; j < ng; ++j) s += red[i * 64 + lane + j];
                v[i] = s;
            }
        }
        __syncthreads();
    }
}

}  // namespace rrtmg
