// Input validation of device-pointer calls on the device (opt-in,
// rrtmg_hip_set_checks): the same checks the host-pointer path makes on the
// host before a call -- every input finite, plev decreasing strictly upward
// and positive at the top.  Each check lowers a flag to the number of the
// first array that fails, so the host can name it.
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.hpp"

namespace rrtmg {

__global__ void flag_init_kernel(int *flag) { *flag = INT_MAX; }

__global__ __launch_bounds__(256) void check_finite_kernel(const double *__restrict__ a, size_t n, int code, int *flag)
{
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !isfinite(a[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicMin(flag, code);
}

// plev (ncol, nlay+1): plev(i, k) > plev(i, k+1) for every k, plev(i, nlay) > 0
__global__ __launch_bounds__(256) void check_plev_kernel(const double *__restrict__ plev, int ncol, int nlay, int code,
                                                         int *flag)
{
    bool bad = false;
    const size_t n = (size_t)ncol * nlay;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n + ncol; i += (size_t)gridDim.x * blockDim.x) {
        if (i < n)
            bad |= !(plev[i] > plev[i + ncol]);
        else
            bad |= !(plev[i] > 0.0);
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicMin(flag, code);
}

hipError_t launch_input_checks(const DeviceArray *arrays, int narr, const double *plev, int ncol, int nlay, int *flag,
                               hipStream_t s)
{
    hipLaunchKernelGGL(flag_init_kernel, dim3(1), dim3(1), 0, s, flag);
    for (int k = 0; k < narr; ++k) {
        if (!arrays[k].p || !arrays[k].n) continue;
        const size_t blocks = (arrays[k].n + 255) / 256;
        hipLaunchKernelGGL(check_finite_kernel, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(256), 0, s,
                           arrays[k].p, arrays[k].n, k, flag);
    }
    if (plev) {
        const size_t n = (size_t)ncol * (nlay + 1), blocks = (n + 255) / 256;
        hipLaunchKernelGGL(check_plev_kernel, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(256), 0, s, plev,
                           ncol, nlay, narr, flag);
    }
    return hipGetLastError();
}

}  // namespace rrtmg
