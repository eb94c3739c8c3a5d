// Evaluation of the smooth functions that the lookup tables stand for (the
// smooth FP64 mode's forward and every mode's backward differentiate these),
// at given arguments, for parity checks against the tables at their nodes:
// LW 1 - exp(-tau) and the tau transition function (lw_rt_exact; tables of
// rrtmg_lw_init.f90:106-123) and SW exp(-x) floored at 1e-20 (tbl_exp with no
// table; rrtmg_sw_init.f90:94-104), each with its derivative.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lw_optics.hpp"
#include "sw_optics.hpp"

namespace rrtmg {

__global__ __launch_bounds__(256) void smooth_functions_kernel(int n, const double *__restrict__ x,
                                                               double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LwRT<double> r = lw_rt_exact(x[i], true);
    const Dual<double, 1> e = tbl_exp(Dual<double, 1>::seed(x[i], 0), (const double *)nullptr);
    double *o = out + (size_t)6 * i;
    o[0] = r.atrans;
    o[1] = r.tf;
    o[2] = r.datr;
    o[3] = r.dtf;
    o[4] = e.v;
    o[5] = e.d[0];
}

hipError_t launch_smooth_functions(int n, const double *x, double *out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(smooth_functions_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, x, out);
    return hipGetLastError();
}

}  // namespace rrtmg
