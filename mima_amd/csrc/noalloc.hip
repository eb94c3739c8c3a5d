// rrtmg_hip_sw_noalloc: the SW forward (rrtmg_hip_sw's kernels) for callers
// that cannot allocate device memory -- code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc.  Its workspace (coefficient
// records, table indices, band fluxes, the band kernel's scratch rows and the
// input-check flag) is carved from `scratch`, a file-scope device array that
// every GPU holds for the life of the process.  No reference counterpart.
//
// `scratch` named in host code is a host-side stand-in, not a device
// address; the kernels only ever receive the address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device by rrtmg_hip_sw_init.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rrtmg_hip.h"
#include "capi_internal.hpp"

static __device__ __align__(256) unsigned char scratch[RRTMG_HIP_NOALLOC_BYTES];

namespace {

constexpr int MAX_DEVICES = 256;
std::mutex g_mu;
void *g_scratch[MAX_DEVICES] = {};  // per-device address of `scratch`

}  // namespace

namespace rrtmg {

int noalloc_attach()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_scratch[dev] = p;
    return 0;
}

}  // namespace rrtmg

extern "C" int rrtmg_hip_sw_noalloc(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev,
                                    const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                                    const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                                    const double *n2ovmr, const double *o2vmr, const double *asdir,
                                    const double *asdif, const double *aldir, const double *aldif,
                                    const double *coszen, double adjes, int dyofyr, double scon, double *swuflx,
                                    double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc,
                                    void *stream)
{
    int dev = 0;
    void *p = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < MAX_DEVICES) {
        std::lock_guard<std::mutex> lk(g_mu);
        p = g_scratch[dev];
    }
    const rrtmg::FixedWorkspace fx{p, (size_t)RRTMG_HIP_NOALLOC_BYTES};
    return rrtmg::sw_forward(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                             n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx, swdflx,
                             swhr, swuflxc, swdflxc, swhrc, RRTMG_HIP_DEVICE, stream, &fx);
}
