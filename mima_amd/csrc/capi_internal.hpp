// Library-internal entry points shared between the C ABI's translation units
// (capi.cpp and noalloc.hip); not exported through include/rrtmg_hip.h.
#pragma once
#include <cstddef>

namespace rrtmg {

struct SwClouds;

// A device region that a call takes its whole workspace from instead of the
// library's growing per-stream buffers.
struct FixedWorkspace {
    void *p;     // device address (256-byte aligned)
    size_t cap;  // bytes
};

// rrtmg_hip_sw's body; fx NULL: the per-stream workspace (allocated on demand);
// fx given: device pointers only, no allocation, workspace carved from fx.
// cl: the cloud arrays of rrtmg_hip_sw_cloud (icld >= 1), or NULL (clear
// sky); not with fx.
int sw_forward(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev, const double *tlay,
               const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
               const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
               const double *asdir, const double *asdif, const double *aldir, const double *aldif,
               const double *coszen, double adjes, int dyofyr, double scon, double *swuflx, double *swdflx,
               double *swhr, double *swuflxc, double *swdflxc, double *swhrc, int ptr_kind, void *stream,
               const FixedWorkspace *fx, const SwClouds *cl = nullptr);

// Looks up the current device's address of the no-allocation scratch array
// (once per device; called by rrtmg_hip_sw_init with the device current).
// 0 on success.
int noalloc_attach();

}  // namespace rrtmg
