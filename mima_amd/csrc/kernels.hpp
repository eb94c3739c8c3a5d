// Kernel launch interfaces shared by the C ABI (capi.cpp) and the device code.
#pragma once
#include <hip/hip_runtime.h>

#include "rrtmg_types.hpp"

namespace rrtmg {

// Device-resident FP64 inputs of one call, (ncol, nlay) column-major.
struct ColumnInputs {
    int ncol, nlay;
    const double *play, *plev, *tlay, *tlev, *tsfc;
    const double *h2o, *o3, *co2, *ch4, *n2o, *o2;
    // SW only
    const double *asdir, *asdif, *aldir, *aldif, *coszen;
    double adjflux;  // adjes (or earth_sun) * scon / 1368.22
    // LW only
    const double *cfc11, *cfc12, *cfc22, *ccl4, *emis, *tauaer;
};

struct SwOutputs {
    double *uflx, *dflx, *hr, *uflxc, *dflxc, *hrc;
};
struct LwOutputs {
    double *uflx, *dflx, *hr, *uflxc, *dflxc, *hrc;
};

// Per-call device workspace (owned by the library, grown on demand).
struct SwWork {
    void *coef = nullptr;     // [SWF_COUNT][nlay][ncol] of Real
    int *ipk = nullptr;       // [nlay][ncol]
    int *laytrop = nullptr;   // [ncol]
    int *laysolfr = nullptr;  // [14][ncol]
    void *bflux = nullptr;    // [14][2][nlay+1][ncol] of Real
};
struct LwWork {
    void *coef = nullptr;     // [LWF_COUNT][nlay][ncol] of Real
    void *tlev = nullptr;     // [nlay+1][ncol] of Real (level Planck interpolation fraction)
    int *ipk = nullptr;       // [nlay][ncol]
    int *ilev = nullptr;      // [nlay+1][ncol] Planck table index of levels
    int *laytrop = nullptr;   // [ncol]
    double *colscal = nullptr;// [3][ncol]: pwvcm, tbound, (spare)
    void *bflux = nullptr;    // [16][2][nlay+1][ncol] of Real
};

struct SwDevTables {
    SwTables desc;
    const SwBand *bands;      // device copy of desc.band
    const void *tab;          // packed tables in the mode's precision
    const double *preflog, *tref;  // [59] FP64
    const WaveUnit *units;
    double heatfac;
};
struct LwDevTables {
    LwTables desc;
    const LwBand *bands;      // device copy of desc.band
    const void *tab;
    const LwConst *cst;       // device copy
    const WaveUnit *units;
    double heatfac;
};

// mixed = 0: FP64 everywhere; 1: FP32 with FP64 islands.  ev: NULL, or four
// events recorded on `s` before prep, after prep, after the band kernel and
// after finish.
hipError_t launch_sw(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t,
                     int mixed, hipStream_t s, hipEvent_t *ev);
hipError_t launch_lw(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t,
                     int mixed, hipStream_t s, hipEvent_t *ev);

}  // namespace rrtmg
