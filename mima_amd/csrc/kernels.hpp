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

// ---------------------------------------------------------------------------
// backward (vector-Jacobian product) of one call
struct VjpCot {
    const double *uflx, *dflx, *hr;  // cotangents of the outputs (NULL = zero)
};
struct LwGrad {
    double *tlay, *tlev, *h2o, *o3, *tsfc;  // overwritten
};
struct SwGrad {
    double *tlay, *h2o, *o3;
};
struct LwVjpWork {
    void *dcoef = nullptr;    // [LWF_COUNT][4][nlay][ncol] of Real: value, d/dT, d/dq, d/dO3
    void *tlev = nullptr;     // [nlay+1][ncol] of Real
    int *ipk = nullptr;       // [nlay][ncol]
    int *ilev = nullptr;      // [nlay+1][ncol]
    int *laytrop = nullptr;   // [ncol]
    double *colscal = nullptr;// [2][ncol]: pwvcm, tbound
    void *cot = nullptr;      // [2][nlay+1][ncol] of Real: up / down flux cotangents incl. heating
    void *gpart = nullptr;    // [16][3][nlay][ncol] of Real: per-band gradient (T, q, O3)
    void *glev = nullptr;     // [16][nlay+1][ncol] of Real: per-band gradient wrt level temperature
    void *gsc = nullptr;      // [16][2][ncol] of Real: per-band d/dtsfc and d/dpwvcm
};
struct SwVjpWork {
    void *dcoef = nullptr;    // [SWF_COUNT][4][nlay][ncol] of Real
    int *ipk = nullptr;
    int *laytrop = nullptr;
    int *laysolfr = nullptr;  // [14][ncol]
    void *cot = nullptr;      // [2][nlay+1][ncol] of Real
    void *gpart = nullptr;    // [14][3][nlay][ncol] of Real
};
// LDS bytes of one backward band workgroup
inline size_t lw_vjp_lds_bytes(int nlay, size_t real_size) { return real_size * ((size_t)3 * nlay * 64 + 4 * 64); }
inline size_t sw_vjp_lds_bytes(int nlay, size_t real_size) { return real_size * ((size_t)5 * (nlay + 1) * 64 + 4 * 64); }

hipError_t launch_lw_vjp(const ColumnInputs &in, const VjpCot &cot, const LwGrad &out, const LwVjpWork &w,
                         const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev);
hipError_t launch_sw_vjp(const ColumnInputs &in, const VjpCot &cot, const SwGrad &out, const SwVjpWork &w,
                         const SwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev);

// mode (the C ABI's precision): 0 FP64 everywhere; 1 FP32 with FP64 islands;
// 2 FP64 with exp(-x) and the tau transition function evaluated directly
// instead of read from the 10000-point tables (smooth in the inputs).  ev: NULL, or four
// events recorded on `s` before prep, after prep, after the band kernel and
// after finish.
hipError_t launch_sw(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t,
                     int mode, hipStream_t s, hipEvent_t *ev);
hipError_t launch_lw(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t,
                     int mode, hipStream_t s, hipEvent_t *ev);

}  // namespace rrtmg
