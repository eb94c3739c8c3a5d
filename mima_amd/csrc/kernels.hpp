// Kernel launch interfaces shared by the C ABI (capi.cpp) and the device code.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rrtmg_hip.h"
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
    // 1: h2o and o3 are volume mixing ratios (the McICA LW driver's inatm);
    // 0: specific humidity and O3 mass mixing ratio (MiMA's nomcica drivers)
    int vmr;
};

struct SwOutputs {
    double *uflx, *dflx, *hr, *uflxc, *dflxc, *hrc;
};
struct LwOutputs {
    double *uflx, *dflx, *hr, *uflxc, *dflxc, *hrc;
};

// Per-call device workspace (owned by the library, grown on demand).
struct SwWork {
    void *coef = nullptr;     // [nlay][ncol][SWF_STRIDE] of Real (one record per layer and column)
    int *ipk = nullptr;       // [nlay][ncol]
    int *laytrop = nullptr;   // [ncol]
    int *laysolfr = nullptr;  // [14][ncol]
    void *bflux = nullptr;    // [14][2][nlay+1][ncol] of Real
    void *dstate = nullptr;   // [chunk][nlay+1][64] of Real: prupd of each wave of a band-kernel launch
    int chunk = 0;            // wave units per band-kernel launch (rows of dstate)
};
struct LwWork {
    void *coef = nullptr;     // [nlay][ncol][LWF_STRIDE] of Real (one record per layer and column)
    void *tlev = nullptr;     // [nlay+1][ncol] of Real (level Planck interpolation fraction)
    int *ipk = nullptr;       // [nlay][ncol]
    int *ilev = nullptr;      // [nlay+1][ncol] Planck table index of levels
    int *laytrop = nullptr;   // [ncol]
    double *colscal = nullptr;// [3][ncol]: pwvcm, tbound, (spare)
    void *bflux = nullptr;    // [16][2][nlay+1][ncol] of Real
    void *dstate = nullptr;   // [chunk][nlay][64] of Real: upward source of each wave of a band-kernel launch
    int chunk = 0;            // wave units per band-kernel launch (rows of dstate)
    int *items = nullptr;     // work-item counter of the LDS-table band kernel
};

struct SwDevTables {
    SwTables desc;
    const SwBand *bands;      // device copy of desc.band
    const void *tab;          // packed tables in the mode's precision
    const double *preflog, *tref;  // [59] FP64
    const WaveUnit *units;
    double heatfac;
};
// Schedule of the FP32 LW band kernel that keeps one band's k-tables in LDS:
// persistent workgroups of LW_WGW waves; work item = (group of LW_WGW tiles,
// unit r of a tile's unit list), whose LW_WGW waves take the group's units of
// the band of unit r.
constexpr int LW_WGW = 16;
struct LwLdsPlan {
    int tab0[17];   // first element of each band's k-tables in the packed buffer; [16]: end of band 16's
    int ubeg[16];   // first unit of each band in a tile's unit list
    int upt[16];    // units of each band per tile
    int lds_bytes;  // largest band's k-tables in FP32
    int ncu;        // compute units: workgroups of one launch
};
// rows of per-wave global scratch (transmittance and upward source, nlay x 64
// each) the LDS-table kernel uses: one per wave of its persistent grid
inline int lw_lds_grid(const LwLdsPlan &p, int ncol, int nwaves)
{
    const int ntiles = (ncol + TILE_COLS - 1) / TILE_COLS;
    const int nitems = (ntiles + LW_WGW - 1) / LW_WGW * nwaves;
    return nitems < p.ncu ? nitems : p.ncu;
}
inline size_t lw_lds_state_bytes(int nlay) { return sizeof(float) * 2 * (size_t)nlay * 64; }
// the backward's LDS-table kernel: LWV_WGW waves per workgroup (at most 168
// VGPRs, three waves per SIMD), per-wave rows of 3 x nlay x 64
constexpr int LWV_WGW = 12;
inline int lw_vjp_lds_grid(const LwLdsPlan &p, int ncol, int nwaves)
{
    const int ntiles = (ncol + TILE_COLS - 1) / TILE_COLS;
    const int nitems = (ntiles + LWV_WGW - 1) / LWV_WGW * nwaves;
    return nitems < p.ncu ? nitems : p.ncu;
}
__host__ __device__ inline size_t lw_vjp_lds_state_bytes(int nlay) { return sizeof(float) * 3 * (size_t)nlay * 64; }

struct LwDevTables {
    LwTables desc;
    const LwBand *bands;      // device copy of desc.band
    const void *tab;
    const LwConst *cst;       // device copy
    const WaveUnit *units;
    double heatfac;
    const LwBand *bands_lds = nullptr;  // desc.band with k-table offsets relative to the band's first element
    LwLdsPlan plan{};
    bool lds = false;                   // FP32 forward uses the LDS-table kernel
};

// ---------------------------------------------------------------------------
// LW with clouds (rrtmg_lw, icld = 1..3, inflglw = 0), device FP64 inputs
struct LwClouds {
    int icld;                         // 1 random overlap (rtrn), 2 and 3 maximum-random (rtrnmr), 4 McICA (rtrnmc)
    const double *cldfr;              // [nlay][ncol]
    const double *taucld;             // [nlay][ncol][16]: the reference's taucld(nbndlw, ncol, nlay)
    const double *cicewp, *cliqwp;    // [nlay][ncol] or NULL (zero)
    const double *cldfmc, *taucmc;    // McICA subcolumns [nlay][ncol][140]: the reference's (ngptlw, ncol, nlay)
    // slant != 0: taucld holds the cloud's depth on each band's diffusivity
    // path, odcld of the transfer (zero where cldprop gave none), as
    // lw_cloud_wp_kernel forms it from water paths
    int slant;
    // cldprop's input mode (inflglw, iceflglw, liqflglw) and radii [nlay][ncol]
    int inflag, iceflag, liqflag;
    const double *reice, *reliq;
    const double *ciwpmc, *clwpmc;  // McICA with inflag 2: subcolumn water paths [nlay][ncol][140]
};
// Water-path cloud optics (inflglw 1, 2): cldprop's inputs and lwcldpr's
// coefficients, packed as abscld1, absice0[2], absice1[5][2],
// absice2[16][43], absice3[16][46], absliq0, absliq1[16][58] (LW_CLD_TAB
// doubles)
constexpr int LW_CLD_TAB = 1 + 2 + 10 + 688 + 736 + 1 + 928;
struct LwCloudWp {
    int inflag, iceflag, liqflag;
    const double *cldfr, *taucld, *cicewp, *cliqwp, *reice, *reliq;  // taucld, paths, radii may be NULL (zero)
    const double *tab;                                                // LW_CLD_TAB doubles
    double *od;                                                       // out: [nlay][ncol][16] slant depths
    int *bad;                                                         // out: layers with a radius the reference stops on
};
// after launch_lw_prep (the columns' precipitable water sets the
// diffusivities); a.bad is cleared first
hipError_t launch_lw_cloud_wp(const ColumnInputs &in, const LwWork &w, const LwCloudWp &a, const LwBand *bands,
                              hipStream_t s);
constexpr int LW_MCICA = 4;  // LwClouds::icld of the McICA driver's transfer
// McICA subcolumns of mcica_subcol_lw / mcica_subcol_sw with the KISS
// generator: inputs [nlay][ncol] (tauc, ssac, asmc, fsfc [nlay][ncol][16 or
// 14]), subcolumn outputs [nlay][ncol][nsub]
struct McicaSubcolArgs {
    int ncol, nlay, icld, permuteseed;
    int nsub;  // 140 (LW) or 112 (SW)
    const double *play, *cldfrac, *ciwp, *clwp, *rei, *rel, *tauc;  // ciwp, clwp, rei, rel may be NULL (zero)
    const double *ssac, *asmc, *fsfc;                               // SW only
    double *cldfmcl, *ciwpmcl, *clwpmcl, *reicmcl, *relqmcl, *taucmcl;  // only cldfmcl and taucmcl required
    double *ssacmcl, *asmcmcl, *fsfcmcl;                                // SW only, required there
    int *bad;  // columns whose pressure rises from the first layer to the second; NULL: not counted
    uint32_t *mt;  // irng = 1: the Mersenne Twister stream (mcica_mt_count values); NULL: KISS
};
// length of the Mersenne Twister stream one generator call draws: one value
// per (subcolumn, column, layer), one per (subcolumn, column) for icld 3
inline size_t mcica_mt_count(int ncol, int nlay, int icld, int nsub)
{
    return (size_t)nsub * ncol * (icld == 3 ? 1 : nlay);
}
// jump constants of one subcolumn's skip over the generator's sequence
struct McicaJump {
    uint32_t a1, c1;  // congruential part: s -> a1 s + c1
    uint32_t p3, p4;  // multiply-with-carry parts: a^skip mod (a 2^16 - 1)
    int band, pad[3];
};
constexpr int KISS_XS_JUMPS = 34;  // squarings M^(2^k) of the 3-shift register's matrix, k < 34
// the KISS_XS_JUMPS x 32 words of those squarings (column j of M^(2^k) at [k*32 + j]); host code
void kiss_xorshift_jumps(uint32_t *t);
// jump: device scratch of nsub McicaJump; xs_jump: the squarings on the device
hipError_t launch_mcica_subcol(const McicaSubcolArgs &a, const uint32_t *xs_jump, McicaJump *jump, hipStream_t s);
// Per-(layer, column) record of the cloud column kernel, in the band
// kernel's precision: the layer's fraction, then rtrnmr's six upward
// overlap factors of index lay+1 and its six downward ones of index lay-1
// (1-based lay; rrtmg_lw_rtrnmr.f90:325-476).
enum LwCloudField {
    LWC_CF,
    LWC_CLD1, LWC_CLD2, LWC_CLR1, LWC_CLR2, LWC_CMB1, LWC_CMB2,
    LWC_CLD1D, LWC_CLD2D, LWC_CLR1D, LWC_CLR2D, LWC_CMB1D, LWC_CMB2D,
    LWC_COUNT
};
constexpr int LWC_STRIDE = 16;
// flag bits per (layer, column): cloudy for the transfer (fraction >= 1e-6),
// first cloudy layer of a block going up (istcld) and going down (istcldd),
// and cldprop's cloud test (the layer's taucld is used)
constexpr int LWC_CLOUDY = 1, LWC_TOP_UP = 2, LWC_TOP_DN = 4, LWC_TAU = 8;
struct LwCloudWork {
    int *flags = nullptr;     // [nlay][ncol]
    void *fac = nullptr;      // [nlay][ncol][LWC_STRIDE] of Real
    void *bflux = nullptr;    // [16][4][nlay+1][ncol] of Real: total up, total down, clear up, clear down
    void *rows = nullptr;     // [chunk][4][nlay][64] of Real: per-layer transmittances and sources of a wave
    int chunk = 0;            // wave units per band-kernel launch (rows)
    double *od = nullptr;     // water-path clouds: [nlay][ncol][16] slant depths (McICA: [nlay][ncol][140] depths)
    int *bad = nullptr;       // water-path clouds: radius count, or NULL
    const double *cldtab = nullptr;  // water-path clouds: lwcldpr's coefficients on the device
};
inline size_t lw_cloud_row_bytes(int nlay, size_t real_size) { return real_size * 4 * (size_t)nlay * 64; }
// mode 0 (FP64) or 1 (FP32 with FP64 islands); w: the forward's prep
// workspace (coef, tlev, ipk, ilev, laytrop, colscal)
hipError_t launch_lw_cloud(const ColumnInputs &in, const LwClouds &cl, const LwOutputs &out, const LwWork &w,
                           const LwCloudWork &cw, const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev);
// the clear-sky forward's prep and column kernels alone
hipError_t launch_lw_prep(const ColumnInputs &in, const LwWork &w, const LwDevTables &t, int mode, hipStream_t s);

// SW with clouds (rrtmg_sw, icld >= 1, inflgsw = 0, iaer = 0), device FP64
// inputs
struct SwClouds {
    const double *cldfr;                              // [nlay][ncol]
    const double *taucld, *ssacld, *asmcld, *fsfcld;  // [nlay][ncol][14]: the reference's (nbndsw, ncol, nlay)
    const double *cicewp, *cliqwp;                    // [nlay][ncol] or NULL (zero)
    // McICA (spcvmc_sw): per-g-point subcolumns [nlay][ncol][112] in place
    // of the fields above (water paths NULL for zero)
    int mcica;
    const double *cldfmc, *taucmc, *ssacmc, *asmcmc, *fsfcmc, *ciwpmc, *clwpmc;
    // scaled != 0: taucld, ssacld, asmcld hold cldprop's final (delta-scaled)
    // cloud optics, as sw_cloud_wp_kernel forms them from water paths
    int scaled;
    int inflag, iceflag, liqflag;
    const double *reice, *reliq;  // [nlay][ncol]
};
// swcldpr's coefficients packed as extliq1, ssaliq1, asyliq1 [14][58],
// extice2, ssaice2, asyice2 [14][43], extice3, ssaice3, asyice3, fdlice3
// [14][46], abari..fbari [6][5], wavenum2 [14]
constexpr int SW_CLD_TAB = 3 * 812 + 3 * 602 + 4 * 644 + 30 + 14;
// water-path cloud optics (inflgsw 1, 2) into tau, ssa, asm [nlay][ncol][14];
// bad counts the layers where the reference stops (cleared first)
hipError_t launch_sw_cloud_wp(const ColumnInputs &in, const SwClouds &cl, const double *tab, double *tau,
                              double *ssa, double *asm_, int *bad, hipStream_t s);
// cldprmc_sw's water-path optics per subcolumn (tau, ssa, asm_: nlay x ncol x
// 112, g fastest), *bad counting where the reference stops
hipError_t launch_sw_mcica_wp(const ColumnInputs &in, const SwClouds &cl, const double *tab, double *tau,
                              double *ssa, double *asm_, int *bad, hipStream_t s);
// flag bits per (layer, column): a fraction above 0 (the total stream
// blends the layer's cloudy optics), above 1e-12 (reftra of the cloudy
// optics, rrtmg_sw_spcvrt.f90:388), cldprop_sw's cloud test (the layer's
// optics are the input's, delta-M scaled; otherwise depth 0, albedo 1)
constexpr int SWC_BLEND = 1, SWC_REFTRA = 2, SWC_OPTICS = 4;
struct SwCloudWork {
    int *flags = nullptr;     // [nlay][ncol]
    void *bflux = nullptr;    // [14][4][nlay+1][ncol] of Real: total up, total down, clear up, clear down
    void *rows = nullptr;     // chunk rows of sw_cloud_row_bytes: prup, prupd of both streams (FP32: + kept values)
    int chunk = 0;
    int *partial = nullptr;   // device int: layers of day columns with a fraction in (1e-6, 1 - 1e-6)
};
// scratch row of one wave unit of the cloudy SW band kernel: prup, prupd of
// both streams, and in FP32 the five kept values of each cloudy layer
__host__ __device__ inline size_t sw_cloud_row_bytes(int nlay, size_t real_size)
{
    return real_size * (4 * (size_t)(nlay + 1) + (real_size == 4 ? 5 * (size_t)nlay : 0)) * 64;
}
// flags (and the partial-cloud count, zeroed first) of a cloudy SW call
// McICA: o[i] = max(c[i], 1e-10) (the driver's zenith-cosine floor)
hipError_t launch_sw_coszen_floor(int ncol, const double *c, double *o, hipStream_t s);
hipError_t launch_sw_cloud_flags(const ColumnInputs &in, const SwClouds &cl, const SwCloudWork &cw, hipStream_t s);
// the rest of the call (mode 0 FP64, 1 FP32 with FP64 islands): the
// clear-sky prep and column kernels, the cloudy band kernel, the finish
hipError_t launch_sw_cloud(const ColumnInputs &in, const SwClouds &cl, const SwOutputs &out, const SwWork &w,
                           const SwCloudWork &cw, const SwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev);
// the clear-sky forward's prep and column kernels alone
hipError_t launch_sw_prep(const ColumnInputs &in, const SwWork &w, const SwDevTables &t, int mode, hipStream_t s);

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
    void *dcoef = nullptr;    // [nlay][ncol][LWF_STRIDE] of Real: the field values
    void *dder = nullptr;     // [nlay][ncol][LW_DSTRIDE] of Real: their nonzero derivatives (LW_DMASK)
    void *tlev = nullptr;     // [nlay+1][ncol] of Real
    int *ipk = nullptr;       // [nlay][ncol]
    int *ilev = nullptr;      // [nlay+1][ncol]
    int *laytrop = nullptr;   // [ncol]
    double *colscal = nullptr;// [4][ncol]: pwvcm, tbound, column water and dry-air totals
    void *cot = nullptr;      // [2][nlay+1][ncol] of Real: up / down flux cotangents incl. heating
    void *gpart = nullptr;    // [16][3][nlay][ncol] of Real: per-band gradient (T, q, O3)
    void *glev = nullptr;     // [16][nlay+1][ncol] of Real: per-band gradient wrt level temperature
    void *gsc = nullptr;      // [16][2][ncol] of Real: per-band d/dtsfc and d/dpwvcm
    void *dstate = nullptr;   // [rows][2][nlay][64] of Real: incoming downward radiance and upward source per wave
                              // ([rows][3][nlay][64] with the transmittance for the LDS-table kernel)
    int chunk = 0;            // wave units per backward band-kernel launch (rows of dstate)
    void *bflux = nullptr;    // [16][2][nlay+1][ncol] of Real: band fluxes of the recomputed forward, or NULL
    int *items = nullptr;     // work-item counter of the LDS-table band kernel
};
struct SwVjpWork {
    void *dcoef = nullptr;    // [nlay][ncol][SWF_STRIDE] of Real: the field values
    void *dder = nullptr;     // [nlay][ncol][SW_DSTRIDE] of Real: their nonzero derivatives (SW_DMASK)
    int *ipk = nullptr;
    int *laytrop = nullptr;
    int *laysolfr = nullptr;  // [14][ncol]
    void *cot = nullptr;      // [2][nlay+1][ncol] of Real
    void *gpart = nullptr;    // [14][3][nlay][ncol] of Real
    void *dstate = nullptr;   // [chunk][3 (nlay+1) + 5 nlay][64] of Real: downward-sweep states and layer adjoints
    int chunk = 0;            // wave units per backward band-kernel launch (rows of dstate)
    void *bflux = nullptr;    // [14][2][nlay+1][ncol] of Real: band fluxes of the recomputed forward, or NULL
};
// The band kernels keep part of the per-level state that one sweep writes and
// the next reads in the other direction in a global scratch row per wave
// unit, the rest in LDS, so that LDS does not cap occupancy below the VGPRs; a launch covers at most `chunk` wave units so the
// scratch stays within VJP_SCRATCH_BYTES (2 GiB), and longer grids are
// launched in chunks that reuse the rows.
constexpr size_t VJP_SCRATCH_BYTES = (size_t)2 << 30;
inline int vjp_chunk(int units, size_t bytes_per_unit)
{
    size_t c = VJP_SCRATCH_BYTES / bytes_per_unit;
    if (c < 8) c = 8;
    return units < (int)c ? units : (int)c;
}
inline size_t sw_fwd_state_bytes(int nlay, size_t real_size) { return real_size * 2 * (size_t)(nlay + 1) * 64; }
inline size_t lw_fwd_state_bytes(int nlay, size_t real_size) { return real_size * (size_t)nlay * 64; }
inline size_t sw_fwd_lds_bytes(int, size_t real_size) { return real_size * 64; }
inline size_t lw_fwd_lds_bytes(int nlay, size_t real_size) { return real_size * ((size_t)nlay * 64 + 64); }
__host__ __device__ inline size_t sw_vjp_state_bytes(int nlay, size_t real_size)
{
    return real_size * (4 * (size_t)(nlay + 1) + 5 * (size_t)nlay) * 64;
}
inline size_t lw_vjp_state_bytes(int nlay, size_t real_size) { return real_size * 2 * (size_t)nlay * 64; }
// LDS bytes of one backward band workgroup
inline size_t lw_vjp_lds_bytes(int nlay, size_t real_size) { return real_size * ((size_t)nlay * 64 + 4 * 64); }
__host__ __device__ inline size_t sw_vjp_lds_bytes(int nlay, size_t real_size) { return real_size * ((size_t)(nlay + 1) * 64 + 4 * 64); }

// fwd: NULL, or the forward's outputs, formed from w.bflux (which the
// backward band kernel then fills from the forward it recomputes) by the
// forward's finish kernel
hipError_t launch_lw_vjp(const ColumnInputs &in, const VjpCot &cot, const LwGrad &out, const LwVjpWork &w,
                         const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev,
                         const LwOutputs *fwd = nullptr);
hipError_t launch_sw_vjp(const ColumnInputs &in, const VjpCot &cot, const SwGrad &out, const SwVjpWork &w,
                         const SwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev,
                         const SwOutputs *fwd = nullptr);
// the three phases of launch_lw_vjp / launch_sw_vjp: prep (prep and column
// kernels), band (the band kernels), finish (the backward's finish and, with
// fwd given, the forward's)
hipError_t launch_lw_vjp_prep(const ColumnInputs &in, const VjpCot &cot, const LwVjpWork &w, const LwDevTables &t,
                              int mode, hipStream_t s);
hipError_t launch_lw_vjp_band(const ColumnInputs &in, const LwVjpWork &w, const LwDevTables &t, int mode,
                              hipStream_t s);
hipError_t launch_lw_vjp_finish(const ColumnInputs &in, const LwGrad &out, const LwVjpWork &w, const LwDevTables &t,
                                int mode, hipStream_t s, const LwOutputs *fwd);
hipError_t launch_sw_vjp_prep(const ColumnInputs &in, const VjpCot &cot, const SwVjpWork &w, const SwDevTables &t,
                              int mode, hipStream_t s);
hipError_t launch_sw_vjp_band(const ColumnInputs &in, const SwVjpWork &w, const SwDevTables &t, int mode,
                              hipStream_t s);
hipError_t launch_sw_vjp_finish(const ColumnInputs &in, const SwGrad &out, const SwVjpWork &w, const SwDevTables &t,
                                int mode, hipStream_t s, const SwOutputs *fwd);

// The band phase of one LW and one SW backward (FP32, LW k-tables staged in
// LDS) as one persistent kernel: workgroups of FUSED_WGW waves claim work
// items from one counter, first the LW items (as the LW LDS-table kernel's,
// FUSED_WGW tiles per group), then SW items of FUSED_WGW wave units each.
// Each wave has one global scratch row (row_floats floats) for either side.
constexpr int FUSED_WGW = 8;
struct FusedBand {
    ColumnInputs lin;
    LwVjpWork lw;
    const LwBand *lbands;       // LDS-relative k-table offsets
    LwLdsPlan plan;
    int lnwaves, litems;        // LW units per tile; LW items
    int exp_off, tfn_off, plk_off;
    const float *ltab;
    const LwConst *cst;
    const WaveUnit *lunits;
    ColumnInputs sin;
    SwVjpWork sw;
    const SwBand *sbands;
    int snwaves, snunits;       // SW units per tile; SW units in all
    int sexp_off;
    const float *stab;
    const WaveUnit *sunits;
    float *rows;
    size_t row_floats;
    int *items;
    int nitems;
};
// LDS bytes and scratch-row floats of the fused kernel, and its grid
size_t fused_lds_bytes(const LwLdsPlan &plan, int sw_nlay);
size_t fused_row_floats(int lw_nlay, int sw_nlay);
int fused_grid(const LwLdsPlan &plan, int nitems);
hipError_t launch_fused_band(const FusedBand &a, hipStream_t s);

// the forward's finish kernels alone, on band fluxes [band][2][nlay+1][ncol]
hipError_t launch_lw_finish(const ColumnInputs &in, const LwOutputs &out, void *bflux, const LwDevTables &t,
                            int mode, hipStream_t s);
hipError_t launch_sw_finish(const ColumnInputs &in, const SwOutputs &out, void *bflux, const SwDevTables &t,
                            int mode, hipStream_t s);

// one device array of n doubles, for launch_input_checks
struct DeviceArray {
    const double *p;
    size_t n;
};
// Checks on the device that every array is finite and, if plev is given,
// that plev (ncol, nlay+1) decreases strictly upward and is positive at the
// top.  flag (device int) ends as INT_MAX, the index of the first non-finite
// array, or narr for plev.
hipError_t launch_input_checks(const DeviceArray *arrays, int narr, const double *plev, int ncol, int nlay, int *flag,
                               hipStream_t s);

// x[n] (device) -> out[n][6] (device): LW 1 - exp(-x), tfn(x), their
// derivatives, SW exp(-x) floored at 1e-20 and its derivative, as the smooth
// mode and the backward evaluate them
hipError_t launch_smooth_functions(int n, const double *x, double *out, hipStream_t s);

// ---------------------------------------------------------------------------
// MiMA's radiation driver around the RRTM calls (driver.hip)
struct ZenithScalars {
    double radsec, dt_pi, radpersec, dec_sin, dec_cos, dec_tan;  // per call, as compute_zenith forms them
    int dt;
};
struct DriverDims {
    int nlon, nlat, nlev, lonstep;
};
typedef RrtmgHipDriverConfig DriverConfigD;
struct DriverPre {
    DriverConfigD cf;
    const double *lat, *p_full, *p_half, *t, *t_half, *q, *o3f, *albedo, *precip, *t_surf, *coszen;
    double *play, *plev, *tlay, *tlev, *h2o, *o3, *co2, *albedo_rr, *tsfc, *coszen_rr;
    double *zq = nullptr, *zp = nullptr;  // workspace: zonal-mean q, zonal precipitation sums
    int *flag = nullptr;                  // workspace: top interface to be replaced
};
struct DriverPost {
    int do_zm_rad;
    const double *swhr, *hr, *swuflx, *swdflx, *uflx, *dflx;
    double *tdt, *tdt_rad, *flux_sw, *flux_lw, *tdt_sw, *tdt_lw, *olr, *isr;
};
// workspace doubles of launch_rrtmg_pre / launch_rrtmg_post
inline size_t driver_work_doubles(int nlat, int nlev) { return (size_t)nlat * (nlev + 3) + 1; }
hipError_t launch_zenith(int n, const double *lat, const double *lon, const ZenithScalars &z, double *cosz,
                         hipStream_t s);
hipError_t launch_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                              const double *t_surf, const double *t, double *t_half, hipStream_t s);
hipError_t launch_rrtmg_pre(const DriverDims &d, DriverPre a, double *work, hipStream_t s);
hipError_t launch_rrtmg_post(const DriverDims &d, const DriverPost &a, double *work, hipStream_t s);

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
