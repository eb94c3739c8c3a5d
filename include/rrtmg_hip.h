/*
 * rrtmg_hip.h -- C ABI of the MI355X (gfx950) clear-sky RRTMG column radiation.
 *
 * Drop-in boundary for MiMA's radiation driver.  Each entry point replaces one
 * reference Fortran entry point with the same argument meaning, order and
 * units; a thin iso_c_binding shim (mima_amd/fortran/) re-provides the
 * reference module procedures on top of these symbols.  Paths below are
 * relative to src/atmos_param/rrtm_radiation/ of mjucker/MiMA.
 *
 * Conventions (as in the reference drivers):
 *   - (ncol, nlay) arrays are column-major: element (i, k) at [i + ncol*k];
 *     k = 0 is the surface layer; level arrays have nlay+1 entries.
 *   - pressures in hPa, temperatures in K, fluxes in W/m2, heating in K/day;
 *     h2ovmr is specific humidity and o3vmr an O3 mass mixing ratio (MiMA's
 *     usage, converted by the library exactly as inatm does).
 *   - ptr_kind = RRTMG_HIP_HOST: all array arguments are host pointers; the
 *     call is synchronous.  ptr_kind = RRTMG_HIP_DEVICE: all array arguments
 *     are device pointers (FP64) on the current HIP device; work is enqueued
 *     on `stream` (a hipStream_t, NULL = default stream) and the call returns
 *     without synchronising.
 *   - Return 0 on success, a negative RRTMG_HIP_E* code on failure; outputs
 *     are then left untouched and rrtmg_hip_last_error() describes the cause.
 *     Contract: rrtmg_hip_sw and rrtmg_hip_lw are clear sky (MiMA passes
 *     icld = idrv = iaer = 0, RR/rrtm_radiation.f90:106-108); icld != 0,
 *     SW iaer != 0 or LW idrv != 0 return RRTMG_HIP_ENOTSUP there.  Clouds
 *     given by their optical properties (inflglw = inflgsw = 0) take
 *     rrtmg_hip_lw_cloud and rrtmg_hip_sw_cloud; McICA takes
 *     rrtmg_hip_mcica_subcol_lw / rrtmg_hip_lw_mcica and
 *     rrtmg_hip_mcica_subcol_sw / rrtmg_hip_sw_mcica.
 */
#ifndef RRTMG_HIP_H
#define RRTMG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define RRTMG_HIP_HOST 0
#define RRTMG_HIP_DEVICE 1

#define RRTMG_HIP_PREC_FP64 0  /* every kernel in FP64 (the GPU oracle path) */
#define RRTMG_HIP_PREC_MIXED 1 /* FP32 gas optics/transfer with FP64 islands  */
#define RRTMG_HIP_PREC_FP64_SMOOTH 2 /* FP64 with exp(-x) and the tau transition
                                      * function evaluated instead of read from
                                      * the 10000-point tables: smooth in the
                                      * inputs, for finite-difference checks */

#define RRTMG_HIP_OK 0
#define RRTMG_HIP_EINVAL (-22)   /* bad size, null array, NaN/Inf input, non-monotone plev */
#define RRTMG_HIP_ENOTSUP (-95)  /* outside the clear-sky contract */
#define RRTMG_HIP_ENOINIT (-2)   /* called before the matching *_init */
#define RRTMG_HIP_ENOMEM (-12)   /* device allocation failed */
#define RRTMG_HIP_EDEVICE (-5)   /* a HIP runtime call failed */
#define RRTMG_HIP_ENOENT (-3)    /* table file missing or malformed */

#define RRTMG_HIP_MAX_NLAY 200   /* mxlay = 203 in parrrtm.f90 / parrrsw.f90 */

/* Replaces rrtmg_sw_ini(cpdair) (rrtmg_sw/gcm_model/src/rrtmg_sw_init.f90:28):
 * loads the SW tables from table_dir (NULL: the package's tables directory),
 * reduces 224 -> 112 g-points, builds the exp table and uploads everything to
 * the current device.  Idempotent per device. */
int rrtmg_hip_sw_init(double cpdair, const char *table_dir);

/* Replaces rrtmg_lw_ini(cpdair) (rrtmg_lw/gcm_model/src/rrtmg_lw_init.f90:28):
 * 256 -> 140 g-points, tau/exp/tfn tables, Planck table.  Idempotent.
 * Environment RRTMG_HIP_LW_LDS=0, read at the first call on a device, selects
 * the one-wave FP32 band kernel instead of the LDS-table kernel (bitwise the
 * same results). */
int rrtmg_hip_lw_init(double cpdair, const char *table_dir);

/* Selects the arithmetic of subsequent calls (RRTMG_HIP_PREC_*); returns the
 * previous mode.  No reference counterpart (the reference is FP64 only). */
int rrtmg_hip_set_precision(int mode);

/* Replaces rrtmg_sw (rrtmg_sw/gcm_model/src/rrtmg_sw_rad.nomcica.f90:78-88).
 * Same leading arguments; the cloud and aerosol arrays of the reference are
 * not read because icld = iaer = 0 (the reference ignores them then too).
 * Columns with coszen < 1e-10 (negative values included) are night columns
 * with all outputs zero, as in the reference (:502-510). */
int rrtmg_hip_sw(int ncol, int nlay, int icld, int iaer,
                 const double *play, const double *plev, const double *tlay,
                 const double *tlev, const double *tsfc,
                 const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                 const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                 const double *asdir, const double *asdif, const double *aldir,
                 const double *aldif, const double *coszen,
                 double adjes, int dyofyr, double scon,
                 double *swuflx, double *swdflx, double *swhr,
                 double *swuflxc, double *swdflxc, double *swhrc,
                 int ptr_kind, void *stream);

/* Replaces rrtmg_lw (rrtmg_lw/gcm_model/src/rrtmg_lw_rad.nomcica.f90:80-89).
 * emis is (ncol, 16).  tauaer is the reference's (ncol, nlay, 16) aerosol
 * optical depth, added to the gas optical depth as iaer = 10 does; NULL means
 * zero (what MiMA passes).  The cloud arrays are not read (icld = 0). */
int rrtmg_hip_lw(int ncol, int nlay, int icld, int idrv,
                 const double *play, const double *plev, const double *tlay,
                 const double *tlev, const double *tsfc,
                 const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                 const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                 const double *cfc11vmr, const double *cfc12vmr,
                 const double *cfc22vmr, const double *ccl4vmr,
                 const double *emis, const double *tauaer,
                 double *uflx, double *dflx, double *hr,
                 double *uflxc, double *dflxc, double *hrc,
                 int ptr_kind, void *stream);

/* Replaces rrtmg_sw with clouds (rrtmg_sw/gcm_model/src/rrtmg_sw_rad.nomcica.f90:78-88):
 * the reference's whole argument list in its order.  icld >= 1 (outside
 * 0..3 taken as 2, :471; the path without McICA treats every icld >= 1
 * alike); icld = 0 is rrtmg_hip_sw (cloud arrays unread).  cldprop_sw
 * (rrtmg_sw_cldprop.f90:118-338): a layer has cloud optics where cldfr
 * (ncol, nlay) >= 1e-20 and cicewp + cliqwp (NULL = 0) or its summed taucld
 * >= 1e-20.  inflgsw = 0: taucld, ssacld, asmcld, fsfcld (14, ncol, nlay),
 * band fastest, are the cloud optical depth, single-scattering albedo,
 * asymmetry and forward-scattering fraction, delta-M scaled by the library.
 * inflgsw = 2: ice optics by iceflgsw 1-3 and liquid by liqflgsw 1 from
 * swcldpr's coefficients (read from rrtmg_sw_cld.tbl beside rrtmg_sw.tbl)
 * with the radii reice, reliq (ncol, nlay, microns); ssacld, asmcld, fsfcld
 * are not read and taucld may be NULL; optics the reference stops on
 * (a radius out of range, optics out of [0, 1]) return RRTMG_HIP_EINVAL
 * (host pointers always, device pointers with rrtmg_hip_set_checks on).
 * inflgsw = 1: clouds without optics, as the reference.  Day columns (coszen >= 1e-10) take overcast or clear layers
 * only, as the reference (it stops on a fraction in (1e-6, 1 - 1e-6),
 * :535-539): EINVAL, checked for host pointers always and for device
 * pointers when rrtmg_hip_set_checks is on (otherwise the layer's clear and
 * cloudy values are blended by the fraction, the reference's formula).
 * The aerosol arrays are not read.  Outputs: total-sky swuflx, swdflx,
 * swhr and clear-sky swuflxc, swdflxc, swhrc.  inflgsw outside 0..2, or
 * iceflgsw / liqflgsw outside 1..3 / 1 with inflgsw = 2, return
 * RRTMG_HIP_EINVAL; iaer != 0 and RRTMG_HIP_PREC_FP64_SMOOTH return
 * RRTMG_HIP_ENOTSUP. */
int rrtmg_hip_sw_cloud(int ncol, int nlay, int icld, int iaer,
                       const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc,
                       const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                       const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *asdir, const double *asdif, const double *aldir,
                       const double *aldif, const double *coszen,
                       double adjes, int dyofyr, double scon,
                       int inflgsw, int iceflgsw, int liqflgsw,
                       const double *cldfr, const double *taucld, const double *ssacld,
                       const double *asmcld, const double *fsfcld,
                       const double *cicewp, const double *cliqwp,
                       const double *reice, const double *reliq,
                       const double *tauaer, const double *ssaaer, const double *asmaer,
                       const double *ecaer,
                       double *swuflx, double *swdflx, double *swhr,
                       double *swuflxc, double *swdflxc, double *swhrc,
                       int ptr_kind, void *stream);

/* Replaces rrtmg_lw with clouds (rrtmg_lw/gcm_model/src/rrtmg_lw_rad.nomcica.f90:80-89):
 * the reference's whole argument list in its order.  icld = 1: random
 * overlap (rrtmg_lw_rtrn.f90); 2 or 3: maximum-random overlap
 * (rrtmg_lw_rtrnmr.f90; the nomcica driver runs 3 as 2, :527-540); icld
 * outside 0..3 is taken as 2 (:437); icld = 0 is rrtmg_hip_lw (cloud arrays
 * unread).  cldprop (rrtmg_lw_cldprop.f90:118-291): a layer has cloud
 * optics where cldfr (ncol, nlay) >= 1e-20 and cicewp + cliqwp (ncol, nlay;
 * NULL = 0) or its summed taucld >= 1e-20, and is cloudy for the transfer
 * where cldfr >= 1e-6.  inflglw = 0: taucld (16, ncol, nlay), band fastest,
 * holds the in-cloud optical depths; iceflglw, liqflglw, reice, reliq are
 * not read.  inflglw = 1: depths abscld1 times the water path.  inflglw =
 * 2: ice absorption by iceflglw 0-3 and liquid by liqflglw 0-1 from
 * lwcldpr's coefficients (rrtmg_lw_init.f90; the library reads them from
 * rrtmg_lw_cld.tbl beside rrtmg_lw.tbl) with the effective ice radius /
 * size reice and liquid radius reliq (ncol, nlay, microns); taucld may then
 * be NULL (it only enters the cloud test).  A radius outside its
 * parameterisation's range in a layer with cloud optics returns
 * RRTMG_HIP_EINVAL (the reference stops; device pointers: with
 * rrtmg_hip_set_checks on).  Outputs: total-sky uflx, dflx, hr and
 * clear-sky uflxc, dflxc, hrc.  inflglw outside 0..2 returns
 * RRTMG_HIP_EINVAL; idrv != 0 and RRTMG_HIP_PREC_FP64_SMOOTH return
 * RRTMG_HIP_ENOTSUP. */
int rrtmg_hip_lw_cloud(int ncol, int nlay, int icld, int idrv,
                       const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc,
                       const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                       const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *cfc11vmr, const double *cfc12vmr,
                       const double *cfc22vmr, const double *ccl4vmr,
                       const double *emis, int inflglw, int iceflglw, int liqflglw,
                       const double *cldfr, const double *taucld,
                       const double *cicewp, const double *cliqwp,
                       const double *reice, const double *reliq, const double *tauaer,
                       double *uflx, double *dflx, double *hr,
                       double *uflxc, double *dflxc, double *hrc,
                       int ptr_kind, void *stream);

/* Replaces mcica_subcol_lw (rrtmg_lw/gcm_model/src/mcica_subcol_gen_lw.f90:48-154):
 * the McICA stochastic cloud subcolumns, one per LW g-point (140), with the
 * reference's whole argument list in its order.  irng = 0: the KISS
 * generator (:323-358, 530-562), seeds from the fractional Pa of the four
 * lowest layers of play (ncol, nlay), advanced permuteseed times; nlay < 4
 * and a column whose second layer has the higher pressure return
 * RRTMG_HIP_EINVAL (the reference stops, :327-329; checked on the host for
 * host pointers, on the device for device pointers with
 * rrtmg_hip_set_checks on).  irng != 0: the Mersenne Twister
 * (mcica_random_numbers.f90's MT19937, :305, 338-339), one stream per call
 * seeded with permuteseed and drawn in the order subcolumn, column, layer,
 * so a column's subcolumns depend on ncol; the stream is serial (one
 * workgroup), 4 bytes of scratch per draw.  icld = 1 random, 2
 * maximum-random, 3 maximum overlap (:364-445); icld = 0 returns at once
 * with the outputs untouched; icld outside 0..3 returns RRTMG_HIP_EINVAL.
 * cldfrac, ciwp, clwp, rei, rel (ncol, nlay); tauc (16, ncol, nlay), band
 * fastest.  Outputs cldfmcl, ciwpmcl, clwpmcl, taucmcl (140, ncol, nlay),
 * g-point fastest, and reicmcl, relqmcl (ncol, nlay) copies of rei, rel.
 * ciwp, clwp, rei and rel may be NULL (read as zero) and the outputs other
 * than cldfmcl and taucmcl NULL (not written).  iplon is not read. */
int rrtmg_hip_mcica_subcol_lw(int iplon, int ncol, int nlay, int icld, int permuteseed, int irng,
                              const double *play, const double *cldfrac,
                              const double *ciwp, const double *clwp,
                              const double *rei, const double *rel, const double *tauc,
                              double *cldfmcl, double *ciwpmcl, double *clwpmcl,
                              double *reicmcl, double *relqmcl, double *taucmcl,
                              int ptr_kind, void *stream);

/* Replaces rrtmg_lw of the McICA driver (rrtmg_lw/gcm_model/src/rrtmg_lw_rad.f90:80-89):
 * the reference's whole argument list in its order, on the subcolumns of
 * rrtmg_hip_mcica_subcol_lw.  The transfer is rtrnmc (rrtmg_lw_rtrnmc.f90):
 * random overlap per g-point with the subcolumn's cloud fraction, a layer
 * taking the cloudy branch for every g-point where any g-point's cldfmcl is
 * 1.  cldprmc (rrtmg_lw_cldprmc.f90:150-230): inflglw = 0, taucmcl holds
 * the subcolumns' in-cloud optical depths (ciwpmcl, clwpmcl, reicmcl,
 * relqmcl, iceflglw, liqflglw not read); inflglw = 2, a subcolumn passing
 * the cloud test takes its depth from its water paths ciwpmcl, clwpmcl
 * (140, ncol, nlay; NULL = 0) and the layer's radii reicmcl, relqmcl
 * (ncol, nlay) with lwcldpr's coefficients (rrtmg_lw_cld.tbl), by
 * iceflglw 0-3 and liqflglw 0-1, a radius outside its range returning
 * RRTMG_HIP_EINVAL as for rrtmg_hip_lw_cloud; inflglw = 1 returns
 * RRTMG_HIP_EINVAL (the reference stops).  cldfmcl, taucmcl (140, ncol,
 * nlay), g-point fastest.  Unlike the other LW entry points,
 * h2ovmr and o3vmr are volume mixing ratios: this driver's inatm takes them
 * as given (rrtmg_lw_rad.f90:778, 784).  icld outside 0..3 is taken as 2
 * (:451); icld = 0 leaves the subcolumns at zero (:743-744, 874) and runs
 * the clear-sky kernels.  idrv != 0 and RRTMG_HIP_PREC_FP64_SMOOTH return
 * RRTMG_HIP_ENOTSUP. */
int rrtmg_hip_lw_mcica(int ncol, int nlay, int icld, int idrv,
                       const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc,
                       const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                       const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *cfc11vmr, const double *cfc12vmr,
                       const double *cfc22vmr, const double *ccl4vmr,
                       const double *emis, int inflglw, int iceflglw, int liqflglw,
                       const double *cldfmcl, const double *taucmcl,
                       const double *ciwpmcl, const double *clwpmcl,
                       const double *reicmcl, const double *relqmcl, const double *tauaer,
                       double *uflx, double *dflx, double *hr,
                       double *uflxc, double *dflxc, double *hrc,
                       int ptr_kind, void *stream);

/* Replaces mcica_subcol_sw (rrtmg_sw/gcm_model/src/mcica_subcol_gen_sw.f90:47-159):
 * the McICA subcolumns of the SW g-points (112), the reference's whole
 * argument list in its order; the generator, its limits and its error
 * behaviour as rrtmg_hip_mcica_subcol_lw.  tauc, ssac, asmc, fsfc (14,
 * ncol, nlay), band fastest; outputs cldfmcl, ciwpmcl, clwpmcl, taucmcl,
 * ssacmcl, asmcmcl, fsfcmcl (112, ncol, nlay), g-point fastest (a clear
 * subcolumn: depth 0, albedo 1, asymmetry 0, forward fraction 0,
 * :477-510), reicmcl, relqmcl (ncol, nlay).  ssac, asmc, fsfc and their
 * outputs are required. */
int rrtmg_hip_mcica_subcol_sw(int iplon, int ncol, int nlay, int icld, int permuteseed, int irng,
                              const double *play, const double *cldfrac,
                              const double *ciwp, const double *clwp,
                              const double *rei, const double *rel, const double *tauc,
                              const double *ssac, const double *asmc, const double *fsfc,
                              double *cldfmcl, double *ciwpmcl, double *clwpmcl,
                              double *reicmcl, double *relqmcl, double *taucmcl,
                              double *ssacmcl, double *asmcmcl, double *fsfcmcl,
                              int ptr_kind, void *stream);

/* Replaces rrtmg_sw of the McICA driver (rrtmg_sw/gcm_model/src/rrtmg_sw_rad.f90:80-90):
 * the reference's whole argument list in its order, on the subcolumns of
 * rrtmg_hip_mcica_subcol_sw.  cldprmc_sw with inflgsw = 0
 * (rrtmg_sw_cldprmc.f90:127-143): a subcolumn whose fraction and water path
 * or depth reach 1e-20 has its depth, albedo and asymmetry delta-M scaled
 * by its forward fraction, any other keeps them as given; spcvmc_sw: the
 * clear stream and the total one blending each g-point's clear and cloudy
 * layer values by its subcolumn's fraction, which may be partial.  Unlike
 * rrtmg_hip_sw: h2ovmr and o3vmr are volume mixing ratios (this driver's
 * inatm, :1015, 1021), and night columns are computed with the zenith
 * cosine floored at 1e-10 rather than zeroed (:571-572), as the reference
 * does.  cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl (112, ncol, nlay),
 * ciwpmcl, clwpmcl likewise or NULL (zero).  inflgsw = 2
 * (rrtmg_sw_cldprmc.f90:144-330): a subcolumn passing that test takes
 * cldprop_sw's final optics of its band ngb(g) from its own paths and the
 * layer radii reicmcl, relqmcl (ncol, nlay; iceflgsw 1-3, liqflgsw 1; the
 * table rrtmg_sw_cld.tbl beside rrtmg_sw.tbl at init), any other keeps its
 * depth, albedo and asymmetry as given; fsfcmcl is not read.  A radius or
 * optics the reference stops on in a cloudy subcolumn returns
 * RRTMG_HIP_EINVAL (host pointers; device pointers under
 * rrtmg_hip_set_checks).  inflgsw 1 (the reference stops) or outside 0..2
 * returns RRTMG_HIP_EINVAL.  The aerosol arrays are not read.  icld
 * outside 0..3 is taken as 2 (:500); icld = 0 uses no subcolumns
 * (:591-597).  iaer != 0 and RRTMG_HIP_PREC_FP64_SMOOTH return
 * RRTMG_HIP_ENOTSUP. */
int rrtmg_hip_sw_mcica(int ncol, int nlay, int icld, int iaer,
                       const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc,
                       const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                       const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *asdir, const double *asdif, const double *aldir,
                       const double *aldif, const double *coszen,
                       double adjes, int dyofyr, double scon,
                       int inflgsw, int iceflgsw, int liqflgsw,
                       const double *cldfmcl, const double *taucmcl, const double *ssacmcl,
                       const double *asmcmcl, const double *fsfcmcl,
                       const double *ciwpmcl, const double *clwpmcl,
                       const double *reicmcl, const double *relqmcl,
                       const double *tauaer, const double *ssaaer, const double *asmaer,
                       const double *ecaer,
                       double *swuflx, double *swdflx, double *swhr,
                       double *swuflxc, double *swdflxc, double *swhrc,
                       int ptr_kind, void *stream);

/* Backward (vector-Jacobian product) of rrtmg_lw.  No reference counterpart:
 * the reference's only derivative is idrv = 1, dF_up/dTs of the surface
 * Planck term (rrtmg_lw/gcm_model/src/rrtmg_lw_rtrnmr.f90:629-746).  Takes
 * rrtmg_lw's inputs (same layout, icld = idrv = 0) and the cotangents of its
 * outputs uflx, dflx (ncol, nlay+1) and hr (ncol, nlay), each NULL for zero,
 * and overwrites the gradients of  sum(uflx_bar*uflx + dflx_bar*dflx +
 * hr_bar*hr)  with respect to tlay, h2ovmr, o3vmr (ncol, nlay), tlev
 * (ncol, nlay+1) and tsfc (ncol).  The forward is recomputed inside (no
 * state is kept from a previous rrtmg_hip_lw call); the arithmetic follows
 * the precision mode.  The derivatives are those of the smooth functions
 * that the exp and tau-transition tables and the thin-layer series stand
 * for, evaluated along the forward's values; jp/jt/jt1 and the other table
 * indices are held fixed (piecewise differentiation). */
int rrtmg_hip_lw_vjp(int ncol, int nlay,
                     const double *play, const double *plev, const double *tlay,
                     const double *tlev, const double *tsfc,
                     const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                     const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                     const double *cfc11vmr, const double *cfc12vmr,
                     const double *cfc22vmr, const double *ccl4vmr,
                     const double *emis, const double *tauaer,
                     const double *uflx_bar, const double *dflx_bar, const double *hr_bar,
                     double *tlay_bar, double *tlev_bar, double *h2ovmr_bar,
                     double *o3vmr_bar, double *tsfc_bar,
                     int ptr_kind, void *stream);

/* Backward of rrtmg_sw (no reference counterpart), as rrtmg_hip_lw_vjp:
 * cotangents of swuflx, swdflx, swhr (each NULL for zero); overwrites the
 * gradients with respect to tlay, h2ovmr and o3vmr (ncol, nlay).  Night
 * columns (coszen < 1e-10) get zero gradients, as their outputs are zero. */
int rrtmg_hip_sw_vjp(int ncol, int nlay,
                     const double *play, const double *plev, const double *tlay,
                     const double *tlev, const double *tsfc,
                     const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                     const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                     const double *asdir, const double *asdif, const double *aldir,
                     const double *aldif, const double *coszen,
                     double adjes, int dyofyr, double scon,
                     const double *swuflx_bar, const double *swdflx_bar, const double *swhr_bar,
                     double *tlay_bar, double *h2ovmr_bar, double *o3vmr_bar,
                     int ptr_kind, void *stream);

/* Forward and backward in one pass (no reference counterpart): as
 * rrtmg_hip_lw_vjp / rrtmg_hip_sw_vjp, and also the forward's outputs (uflx,
 * dflx (ncol, nlay+1), hr (ncol, nlay); the clear-sky copies of rrtmg_lw /
 * rrtmg_sw are the same values), taken from the forward that the backward
 * recomputes anyway.  For callers that know the cotangents before the
 * forward, e.g. the sensitivity of a fixed functional of the heating rates;
 * equal to rrtmg_hip_lw + rrtmg_hip_lw_vjp (SW likewise) at about the cost of
 * the backward alone.  Device-pointer calls launch kernels only (no
 * allocation once the workspace has grown to the call's size, no memset,
 * no synchronisation), so a warmed-up call can be captured in a HIP graph. */
int rrtmg_hip_lw_fwd_vjp(int ncol, int nlay,
                         const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc,
                         const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                         const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *cfc11vmr, const double *cfc12vmr,
                         const double *cfc22vmr, const double *ccl4vmr,
                         const double *emis, const double *tauaer,
                         const double *uflx_bar, const double *dflx_bar, const double *hr_bar,
                         double *uflx, double *dflx, double *hr,
                         double *tlay_bar, double *tlev_bar, double *h2ovmr_bar,
                         double *o3vmr_bar, double *tsfc_bar,
                         int ptr_kind, void *stream);
int rrtmg_hip_sw_fwd_vjp(int ncol, int nlay,
                         const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc,
                         const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                         const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *asdir, const double *asdif, const double *aldir,
                         const double *aldif, const double *coszen,
                         double adjes, int dyofyr, double scon,
                         const double *swuflx_bar, const double *swdflx_bar, const double *swhr_bar,
                         double *swuflx, double *swdflx, double *swhr,
                         double *tlay_bar, double *h2ovmr_bar, double *o3vmr_bar,
                         int ptr_kind, void *stream);

/* Fused session: the one-pass forward + backward of both sides with their
 * band work in one persistent kernel.  No counterpart in the reference (its
 * forward and backward are separate, and it has no backward).  Between
 * rrtmg_hip_fused_begin(stream) and rrtmg_hip_fused_end(stream), device-pointer
 * calls of rrtmg_hip_lw_vjp / rrtmg_hip_lw_fwd_vjp and rrtmg_hip_sw_vjp /
 * rrtmg_hip_sw_fwd_vjp (at most one per side) on that stream enqueue their
 * prep kernels only; rrtmg_hip_fused_end enqueues the band work of both and
 * then their finish kernels, so their outputs are complete once the stream
 * reaches that point.  With both sides pending in the FP32-mixed precision
 * the band work of both runs as one persistent kernel whose workgroups claim
 * LW and then SW work items from one counter; otherwise (one side, FP64, or
 * LDS too small for the level count) each side's own band kernels run.
 * The LW results are bitwise those of the separate calls; the SW results
 * agree with them to FP32 rounding (the fused kernel calls the SW body as a
 * function where the SW kernel inlines it).  At 131072 x 60 the single
 * kernel is slower than the two sides' kernels on two streams (88.1 against
 * 81.2 ms per step).  Host-pointer calls inside a session return
 * RRTMG_HIP_EINVAL.  Warmed-up sessions allocate nothing, so a session can
 * be captured in a HIP graph. */
int rrtmg_hip_fused_begin(void *stream);
int rrtmg_hip_fused_end(void *stream);
/* 1 if the last rrtmg_hip_fused_end ran the single fused kernel, else 0 */
int rrtmg_hip_fused_last_kernel(void);

/* ---------------------------------------------------------------------------
 * MLP surrogate of the LW+SW heating-rate profile (no reference counterpart;
 * it learns what rrtmg_hip_sw/rrtmg_hip_lw compute and is judged against
 * them).  All pointers are device pointers; work is enqueued on `stream`.
 * BF16 arrays are passed as void* (IEEE bfloat16 bit patterns). */
#define RRTMG_HIP_MLP_MAX_LAYERS 8

/* A trained network: layer l maps width_{l-1} -> width[l] (width_{-1} = ldx,
 * the padded feature count), weights w[l] BF16 row-major [width[l]][width_{l-1}],
 * biases b[l] FP32 [width[l]]; ReLU after every layer but the last.  Features
 * of a column: tlay (nlay), ln h2ovmr (nlay), o3vmr x 1e6 (nlay), tsfc, coszen,
 * normalised as (x - in_mean) * in_istd; outputs are (hr_lw, hr_sw) normalised
 * as (hr - out_mean) / out_std.  in_lo/in_hi bound the training envelope of
 * the raw features (NULL: no envelope test). */
typedef struct {
    int nlay, nin, nout, ldx, nlayers;
    int width[RRTMG_HIP_MLP_MAX_LAYERS];
    const void *w[RRTMG_HIP_MLP_MAX_LAYERS];
    const float *b[RRTMG_HIP_MLP_MAX_LAYERS];
    const float *in_mean, *in_istd, *in_lo, *in_hi;
    const float *out_mean, *out_std;
} RrtmgHipMlp;

/* Surrogate inference: heating rates (ncol, nlay) in K/day of the columns
 * (tlay, h2ovmr, o3vmr as rrtmg_sw/rrtmg_lw take them; tsfc, coszen (ncol))
 * and out_of_envelope[ncol] = 1 where some feature lies outside
 * [in_lo, in_hi] (the caller recomputes those columns with the exact entry
 * points).  Uses a library-owned workspace of ncol x (2 max width BF16 +
 * output FP32). */
int rrtmg_hip_surrogate(int ncol, int nlay, const double *tlay, const double *h2ovmr,
                        const double *o3vmr, const double *tsfc, const double *coszen,
                        const RrtmgHipMlp *net, double *hr_lw, double *hr_sw,
                        int *out_of_envelope, void *stream);

/* Training kernels.  C = op(A) op(B) in BF16 with FP32 accumulation:
 * op(A)(m,k) = A[m*lda + k] (transA = 0) or A[k*lda + m]; op(B)(k,n) =
 * B[n*ldb + k] (transB = 0, the weight layout) or B[k*ldb + n].  Epilogue:
 * v = C + bias[n] (bias may be NULL); out_pre = v; v = max(v, 0) if relu;
 * out_f32 = v (+= with accumulate); out_bf16 = bf16(v); outputs row pitch ldc,
 * each may be NULL.  With both operands transposed and few 64 x 64 output
 * tiles (the weight gradient dY^T X) K is split over a library-owned
 * workspace of up to 256 x 64 x 64 floats per device and stream. */
int rrtmg_hip_gemm_bf16(int M, int N, int K, const void *A, int lda, int transA,
                        const void *B, int ldb, int transB, const float *bias, int relu,
                        float *out_pre, float *out_f32, void *out_bf16, int ldc,
                        int accumulate, void *stream);
/* dz = (pre[i] > 0 ? dy[i] : 0), i < n, into dz_bf16 and dz_f32 (may be NULL) */
int rrtmg_hip_mlp_relu_backward(long n, const float *dy, const float *pre, void *dz_bf16, float *dz_f32,
                                void *stream);
/* out[n] = sum_m x[m*ldx + n] */
int rrtmg_hip_mlp_col_sum(int M, int N, const float *x, int ldx, float *out, void *stream);
/* loss_partial[b] = per-block partial sums of sum (y - t)^2 over M x N (nblocks
 * blocks); dy (FP32) and dy_bf16 (each may be NULL) = 2 (y - t) / (M N) */
int rrtmg_hip_mlp_mse(int M, int N, const float *y, int ldy, const float *t, int ldt, float *dy,
                      void *dy_bf16, int ldd, float *loss_partial, int nblocks, void *stream);
/* one Adam step (bias-corrected, step >= 1) on FP32 master weights w with
 * moments m, v and gradient g; writes the BF16 copy w_bf16 (may be NULL) */
int rrtmg_hip_mlp_adam(long n, float *w, float *m, float *v, const float *g, void *w_bf16,
                       float lr, float beta1, float beta2, float eps, float weight_decay,
                       int step, void *stream);
/* features of columns as rrtmg_hip_surrogate forms them (ldx even, at most
 * 240; the same bound applies to RrtmgHipMlp.ldx, and nlay <= 120 there): x_bf16 [ncol][ldx]
 * (normalised when mean/istd are given), xraw [ncol][ldx] FP32 raw features
 * (may be NULL), out_of_envelope (may be NULL; needs lo/hi) */
int rrtmg_hip_mlp_features(int ncol, int nlay, const double *tlay, const double *h2ovmr,
                           const double *o3vmr, const double *tsfc, const double *coszen,
                           const float *mean, const float *istd, const float *lo, const float *hi,
                           int ldx, void *x_bf16, float *xraw, int *out_of_envelope, void *stream);
/* training targets t [ncol][ldt] (ldt at most 240) FP32 from heating rates (ncol, nlay),
 * normalised when mean/istd are given */
int rrtmg_hip_mlp_targets(int ncol, int nlay, const double *hr_lw, const double *hr_sw,
                          const float *mean, const float *istd, int ldt, float *t, void *stream);

/* ---------------------------------------------------------------------------
 * MiMA's radiation driver around rrtmg_sw/rrtmg_lw, on the device, so that a
 * model whose state lives on the GPU calls the radiation without host copies.
 * GCM arrays are FMS physics-window arrays (nlon, nlat, nlev) in Fortran
 * order (longitude fastest), level 1 at the top of the atmosphere; RRTM
 * arrays are the (ncol, nlay) arrays rrtmg_hip_sw/rrtmg_hip_lw take, ncol =
 * nlon*nlat/lonstep.  All array arguments are device pointers (FP64); work is
 * enqueued on `stream`.  Results are the reference's arithmetic operation by
 * operation (bitwise for the pre- and post-processing; compute_zenith up to
 * the last-place differences of the device's trigonometric functions). */

/* Replaces compute_zenith(Time, equinox_day, dt, lat, lon, cosz, dyofyr)
 * (RR/astro.f90:59-248, RR = src/atmos_param/rrtm_radiation): cosz[n] at lat,
 * lon (radians).  Time as get_time returns it (seconds of the day, days),
 * daysperyear = length_of_year()'s days, obliq in degrees (astro_nml).  dt = 0:
 * instantaneous; 0 < dt < 86400: averaged over [t, t + dt] seconds (the
 * reference's cases assume dt <= 43200, as its comments state); dt >= 86400:
 * daily mean.  *dyofyr (host int) receives the day of the year. */
int rrtmg_hip_compute_zenith(int n, const double *lat, const double *lon, int seconds, int days, int daysperyear,
                             double equinox_day, double obliq, int dt, double *cosz, int *dyofyr, void *stream);

/* Replaces interp_temp(z_full, z_half, t_surf_rad, t) (RR/rrtm_radiation.f90:
 * 422-461): interface temperatures t_half (nlon, nlat, nlev+1) from full-level
 * temperatures t and heights z_full (nlon, nlat, nlev), z_half (nlon, nlat,
 * nlev+1) in metres; top extrapolated, bottom = t_surf (nlon, nlat). */
int rrtmg_hip_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                          const double *t_surf, const double *t, double *t_half, void *stream);

/* run_rrtmg's options (rrtm_radiation_nml, RR/rrtm_radiation.f90:111-173)
 * and the physics window. */
typedef struct {
    int nlon, nlat, nlev, lonstep; /* nlon % lonstep == 0 */
    int do_zm_tracers;             /* q replaced by its zonal mean */
    int do_fixed_water;            /* q = fixed_water where p_full <= fixed_water_pres*100 and
                                    * |lat| <= fixed_water_lat (lat in radians against the namelist's
                                    * number, as the reference compares them) */
    int do_precip_albedo, num_precip; /* albedo + (precip_albedo - albedo) * zonal mean precip / num_precip */
    int do_read_ozone;             /* o3 = max(0, o3f * scale_ozone), else o3_val */
    int do_zm_rad;                 /* zonal-mean heating and surface fluxes */
    double fixed_water, fixed_water_pres, fixed_water_lat, precip_albedo, precip_lat;
    double scale_ozone, o3_val, co2ppmv, h2o_lower_limit, temp_lower_limit, temp_upper_limit;
} RrtmgHipDriverConfig;

/* run_rrtmg from the model state to the RRTM inputs (RR/rrtm_radiation.f90:
 * 581-677): optional zonal-mean water vapour, fixed water, precipitation
 * albedo; every lonstep-th longitude, levels reversed to surface-first,
 * pressures Pa -> hPa, a top interface at or below zero in any column
 * replaced by half the top full-level pressure in all columns, water vapour
 * and temperatures clamped.  Inputs: lat (nlon, nlat) radians, p_full, t, q
 * (and o3f with do_read_ozone) (nlon, nlat, nlev), p_half, t_half
 * (nlon, nlat, nlev+1), albedo, t_surf, coszen (and precip with
 * do_precip_albedo: the accumulated precipitation, which the caller resets)
 * (nlon, nlat).  Outputs: play, tlay, h2ovmr, o3vmr, co2vmr (ncol, nlev),
 * plev, tlev (ncol, nlev+1), albedo_rr, tsfc, coszen_rr (ncol). */
int rrtmg_hip_rrtmg_pre(const RrtmgHipDriverConfig *cfg, const double *lat, const double *p_full,
                        const double *p_half, const double *t, const double *t_half, const double *q,
                        const double *o3f, const double *albedo, const double *precip, const double *t_surf,
                        const double *coszen, double *play, double *plev, double *tlay, double *tlev,
                        double *h2ovmr, double *o3vmr, double *co2vmr, double *albedo_rr, double *tsfc,
                        double *coszen_rr, void *stream);

/* run_rrtmg from the RRTM outputs back to the model (RR/rrtm_radiation.f90:
 * 715-826): heating rates (ncol, nlev) in K/day -> K/s, levels reversed,
 * interpolated linearly in longitude between the RRTM columns (closing
 * toroidally) or zonally averaged (do_zm_rad), added to tdt (nlon, nlat,
 * nlev).  Optional outputs (NULL to skip): tdt_rad, tdt_sw, tdt_lw (the
 * tendency, its SW and LW parts), flux_sw (net down SW at the surface) and
 * flux_lw (down LW at the surface) -- both or neither, as the reference --,
 * olr and isr (nlon, nlat). */
int rrtmg_hip_rrtmg_post(const RrtmgHipDriverConfig *cfg, const double *swhr, const double *hr,
                         const double *swuflx, const double *swdflx, const double *uflx, const double *dflx,
                         double *tdt, double *tdt_rad, double *flux_sw, double *flux_lw, double *tdt_sw,
                         double *tdt_lw, double *olr, double *isr, void *stream);

/* Input checks of device-pointer calls (no reference counterpart; the
 * reference checks nothing).  Host-pointer calls always check that every
 * input is finite and that plev decreases strictly upward with a positive
 * top, and return RRTMG_HIP_EINVAL naming the array otherwise.  With checks on
 * (on != 0), device-pointer calls of rrtmg_hip_sw, rrtmg_hip_lw and the
 * backward entry points make the same checks on the device before any
 * kernel of the call runs; the call then waits for them, so it is no longer
 * asynchronous and cannot be captured in a graph.  Off by default; returns
 * the previous setting. */
int rrtmg_hip_set_checks(int on);

/* Kernel timing (no reference counterpart).  With timing on, each call
 * records HIP events on its stream around its three kernels (prep, band,
 * finish); rrtmg_hip_last_kernel_ms waits for the last call of the given kind
 * on this device (0 rrtmg_hip_sw, 1 rrtmg_hip_lw, 2 rrtmg_hip_sw_vjp,
 * 3 rrtmg_hip_lw_vjp) and writes the three durations in milliseconds. */
int rrtmg_hip_set_timing(int on);
int rrtmg_hip_last_kernel_ms(int kind, float *ms);

/* Workspaces (all entry points): the library keeps its device workspaces per
 * device and per stream, allocated on first use and grown on demand; calls
 * on one stream reuse them in stream order, calls on different streams use
 * different workspaces and may overlap.  Growing a workspace synchronises
 * the device. */

/* rrtmg_hip_sw without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as rrtmg_hip_sw with ptr_kind = RRTMG_HIP_DEVICE (bitwise the same
 * outputs: the same kernels), work enqueued on `stream`.  The workspace is a
 * fixed device array of RRTMG_HIP_NOALLOC_BYTES per GPU, held from
 * rrtmg_hip_sw_init until the process ends; the call allocates nothing.  Any
 * ncol x (nlay + 1) <= RRTMG_HIP_NOALLOC_MAX_CELLS fits in every precision
 * mode; a call that does not fit returns RRTMG_HIP_EINVAL naming the size.
 * Every call on one GPU shares that GPU's array, whatever its stream: calls
 * that overlap on one GPU (two streams, two host threads) race, so order them.
 * With rrtmg_hip_set_checks on, the input checks run as for rrtmg_hip_sw
 * (their flag is in the same array; the call then waits for them). */
#define RRTMG_HIP_NOALLOC_BYTES (128u << 20)
#define RRTMG_HIP_NOALLOC_MAX_CELLS 262144
int rrtmg_hip_sw_noalloc(int ncol, int nlay, int icld, int iaer,
                         const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc,
                         const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                         const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *asdir, const double *asdif, const double *aldir,
                         const double *aldif, const double *coszen,
                         double adjes, int dyofyr, double scon,
                         double *swuflx, double *swdflx, double *swhr,
                         double *swuflxc, double *swdflxc, double *swhrc,
                         void *stream);

/* Message for the last failing call on this host thread ("" if none). */
const char *rrtmg_hip_last_error(void);

/* Debug/parity accessors (no reference counterpart).
 * Per-layer table indices of the last call of the given side on this device,
 * copied to host int arrays of (ncol, nlay): jp, jt, jt1 (1-based, as the
 * reference's setcoef), and laytrop per column.  0 on success. */
int rrtmg_hip_debug_indices(int lw, int *jp, int *jt, int *jt1, int *laytrop);
/* The smooth functions that the lookup tables stand for, at n host
 * arguments x (optical depths), as the smooth FP64 mode and the backward
 * evaluate them on the device: out[6 i + 0..5] = LW 1 - exp(-x), LW tau
 * transition function tfn(x) (rrtmg_lw_init.f90:106-123), their derivatives,
 * SW exp(-x) floored at 1e-20 (rrtmg_sw_init.f90:94-104) and its derivative.
 * Host arrays; synchronous. */
int rrtmg_hip_debug_smooth(int n, const double *x, double *out);
/* A reduced table by name ("sw/b16/ka", "lw/b03/fracrefa", ...), Fortran
 * order, FP64, as built at init; returns the element count or -1. */
long rrtmg_hip_debug_table(const char *name, double *out, long n);
/* Builds (host memory only, no device needed) the reduced tables that
 * rrtmg_hip_debug_table then returns; used to check the reduction on CPU. */
int rrtmg_hip_debug_host_tables(int lw, const char *table_dir, double cpdair);

#ifdef __cplusplus
}
#endif
#endif
