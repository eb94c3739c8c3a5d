/*
 * CPU FP64 oracle for MiMA's clear-sky RRTMG column radiation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the tests, the smoke
 * check and the CPU-baseline leg of the benchmark compare the HIP library
 * against.  It is never linked into, loaded by, or called from the product
 * path (mima_amd/), which fails loudly when its HIP library is missing.
 *
 * It restates, in plain C and FP64, the arithmetic of the reference Fortran
 * (AER RRTMG_LW v4.85 / RRTMG_SW "gcm_model" non-McICA drivers as built by
 * MiMA with -r8).  Every function cites the reference file:line it follows;
 * paths are relative to src/atmos_param/rrtm_radiation/ of mjucker/MiMA.
 *
 * Array conventions follow the reference drivers: (ncol, nlay) arrays are
 * Fortran column-major, element (i, k) at [i + ncol*k], index k = 0 is the
 * surface layer; level arrays have nlay+1 entries; pressures in hPa.
 */
#ifndef RRTMG_ORACLE_H
#define RRTMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Load an RRTMGTB1 table file, reduce the g-points and build lookup tables
 * (rrtmg_sw_init.f90:28-154 / rrtmg_lw_init.f90:28-175).  0 on success. */
int oracle_sw_init(const char *table_path, double cpdair);
int oracle_lw_init(const char *table_path, double cpdair);

/* rrtmg_sw (rrtmg_sw_rad.nomcica.f90:78-731), clear sky, iaer = 0.
 * Outputs are (ncol, nlay+1) fluxes in W/m2 and (ncol, nlay) heating in K/day. */
int oracle_rrtmg_sw(int ncol, int nlay,
                    const double *play, const double *plev, const double *tlay,
                    const double *tlev, const double *tsfc,
                    const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                    const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                    const double *asdir, const double *asdif, const double *aldir,
                    const double *aldif, const double *coszen,
                    double adjes, int dyofyr, double scon,
                    double *swuflx, double *swdflx, double *swhr,
                    double *swuflxc, double *swdflxc, double *swhrc);

/* rrtmg_lw (rrtmg_lw_rad.nomcica.f90:80-569), clear sky, idrv = 0.
 * emis is (ncol, 16). */
int oracle_rrtmg_lw(int ncol, int nlay,
                    const double *play, const double *plev, const double *tlay,
                    const double *tlev, const double *tsfc,
                    const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                    const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                    const double *cfc11vmr, const double *cfc12vmr,
                    const double *cfc22vmr, const double *ccl4vmr,
                    const double *emis,
                    double *uflx, double *dflx, double *hr,
                    double *uflxc, double *dflxc, double *hrc);

/* Per-column intermediates for index/optics parity tests (single column,
 * arrays for that column only, surface-first).  jp/jt/jt1 are 1-based as in
 * the reference.  Returns laytrop. */
int oracle_sw_setcoef(int nlay, const double *play, const double *plev,
                      const double *tlay, const double *h2ovmr, const double *o3vmr,
                      const double *co2vmr, const double *ch4vmr,
                      const double *n2ovmr, const double *o2vmr,
                      int *jp, int *jt, int *jt1,
                      double *fac00, double *fac01, double *fac10, double *fac11,
                      double *colmol);
int oracle_lw_setcoef(int nlay, const double *play, const double *plev,
                      const double *tlay, const double *tlev, double tsfc,
                      const double *h2ovmr, const double *o3vmr,
                      const double *co2vmr, const double *ch4vmr,
                      const double *n2ovmr, const double *o2vmr,
                      int *jp, int *jt, int *jt1,
                      double *fac00, double *fac01, double *fac10, double *fac11,
                      double *planklay /* (nlay,16) */, double *planklev /* (nlay+1,16) */,
                      double *plankbnd /* 16 */);
/* taumol for one column: taug/taur (nlay, 112), sfluxzen (112). */
int oracle_sw_taumol(int nlay, const double *play, const double *plev,
                     const double *tlay, const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr,
                     const double *n2ovmr, const double *o2vmr,
                     double *taug, double *taur, double *sfluxzen);
/* taumol for one column: taug/fracs (nlay, 140). */
int oracle_lw_taumol(int nlay, const double *play, const double *plev,
                     const double *tlay, const double *tlev, double tsfc,
                     const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr,
                     const double *n2ovmr, const double *o2vmr,
                     const double *cfc11vmr, const double *cfc12vmr,
                     const double *cfc22vmr, const double *ccl4vmr,
                     double *taug, double *fracs);
/* inatm of one column: coldry (nlay), molecule columns wkl (7, nlay) in
 * molecules/cm2 and precipitable water pwvcm (cm). */
int oracle_lw_inatm(int nlay, const double *play, const double *plev,
                    const double *tlay, const double *tlev, double tsfc,
                    const double *h2ovmr, const double *o3vmr,
                    const double *co2vmr, const double *ch4vmr,
                    const double *n2ovmr, const double *o2vmr,
                    double *coldry, double *wkl, double *pwvcm);
/* reftra_sw for nlay layers (rrtmg_sw_reftra.f90:29-305), all layers active. */
void oracle_sw_reftra(int nlay, const double *gg, double prmuz, const double *tau,
                      const double *w, double *pref, double *prefd,
                      double *ptra, double *ptrad);
/* rrtmg_sw with clouds (rrtmg_sw_rad.nomcica.f90:78-731): cldprop_sw with
 * inflgsw = 0 (delta-M scaled input optics), spcvrt_sw's clear and total
 * streams; iaer = 0.  cldfr, cicewp, cliqwp (ncol, nlay) (NULL water paths =
 * 0); taucld, ssacld, asmcld, fsfcld (14, ncol, nlay), band fastest.  0 on
 * success, -5 on partial cloud in a day column (the reference stops). */
int oracle_rrtmg_sw_cloud(int ncol, int nlay, int icld, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *asdir, const double *asdif,
                          const double *aldir, const double *aldif, const double *coszen, double adjes,
                          int dyofyr, double scon, const double *cldfr, const double *taucld,
                          const double *ssacld, const double *asmcld, const double *fsfcld,
                          const double *cicewp, const double *cliqwp, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc);

/* rrtmg_lw with clouds (rrtmg_lw_rad.nomcica.f90:80-569): cldprop with
 * inflglw = 0, then rtrn (icld = 1) or rtrnmr (icld = 2, 3); idrv = 0.
 * cldfr, cicewp, cliqwp (ncol, nlay) (NULL water paths = 0); taucld
 * (16, ncol, nlay) band fastest; tauaer (ncol, nlay, 16) or NULL.  Total-sky
 * and clear-sky outputs as rrtmg_lw.  0 on success. */
int oracle_rrtmg_lw_cloud(int ncol, int nlay, int icld, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                          const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                          const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                          const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                          const double *emis, const double *cldfr, const double *taucld, const double *cicewp,
                          const double *cliqwp, const double *tauaer, double *uflx, double *dflx, double *hr,
                          double *uflxc, double *dflxc, double *hrc);

/* rrtmg_lw with clouds given by water paths (inflglw = 1: abscld1 times
 * the path; 2: ice and liquid absorption coefficients from lwcldpr's
 * tables by iceflglw 0-3 and liqflglw 0-1, rrtmg_lw_cldprop.f90:118-291),
 * the tables loaded by oracle_lw_cld_init from rrtmg_lw_cld.tbl; reice,
 * reliq (ncol, nlay) effective radii / sizes in microns.  taucld (16, ncol,
 * nlay) enters only the cloud test.  -8 for a radius the reference stops
 * on. */
int oracle_lw_cld_init(const char *path);
int oracle_rrtmg_lw_cloud_wp(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play,
                             const double *plev, const double *tlay, const double *tlev, const double *tsfc,
                             const double *h2ovmr, const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                             const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                             const double *emis, const double *cldfr, const double *taucld, const double *cicewp,
                             const double *cliqwp, const double *reice, const double *reliq, const double *tauaer,
                             double *uflx, double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc);

/* McICA LW: the stochastic subcolumns of mcica_subcol_lw with the KISS
 * generator (irng = 0; mcica_subcol_gen_lw.f90:48-520) and rrtmg_lw of the
 * McICA driver on them (rrtmg_lw_rad.f90:80-572: cldprmc with inflglw = 0,
 * rtrnmc).  Cloud inputs (ncol, nlay), tauc (16, ncol, nlay); subcolumn
 * arrays (140, ncol, nlay).  0 on success. */
int oracle_lw_mcica_subcol(int ncol, int nlay, int icld, int permuteseed, const double *play,
                           const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                           double *cldfmcl, double *ciwpmcl, double *clwpmcl, double *taucmcl);
int oracle_rrtmg_lw_mcica(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                          const double *cfc22vmr, const double *ccl4vmr, const double *emis,
                          const double *cldfmcl, const double *taucmcl, const double *tauaer, double *uflx,
                          double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc);

/* The KISS generator's cloudy-subcolumn mask of one column (both sides'
 * generators, mcica_subcol_gen_lw.f90:158-520 / mcica_subcol_gen_sw.f90:
 * 163-536, nsub 140 or 112): mask[ig * 256 + l].  -6 on rising pressure. */
int oracle_mcica_mask(int i, int ncol, int nlay, int icld, int permuteseed, int nsub, const double *play,
                      const double *cldfrac, unsigned char *mask);
/* MT19937's first `count` tempered words for init_genrand(seed)
 * (mcica_random_numbers.f90:134-185, 243-258). */
void oracle_mt_stream(uint32_t seed, long count, uint32_t *out);
/* the Mersenne Twister stream oracle_mcica_mask draws from (NULL: KISS) */
void oracle_mcica_set_stream(const uint32_t *mt);
/* the subcolumn generators with the Mersenne Twister (irng = 1), seeded
 * with permuteseed; arguments as oracle_lw_mcica_subcol /
 * oracle_sw_mcica_subcol.  Not thread-safe (one stream per call). */
int oracle_lw_mcica_subcol_mt(int ncol, int nlay, int icld, int permuteseed, const double *play,
                              const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                              double *cldfmcl, double *ciwpmcl, double *clwpmcl, double *taucmcl);
int oracle_sw_mcica_subcol_mt(int ncol, int nlay, int icld, int permuteseed, const double *play,
                              const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                              const double *ssac, const double *asmc, const double *fsfc, double *cldfmcl,
                              double *ciwpmcl, double *clwpmcl, double *taucmcl, double *ssacmcl, double *asmcmcl,
                              double *fsfcmcl);
/* McICA SW: mcica_subcol_sw with the KISS generator and rrtmg_sw of the
 * McICA driver (rrtmg_sw_rad.f90:80-745: volume mixing ratios, zenith
 * cosine floored at 1e-10, cldprmc_sw with inflgsw = 0, spcvmc_sw).
 * Cloud inputs (ncol, nlay), optics (14, ncol, nlay); subcolumns (112,
 * ncol, nlay).  0 on success. */
int oracle_sw_mcica_subcol(int ncol, int nlay, int icld, int permuteseed, const double *play,
                           const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                           const double *ssac, const double *asmc, const double *fsfc, double *cldfmcl,
                           double *ciwpmcl, double *clwpmcl, double *taucmcl, double *ssacmcl, double *asmcmcl,
                           double *fsfcmcl);
int oracle_rrtmg_sw_mcica(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                          const double *ch4vmr, const double *n2ovmr, const double *o2vmr, const double *asdir,
                          const double *asdif, const double *aldir, const double *aldif, const double *coszen,
                          double adjes, int dyofyr, double scon, const double *cldfmcl, const double *taucmcl,
                          const double *ssacmcl, const double *asmcmcl, const double *fsfcmcl,
                          const double *ciwpmcl, const double *clwpmcl, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc);

/* rrtmg_sw with clouds given by water paths (inflgsw = 2, cldprop_sw's
 * ice parameterisations 1-3 and liquid 1 on swcldpr's coefficients,
 * rrtmg_sw_cldprop.f90:143-338, loaded by oracle_sw_cld_init from
 * rrtmg_sw_cld.tbl; inflgsw = 1 gives the clouds no optics, as the
 * reference).  reice, reliq (ncol, nlay); taucld (14, ncol, nlay) only in
 * the cloud test (NULL = 0).  -8 where the reference stops. */
int oracle_sw_cld_init(const char *path);
int oracle_rrtmg_sw_cloud_wp(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play,
                             const double *plev, const double *tlay, const double *tlev, const double *h2ovmr,
                             const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *asdir, const double *asdif,
                             const double *aldir, const double *aldif, const double *coszen, double adjes,
                             int dyofyr, double scon, const double *cldfr, const double *taucld,
                             const double *cicewp, const double *cliqwp, const double *reice, const double *reliq,
                             double *swuflx, double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                             double *swhrc);

/* oracle_rrtmg_sw_mcica with cldprmc_sw's water-path optics (inflgsw = 2,
 * rrtmg_sw_cldprmc.f90:144-330: cldprop_sw's per-band optics for each
 * subcolumn passing the cloud test, from its paths and the layer's radii
 * reicmcl, relqmcl (ncol, nlay); any other keeps its input optics), after
 * oracle_sw_cld_init.  -8 where the reference stops. */
int oracle_rrtmg_sw_mcica_wp(int ncol, int nlay, int iceflag, int liqflag, const double *play, const double *plev,
                             const double *tlay, const double *tlev, const double *h2ovmr, const double *o3vmr,
                             const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                             const double *o2vmr, const double *asdir, const double *asdif, const double *aldir,
                             const double *aldif, const double *coszen, double adjes, int dyofyr, double scon,
                             const double *cldfmcl, const double *taucmcl, const double *ssacmcl,
                             const double *asmcmcl, const double *fsfcmcl, const double *ciwpmcl,
                             const double *clwpmcl, const double *reicmcl, const double *relqmcl, double *swuflx,
                             double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc);

/* oracle_rrtmg_lw_mcica with cldprmc's water-path optics (inflglw = 2,
 * rrtmg_lw_cldprmc.f90:150-230): per subcolumn that passes the cloud test,
 * its depth from ciwpmcl, clwpmcl (140, ncol, nlay) and the layer's radii
 * reicmcl, relqmcl (ncol, nlay) with its g-point's band's coefficients
 * (after oracle_lw_cld_init).  -8 on a radius the reference stops on. */
int oracle_rrtmg_lw_mcica_wp(int ncol, int nlay, int iceflag, int liqflag, const double *play, const double *plev,
                             const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                             const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                             const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                             const double *emis, const double *cldfmcl, const double *taucmcl,
                             const double *ciwpmcl, const double *clwpmcl, const double *reicmcl,
                             const double *relqmcl, const double *tauaer, double *uflx, double *dflx, double *hr,
                             double *uflxc, double *dflxc, double *hrc);

/* A reduced table by name ("sw/b16/ka", "lw/b03/fracrefa", ...): copies up to
 * n values (Fortran order) into out and returns the element count, or -1. */
long oracle_reduced(const char *name, double *out, long n);
/* exp_tbl / tfn_tbl entries (0..10000) of the SW or LW init. */
double oracle_sw_exp_tbl(int i);
double oracle_lw_tbl(int which /*0 tau, 1 exp, 2 tfn*/, int i);
double oracle_heatfac(int lw);
/* Threads used by the column loops (OpenMP); returns the previous value. */
int oracle_set_threads(int n);

/* ---- MiMA's radiation driver around rrtmg_sw/rrtmg_lw (oracle_driver.c).
 * GCM arrays: (nlon, nlat, nlev) Fortran order, longitude fastest, level 1
 * at the top; RRTMG arrays (ncol, nlay), ncol = nlon*nlat/lonstep. */
typedef struct {
    int nlon, nlat, nlev, lonstep;
    int do_zm_tracers, do_fixed_water, do_precip_albedo, num_precip, do_read_ozone, do_zm_rad;
    double fixed_water, fixed_water_pres, fixed_water_lat, precip_albedo, precip_lat;
    double scale_ozone, o3_val, co2ppmv, h2o_lower_limit, temp_lower_limit, temp_upper_limit;
} OracleDriverConfig;

/* compute_zenith (RR/astro.f90:59-248) */
int oracle_compute_zenith(int n, const double *lat, const double *lon, int seconds, int days, int daysperyear,
                          double equinox_day, double obliq, int dt, double *cosz, int *dyofyr);
/* interp_temp (RR/rrtm_radiation.f90:422-461) */
int oracle_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                       const double *t_surf, const double *t, double *t_half);
/* run_rrtmg before the RRTM calls (RR/rrtm_radiation.f90:581-677) */
int oracle_rrtmg_pre(const OracleDriverConfig *cf, const double *lat, const double *p_full, const double *p_half,
                     const double *t, const double *t_half, const double *q, const double *o3f,
                     const double *albedo, const double *precip, const double *t_surf, const double *coszen,
                     double *play, double *plev, double *tlay, double *tlev, double *h2ovmr, double *o3vmr,
                     double *co2vmr, double *albedo_rr, double *tsfc, double *coszen_rr);
/* run_rrtmg after the RRTM calls (RR/rrtm_radiation.f90:715-826) */
int oracle_rrtmg_post(const OracleDriverConfig *cf, const double *swhr, const double *hr, const double *swuflx,
                      const double *swdflx, const double *uflx, const double *dflx, double *tdt, double *tdt_rad,
                      double *flux_sw, double *flux_lw, double *tdt_sw, double *tdt_lw, double *olr, double *isr);

#ifdef __cplusplus
}
#endif
#endif
