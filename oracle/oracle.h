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
/* A reduced table by name ("sw/b16/ka", "lw/b03/fracrefa", ...): copies up to
 * n values (Fortran order) into out and returns the element count, or -1. */
long oracle_reduced(const char *name, double *out, long n);
/* exp_tbl / tfn_tbl entries (0..10000) of the SW or LW init. */
double oracle_sw_exp_tbl(int i);
double oracle_lw_tbl(int which /*0 tau, 1 exp, 2 tfn*/, int i);
double oracle_heatfac(int lw);
/* Threads used by the column loops (OpenMP); returns the previous value. */
int oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
