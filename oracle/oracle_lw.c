/*
 * Oracle: RRTMG_LW clear-sky column model in FP64 (TEST INFRASTRUCTURE ONLY;
 * see oracle.h).  Restates, in order of evaluation:
 *   rrtmg_lw_init.f90:28-175 (tables, rwgt), :178-363, :366-2015 (cmbgb1..16)
 *   rrtmg_lw_rad.nomcica.f90:419-567 (column loop), :572-901 (inatm)
 *   rrtmg_lw_setcoef.f90:31-415 (setcoef)
 *   rrtmg_lw_taumol.f90:31-3149 (taugb1..16)
 *   rrtmg_lw_rtrnmr.f90:32-779 (rtrnmr, clear-sky branch: icldlyr = 0)
 * LW k-tables are seeded synthetic values (AER's rrtmg_lw_k_g.f90 is absent);
 * the arithmetic is the reference's regardless of the table values.  Pinned
 * against the reference LW column code itself (inatm/setcoef/taumol/rtrnmr
 * compiled from its sources, oracle/ref_lw_driver.f90) on those tables.
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NBLW 16
#define NGLW 140
#define MAXLAY 256

typedef struct {
    double *p;
    int ext[4];
    int nd;
} rarr_t; /* reduced array, Fortran order, g-point dimension last (fracref: first) */

typedef struct {
    int ng, nspa, nspb;
    rarr_t absa, absb, selfref, forref, fracrefa, fracrefb;
    rarr_t ka_mn2, kb_mn2, ka_mn2o, kb_mn2o, ka_mo3, kb_mo3, ka_mco2, kb_mco2, ka_mco, ka_mo2, kb_mo2;
    rarr_t ccl4, cfc11adj, cfc12, cfc22adj;
} lwband_t;

static struct {
    int ready;
    double heatfac, grav, avogad;
    double tau_tbl[ORACLE_NTBL + 1], exp_tbl[ORACLE_NTBL + 1], tfn_tbl[ORACLE_NTBL + 1];
    double bpade;
    int ngc[NBLW], ngs[NBLW];
    double delwave[NBLW];
    double rwgt[NBLW * 16];
    double pref[59], preflog[59], tref[59];
    double chi_mls[7 * 59];
    double totplnk[181 * 16];
    double a0[16], a1[16], a2[16];
    lwband_t b[NBLW];
} L;

#define CHI(i, j) L.chi_mls[((j)-1) * 7 + ((i)-1)]
#define TOTPLNK(i, ib) L.totplnk[((ib)-1) * 181 + ((i)-1)]

/* 2-D and 3-D accessors with the g-point index last */
#define A2(r, i, ig) ((r).p[(long)((ig)-1) * (r).ext[0] + ((i)-1)])
#define A3(r, i, j, ig) ((r).p[((long)((ig)-1) * (r).ext[1] + ((j)-1)) * (r).ext[0] + ((i)-1)])
#define G1(r, ig) ((r).p[(ig)-1])
/* fracref(ng, neta): g-point first */
#define FR2(r, ig, j) ((r).p[((j)-1) * (r).ext[0] + ((ig)-1)])

static const double oneminus_lw = 1.0 - 1.e-6; /* rrtmg_lw_rad.nomcica.f90:419 */

static int reduce_lw(const otbl_t *t, int band, const char *oname, const char *rname, int gfirst,
                     int weighted, const int *ngn, rarr_t *out)
{
    char key[64];
    snprintf(key, sizeof key, "lw/b%02d/%s", band, oname);
    int nd, ext[4], lb[4], isint;
    const double *x = (const double *)otbl_get(t, key, &nd, ext, lb, &isint);
    memset(out, 0, sizeof *out);
    if (!x) return 0;
    const int *first_ngn = ngn + (band == 1 ? 0 : L.ngs[band - 2]);
    int ngc = L.ngc[band - 1];
    out->p = oracle_reduce(x, nd, ext, gfirst ? 0 : 1, ngc, first_ngn, L.rwgt + (band - 1) * 16, weighted);
    out->nd = nd;
    for (int d = 0; d < 4; ++d) out->ext[d] = ext[d];
    out->ext[gfirst ? 0 : nd - 1] = ngc;
    long n = 1;
    for (int d = 0; d < nd; ++d) n *= out->ext[d];
    snprintf(key, sizeof key, "lw/b%02d/%s", band, rname);
    oracle_register(key, out->p, n);
    return 1;
}

int oracle_lw_init(const char *path, double cpdair)
{
    otbl_t t;
    if (otbl_load(&t, path) != 0) return -1;
    /* lwdatinit (rrtmg_lw_init.f90:178-281) */
    L.grav = 9.8066;
    L.avogad = 6.02214199e+23;
    L.heatfac = L.grav * 8.6400e4 / (cpdair * 1.e2);
    const int *ngc = otbl_i(&t, "lw/ngc", NULL), *ngs = otbl_i(&t, "lw/ngs", NULL);
    const int *ngn = otbl_i(&t, "lw/ngn", NULL), *ngm = otbl_i(&t, "lw/ngm", NULL);
    const int *nspa = otbl_i(&t, "lw/nspa", NULL), *nspb = otbl_i(&t, "lw/nspb", NULL);
    const double *wt = otbl_d(&t, "lw/wt", NULL);
    if (!ngc || !ngs || !ngn || !ngm || !nspa || !nspb || !wt) return -2;
    memcpy(L.ngc, ngc, sizeof L.ngc);
    memcpy(L.ngs, ngs, sizeof L.ngs);
    memcpy(L.delwave, otbl_d(&t, "lw/delwave", NULL), sizeof L.delwave);
    memcpy(L.pref, otbl_d(&t, "lw/pref", NULL), sizeof L.pref);
    memcpy(L.preflog, otbl_d(&t, "lw/preflog", NULL), sizeof L.preflog);
    memcpy(L.tref, otbl_d(&t, "lw/tref", NULL), sizeof L.tref);
    memcpy(L.chi_mls, otbl_d(&t, "lw/chi_mls", NULL), sizeof L.chi_mls);
    memcpy(L.totplnk, otbl_d(&t, "lw/totplnk", NULL), sizeof L.totplnk);
    memcpy(L.a0, otbl_d(&t, "lw/a0", NULL), sizeof L.a0);
    memcpy(L.a1, otbl_d(&t, "lw/a1", NULL), sizeof L.a1);
    memcpy(L.a2, otbl_d(&t, "lw/a2", NULL), sizeof L.a2);

    /* lookup tables (rrtmg_lw_init.f90:106-123) */
    const double expeps = 1.e-20;
    L.tau_tbl[0] = 0.0;
    L.tau_tbl[ORACLE_NTBL] = 1.e10;
    L.exp_tbl[0] = 1.0;
    L.exp_tbl[ORACLE_NTBL] = expeps;
    L.tfn_tbl[0] = 0.0;
    L.tfn_tbl[ORACLE_NTBL] = 1.0;
    L.bpade = 1.0 / 0.278;
    for (int itr = 1; itr <= ORACLE_NTBL - 1; ++itr) {
        double tfn = (double)itr / (double)ORACLE_NTBL;
        L.tau_tbl[itr] = L.bpade * tfn / (1.0 - tfn);
        L.exp_tbl[itr] = exp(-L.tau_tbl[itr]);
        if (L.exp_tbl[itr] <= expeps) L.exp_tbl[itr] = expeps;
        if (L.tau_tbl[itr] < 0.06)
            L.tfn_tbl[itr] = L.tau_tbl[itr] / 6.0;
        else
            L.tfn_tbl[itr] = 1.0 - 2.0 * ((1.0 / L.tau_tbl[itr]) - (L.exp_tbl[itr] / (1. - L.exp_tbl[itr])));
    }
    oracle_make_rwgt(NBLW, ngc, ngn, ngm, wt, L.rwgt);
    oracle_register("lw/rwgt", L.rwgt, NBLW * 16);

    static const char *kn[][2] = {{"kao_mn2", "ka_mn2"},   {"kbo_mn2", "kb_mn2"},   {"kao_mn2o", "ka_mn2o"},
                                  {"kbo_mn2o", "kb_mn2o"}, {"kao_mo3", "ka_mo3"},   {"kbo_mo3", "kb_mo3"},
                                  {"kao_mco2", "ka_mco2"}, {"kbo_mco2", "kb_mco2"}, {"kao_mco", "ka_mco"},
                                  {"kao_mo2", "ka_mo2"},   {"kbo_mo2", "kb_mo2"},   {"ccl4o", "ccl4"},
                                  {"cfc11adjo", "cfc11adj"}, {"cfc12o", "cfc12"},   {"cfc22adjo", "cfc22adj"}};
    for (int band = 1; band <= NBLW; ++band) {
        lwband_t *bb = &L.b[band - 1];
        memset(bb, 0, sizeof *bb);
        bb->ng = ngc[band - 1];
        bb->nspa = nspa[band - 1];
        bb->nspb = nspb[band - 1];
        reduce_lw(&t, band, "kao", "ka", 0, 1, ngn, &bb->absa);
        /* absa(65*nspa, ng): flatten the leading dims */
        if (bb->absa.p) {
            int rows = 1;
            for (int d = 0; d < bb->absa.nd - 1; ++d) rows *= bb->absa.ext[d];
            bb->absa.ext[0] = rows;
        }
        reduce_lw(&t, band, "kbo", "kb", 0, 1, ngn, &bb->absb);
        if (bb->absb.p) {
            int rows = 1;
            for (int d = 0; d < bb->absb.nd - 1; ++d) rows *= bb->absb.ext[d];
            bb->absb.ext[0] = rows;
        }
        reduce_lw(&t, band, "selfrefo", "selfref", 0, 1, ngn, &bb->selfref);
        reduce_lw(&t, band, "forrefo", "forref", 0, 1, ngn, &bb->forref);
        reduce_lw(&t, band, "fracrefao", "fracrefa", 1, 0, ngn, &bb->fracrefa);
        reduce_lw(&t, band, "fracrefbo", "fracrefb", 1, 0, ngn, &bb->fracrefb);
        rarr_t *slots[] = {&bb->ka_mn2,  &bb->kb_mn2,  &bb->ka_mn2o, &bb->kb_mn2o,   &bb->ka_mo3,
                           &bb->kb_mo3,  &bb->ka_mco2, &bb->kb_mco2, &bb->ka_mco,    &bb->ka_mo2,
                           &bb->kb_mo2,  &bb->ccl4,    &bb->cfc11adj, &bb->cfc12,    &bb->cfc22adj};
        for (int s = 0; s < 15; ++s) reduce_lw(&t, band, kn[s][0], kn[s][1], 0, 1, ngn, slots[s]);
    }
    otbl_free(&t);
    L.ready = 1;
    return 0;
}

double oracle_lw_tbl(int which, int i)
{
    return which == 0 ? L.tau_tbl[i] : which == 1 ? L.exp_tbl[i] : L.tfn_tbl[i];
}

double oracle_heatfac(int lw);
double oracle_heatfac_sw(void);
double oracle_heatfac(int lw) { return lw ? L.heatfac : oracle_heatfac_sw(); }

/* ------------------------------------------------------------------------ */

typedef struct {
    int nlayers, laytrop;
    double pavel[MAXLAY + 1], tavel[MAXLAY + 1], pz[MAXLAY + 1], tz[MAXLAY + 1], tbound;
    double semiss[17], coldry[MAXLAY + 1], wkl[8][MAXLAY + 1], wbrodl[MAXLAY + 1], wx[5][MAXLAY + 1];
    double pwvcm;
    int jp[MAXLAY + 1], jt[MAXLAY + 1], jt1[MAXLAY + 1];
    double planklay[MAXLAY + 1][17], planklev[MAXLAY + 1][17], plankbnd[17];
    double colh2o[MAXLAY + 1], colco2[MAXLAY + 1], colo3[MAXLAY + 1], coln2o[MAXLAY + 1];
    double colco[MAXLAY + 1], colch4[MAXLAY + 1], colo2[MAXLAY + 1], colbrd[MAXLAY + 1];
    double fac00[MAXLAY + 1], fac01[MAXLAY + 1], fac10[MAXLAY + 1], fac11[MAXLAY + 1];
    double rat_h2oco2[MAXLAY + 1], rat_h2oco2_1[MAXLAY + 1], rat_h2oo3[MAXLAY + 1], rat_h2oo3_1[MAXLAY + 1];
    double rat_h2on2o[MAXLAY + 1], rat_h2on2o_1[MAXLAY + 1], rat_h2och4[MAXLAY + 1], rat_h2och4_1[MAXLAY + 1];
    double rat_n2oco2[MAXLAY + 1], rat_n2oco2_1[MAXLAY + 1], rat_o3co2[MAXLAY + 1], rat_o3co2_1[MAXLAY + 1];
    double selffac[MAXLAY + 1], selffrac[MAXLAY + 1], forfac[MAXLAY + 1], forfrac[MAXLAY + 1];
    int indself[MAXLAY + 1], indfor[MAXLAY + 1], indminor[MAXLAY + 1];
    double minorfrac[MAXLAY + 1], scaleminor[MAXLAY + 1], scaleminorn2[MAXLAY + 1];
    double taug[MAXLAY + 1][NGLW + 1], fracs[MAXLAY + 1][NGLW + 1];
} lwcol_t;

/* inatm (rrtmg_lw_rad.nomcica.f90:729-852), clear sky, iaer handled by caller */
static void inatm_u(lwcol_t *c, int iplon, int ncol, int nlay, const double *play, const double *plev,
                    const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                    const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                    const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                    const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                    const double *emis, int vmr);
/* MiMA's nomcica inatm: specific humidity and O3 mass mixing ratio in */
static void inatm(lwcol_t *c, int iplon, int ncol, int nlay, const double *play, const double *plev,
                  const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                  const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                  const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                  const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                  const double *emis)
{
    inatm_u(c, iplon, ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
            cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, 0);
}

/* vmr = 0: MiMA's units, converted as its nomcica inatm does
 * (rrtmg_lw_rad.nomcica.f90:775-778); vmr = 1: volume mixing ratios as the
 * McICA driver's inatm takes them (rrtmg_lw_rad.f90:778, 784) */
static void inatm_u(lwcol_t *c, int iplon, int ncol, int nlay, const double *play, const double *plev,
                    const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                    const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                    const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                    const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                    const double *emis, int vmr)
{
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
#define IN(a, i, k) a[(i) + (long)ncol * ((k)-1)]
    c->nlayers = nlay;
    for (int m = 0; m < 8; ++m)
        for (int l = 0; l <= nlay; ++l) c->wkl[m][l] = 0.0;
    for (int m = 0; m < 5; ++m)
        for (int l = 0; l <= nlay; ++l) c->wx[m][l] = 0.0;
    double amttl = 0.0, wvttl = 0.0;
    c->tbound = tsfc[iplon];
    c->pz[0] = IN(plev, iplon, 1);
    c->tz[0] = IN(tlev, iplon, 1);
    for (int l = 1; l <= nlay; ++l) {
        c->pavel[l] = IN(play, iplon, l);
        c->tavel[l] = IN(tlay, iplon, l);
        c->pz[l] = IN(plev, iplon, l + 1);
        c->tz[l] = IN(tlev, iplon, l + 1);
        c->wkl[1][l] = vmr ? IN(h2ovmr, iplon, l) : (IN(h2ovmr, iplon, l) / (1.0 - IN(h2ovmr, iplon, l))) * amdw;
        c->wkl[2][l] = IN(co2vmr, iplon, l);
        c->wkl[3][l] = vmr ? IN(o3vmr, iplon, l) : IN(o3vmr, iplon, l) * amdo;
        c->wkl[4][l] = IN(n2ovmr, iplon, l);
        c->wkl[6][l] = IN(ch4vmr, iplon, l);
        c->wkl[7][l] = IN(o2vmr, iplon, l);
        double amm = (1.0 - c->wkl[1][l]) * amd + c->wkl[1][l] * amw;
        c->coldry[l] = (c->pz[l - 1] - c->pz[l]) * 1.e3 * L.avogad / (1.e2 * L.grav * amm * (1.0 + c->wkl[1][l]));
    }
    for (int l = 1; l <= nlay; ++l) {
        c->wx[1][l] = IN(ccl4vmr, iplon, l);
        c->wx[2][l] = IN(cfc11vmr, iplon, l);
        c->wx[3][l] = IN(cfc12vmr, iplon, l);
        c->wx[4][l] = IN(cfc22vmr, iplon, l);
    }
    for (int l = 1; l <= nlay; ++l) {
        double summol = 0.0;
        for (int imol = 2; imol <= 7; ++imol) summol = summol + c->wkl[imol][l];
        c->wbrodl[l] = c->coldry[l] * (1.0 - summol);
        for (int imol = 1; imol <= 7; ++imol) c->wkl[imol][l] = c->coldry[l] * c->wkl[imol][l];
        amttl = amttl + c->coldry[l] + c->wkl[1][l];
        wvttl = wvttl + c->wkl[1][l];
        for (int ix = 1; ix <= 4; ++ix) c->wx[ix][l] = c->coldry[l] * c->wx[ix][l] * 1.e-20;
    }
    double wvsh = (amw * wvttl) / (amd * amttl);
    c->pwvcm = wvsh * (1.e3 * c->pz[0]) / (1.e2 * L.grav);
    for (int n = 1; n <= 16; ++n) c->semiss[n] = emis[iplon + (long)ncol * (n - 1)];
#undef IN
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* setcoef (rrtmg_lw_setcoef.f90:150-413), istart = 1, idrv = 0 */
static void setcoef(lwcol_t *c)
{
    const double stpfac = 296.0 / 1013.0;
    int indbound = clampi((int)(c->tbound - 159.0), 1, 180);
    double tbndfrac = c->tbound - 159.0 - (double)indbound;
    int indlev0 = clampi((int)(c->tz[0] - 159.0), 1, 180);
    double t0frac = c->tz[0] - 159.0 - (double)indlev0;
    c->laytrop = 0;
    for (int lay = 1; lay <= c->nlayers; ++lay) {
        int indlay = clampi((int)(c->tavel[lay] - 159.0), 1, 180);
        double tlayfrac = c->tavel[lay] - 159.0 - (double)indlay;
        int indlev = clampi((int)(c->tz[lay] - 159.0), 1, 180);
        double tlevfrac = c->tz[lay] - 159.0 - (double)indlev;
        /* bands 1-15, then band 16 through the same totplnk path (istart = 1) */
        for (int iband = 1; iband <= 16; ++iband) {
            double dbdtlev, dbdtlay;
            if (lay == 1) {
                dbdtlev = TOTPLNK(indbound + 1, iband) - TOTPLNK(indbound, iband);
                c->plankbnd[iband] = c->semiss[iband] * (TOTPLNK(indbound, iband) + tbndfrac * dbdtlev);
                dbdtlev = TOTPLNK(indlev0 + 1, iband) - TOTPLNK(indlev0, iband);
                c->planklev[0][iband] = TOTPLNK(indlev0, iband) + t0frac * dbdtlev;
            }
            dbdtlev = TOTPLNK(indlev + 1, iband) - TOTPLNK(indlev, iband);
            dbdtlay = TOTPLNK(indlay + 1, iband) - TOTPLNK(indlay, iband);
            c->planklay[lay][iband] = TOTPLNK(indlay, iband) + tlayfrac * dbdtlay;
            c->planklev[lay][iband] = TOTPLNK(indlev, iband) + tlevfrac * dbdtlev;
        }
        double plog = log(c->pavel[lay]);
        c->jp[lay] = clampi((int)(36.0 - 5 * (plog + 0.04)), 1, 58);
        int jp1 = c->jp[lay] + 1;
        double fp = 5.0 * (L.preflog[c->jp[lay] - 1] - plog);
        c->jt[lay] = clampi((int)(3.0 + (c->tavel[lay] - L.tref[c->jp[lay] - 1]) / 15.0), 1, 4);
        double ft = ((c->tavel[lay] - L.tref[c->jp[lay] - 1]) / 15.0) - (double)(c->jt[lay] - 3);
        c->jt1[lay] = clampi((int)(3.0 + (c->tavel[lay] - L.tref[jp1 - 1]) / 15.0), 1, 4);
        double ft1 = ((c->tavel[lay] - L.tref[jp1 - 1]) / 15.0) - (double)(c->jt1[lay] - 3);
        double water = c->wkl[1][lay] / c->coldry[lay];
        double scalefac = c->pavel[lay] * stpfac / c->tavel[lay];
        const int jp_ = c->jp[lay];
        if (plog > 4.56) {
            c->laytrop = c->laytrop + 1;
            c->forfac[lay] = scalefac / (1. + water);
            double factor = (332.0 - c->tavel[lay]) / 36.0;
            c->indfor[lay] = clampi((int)factor, 1, 2);
            c->forfrac[lay] = factor - (double)c->indfor[lay];
            c->selffac[lay] = water * c->forfac[lay];
            factor = (c->tavel[lay] - 188.0) / 7.2;
            c->indself[lay] = clampi((int)factor - 7, 1, 9);
            c->selffrac[lay] = factor - (double)(c->indself[lay] + 7);
            c->scaleminor[lay] = c->pavel[lay] / c->tavel[lay];
            c->scaleminorn2[lay] = (c->pavel[lay] / c->tavel[lay]) * (c->wbrodl[lay] / (c->coldry[lay] + c->wkl[1][lay]));
            factor = (c->tavel[lay] - 180.8) / 7.2;
            c->indminor[lay] = clampi((int)factor, 1, 18);
            c->minorfrac[lay] = factor - (double)c->indminor[lay];
            c->rat_h2oco2[lay] = CHI(1, jp_) / CHI(2, jp_);
            c->rat_h2oco2_1[lay] = CHI(1, jp_ + 1) / CHI(2, jp_ + 1);
            c->rat_h2oo3[lay] = CHI(1, jp_) / CHI(3, jp_);
            c->rat_h2oo3_1[lay] = CHI(1, jp_ + 1) / CHI(3, jp_ + 1);
            c->rat_h2on2o[lay] = CHI(1, jp_) / CHI(4, jp_);
            c->rat_h2on2o_1[lay] = CHI(1, jp_ + 1) / CHI(4, jp_ + 1);
            c->rat_h2och4[lay] = CHI(1, jp_) / CHI(6, jp_);
            c->rat_h2och4_1[lay] = CHI(1, jp_ + 1) / CHI(6, jp_ + 1);
            c->rat_n2oco2[lay] = CHI(4, jp_) / CHI(2, jp_);
            c->rat_n2oco2_1[lay] = CHI(4, jp_ + 1) / CHI(2, jp_ + 1);
        } else {
            c->forfac[lay] = scalefac / (1. + water);
            double factor = (c->tavel[lay] - 188.0) / 36.0;
            c->indfor[lay] = 3;
            c->forfrac[lay] = factor - 1.0;
            c->selffac[lay] = water * c->forfac[lay];
            c->scaleminor[lay] = c->pavel[lay] / c->tavel[lay];
            c->scaleminorn2[lay] = (c->pavel[lay] / c->tavel[lay]) * (c->wbrodl[lay] / (c->coldry[lay] + c->wkl[1][lay]));
            factor = (c->tavel[lay] - 180.8) / 7.2;
            c->indminor[lay] = clampi((int)factor, 1, 18);
            c->minorfrac[lay] = factor - (double)c->indminor[lay];
            c->rat_h2oco2[lay] = CHI(1, jp_) / CHI(2, jp_);
            c->rat_h2oco2_1[lay] = CHI(1, jp_ + 1) / CHI(2, jp_ + 1);
            c->rat_o3co2[lay] = CHI(3, jp_) / CHI(2, jp_);
            c->rat_o3co2_1[lay] = CHI(3, jp_ + 1) / CHI(2, jp_ + 1);
            c->indself[lay] = 0; /* not set by the reference above laytrop; unused there */
            c->selffrac[lay] = 0.0;
        }
        c->colh2o[lay] = 1.e-20 * c->wkl[1][lay];
        c->colco2[lay] = 1.e-20 * c->wkl[2][lay];
        c->colo3[lay] = 1.e-20 * c->wkl[3][lay];
        c->coln2o[lay] = 1.e-20 * c->wkl[4][lay];
        c->colco[lay] = 1.e-20 * c->wkl[5][lay];
        c->colch4[lay] = 1.e-20 * c->wkl[6][lay];
        c->colo2[lay] = 1.e-20 * c->wkl[7][lay];
        if (c->colco2[lay] == 0.0) c->colco2[lay] = 1.e-32 * c->coldry[lay];
        if (c->colo3[lay] == 0.0) c->colo3[lay] = 1.e-32 * c->coldry[lay];
        if (c->coln2o[lay] == 0.0) c->coln2o[lay] = 1.e-32 * c->coldry[lay];
        if (c->colco[lay] == 0.0) c->colco[lay] = 1.e-32 * c->coldry[lay];
        if (c->colch4[lay] == 0.0) c->colch4[lay] = 1.e-32 * c->coldry[lay];
        c->colbrd[lay] = 1.e-20 * c->wbrodl[lay];
        double compfp = 1. - fp;
        c->fac10[lay] = compfp * ft;
        c->fac00[lay] = compfp * (1.0 - ft);
        c->fac11[lay] = fp * ft1;
        c->fac01[lay] = fp * (1.0 - ft1);
        c->selffac[lay] = c->colh2o[lay] * c->selffac[lay];
        c->forfac[lay] = c->colh2o[lay] * c->forfac[lay];
    }
}

/* ------------------------------------------------------------------------ */
/* taumol helpers                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    double speccomb, specparm;
    int js;
    double fs;
} eta_t;

/* the speccomb/specparm/js/fs block that opens every binary-species band */
static void eta(double cola, double colb, double rat, double mult, eta_t *e)
{
    e->speccomb = cola + rat * colb;
    e->specparm = cola / e->speccomb;
    if (e->specparm >= oneminus_lw) e->specparm = oneminus_lw;
    double specmult = mult * (e->specparm);
    e->js = 1 + (int)specmult;
    e->fs = fmod(specmult, 1.0);
}

/* lower-atmosphere major absorber with the 3-point eta scheme near the edges
 * (e.g. rrtmg_lw_taumol.f90:516-668): one of the two pressure neighbours.
 * facA = fac00 or fac01, facB = fac10 or fac11 of the layer. */
static double major3(const rarr_t *absa, const eta_t *e, int ind, double facA, double facB, int ig)
{
    double f0, f1, f2, g0, g1, g2;
#define AB(i) A2(*absa, (i), ig)
    if (e->specparm < 0.125) {
        double p = e->fs - 1;
        double p2 = p * p, p4 = p2 * p2;
        double fk0 = p4, fk1 = 1 - p - 2.0 * p4, fk2 = p + p4;
        f0 = fk0 * facA; f1 = fk1 * facA; f2 = fk2 * facA;
        g0 = fk0 * facB; g1 = fk1 * facB; g2 = fk2 * facB;
        return e->speccomb * (f0 * AB(ind) + f1 * AB(ind + 1) + f2 * AB(ind + 2) + g0 * AB(ind + 9) +
                              g1 * AB(ind + 10) + g2 * AB(ind + 11));
    } else if (e->specparm > 0.875) {
        double p = -e->fs;
        double p2 = p * p, p4 = p2 * p2;
        double fk0 = p4, fk1 = 1 - p - 2.0 * p4, fk2 = p + p4;
        f0 = fk0 * facA; f1 = fk1 * facA; f2 = fk2 * facA;
        g0 = fk0 * facB; g1 = fk1 * facB; g2 = fk2 * facB;
        return e->speccomb * (f2 * AB(ind - 1) + f1 * AB(ind) + f0 * AB(ind + 1) + g2 * AB(ind + 8) +
                              g1 * AB(ind + 9) + g0 * AB(ind + 10));
    } else {
        f0 = (1.0 - e->fs) * facA;
        g0 = (1.0 - e->fs) * facB;
        f1 = e->fs * facA;
        g1 = e->fs * facB;
        return e->speccomb * (f0 * AB(ind) + f1 * AB(ind + 1) + g0 * AB(ind + 9) + g1 * AB(ind + 10));
    }
#undef AB
}

/* upper-atmosphere binary-species major term (e.g. rrtmg_lw_taumol.f90:740-749) */
static double major_b(const rarr_t *absb, const eta_t *e, int ind, double facA, double facB, int ig)
{
    double f0 = (1.0 - e->fs) * facA, g0 = (1.0 - e->fs) * facB;
    double f1 = e->fs * facA, g1 = e->fs * facB;
#define AB(i) A2(*absb, (i), ig)
    return e->speccomb * (f0 * AB(ind) + f1 * AB(ind + 1) + g0 * AB(ind + 5) + g1 * AB(ind + 6));
#undef AB
}

static double four(const rarr_t *ab, const lwcol_t *c, int lay, int ind0, int ind1, int ig)
{
    return c->fac00[lay] * A2(*ab, ind0, ig) + c->fac10[lay] * A2(*ab, ind0 + 1, ig) +
           c->fac01[lay] * A2(*ab, ind1, ig) + c->fac11[lay] * A2(*ab, ind1 + 1, ig);
}

static double tself(const lwband_t *bb, const lwcol_t *c, int lay, int ig)
{
    int inds = c->indself[lay];
    return c->selffac[lay] *
           (A2(bb->selfref, inds, ig) + c->selffrac[lay] * (A2(bb->selfref, inds + 1, ig) - A2(bb->selfref, inds, ig)));
}

static double tfor(const lwband_t *bb, const lwcol_t *c, int lay, int ig)
{
    int indf = c->indfor[lay];
    return c->forfac[lay] *
           (A2(bb->forref, indf, ig) + c->forfrac[lay] * (A2(bb->forref, indf + 1, ig) - A2(bb->forref, indf, ig)));
}

/* minor absorber interpolated in temperature only, k(19, ng) */
static double minor1(const rarr_t *k, const lwcol_t *c, int lay, int ig)
{
    int indm = c->indminor[lay];
    return A2(*k, indm, ig) + c->minorfrac[lay] * (A2(*k, indm + 1, ig) - A2(*k, indm, ig));
}

/* minor absorber interpolated in eta and temperature, k(neta, 19, ng) */
static double minor2(const rarr_t *k, const lwcol_t *c, int lay, int jm, double fm, int ig)
{
    int indm = c->indminor[lay];
    double m1 = A3(*k, jm, indm, ig) + fm * (A3(*k, jm + 1, indm, ig) - A3(*k, jm, indm, ig));
    double m2 = A3(*k, jm, indm + 1, ig) + fm * (A3(*k, jm + 1, indm + 1, ig) - A3(*k, jm, indm + 1, ig));
    return m1 + c->minorfrac[lay] * (m2 - m1);
}

static double planck_frac(const rarr_t *fr, const eta_t *e, int ig)
{
    return FR2(*fr, ig, e->js) + e->fs * (FR2(*fr, ig, e->js + 1) - FR2(*fr, ig, e->js));
}

/* column-amount adjustments of the minor CO2 / N2O (e.g. rrtmg_lw_taumol.f90:1314-1320) */
static double adjcol(double col, double coldry, double chi_ref, double thresh, double base, double ex)
{
    double chi = col / coldry;
    double rat = 1.e20 * chi / chi_ref;
    if (rat > thresh) {
        double adjfac = base + pow(rat - base, ex);
        return adjfac * chi_ref * coldry * 1.e-20;
    }
    return col;
}

static void taumol(lwcol_t *c)
{
    const int nl = c->nlayers, lt = c->laytrop;
    const int *jp = c->jp, *jt = c->jt, *jt1 = c->jt1;
    int lay, ig, ind0, ind1;
    const lwband_t *bb;
    eta_t e, e1, em, ep;

#define IND0A(nsp, js_) (((jp[lay] - 1) * 5 + (jt[lay] - 1)) * (nsp) + (js_))
#define IND1A(nsp, js_) ((jp[lay] * 5 + (jt1[lay] - 1)) * (nsp) + (js_))
#define IND0B(nsp, js_) (((jp[lay] - 13) * 5 + (jt[lay] - 1)) * (nsp) + (js_))
#define IND1B(nsp, js_) (((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * (nsp) + (js_))

    /* band 1: rrtmg_lw_taumol.f90:280-373 */
    bb = &L.b[0];
    for (lay = 1; lay <= lt; ++lay) {
        ind0 = IND0A(bb->nspa, 1);
        ind1 = IND1A(bb->nspa, 1);
        double pp = c->pavel[lay];
        double corradj = 1.;
        if (pp < 250.0) corradj = 1.0 - 0.15 * (250.0 - pp) / 154.4;
        double scalen2 = c->colbrd[lay] * c->scaleminorn2[lay];
        for (ig = 1; ig <= bb->ng; ++ig) {
            double taun2 = scalen2 * minor1(&bb->ka_mn2, c, lay, ig);
            c->taug[lay][ig] = corradj * (c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                          tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) + taun2);
            c->fracs[lay][ig] = G1(bb->fracrefa, ig);
        }
    }
    for (lay = lt + 1; lay <= nl; ++lay) {
        ind0 = IND0B(bb->nspb, 1);
        ind1 = IND1B(bb->nspb, 1);
        double pp = c->pavel[lay];
        double corradj = 1.0 - 0.15 * (pp / 95.6);
        double scalen2 = c->colbrd[lay] * c->scaleminorn2[lay];
        for (ig = 1; ig <= bb->ng; ++ig) {
            double taun2 = scalen2 * minor1(&bb->kb_mn2, c, lay, ig);
            c->taug[lay][ig] = corradj * (c->colh2o[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) +
                                          tfor(bb, c, lay, ig) + taun2);
            c->fracs[lay][ig] = G1(bb->fracrefb, ig);
        }
    }

    /* band 2: rrtmg_lw_taumol.f90:376-445 */
    bb = &L.b[1];
    {
        const int ngs = 10;
        for (lay = 1; lay <= lt; ++lay) {
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            double pp = c->pavel[lay];
            double corradj = 1.0 - .05 * (pp - 100.0) / 900.0;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = corradj * (c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                                    tself(bb, c, lay, ig) + tfor(bb, c, lay, ig));
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 3: rrtmg_lw_taumol.f90:448-760 */
    bb = &L.b[2];
    {
        const int ngs = 22;
        const double refrat_planck_a = CHI(1, 9) / CHI(2, 9), refrat_planck_b = CHI(1, 13) / CHI(2, 13);
        const double refrat_m_a = CHI(1, 3) / CHI(2, 3), refrat_m_b = CHI(1, 13) / CHI(2, 13);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colco2[lay], refrat_m_a, 8.0, &em);
            double adjcoln2o = adjcol(c->coln2o[lay], c->coldry[lay], CHI(4, jp[lay] + 1), 1.5, 0.5, 0.65);
            eta(c->colh2o[lay], c->colco2[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absn2o = minor2(&bb->ka_mn2o, c, lay, em.js, em.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) +
                                         adjcoln2o * absn2o;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2[lay], 4.0, &e);
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2_1[lay], 4.0, &e1);
            eta(c->colh2o[lay], c->colco2[lay], refrat_m_b, 4.0, &em);
            double adjcoln2o = adjcol(c->coln2o[lay], c->coldry[lay], CHI(4, jp[lay] + 1), 1.5, 0.5, 0.65);
            eta(c->colh2o[lay], c->colco2[lay], refrat_planck_b, 4.0, &ep);
            ind0 = IND0B(bb->nspb, e.js);
            ind1 = IND1B(bb->nspb, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absn2o = minor2(&bb->kb_mn2o, c, lay, em.js, em.fs, ig);
                c->taug[lay][ngs + ig] = major_b(&bb->absb, &e, ind0, c->fac00[lay], c->fac10[lay], ig) +
                                         major_b(&bb->absb, &e1, ind1, c->fac01[lay], c->fac11[lay], ig) +
                                         tfor(bb, c, lay, ig) + adjcoln2o * absn2o;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefb, &ep, ig);
            }
        }
    }

    /* band 4: rrtmg_lw_taumol.f90:763-1019 */
    bb = &L.b[3];
    {
        const int ngs = 38;
        const double refrat_planck_a = CHI(1, 11) / CHI(2, 11), refrat_planck_b = CHI(3, 13) / CHI(2, 13);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colco2[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            eta(c->colo3[lay], c->colco2[lay], c->rat_o3co2[lay], 4.0, &e);
            eta(c->colo3[lay], c->colco2[lay], c->rat_o3co2_1[lay], 4.0, &e1);
            eta(c->colo3[lay], c->colco2[lay], refrat_planck_b, 4.0, &ep);
            ind0 = IND0B(bb->nspb, e.js);
            ind1 = IND1B(bb->nspb, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = major_b(&bb->absb, &e, ind0, c->fac00[lay], c->fac10[lay], ig) +
                                         major_b(&bb->absb, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefb, &ep, ig);
            }
            c->taug[lay][ngs + 8] = c->taug[lay][ngs + 8] * 0.92;
            c->taug[lay][ngs + 9] = c->taug[lay][ngs + 9] * 0.88;
            c->taug[lay][ngs + 10] = c->taug[lay][ngs + 10] * 1.07;
            c->taug[lay][ngs + 11] = c->taug[lay][ngs + 11] * 1.1;
            c->taug[lay][ngs + 12] = c->taug[lay][ngs + 12] * 0.99;
            c->taug[lay][ngs + 13] = c->taug[lay][ngs + 13] * 0.88;
            c->taug[lay][ngs + 14] = c->taug[lay][ngs + 14] * 0.943;
        }
    }

    /* band 5: rrtmg_lw_taumol.f90:1022-1294 */
    bb = &L.b[4];
    {
        const int ngs = 52;
        const double refrat_planck_a = CHI(1, 5) / CHI(2, 5), refrat_planck_b = CHI(3, 43) / CHI(2, 43);
        const double refrat_m_a = CHI(1, 7) / CHI(2, 7);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colco2[lay], refrat_m_a, 8.0, &em);
            eta(c->colh2o[lay], c->colco2[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double abso3 = minor2(&bb->ka_mo3, c, lay, em.js, em.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) +
                                         abso3 * c->colo3[lay] + c->wx[1][lay] * G1(bb->ccl4, ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            eta(c->colo3[lay], c->colco2[lay], c->rat_o3co2[lay], 4.0, &e);
            eta(c->colo3[lay], c->colco2[lay], c->rat_o3co2_1[lay], 4.0, &e1);
            eta(c->colo3[lay], c->colco2[lay], refrat_planck_b, 4.0, &ep);
            ind0 = IND0B(bb->nspb, e.js);
            ind1 = IND1B(bb->nspb, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = major_b(&bb->absb, &e, ind0, c->fac00[lay], c->fac10[lay], ig) +
                                         major_b(&bb->absb, &e1, ind1, c->fac01[lay], c->fac11[lay], ig) +
                                         c->wx[1][lay] * G1(bb->ccl4, ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefb, &ep, ig);
            }
        }
    }

    /* band 6: rrtmg_lw_taumol.f90:1297-1380 */
    bb = &L.b[5];
    {
        const int ngs = 68;
        for (lay = 1; lay <= lt; ++lay) {
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], CHI(2, jp[lay] + 1), 3.0, 2.0, 0.77);
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor1(&bb->ka_mco2, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                         tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) + adjcolco2 * absco2 +
                                         c->wx[2][lay] * G1(bb->cfc11adj, ig) + c->wx[3][lay] * G1(bb->cfc12, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = 0.0 + c->wx[2][lay] * G1(bb->cfc11adj, ig) + c->wx[3][lay] * G1(bb->cfc12, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
    }

    /* band 7: rrtmg_lw_taumol.f90:1383-1654 */
    bb = &L.b[6];
    {
        const int ngs = 76;
        const double refrat_planck_a = CHI(1, 3) / CHI(3, 3), refrat_m_a = CHI(1, 3) / CHI(3, 3);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colo3[lay], c->rat_h2oo3[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colo3[lay], c->rat_h2oo3_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colo3[lay], refrat_m_a, 8.0, &em);
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], CHI(2, jp[lay] + 1), 3.0, 3.0, 0.79);
            eta(c->colh2o[lay], c->colo3[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor2(&bb->ka_mco2, c, lay, em.js, em.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) +
                                         adjcolco2 * absco2;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], CHI(2, jp[lay] + 1), 3.0, 2.0, 0.79);
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor1(&bb->kb_mco2, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colo3[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) + adjcolco2 * absco2;
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
            c->taug[lay][ngs + 6] = c->taug[lay][ngs + 6] * 0.92;
            c->taug[lay][ngs + 7] = c->taug[lay][ngs + 7] * 0.88;
            c->taug[lay][ngs + 8] = c->taug[lay][ngs + 8] * 1.07;
            c->taug[lay][ngs + 9] = c->taug[lay][ngs + 9] * 1.1;
            c->taug[lay][ngs + 10] = c->taug[lay][ngs + 10] * 0.99;
            c->taug[lay][ngs + 11] = c->taug[lay][ngs + 11] * 0.855;
        }
    }

    /* band 8: rrtmg_lw_taumol.f90:1657-1777 */
    bb = &L.b[7];
    {
        const int ngs = 88;
        for (lay = 1; lay <= lt; ++lay) {
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], CHI(2, jp[lay] + 1), 3.0, 2.0, 0.65);
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor1(&bb->ka_mco2, c, lay, ig);
                double abso3 = minor1(&bb->ka_mo3, c, lay, ig);
                double absn2o = minor1(&bb->ka_mn2o, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                         tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) + adjcolco2 * absco2 +
                                         c->colo3[lay] * abso3 + c->coln2o[lay] * absn2o +
                                         c->wx[3][lay] * G1(bb->cfc12, ig) + c->wx[4][lay] * G1(bb->cfc22adj, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], CHI(2, jp[lay] + 1), 3.0, 2.0, 0.65);
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor1(&bb->kb_mco2, c, lay, ig);
                double absn2o = minor1(&bb->kb_mn2o, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colo3[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) +
                                         adjcolco2 * absco2 + c->coln2o[lay] * absn2o +
                                         c->wx[3][lay] * G1(bb->cfc12, ig) + c->wx[4][lay] * G1(bb->cfc22adj, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 9: rrtmg_lw_taumol.f90:1780-2040 */
    bb = &L.b[8];
    {
        const int ngs = 96;
        const double refrat_planck_a = CHI(1, 9) / CHI(6, 9), refrat_m_a = CHI(1, 3) / CHI(6, 3);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colch4[lay], c->rat_h2och4[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colch4[lay], c->rat_h2och4_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colch4[lay], refrat_m_a, 8.0, &em);
            double adjcoln2o = adjcol(c->coln2o[lay], c->coldry[lay], CHI(4, jp[lay] + 1), 1.5, 0.5, 0.65);
            eta(c->colh2o[lay], c->colch4[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absn2o = minor2(&bb->ka_mn2o, c, lay, em.js, em.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) +
                                         adjcoln2o * absn2o;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            double adjcoln2o = adjcol(c->coln2o[lay], c->coldry[lay], CHI(4, jp[lay] + 1), 1.5, 0.5, 0.65);
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absn2o = minor1(&bb->kb_mn2o, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colch4[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) + adjcoln2o * absn2o;
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 10: rrtmg_lw_taumol.f90:2043-2107 */
    bb = &L.b[9];
    {
        const int ngs = 108;
        for (lay = 1; lay <= lt; ++lay) {
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                         tself(bb, c, lay, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 11: rrtmg_lw_taumol.f90:2110-2187 */
    bb = &L.b[10];
    {
        const int ngs = 114;
        for (lay = 1; lay <= lt; ++lay) {
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            double scaleo2 = c->colo2[lay] * c->scaleminor[lay];
            for (ig = 1; ig <= bb->ng; ++ig) {
                double tauo2 = scaleo2 * minor1(&bb->ka_mo2, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                         tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) + tauo2;
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            double scaleo2 = c->colo2[lay] * c->scaleminor[lay];
            for (ig = 1; ig <= bb->ng; ++ig) {
                double tauo2 = scaleo2 * minor1(&bb->kb_mo2, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colh2o[lay] * four(&bb->absb, c, lay, ind0, ind1, ig) +
                                         tfor(bb, c, lay, ig) + tauo2;
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 12: rrtmg_lw_taumol.f90:2190-2392 */
    bb = &L.b[11];
    {
        const int ngs = 122;
        const double refrat_planck_a = CHI(1, 10) / CHI(2, 10);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colco2[lay], c->rat_h2oco2_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colco2[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = 0.0;
                c->fracs[lay][ngs + ig] = 0.0;
            }
    }

    /* band 13: rrtmg_lw_taumol.f90:2395-2652 */
    bb = &L.b[12];
    {
        const int ngs = 130;
        const double refrat_planck_a = CHI(1, 5) / CHI(4, 5), refrat_m_a = CHI(1, 1) / CHI(4, 1);
        const double refrat_m_a3 = CHI(1, 3) / CHI(4, 3);
        eta_t emco;
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->coln2o[lay], c->rat_h2on2o[lay], 8.0, &e);
            eta(c->colh2o[lay], c->coln2o[lay], c->rat_h2on2o_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->coln2o[lay], refrat_m_a, 8.0, &em);
            double adjcolco2 = adjcol(c->colco2[lay], c->coldry[lay], 3.55e-4, 3.0, 2.0, 0.68);
            eta(c->colh2o[lay], c->coln2o[lay], refrat_m_a3, 8.0, &emco);
            eta(c->colh2o[lay], c->coln2o[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double absco2 = minor2(&bb->ka_mco2, c, lay, em.js, em.fs, ig);
                double absco = minor2(&bb->ka_mco, c, lay, emco.js, emco.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) +
                                         adjcolco2 * absco2 + c->colco[lay] * absco;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                double abso3 = minor1(&bb->kb_mo3, c, lay, ig);
                c->taug[lay][ngs + ig] = c->colo3[lay] * abso3;
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
    }

    /* band 14: rrtmg_lw_taumol.f90:2655-2713 */
    bb = &L.b[13];
    {
        const int ngs = 134;
        for (lay = 1; lay <= lt; ++lay) {
            ind0 = IND0A(bb->nspa, 1);
            ind1 = IND1A(bb->nspa, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colco2[lay] * four(&bb->absa, c, lay, ind0, ind1, ig) +
                                         tself(bb, c, lay, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefa, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colco2[lay] * four(&bb->absb, c, lay, ind0, ind1, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }

    /* band 15: rrtmg_lw_taumol.f90:2716-2938 */
    bb = &L.b[14];
    {
        const int ngs = 136;
        const double refrat_planck_a = CHI(4, 1) / CHI(2, 1), refrat_m_a = CHI(4, 1) / CHI(2, 1);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->coln2o[lay], c->colco2[lay], c->rat_n2oco2[lay], 8.0, &e);
            eta(c->coln2o[lay], c->colco2[lay], c->rat_n2oco2_1[lay], 8.0, &e1);
            eta(c->coln2o[lay], c->colco2[lay], refrat_m_a, 8.0, &em);
            eta(c->coln2o[lay], c->colco2[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            double scalen2 = c->colbrd[lay] * c->scaleminor[lay];
            for (ig = 1; ig <= bb->ng; ++ig) {
                double taun2 = scalen2 * minor2(&bb->ka_mn2, c, lay, em.js, em.fs, ig);
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig) + taun2;
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = 0.0;
                c->fracs[lay][ngs + ig] = 0.0;
            }
    }

    /* band 16: rrtmg_lw_taumol.f90:2941-3147 (nspb(16) = 0: ind0 = ind1 = 1 aloft) */
    bb = &L.b[15];
    {
        const int ngs = 138;
        const double refrat_planck_a = CHI(1, 6) / CHI(6, 6);
        for (lay = 1; lay <= lt; ++lay) {
            eta(c->colh2o[lay], c->colch4[lay], c->rat_h2och4[lay], 8.0, &e);
            eta(c->colh2o[lay], c->colch4[lay], c->rat_h2och4_1[lay], 8.0, &e1);
            eta(c->colh2o[lay], c->colch4[lay], refrat_planck_a, 8.0, &ep);
            ind0 = IND0A(bb->nspa, e.js);
            ind1 = IND1A(bb->nspa, e1.js);
            for (ig = 1; ig <= bb->ng; ++ig) {
                double tau_major = major3(&bb->absa, &e, ind0, c->fac00[lay], c->fac10[lay], ig);
                double tau_major1 = major3(&bb->absa, &e1, ind1, c->fac01[lay], c->fac11[lay], ig);
                c->taug[lay][ngs + ig] = tau_major + tau_major1 + tself(bb, c, lay, ig) + tfor(bb, c, lay, ig);
                c->fracs[lay][ngs + ig] = planck_frac(&bb->fracrefa, &ep, ig);
            }
        }
        for (lay = lt + 1; lay <= nl; ++lay) {
            ind0 = IND0B(bb->nspb, 1);
            ind1 = IND1B(bb->nspb, 1);
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colch4[lay] * four(&bb->absb, c, lay, ind0, ind1, ig);
                c->fracs[lay][ngs + ig] = G1(bb->fracrefb, ig);
            }
        }
    }
#undef IND0A
#undef IND1A
#undef IND0B
#undef IND1B
}

/* rtrnmr clear-sky branch (rrtmg_lw_rtrnmr.f90:246-777 with icldlyr = 0,
 * iout = 0, idrv = 0); taut = taug + taua with taua = 0 (iaer = 10). */
static void rtrnmr(const lwcol_t *c, double *totuflux, double *totdflux, double *htr)
{
    const int nl = c->nlayers;
    const double wtdiff = 0.5, rec_6 = 0.166667;
    const double pi = 2.0 * asin(1.0);
    const double fluxfac = pi * 2.e4;
    double secdiff[17];
    double urad[MAXLAY + 1], drad[MAXLAY + 1], atrans[MAXLAY + 1], bbugas[MAXLAY + 1];
    double fnet[MAXLAY + 1];
    for (int ibnd = 1; ibnd <= 16; ++ibnd) {
        if (ibnd == 1 || ibnd == 4 || ibnd >= 10) {
            secdiff[ibnd] = 1.66;
        } else {
            secdiff[ibnd] = L.a0[ibnd - 1] + L.a1[ibnd - 1] * exp(L.a2[ibnd - 1] * c->pwvcm);
            if (secdiff[ibnd] > 1.80) secdiff[ibnd] = 1.80;
            if (secdiff[ibnd] < 1.50) secdiff[ibnd] = 1.50;
        }
    }
    for (int lev = 0; lev <= nl; ++lev) {
        urad[lev] = 0.0;
        drad[lev] = 0.0;
        totuflux[lev] = 0.0;
        totdflux[lev] = 0.0;
    }
    int igc = 1;
    for (int iband = 1; iband <= 16; ++iband) {
        for (; igc <= L.ngs[iband - 1]; ++igc) {
            double radld = 0.0;
            for (int lev = nl; lev >= 1; --lev) {
                double plfrac = c->fracs[lev][igc];
                double blay = c->planklay[lev][iband];
                double dplankup = c->planklev[lev][iband] - blay;
                double dplankdn = c->planklev[lev - 1][iband] - blay;
                double odepth = secdiff[iband] * c->taug[lev][igc];
                if (odepth < 0.0) odepth = 0.0;
                double bbd;
                if (odepth <= 0.06) {
                    atrans[lev] = odepth - 0.5 * odepth * odepth;
                    odepth = rec_6 * odepth;
                    bbd = plfrac * (blay + dplankdn * odepth);
                    bbugas[lev] = plfrac * (blay + dplankup * odepth);
                } else {
                    double tblind = odepth / (L.bpade + odepth);
                    int itr = (int)(1.e4 * tblind + 0.5);
                    double transc = L.exp_tbl[itr];
                    atrans[lev] = 1.0 - transc;
                    double tausfac = L.tfn_tbl[itr];
                    bbd = plfrac * (blay + tausfac * dplankdn);
                    bbugas[lev] = plfrac * (blay + tausfac * dplankup);
                }
                radld = radld + (bbd - radld) * atrans[lev];
                drad[lev - 1] = drad[lev - 1] + radld;
            }
            double rad0 = c->fracs[1][igc] * c->plankbnd[iband];
            double reflect = 1.0 - c->semiss[iband];
            double radlu = rad0 + reflect * radld;
            urad[0] = urad[0] + radlu;
            for (int lev = 1; lev <= nl; ++lev) {
                radlu = radlu + (bbugas[lev] - radlu) * atrans[lev];
                urad[lev] = urad[lev] + radlu;
            }
        }
        for (int lev = nl; lev >= 0; --lev) {
            double uflux = urad[lev] * wtdiff;
            double dflux = drad[lev] * wtdiff;
            urad[lev] = 0.0;
            drad[lev] = 0.0;
            totuflux[lev] = totuflux[lev] + uflux * L.delwave[iband - 1];
            totdflux[lev] = totdflux[lev] + dflux * L.delwave[iband - 1];
        }
    }
    totuflux[0] = totuflux[0] * fluxfac;
    totdflux[0] = totdflux[0] * fluxfac;
    fnet[0] = totuflux[0] - totdflux[0];
    for (int lev = 1; lev <= nl; ++lev) {
        totuflux[lev] = totuflux[lev] * fluxfac;
        totdflux[lev] = totdflux[lev] * fluxfac;
        fnet[lev] = totuflux[lev] - totdflux[lev];
        int l = lev - 1;
        htr[l] = L.heatfac * (fnet[l] - fnet[lev]) / (c->pz[l] - c->pz[lev]);
    }
    htr[nl] = 0.0;
}

int oracle_rrtmg_lw(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                    const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                    const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                    const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                    const double *cfc22vmr, const double *ccl4vmr, const double *emis, double *uflx,
                    double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc)
{
    if (!L.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    int err = 0;
#pragma omp parallel
    {
        lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
        double tu[MAXLAY + 1], td[MAXLAY + 1], htr[MAXLAY + 1];
        if (!c) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (!c) continue;
            inatm(c, iplon, ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                  o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis);
            setcoef(c);
            taumol(c);
            rtrnmr(c, tu, td, htr);
            for (int k = 0; k <= nlay; ++k) {
                uflx[iplon + (long)ncol * k] = tu[k];
                dflx[iplon + (long)ncol * k] = td[k];
                uflxc[iplon + (long)ncol * k] = tu[k];
                dflxc[iplon + (long)ncol * k] = td[k];
            }
            for (int k = 0; k < nlay; ++k) {
                hr[iplon + (long)ncol * k] = htr[k];
                hrc[iplon + (long)ncol * k] = htr[k];
            }
        }
        free(c);
    }
    return err;
}

/* ---- cloudy sky (icld = 1..3, inflglw = 0) -------------------------------- */

/* cldprop with inflag = 0 (rrtmg_lw_cldprop.f90:154-176): a layer is cloudy
 * for cldprop when its fraction and its water path or summed optical depth
 * reach cldmin = 1e-20; its in-cloud optical depths are then the input's,
 * otherwise zero.  tauc[lay][ib], 1-based. */
/* lwcldpr's cloud absorption coefficients (rrtmg_lw_init.f90, module
 * rrlw_cld), from rrtmg_lw_cld.tbl; Fortran (k, ib) arrays as [ib][k] */
static struct {
    int ready;
    double abscld1, absice0[2], absice1[5][2], absice2[16][43], absice3[16][46], absliq0, absliq1[16][58];
} LC;

int oracle_lw_cld_init(const char *path)
{
    otbl_t t;
    if (otbl_load(&t, path) != 0) return -1;
    const char *nm[] = {"lwcld/abscld1", "lwcld/absice0", "lwcld/absice1", "lwcld/absice2", "lwcld/absice3",
                        "lwcld/absliq0", "lwcld/absliq1"};
    double *dst[] = {&LC.abscld1, LC.absice0, &LC.absice1[0][0], &LC.absice2[0][0], &LC.absice3[0][0],
                     &LC.absliq0, &LC.absliq1[0][0]};
    const long len[] = {1, 2, 10, 688, 736, 1, 928};
    for (int i = 0; i < 7; ++i) {
        long n = 0;
        const double *v = otbl_d(&t, nm[i], &n);
        if (!v || n != len[i]) return -2;
        memcpy(dst[i], v, sizeof(double) * len[i]);
    }
    LC.ready = 1;
    return 0;
}

/* cldprop (rrtmg_lw_cldprop.f90:118-291) for inflag 1 (abscld1 times the
 * water path) and 2 (ice: iceflag 0 absice0, 1 absice1 (5 bands), 2 absice2,
 * 3 absice3 interpolated in the radius; liquid: liqflag 0 absliq0, 1
 * absliq1 interpolated).  As in the reference, ncbands carries from layer to
 * layer (it starts at 1, and iceflag 0 with liqflag 0 or no liquid leaves
 * it), so only the first ncbands bands of a layer get a depth, and a layer
 * with 5-band ice and no 16-band liquid reads its 5 ice coefficients by
 * band (iceind 1 -> 2).  The final ncbands goes to the transfer, which
 * maps each spectral band to a cloud band by it.  -8 for a radius the
 * reference stops on. */
static int cldprop_wp(int nl, int inflag, int iceflag, int liqflag, const double *cldfrac, double (*tauc)[17],
                      const double *ciwp, const double *clwp, const double *rei, const double *rel,
                      double (*taucloud)[17], int *ncbands_out)
{
    static const int icb[3][16] = {{1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
                                   {1, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5},
                                   {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16}};
    const double cldmin = 1.e-20;
    int ncbands = 1;
    double abscoice[17], abscoliq[17];
    for (int lay = 1; lay <= nl; ++lay) {
        double tauctot = 0.0;
        for (int ib = 1; ib <= 16; ++ib) {
            taucloud[lay][ib] = 0.0;
            tauctot = tauctot + tauc[lay][ib];
        }
        double cwp = ciwp[lay] + clwp[lay];
        if (!(cldfrac[lay] >= cldmin && (cwp >= cldmin || tauctot >= cldmin))) continue;
        if (inflag == 1) {
            ncbands = 16;
            for (int ib = 1; ib <= ncbands; ++ib) taucloud[lay][ib] = LC.abscld1 * cwp;
            continue;
        }
        double radice = rei[lay];
        int iceind = 0, liqind = 0;
        if (ciwp[lay] == 0.0) {
            abscoice[1] = 0.0;
            iceind = 0;
        } else if (iceflag == 0) {
            if (radice < 10.0) return -8;
            abscoice[1] = LC.absice0[0] + LC.absice0[1] / radice;
            iceind = 0;
        } else if (iceflag == 1) {
            if (radice < 13.0 || radice > 130.0) return -8;
            ncbands = 5;
            for (int ib = 1; ib <= ncbands; ++ib) abscoice[ib] = LC.absice1[ib - 1][0] + LC.absice1[ib - 1][1] / radice;
            iceind = 1;
        } else if (iceflag == 2 || iceflag == 3) {
            const int nr = iceflag == 2 ? 43 : 46;
            if (radice < 5.0 || radice > (iceflag == 2 ? 131.0 : 140.0)) return -8;
            ncbands = 16;
            double factor = (radice - 2.0) / 3.0;
            int index = (int)factor;
            if (index == nr) index = nr - 1;
            double fint = factor - (double)index;
            for (int ib = 1; ib <= ncbands; ++ib) {
                const double *a = iceflag == 2 ? LC.absice2[ib - 1] : LC.absice3[ib - 1];
                abscoice[ib] = a[index - 1] + fint * (a[index] - (a[index - 1]));
            }
            iceind = 2;
        }
        if (clwp[lay] == 0.0) {
            abscoliq[1] = 0.0;
            liqind = 0;
            if (iceind == 1) iceind = 2;
        } else if (liqflag == 0) {
            abscoliq[1] = LC.absliq0;
            liqind = 0;
            if (iceind == 1) iceind = 2;
        } else if (liqflag == 1) {
            double radliq = rel[lay];
            if (radliq < 2.5 || radliq > 60.0) return -8;
            int index = (int)(radliq - 1.5);
            if (index == 0) index = 1;
            if (index == 58) index = 57;
            double fint = radliq - 1.5 - (double)index;
            ncbands = 16;
            for (int ib = 1; ib <= ncbands; ++ib) {
                const double *a = LC.absliq1[ib - 1];
                abscoliq[ib] = a[index - 1] + fint * (a[index] - (a[index - 1]));
            }
            liqind = 2;
        }
        for (int ib = 1; ib <= ncbands; ++ib)
            taucloud[lay][ib] = ciwp[lay] * abscoice[icb[iceind][ib - 1]] + clwp[lay] * abscoliq[icb[liqind][ib - 1]];
    }
    *ncbands_out = ncbands;
    return 0;
}

static void cldprop0(int nl, const double *cldfrac, double (*tauc)[17], const double *cwp, double (*taucloud)[17])
{
    const double cldmin = 1.e-20;
    for (int lay = 1; lay <= nl; ++lay) {
        double tauctot = 0.0;
        for (int ib = 1; ib <= 16; ++ib) {
            taucloud[lay][ib] = 0.0;
            tauctot = tauctot + tauc[lay][ib];
        }
        if (cldfrac[lay] >= cldmin && (cwp[lay] >= cldmin || tauctot >= cldmin))
            for (int ib = 1; ib <= 16; ++ib) taucloud[lay][ib] = tauc[lay][ib];
    }
}

/* Maximum-random overlap factors of rtrnmr (rrtmg_lw_rtrnmr.f90:325-476).
 * cldfrac[0] and cldfrac[nl+1] are 0: the reference reads cldfrac(0) at
 * lev = 1 only as a factor of faccld2(1) = facclr2(1) = 0, and the driver's
 * cldfrac(nlay+1) is zero (rrtmg_lw_rad.nomcica.f90:736).  Arrays 0..nl+1. */
typedef struct {
    double faccld1[MAXLAY + 2], faccld2[MAXLAY + 2], facclr1[MAXLAY + 2], facclr2[MAXLAY + 2];
    double faccmb1[MAXLAY + 2], faccmb2[MAXLAY + 2];
    double faccld1d[MAXLAY + 2], faccld2d[MAXLAY + 2], facclr1d[MAXLAY + 2], facclr2d[MAXLAY + 2];
    double faccmb1d[MAXLAY + 2], faccmb2d[MAXLAY + 2];
    int istcld[MAXLAY + 2], istcldd[MAXLAY + 2];
} overlap_t;

static void overlap_mr(int nl, const double *cldfrac, const int *icldlyr, overlap_t *o)
{
    double rat1 = 0.0, rat2 = 0.0, fmax, fmin;
    memset(o, 0, sizeof *o);
    o->istcld[1] = 1;
    o->istcldd[nl] = 1;
    for (int lev = 1; lev <= nl; ++lev) {
        if (icldlyr[lev] == 1) {
            o->istcld[lev + 1] = 0;
            if (lev == nl) {
                o->faccld1[lev + 1] = 0.0;
                o->faccld2[lev + 1] = 0.0;
                o->facclr1[lev + 1] = 0.0;
                o->facclr2[lev + 1] = 0.0;
                o->faccmb1[lev + 1] = 0.0;
                o->faccmb2[lev + 1] = 0.0;
            } else if (cldfrac[lev + 1] >= cldfrac[lev]) {
                o->faccld1[lev + 1] = 0.0;
                o->faccld2[lev + 1] = 0.0;
                if (o->istcld[lev] == 1) {
                    o->facclr1[lev + 1] = 0.0;
                    o->facclr2[lev + 1] = 0.0;
                    if (cldfrac[lev] < 1.0) o->facclr2[lev + 1] = (cldfrac[lev + 1] - cldfrac[lev]) / (1.0 - cldfrac[lev]);
                    o->facclr2[lev] = 0.0;
                    o->faccld2[lev] = 0.0;
                } else {
                    fmax = cldfrac[lev] > cldfrac[lev - 1] ? cldfrac[lev] : cldfrac[lev - 1];
                    if (cldfrac[lev + 1] > fmax) {
                        o->facclr1[lev + 1] = rat2;
                        o->facclr2[lev + 1] = (cldfrac[lev + 1] - fmax) / (1.0 - fmax);
                    } else if (cldfrac[lev + 1] < fmax) {
                        o->facclr1[lev + 1] = (cldfrac[lev + 1] - cldfrac[lev]) / (cldfrac[lev - 1] - cldfrac[lev]);
                        o->facclr2[lev + 1] = 0.0;
                    } else {
                        o->facclr1[lev + 1] = rat2;
                        o->facclr2[lev + 1] = 0.0;
                    }
                }
                if (o->facclr1[lev + 1] > 0.0 || o->facclr2[lev + 1] > 0.0) {
                    rat1 = 1.0;
                    rat2 = 0.0;
                } else {
                    rat1 = 0.0;
                    rat2 = 0.0;
                }
            } else {
                o->facclr1[lev + 1] = 0.0;
                o->facclr2[lev + 1] = 0.0;
                if (o->istcld[lev] == 1) {
                    o->faccld1[lev + 1] = 0.0;
                    o->faccld2[lev + 1] = (cldfrac[lev] - cldfrac[lev + 1]) / cldfrac[lev];
                    o->facclr2[lev] = 0.0;
                    o->faccld2[lev] = 0.0;
                } else {
                    fmin = cldfrac[lev] < cldfrac[lev - 1] ? cldfrac[lev] : cldfrac[lev - 1];
                    if (cldfrac[lev + 1] <= fmin) {
                        o->faccld1[lev + 1] = rat1;
                        o->faccld2[lev + 1] = (fmin - cldfrac[lev + 1]) / fmin;
                    } else {
                        o->faccld1[lev + 1] = (cldfrac[lev] - cldfrac[lev + 1]) / (cldfrac[lev] - fmin);
                        o->faccld2[lev + 1] = 0.0;
                    }
                }
                if (o->faccld1[lev + 1] > 0.0 || o->faccld2[lev + 1] > 0.0) {
                    rat1 = 0.0;
                    rat2 = 1.0;
                } else {
                    rat1 = 0.0;
                    rat2 = 0.0;
                }
            }
            o->faccmb1[lev + 1] = o->facclr1[lev + 1] * o->faccld2[lev] * cldfrac[lev - 1];
            o->faccmb2[lev + 1] = o->faccld1[lev + 1] * o->facclr2[lev] * (1.0 - cldfrac[lev - 1]);
        } else {
            o->istcld[lev + 1] = 1;
        }
    }
    for (int lev = nl; lev >= 1; --lev) {
        if (icldlyr[lev] == 1) {
            o->istcldd[lev - 1] = 0;
            if (lev == 1) {
                o->faccld1d[lev - 1] = 0.0;
                o->faccld2d[lev - 1] = 0.0;
                o->facclr1d[lev - 1] = 0.0;
                o->facclr2d[lev - 1] = 0.0;
                o->faccmb1d[lev - 1] = 0.0;
                o->faccmb2d[lev - 1] = 0.0;
            } else if (cldfrac[lev - 1] >= cldfrac[lev]) {
                o->faccld1d[lev - 1] = 0.0;
                o->faccld2d[lev - 1] = 0.0;
                if (o->istcldd[lev] == 1) {
                    o->facclr1d[lev - 1] = 0.0;
                    o->facclr2d[lev - 1] = 0.0;
                    if (cldfrac[lev] < 1.0) o->facclr2d[lev - 1] = (cldfrac[lev - 1] - cldfrac[lev]) / (1.0 - cldfrac[lev]);
                    o->facclr2d[lev] = 0.0;
                    o->faccld2d[lev] = 0.0;
                } else {
                    fmax = cldfrac[lev] > cldfrac[lev + 1] ? cldfrac[lev] : cldfrac[lev + 1];
                    if (cldfrac[lev - 1] > fmax) {
                        o->facclr1d[lev - 1] = rat2;
                        o->facclr2d[lev - 1] = (cldfrac[lev - 1] - fmax) / (1.0 - fmax);
                    } else if (cldfrac[lev - 1] < fmax) {
                        o->facclr1d[lev - 1] = (cldfrac[lev - 1] - cldfrac[lev]) / (cldfrac[lev + 1] - cldfrac[lev]);
                        o->facclr2d[lev - 1] = 0.0;
                    } else {
                        o->facclr1d[lev - 1] = rat2;
                        o->facclr2d[lev - 1] = 0.0;
                    }
                }
                if (o->facclr1d[lev - 1] > 0.0 || o->facclr2d[lev - 1] > 0.0) {
                    rat1 = 1.0;
                    rat2 = 0.0;
                } else {
                    rat1 = 0.0;
                    rat2 = 0.0;
                }
            } else {
                o->facclr1d[lev - 1] = 0.0;
                o->facclr2d[lev - 1] = 0.0;
                if (o->istcldd[lev] == 1) {
                    o->faccld1d[lev - 1] = 0.0;
                    o->faccld2d[lev - 1] = (cldfrac[lev] - cldfrac[lev - 1]) / cldfrac[lev];
                    o->facclr2d[lev] = 0.0;
                    o->faccld2d[lev] = 0.0;
                } else {
                    fmin = cldfrac[lev] < cldfrac[lev + 1] ? cldfrac[lev] : cldfrac[lev + 1];
                    if (cldfrac[lev - 1] <= fmin) {
                        o->faccld1d[lev - 1] = rat1;
                        o->faccld2d[lev - 1] = (fmin - cldfrac[lev - 1]) / fmin;
                    } else {
                        o->faccld1d[lev - 1] = (cldfrac[lev] - cldfrac[lev - 1]) / (cldfrac[lev] - fmin);
                        o->faccld2d[lev - 1] = 0.0;
                    }
                }
                if (o->faccld1d[lev - 1] > 0.0 || o->faccld2d[lev - 1] > 0.0) {
                    rat1 = 0.0;
                    rat2 = 1.0;
                } else {
                    rat1 = 0.0;
                    rat2 = 0.0;
                }
            }
            o->faccmb1d[lev - 1] = o->facclr1d[lev - 1] * o->faccld2d[lev] * cldfrac[lev + 1];
            o->faccmb2d[lev - 1] = o->faccld1d[lev - 1] * o->facclr2d[lev] * (1.0 - cldfrac[lev + 1]);
        } else {
            o->istcldd[lev - 1] = 1;
        }
    }
}

/* One cloudy layer's gas and gas+cloud terms (the three branches shared by
 * rrtmg_lw_rtrn.f90:349-418 and rrtmg_lw_rtrnmr.f90:497-545). */
typedef struct {
    double atrans, atot, gassrc, bbd, bbdtot, bbugas, bbutot;
} cldlay_t;

static void cloudy_layer(double odepth, double odcld, double plfrac, double blay, double dplankup, double dplankdn,
                         cldlay_t *r)
{
    const double rec_6 = 0.166667;
    double odtot = odepth + odcld;
    if (odtot < 0.06) {
        r->atrans = odepth - 0.5 * odepth * odepth;
        double odepth_rec = rec_6 * odepth;
        r->gassrc = plfrac * (blay + dplankdn * odepth_rec) * r->atrans;
        r->atot = odtot - 0.5 * odtot * odtot;
        double odtot_rec = rec_6 * odtot;
        r->bbdtot = plfrac * (blay + dplankdn * odtot_rec);
        r->bbd = plfrac * (blay + dplankdn * odepth_rec);
        r->bbugas = plfrac * (blay + dplankup * odepth_rec);
        r->bbutot = plfrac * (blay + dplankup * odtot_rec);
    } else if (odepth <= 0.06) {
        r->atrans = odepth - 0.5 * odepth * odepth;
        double odepth_rec = rec_6 * odepth;
        r->gassrc = plfrac * (blay + dplankdn * odepth_rec) * r->atrans;
        odtot = odepth + odcld;
        double tblind = odtot / (L.bpade + odtot);
        int ittot = (int)(1.e4 * tblind + 0.5);
        double tfactot = L.tfn_tbl[ittot];
        r->bbdtot = plfrac * (blay + tfactot * dplankdn);
        r->bbd = plfrac * (blay + dplankdn * odepth_rec);
        r->atot = 1.0 - L.exp_tbl[ittot];
        r->bbugas = plfrac * (blay + dplankup * odepth_rec);
        r->bbutot = plfrac * (blay + tfactot * dplankup);
    } else {
        double tblind = odepth / (L.bpade + odepth);
        int itgas = (int)(1.e4 * tblind + 0.5);
        odepth = L.tau_tbl[itgas];
        r->atrans = 1.0 - L.exp_tbl[itgas];
        double tfacgas = L.tfn_tbl[itgas];
        r->gassrc = r->atrans * plfrac * (blay + tfacgas * dplankdn);
        odtot = odepth + odcld;
        tblind = odtot / (L.bpade + odtot);
        int ittot = (int)(1.e4 * tblind + 0.5);
        double tfactot = L.tfn_tbl[ittot];
        r->bbdtot = plfrac * (blay + tfactot * dplankdn);
        r->bbd = plfrac * (blay + tfacgas * dplankdn);
        r->atot = 1.0 - L.exp_tbl[ittot];
        r->bbugas = plfrac * (blay + tfacgas * dplankup);
        r->bbutot = plfrac * (blay + tfactot * dplankup);
    }
}

/* rtrn (icld = 1, random overlap; rrtmg_lw_rtrn.f90:279-559) and rtrnmr
 * (icld = 2, 3, maximum-random overlap; rrtmg_lw_rtrnmr.f90:283-777) with
 * iout = 0, idrv = 0; taut = taug + taua(ngb) (iaer = 10,
 * rrtmg_lw_rad.nomcica.f90:512-518).  Total-sky and clear-sky fluxes and
 * heating rates. */
static void rtrn_cloudy(const lwcol_t *c, int icld, int ncbands, const double *cldfrac, double (*taucloud)[17],
                        double (*taua)[17], double *totuflux, double *totdflux, double *htr, double *totuclfl,
                        double *totdclfl, double *htrc)
{
    const int nl = c->nlayers;
    const double wtdiff = 0.5, rec_6 = 0.166667;
    const double pi = 2.0 * asin(1.0);
    const double fluxfac = pi * 2.e4;
    double secdiff[17];
    static __thread double odcld[MAXLAY + 2][17], efclfrac[MAXLAY + 2][17];
    double urad[MAXLAY + 2], drad[MAXLAY + 2], clrurad[MAXLAY + 2], clrdrad[MAXLAY + 2];
    double atrans[MAXLAY + 2], atot[MAXLAY + 2], bbugas[MAXLAY + 2], bbutot[MAXLAY + 2];
    double fnet[MAXLAY + 2], fnetc[MAXLAY + 2], cf[MAXLAY + 2];
    int icldlyr[MAXLAY + 2];
    static __thread overlap_t ov;
    /* the cloud band of each spectral band for cldprop's ncbands
     * (rrtmg_lw_rtrnmr.f90:487-493, ipat of :246-248);
     * the cloud depth takes the cloud band's diffusivity (:315-317) */
    static const int ipat[3][17] = {{0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
                                    {0, 1, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5},
                                    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16}};
    const int *cb = ipat[ncbands == 1 ? 0 : ncbands == 5 ? 1 : 2];
    for (int ibnd = 1; ibnd <= 16; ++ibnd) {
        if (ibnd == 1 || ibnd == 4 || ibnd >= 10) {
            secdiff[ibnd] = 1.66;
        } else {
            secdiff[ibnd] = L.a0[ibnd - 1] + L.a1[ibnd - 1] * exp(L.a2[ibnd - 1] * c->pwvcm);
            if (secdiff[ibnd] > 1.80) secdiff[ibnd] = 1.80;
            if (secdiff[ibnd] < 1.50) secdiff[ibnd] = 1.50;
        }
    }
    for (int lev = 0; lev <= nl; ++lev) {
        urad[lev] = drad[lev] = clrurad[lev] = clrdrad[lev] = 0.0;
        totuflux[lev] = totdflux[lev] = totuclfl[lev] = totdclfl[lev] = 0.0;
    }
    cf[0] = 0.0;
    cf[nl + 1] = 0.0;
    for (int lay = 1; lay <= nl; ++lay) {
        cf[lay] = cldfrac[lay];
        icldlyr[lay] = cf[lay] >= 1.e-6 ? 1 : 0;
        for (int ib = 1; ib <= 16; ++ib) {
            if (icldlyr[lay]) {
                odcld[lay][ib] = secdiff[ib] * taucloud[lay][ib];
                double transcld = exp(-odcld[lay][ib]);
                efclfrac[lay][ib] = (1.0 - transcld) * cf[lay];
            } else {
                odcld[lay][ib] = 0.0;
                efclfrac[lay][ib] = 0.0;
            }
        }
    }
    if (icld != 1) overlap_mr(nl, cf, icldlyr, &ov);
    int igc = 1;
    for (int iband = 1; iband <= 16; ++iband) {
        for (; igc <= L.ngs[iband - 1]; ++igc) {
            double radld = 0.0, radclrd = 0.0;
            double cldradd = 0.0, clrradd = 0.0, rad = 0.0, oldcld, oldclr;
            int iclddn = 0;
            for (int lev = nl; lev >= 1; --lev) {
                double plfrac = c->fracs[lev][igc];
                double blay = c->planklay[lev][iband];
                double dplankup = c->planklev[lev][iband] - blay;
                double dplankdn = c->planklev[lev - 1][iband] - blay;
                double taut = c->taug[lev][igc] + (taua ? taua[lev][iband] : 0.0);
                double odepth = secdiff[iband] * taut;
                if (odepth < 0.0) odepth = 0.0;
                double bbd;
                if (icldlyr[lev] == 1) {
                    iclddn = 1;
                    cldlay_t r;
                    cloudy_layer(odepth, odcld[lev][cb[iband]], plfrac, blay, dplankup, dplankdn, &r);
                    atrans[lev] = r.atrans;
                    atot[lev] = r.atot;
                    bbugas[lev] = r.bbugas;
                    bbutot[lev] = r.bbutot;
                    bbd = r.bbd;
                    if (icld == 1) {
                        radld = radld - radld * (r.atrans + efclfrac[lev][cb[iband]] * (1.0 - r.atrans)) + r.gassrc +
                                cf[lev] * (r.bbdtot * r.atot - r.gassrc);
                    } else {
                        if (ov.istcldd[lev] == 1) {
                            cldradd = cf[lev] * radld;
                            clrradd = radld - cldradd;
                            rad = 0.0;
                        }
                        double ttot = 1.0 - r.atot;
                        double cldsrc = r.bbdtot * r.atot;
                        cldradd = cldradd * ttot + cf[lev] * cldsrc;
                        clrradd = clrradd * (1.0 - r.atrans) + (1.0 - cf[lev]) * r.gassrc;
                        radld = cldradd + clrradd;
                        double radmod = rad * (ov.facclr1d[lev - 1] * (1.0 - r.atrans) + ov.faccld1d[lev - 1] * ttot) -
                                        ov.faccmb1d[lev - 1] * r.gassrc + ov.faccmb2d[lev - 1] * cldsrc;
                        oldcld = cldradd - radmod;
                        oldclr = clrradd + radmod;
                        rad = -radmod + ov.facclr2d[lev - 1] * oldclr - ov.faccld2d[lev - 1] * oldcld;
                        cldradd = cldradd + rad;
                        clrradd = clrradd - rad;
                    }
                    drad[lev - 1] = drad[lev - 1] + radld;
                } else {
                    if (odepth <= 0.06) {
                        atrans[lev] = odepth - 0.5 * odepth * odepth;
                        odepth = rec_6 * odepth;
                        bbd = plfrac * (blay + dplankdn * odepth);
                        bbugas[lev] = plfrac * (blay + dplankup * odepth);
                    } else {
                        double tblind = odepth / (L.bpade + odepth);
                        int itr = (int)(1.e4 * tblind + 0.5);
                        double transc = L.exp_tbl[itr];
                        atrans[lev] = 1.0 - transc;
                        double tausfac = L.tfn_tbl[itr];
                        bbd = plfrac * (blay + tausfac * dplankdn);
                        bbugas[lev] = plfrac * (blay + tausfac * dplankup);
                    }
                    radld = radld + (bbd - radld) * atrans[lev];
                    drad[lev - 1] = drad[lev - 1] + radld;
                }
                if (iclddn == 1) {
                    radclrd = radclrd + (bbd - radclrd) * atrans[lev];
                    clrdrad[lev - 1] = clrdrad[lev - 1] + radclrd;
                } else {
                    radclrd = radld;
                    clrdrad[lev - 1] = drad[lev - 1];
                }
            }
            double rad0 = c->fracs[1][igc] * c->plankbnd[iband];
            double reflect = 1.0 - c->semiss[iband];
            double radlu = rad0 + reflect * radld;
            double radclru = rad0 + reflect * radclrd;
            double cldradu = 0.0, clrradu = 0.0;
            urad[0] = urad[0] + radlu;
            clrurad[0] = clrurad[0] + radclru;
            for (int lev = 1; lev <= nl; ++lev) {
                if (icldlyr[lev] == 1) {
                    double gassrc = bbugas[lev] * atrans[lev];
                    if (icld == 1) {
                        radlu = radlu - radlu * (atrans[lev] + efclfrac[lev][cb[iband]] * (1.0 - atrans[lev])) + gassrc +
                                cf[lev] * (bbutot[lev] * atot[lev] - gassrc);
                    } else {
                        if (ov.istcld[lev] == 1) {
                            cldradu = cf[lev] * radlu;
                            clrradu = radlu - cldradu;
                            rad = 0.0;
                        }
                        double ttot = 1.0 - atot[lev];
                        double cldsrc = bbutot[lev] * atot[lev];
                        cldradu = cldradu * ttot + cf[lev] * cldsrc;
                        clrradu = clrradu * (1.0 - atrans[lev]) + (1.0 - cf[lev]) * gassrc;
                        radlu = cldradu + clrradu;
                        double radmod = rad * (ov.facclr1[lev + 1] * (1.0 - atrans[lev]) + ov.faccld1[lev + 1] * ttot) -
                                        ov.faccmb1[lev + 1] * gassrc + ov.faccmb2[lev + 1] * cldsrc;
                        oldcld = cldradu - radmod;
                        oldclr = clrradu + radmod;
                        rad = -radmod + ov.facclr2[lev + 1] * oldclr - ov.faccld2[lev + 1] * oldcld;
                        cldradu = cldradu + rad;
                        clrradu = clrradu - rad;
                    }
                } else {
                    radlu = radlu + (bbugas[lev] - radlu) * atrans[lev];
                }
                urad[lev] = urad[lev] + radlu;
                if (iclddn == 1) {
                    radclru = radclru + (bbugas[lev] - radclru) * atrans[lev];
                    clrurad[lev] = clrurad[lev] + radclru;
                } else {
                    radclru = radlu;
                    clrurad[lev] = urad[lev];
                }
            }
        }
        for (int lev = nl; lev >= 0; --lev) {
            double dw = L.delwave[iband - 1];
            totuflux[lev] = totuflux[lev] + urad[lev] * wtdiff * dw;
            totdflux[lev] = totdflux[lev] + drad[lev] * wtdiff * dw;
            totuclfl[lev] = totuclfl[lev] + clrurad[lev] * wtdiff * dw;
            totdclfl[lev] = totdclfl[lev] + clrdrad[lev] * wtdiff * dw;
            urad[lev] = drad[lev] = clrurad[lev] = clrdrad[lev] = 0.0;
        }
    }
    for (int lev = 0; lev <= nl; ++lev) {
        totuflux[lev] = totuflux[lev] * fluxfac;
        totdflux[lev] = totdflux[lev] * fluxfac;
        fnet[lev] = totuflux[lev] - totdflux[lev];
        totuclfl[lev] = totuclfl[lev] * fluxfac;
        totdclfl[lev] = totdclfl[lev] * fluxfac;
        fnetc[lev] = totuclfl[lev] - totdclfl[lev];
    }
    for (int lev = 1; lev <= nl; ++lev) {
        int l = lev - 1;
        htr[l] = L.heatfac * (fnet[l] - fnet[lev]) / (c->pz[l] - c->pz[lev]);
        htrc[l] = L.heatfac * (fnetc[l] - fnetc[lev]) / (c->pz[l] - c->pz[lev]);
    }
    htr[nl] = 0.0;
    htrc[nl] = 0.0;
}

/* rrtmg_lw with clouds (rrtmg_lw_rad.nomcica.f90:80-569): icld 0..3 (out
 * of range -> 2, :437; 0 = clear sky, cloud arrays unread), inflglw = 0,
 * idrv = 0.  cldfr, cicewp, cliqwp
 * (ncol, nlay); taucld (16, ncol, nlay), band fastest; tauaer (ncol, nlay,
 * 16) or NULL; cicewp/cliqwp NULL = 0. */
static int lw_cloud_driver(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                          const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                          const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                          const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                          const double *emis, const double *cldfr, const double *taucld, const double *cicewp,
                          const double *cliqwp, const double *reice, const double *reliq, const double *tauaer, double *uflx, double *dflx, double *hr,
                          double *uflxc, double *dflxc, double *hrc)
{
    if (!L.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    if (icld < 0 || icld > 3) icld = 2;
    /* icld = 0: inatm leaves cldfrac = 0 (:736, :868), every layer clear */
    const int clouds = icld != 0;
    int err = 0;
#pragma omp parallel
    {
        lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
        double (*tauc)[17] = malloc(sizeof(double[MAXLAY + 2][17]));
        double (*taucloud)[17] = malloc(sizeof(double[MAXLAY + 2][17]));
        double (*taua)[17] = malloc(sizeof(double[MAXLAY + 2][17]));
        double cf[MAXLAY + 2], cwp[MAXLAY + 2], ciw[MAXLAY + 2], clw[MAXLAY + 2], rei[MAXLAY + 2], rel[MAXLAY + 2];
        double tu[MAXLAY + 2], td[MAXLAY + 2], tuc[MAXLAY + 2], tdc[MAXLAY + 2], htr[MAXLAY + 2],
            htrc[MAXLAY + 2];
        if (!c || !tauc || !taucloud || !taua) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (err) continue;
            inatm(c, iplon, ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                  o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis);
            for (int l = 1; l <= nlay; ++l) {
                long ix = iplon + (long)ncol * (l - 1);
                cf[l] = clouds ? cldfr[ix] : 0.0;
                ciw[l] = clouds && cicewp ? cicewp[ix] : 0.0;
                clw[l] = clouds && cliqwp ? cliqwp[ix] : 0.0;
                rei[l] = clouds && reice ? reice[ix] : 0.0;
                rel[l] = clouds && reliq ? reliq[ix] : 0.0;
                cwp[l] = clouds ? (cicewp ? cicewp[ix] : 0.0) + (cliqwp ? cliqwp[ix] : 0.0) : 0.0;
                for (int ib = 1; ib <= 16; ++ib) {
                    tauc[l][ib] = clouds && taucld ? taucld[(ib - 1) + 16 * ix] : 0.0;
                    taua[l][ib] = tauaer ? tauaer[iplon + (long)ncol * ((l - 1) + (long)nlay * (ib - 1))] : 0.0;
                }
            }
            int ncbands = 16;
            if (inflag == 0 || !clouds) {
                cldprop0(nlay, cf, tauc, cwp, taucloud);
            } else if (cldprop_wp(nlay, inflag, iceflag, liqflag, cf, tauc, ciw, clw, rei, rel, taucloud,
                                  &ncbands) != 0) {
#pragma omp atomic write
                err = -8;
                continue;
            }
            setcoef(c);
            taumol(c);
            rtrn_cloudy(c, clouds ? icld : 2, ncbands, cf, taucloud, tauaer ? taua : NULL, tu, td, htr, tuc, tdc, htrc);
            for (int k = 0; k <= nlay; ++k) {
                uflx[iplon + (long)ncol * k] = tu[k];
                dflx[iplon + (long)ncol * k] = td[k];
                uflxc[iplon + (long)ncol * k] = tuc[k];
                dflxc[iplon + (long)ncol * k] = tdc[k];
            }
            for (int k = 0; k < nlay; ++k) {
                hr[iplon + (long)ncol * k] = htr[k];
                hrc[iplon + (long)ncol * k] = htrc[k];
            }
        }
        free(c);
        free(tauc);
        free(taucloud);
        free(taua);
    }
    return err;
}

int oracle_rrtmg_lw_cloud(int ncol, int nlay, int icld, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                          const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                          const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                          const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                          const double *emis, const double *cldfr, const double *taucld, const double *cicewp,
                          const double *cliqwp, const double *tauaer, double *uflx, double *dflx, double *hr,
                          double *uflxc, double *dflxc, double *hrc)
{
    return lw_cloud_driver(ncol, nlay, icld, 0, 0, 0, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                           n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, cldfr, taucld, cicewp, cliqwp,
                           NULL, NULL, tauaer, uflx, dflx, hr, uflxc, dflxc, hrc);
}

int oracle_rrtmg_lw_cloud_wp(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play,
                             const double *plev, const double *tlay, const double *tlev, const double *tsfc,
                             const double *h2ovmr, const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                             const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                             const double *emis, const double *cldfr, const double *taucld, const double *cicewp,
                             const double *cliqwp, const double *reice, const double *reliq, const double *tauaer,
                             double *uflx, double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc)
{
    if (inflag != 0 && !LC.ready) return -1;
    return lw_cloud_driver(ncol, nlay, icld, inflag, iceflag, liqflag, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, cldfr, taucld,
                           cicewp, cliqwp, reice, reliq, tauaer, uflx, dflx, hr, uflxc, dflxc, hrc);
}

/* ---- McICA (rrtmg_lw_rad.f90 with mcica_subcol_gen_lw.f90) ----------------- */

/* One step of the KISS generator (mcica_subcol_gen_lw.f90:530-562): a
 * congruential, a 3-shift register and two multiply-with-carry generators
 * in 32-bit wrap-around arithmetic (the reference's default integers),
 * their sum read as a signed integer and mapped to (0, 1). */
static double kiss_step(uint32_t s[4])
{
    s[0] = 69069u * s[0] + 1327217885u;
    uint32_t k = s[1];
    k ^= k << 13;
    k ^= k >> 17;
    k ^= k << 5;
    s[1] = k;
    s[2] = 18000u * (s[2] & 65535u) + (s[2] >> 16);
    s[3] = 30903u * (s[3] & 65535u) + (s[3] >> 16);
    uint32_t sum = s[0] + s[1] + (s[2] << 16) + s[3];
    return (double)(int32_t)sum * 2.328306e-10 + 0.5;
}

/* The KISS generator's cloudy-subcolumn mask of column i (irng = 0;
 * generate_stochastic_clouds, mcica_subcol_gen_lw.f90:158-520, and its SW
 * twin generate_stochastic_clouds_sw, mcica_subcol_gen_sw.f90:163-536, which
 * differ only in nsub): seeds from the fractional Pa of the four lowest
 * layers, advanced permuteseed times; one random number per (subcolumn,
 * layer) for random (icld = 1) and maximum-random (2) overlap, the latter
 * then carried up through cloudy layers; one per subcolumn for maximum
 * overlap (3).  Subcolumn ig of layer l is cloudy where its number >= 1 -
 * cldfrac: mask[ig * MAXLAY + l] (0-based).  -6 when the second layer's
 * pressure is the higher (the reference stops). */
/* MT19937 (Matsumoto and Nishimura, ACM TOMACS 8, 1998), as
 * mcica_random_numbers.f90 states it: init_genrand's scalar seeding
 * (:169-185), the twist of all 624 words once they are used up (:134-152),
 * the tempering (:154-165).  out receives the first `count` tempered words
 * (genrand_int32). */
void oracle_mt_stream(uint32_t seed, long count, uint32_t *out)
{
    uint32_t mt[624];
    mt[0] = seed;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int idx = 624;
    for (long n = 0; n < count; ++n) {
        if (idx >= 624) {
            for (int k = 0; k < 624; ++k) {
                uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((mt[(k + 1) % 624] & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        out[n] = y;
    }
}

/* the Mersenne Twister stream of the current generator call (irng = 1), or
 * NULL for KISS; set around the column loop by the subcolumn drivers */
static const uint32_t *mcica_mt;

void oracle_mcica_set_stream(const uint32_t *mt) { mcica_mt = mt; }

int oracle_mcica_mask(int i, int ncol, int nlay, int icld, int permuteseed, int nsub, const double *play,
                      const double *cldfrac, unsigned char *mask)
{
    if (mcica_mt) {
        /* irng = 1 (mcica_subcol_gen_lw.f90:360-367, 385-392, 418-425): the
         * stream in the order subcolumn, column, layer (one per subcolumn and
         * column for icld 3), each word w as getRandomReal's w / (2^32 - 1) */
        double cldf[MAXLAY + 1];
        for (int l = 0; l < nlay; ++l) {
            double f = cldfrac[i + (long)ncol * l];
            cldf[l] = f < 1.e-20 ? 0.0 : f;
        }
        for (int ig = 0; ig < nsub; ++ig) {
            double cdf[MAXLAY + 1];
            for (int l = 0; l < nlay; ++l) {
                long k = icld == 3 ? (long)ig * ncol + i : ((long)ig * ncol + i) * nlay + l;
                cdf[l] = (double)mcica_mt[k] / 4294967295.0;
            }
            if (icld == 2)
                for (int l = 1; l < nlay; ++l) {
                    if (cdf[l - 1] > 1.0 - cldf[l - 1])
                        cdf[l] = cdf[l - 1];
                    else
                        cdf[l] = cdf[l] * (1.0 - cldf[l - 1]);
                }
            for (int l = 0; l < nlay; ++l) mask[ig * MAXLAY + l] = cdf[l] >= 1.0 - cldf[l];
        }
        return 0;
    }
    if (play[i] * 1.e2 < play[i + ncol] * 1.e2) return -6;
    uint32_t s[4];
    for (int k = 0; k < 4; ++k) {
        double p = play[i + (long)ncol * k] * 1.e2;
        s[k] = (uint32_t)(int32_t)((p - (double)(int32_t)p) * 1000000000.0);
    }
    for (int n = 0; n < permuteseed; ++n) (void)kiss_step(s);
    double cldf[MAXLAY + 1];
    for (int l = 0; l < nlay; ++l) {
        double f = cldfrac[i + (long)ncol * l];
        cldf[l] = f < 1.e-20 ? 0.0 : f;
    }
    for (int ig = 0; ig < nsub; ++ig) {
        double cdf[MAXLAY + 1];
        if (icld == 3) {
            double r = kiss_step(s);
            for (int l = 0; l < nlay; ++l) cdf[l] = r;
        } else {
            for (int l = 0; l < nlay; ++l) cdf[l] = kiss_step(s);
        }
        if (icld == 2)
            for (int l = 1; l < nlay; ++l) {
                if (cdf[l - 1] > 1.0 - cldf[l - 1])
                    cdf[l] = cdf[l - 1];
                else
                    cdf[l] = cdf[l] * (1.0 - cldf[l - 1]);
            }
        for (int l = 0; l < nlay; ++l) mask[ig * MAXLAY + l] = cdf[l] >= 1.0 - cldf[l];
    }
    return 0;
}

/* mcica_subcol_lw with the KISS generator (irng = 0;
 * mcica_subcol_gen_lw.f90:48-152): the mask of oracle_mcica_mask; a cloudy
 * subcolumn takes the band's in-cloud depth and the layer's water paths.
 * Arrays: cldfrac, ciwp, clwp (ncol, nlay); tauc (16, ncol, nlay); outputs
 * (140, ncol, nlay).  -6 for a column whose layer pressures do not fall from
 * the surface (the reference stops), -7 for nlay < 4. */
int oracle_lw_mcica_subcol(int ncol, int nlay, int icld, int permuteseed, const double *play,
                           const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                           double *cldfmcl, double *ciwpmcl, double *clwpmcl, double *taucmcl)
{
    if (!L.ready) return -1;
    if (nlay < (mcica_mt ? 2 : 4) || nlay > MAXLAY) return -7;
    if (icld < 1 || icld > 3) return -2;
    for (int i = 0; i < ncol && !mcica_mt; ++i)
        if (play[i] * 1.e2 < play[i + ncol] * 1.e2) return -6;
    int band_of[NGLW];
    for (int b = 0, ig = 0; b < 16; ++b)
        for (; ig < L.ngs[b]; ++ig) band_of[ig] = b;
    static __thread unsigned char mask[NGLW * MAXLAY];
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ncol; ++i) {
        oracle_mcica_mask(i, ncol, nlay, icld, permuteseed, NGLW, play, cldfrac, mask);
        for (int l = 0; l < nlay; ++l) {
            long ix = i + (long)ncol * l;
            for (int ig = 0; ig < NGLW; ++ig) {
                long o = ig + (long)NGLW * ix;
                int cloudy = mask[ig * MAXLAY + l];
                cldfmcl[o] = cloudy ? 1.0 : 0.0;
                clwpmcl[o] = cloudy ? clwp[ix] : 0.0;
                ciwpmcl[o] = cloudy ? ciwp[ix] : 0.0;
                taucmcl[o] = cloudy ? tauc[band_of[ig] + 16 * ix] : 0.0;
            }
        }
    }
    return 0;
}

/* oracle_lw_mcica_subcol with the Mersenne Twister (irng = 1), seeded with
 * permuteseed (mcica_subcol_gen_lw.f90:338-339); no pressure-order or
 * four-layer condition (those belong to the KISS seeds). */
int oracle_lw_mcica_subcol_mt(int ncol, int nlay, int icld, int permuteseed, const double *play,
                              const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                              double *cldfmcl, double *ciwpmcl, double *clwpmcl, double *taucmcl)
{
    if (icld < 1 || icld > 3) return -2;
    long n = (long)NGLW * ncol * (icld == 3 ? 1 : nlay);
    uint32_t *mt = malloc(sizeof(uint32_t) * (size_t)n);
    if (!mt) return -3;
    oracle_mt_stream((uint32_t)permuteseed, n, mt);
    oracle_mcica_set_stream(mt);
    int rc = oracle_lw_mcica_subcol(ncol, nlay, icld, permuteseed, play, cldfrac, ciwp, clwp, tauc, cldfmcl,
                                    ciwpmcl, clwpmcl, taucmcl);
    oracle_mcica_set_stream(NULL);
    free(mt);
    return rc;
}

/* rtrnmc (rrtmg_lw_rtrnmc.f90:225-579; iout = 0, idrv = 0): rtrn's random
 * overlap with a cloud fraction of 0 or 1 per (g-point, layer); a layer with
 * any cloudy g-point takes the cloudy branch for all of them. */
static void rtrnmc(const lwcol_t *c, double (*cldfmc)[NGLW + 1], double (*taucmc)[NGLW + 1],
                   double (*taua)[17], double *totuflux, double *totdflux, double *htr, double *totuclfl,
                   double *totdclfl, double *htrc)
{
    const int nl = c->nlayers;
    const double wtdiff = 0.5, rec_6 = 0.166667;
    const double pi = 2.0 * asin(1.0);
    const double fluxfac = pi * 2.e4;
    double secdiff[17];
    static __thread double efclfrac[MAXLAY + 2][NGLW + 1], odcld[MAXLAY + 2][NGLW + 1];
    double urad[MAXLAY + 2], drad[MAXLAY + 2], clrurad[MAXLAY + 2], clrdrad[MAXLAY + 2];
    double atrans[MAXLAY + 2], atot[MAXLAY + 2], bbugas[MAXLAY + 2], bbutot[MAXLAY + 2];
    double fnet[MAXLAY + 2], fnetc[MAXLAY + 2];
    int icldlyr[MAXLAY + 2], ngb[NGLW + 1];
    for (int b = 1, ig = 1; b <= 16; ++b)
        for (; ig <= L.ngs[b - 1]; ++ig) ngb[ig] = b;
    for (int ibnd = 1; ibnd <= 16; ++ibnd) {
        if (ibnd == 1 || ibnd == 4 || ibnd >= 10) {
            secdiff[ibnd] = 1.66;
        } else {
            secdiff[ibnd] = L.a0[ibnd - 1] + L.a1[ibnd - 1] * exp(L.a2[ibnd - 1] * c->pwvcm);
            if (secdiff[ibnd] > 1.80) secdiff[ibnd] = 1.80;
            if (secdiff[ibnd] < 1.50) secdiff[ibnd] = 1.50;
        }
    }
    for (int lev = 0; lev <= nl; ++lev) {
        urad[lev] = drad[lev] = clrurad[lev] = clrdrad[lev] = 0.0;
        totuflux[lev] = totdflux[lev] = totuclfl[lev] = totdclfl[lev] = 0.0;
    }
    for (int lay = 1; lay <= nl; ++lay) {
        icldlyr[lay] = 0;
        for (int ig = 1; ig <= NGLW; ++ig) {
            if (cldfmc[lay][ig] == 1.0) {
                odcld[lay][ig] = secdiff[ngb[ig]] * taucmc[lay][ig];
                double transcld = exp(-odcld[lay][ig]);
                efclfrac[lay][ig] = (1.0 - transcld) * cldfmc[lay][ig];
                icldlyr[lay] = 1;
            } else {
                odcld[lay][ig] = 0.0;
                efclfrac[lay][ig] = 0.0;
            }
        }
    }
    int igc = 1;
    for (int iband = 1; iband <= 16; ++iband) {
        for (; igc <= L.ngs[iband - 1]; ++igc) {
            double radld = 0.0, radclrd = 0.0;
            int iclddn = 0;
            for (int lev = nl; lev >= 1; --lev) {
                double plfrac = c->fracs[lev][igc];
                double blay = c->planklay[lev][iband];
                double dplankup = c->planklev[lev][iband] - blay;
                double dplankdn = c->planklev[lev - 1][iband] - blay;
                double taut = c->taug[lev][igc] + (taua ? taua[lev][iband] : 0.0);
                double odepth = secdiff[iband] * taut;
                if (odepth < 0.0) odepth = 0.0;
                double bbd;
                if (icldlyr[lev] == 1) {
                    iclddn = 1;
                    cldlay_t r;
                    cloudy_layer(odepth, odcld[lev][igc], plfrac, blay, dplankup, dplankdn, &r);
                    atrans[lev] = r.atrans;
                    atot[lev] = r.atot;
                    bbugas[lev] = r.bbugas;
                    bbutot[lev] = r.bbutot;
                    bbd = r.bbd;
                    radld = radld - radld * (r.atrans + efclfrac[lev][igc] * (1.0 - r.atrans)) + r.gassrc +
                            cldfmc[lev][igc] * (r.bbdtot * r.atot - r.gassrc);
                } else {
                    if (odepth <= 0.06) {
                        atrans[lev] = odepth - 0.5 * odepth * odepth;
                        odepth = rec_6 * odepth;
                        bbd = plfrac * (blay + dplankdn * odepth);
                        bbugas[lev] = plfrac * (blay + dplankup * odepth);
                    } else {
                        double tblind = odepth / (L.bpade + odepth);
                        int itr = (int)(1.e4 * tblind + 0.5);
                        atrans[lev] = 1.0 - L.exp_tbl[itr];
                        double tausfac = L.tfn_tbl[itr];
                        bbd = plfrac * (blay + tausfac * dplankdn);
                        bbugas[lev] = plfrac * (blay + tausfac * dplankup);
                    }
                    radld = radld + (bbd - radld) * atrans[lev];
                }
                drad[lev - 1] = drad[lev - 1] + radld;
                if (iclddn == 1) {
                    radclrd = radclrd + (bbd - radclrd) * atrans[lev];
                    clrdrad[lev - 1] = clrdrad[lev - 1] + radclrd;
                } else {
                    radclrd = radld;
                    clrdrad[lev - 1] = drad[lev - 1];
                }
            }
            double rad0 = c->fracs[1][igc] * c->plankbnd[iband];
            double reflect = 1.0 - c->semiss[iband];
            double radlu = rad0 + reflect * radld;
            double radclru = rad0 + reflect * radclrd;
            urad[0] = urad[0] + radlu;
            clrurad[0] = clrurad[0] + radclru;
            for (int lev = 1; lev <= nl; ++lev) {
                if (icldlyr[lev] == 1) {
                    double gassrc = bbugas[lev] * atrans[lev];
                    radlu = radlu - radlu * (atrans[lev] + efclfrac[lev][igc] * (1.0 - atrans[lev])) + gassrc +
                            cldfmc[lev][igc] * (bbutot[lev] * atot[lev] - gassrc);
                } else {
                    radlu = radlu + (bbugas[lev] - radlu) * atrans[lev];
                }
                urad[lev] = urad[lev] + radlu;
                if (iclddn == 1) {
                    radclru = radclru + (bbugas[lev] - radclru) * atrans[lev];
                    clrurad[lev] = clrurad[lev] + radclru;
                } else {
                    radclru = radlu;
                    clrurad[lev] = urad[lev];
                }
            }
        }
        for (int lev = nl; lev >= 0; --lev) {
            double dw = L.delwave[iband - 1];
            totuflux[lev] = totuflux[lev] + urad[lev] * wtdiff * dw;
            totdflux[lev] = totdflux[lev] + drad[lev] * wtdiff * dw;
            totuclfl[lev] = totuclfl[lev] + clrurad[lev] * wtdiff * dw;
            totdclfl[lev] = totdclfl[lev] + clrdrad[lev] * wtdiff * dw;
            urad[lev] = drad[lev] = clrurad[lev] = clrdrad[lev] = 0.0;
        }
    }
    for (int lev = 0; lev <= nl; ++lev) {
        totuflux[lev] = totuflux[lev] * fluxfac;
        totdflux[lev] = totdflux[lev] * fluxfac;
        fnet[lev] = totuflux[lev] - totdflux[lev];
        totuclfl[lev] = totuclfl[lev] * fluxfac;
        totdclfl[lev] = totdclfl[lev] * fluxfac;
        fnetc[lev] = totuclfl[lev] - totdclfl[lev];
    }
    for (int lev = 1; lev <= nl; ++lev) {
        int l = lev - 1;
        htr[l] = L.heatfac * (fnet[l] - fnet[lev]) / (c->pz[l] - c->pz[lev]);
        htrc[l] = L.heatfac * (fnetc[l] - fnetc[lev]) / (c->pz[l] - c->pz[lev]);
    }
    htr[nl] = 0.0;
    htrc[nl] = 0.0;
}

/* rrtmg_lw of the McICA driver (rrtmg_lw_rad.f90:80-572) on given
 * subcolumns, icld 1..3, inflglw = 0 (cldprmc leaves taucmc as given),
 * idrv = 0; h2ovmr and o3vmr are volume mixing ratios, as that driver's
 * inatm takes them (it lacks MiMA's conversions).  cldfmcl, taucmcl (140, ncol, nlay); tauaer (ncol, nlay, 16) or
 * NULL. */
static int lw_mcica_driver(int ncol, int nlay, int iceflag, int liqflag, const double *ciwpmcl,
                           const double *clwpmcl, const double *reicmcl, const double *relqmcl, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                          const double *cfc22vmr, const double *ccl4vmr, const double *emis,
                          const double *cldfmcl, const double *taucmcl, const double *tauaer, double *uflx,
                          double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc)
{
    if (!L.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    int err = 0;
#pragma omp parallel
    {
        lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
        double (*cf)[NGLW + 1] = malloc(sizeof(double[MAXLAY + 2][NGLW + 1]));
        double (*tc)[NGLW + 1] = malloc(sizeof(double[MAXLAY + 2][NGLW + 1]));
        double (*taua)[17] = malloc(sizeof(double[MAXLAY + 2][17]));
        double tu[MAXLAY + 2], td[MAXLAY + 2], tuc[MAXLAY + 2], tdc[MAXLAY + 2], htr[MAXLAY + 2],
            htrc[MAXLAY + 2];
        int bad = 0;
        if (!c || !cf || !tc || !taua) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (err) continue;
            inatm_u(c, iplon, ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                    o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, 1);
            for (int l = 1; l <= nlay; ++l) {
                long ix = iplon + (long)ncol * (l - 1);
                for (int ig = 1; ig <= NGLW; ++ig) {
                    cf[l][ig] = cldfmcl[(ig - 1) + (long)NGLW * ix];
                    tc[l][ig] = taucmcl[(ig - 1) + (long)NGLW * ix];
                }
                /* cldprmc with inflag 2 (rrtmg_lw_cldprmc.f90:150-230): per
                 * g-point, where the subcolumn passes the cloud test, its
                 * depth from its water paths and the layer's radii with the
                 * coefficients of the g-point's band */
                if (iceflag >= 0) {
                    static const int icb[16] = {1, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5};
                    for (int ig = 1; ig <= NGLW; ++ig) {
                        long o = (ig - 1) + (long)NGLW * ix;
                        double ci = ciwpmcl ? ciwpmcl[o] : 0.0, cw = clwpmcl ? clwpmcl[o] : 0.0;
                        if (!(cf[l][ig] >= 1.e-20 && (ci + cw >= 1.e-20 || tc[l][ig] >= 1.e-20))) continue;
                        int ib = 1;
                        while (ig > L.ngs[ib - 1]) ++ib;
                        double radice = reicmcl ? reicmcl[ix] : 0.0, aice = 0.0, aliq = 0.0;
                        if (ci == 0.0) {
                            aice = 0.0;
                        } else if (iceflag == 0) {
                            if (radice < 10.0) bad = 1;
                            aice = LC.absice0[0] + LC.absice0[1] / radice;
                        } else if (iceflag == 1) {
                            if (radice < 13.0 || radice > 130.0) bad = 1;
                            int b5 = icb[ib - 1];
                            aice = LC.absice1[b5 - 1][0] + LC.absice1[b5 - 1][1] / radice;
                        } else {
                            const int nr = iceflag == 2 ? 43 : 46;
                            if (radice < 5.0 || radice > (iceflag == 2 ? 131.0 : 140.0)) bad = 1;
                            double factor = (radice - 2.0) / 3.0;
                            int index = (int)factor;
                            if (index == nr) index = nr - 1;
                            double fint = factor - (double)index;
                            const double *a = iceflag == 2 ? LC.absice2[ib - 1] : LC.absice3[ib - 1];
                            aice = a[index - 1] + fint * (a[index] - (a[index - 1]));
                        }
                        if (cw == 0.0) {
                            aliq = 0.0;
                        } else if (liqflag == 0) {
                            aliq = LC.absliq0;
                        } else {
                            double radliq = relqmcl ? relqmcl[ix] : 0.0;
                            if (radliq < 2.5 || radliq > 60.0) bad = 1;
                            int index = (int)(radliq - 1.5);
                            if (index == 0) index = 1;
                            if (index == 58) index = 57;
                            double fint = radliq - 1.5 - (double)index;
                            const double *a = LC.absliq1[ib - 1];
                            aliq = a[index - 1] + fint * (a[index] - (a[index - 1]));
                        }
                        tc[l][ig] = ci * aice + cw * aliq;
                    }
                }
                for (int ib = 1; ib <= 16; ++ib)
                    taua[l][ib] = tauaer ? tauaer[iplon + (long)ncol * ((l - 1) + (long)nlay * (ib - 1))] : 0.0;
            }
            if (bad) {
#pragma omp atomic write
                err = -8;
                continue;
            }
            setcoef(c);
            taumol(c);
            rtrnmc(c, cf, tc, tauaer ? taua : NULL, tu, td, htr, tuc, tdc, htrc);
            for (int k = 0; k <= nlay; ++k) {
                uflx[iplon + (long)ncol * k] = tu[k];
                dflx[iplon + (long)ncol * k] = td[k];
                uflxc[iplon + (long)ncol * k] = tuc[k];
                dflxc[iplon + (long)ncol * k] = tdc[k];
            }
            for (int k = 0; k < nlay; ++k) {
                hr[iplon + (long)ncol * k] = htr[k];
                hrc[iplon + (long)ncol * k] = htrc[k];
            }
        }
        free(c);
        free(cf);
        free(tc);
        free(taua);
    }
    return err;
}

int oracle_rrtmg_lw_mcica(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                          const double *cfc22vmr, const double *ccl4vmr, const double *emis,
                          const double *cldfmcl, const double *taucmcl, const double *tauaer, double *uflx,
                          double *dflx, double *hr, double *uflxc, double *dflxc, double *hrc)
{
    return lw_mcica_driver(ncol, nlay, -1, 0, NULL, NULL, NULL, NULL, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, cldfmcl,
                           taucmcl, tauaer, uflx, dflx, hr, uflxc, dflxc, hrc);
}

int oracle_rrtmg_lw_mcica_wp(int ncol, int nlay, int iceflag, int liqflag, const double *play, const double *plev,
                             const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                             const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *cfc11vmr,
                             const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                             const double *emis, const double *cldfmcl, const double *taucmcl,
                             const double *ciwpmcl, const double *clwpmcl, const double *reicmcl,
                             const double *relqmcl, const double *tauaer, double *uflx, double *dflx, double *hr,
                             double *uflxc, double *dflxc, double *hrc)
{
    if (!LC.ready || iceflag < 0 || iceflag > 3) return -1;
    return lw_mcica_driver(ncol, nlay, iceflag, liqflag, ciwpmcl, clwpmcl, reicmcl, relqmcl, play, plev, tlay,
                           tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr,
                           ccl4vmr, emis, cldfmcl, taucmcl, tauaer, uflx, dflx, hr, uflxc, dflxc, hrc);
}

static void one_lw(lwcol_t *c, int nlay, const double *play, const double *plev, const double *tlay,
                   const double *tlev, double tsfc, const double *h2ovmr, const double *o3vmr,
                   const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                   const double *cfc11, const double *cfc12, const double *cfc22, const double *ccl4)
{
    double emis[16];
    for (int i = 0; i < 16; ++i) emis[i] = 1.0;
    double zeros[MAXLAY];
    memset(zeros, 0, sizeof zeros);
    inatm(c, 0, 1, nlay, play, plev, tlay, tlev, &tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
          cfc11 ? cfc11 : zeros, cfc12 ? cfc12 : zeros, cfc22 ? cfc22 : zeros, ccl4 ? ccl4 : zeros, emis);
    setcoef(c);
}

int oracle_lw_setcoef(int nlay, const double *play, const double *plev, const double *tlay,
                      const double *tlev, double tsfc, const double *h2ovmr, const double *o3vmr,
                      const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                      const double *o2vmr, int *jp, int *jt, int *jt1, double *fac00, double *fac01,
                      double *fac10, double *fac11, double *planklay, double *planklev, double *plankbnd)
{
    if (!L.ready || nlay > MAXLAY) return -1;
    lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
    one_lw(c, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, NULL, NULL,
           NULL, NULL);
    for (int k = 1; k <= nlay; ++k) {
        jp[k - 1] = c->jp[k];
        jt[k - 1] = c->jt[k];
        jt1[k - 1] = c->jt1[k];
        fac00[k - 1] = c->fac00[k];
        fac01[k - 1] = c->fac01[k];
        fac10[k - 1] = c->fac10[k];
        fac11[k - 1] = c->fac11[k];
        for (int b = 1; b <= 16; ++b) planklay[(b - 1) * nlay + (k - 1)] = c->planklay[k][b];
    }
    for (int k = 0; k <= nlay; ++k)
        for (int b = 1; b <= 16; ++b) planklev[(b - 1) * (nlay + 1) + k] = c->planklev[k][b];
    for (int b = 1; b <= 16; ++b) plankbnd[b - 1] = c->plankbnd[b];
    int lt = c->laytrop;
    free(c);
    return lt;
}

int oracle_lw_taumol(int nlay, const double *play, const double *plev, const double *tlay,
                     const double *tlev, double tsfc, const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                     const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                     const double *cfc22vmr, const double *ccl4vmr, double *taug, double *fracs)
{
    if (!L.ready || nlay > MAXLAY) return -1;
    lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
    one_lw(c, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr,
           cfc12vmr, cfc22vmr, ccl4vmr);
    taumol(c);
    for (int k = 1; k <= nlay; ++k)
        for (int ig = 1; ig <= NGLW; ++ig) {
            taug[(ig - 1) * nlay + (k - 1)] = c->taug[k][ig];
            fracs[(ig - 1) * nlay + (k - 1)] = c->fracs[k][ig];
        }
    int lt = c->laytrop;
    free(c);
    return lt;
}

/* inatm of one column (rrtmg_lw_rad.nomcica.f90:729-852): dry-air column
 * coldry (nlay) and the molecule columns wkl(1..7) (7 x nlay, molecules/cm2),
 * in molec/cm2, and the precipitable water pwvcm.  For the hydrostatic and
 * unit checks of the CPU tests. */
int oracle_lw_inatm(int nlay, const double *play, const double *plev, const double *tlay,
                    const double *tlev, double tsfc, const double *h2ovmr, const double *o3vmr,
                    const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                    const double *o2vmr, double *coldry, double *wkl, double *pwvcm)
{
    if (!L.ready || nlay > MAXLAY) return -1;
    lwcol_t *c = (lwcol_t *)malloc(sizeof(lwcol_t));
    if (!c) return -3;
    one_lw(c, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, NULL, NULL,
           NULL, NULL);
    for (int k = 1; k <= nlay; ++k) {
        coldry[k - 1] = c->coldry[k];
        for (int m = 1; m <= 7; ++m) wkl[(m - 1) * nlay + (k - 1)] = c->wkl[m][k];
    }
    *pwvcm = c->pwvcm;
    free(c);
    return 0;
}
