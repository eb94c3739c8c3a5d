/*
 * Oracle: RRTMG_SW clear-sky column model in FP64 (TEST INFRASTRUCTURE ONLY;
 * see oracle.h).  Restates, in order of evaluation:
 *   rrtmg_sw_init.f90:28-154 (init, exp_tbl, rwgt, cmbgb16s..29)
 *   rrtmg_sw_rad.nomcica.f90:500-729 (column loop), :761-1101 (inatm_sw)
 *   rrtmg_sw_setcoef.f90:30-286 (setcoef_sw)
 *   rrtmg_sw_taumol.f90:31-1538 (taumol16..29)
 *   rrtmg_sw_spcvrt.f90:34-626 (spcvrt_sw), rrtmg_sw_reftra.f90:29-305,
 *   rrtmg_sw_vrtqdr.f90:28-152.
 * Arrays local to a column use 1-based indices as in the Fortran.
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NBSW 14
#define NGSW 112
#define MAXLAY 256

typedef struct {
    int ng, nspa, nspb;
    double *absa, *absb; /* (nrow, ng) */
    int nrow_a, nrow_b;
    double *selfref, *forref; /* (10, ng), (nfor, ng) */
    int nfor;
    double *sfluxref; /* (ng, nsf) */
    int nsf;
    double rayl;           /* scalar Rayleigh */
    double *raylg;         /* rayl(ng) of bands 23, 25, 26, 27 */
    double *rayla, *raylb; /* band 24: rayla(ng, 9), raylb(ng) */
    double *abso3a, *abso3b, *absch4, *absco2, *absh2o;
} swband_t;

static struct {
    int ready;
    double heatfac, grav, avogad;
    double exp_tbl[ORACLE_NTBL + 1];
    double bpade;
    int ngc[NBSW], ngs[NBSW];
    double rwgt[NBSW * 16];
    double pref[59], preflog[59], tref[59];
    swband_t b[NBSW]; /* index band-16 */
} S;

#define B(jb) (&S.b[(jb)-16])
#define ABSA(bb, ind, ig) ((bb)->absa[(long)((ig)-1) * (bb)->nrow_a + ((ind)-1)])
#define ABSB(bb, ind, ig) ((bb)->absb[(long)((ig)-1) * (bb)->nrow_b + ((ind)-1)])
#define SELFREF(bb, i, ig) ((bb)->selfref[((ig)-1) * 10 + ((i)-1)])
#define FORREF(bb, i, ig) ((bb)->forref[((ig)-1) * (bb)->nfor + ((i)-1)])
#define SFLUXREF(bb, ig, j) ((bb)->sfluxref[((j)-1) * (bb)->ng + ((ig)-1)])
#define RAYLA(bb, ig, j) ((bb)->rayla[((j)-1) * (bb)->ng + ((ig)-1)])

static const double oneminus_sw = 1.0 - 1.e-6; /* rrtmg_sw_rad.nomcica.f90:445-447 */

/* ------------------------------------------------------------------------ */
/* init                                                                     */
/* ------------------------------------------------------------------------ */

static double *reduce_entry(const otbl_t *t, int jb, const char *nm, int gdim, int weighted,
                            const int *ngn_all, int *other_out)
{
    char key[64];
    snprintf(key, sizeof key, "sw/b%d/%s", jb, nm);
    int nd, ext[4], lb[4], isint;
    const double *x = (const double *)otbl_get(t, key, &nd, ext, lb, &isint);
    if (!x) return NULL;
    int ibnd = jb - 15;
    const int *first_ngn = ngn_all + (ibnd == 1 ? 0 : S.ngs[ibnd - 2]);
    long total = 1;
    for (int d = 0; d < nd; ++d) total *= ext[d];
    if (other_out) *other_out = (int)(total / 16);
    double *r = oracle_reduce(x, nd, ext, gdim, S.ngc[ibnd - 1], first_ngn,
                              S.rwgt + (ibnd - 1) * 16, weighted);
    char rkey[64];
    /* register under the reduced Fortran name */
    const char *rn = nm;
    char buf[32];
    size_t L = strlen(nm);
    if (L > 1 && nm[L - 1] == 'o') {
        memcpy(buf, nm, L - 1);
        buf[L - 1] = 0;
        if (!strcmp(buf, "ka")) rn = "ka";
        else if (!strcmp(buf, "kb")) rn = "kb";
        else rn = buf;
    }
    snprintf(rkey, sizeof rkey, "sw/b%d/%s", jb, rn);
    oracle_register(rkey, r, (total / 16) * S.ngc[ibnd - 1]);
    return r;
}

int oracle_sw_init(const char *path, double cpdair)
{
    otbl_t t;
    if (otbl_load(&t, path) != 0) return -1;
    /* swdatinit (rrtmg_sw_init.f90:157-241) */
    S.grav = 9.8066;
    S.avogad = 6.02214199e+23;
    S.heatfac = S.grav * 8.6400e4 / (cpdair * 1.e2);
    const int *ngc = otbl_i(&t, "sw/ngc", NULL), *ngs = otbl_i(&t, "sw/ngs", NULL);
    const int *ngn = otbl_i(&t, "sw/ngn", NULL), *ngm = otbl_i(&t, "sw/ngm", NULL);
    const int *nspa = otbl_i(&t, "sw/nspa", NULL), *nspb = otbl_i(&t, "sw/nspb", NULL);
    const double *wt = otbl_d(&t, "sw/wt", NULL);
    if (!ngc || !ngs || !ngn || !ngm || !wt || !nspa || !nspb) return -2;
    memcpy(S.ngc, ngc, sizeof S.ngc);
    memcpy(S.ngs, ngs, sizeof S.ngs);
    memcpy(S.pref, otbl_d(&t, "sw/pref", NULL), sizeof S.pref);
    memcpy(S.preflog, otbl_d(&t, "sw/preflog", NULL), sizeof S.preflog);
    memcpy(S.tref, otbl_d(&t, "sw/tref", NULL), sizeof S.tref);

    /* exp_tbl (rrtmg_sw_init.f90:94-104) */
    const double expeps = 1.e-20;
    S.exp_tbl[0] = 1.0;
    S.exp_tbl[ORACLE_NTBL] = expeps;
    S.bpade = 1.0 / 0.278;
    for (int itr = 1; itr <= ORACLE_NTBL - 1; ++itr) {
        double tfn = (double)itr / (double)ORACLE_NTBL;
        double tau = S.bpade * tfn / (1.0 - tfn);
        S.exp_tbl[itr] = exp(-tau);
        if (S.exp_tbl[itr] <= expeps) S.exp_tbl[itr] = expeps;
    }
    oracle_make_rwgt(NBSW, ngc, ngn, ngm, wt, S.rwgt);
    oracle_register("sw/rwgt", S.rwgt, NBSW * 16);

    for (int jb = 16; jb <= 29; ++jb) {
        swband_t *bb = B(jb);
        memset(bb, 0, sizeof *bb);
        bb->ng = ngc[jb - 16];
        bb->nspa = nspa[jb - 16];
        bb->nspb = nspb[jb - 16];
        int other;
        bb->absa = reduce_entry(&t, jb, "kao", 1, 1, ngn, &other);
        bb->nrow_a = bb->absa ? other : 0;
        bb->absb = reduce_entry(&t, jb, "kbo", 1, 1, ngn, &other);
        bb->nrow_b = bb->absb ? other : 0;
        bb->selfref = reduce_entry(&t, jb, "selfrefo", 1, 1, ngn, NULL);
        bb->forref = reduce_entry(&t, jb, "forrefo", 1, 1, ngn, &other);
        bb->nfor = bb->forref ? other : 0;
        bb->sfluxref = reduce_entry(&t, jb, "sfluxrefo", 0, 0, ngn, &other);
        bb->nsf = other;
        char key[64];
        snprintf(key, sizeof key, "sw/b%d/rayl", jb);
        const double *rs = otbl_d(&t, key, NULL);
        bb->rayl = rs ? rs[0] : 0.0;
        bb->raylg = reduce_entry(&t, jb, "raylo", 0, 1, ngn, NULL);
        bb->rayla = reduce_entry(&t, jb, "raylao", 0, 1, ngn, NULL);
        bb->raylb = reduce_entry(&t, jb, "raylbo", 0, 1, ngn, NULL);
        bb->abso3a = reduce_entry(&t, jb, "abso3ao", 0, 1, ngn, NULL);
        bb->abso3b = reduce_entry(&t, jb, "abso3bo", 0, 1, ngn, NULL);
        bb->absch4 = reduce_entry(&t, jb, "absch4o", 0, 1, ngn, NULL);
        bb->absco2 = reduce_entry(&t, jb, "absco2o", 0, 1, ngn, NULL);
        bb->absh2o = reduce_entry(&t, jb, "absh2oo", 0, 1, ngn, NULL);
    }
    otbl_free(&t);
    S.ready = 1;
    return 0;
}

double oracle_sw_exp_tbl(int i) { return S.exp_tbl[i]; }

/* ------------------------------------------------------------------------ */
/* per-column state                                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nlayers, laytrop, layswtch, laylow;
    double pavel[MAXLAY + 1], tavel[MAXLAY + 1], pz[MAXLAY + 1], tz[MAXLAY + 1], pdp[MAXLAY + 1];
    double coldry[MAXLAY + 1], wkl[8][MAXLAY + 1];
    int jp[MAXLAY + 2], jt[MAXLAY + 1], jt1[MAXLAY + 1];
    double co2mult[MAXLAY + 1], colch4[MAXLAY + 1], colco2[MAXLAY + 1], colh2o[MAXLAY + 1];
    double colmol[MAXLAY + 1], coln2o[MAXLAY + 1], colo2[MAXLAY + 1], colo3[MAXLAY + 1];
    double fac00[MAXLAY + 1], fac01[MAXLAY + 1], fac10[MAXLAY + 1], fac11[MAXLAY + 1];
    double selffac[MAXLAY + 1], selffrac[MAXLAY + 1], forfac[MAXLAY + 1], forfrac[MAXLAY + 1];
    int indself[MAXLAY + 1], indfor[MAXLAY + 1];
    double sfluxzen[NGSW + 1];
    double taug[MAXLAY + 1][NGSW + 1], taur[MAXLAY + 1][NGSW + 1];
} swcol_t;

/* inatm_sw (rrtmg_sw_rad.nomcica.f90:761-1101), the clear-sky / iaer=0 part.
 * vmr = 0: h2o as specific humidity and O3 as a mass mixing ratio, MiMA's
 * usage (:1004, 1007); vmr = 1: both as volume mixing ratios, as the McICA
 * driver's inatm takes them (rrtmg_sw_rad.f90:1015, 1021) */
static void inatm_sw(swcol_t *c, int iplon, int ncol, int nlay, const double *play,
                     const double *plev, const double *tlay, const double *tlev,
                     const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                     const double *ch4vmr, const double *n2ovmr, const double *o2vmr, int vmr)
{
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
#define A2(a, i, k) a[(i) + (long)ncol * ((k)-1)]
    c->nlayers = nlay;
    for (int m = 0; m < 8; ++m)
        for (int l = 0; l <= nlay; ++l) c->wkl[m][l] = 0.0;
    c->pz[0] = A2(plev, iplon, 1);
    c->tz[0] = A2(tlev, iplon, 1);
    for (int l = 1; l <= nlay; ++l) {
        c->pavel[l] = A2(play, iplon, l);
        c->tavel[l] = A2(tlay, iplon, l);
        c->pz[l] = A2(plev, iplon, l + 1);
        c->tz[l] = A2(tlev, iplon, l + 1);
        c->pdp[l] = c->pz[l - 1] - c->pz[l];
        c->wkl[1][l] = vmr ? A2(h2ovmr, iplon, l) : (A2(h2ovmr, iplon, l) / (1.0 - A2(h2ovmr, iplon, l))) * amdw;
        c->wkl[2][l] = A2(co2vmr, iplon, l);
        c->wkl[3][l] = vmr ? A2(o3vmr, iplon, l) : A2(o3vmr, iplon, l) * amdo;
        c->wkl[4][l] = A2(n2ovmr, iplon, l);
        c->wkl[6][l] = A2(ch4vmr, iplon, l);
        c->wkl[7][l] = A2(o2vmr, iplon, l);
        double amm = (1.0 - c->wkl[1][l]) * amd + c->wkl[1][l] * amw;
        c->coldry[l] = (c->pz[l - 1] - c->pz[l]) * 1.e3 * S.avogad /
                       (1.e2 * S.grav * amm * (1.0 + c->wkl[1][l]));
    }
    for (int l = 1; l <= nlay; ++l)
        for (int imol = 1; imol <= 7; ++imol) c->wkl[imol][l] = c->coldry[l] * c->wkl[imol][l];
#undef A2
}

/* setcoef_sw (rrtmg_sw_setcoef.f90:129-284) */
static void setcoef_sw(swcol_t *c)
{
    const double stpfac = 296.0 / 1013.0;
    c->laytrop = 0;
    c->layswtch = 0;
    c->laylow = 0;
    for (int lay = 1; lay <= c->nlayers; ++lay) {
        double plog = log(c->pavel[lay]);
        c->jp[lay] = (int)(36.0 - 5 * (plog + 0.04));
        if (c->jp[lay] < 1) c->jp[lay] = 1;
        else if (c->jp[lay] > 58) c->jp[lay] = 58;
        int jp1 = c->jp[lay] + 1;
        double fp = 5.0 * (S.preflog[c->jp[lay] - 1] - plog);
        c->jt[lay] = (int)(3.0 + (c->tavel[lay] - S.tref[c->jp[lay] - 1]) / 15.0);
        if (c->jt[lay] < 1) c->jt[lay] = 1;
        else if (c->jt[lay] > 4) c->jt[lay] = 4;
        double ft = ((c->tavel[lay] - S.tref[c->jp[lay] - 1]) / 15.0) - (double)(c->jt[lay] - 3);
        c->jt1[lay] = (int)(3.0 + (c->tavel[lay] - S.tref[jp1 - 1]) / 15.0);
        if (c->jt1[lay] < 1) c->jt1[lay] = 1;
        else if (c->jt1[lay] > 4) c->jt1[lay] = 4;
        double ft1 = ((c->tavel[lay] - S.tref[jp1 - 1]) / 15.0) - (double)(c->jt1[lay] - 3);
        double water = c->wkl[1][lay] / c->coldry[lay];
        double scalefac = c->pavel[lay] * stpfac / c->tavel[lay];
        if (plog > 4.56) {
            c->laytrop += 1;
            if (plog >= 6.62) c->laylow += 1;
            c->forfac[lay] = scalefac / (1. + water);
            double factor = (332.0 - c->tavel[lay]) / 36.0;
            int ifac = (int)factor;
            c->indfor[lay] = ifac < 1 ? 1 : (ifac > 2 ? 2 : ifac);
            c->forfrac[lay] = factor - (double)c->indfor[lay];
            c->selffac[lay] = water * c->forfac[lay];
            factor = (c->tavel[lay] - 188.0) / 7.2;
            int is = (int)factor - 7;
            c->indself[lay] = is < 1 ? 1 : (is > 9 ? 9 : is);
            c->selffrac[lay] = factor - (double)(c->indself[lay] + 7);
        } else {
            c->forfac[lay] = scalefac / (1. + water);
            double factor = (c->tavel[lay] - 188.0) / 36.0;
            c->indfor[lay] = 3;
            c->forfrac[lay] = factor - 1.0;
            c->selffac[lay] = 0.0;
            c->selffrac[lay] = 0.0;
            c->indself[lay] = 0;
        }
        c->colh2o[lay] = 1.e-20 * c->wkl[1][lay];
        c->colco2[lay] = 1.e-20 * c->wkl[2][lay];
        c->colo3[lay] = 1.e-20 * c->wkl[3][lay];
        c->coln2o[lay] = 1.e-20 * c->wkl[4][lay];
        c->colch4[lay] = 1.e-20 * c->wkl[6][lay];
        c->colo2[lay] = 1.e-20 * c->wkl[7][lay];
        c->colmol[lay] = 1.e-20 * c->coldry[lay] + c->colh2o[lay];
        if (c->colco2[lay] == 0.0) c->colco2[lay] = 1.e-32 * c->coldry[lay];
        if (c->coln2o[lay] == 0.0) c->coln2o[lay] = 1.e-32 * c->coldry[lay];
        if (c->colch4[lay] == 0.0) c->colch4[lay] = 1.e-32 * c->coldry[lay];
        if (c->colo2[lay] == 0.0) c->colo2[lay] = 1.e-32 * c->coldry[lay];
        double co2reg = 3.55e-24 * c->coldry[lay];
        c->co2mult[lay] = (c->colco2[lay] - co2reg) * 272.63 * exp(-1919.4 / c->tavel[lay]) /
                          (8.7604e-4 * c->tavel[lay]);
        double compfp = 1.0 - fp;
        c->fac10[lay] = compfp * ft;
        c->fac00[lay] = compfp * (1.0 - ft);
        c->fac11[lay] = fp * ft1;
        c->fac01[lay] = fp * (1.0 - ft1);
    }
}

/* ------------------------------------------------------------------------ */
/* taumol_sw (rrtmg_sw_taumol.f90)                                          */
/* ------------------------------------------------------------------------ */

/* helpers for the recurring 8-point (eta, T, p) interpolation pieces */
typedef struct {
    double fac000, fac001, fac010, fac011, fac100, fac101, fac110, fac111;
} fac8_t;

static void fac8(const swcol_t *c, int lay, double fs, fac8_t *f)
{
    f->fac000 = (1.0 - fs) * c->fac00[lay];
    f->fac010 = (1.0 - fs) * c->fac10[lay];
    f->fac100 = fs * c->fac00[lay];
    f->fac110 = fs * c->fac10[lay];
    f->fac001 = (1.0 - fs) * c->fac01[lay];
    f->fac011 = (1.0 - fs) * c->fac11[lay];
    f->fac101 = fs * c->fac01[lay];
    f->fac111 = fs * c->fac11[lay];
}

#define MAJ8(ARR, f, ind0, ind1, ig, S1, S2)                                                      \
    ((f).fac000 * ARR(bb, ind0, ig) + (f).fac100 * ARR(bb, (ind0) + 1, ig) +                      \
     (f).fac010 * ARR(bb, (ind0) + (S1), ig) + (f).fac110 * ARR(bb, (ind0) + (S2), ig) +          \
     (f).fac001 * ARR(bb, ind1, ig) + (f).fac101 * ARR(bb, (ind1) + 1, ig) +                      \
     (f).fac011 * ARR(bb, (ind1) + (S1), ig) + (f).fac111 * ARR(bb, (ind1) + (S2), ig))
#define MAJ4(ARR, c, lay, ind0, ind1, ig)                                                          \
    ((c)->fac00[lay] * ARR(bb, ind0, ig) + (c)->fac10[lay] * ARR(bb, (ind0) + 1, ig) +            \
     (c)->fac01[lay] * ARR(bb, ind1, ig) + (c)->fac11[lay] * ARR(bb, (ind1) + 1, ig))
#define SELF(c, lay, inds, ig)                                                                     \
    ((c)->selffac[lay] * (SELFREF(bb, inds, ig) +                                                  \
                          (c)->selffrac[lay] * (SELFREF(bb, (inds) + 1, ig) - SELFREF(bb, inds, ig))))
#define FOR(c, lay, indf, ig)                                                                      \
    ((c)->forfac[lay] *                                                                            \
     (FORREF(bb, indf, ig) + (c)->forfrac[lay] * (FORREF(bb, (indf) + 1, ig) - FORREF(bb, indf, ig))))

static void eta(double cola, double colb, double strrat, double mult, double *speccomb, int *js,
                double *fs)
{
    *speccomb = cola + strrat * colb;
    double specparm = cola / *speccomb;
    if (specparm >= oneminus_sw) specparm = oneminus_sw;
    double specmult = mult * (specparm);
    *js = 1 + (int)specmult;
    *fs = fmod(specmult, 1.0);
}

static void taumol_sw(swcol_t *c)
{
    const int nlayers = c->nlayers, laytrop = c->laytrop;
    const int *jp = c->jp, *jt = c->jt, *jt1 = c->jt1;
    int lay, ig, ind0, ind1, inds, indf, js, laysolfr, layreffr;
    double speccomb, fs, tauray;
    fac8_t f;
    const swband_t *bb;
    /* jp(nlayers+1) is read by the lower-atmosphere reference-layer search of
     * bands 18-25 only when laytrop == nlayers; keep it defined. */
    c->jp[nlayers + 1] = c->jp[nlayers];

    /* band 16: rrtmg_sw_taumol.f90:243-339 */
    bb = B(16);
    {
        const double strrat1 = 252.131;
        layreffr = 18;
        for (lay = 1; lay <= laytrop; ++lay) {
            eta(c->colh2o[lay], c->colch4[lay], strrat1, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                   c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                c->taur[lay][ig] = tauray;
            }
        }
        laysolfr = nlayers;
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            if (jp[lay - 1] < layreffr && jp[lay] >= layreffr) laysolfr = lay;
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ig] = c->colch4[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig);
                if (lay == laysolfr) c->sfluxzen[ig] = bb->sfluxref[ig - 1];
                c->taur[lay][ig] = tauray;
            }
        }
    }

    /* band 17: rrtmg_sw_taumol.f90:342-462 */
    bb = B(17);
    {
        const int ngs = 6;
        const double strrat = 0.364641;
        layreffr = 30;
        for (lay = 1; lay <= laytrop; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        laysolfr = nlayers;
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            if (jp[lay - 1] < layreffr && jp[lay] >= layreffr) laysolfr = lay;
            eta(c->colh2o[lay], c->colco2[lay], strrat, 4.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + js;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + js;
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSB, f, ind0, ind1, ig, 5, 6) +
                                         c->colh2o[lay] * FOR(c, lay, indf, ig);
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* bands 18, 19, 21, 22 (lower) share the h2o/x eta structure; each is
     * written out because the upper-atmosphere treatment differs.
     * band 18: rrtmg_sw_taumol.f90:465-561 */
    bb = B(18);
    {
        const int ngs = 18;
        const double strrat = 38.9589;
        layreffr = 6;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            eta(c->colh2o[lay], c->colch4[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colch4[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig);
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 19: rrtmg_sw_taumol.f90:564-660 */
    bb = B(19);
    {
        const int ngs = 26;
        const double strrat = 5.49281;
        layreffr = 3;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            eta(c->colh2o[lay], c->colco2[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colco2[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig);
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 20: rrtmg_sw_taumol.f90:663-746 */
    bb = B(20);
    {
        const int ngs = 34;
        layreffr = 3;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + 1;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + 1;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] =
                    c->colh2o[lay] * ((MAJ4(ABSA, c, lay, ind0, ind1, ig)) + SELF(c, lay, inds, ig) +
                                      FOR(c, lay, indf, ig)) +
                    c->colch4[lay] * bb->absch4[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = bb->sfluxref[ig - 1];
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] =
                    c->colh2o[lay] * (MAJ4(ABSB, c, lay, ind0, ind1, ig) + FOR(c, lay, indf, ig)) +
                    c->colch4[lay] * bb->absch4[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 21: rrtmg_sw_taumol.f90:749-868 */
    bb = B(21);
    {
        const int ngs = 44;
        const double strrat = 0.0045321;
        layreffr = 8;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            eta(c->colh2o[lay], c->colco2[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            eta(c->colh2o[lay], c->colco2[lay], strrat, 4.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + js;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + js;
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSB, f, ind0, ind1, ig, 5, 6) +
                                         c->colh2o[lay] * FOR(c, lay, indf, ig);
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 22: rrtmg_sw_taumol.f90:871-977 */
    bb = B(22);
    {
        const int ngs = 54;
        const double o2adj = 1.6, strrat = 0.022708;
        layreffr = 2;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            double o2cont = 4.35e-4 * c->colo2[lay] / (350.0 * 2.0);
            eta(c->colh2o[lay], c->colo2[lay], o2adj * strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig)) +
                                         o2cont;
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            double o2cont = 4.35e-4 * c->colo2[lay] / (350.0 * 2.0);
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = c->colo2[lay] * o2adj * MAJ4(ABSB, c, lay, ind0, ind1, ig) + o2cont;
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 23: rrtmg_sw_taumol.f90:980-1051 */
    bb = B(23);
    {
        const int ngs = 56;
        const double givfac = 1.029;
        layreffr = 6;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + 1;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + 1;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylg[ig - 1];
                c->taug[lay][ngs + ig] =
                    c->colh2o[lay] * (givfac * MAJ4(ABSA, c, lay, ind0, ind1, ig) + SELF(c, lay, inds, ig) +
                                      FOR(c, lay, indf, ig));
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = bb->sfluxref[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = 0.0;
                c->taur[lay][ngs + ig] = c->colmol[lay] * bb->raylg[ig - 1];
            }
    }

    /* band 24: rrtmg_sw_taumol.f90:1054-1153 */
    bb = B(24);
    {
        const int ngs = 66;
        const double strrat = 0.124692;
        layreffr = 1;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            eta(c->colh2o[lay], c->colo2[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * (RAYLA(bb, ig, js) + fs * (RAYLA(bb, ig, js + 1) - RAYLA(bb, ig, js)));
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10) +
                                         c->colo3[lay] * bb->abso3a[ig - 1] +
                                         c->colh2o[lay] * (SELF(c, lay, inds, ig) + FOR(c, lay, indf, ig));
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylb[ig - 1];
                c->taug[lay][ngs + ig] =
                    c->colo2[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig) + c->colo3[lay] * bb->abso3b[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 25: rrtmg_sw_taumol.f90:1156-1217 */
    bb = B(25);
    {
        const int ngs = 74;
        layreffr = 2;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay) {
            if (jp[lay] < layreffr && jp[lay + 1] >= layreffr) laysolfr = (lay + 1 < laytrop ? lay + 1 : laytrop);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + 1;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + 1;
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylg[ig - 1];
                c->taug[lay][ngs + ig] =
                    c->colh2o[lay] * MAJ4(ABSA, c, lay, ind0, ind1, ig) + c->colo3[lay] * bb->abso3a[ig - 1];
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = bb->sfluxref[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        for (lay = laytrop + 1; lay <= nlayers; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylg[ig - 1];
                c->taug[lay][ngs + ig] = c->colo3[lay] * bb->abso3b[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
    }

    /* band 26: rrtmg_sw_taumol.f90:1220-1268 */
    bb = B(26);
    {
        const int ngs = 80;
        laysolfr = laytrop;
        for (lay = 1; lay <= laytrop; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = bb->sfluxref[ig - 1];
                c->taug[lay][ngs + ig] = 0.0;
                c->taur[lay][ngs + ig] = c->colmol[lay] * bb->raylg[ig - 1];
            }
        for (lay = laytrop + 1; lay <= nlayers; ++lay)
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = 0.0;
                c->taur[lay][ngs + ig] = c->colmol[lay] * bb->raylg[ig - 1];
            }
    }

    /* band 27: rrtmg_sw_taumol.f90:1271-1347 */
    bb = B(27);
    {
        const int ngs = 86;
        const double scalekur = 50.15 / 48.37;
        layreffr = 32;
        for (lay = 1; lay <= laytrop; ++lay) {
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + 1;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + 1;
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylg[ig - 1];
                c->taug[lay][ngs + ig] = c->colo3[lay] * MAJ4(ABSA, c, lay, ind0, ind1, ig);
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        laysolfr = nlayers;
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            if (jp[lay - 1] < layreffr && jp[lay] >= layreffr) laysolfr = lay;
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            for (ig = 1; ig <= bb->ng; ++ig) {
                tauray = c->colmol[lay] * bb->raylg[ig - 1];
                c->taug[lay][ngs + ig] = c->colo3[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig);
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = scalekur * bb->sfluxref[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 28: rrtmg_sw_taumol.f90:1350-1455 */
    bb = B(28);
    {
        const int ngs = 94;
        const double strrat = 6.67029e-07;
        layreffr = 58;
        for (lay = 1; lay <= laytrop; ++lay) {
            eta(c->colo3[lay], c->colo2[lay], strrat, 8.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + js;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + js;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSA, f, ind0, ind1, ig, 9, 10);
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        laysolfr = nlayers;
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            if (jp[lay - 1] < layreffr && jp[lay] >= layreffr) laysolfr = lay;
            eta(c->colo3[lay], c->colo2[lay], strrat, 4.0, &speccomb, &js, &fs);
            fac8(c, lay, fs, &f);
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + js;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + js;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] = speccomb * MAJ8(ABSB, f, ind0, ind1, ig, 5, 6);
                if (lay == laysolfr)
                    c->sfluxzen[ngs + ig] =
                        SFLUXREF(bb, ig, js) + fs * (SFLUXREF(bb, ig, js + 1) - SFLUXREF(bb, ig, js));
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }

    /* band 29: rrtmg_sw_taumol.f90:1458-1536 */
    bb = B(29);
    {
        const int ngs = 100;
        layreffr = 49;
        for (lay = 1; lay <= laytrop; ++lay) {
            ind0 = ((jp[lay] - 1) * 5 + (jt[lay] - 1)) * bb->nspa + 1;
            ind1 = (jp[lay] * 5 + (jt1[lay] - 1)) * bb->nspa + 1;
            inds = c->indself[lay];
            indf = c->indfor[lay];
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] =
                    c->colh2o[lay] * ((MAJ4(ABSA, c, lay, ind0, ind1, ig)) + SELF(c, lay, inds, ig) +
                                      FOR(c, lay, indf, ig)) +
                    c->colco2[lay] * bb->absco2[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
        laysolfr = nlayers;
        for (lay = laytrop + 1; lay <= nlayers; ++lay) {
            if (jp[lay - 1] < layreffr && jp[lay] >= layreffr) laysolfr = lay;
            ind0 = ((jp[lay] - 13) * 5 + (jt[lay] - 1)) * bb->nspb + 1;
            ind1 = ((jp[lay] - 12) * 5 + (jt1[lay] - 1)) * bb->nspb + 1;
            tauray = c->colmol[lay] * bb->rayl;
            for (ig = 1; ig <= bb->ng; ++ig) {
                c->taug[lay][ngs + ig] =
                    c->colco2[lay] * MAJ4(ABSB, c, lay, ind0, ind1, ig) + c->colh2o[lay] * bb->absh2o[ig - 1];
                if (lay == laysolfr) c->sfluxzen[ngs + ig] = bb->sfluxref[ig - 1];
                c->taur[lay][ngs + ig] = tauray;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* reftra_sw (rrtmg_sw_reftra.f90:127-303), kmodts = 2                      */
/* ------------------------------------------------------------------------ */

static double tbl_exp(double x)
{
    /* table lookup used throughout the SW code (e.g. rrtmg_sw_reftra.f90:181-184) */
    double tblind = x / (S.bpade + x);
    int itind = (int)(1.e4 * tblind + 0.5);
    return S.exp_tbl[itind];
}

static void reftra_sw(int nlayers, const int *lrtchk, const double *pgg, double prmuz,
                      const double *ptau, const double *pw, double *pref, double *prefd,
                      double *ptra, double *ptrad)
{
    const double eps = 1.e-08, od_lo = 0.06, zwcrit = 0.9999995;
    for (int jk = 1; jk <= nlayers; ++jk) {
        if (!lrtchk[jk]) {
            pref[jk] = 0.0;
            ptra[jk] = 1.0;
            prefd[jk] = 0.0;
            ptrad[jk] = 1.0;
            continue;
        }
        double zto1 = ptau[jk], zw = pw[jk], zg = pgg[jk];
        double zg3 = 3.0 * zg;
        double zgamma1 = (8.0 - zw * (5.0 + zg3)) * 0.25;
        double zgamma2 = 3.0 * (zw * (1.0 - zg)) * 0.25;
        double zgamma3 = (2.0 - zg3 * prmuz) * 0.25;
        double zgamma4 = 1.0 - zgamma3;
        double q = zg / (1.0 - zg);
        double zwo = zw / (1.0 - (1.0 - zw) * (q * q));
        if (zwo >= zwcrit) {
            double za = zgamma1 * prmuz;
            double za1 = za - zgamma3;
            double zgt = zgamma1 * zto1;
            double ze1 = fmin(zto1 / prmuz, 500.0);
            double ze2;
            if (ze1 <= od_lo)
                ze2 = 1.0 - ze1 + 0.5 * ze1 * ze1;
            else
                ze2 = tbl_exp(ze1);
            pref[jk] = (zgt - za1 * (1.0 - ze2)) / (1.0 + zgt);
            ptra[jk] = 1.0 - pref[jk];
            prefd[jk] = zgt / (1.0 + zgt);
            ptrad[jk] = 1.0 - prefd[jk];
            if (ze2 == 1.0) {
                pref[jk] = 0.0;
                ptra[jk] = 1.0;
                prefd[jk] = 0.0;
                ptrad[jk] = 1.0;
            }
        } else {
            double za1 = zgamma1 * zgamma4 + zgamma2 * zgamma3;
            double za2 = zgamma1 * zgamma3 + zgamma2 * zgamma4;
            double zrk = sqrt(zgamma1 * zgamma1 - zgamma2 * zgamma2);
            double zrp = zrk * prmuz;
            double zrp1 = 1.0 + zrp;
            double zrm1 = 1.0 - zrp;
            double zrk2 = 2.0 * zrk;
            double zrpp = 1.0 - zrp * zrp;
            double zrkg = zrk + zgamma1;
            double zr1 = zrm1 * (za2 + zrk * zgamma3);
            double zr2 = zrp1 * (za2 - zrk * zgamma3);
            double zr3 = zrk2 * (zgamma3 - za2 * prmuz);
            double zr4 = zrpp * zrkg;
            double zr5 = zrpp * (zrk - zgamma1);
            double zt1 = zrp1 * (za1 + zrk * zgamma4);
            double zt2 = zrm1 * (za1 - zrk * zgamma4);
            double zt3 = zrk2 * (zgamma4 + za1 * prmuz);
            double zt4 = zr4;
            double zt5 = zr5;
            double zbeta = (zgamma1 - zrk) / zrkg;
            double ze1 = fmin(zrk * zto1, 500.0);
            double ze2 = fmin(zto1 / prmuz, 500.0);
            double zem1, zep1, zem2, zep2;
            if (ze1 <= od_lo) {
                zem1 = 1.0 - ze1 + 0.5 * ze1 * ze1;
                zep1 = 1.0 / zem1;
            } else {
                zem1 = tbl_exp(ze1);
                zep1 = 1.0 / zem1;
            }
            if (ze2 <= od_lo) {
                zem2 = 1.0 - ze2 + 0.5 * ze2 * ze2;
                zep2 = 1.0 / zem2;
            } else {
                zem2 = tbl_exp(ze2);
                zep2 = 1.0 / zem2;
            }
            double zdenr = zr4 * zep1 + zr5 * zem1;
            double zdent = zt4 * zep1 + zt5 * zem1;
            if (zdenr >= -eps && zdenr <= eps) {
                pref[jk] = eps;
                ptra[jk] = zem2;
            } else {
                pref[jk] = zw * (zr1 * zep1 - zr2 * zem1 - zr3 * zem2) / zdenr;
                ptra[jk] = zem2 - zem2 * zw * (zt1 * zep1 - zt2 * zem1 - zt3 * zep2) / zdent;
            }
            double zemm = zem1 * zem1;
            double zdend = 1.0 / ((1.0 - zbeta * zemm) * zrkg);
            prefd[jk] = zgamma2 * (1.0 - zemm) * zdend;
            ptrad[jk] = zrk2 * zem1 * zdend;
        }
    }
}

void oracle_sw_reftra(int nlay, const double *gg, double prmuz, const double *tau, const double *w,
                      double *pref, double *prefd, double *ptra, double *ptrad)
{
    int chk[MAXLAY + 2];
    double g1[MAXLAY + 2], t1[MAXLAY + 2], w1[MAXLAY + 2];
    double r[MAXLAY + 2], rd[MAXLAY + 2], tr[MAXLAY + 2], trd[MAXLAY + 2];
    for (int k = 1; k <= nlay; ++k) {
        chk[k] = 1;
        g1[k] = gg[k - 1];
        t1[k] = tau[k - 1];
        w1[k] = w[k - 1];
    }
    reftra_sw(nlay, chk, g1, prmuz, t1, w1, r, rd, tr, trd);
    for (int k = 1; k <= nlay; ++k) {
        pref[k - 1] = r[k];
        prefd[k - 1] = rd[k];
        ptra[k - 1] = tr[k];
        ptrad[k - 1] = trd[k];
    }
}

/* vrtqdr_sw (rrtmg_sw_vrtqdr.f90:103-150); pfd/pfu for one g-point */
static void vrtqdr_sw(int klev, const double *pref, const double *prefd, const double *ptra,
                      const double *ptrad, const double *pdbt, double *prdnd, double *prup,
                      double *prupd, const double *ptdbt, double *pfd, double *pfu)
{
    double ztdn[MAXLAY + 3];
    double zreflect = 1.0 / (1.0 - prefd[klev + 1] * prefd[klev]);
    prup[klev] = pref[klev] + (ptrad[klev] * ((ptra[klev] - pdbt[klev]) * prefd[klev + 1] +
                                              pdbt[klev] * pref[klev + 1])) * zreflect;
    prupd[klev] = prefd[klev] + ptrad[klev] * ptrad[klev] * prefd[klev + 1] * zreflect;
    for (int jk = 1; jk <= klev - 1; ++jk) {
        int ikp = klev + 1 - jk;
        int ikx = ikp - 1;
        zreflect = 1.0 / (1.0 - prupd[ikp] * prefd[ikx]);
        prup[ikx] = pref[ikx] + (ptrad[ikx] * ((ptra[ikx] - pdbt[ikx]) * prupd[ikp] +
                                               pdbt[ikx] * prup[ikp])) * zreflect;
        prupd[ikx] = prefd[ikx] + ptrad[ikx] * ptrad[ikx] * prupd[ikp] * zreflect;
    }
    ztdn[1] = 1.0;
    prdnd[1] = 0.0;
    ztdn[2] = ptra[1];
    prdnd[2] = prefd[1];
    for (int jk = 2; jk <= klev; ++jk) {
        int ikp = jk + 1;
        zreflect = 1.0 / (1.0 - prefd[jk] * prdnd[jk]);
        ztdn[ikp] = ptdbt[jk] * ptra[jk] +
                    (ptrad[jk] * ((ztdn[jk] - ptdbt[jk]) + ptdbt[jk] * pref[jk] * prdnd[jk])) * zreflect;
        prdnd[ikp] = prefd[jk] + ptrad[jk] * ptrad[jk] * prdnd[jk] * zreflect;
    }
    for (int jk = 1; jk <= klev + 1; ++jk) {
        zreflect = 1.0 / (1.0 - prdnd[jk] * prupd[jk]);
        pfu[jk] = (ptdbt[jk] * prup[jk] + (ztdn[jk] - ptdbt[jk]) * prupd[jk]) * zreflect;
        pfd[jk] = ptdbt[jk] + (ztdn[jk] - ptdbt[jk] + ptdbt[jk] * prup[jk] * prdnd[jk]) * zreflect;
    }
}

/* spcvrt_sw clear-sky path (rrtmg_sw_spcvrt.f90:259-621).  With no clouds and
 * no aerosol the reference's total-sky pass sees identical reftra/vrtqdr
 * inputs (lrtchkcld is false for every layer, zclear = 1, zcloud = 0), so the
 * total-sky fluxes equal the clear-sky ones bit for bit; they are computed
 * once here. */
static void spcvrt_sw(swcol_t *c, double prmu0, const double *palbd, const double *palbp,
                      const double *adjflux, double *pbbcd, double *pbbcu)
{
    const int klev = c->nlayers;
    const double od_lo = 0.06;
    double zrefc[MAXLAY + 3], zrefdc[MAXLAY + 3], ztrac[MAXLAY + 3], ztradc[MAXLAY + 3];
    double zdbtc[MAXLAY + 3], ztdbtc[MAXLAY + 3], zrdndc[MAXLAY + 3], zrupc[MAXLAY + 3], zrupdc[MAXLAY + 3];
    double ztauc[MAXLAY + 3], zomcc[MAXLAY + 3], zgcc[MAXLAY + 3];
    double zcd[MAXLAY + 3], zcu[MAXLAY + 3];
    int lrtchkclr[MAXLAY + 3];
    for (int jk = 1; jk <= klev + 1; ++jk) {
        pbbcd[jk] = 0.0;
        pbbcu[jk] = 0.0;
    }
    taumol_sw(c);
    int iw = 0;
    for (int jb = 16; jb <= 29; ++jb) {
        int ibm = jb - 15;
        int igt = S.ngc[ibm - 1];
        for (int jg = 1; jg <= igt; ++jg) {
            iw = iw + 1;
            double zincflx = adjflux[jb] * c->sfluxzen[iw] * prmu0;
            ztdbtc[1] = 1.0;
            zdbtc[klev + 1] = 0.0;
            ztrac[klev + 1] = 0.0;
            ztradc[klev + 1] = 0.0;
            zrefc[klev + 1] = palbp[ibm];
            zrefdc[klev + 1] = palbd[ibm];
            zrupc[klev + 1] = palbp[ibm];
            zrupdc[klev + 1] = palbd[ibm];
            for (int jk = 1; jk <= klev; ++jk) {
                int ikl = klev + 1 - jk;
                lrtchkclr[jk] = 1;
                /* ptaua = 0, pomga = 1, pasya = 0 (iaer = 0) */
                ztauc[jk] = c->taur[ikl][iw] + c->taug[ikl][iw] + 0.0;
                zomcc[jk] = c->taur[ikl][iw] * 1.0 + 0.0 * 1.0;
                zgcc[jk] = 0.0 * 1.0 * 0.0 / zomcc[jk];
                zomcc[jk] = zomcc[jk] / ztauc[jk];
                double zf = zgcc[jk] * zgcc[jk];
                double zwf = zomcc[jk] * zf;
                ztauc[jk] = (1.0 - zwf) * ztauc[jk];
                zomcc[jk] = (zomcc[jk] - zwf) / (1.0 - zwf);
                zgcc[jk] = (zgcc[jk] - zf) / (1.0 - zf);
            }
            reftra_sw(klev, lrtchkclr, zgcc, prmu0, ztauc, zomcc, zrefc, zrefdc, ztrac, ztradc);
            for (int jk = 1; jk <= klev; ++jk) {
                double ze1 = ztauc[jk] / prmu0;
                double zdbtmc;
                if (ze1 <= od_lo)
                    zdbtmc = 1.0 - ze1 + 0.5 * ze1 * ze1;
                else
                    zdbtmc = tbl_exp(ze1);
                zdbtc[jk] = zdbtmc;
                ztdbtc[jk + 1] = zdbtc[jk] * ztdbtc[jk];
            }
            vrtqdr_sw(klev, zrefc, zrefdc, ztrac, ztradc, zdbtc, zrdndc, zrupc, zrupdc, ztdbtc, zcd, zcu);
            for (int jk = 1; jk <= klev + 1; ++jk) {
                int ikl = klev + 2 - jk;
                pbbcu[ikl] = pbbcu[ikl] + zincflx * zcu[jk];
                pbbcd[ikl] = pbbcd[ikl] + zincflx * zcd[jk];
            }
        }
    }
}

/* cldprop_sw with inflag = 0 (rrtmg_sw_cldprop.f90:118-157): a layer that
 * passes the cloud test (fraction >= 1e-20 and water path or summed depth
 * >= 1e-20) gets delta-M scaled optics from the input optical depth, single
 * scattering albedo, asymmetry and forward-scattering fraction; every other
 * layer has depth 0, albedo 1, asymmetry 0.  Arrays [lay][ib], ib = 1..14. */
/* swcldpr's cloud optics coefficients (rrtmg_sw_init.f90, module rrsw_cld)
 * from rrtmg_sw_cld.tbl; Fortran (k, ib) arrays, bands 16-29, as [ib][k] */
static struct {
    int ready;
    double extliq1[14][58], ssaliq1[14][58], asyliq1[14][58];
    double extice2[14][43], ssaice2[14][43], asyice2[14][43];
    double extice3[14][46], ssaice3[14][46], asyice3[14][46], fdlice3[14][46];
    double abari[5], bbari[5], cbari[5], dbari[5], ebari[5], fbari[5], wavenum2[14];
} SC;

int oracle_sw_cld_init(const char *path)
{
    otbl_t t;
    if (otbl_load(&t, path) != 0) return -1;
    const char *nm[] = {"swcld/extliq1", "swcld/ssaliq1", "swcld/asyliq1", "swcld/extice2", "swcld/ssaice2",
                        "swcld/asyice2", "swcld/extice3", "swcld/ssaice3", "swcld/asyice3", "swcld/fdlice3",
                        "swcld/abari",   "swcld/bbari",   "swcld/cbari",   "swcld/dbari",   "swcld/ebari",
                        "swcld/fbari",   "swcld/wavenum2"};
    double *dst[] = {&SC.extliq1[0][0], &SC.ssaliq1[0][0], &SC.asyliq1[0][0], &SC.extice2[0][0],
                     &SC.ssaice2[0][0], &SC.asyice2[0][0], &SC.extice3[0][0], &SC.ssaice3[0][0],
                     &SC.asyice3[0][0], &SC.fdlice3[0][0], SC.abari, SC.bbari, SC.cbari, SC.dbari, SC.ebari,
                     SC.fbari, SC.wavenum2};
    const long len[] = {812, 812, 812, 602, 602, 602, 644, 644, 644, 644, 5, 5, 5, 5, 5, 5, 14};
    for (int i = 0; i < 17; ++i) {
        long n = 0;
        const double *v = otbl_d(&t, nm[i], &n);
        if (!v || n != len[i]) return -2;
        memcpy(dst[i], v, sizeof(double) * len[i]);
    }
    SC.ready = 1;
    return 0;
}

/* cldprop_sw's water-path optics of one band (0-based ib) for the given
 * paths and radii (rrtmg_sw_cldprop.f90:169-337, per band): r = final
 * depth, albedo, asymmetry.  -8 where the reference stops. */
static int sw_wp_optics(int ib, int iceflag, int liqflag, double ciwp, double clwp, double radice, double radliq,
                        double r[3])
{
    const double eps = 1.e-06, cldmin = 1.e-20;
    double extcoice = 0.0, ssacoice = 0.0, gice = 0.0, forwice = 0.0;
    double extcoliq = 0.0, ssacoliq = 0.0, gliq = 0.0, forwliq = 0.0;
    if (ciwp != 0.0) {
        if (iceflag == 1) {
            if (radice < 13.0 || radice > 130.0) return -8;
            double w2 = SC.wavenum2[ib];
            int icx = w2 > 1.43e04 ? 1 : w2 > 7.7e03 ? 2 : w2 > 5.3e03 ? 3 : w2 > 4.0e03 ? 4 : 5;
            extcoice = (SC.abari[icx - 1] + SC.bbari[icx - 1] / radice);
            ssacoice = 1.0 - SC.cbari[icx - 1] - SC.dbari[icx - 1] * radice;
            gice = SC.ebari[icx - 1] + SC.fbari[icx - 1] * radice;
            if (gice >= 1.0) gice = 1.0 - eps;
            forwice = gice * gice;
        } else if (iceflag == 2 || iceflag == 3) {
            const int nr = iceflag == 2 ? 43 : 46;
            if (radice < 5.0 || radice > (iceflag == 2 ? 131.0 : 140.0)) return -8;
            double factor = (radice - 2.0) / 3.0;
            int index = (int)factor;
            if (index == nr) index = nr - 1;
            double fint = factor - (double)index;
            const double *ex = iceflag == 2 ? SC.extice2[ib] : SC.extice3[ib];
            const double *ss = iceflag == 2 ? SC.ssaice2[ib] : SC.ssaice3[ib];
            const double *as = iceflag == 2 ? SC.asyice2[ib] : SC.asyice3[ib];
            extcoice = ex[index - 1] + fint * (ex[index] - ex[index - 1]);
            ssacoice = ss[index - 1] + fint * (ss[index] - ss[index - 1]);
            gice = as[index - 1] + fint * (as[index] - as[index - 1]);
            if (iceflag == 2) {
                forwice = gice * gice;
            } else {
                const double *fd = SC.fdlice3[ib];
                double fdelta = fd[index - 1] + fint * (fd[index] - fd[index - 1]);
                if (fdelta < 0.0 || fdelta > 1.0) return -8;
                forwice = fdelta + 0.5 / ssacoice;
                if (forwice > gice) forwice = gice;
            }
        } else {
            return -8;
        }
        if (extcoice < 0.0 || ssacoice > 1.0 || ssacoice < 0.0 || gice > 1.0 || gice < 0.0) return -8;
    }
    if (clwp != 0.0) {
        if (liqflag != 1 || radliq < 2.5 || radliq > 60.0) return -8;
        int index = (int)(radliq - 1.5);
        if (index == 0) index = 1;
        if (index == 58) index = 57;
        double fint = radliq - 1.5 - (double)index;
        const double *ex = SC.extliq1[ib], *ss = SC.ssaliq1[ib], *as = SC.asyliq1[ib];
        extcoliq = ex[index - 1] + fint * (ex[index] - ex[index - 1]);
        ssacoliq = ss[index - 1] + fint * (ss[index] - ss[index - 1]);
        if (fint < 0.0 && ssacoliq > 1.0) ssacoliq = ss[index - 1];
        gliq = as[index - 1] + fint * (as[index] - as[index - 1]);
        forwliq = gliq * gliq;
        if (extcoliq < 0.0 || ssacoliq > 1.0 || ssacoliq < 0.0 || gliq > 1.0 || gliq < 0.0) return -8;
    }
    double tauliqorig = clwp * extcoliq, tauiceorig = ciwp * extcoice;
    double ssaliq = ssacoliq * (1.0 - forwliq) / (1.0 - forwliq * ssacoliq);
    double tauliq = (1.0 - forwliq * ssacoliq) * tauliqorig;
    double ssaice = ssacoice * (1.0 - forwice) / (1.0 - forwice * ssacoice);
    double tauice = (1.0 - forwice * ssacoice) * tauiceorig;
    double scatliq = ssaliq * tauliq, scatice = ssaice * tauice;
    double t = tauliq + tauice;
    if (t == 0.0) t = cldmin;
    if (scatice == 0.0) scatice = cldmin;
    r[0] = t;
    r[1] = (scatliq + scatice) / t;
    if (iceflag == 3)
        r[2] = (1.0 / (scatliq + scatice)) *
               (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * ((gice - forwice) / (1.0 - forwice)));
    else
        r[2] = (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * (gice - forwice) / (1.0 - forwice)) /
               (scatliq + scatice);
    return 0;
}

/* cldprop_sw with inflag 2 (rrtmg_sw_cldprop.f90:143-338): per layer that
 * passes the cloud test, ice extinction, albedo and asymmetry by iceflag 1
 * (the abari..fbari fits in the radius, band group by the band's upper
 * wavenumber), 2 or 3 (tables interpolated in the radius; 3 with the delta
 * fraction fdlice3), liquid by liqflag 1 (tables in the radius); each
 * phase delta-scaled by its forward fraction, then combined, with the
 * reference's floors (cldmin for a zero total depth or ice scattering).
 * taucloud, ssacloud, asmcloud as cldprop_sw0's (0, 1, 0 where the test
 * fails).  -8 where the reference stops (radius out of range, optics out of
 * [0, 1], a negative extinction); iceflag 1-3 and liqflag 1 only. */
static int cldprop_sw2(int nl, int iceflag, int liqflag, const double *cldfrac, double (*tauc)[15],
                       const double *ciwp, const double *clwp, const double *rei, const double *rel,
                       double (*taucloud)[15], double (*ssacloud)[15], double (*asmcloud)[15])
{
    const double eps = 1.e-06, cldmin = 1.e-20;
    double extcoice[15], ssacoice[15], gice[15], forwice[15], extcoliq[15], ssacoliq[15], gliq[15], forwliq[15];
    for (int lay = 1; lay <= nl; ++lay) {
        double tauctot = 0.0;
        for (int ib = 1; ib <= 14; ++ib) {
            taucloud[lay][ib] = 0.0;
            ssacloud[lay][ib] = 1.0;
            asmcloud[lay][ib] = 0.0;
            tauctot = tauctot + tauc[lay][ib];
        }
        double cwp = ciwp[lay] + clwp[lay];
        if (!(cldfrac[lay] >= cldmin && (cwp >= cldmin || tauctot >= cldmin))) continue;
        double radice = rei[lay];
        if (ciwp[lay] == 0.0) {
            for (int ib = 1; ib <= 14; ++ib) extcoice[ib] = ssacoice[ib] = gice[ib] = forwice[ib] = 0.0;
        } else if (iceflag == 1) {
            if (radice < 13.0 || radice > 130.0) return -8;
            for (int ib = 1; ib <= 14; ++ib) {
                double w2 = SC.wavenum2[ib - 1];
                int icx = w2 > 1.43e04 ? 1 : w2 > 7.7e03 ? 2 : w2 > 5.3e03 ? 3 : w2 > 4.0e03 ? 4 : 5;
                extcoice[ib] = SC.abari[icx - 1] + SC.bbari[icx - 1] / radice;
                ssacoice[ib] = 1.0 - SC.cbari[icx - 1] - SC.dbari[icx - 1] * radice;
                gice[ib] = SC.ebari[icx - 1] + SC.fbari[icx - 1] * radice;
                if (gice[ib] >= 1.0) gice[ib] = 1.0 - eps;
                forwice[ib] = gice[ib] * gice[ib];
                if (extcoice[ib] < 0.0 || ssacoice[ib] > 1.0 || ssacoice[ib] < 0.0 || gice[ib] > 1.0 || gice[ib] < 0.0)
                    return -8;
            }
        } else if (iceflag == 2 || iceflag == 3) {
            const int nr = iceflag == 2 ? 43 : 46;
            if (radice < 5.0 || radice > (iceflag == 2 ? 131.0 : 140.0)) return -8;
            double factor = (radice - 2.0) / 3.0;
            int index = (int)factor;
            if (index == nr) index = nr - 1;
            double fint = factor - (double)index;
            for (int ib = 1; ib <= 14; ++ib) {
                const double *ex = iceflag == 2 ? SC.extice2[ib - 1] : SC.extice3[ib - 1];
                const double *ss = iceflag == 2 ? SC.ssaice2[ib - 1] : SC.ssaice3[ib - 1];
                const double *as = iceflag == 2 ? SC.asyice2[ib - 1] : SC.asyice3[ib - 1];
                extcoice[ib] = ex[index - 1] + fint * (ex[index] - ex[index - 1]);
                ssacoice[ib] = ss[index - 1] + fint * (ss[index] - ss[index - 1]);
                gice[ib] = as[index - 1] + fint * (as[index] - as[index - 1]);
                if (iceflag == 2) {
                    forwice[ib] = gice[ib] * gice[ib];
                } else {
                    const double *fd = SC.fdlice3[ib - 1];
                    double fdelta = fd[index - 1] + fint * (fd[index] - fd[index - 1]);
                    if (fdelta < 0.0 || fdelta > 1.0) return -8;
                    forwice[ib] = fdelta + 0.5 / ssacoice[ib];
                    if (forwice[ib] > gice[ib]) forwice[ib] = gice[ib];
                }
                if (extcoice[ib] < 0.0 || ssacoice[ib] > 1.0 || ssacoice[ib] < 0.0 || gice[ib] > 1.0 || gice[ib] < 0.0)
                    return -8;
            }
        } else {
            return -8;
        }
        if (clwp[lay] == 0.0) {
            for (int ib = 1; ib <= 14; ++ib) extcoliq[ib] = ssacoliq[ib] = gliq[ib] = forwliq[ib] = 0.0;
        } else if (liqflag == 1) {
            double radliq = rel[lay];
            if (radliq < 2.5 || radliq > 60.0) return -8;
            int index = (int)(radliq - 1.5);
            if (index == 0) index = 1;
            if (index == 58) index = 57;
            double fint = radliq - 1.5 - (double)index;
            for (int ib = 1; ib <= 14; ++ib) {
                const double *ex = SC.extliq1[ib - 1], *ss = SC.ssaliq1[ib - 1], *as = SC.asyliq1[ib - 1];
                extcoliq[ib] = ex[index - 1] + fint * (ex[index] - ex[index - 1]);
                ssacoliq[ib] = ss[index - 1] + fint * (ss[index] - ss[index - 1]);
                if (fint < 0.0 && ssacoliq[ib] > 1.0) ssacoliq[ib] = ss[index - 1];
                gliq[ib] = as[index - 1] + fint * (as[index] - as[index - 1]);
                forwliq[ib] = gliq[ib] * gliq[ib];
                if (extcoliq[ib] < 0.0 || ssacoliq[ib] > 1.0 || ssacoliq[ib] < 0.0 || gliq[ib] > 1.0 || gliq[ib] < 0.0)
                    return -8;
            }
        } else {
            return -8;
        }
        for (int ib = 1; ib <= 14; ++ib) {
            double tauliqorig = clwp[lay] * extcoliq[ib];
            double tauiceorig = ciwp[lay] * extcoice[ib];
            double ssaliq = ssacoliq[ib] * (1.0 - forwliq[ib]) / (1.0 - forwliq[ib] * ssacoliq[ib]);
            double tauliq = (1.0 - forwliq[ib] * ssacoliq[ib]) * tauliqorig;
            double ssaice = ssacoice[ib] * (1.0 - forwice[ib]) / (1.0 - forwice[ib] * ssacoice[ib]);
            double tauice = (1.0 - forwice[ib] * ssacoice[ib]) * tauiceorig;
            double scatliq = ssaliq * tauliq;
            double scatice = ssaice * tauice;
            taucloud[lay][ib] = tauliq + tauice;
            if (taucloud[lay][ib] == 0.0) taucloud[lay][ib] = cldmin;
            if (scatice == 0.0) scatice = cldmin;
            ssacloud[lay][ib] = (scatliq + scatice) / taucloud[lay][ib];
            if (iceflag == 3)
                asmcloud[lay][ib] = (1.0 / (scatliq + scatice)) *
                                    (scatliq * (gliq[ib] - forwliq[ib]) / (1.0 - forwliq[ib]) +
                                     scatice * ((gice[ib] - forwice[ib]) / (1.0 - forwice[ib])));
            else
                asmcloud[lay][ib] = (scatliq * (gliq[ib] - forwliq[ib]) / (1.0 - forwliq[ib]) +
                                     scatice * (gice[ib] - forwice[ib]) / (1.0 - forwice[ib])) /
                                    (scatliq + scatice);
        }
    }
    return 0;
}

static void cldprop_sw0(int nl, const double *cldfrac, double (*tauc)[15], double (*ssac)[15],
                        double (*asmc)[15], double (*fsfc)[15], const double *cwp, double (*taucloud)[15],
                        double (*ssacloud)[15], double (*asmcloud)[15])
{
    const double cldmin = 1.e-20;
    for (int lay = 1; lay <= nl; ++lay) {
        double tauctot = 0.0;
        for (int ib = 1; ib <= 14; ++ib) {
            taucloud[lay][ib] = 0.0;
            ssacloud[lay][ib] = 1.0;
            asmcloud[lay][ib] = 0.0;
            tauctot = tauctot + tauc[lay][ib];
        }
        if (cldfrac[lay] >= cldmin && (cwp[lay] >= cldmin || tauctot >= cldmin)) {
            for (int ib = 1; ib <= 14; ++ib) {
                double ffp = fsfc[lay][ib];
                double ffp1 = 1.0 - ffp;
                double ffpssa = 1.0 - ffp * ssac[lay][ib];
                ssacloud[lay][ib] = ffp1 * ssac[lay][ib] / ffpssa;
                taucloud[lay][ib] = ffpssa * tauc[lay][ib];
                asmcloud[lay][ib] = (asmc[lay][ib] - ffp) / ffp1;
            }
        }
    }
}

/* spcvrt_sw with clouds (rrtmg_sw_spcvrt.f90:302-600; icpr = 1, idelm = 0,
 * no aerosol): the clear optics of each layer as in the clear-sky path, the
 * cloudy optics (clear plus delta-scaled cloud, :459-466) through reftra
 * where the fraction exceeds 1e-12, the total-sky layer values blended by
 * the fraction (:502-549), and vrtqdr for the clear and the total stream.
 * The cloud inputs are per (layer, g-point): spcvmc_sw
 * (rrtmg_sw_spcvmc.f90:302-600) is spcvrt_sw with pcldfmc(ikl, iw),
 * ptaucmc(ikl, iw), ... in place of pclfr(ikl), ptauc(ikl, ibm), ..., so
 * the nomcica driver passes its per-layer fraction and per-band optics
 * repeated over the g-points and the McICA driver its subcolumns. */
static void spcvrt_sw_cloud(swcol_t *c, double prmu0, const double *palbd, const double *palbp,
                            const double *adjflux, double (*pcf)[NGSW + 1], double (*ptauc)[NGSW + 1],
                            double (*pasyc)[NGSW + 1], double (*pomgc)[NGSW + 1], double *pbbfd, double *pbbfu,
                            double *pbbcd, double *pbbcu)
{
    const int klev = c->nlayers;
    const double od_lo = 0.06, repclc = 1.e-12;
    static __thread double zrefc[MAXLAY + 3], zrefdc[MAXLAY + 3], ztrac[MAXLAY + 3], ztradc[MAXLAY + 3];
    static __thread double zrefo[MAXLAY + 3], zrefdo[MAXLAY + 3], ztrao[MAXLAY + 3], ztrado[MAXLAY + 3];
    static __thread double zref[MAXLAY + 3], zrefd[MAXLAY + 3], ztra[MAXLAY + 3], ztrad[MAXLAY + 3];
    static __thread double zdbtc[MAXLAY + 3], ztdbtc[MAXLAY + 3], zdbt[MAXLAY + 3], ztdbt[MAXLAY + 3];
    static __thread double zrdndc[MAXLAY + 3], zrupc[MAXLAY + 3], zrupdc[MAXLAY + 3];
    static __thread double zrdnd[MAXLAY + 3], zrup[MAXLAY + 3], zrupd[MAXLAY + 3];
    static __thread double ztauc[MAXLAY + 3], zomcc[MAXLAY + 3], zgcc[MAXLAY + 3];
    static __thread double ztauo[MAXLAY + 3], zomco[MAXLAY + 3], zgco[MAXLAY + 3];
    static __thread double zcd[MAXLAY + 3], zcu[MAXLAY + 3], zfd[MAXLAY + 3], zfu[MAXLAY + 3];
    int lrtchkclr[MAXLAY + 3], lrtchkcld[MAXLAY + 3];
    for (int jk = 1; jk <= klev + 1; ++jk) pbbcd[jk] = pbbcu[jk] = pbbfd[jk] = pbbfu[jk] = 0.0;
    taumol_sw(c);
    int iw = 0;
    for (int jb = 16; jb <= 29; ++jb) {
        int ibm = jb - 15;
        int igt = S.ngc[ibm - 1];
        for (int jg = 1; jg <= igt; ++jg) {
            iw = iw + 1;
            double zincflx = adjflux[jb] * c->sfluxzen[iw] * prmu0;
            ztdbtc[1] = 1.0;
            zdbtc[klev + 1] = 0.0;
            ztrac[klev + 1] = 0.0;
            ztradc[klev + 1] = 0.0;
            zrefc[klev + 1] = palbp[ibm];
            zrefdc[klev + 1] = palbd[ibm];
            zrupc[klev + 1] = palbp[ibm];
            zrupdc[klev + 1] = palbd[ibm];
            ztrao[klev + 1] = 0.0;
            ztrado[klev + 1] = 0.0;
            zrefo[klev + 1] = palbp[ibm];
            zrefdo[klev + 1] = palbd[ibm];
            ztdbt[1] = 1.0;
            zdbt[klev + 1] = 0.0;
            ztra[klev + 1] = 0.0;
            ztrad[klev + 1] = 0.0;
            zref[klev + 1] = palbp[ibm];
            zrefd[klev + 1] = palbd[ibm];
            zrup[klev + 1] = palbp[ibm];
            zrupd[klev + 1] = palbd[ibm];
            for (int jk = 1; jk <= klev; ++jk) {
                int ikl = klev + 1 - jk;
                lrtchkclr[jk] = 1;
                lrtchkcld[jk] = pcf[ikl][iw] > repclc;
                /* ptaua = 0, pomga = 1, pasya = 0 (iaer = 0) */
                ztauc[jk] = c->taur[ikl][iw] + c->taug[ikl][iw] + 0.0;
                zomcc[jk] = c->taur[ikl][iw] * 1.0 + 0.0 * 1.0;
                zgcc[jk] = 0.0 * 1.0 * 0.0 / zomcc[jk];
                zomcc[jk] = zomcc[jk] / ztauc[jk];
                double zf = zgcc[jk] * zgcc[jk];
                double zwf = zomcc[jk] * zf;
                ztauc[jk] = (1.0 - zwf) * ztauc[jk];
                zomcc[jk] = (zomcc[jk] - zwf) / (1.0 - zwf);
                zgcc[jk] = (zgcc[jk] - zf) / (1.0 - zf);
                ztauo[jk] = ztauc[jk] + ptauc[ikl][iw];
                zomco[jk] = ztauc[jk] * zomcc[jk] + ptauc[ikl][iw] * pomgc[ikl][iw];
                zgco[jk] = (ptauc[ikl][iw] * pomgc[ikl][iw] * pasyc[ikl][iw] + ztauc[jk] * zomcc[jk] * zgcc[jk]) /
                           zomco[jk];
                zomco[jk] = zomco[jk] / ztauo[jk];
            }
            reftra_sw(klev, lrtchkclr, zgcc, prmu0, ztauc, zomcc, zrefc, zrefdc, ztrac, ztradc);
            reftra_sw(klev, lrtchkcld, zgco, prmu0, ztauo, zomco, zrefo, zrefdo, ztrao, ztrado);
            for (int jk = 1; jk <= klev; ++jk) {
                int ikl = klev + 1 - jk;
                double zclear = 1.0 - pcf[ikl][iw];
                double zcloud = pcf[ikl][iw];
                zref[jk] = zclear * zrefc[jk] + zcloud * zrefo[jk];
                zrefd[jk] = zclear * zrefdc[jk] + zcloud * zrefdo[jk];
                ztra[jk] = zclear * ztrac[jk] + zcloud * ztrao[jk];
                ztrad[jk] = zclear * ztradc[jk] + zcloud * ztrado[jk];
                double ze1 = ztauc[jk] / prmu0;
                double zdbtmc = ze1 <= od_lo ? 1.0 - ze1 + 0.5 * ze1 * ze1 : tbl_exp(ze1);
                zdbtc[jk] = zdbtmc;
                ztdbtc[jk + 1] = zdbtc[jk] * ztdbtc[jk];
                ze1 = ztauo[jk] / prmu0;
                double zdbtmo = ze1 <= od_lo ? 1.0 - ze1 + 0.5 * ze1 * ze1 : tbl_exp(ze1);
                zdbt[jk] = zclear * zdbtmc + zcloud * zdbtmo;
                ztdbt[jk + 1] = zdbt[jk] * ztdbt[jk];
            }
            vrtqdr_sw(klev, zrefc, zrefdc, ztrac, ztradc, zdbtc, zrdndc, zrupc, zrupdc, ztdbtc, zcd, zcu);
            vrtqdr_sw(klev, zref, zrefd, ztra, ztrad, zdbt, zrdnd, zrup, zrupd, ztdbt, zfd, zfu);
            for (int jk = 1; jk <= klev + 1; ++jk) {
                int ikl = klev + 2 - jk;
                pbbfu[ikl] = pbbfu[ikl] + zincflx * zfu[jk];
                pbbfd[ikl] = pbbfd[ikl] + zincflx * zfd[jk];
                pbbcu[ikl] = pbbcu[ikl] + zincflx * zcu[jk];
                pbbcd[ikl] = pbbcd[ikl] + zincflx * zcd[jk];
            }
        }
    }
}

static void sw_column_u(swcol_t *c, int iplon, int ncol, int nlay, const double *play,
                        const double *plev, const double *tlay, const double *tlev,
                        const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                        const double *ch4vmr, const double *n2ovmr, const double *o2vmr, int vmr)
{
    inatm_sw(c, iplon, ncol, nlay, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, vmr);
    setcoef_sw(c);
}

static void sw_column(swcol_t *c, int iplon, int ncol, int nlay, const double *play,
                      const double *plev, const double *tlay, const double *tlev,
                      const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                      const double *ch4vmr, const double *n2ovmr, const double *o2vmr)
{
    inatm_sw(c, iplon, ncol, nlay, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, 0);
    setcoef_sw(c);
}

/* earth_sun (rrtmg_sw_rad.nomcica.f90:734-758) */
static double earth_sun(int idn)
{
    const double pi = 2.0 * asin(1.0);
    double gamma = 2.0 * pi * (idn - 1) / 365.0;
    return 1.000110 + .034221 * cos(gamma) + .001289 * sin(gamma) + .000719 * cos(2.0 * gamma) +
           .000077 * sin(2.0 * gamma);
}

int oracle_rrtmg_sw(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                    const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                    const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                    const double *o2vmr, const double *asdir, const double *asdif,
                    const double *aldir, const double *aldif, const double *coszen, double adjes,
                    int dyofyr, double scon, double *swuflx, double *swdflx, double *swhr,
                    double *swuflxc, double *swdflxc, double *swhrc)
{
    (void)tsfc;
    if (!S.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    const double zepzen = 1.e-10;
    double adjflx = adjes;
    if (dyofyr > 0) adjflx = earth_sun(dyofyr);
    double adjflux[30];
    for (int ib = 16; ib <= 29; ++ib) adjflux[ib] = adjflx * (scon / 1.36822e+03);
    int err = 0;
#pragma omp parallel
    {
        swcol_t *c = (swcol_t *)malloc(sizeof(swcol_t));
        double zbbcu[MAXLAY + 3], zbbcd[MAXLAY + 3], swnflxc[MAXLAY + 3];
        if (!c) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (!c) continue;
#define O2(a, i, k) a[(i) + (long)ncol * ((k)-1)]
            if (coszen[iplon] < zepzen) {
                for (int k = 1; k <= nlay + 1; ++k) {
                    O2(swuflx, iplon, k) = 0.0;
                    O2(swdflx, iplon, k) = 0.0;
                    O2(swuflxc, iplon, k) = 0.0;
                    O2(swdflxc, iplon, k) = 0.0;
                }
                for (int k = 1; k <= nlay; ++k) {
                    O2(swhr, iplon, k) = 0.0;
                    O2(swhrc, iplon, k) = 0.0;
                }
                continue;
            }
            sw_column(c, iplon, ncol, nlay, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr,
                      n2ovmr, o2vmr);
            double cossza = coszen[iplon];
            if (cossza < zepzen) cossza = zepzen;
            double albdir[15], albdif[15];
            for (int ib = 1; ib <= 9; ++ib) {
                albdir[ib] = aldir[iplon];
                albdif[ib] = aldif[iplon];
            }
            albdir[14] = aldir[iplon];
            albdif[14] = aldif[iplon];
            for (int ib = 10; ib <= 13; ++ib) {
                albdir[ib] = asdir[iplon];
                albdif[ib] = asdif[iplon];
            }
            spcvrt_sw(c, cossza, albdif, albdir, adjflux, zbbcd, zbbcu);
            for (int i = 1; i <= nlay + 1; ++i) {
                O2(swuflxc, iplon, i) = zbbcu[i];
                O2(swdflxc, iplon, i) = zbbcd[i];
                O2(swuflx, iplon, i) = zbbcu[i];
                O2(swdflx, iplon, i) = zbbcd[i];
                swnflxc[i] = zbbcd[i] - zbbcu[i];
            }
            for (int i = 1; i <= nlay; ++i) {
                double zdpgcp = S.heatfac / c->pdp[i];
                O2(swhrc, iplon, i) = (swnflxc[i + 1] - swnflxc[i]) * zdpgcp;
                O2(swhr, iplon, i) = O2(swhrc, iplon, i);
            }
            O2(swhrc, iplon, nlay) = 0.0;
            O2(swhr, iplon, nlay) = 0.0;
#undef O2
        }
        free(c);
    }
    return err;
}

/* rrtmg_sw with clouds (rrtmg_sw_rad.nomcica.f90:78-731): icld 1..3 (out of
 * range -> 2; the nomcica path ignores the overlap choice), inflgsw = 0,
 * iaer = 0.  cldfr, cicewp, cliqwp (ncol, nlay) (NULL water paths = 0);
 * taucld, ssacld, asmcld, fsfcld (14, ncol, nlay), band fastest.  Returns -5
 * when a day column has a fraction in (1e-6, 1 - 1e-6) (the reference
 * stops: partial cloud is not allowed, :535-539). */
static int sw_cloud_driver(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *asdir, const double *asdif,
                          const double *aldir, const double *aldif, const double *coszen, double adjes,
                          int dyofyr, double scon, const double *cldfr, const double *taucld,
                          const double *ssacld, const double *asmcld, const double *fsfcld,
                          const double *cicewp, const double *cliqwp, const double *reice, const double *reliq, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc)
{
    if (!S.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    if (icld < 0 || icld > 3) icld = 2;
    const int clouds = icld != 0;
    const double zepzen = 1.e-10, zepsec = 1.e-6;
    double adjflx = adjes;
    if (dyofyr > 0) adjflx = earth_sun(dyofyr);
    double adjflux[30];
    for (int ib = 16; ib <= 29; ++ib) adjflux[ib] = adjflx * (scon / 1.36822e+03);
    int err = 0;
#pragma omp parallel
    {
        swcol_t *c = (swcol_t *)malloc(sizeof(swcol_t));
        double (*a)[MAXLAY + 2][15] = malloc(sizeof(double[7][MAXLAY + 2][15]));
        double (*gq)[MAXLAY + 2][NGSW + 1] = malloc(sizeof(double[4][MAXLAY + 2][NGSW + 1]));
        double cf[MAXLAY + 3], cwp[MAXLAY + 3], ciw[MAXLAY + 3], clw[MAXLAY + 3], rei[MAXLAY + 3], rel[MAXLAY + 3];
        double zbbcu[MAXLAY + 3], zbbcd[MAXLAY + 3], zbbfu[MAXLAY + 3], zbbfd[MAXLAY + 3];
        double swnflxc[MAXLAY + 3], swnflx[MAXLAY + 3];
        if (!c || !a || !gq) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (err == -3) continue;
#define O2(x, i, k) x[(i) + (long)ncol * ((k)-1)]
            if (coszen[iplon] < zepzen) {
                for (int k = 1; k <= nlay + 1; ++k)
                    O2(swuflx, iplon, k) = O2(swdflx, iplon, k) = O2(swuflxc, iplon, k) = O2(swdflxc, iplon, k) = 0.0;
                for (int k = 1; k <= nlay; ++k) O2(swhr, iplon, k) = O2(swhrc, iplon, k) = 0.0;
                continue;
            }
            double (*tauc)[15] = a[0], (*ssac)[15] = a[1], (*asmc)[15] = a[2], (*fsfc)[15] = a[3];
            double (*ptauc)[15] = a[4], (*pomgc)[15] = a[5], (*pasyc)[15] = a[6];
            int partial = 0;
            for (int l = 1; l <= nlay; ++l) {
                long ix = iplon + (long)ncol * (l - 1);
                cf[l] = clouds ? cldfr[ix] : 0.0;
                cwp[l] = clouds ? (cicewp ? cicewp[ix] : 0.0) + (cliqwp ? cliqwp[ix] : 0.0) : 0.0;
                ciw[l] = clouds && cicewp ? cicewp[ix] : 0.0;
                clw[l] = clouds && cliqwp ? cliqwp[ix] : 0.0;
                rei[l] = clouds && reice ? reice[ix] : 0.0;
                rel[l] = clouds && reliq ? reliq[ix] : 0.0;
                for (int ib = 1; ib <= 14; ++ib) {
                    long jx = (ib - 1) + 14 * ix;
                    tauc[l][ib] = clouds && taucld ? taucld[jx] : 0.0;
                    ssac[l][ib] = clouds && ssacld ? ssacld[jx] : 0.0;
                    asmc[l][ib] = clouds && asmcld ? asmcld[jx] : 0.0;
                    fsfc[l][ib] = clouds && fsfcld ? fsfcld[jx] : 0.0;
                }
                if (cf[l] > zepsec && cf[l] < 1.0 - zepsec) partial = 1;
            }
            if (partial) {
#pragma omp atomic write
                err = -5;
                continue;
            }
            if (inflag == 2 && clouds) {
                if (cldprop_sw2(nlay, iceflag, liqflag, cf, tauc, ciw, clw, rei, rel, ptauc, pomgc, pasyc) != 0) {
#pragma omp atomic write
                    err = -8;
                    continue;
                }
            } else if (inflag == 1 && clouds) {
                /* cldprop_sw has no inflag 1 branch: every layer keeps depth 0,
                 * albedo 1, asymmetry 0 (rrtmg_sw_cldprop.f90:126-133) */
                for (int l = 1; l <= nlay; ++l)
                    for (int ib = 1; ib <= 14; ++ib) ptauc[l][ib] = 0.0, pomgc[l][ib] = 1.0, pasyc[l][ib] = 0.0;
            } else {
                cldprop_sw0(nlay, cf, tauc, ssac, asmc, fsfc, cwp, ptauc, pomgc, pasyc);
            }
            sw_column(c, iplon, ncol, nlay, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr);
            double cossza = coszen[iplon];
            double albdir[15], albdif[15];
            for (int ib = 1; ib <= 9; ++ib) {
                albdir[ib] = aldir[iplon];
                albdif[ib] = aldif[iplon];
            }
            albdir[14] = aldir[iplon];
            albdif[14] = aldif[iplon];
            for (int ib = 10; ib <= 13; ++ib) {
                albdir[ib] = asdir[iplon];
                albdif[ib] = asdif[iplon];
            }
            for (int l = 1; l <= nlay; ++l)
                for (int ib = 1, iw = 1; ib <= 14; ++ib)
                    for (int jg = 1; jg <= S.ngc[ib - 1]; ++jg, ++iw) {
                        gq[0][l][iw] = cf[l];
                        gq[1][l][iw] = ptauc[l][ib];
                        gq[2][l][iw] = pasyc[l][ib];
                        gq[3][l][iw] = pomgc[l][ib];
                    }
            spcvrt_sw_cloud(c, cossza, albdif, albdir, adjflux, gq[0], gq[1], gq[2], gq[3], zbbfd, zbbfu, zbbcd,
                            zbbcu);
            for (int i = 1; i <= nlay + 1; ++i) {
                O2(swuflxc, iplon, i) = zbbcu[i];
                O2(swdflxc, iplon, i) = zbbcd[i];
                O2(swuflx, iplon, i) = zbbfu[i];
                O2(swdflx, iplon, i) = zbbfd[i];
                swnflxc[i] = zbbcd[i] - zbbcu[i];
                swnflx[i] = zbbfd[i] - zbbfu[i];
            }
            for (int i = 1; i <= nlay; ++i) {
                double zdpgcp = S.heatfac / c->pdp[i];
                O2(swhrc, iplon, i) = (swnflxc[i + 1] - swnflxc[i]) * zdpgcp;
                O2(swhr, iplon, i) = (swnflx[i + 1] - swnflx[i]) * zdpgcp;
            }
            O2(swhrc, iplon, nlay) = 0.0;
            O2(swhr, iplon, nlay) = 0.0;
#undef O2
        }
        free(c);
        free(a);
        free(gq);
    }
    return err;
}

int oracle_rrtmg_sw_cloud(int ncol, int nlay, int icld, const double *play, const double *plev,
                          const double *tlay, const double *tlev, const double *h2ovmr, const double *o3vmr,
                          const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                          const double *o2vmr, const double *asdir, const double *asdif,
                          const double *aldir, const double *aldif, const double *coszen, double adjes,
                          int dyofyr, double scon, const double *cldfr, const double *taucld,
                          const double *ssacld, const double *asmcld, const double *fsfcld,
                          const double *cicewp, const double *cliqwp, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc)
{
    return sw_cloud_driver(ncol, nlay, icld, 0, 0, 0, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                           o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, cldfr, taucld, ssacld,
                           asmcld, fsfcld, cicewp, cliqwp, NULL, NULL, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc);
}

int oracle_rrtmg_sw_cloud_wp(int ncol, int nlay, int icld, int inflag, int iceflag, int liqflag, const double *play,
                             const double *plev, const double *tlay, const double *tlev, const double *h2ovmr,
                             const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                             const double *n2ovmr, const double *o2vmr, const double *asdir, const double *asdif,
                             const double *aldir, const double *aldif, const double *coszen, double adjes,
                             int dyofyr, double scon, const double *cldfr, const double *taucld,
                             const double *cicewp, const double *cliqwp, const double *reice, const double *reliq,
                             double *swuflx, double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                             double *swhrc)
{
    if (inflag == 2 && !SC.ready) return -1;
    return sw_cloud_driver(ncol, nlay, icld, inflag, iceflag, liqflag, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr,
                           ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, cldfr,
                           taucld, NULL, NULL, NULL, cicewp, cliqwp, reice, reliq, swuflx, swdflx, swhr, swuflxc,
                           swdflxc, swhrc);
}

/* mcica_subcol_sw with the KISS generator (irng = 0;
 * mcica_subcol_gen_sw.f90:47-159, generate_stochastic_clouds_sw :163-536):
 * the mask of oracle_mcica_mask over 112 subcolumns; a cloudy subcolumn
 * takes its band's in-cloud optics and the layer's water paths, a clear one
 * depth 0, albedo 1, asymmetry 0 and forward fraction 0 (:477-510).  Inputs
 * cldfrac, ciwp, clwp (ncol, nlay); tauc, ssac, asmc, fsfc (14, ncol, nlay);
 * outputs (112, ncol, nlay).  -6, -7 as oracle_lw_mcica_subcol. */
/* set while oracle_sw_mcica_subcol_mt runs the generator on a Mersenne
 * Twister stream */
static int sw_mt;

int oracle_sw_mcica_subcol(int ncol, int nlay, int icld, int permuteseed, const double *play,
                           const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                           const double *ssac, const double *asmc, const double *fsfc, double *cldfmcl,
                           double *ciwpmcl, double *clwpmcl, double *taucmcl, double *ssacmcl, double *asmcmcl,
                           double *fsfcmcl)
{
    if (!S.ready) return -1;
    if (nlay < (sw_mt ? 2 : 4) || nlay > MAXLAY) return -7;
    if (icld < 1 || icld > 3) return -2;
    for (int i = 0; i < ncol && !sw_mt; ++i)
        if (play[i] * 1.e2 < play[i + ncol] * 1.e2) return -6;
    int band_of[NGSW];
    for (int b = 0, ig = 0; b < 14; ++b)
        for (int k = 0; k < S.ngc[b]; ++k) band_of[ig++] = b;
    static __thread unsigned char mask[NGSW * MAXLAY];
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < ncol; ++i) {
        oracle_mcica_mask(i, ncol, nlay, icld, permuteseed, NGSW, play, cldfrac, mask);
        for (int l = 0; l < nlay; ++l) {
            long ix = i + (long)ncol * l;
            for (int ig = 0; ig < NGSW; ++ig) {
                long o = ig + (long)NGSW * ix, j = band_of[ig] + 14 * ix;
                int cloudy = mask[ig * MAXLAY + l];
                cldfmcl[o] = cloudy ? 1.0 : 0.0;
                clwpmcl[o] = cloudy ? clwp[ix] : 0.0;
                ciwpmcl[o] = cloudy ? ciwp[ix] : 0.0;
                taucmcl[o] = cloudy ? tauc[j] : 0.0;
                ssacmcl[o] = cloudy ? ssac[j] : 1.0;
                asmcmcl[o] = cloudy ? asmc[j] : 0.0;
                fsfcmcl[o] = cloudy ? fsfc[j] : 0.0;
            }
        }
    }
    return 0;
}

/* oracle_sw_mcica_subcol with the Mersenne Twister (irng = 1, seeded with
 * permuteseed; mcica_subcol_gen_sw.f90 as the LW generator). */
int oracle_sw_mcica_subcol_mt(int ncol, int nlay, int icld, int permuteseed, const double *play,
                              const double *cldfrac, const double *ciwp, const double *clwp, const double *tauc,
                              const double *ssac, const double *asmc, const double *fsfc, double *cldfmcl,
                              double *ciwpmcl, double *clwpmcl, double *taucmcl, double *ssacmcl, double *asmcmcl,
                              double *fsfcmcl)
{
    if (icld < 1 || icld > 3) return -2;
    long n = (long)NGSW * ncol * (icld == 3 ? 1 : nlay);
    uint32_t *mt = malloc(sizeof(uint32_t) * (size_t)n);
    if (!mt) return -3;
    oracle_mt_stream((uint32_t)permuteseed, n, mt);
    oracle_mcica_set_stream(mt);
    sw_mt = 1;
    int rc = oracle_sw_mcica_subcol(ncol, nlay, icld, permuteseed, play, cldfrac, ciwp, clwp, tauc, ssac, asmc,
                                    fsfc, cldfmcl, ciwpmcl, clwpmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl);
    sw_mt = 0;
    oracle_mcica_set_stream(NULL);
    free(mt);
    return rc;
}

/* rrtmg_sw of the McICA driver (rrtmg_sw_rad.f90:80-745; iaer = 0,
 * inflgsw = 0): inatm with h2o and O3 as volume mixing ratios; no night
 * early exit, the zenith cosine floored at 1e-10 instead (:571-572);
 * cldprmc_sw (rrtmg_sw_cldprmc.f90:112-143): a subcolumn whose fraction
 * and water path or depth reach 1e-20 has its optics delta-M scaled by its
 * forward fraction, any other keeps its input optics; spcvmc_sw on them.
 * Subcolumn inputs (112, ncol, nlay) (NULL water paths = 0). */
static int sw_mcica_driver(int ncol, int nlay, int iceflag, int liqflag, const double *reicmcl,
                           const double *relqmcl, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                          const double *ch4vmr, const double *n2ovmr, const double *o2vmr, const double *asdir,
                          const double *asdif, const double *aldir, const double *aldif, const double *coszen,
                          double adjes, int dyofyr, double scon, const double *cldfmcl, const double *taucmcl,
                          const double *ssacmcl, const double *asmcmcl, const double *fsfcmcl,
                          const double *ciwpmcl, const double *clwpmcl, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc)
{
    if (!S.ready) return -1;
    if (nlay < 2 || nlay > MAXLAY) return -2;
    const double zepzen = 1.e-10, cldmin = 1.e-20;
    double adjflx = adjes;
    if (dyofyr > 0) adjflx = earth_sun(dyofyr);
    double adjflux[30];
    for (int ib = 16; ib <= 29; ++ib) adjflux[ib] = adjflx * (scon / 1.36822e+03);
    int err = 0;
#pragma omp parallel
    {
        swcol_t *c = (swcol_t *)malloc(sizeof(swcol_t));
        double (*gq)[MAXLAY + 2][NGSW + 1] = malloc(sizeof(double[4][MAXLAY + 2][NGSW + 1]));
        double zbbcu[MAXLAY + 3], zbbcd[MAXLAY + 3], zbbfu[MAXLAY + 3], zbbfd[MAXLAY + 3];
        double swnflxc[MAXLAY + 3], swnflx[MAXLAY + 3];
        int bad = 0;
        if (!c || !gq) {
#pragma omp atomic write
            err = -3;
        }
#pragma omp for schedule(dynamic, 16)
        for (int iplon = 0; iplon < ncol; ++iplon) {
            if (err == -3) continue;
#define O2(x, i, k) x[(i) + (long)ncol * ((k)-1)]
            for (int l = 1; l <= nlay; ++l) {
                long ix = iplon + (long)ncol * (l - 1);
                for (int iw = 1; iw <= NGSW; ++iw) {
                    long o = (iw - 1) + (long)NGSW * ix;
                    double cf = cldfmcl[o], tau = taucmcl[o], ssa = ssacmcl[o], asm_ = asmcmcl[o];
                    double cwp = (ciwpmcl ? ciwpmcl[o] : 0.0) + (clwpmcl ? clwpmcl[o] : 0.0);
                    if (cf >= cldmin && (cwp >= cldmin || tau >= cldmin)) {
                        if (iceflag > 0) {
                            /* cldprmc_sw with inflag 2 (rrtmg_sw_cldprmc.f90:144-330):
                             * cldprop_sw's optics per subcolumn, band ngb(ig) */
                            int ib = 0;
                            for (int acc = S.ngc[0]; iw > acc; acc += S.ngc[++ib]) {
                            }
                            double ci = ciwpmcl ? ciwpmcl[o] : 0.0, cw = clwpmcl ? clwpmcl[o] : 0.0;
                            double r[3] = {0.0, 1.0, 0.0};
                            if (sw_wp_optics(ib, iceflag, liqflag, ci, cw, reicmcl ? reicmcl[ix] : 0.0,
                                             relqmcl ? relqmcl[ix] : 0.0, r) != 0)
                                bad = 1;
                            tau = r[0];
                            ssa = r[1];
                            asm_ = r[2];
                        } else {
                            double ffp = fsfcmcl[o], ffp1 = 1.0 - ffp, ffpssa = 1.0 - ffp * ssa;
                            ssa = ffp1 * ssa / ffpssa;
                            tau = ffpssa * tau;
                            asm_ = (asm_ - ffp) / ffp1;
                        }
                    }
                    gq[0][l][iw] = cf;
                    gq[1][l][iw] = tau;
                    gq[2][l][iw] = asm_;
                    gq[3][l][iw] = ssa;
                }
            }
            if (bad) {
#pragma omp atomic write
                err = -8;
                continue;
            }
            sw_column_u(c, iplon, ncol, nlay, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                        1);
            double cossza = coszen[iplon];
            if (cossza < zepzen) cossza = zepzen;
            double albdir[15], albdif[15];
            for (int ib = 1; ib <= 9; ++ib) {
                albdir[ib] = aldir[iplon];
                albdif[ib] = aldif[iplon];
            }
            albdir[14] = aldir[iplon];
            albdif[14] = aldif[iplon];
            for (int ib = 10; ib <= 13; ++ib) {
                albdir[ib] = asdir[iplon];
                albdif[ib] = asdif[iplon];
            }
            spcvrt_sw_cloud(c, cossza, albdif, albdir, adjflux, gq[0], gq[1], gq[2], gq[3], zbbfd, zbbfu, zbbcd,
                            zbbcu);
            for (int i = 1; i <= nlay + 1; ++i) {
                O2(swuflxc, iplon, i) = zbbcu[i];
                O2(swdflxc, iplon, i) = zbbcd[i];
                O2(swuflx, iplon, i) = zbbfu[i];
                O2(swdflx, iplon, i) = zbbfd[i];
                swnflxc[i] = zbbcd[i] - zbbcu[i];
                swnflx[i] = zbbfd[i] - zbbfu[i];
            }
            for (int i = 1; i <= nlay; ++i) {
                double zdpgcp = S.heatfac / c->pdp[i];
                O2(swhrc, iplon, i) = (swnflxc[i + 1] - swnflxc[i]) * zdpgcp;
                O2(swhr, iplon, i) = (swnflx[i + 1] - swnflx[i]) * zdpgcp;
            }
            O2(swhrc, iplon, nlay) = 0.0;
            O2(swhr, iplon, nlay) = 0.0;
#undef O2
        }
        free(c);
        free(gq);
    }
    return err;
}

int oracle_rrtmg_sw_mcica(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                          const double *tlev, const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                          const double *ch4vmr, const double *n2ovmr, const double *o2vmr, const double *asdir,
                          const double *asdif, const double *aldir, const double *aldif, const double *coszen,
                          double adjes, int dyofyr, double scon, const double *cldfmcl, const double *taucmcl,
                          const double *ssacmcl, const double *asmcmcl, const double *fsfcmcl,
                          const double *ciwpmcl, const double *clwpmcl, double *swuflx, double *swdflx,
                          double *swhr, double *swuflxc, double *swdflxc, double *swhrc)
{
    return sw_mcica_driver(ncol, nlay, 0, 0, NULL, NULL, play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, ch4vmr,
                           n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, cldfmcl, taucmcl,
                           ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, swuflx, swdflx, swhr, swuflxc, swdflxc,
                           swhrc);
}

int oracle_rrtmg_sw_mcica_wp(int ncol, int nlay, int iceflag, int liqflag, const double *play, const double *plev,
                             const double *tlay, const double *tlev, const double *h2ovmr, const double *o3vmr,
                             const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                             const double *o2vmr, const double *asdir, const double *asdif, const double *aldir,
                             const double *aldif, const double *coszen, double adjes, int dyofyr, double scon,
                             const double *cldfmcl, const double *taucmcl, const double *ssacmcl,
                             const double *asmcmcl, const double *fsfcmcl, const double *ciwpmcl,
                             const double *clwpmcl, const double *reicmcl, const double *relqmcl, double *swuflx,
                             double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc)
{
    if (!SC.ready || iceflag < 1 || iceflag > 3) return -1;
    return sw_mcica_driver(ncol, nlay, iceflag, liqflag, reicmcl, relqmcl, play, plev, tlay, tlev, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon,
                           cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, swuflx, swdflx, swhr,
                           swuflxc, swdflxc, swhrc);
}

/* ---- single-column intermediates ---- */

static void one_column(swcol_t *c, int nlay, const double *play, const double *plev,
                       const double *tlay, const double *h2ovmr, const double *o3vmr,
                       const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr)
{
    sw_column(c, 0, 1, nlay, play, plev, tlay, plev /* tlev unused by SW */, h2ovmr, o3vmr, co2vmr,
              ch4vmr, n2ovmr, o2vmr);
}

int oracle_sw_setcoef(int nlay, const double *play, const double *plev, const double *tlay,
                      const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                      const double *ch4vmr, const double *n2ovmr, const double *o2vmr, int *jp,
                      int *jt, int *jt1, double *fac00, double *fac01, double *fac10, double *fac11,
                      double *colmol)
{
    if (!S.ready || nlay > MAXLAY) return -1;
    swcol_t *c = (swcol_t *)malloc(sizeof(swcol_t));
    one_column(c, nlay, play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr);
    for (int k = 1; k <= nlay; ++k) {
        jp[k - 1] = c->jp[k];
        jt[k - 1] = c->jt[k];
        jt1[k - 1] = c->jt1[k];
        fac00[k - 1] = c->fac00[k];
        fac01[k - 1] = c->fac01[k];
        fac10[k - 1] = c->fac10[k];
        fac11[k - 1] = c->fac11[k];
        colmol[k - 1] = c->colmol[k];
    }
    int lt = c->laytrop;
    free(c);
    return lt;
}

int oracle_sw_taumol(int nlay, const double *play, const double *plev, const double *tlay,
                     const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                     const double *ch4vmr, const double *n2ovmr, const double *o2vmr, double *taug,
                     double *taur, double *sfluxzen)
{
    if (!S.ready || nlay > MAXLAY) return -1;
    swcol_t *c = (swcol_t *)malloc(sizeof(swcol_t));
    one_column(c, nlay, play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr);
    taumol_sw(c);
    for (int k = 1; k <= nlay; ++k)
        for (int ig = 1; ig <= NGSW; ++ig) {
            taug[(ig - 1) * nlay + (k - 1)] = c->taug[k][ig];
            taur[(ig - 1) * nlay + (k - 1)] = c->taur[k][ig];
        }
    for (int ig = 1; ig <= NGSW; ++ig) sfluxzen[ig - 1] = c->sfluxzen[ig];
    int lt = c->laytrop;
    free(c);
    return lt;
}

double oracle_heatfac_sw(void) { return S.heatfac; }
