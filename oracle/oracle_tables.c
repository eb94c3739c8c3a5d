/*
 * Oracle table loading and g-point reduction (TEST INFRASTRUCTURE ONLY; see
 * oracle.h).  Reads the RRTMGTB1 files written by
 * mima_amd/tables/extract_tables.py and restates the reference reduction:
 *   rwgt:   rrtmg_sw_init.f90:106-130, rrtmg_lw_init.f90:126-154
 *   cmbgb*: rrtmg_sw_init.f90:473-1516, rrtmg_lw_init.f90:366-2015
 * In every cmbgb routine the absorption-type arrays (k, self, foreign,
 * minor, cross-section, Rayleigh) are reduced as sum(k*rwgt) over the
 * original g-points of a reduced g-point, and the source/fraction arrays
 * (sfluxref, fracref) as a plain sum.
 */
#include "oracle_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    char name[48];
    unsigned dtype, ndim;
    int lb[4], ext[4];
    unsigned long long off, nbytes;
} tbl_entry_t;

int otbl_load(otbl_t *t, const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    t->raw = (unsigned char *)malloc((size_t)sz);
    if (!t->raw || fread(t->raw, 1, (size_t)sz, f) != (size_t)sz) {
        fclose(f);
        return -2;
    }
    fclose(f);
    if (memcmp(t->raw, "RRTMGTB1", 8) != 0) return -3;
    memcpy(&t->n, t->raw + 8, 4);
    memcpy(&t->flags, t->raw + 12, 4);
    t->size = sz;
    return 0;
}

void otbl_free(otbl_t *t)
{
    free(t->raw);
    t->raw = NULL;
}

/* Entry lookup: returns pointer to the data (double or int) and fills dims. */
const void *otbl_get(const otbl_t *t, const char *name, int *ndim, int *ext, int *lb, int *is_int)
{
    const size_t esz = 48 + 4 + 4 + 16 + 16 + 8 + 8;
    for (unsigned i = 0; i < t->n; ++i) {
        const unsigned char *e = t->raw + 16 + i * esz;
        if (strncmp((const char *)e, name, 48) != 0) continue;
        tbl_entry_t te;
        memcpy(te.name, e, 48);
        memcpy(&te.dtype, e + 48, 4);
        memcpy(&te.ndim, e + 52, 4);
        memcpy(te.lb, e + 56, 16);
        memcpy(te.ext, e + 72, 16);
        memcpy(&te.off, e + 88, 8);
        memcpy(&te.nbytes, e + 96, 8);
        if (ndim) *ndim = (int)te.ndim;
        for (int d = 0; d < 4; ++d) {
            if (ext) ext[d] = te.ext[d];
            if (lb) lb[d] = te.lb[d];
        }
        if (is_int) *is_int = (int)te.dtype;
        return t->raw + te.off;
    }
    return NULL;
}

const double *otbl_d(const otbl_t *t, const char *name, long *count)
{
    int nd, ext[4], isint;
    const void *p = otbl_get(t, name, &nd, ext, NULL, &isint);
    if (!p || isint) return NULL;
    if (count) *count = (long)ext[0] * ext[1] * ext[2] * ext[3];
    return (const double *)p;
}

const int *otbl_i(const otbl_t *t, const char *name, long *count)
{
    int nd, ext[4], isint;
    const void *p = otbl_get(t, name, &nd, ext, NULL, &isint);
    if (!p || !isint) return NULL;
    if (count) *count = (long)ext[0] * ext[1] * ext[2] * ext[3];
    return (const int *)p;
}

/* rwgt for all bands (rrtmg_sw_init.f90:106-130 / rrtmg_lw_init.f90:126-154).
 * nb bands of 16 original g-points; ngc/ngn/ngm as in the reference. */
void oracle_make_rwgt(int nb, const int *ngc, const int *ngn, const int *ngm,
                      const double *wt, double *rwgt)
{
    int igcsm = 0;
    for (int ibnd = 1; ibnd <= nb; ++ibnd) {
        int iprsm = 0;
        double wtsm[16];
        if (ngc[ibnd - 1] < 16) {
            for (int igc = 1; igc <= ngc[ibnd - 1]; ++igc) {
                igcsm += 1;
                double wtsum = 0.0;
                for (int ipr = 1; ipr <= ngn[igcsm - 1]; ++ipr) {
                    iprsm += 1;
                    wtsum = wtsum + wt[iprsm - 1];
                }
                wtsm[igc - 1] = wtsum;
            }
            for (int ig = 1; ig <= 16; ++ig) {
                int ind = (ibnd - 1) * 16 + ig;
                rwgt[ind - 1] = wt[ig - 1] / wtsm[ngm[ind - 1] - 1];
            }
        } else {
            for (int ig = 1; ig <= 16; ++ig) {
                igcsm += 1;
                rwgt[(ibnd - 1) * 16 + ig - 1] = 1.0;
            }
        }
    }
}

/* Reduce an original-resolution array whose g-point dimension (16 entries)
 * is `gdim` (0 = first, 1 = last) to ngc reduced g-points.  weighted != 0:
 * sum(x*rwgt) (the cmbgb treatment of k-type arrays); 0: plain sum (sfluxref,
 * fracref).  `first_ngn` points at ngn(ngs(band-1)+1), `rw` at
 * rwgt((band-1)*16+1).  Returns a malloc'ed array with the same layout and
 * the g dimension shrunk to ngc. */
double *oracle_reduce(const double *x, int nd, const int *ext, int gdim, int ngc,
                      const int *first_ngn, const double *rw, int weighted)
{
    long total = 1;
    for (int d = 0; d < nd; ++d) total *= ext[d];
    long other = total / 16;
    double *out = (double *)malloc(sizeof(double) * (size_t)(other * ngc));
    for (long o = 0; o < other; ++o) {
        int iprsm = 0;
        for (int igc = 1; igc <= ngc; ++igc) {
            double s = 0.0;
            for (int ipr = 1; ipr <= first_ngn[igc - 1]; ++ipr) {
                iprsm += 1;
                double v = gdim == 1 ? x[(long)(iprsm - 1) * other + o] : x[o * 16 + (iprsm - 1)];
                if (weighted)
                    s = s + v * rw[iprsm - 1];
                else
                    s = s + v;
            }
            if (gdim == 1)
                out[(long)(igc - 1) * other + o] = s;
            else
                out[o * ngc + (igc - 1)] = s;
        }
    }
    return out;
}
