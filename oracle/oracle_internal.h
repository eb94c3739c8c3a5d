/* Internal declarations of the CPU oracle (TEST INFRASTRUCTURE ONLY). */
#ifndef RRTMG_ORACLE_INTERNAL_H
#define RRTMG_ORACLE_INTERNAL_H

#include "oracle.h"

typedef struct {
    unsigned char *raw;
    long size;
    unsigned n, flags;
} otbl_t;

int otbl_load(otbl_t *t, const char *path);
void otbl_free(otbl_t *t);
const void *otbl_get(const otbl_t *t, const char *name, int *ndim, int *ext, int *lb, int *is_int);
const double *otbl_d(const otbl_t *t, const char *name, long *count);
const int *otbl_i(const otbl_t *t, const char *name, long *count);
void oracle_make_rwgt(int nb, const int *ngc, const int *ngn, const int *ngm,
                      const double *wt, double *rwgt);
double *oracle_reduce(const double *x, int nd, const int *ext, int gdim, int ngc,
                      const int *first_ngn, const double *rw, int weighted);

/* reduced-table registry for oracle_reduced() */
void oracle_register(const char *name, const double *p, long n);

#define ORACLE_NTBL 10000

#endif
