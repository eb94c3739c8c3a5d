/* Oracle bookkeeping shared by the SW and LW restatements (TEST
 * INFRASTRUCTURE ONLY; see oracle.h). */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NREG 512
static struct {
    char name[48];
    const double *p;
    long n;
} reg[NREG];
static int nreg;

void oracle_register(const char *name, const double *p, long n)
{
    for (int i = 0; i < nreg; ++i)
        if (!strncmp(reg[i].name, name, 47)) {
            reg[i].p = p;
            reg[i].n = n;
            return;
        }
    if (nreg >= NREG) return;
    strncpy(reg[nreg].name, name, 47);
    reg[nreg].p = p;
    reg[nreg].n = n;
    ++nreg;
}

long oracle_reduced(const char *name, double *out, long n)
{
    for (int i = 0; i < nreg; ++i)
        if (!strncmp(reg[i].name, name, 47)) {
            long m = reg[i].n < n ? reg[i].n : n;
            if (out) memcpy(out, reg[i].p, sizeof(double) * (size_t)m);
            return reg[i].n;
        }
    return -1;
}

int oracle_set_threads(int n)
{
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
#else
    (void)n;
    return 1;
#endif
}
