// Longwave RRTMG with clouds on gfx950: rrtmg_lw with icld = 1..3 and
// inflglw = 0 (rrtmg_lw/gcm_model/src/rrtmg_lw_rad.nomcica.f90:458-540):
// cldprop's in-cloud optical depths (rrtmg_lw_cldprop.f90:154-176), then
// rtrn's random overlap (icld = 1, rrtmg_lw_rtrn.f90) or rtrnmr's
// maximum-random overlap (icld = 2, 3, rrtmg_lw_rtrnmr.f90), with the
// clear-sky fluxes beside the total-sky ones.
//
// The clear-sky forward's prep and column kernels supply the gas-optics
// coefficients.  lw_cloud_column_kernel (one thread per column, FP64) runs
// cldprop's cloud test, rtrnmr's cloudy-layer test and its two overlap-factor
// sweeps, which carry state from layer to layer, and stores one record per
// (layer, column).  The band kernels follow the clear-sky band kernels'
// decomposition (one wave per work unit, lane = (column, g); FP32 with the
// band's k-tables in LDS in persistent workgroups, FP64 one wave per
// workgroup) with the cloudy branches of the transfer; a cloudy layer keeps
// four per-layer terms for the upward sweep instead of two, in a global
// scratch row per wave, and each level sums four radiances over g-points
// (total and clear, up and down).  The cloud path is separate from the clear-sky kernels, whose code
// it leaves unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"
#include "lw_optics.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// cloud column kernel: flags and overlap factors of every layer of a column
// (rrtmg_lw_cldprop.f90:154-176, rrtmg_lw_rtrnmr.f90:314-476), FP64.
// The reference's factor arrays are indexed by level; each cloudy layer
// writes the upward factors of index lay+1 and the downward ones of index
// lay-1, which only that layer's iteration writes and only that layer's
// transfer reads, so they are stored with the layer.  The reference also
// reads faccld2/facclr2 of index lay (written by the layer below when it is
// cloudy, zero otherwise) for the combined factors; `c2`, `l2` carry them.
// cldfrac(0) (read at lay = 1 only as a factor of zero) and
// cldfrac(nlay+1) are taken as 0, the latter as the driver sets it
// (rrtmg_lw_rad.nomcica.f90:736).
template <typename Real>
__global__ __launch_bounds__(64) void lw_cloud_column_kernel(int ncol, int nlay, LwClouds cl, LwCloudWork cw)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    Real *fac = (Real *)cw.fac;
    auto cf = [&](int lay) { return lay < 1 || lay > nlay ? 0.0 : cl.cldfr[(size_t)(lay - 1) * ncol + col]; };
    auto cloudy = [&](int lay) { return lay >= 1 && lay <= nlay && cf(lay) >= 1.e-6; };
    auto put = [&](int lay, int f, double v) { fac[((size_t)(lay - 1) * ncol + col) * LWC_STRIDE + f] = (Real)v; };
    for (int lay = 1; lay <= nlay; ++lay) {
        const size_t ix = (size_t)(lay - 1) * ncol + col;
        double tauctot = 0.0;
        for (int b = 0; b < 16; ++b) tauctot = tauctot + cl.taucld[ix * 16 + b];
        const double cwp = (cl.cicewp ? cl.cicewp[ix] : 0.0) + (cl.cliqwp ? cl.cliqwp[ix] : 0.0);
        const double f = cf(lay);
        int fl = 0;
        if (cloudy(lay)) fl |= LWC_CLOUDY;
        if (!cloudy(lay - 1)) fl |= LWC_TOP_UP;
        if (!cloudy(lay + 1)) fl |= LWC_TOP_DN;
        if (cl.slant || (f >= 1.e-20 && (cwp >= 1.e-20 || tauctot >= 1.e-20))) fl |= LWC_TAU;
        cw.flags[ix] = fl;
        put(lay, LWC_CF, f);
        for (int k = LWC_CLD1; k < LWC_COUNT; ++k) put(lay, k, 0.0);
    }
    if (cl.icld == 1) return;
    // upward sweep (rrtmg_lw_rtrnmr.f90:329-399)
    double rat1 = 0.0, rat2 = 0.0, c2 = 0.0, l2 = 0.0;
    for (int lev = 1; lev <= nlay; ++lev) {
        if (!cloudy(lev)) {
            c2 = l2 = 0.0;
            continue;
        }
        const bool top = !cloudy(lev - 1);
        const double fd = cf(lev - 1), f0 = cf(lev), fu = cf(lev + 1);
        double cld1 = 0.0, cld2 = 0.0, clr1 = 0.0, clr2 = 0.0, cmb1 = 0.0, cmb2 = 0.0;
        if (lev == nlay) {
            // all six zero; rat1, rat2 unchanged
        } else if (fu >= f0) {
            if (top) {
                if (f0 < 1.0) clr2 = (fu - f0) / (1.0 - f0);
                c2 = l2 = 0.0;
            } else {
                const double fmax = std::max(f0, fd);
                if (fu > fmax) {
                    clr1 = rat2;
                    clr2 = (fu - fmax) / (1.0 - fmax);
                } else if (fu < fmax) {
                    clr1 = (fu - f0) / (fd - f0);
                } else {
                    clr1 = rat2;
                }
            }
            rat1 = clr1 > 0.0 || clr2 > 0.0 ? 1.0 : 0.0;
            rat2 = 0.0;
        } else {
            if (top) {
                cld2 = (f0 - fu) / f0;
                c2 = l2 = 0.0;
            } else {
                const double fmin = std::min(f0, fd);
                if (fu <= fmin) {
                    cld1 = rat1;
                    cld2 = (fmin - fu) / fmin;
                } else {
                    cld1 = (f0 - fu) / (f0 - fmin);
                }
            }
            rat1 = 0.0;
            rat2 = cld1 > 0.0 || cld2 > 0.0 ? 1.0 : 0.0;
        }
        cmb1 = clr1 * c2 * fd;
        cmb2 = cld1 * l2 * (1.0 - fd);
        put(lev, LWC_CLD1, cld1);
        put(lev, LWC_CLD2, cld2);
        put(lev, LWC_CLR1, clr1);
        put(lev, LWC_CLR2, clr2);
        put(lev, LWC_CMB1, cmb1);
        put(lev, LWC_CMB2, cmb2);
        c2 = cld2;
        l2 = clr2;
    }
    // downward sweep (rrtmg_lw_rtrnmr.f90:401-476); rat1, rat2 carry over
    c2 = l2 = 0.0;
    for (int lev = nlay; lev >= 1; --lev) {
        if (!cloudy(lev)) {
            c2 = l2 = 0.0;
            continue;
        }
        const bool top = !cloudy(lev + 1);
        const double fa = cf(lev + 1), f0 = cf(lev), fb = cf(lev - 1);
        double cld1 = 0.0, cld2 = 0.0, clr1 = 0.0, clr2 = 0.0, cmb1 = 0.0, cmb2 = 0.0;
        if (lev == 1) {
            // all six zero; rat1, rat2 unchanged
        } else if (fb >= f0) {
            if (top) {
                if (f0 < 1.0) clr2 = (fb - f0) / (1.0 - f0);
                c2 = l2 = 0.0;
            } else {
                const double fmax = std::max(f0, fa);
                if (fb > fmax) {
                    clr1 = rat2;
                    clr2 = (fb - fmax) / (1.0 - fmax);
                } else if (fb < fmax) {
                    clr1 = (fb - f0) / (fa - f0);
                } else {
                    clr1 = rat2;
                }
            }
            rat1 = clr1 > 0.0 || clr2 > 0.0 ? 1.0 : 0.0;
            rat2 = 0.0;
        } else {
            if (top) {
                cld2 = (f0 - fb) / f0;
                c2 = l2 = 0.0;
            } else {
                const double fmin = std::min(f0, fa);
                if (fb <= fmin) {
                    cld1 = rat1;
                    cld2 = (fmin - fb) / fmin;
                } else {
                    cld1 = (f0 - fb) / (f0 - fmin);
                }
            }
            rat1 = 0.0;
            rat2 = cld1 > 0.0 || cld2 > 0.0 ? 1.0 : 0.0;
        }
        cmb1 = clr1 * c2 * fa;
        cmb2 = cld1 * l2 * (1.0 - fa);
        put(lev, LWC_CLD1D, cld1);
        put(lev, LWC_CLD2D, cld2);
        put(lev, LWC_CLR1D, clr1);
        put(lev, LWC_CLR2D, clr2);
        put(lev, LWC_CMB1D, cmb1);
        put(lev, LWC_CMB2D, cmb2);
        c2 = cld2;
        l2 = clr2;
    }
}

// ---------------------------------------------------------------------------
// Water-path cloud optics (inflglw = 1, 2)

// cldprop with inflag 1 or 2 (rrtmg_lw_cldprop.f90:118-291) and the
// transfer's use of its result (rrtmg_lw_rtrnmr.f90:315-317, 487-493), one
// thread per column in FP64: per layer, the cloud test on the fraction, the
// water path and the input depths; inflag 1: abscld1 times the path in
// every band; inflag 2: ice absorption by iceflag 0 (absice0, one
// coefficient), 1 (absice1, five bands), 2 or 3 (absice2 / absice3
// interpolated in the radius) and liquid by liqflag 0 (absliq0) or 1
// (absliq1 interpolated), combined over ncbands bands.  As in the
// reference, ncbands starts at 1 and carries from layer to layer (iceflag 0
// with no 16-band liquid leaves it), five-band ice without 16-band liquid
// reads its coefficients by band (iceind 1 -> 2), and after the column the
// final ncbands maps each spectral band to a cloud band (all to 1, the
// five ice bands, or one to one), whose diffusivity multiplies that cloud
// band's depth.  Out: the slant depths odcld per (layer, spectral band).
// A radius the reference stops on is counted in *bad and clamped into the
// table (the call then fails).  Contraction is off, as the reference
// rounds each product and sum.
__global__ __launch_bounds__(64) void lw_cloud_wp_kernel(ColumnInputs in, LwWork w, LwCloudWp a,
                                                         const LwBand *__restrict__ bands)
{
#pragma clang fp contract(off)
    const int ncol = in.ncol, nlay = in.nlay;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const double *T = a.tab;
    const double abscld1 = T[0], *absice0 = T + 1, *absice1 = T + 3, *absice2 = T + 13, *absice3 = T + 701;
    const double absliq0 = T[1437], *absliq1 = T + 1438;
    const unsigned char icb1[16] = {1, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5};
    auto icb = [&](int ib, int ind) { return ind == 0 ? 1 : ind == 1 ? (int)icb1[ib - 1] : ib; };
    int ncbands = 1, nbad = 0;
    double abscoice[17], abscoliq[17];
    for (int lay = 0; lay < nlay; ++lay) {
        const size_t ix = (size_t)lay * ncol + col;
        double *tc = a.od + ix * 16;
        double tauctot = 0.0;
        for (int b = 0; b < 16; ++b) {
            tc[b] = 0.0;
            tauctot = tauctot + (a.taucld ? a.taucld[ix * 16 + b] : 0.0);
        }
        const double ciwp = a.cicewp ? a.cicewp[ix] : 0.0, clwp = a.cliqwp ? a.cliqwp[ix] : 0.0;
        const double cwp = ciwp + clwp;
        if (!(a.cldfr[ix] >= 1.e-20 && (cwp >= 1.e-20 || tauctot >= 1.e-20))) continue;
        if (a.inflag == 1) {
            ncbands = 16;
            for (int ib = 0; ib < 16; ++ib) tc[ib] = abscld1 * cwp;
            continue;
        }
        const double radice = a.reice ? a.reice[ix] : 0.0;
        int iceind = 0, liqind = 0;
        if (ciwp == 0.0) {
            abscoice[1] = 0.0;
        } else if (a.iceflag == 0) {
            nbad += radice < 10.0;
            abscoice[1] = absice0[0] + absice0[1] / radice;
        } else if (a.iceflag == 1) {
            nbad += radice < 13.0 || radice > 130.0;
            ncbands = 5;
            for (int ib = 1; ib <= 5; ++ib) abscoice[ib] = absice1[(ib - 1) * 2] + absice1[(ib - 1) * 2 + 1] / radice;
            iceind = 1;
        } else {
            const int nr = a.iceflag == 2 ? 43 : 46;
            nbad += radice < 5.0 || radice > (a.iceflag == 2 ? 131.0 : 140.0);
            ncbands = 16;
            const double factor = (radice - 2.0) / 3.0;
            int index = (int)factor;
            if (index == nr) index = nr - 1;
            index = index < 1 ? 1 : (index > nr - 1 ? nr - 1 : index);  // in-table only for valid radii
            const double fint = factor - (double)index;
            const double *tb = a.iceflag == 2 ? absice2 : absice3;
            for (int ib = 1; ib <= 16; ++ib) {
                const double *r = tb + (ib - 1) * nr;
                abscoice[ib] = r[index - 1] + fint * (r[index] - (r[index - 1]));
            }
            iceind = 2;
        }
        if (clwp == 0.0) {
            abscoliq[1] = 0.0;
            if (iceind == 1) iceind = 2;
        } else if (a.liqflag == 0) {
            abscoliq[1] = absliq0;
            if (iceind == 1) iceind = 2;
        } else {
            const double radliq = a.reliq ? a.reliq[ix] : 0.0;
            nbad += radliq < 2.5 || radliq > 60.0;
            int index = (int)(radliq - 1.5);
            if (index == 0) index = 1;
            if (index == 58) index = 57;
            index = index < 1 ? 1 : (index > 57 ? 57 : index);
            const double fint = radliq - 1.5 - (double)index;
            ncbands = 16;
            for (int ib = 1; ib <= 16; ++ib) {
                const double *r = absliq1 + (ib - 1) * 58;
                abscoliq[ib] = r[index - 1] + fint * (r[index] - (r[index - 1]));
            }
            liqind = 2;
        }
        for (int ib = 1; ib <= ncbands; ++ib)
            tc[ib - 1] = ciwp * abscoice[icb(ib, iceind)] + clwp * abscoliq[icb(ib, liqind)];
    }
    if (nbad && a.bad) atomicAdd(a.bad, nbad);
    // the transfer: spectral band iband takes cloud band ib = ipat(iband)
    // with ib's diffusivity (rrtmg_lw_rtrnmr.f90:246-262, 315-317)
    const double pwvcm = w.colscal[col];
    double sd[16];
    for (int b = 0; b < 16; ++b) {
        if (b == 0 || b == 3 || b >= 9) {
            sd[b] = 1.66;
        } else {
            const LwBand &B = bands[b];
            double v = B.a0 + B.a1 * exp(B.a2 * pwvcm);
            sd[b] = v > 1.80 ? 1.80 : (v < 1.50 ? 1.50 : v);
        }
    }
    for (int lay = 0; lay < nlay; ++lay) {
        double *tc = a.od + ((size_t)lay * ncol + col) * 16;
        double t[16];
        for (int b = 0; b < 16; ++b) t[b] = tc[b];
        for (int b = 0; b < 16; ++b) {
            const int ib = ncbands == 1 ? 0 : ncbands == 5 ? (int)icb1[b] - 1 : b;
            tc[b] = sd[ib] * t[ib];
        }
    }
}

// cldprmc with inflag 2 (rrtmg_lw_cldprmc.f90:150-230), one thread per
// (layer, column, g-point): where the subcolumn passes the cloud test, its
// in-cloud depth from its water paths and the layer's radii with the
// coefficients of the g-point's band (five-band ice by icb of the band);
// elsewhere the input depth.  Contraction off, as above.
__global__ __launch_bounds__(256) void lw_mcica_wp_kernel(int ncol, int nlay, LwClouds cl, const double *tab,
                                                          double *out, int *bad)
{
#pragma clang fp contract(off)
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nlay * ncol * 140) return;
    const size_t ix = t / 140;
    const int ig = (int)(t - ix * 140);
    const double *T = tab;
    const double *absice0 = T + 1, *absice1 = T + 3, *absice2 = T + 13, *absice3 = T + 701;
    const double absliq0 = T[1437], *absliq1 = T + 1438;
    const double cf = cl.cldfmc[t], tin = cl.taucmc[t];
    const double ci = cl.ciwpmc ? cl.ciwpmc[t] : 0.0, cw = cl.clwpmc ? cl.clwpmc[t] : 0.0;
    double tau = tin;
    int nbad = 0;
    if (cf >= 1.e-20 && (ci + cw >= 1.e-20 || tin >= 1.e-20)) {
        int b = 0, end = lw_ng(1);
        while (ig >= end) end += lw_ng(++b + 1);
        const int ib = b + 1;
        const double radice = cl.reice ? cl.reice[ix] : 0.0;
        double aice = 0.0, aliq = 0.0;
        if (ci == 0.0) {
            aice = 0.0;
        } else if (cl.iceflag == 0) {
            nbad += radice < 10.0;
            aice = absice0[0] + absice0[1] / radice;
        } else if (cl.iceflag == 1) {
            nbad += radice < 13.0 || radice > 130.0;
            const int b5 = ib <= 1 ? 1 : ib == 2 ? 2 : ib <= 5 ? 3 : ib <= 8 ? 4 : 5;
            aice = absice1[(b5 - 1) * 2] + absice1[(b5 - 1) * 2 + 1] / radice;
        } else {
            const int nr = cl.iceflag == 2 ? 43 : 46;
            nbad += radice < 5.0 || radice > (cl.iceflag == 2 ? 131.0 : 140.0);
            const double factor = (radice - 2.0) / 3.0;
            int index = (int)factor;
            if (index == nr) index = nr - 1;
            index = index < 1 ? 1 : (index > nr - 1 ? nr - 1 : index);
            const double fint = factor - (double)index;
            const double *r = (cl.iceflag == 2 ? absice2 : absice3) + (ib - 1) * nr;
            aice = r[index - 1] + fint * (r[index] - (r[index - 1]));
        }
        if (cw == 0.0) {
            aliq = 0.0;
        } else if (cl.liqflag == 0) {
            aliq = absliq0;
        } else {
            const double radliq = cl.reliq ? cl.reliq[ix] : 0.0;
            nbad += radliq < 2.5 || radliq > 60.0;
            int index = (int)(radliq - 1.5);
            if (index == 0) index = 1;
            if (index == 58) index = 57;
            index = index < 1 ? 1 : (index > 57 ? 57 : index);
            const double fint = radliq - 1.5 - (double)index;
            const double *r = absliq1 + (ib - 1) * 58;
            aliq = r[index - 1] + fint * (r[index] - (r[index - 1]));
        }
        tau = ci * aice + cw * aliq;
    }
    out[t] = tau;
    if (nbad && bad) atomicAdd(bad, nbad);
}

__global__ void lw_cloud_wp_zero_kernel(int *bad) { *bad = 0; }

hipError_t launch_lw_cloud_wp(const ColumnInputs &in, const LwWork &w, const LwCloudWp &a, const LwBand *bands,
                              hipStream_t s)
{
    if (a.bad) hipLaunchKernelGGL(lw_cloud_wp_zero_kernel, dim3(1), dim3(1), 0, s, a.bad);
    hipLaunchKernelGGL(lw_cloud_wp_kernel, dim3((in.ncol + 63) / 64), dim3(64), 0, s, in, w, a, bands);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// McICA transfer (rtrnmc); the subcolumn generator is in mcica.hip

// the McICA transfer's layer flag (rrtmg_lw_rtrnmc.f90:297-313): cloudy
// where any g-point's subcolumn is.  A block reads the 140 fractions of
// MCICA_FLAG_PAIRS consecutive (layer, column) pairs as one contiguous run
// (coalesced loads; one thread per pair would stride 1120 bytes across a
// wave) and marks the pairs in LDS.
constexpr int MCICA_FLAG_PAIRS = 32;
__global__ __launch_bounds__(256) void lw_mcica_flags_kernel(int ncol, int nlay, LwClouds cl, LwCloudWork cw)
{
    __shared__ int s_fl[MCICA_FLAG_PAIRS];
    const size_t n2 = (size_t)nlay * ncol, p0 = (size_t)blockIdx.x * MCICA_FLAG_PAIRS;
    if (threadIdx.x < MCICA_FLAG_PAIRS) s_fl[threadIdx.x] = 0;
    __syncthreads();
    const size_t np = n2 - p0 < (size_t)MCICA_FLAG_PAIRS ? n2 - p0 : (size_t)MCICA_FLAG_PAIRS;
    const double *f = cl.cldfmc + p0 * 140;
    for (int i = threadIdx.x; i < (int)(np * 140); i += blockDim.x)
        if (f[i] == 1.0) s_fl[i / 140] = LWC_CLOUDY;
    __syncthreads();
    if (threadIdx.x < np) cw.flags[p0 + threadIdx.x] = s_fl[threadIdx.x];
}

// tau_tbl(itr) = bpade t / (1 - t), t = itr / 10000, tau_tbl(10000) = 1e10
// (rrtmg_lw_init.f90:106-123): evaluated, not stored; FP64 gives the
// table's value exactly
__device__ __forceinline__ double lw_tau_tbl(int itr)
{
    if (itr >= NTBL) return 1.e10;
    const double t = (double)itr / (double)NTBL;
    return (1.0 / 0.278) * t / (1.0 - t);
}

// The gas-only and gas-and-cloud terms of a cloudy layer: the three
// branches of rrtmg_lw_rtrn.f90:349-418 / rrtmg_lw_rtrnmr.f90:497-545.
template <typename Real>
struct CloudyLayer {
    Real atrans, atot, gassrc, bbd, bbdtot, bbugas, bbutot;
};
template <typename Real>
__device__ __forceinline__ CloudyLayer<Real> lw_cloudy_layer(Real odepth, Real odcld, Real plfrac, Real blay,
                                                             Real dplankup, Real dplankdn,
                                                             const Real *__restrict__ exp_tbl,
                                                             const Real *__restrict__ tfn_tbl)
{
    const Real rec_6 = (Real)0.166667, half = (Real)0.5, one = (Real)1;
    CloudyLayer<Real> r;
    Real odtot = odepth + odcld;
    if (odtot < (Real)0.06) {
        r.atrans = odepth - half * odepth * odepth;
        const Real odepth_rec = rec_6 * odepth;
        r.gassrc = plfrac * (blay + dplankdn * odepth_rec) * r.atrans;
        r.atot = odtot - half * odtot * odtot;
        const Real odtot_rec = rec_6 * odtot;
        r.bbdtot = plfrac * (blay + dplankdn * odtot_rec);
        r.bbd = plfrac * (blay + dplankdn * odepth_rec);
        r.bbugas = plfrac * (blay + dplankup * odepth_rec);
        r.bbutot = plfrac * (blay + dplankup * odtot_rec);
    } else if (odepth <= (Real)0.06) {
        r.atrans = odepth - half * odepth * odepth;
        const Real odepth_rec = rec_6 * odepth;
        r.gassrc = plfrac * (blay + dplankdn * odepth_rec) * r.atrans;
        const int ittot = tbl_index(odtot);
        const Real tfactot = at32(tfn_tbl, (unsigned)ittot);
        r.bbdtot = plfrac * (blay + tfactot * dplankdn);
        r.bbd = plfrac * (blay + dplankdn * odepth_rec);
        r.atot = one - at32(exp_tbl, (unsigned)ittot);
        r.bbugas = plfrac * (blay + dplankup * odepth_rec);
        r.bbutot = plfrac * (blay + tfactot * dplankup);
    } else {
        const int itgas = tbl_index(odepth);
        odepth = (Real)lw_tau_tbl(itgas);
        r.atrans = one - at32(exp_tbl, (unsigned)itgas);
        const Real tfacgas = at32(tfn_tbl, (unsigned)itgas);
        r.gassrc = r.atrans * plfrac * (blay + tfacgas * dplankdn);
        odtot = odepth + odcld;
        const int ittot = tbl_index(odtot);
        const Real tfactot = at32(tfn_tbl, (unsigned)ittot);
        r.bbdtot = plfrac * (blay + tfactot * dplankdn);
        r.bbd = plfrac * (blay + tfacgas * dplankdn);
        r.atot = one - at32(exp_tbl, (unsigned)ittot);
        r.bbugas = plfrac * (blay + tfacgas * dplankup);
        r.bbutot = plfrac * (blay + tfactot * dplankup);
    }
    return r;
}

// Overlap state of one stream direction (rtrnmr): the cloudy and clear
// parts of the radiance in the current cloud block and the exchange term
// `rad` between them.
template <typename Real>
struct Overlap {
    Real cld, clr, rad;
    // one cloudy layer: rrtmg_lw_rtrnmr.f90:547-569 (down) / :653-673 (up);
    // f points at the layer's six factors (cld1, cld2, clr1, clr2, cmb1, cmb2)
    __device__ __forceinline__ Real step(Real radin, bool top, Real cf, Real atrans, Real atot, Real src_tot,
                                         Real gassrc, const Real *f)
    {
        const Real one = (Real)1;
        if (top) {
            cld = cf * radin;
            clr = radin - cld;
            rad = (Real)0;
        }
        const Real ttot = one - atot;
        const Real cldsrc = src_tot * atot;
        cld = cld * ttot + cf * cldsrc;
        clr = clr * (one - atrans) + (one - cf) * gassrc;
        const Real radout = cld + clr;
        const Real radmod = rad * (f[2] * (one - atrans) + f[0] * ttot) - f[4] * gassrc + f[5] * cldsrc;
        const Real oldcld = cld - radmod;
        const Real oldclr = clr + radmod;
        rad = -radmod + f[3] * oldclr - f[1] * oldcld;
        cld = cld + rad;
        clr = clr - rad;
        return radout;
    }
};

// One band of one work unit with clouds.  ICLD 1: random overlap (rtrn);
// 2: maximum-random (rtrnmr).  ktab: the k-tables (B's offsets index it: the
// packed global buffer, or the band's copy in LDS with rebased offsets);
// tab: the packed global buffer (transmittance and Planck tables).
// row: this wave's scratch, [4][nlay][64]
// (gas transmittance, gas upward source, gas-and-cloud absorptance and
// upward source); s_red: 64 LDS entries for the FP64 g-sum.
template <int BAND, typename Real, int ICLD>
__device__ __forceinline__ void lw_cloud_band_body(const ColumnInputs &in, const LwWork &w, const LwClouds &cl,
                                                   const LwCloudWork &cw, const LwBand &B, const Real *__restrict__ ktab,
                                                   const Real *__restrict__ tab, int exp_off, int tfn_off, int plk_off,
                                                   const LwConst *__restrict__ cst, WaveUnit u, int tile, int lane,
                                                   Real *__restrict__ row, Real *s_red)
{
    constexpr int NG = lw_ng(BAND);
    const int ncol = in.ncol, nlay = in.nlay;
    const int cl_ = lane / NG, g = lane - cl_ * NG;
    const GSum<Real, NG> gsum(lane, g);
    const int col0 = tile * TILE_COLS + u.first_col + cl_;
    const bool active = cl_ < u.ncols && col0 < ncol;
    const int col = active ? col0 : 0;
    const Real *coef = (const Real *)w.coef;
    const Real *tlevf = (const Real *)w.tlev;
    const Real *exp_tbl = tab + exp_off;
    const Real *tfn_tbl = tab + tfn_off;
    const Real *plk = tab + plk_off + (BAND - 1) * 181;
    const Real *fac = (const Real *)cw.fac;
    LTab<Real> T{ktab, NG, g};
    const size_t bs = (size_t)(nlay + 1) * ncol;
    Real *bf = (Real *)cw.bflux + (size_t)(BAND - 1) * 4 * bs;  // [4][nlay+1][ncol]
    Real *s_atr = row, *s_bbu = row + (size_t)nlay * 64, *s_atot = row + 2 * (size_t)nlay * 64,
         *s_bbt = row + 3 * (size_t)nlay * 64;
    const Real one = (Real)1;

    const int laytrop = w.laytrop[col];
    const double pwvcm = w.colscal[col];
    const double tbound = w.colscal[(size_t)ncol + col];
    double sd = 1.66;
    if (!(BAND == 1 || BAND == 4 || BAND >= 10)) {
        sd = B.a0 + B.a1 * exp(B.a2 * pwvcm);
        if (sd > 1.80) sd = 1.80;
        if (sd < 1.50) sd = 1.50;
    }
    const Real secdiff = (Real)sd;
    const Real semiss = (Real)in.emis[(size_t)(BAND - 1) * ncol + col];
    int indbound = clampi((int)(tbound - 159.0), 1, 180);
    double tbndfrac = tbound - 159.0 - (double)indbound;
    const Real plankbnd =
        (Real)((double)semiss * ((double)plk[indbound - 1] + tbndfrac * ((double)plk[indbound] - (double)plk[indbound - 1])));
    auto planklev = [&](int lev) {
        int i = at32(w.ilev, (unsigned)(lev * ncol + col));
        Real f = at32(tlevf, (unsigned)(lev * ncol + col));
        return at32(plk, (unsigned)(i - 1)) + f * (at32(plk, (unsigned)i) - at32(plk, (unsigned)(i - 1)));
    };
    // the layer's cloud optical depth on the diffusivity slant path
    // (rrtmg_lw_rtrnmr.f90:314-323: secdiff of the band times cldprop's depth)
    auto odcld_of = [&](int lay, int fl) {
        if constexpr (ICLD == LW_MCICA) {
            // the lane's g-point subcolumn (rrtmg_lw_rtrnmc.f90:297-306)
            const size_t o = ((size_t)lay * ncol + col) * 140 + B.ngs + g;
            return cl.cldfmc[o] == 1.0 ? secdiff * (Real)cl.taucmc[o] : (Real)0;
        }
        if (cl.slant) return (fl & LWC_TAU) ? (Real)cl.taucld[((size_t)lay * ncol + col) * 16 + (BAND - 1)] : (Real)0;
        return (fl & LWC_TAU) ? secdiff * (Real)cl.taucld[((size_t)lay * ncol + col) * 16 + (BAND - 1)] : (Real)0;
    };
    // the cloud fraction the transfer weights a cloudy layer with: the
    // layer's, or in McICA the lane's subcolumn's (0 or 1)
    auto cf_of = [&](int lay, const Real *rec) {
        if constexpr (ICLD == LW_MCICA)
            return (Real)cl.cldfmc[((size_t)lay * ncol + col) * 140 + B.ngs + g];
        return rec[LWC_CF];
    };
    const double *tauaer = in.tauaer;
    auto store = [&](int k, int lev, Real v) {
        if (active && g == 0) bf[(size_t)k * bs + (size_t)lev * ncol + col] = v;
    };

    // downward sweep
    Real radld = (Real)0, radclrd = (Real)0, fracs1 = (Real)0;
    Overlap<Real> ov{(Real)0, (Real)0, (Real)0};
    bool iclddn = false;
    Real plev_up = planklev(nlay);
    store(1, nlay, (Real)0);
    store(3, nlay, (Real)0);
    LwLayer<Real> c;
    for (int lev = nlay; lev >= 1; --lev) {
        lw_load_layer(c, coef, w.ipk, lev - 1, col, ncol, nlay);
        Real taug, plfrac;
        lw_gas<BAND, Real, Real, true>(B, T, cst, c, lev <= laytrop, taug, plfrac);
        if (tauaer) taug = taug + (Real)tauaer[((size_t)(BAND - 1) * nlay + (lev - 1)) * ncol + col];
        Real plev_dn = planklev(lev - 1);
        Real blay = at32(plk, (unsigned)(c.indlay - 1)) +
                    c.tlayfrac * (at32(plk, (unsigned)c.indlay) - at32(plk, (unsigned)(c.indlay - 1)));
        Real dplankup = plev_up - blay, dplankdn = plev_dn - blay;
        Real odepth = secdiff * taug;
        if (odepth < (Real)0) odepth = (Real)0;
        const int fl = cw.flags[(size_t)(lev - 1) * ncol + col];
        const Real *rec = fac + ((size_t)(lev - 1) * ncol + col) * LWC_STRIDE;
        Real atrans, bbd, bbugas, atot = (Real)0, bbutot = (Real)0;
        if (fl & LWC_CLOUDY) {
            iclddn = true;
            const Real cf = cf_of(lev - 1, rec);
            const Real odcld = odcld_of(lev - 1, fl);
            const CloudyLayer<Real> r =
                lw_cloudy_layer(odepth, odcld, plfrac, blay, dplankup, dplankdn, exp_tbl, tfn_tbl);
            atrans = r.atrans;
            atot = r.atot;
            bbd = r.bbd;
            bbugas = r.bbugas;
            bbutot = r.bbutot;
            if constexpr (ICLD == 1 || ICLD == LW_MCICA) {
                const Real efcl = (one - exp(-odcld)) * cf;
                radld = radld - radld * (atrans + efcl * (one - atrans)) + r.gassrc + cf * (r.bbdtot * atot - r.gassrc);
            } else {
                radld = ov.step(radld, fl & LWC_TOP_DN, cf, atrans, atot, r.bbdtot, r.gassrc, rec + LWC_CLD1D);
            }
        } else if (odepth <= (Real)0.06) {
            atrans = odepth - (Real)0.5 * odepth * odepth;
            odepth = (Real)0.166667 * odepth;
            bbd = plfrac * (blay + dplankdn * odepth);
            bbugas = plfrac * (blay + dplankup * odepth);
            radld = radld + (bbd - radld) * atrans;
        } else {
            const int itr = tbl_index(odepth);
            atrans = one - at32(exp_tbl, (unsigned)itr);
            const Real tausfac = at32(tfn_tbl, (unsigned)itr);
            bbd = plfrac * (blay + tausfac * dplankdn);
            bbugas = plfrac * (blay + tausfac * dplankup);
            radld = radld + (bbd - radld) * atrans;
        }
        // the clear stream follows the total one until the first cloud
        radclrd = iclddn ? radclrd + (bbd - radclrd) * atrans : radld;
        const unsigned ix = (unsigned)((lev - 1) * 64 + lane);
        at32(s_atr, ix) = atrans;
        at32(s_bbu, ix) = bbugas;
        if (fl & LWC_CLOUDY) {
            at32(s_atot, ix) = atot;
            at32(s_bbt, ix) = bbutot;
        }
        const Real sd_tot = gsum.sum(active ? radld : (Real)0, s_red);
        store(1, lev - 1, sd_tot);
        // while no column of the wave has met a cloud both streams are equal
        const bool met = __builtin_amdgcn_ballot_w64(iclddn) != 0;
        store(3, lev - 1, met ? gsum.sum(active ? radclrd : (Real)0, s_red) : sd_tot);
        plev_up = plev_dn;
        if (lev == 1) fracs1 = plfrac;
    }
    // surface emission and reflection (rrtmg_lw_rtrnmr.f90:618-637)
    const Real rad0 = fracs1 * plankbnd, reflect = one - semiss;
    Real radlu = rad0 + reflect * radld;
    Real radclru = rad0 + reflect * radclrd;
    // no column of the wave cloudy: the clear stream is the total one throughout
    const bool any_cloud = __builtin_amdgcn_ballot_w64(iclddn) != 0;
    const Real su0 = gsum.sum(active ? radlu : (Real)0, s_red);
    store(0, 0, su0);
    store(2, 0, any_cloud ? gsum.sum(active ? radclru : (Real)0, s_red) : su0);
    ov = Overlap<Real>{(Real)0, (Real)0, (Real)0};
    for (int lev = 1; lev <= nlay; ++lev) {
        const unsigned ix = (unsigned)((lev - 1) * 64 + lane);
        const Real atrans = at32(s_atr, ix), bbugas = at32(s_bbu, ix);
        const int fl = cw.flags[(size_t)(lev - 1) * ncol + col];
        if (fl & LWC_CLOUDY) {
            const Real *rec = fac + ((size_t)(lev - 1) * ncol + col) * LWC_STRIDE;
            const Real cf = cf_of(lev - 1, rec);
            const Real atot = at32(s_atot, ix), bbutot = at32(s_bbt, ix);
            const Real gassrc = bbugas * atrans;
            if constexpr (ICLD == 1 || ICLD == LW_MCICA) {
                const Real efcl = (one - exp(-odcld_of(lev - 1, fl))) * cf;
                radlu = radlu - radlu * (atrans + efcl * (one - atrans)) + gassrc + cf * (bbutot * atot - gassrc);
            } else {
                radlu = ov.step(radlu, fl & LWC_TOP_UP, cf, atrans, atot, bbutot, gassrc, rec + LWC_CLD1);
            }
        } else {
            radlu = radlu + (bbugas - radlu) * atrans;
        }
        radclru = iclddn ? radclru + (bbugas - radclru) * atrans : radlu;
        const Real su = gsum.sum(active ? radlu : (Real)0, s_red);
        store(0, lev, su);
        store(2, lev, any_cloud ? gsum.sum(active ? radclru : (Real)0, s_red) : su);
    }
}

template <typename Real, int ICLD>
__global__ __launch_bounds__(64) void lw_cloud_band_kernel(ColumnInputs in, LwWork w, LwClouds cl, LwCloudWork cw,
                                                           const LwBand *__restrict__ bands, int nwaves, int blk0,
                                                           int nblocks, int exp_off, int tfn_off, int plk_off,
                                                           const Real *__restrict__ tab,
                                                           const LwConst *__restrict__ cst,
                                                           const WaveUnit *__restrict__ units)
{
    __shared__ Real s_red[64];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const LwBand B = bands[u.band];
    Real *row = (Real *)cw.rows + (size_t)blk * 4 * in.nlay * 64;
#define BODY(b)                                                                                                 \
    case b - 1:                                                                                                 \
        lw_cloud_band_body<b, Real, ICLD>(in, w, cl, cw, B, tab, tab, exp_off, tfn_off, plk_off, cst, u, tile,   \
                                          (int)threadIdx.x, row, s_red);                                        \
        break;
    switch (u.band) {
        BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
        BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
    }
#undef BODY
}

// FP32 with the band's k-tables in LDS: the schedule of the clear-sky
// lw_band_lds_kernel (persistent workgroups of LW_WGW waves claiming work
// items from a global counter; item = (group of LW_WGW tiles, unit r of a
// tile's unit list); a band's tables copied into LDS before its first item)
// with the cloudy band body.  Each wave of the persistent grid has one
// scratch row of 4 x nlay x 64 floats.
template <int ICLD>
__global__ __launch_bounds__(64 * LW_WGW) void lw_cloud_band_lds_kernel(ColumnInputs in, LwWork w, LwClouds cl,
                                                                        LwCloudWork cw, const LwBand *__restrict__ bands,
                                                                        LwLdsPlan plan, int nwaves, int nitems,
                                                                        int exp_off, int tfn_off, int plk_off,
                                                                        const float *__restrict__ tab,
                                                                        const LwConst *__restrict__ cst,
                                                                        const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    float *ktab = (float *)smem_raw;
    const int nlay = in.nlay;
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    float *row = (float *)cw.rows + ((size_t)blockIdx.x * LW_WGW + wave) * 4 * nlay * 64;
    __shared__ int s_item;
    int loaded = -1;
    for (;;) {
        const int item = next_item(w.items, &s_item);
        if (item >= nitems) break;
        const int group = item / nwaves, r = item - group * nwaves;
        const int b = units[r].band;
        if (b != loaded) {
            const int t0 = plan.tab0[b], n = plan.tab0[b + 1] - t0;
#pragma unroll 8
            for (int i = threadIdx.x; i < n; i += 64 * LW_WGW) ktab[i] = tab[t0 + i];
            __syncthreads();
            loaded = b;
        }
        const int k = (r - plan.ubeg[b]) * LW_WGW + wave;
        const int tile = group * LW_WGW + k / plan.upt[b];
        if (tile >= ntiles) continue;
        const WaveUnit u = units[plan.ubeg[b] + k % plan.upt[b]];
        const LwBand B = bands[b];
#define BODY(bb)                                                                                              \
    case bb - 1:                                                                                              \
        lw_cloud_band_body<bb, float, ICLD>(in, w, cl, cw, B, ktab, tab, exp_off, tfn_off, plk_off, cst, u,   \
                                            tile, lane, row, nullptr);                                        \
        break;
        switch (b) {
            BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
            BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
        }
#undef BODY
    }
}

// finish: band sums, fluxfac, net fluxes and heating of both streams
// (rrtmg_lw_rtrnmr.f90:719-777), one thread per (layer, column)
template <typename Real>
__global__ __launch_bounds__(256) void lw_cloud_finish_kernel(ColumnInputs in, LwCloudWork cw, LwOutputs out,
                                                              const LwBand *__restrict__ bands, double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    const Real *bflux = (const Real *)cw.bflux;
    const size_t bs = (size_t)(nlay + 1) * ncol;
    const double fluxfac = 2.0 * asin(1.0) * 2.e4;
    double fnet[2][2];
    for (int j = 0; j < 2; ++j) {
        const int k = lay + j;
        double f[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = 0; b < 16; ++b) {
            const double dw = bands[b].delwave;
            for (int s = 0; s < 4; ++s)
                f[s] = f[s] + ((double)bflux[((size_t)b * 4 + s) * bs + (size_t)k * ncol + col] * 0.5) * dw;
        }
        for (int s = 0; s < 4; ++s) f[s] = f[s] * fluxfac;
        if (j == 0 || lay == nlay - 1) {
            const size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = f[0];
            out.dflx[ix] = f[1];
            out.uflxc[ix] = f[2];
            out.dflxc[ix] = f[3];
        }
        fnet[0][j] = f[0] - f[1];
        fnet[1][j] = f[2] - f[3];
    }
    const size_t il = (size_t)lay * ncol + col;
    const double dp = in.plev[il] - in.plev[il + ncol];
    out.hr[il] = heatfac * (fnet[0][0] - fnet[0][1]) / dp;
    out.hrc[il] = heatfac * (fnet[1][0] - fnet[1][1]) / dp;
}

// the band kernels of one transfer (ICLD 1 rtrn, 2 rtrnmr, LW_MCICA rtrnmc)
template <typename Real, int ICLD>
static void launch_lw_cloud_bands(const ColumnInputs &in, const LwClouds &cl, const LwWork &w,
                                  const LwCloudWork &cw, const LwDevTables &t, hipStream_t s)
{
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    if constexpr (sizeof(Real) == 4) {
        if (t.lds) {
            const int nitems = (ntiles + LW_WGW - 1) / LW_WGW * t.desc.nwaves;
            const dim3 grid(lw_lds_grid(t.plan, in.ncol, t.desc.nwaves)), block(64 * LW_WGW);
            hipLaunchKernelGGL(reset_items_kernel, dim3(1), dim3(1), 0, s, w.items);
            hipLaunchKernelGGL(lw_cloud_band_lds_kernel<ICLD>, grid, block, (size_t)t.plan.lds_bytes, s, in, w, cl, cw,
                               t.bands_lds, t.plan, t.desc.nwaves, nitems, t.desc.exp_tbl, t.desc.tfn_tbl,
                               t.desc.totplnk, (const float *)t.tab, t.cst, t.units);
            return;
        }
    }
    const int nblocks = ntiles * t.desc.nwaves;
    for (int b0 = 0; b0 < nblocks; b0 += cw.chunk) {
        const int nb_c = std::min(cw.chunk, nblocks - b0);
        hipLaunchKernelGGL((lw_cloud_band_kernel<Real, ICLD>), dim3(xcd_grid(nb_c)), dim3(64), 0, s, in, w, cl, cw,
                           t.bands, t.desc.nwaves, b0, nb_c, t.desc.exp_tbl, t.desc.tfn_tbl, t.desc.totplnk,
                           (const Real *)t.tab, t.cst, t.units);
    }
}

template <typename Real>
static hipError_t launch_lw_cloud_t(const ColumnInputs &in, const LwClouds &cl_in, const LwOutputs &out,
                                    const LwWork &w, const LwCloudWork &cw, const LwDevTables &t, int mode,
                                    hipStream_t s, hipEvent_t *ev)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipError_t e = launch_lw_prep(in, w, t, mode, s);
    if (e != hipSuccess) return e;
    LwClouds cl = cl_in;
    if (cl.inflag == 2 && cl.icld == LW_MCICA) {
        // McICA with water paths: cldprmc's depths per subcolumn
        const size_t n = (size_t)in.nlay * in.ncol * 140;
        if (cw.bad) hipLaunchKernelGGL(lw_cloud_wp_zero_kernel, dim3(1), dim3(1), 0, s, cw.bad);
        hipLaunchKernelGGL(lw_mcica_wp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in.ncol, in.nlay,
                           cl, cw.cldtab, cw.od, cw.bad);
        cl.taucmc = cw.od;
    } else if (cl.inflag != 0 && cl.icld != LW_MCICA) {
        // water paths: cldprop's optics, then the transfer on its slant depths
        LwCloudWp a{cl.inflag, cl.iceflag, cl.liqflag, cl.cldfr, cl.taucld, cl.cicewp, cl.cliqwp,
                    cl.reice, cl.reliq, cw.cldtab, cw.od, cw.bad};
        e = launch_lw_cloud_wp(in, w, a, t.bands, s);
        if (e != hipSuccess) return e;
        cl.taucld = cw.od;
        cl.slant = 1;
    }
    if (cl.icld == LW_MCICA) {
        const size_t n2 = (size_t)in.nlay * in.ncol;
        hipLaunchKernelGGL(lw_mcica_flags_kernel, dim3((unsigned)((n2 + MCICA_FLAG_PAIRS - 1) / MCICA_FLAG_PAIRS)),
                           dim3(256), 0, s, in.ncol, in.nlay, cl, cw);
    } else {
        // one thread per column, sequential over layers: 64-thread blocks
        // spread the columns over all compute units
        hipLaunchKernelGGL(lw_cloud_column_kernel<Real>, dim3((in.ncol + 63) / 64), dim3(64), 0, s, in.ncol, in.nlay,
                           cl, cw);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (cl.icld == 1)
        launch_lw_cloud_bands<Real, 1>(in, cl, w, cw, t, s);
    else if (cl.icld == LW_MCICA)
        launch_lw_cloud_bands<Real, LW_MCICA>(in, cl, w, cw, t, s);
    else
        launch_lw_cloud_bands<Real, 2>(in, cl, w, cw, t, s);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(lw_cloud_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, cw, out, t.bands, t.heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_lw_cloud(const ColumnInputs &in, const LwClouds &cl, const LwOutputs &out, const LwWork &w,
                           const LwCloudWork &cw, const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev)
{
    return mode == 1 ? launch_lw_cloud_t<float>(in, cl, out, w, cw, t, mode, s, ev)
                     : launch_lw_cloud_t<double>(in, cl, out, w, cw, t, mode, s, ev);
}

}  // namespace rrtmg
