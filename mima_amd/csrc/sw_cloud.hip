// Shortwave RRTMG with clouds on gfx950: rrtmg_sw with icld >= 1 and
// inflgsw = 0 (rrtmg_sw/gcm_model/src/rrtmg_sw_rad.nomcica.f90:498-729):
// cldprop_sw's delta-M scaled cloud optics (rrtmg_sw_cldprop.f90:118-157)
// and spcvrt_sw's two streams (rrtmg_sw_spcvrt.f90:302-600): the clear one,
// and the total one whose layer reflectances, transmittances and direct-beam
// transmittance blend the clear and the cloudy layer values by the cloud
// fraction.  Without McICA the reference takes overcast or clear layers
// only (it stops on partial cloud in a day column, :535-539); the blend is
// its formula either way.
//
// The clear-sky forward's prep and column kernels supply the gas optics.
// sw_cloud_flags_kernel classifies each (layer, column) once.  The band
// kernel follows the clear-sky one (one wave per work unit, lane = (column,
// g); both sweeps evaluate the layer optics) and adds, for layers with
// cloud, the cloudy layer optics: clear plus delta-scaled cloud optical
// depth, albedo and asymmetry (:459-466) through a general-asymmetry reftra
// in FP64 in both precisions (cloud albedos near 1 are where a two-stream
// eigenvalue in FP32 loses its digits).  Each sweep carries the recurrences
// of both streams (the scratch row holds prup/prupd of each), and each level
// sums four fluxes over g.  The clear-sky kernels are left unchanged.
//
// McICA (rrtmg_sw_rad.f90 with cldprmc_sw and spcvmc_sw): spcvmc_sw is
// spcvrt_sw with the fraction and cloud optics of each g-point's subcolumn
// in place of the layer's and the band's (rrtmg_sw_spcvmc.f90:302-600), so
// the same band kernel runs it with each lane reading its own g-point's
// subcolumn (MC = true); the layer flag then only says that some subcolumn
// is cloudy.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "rrtmg_device.hpp"
#include "sw_optics.hpp"

namespace rrtmg {

// flags of each (layer, column) (rrtmg_sw_cldprop.f90:128-138,
// rrtmg_sw_spcvrt.f90:388) and the count of layers of day columns with a
// fraction the nomcica driver refuses (rrtmg_sw_rad.nomcica.f90:501, 535-539)
__global__ __launch_bounds__(256) void sw_cloud_flags_kernel(ColumnInputs in, SwClouds cl, SwCloudWork cw)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int col = (int)(i % ncol);
    const double cf = cl.cldfr[i];
    double tauctot = 0.0;
    for (int b = 0; b < 14; ++b) tauctot = tauctot + cl.taucld[i * 14 + b];
    const double cwp = (cl.cicewp ? cl.cicewp[i] : 0.0) + (cl.cliqwp ? cl.cliqwp[i] : 0.0);
    int fl = 0;
    if (cf > 0.0) fl |= SWC_BLEND;
    if (cf > 1.e-12) fl |= SWC_REFTRA;
    if (cf >= 1.e-20 && (cwp >= 1.e-20 || tauctot >= 1.e-20)) fl |= SWC_OPTICS;
    cw.flags[i] = fl;
    if (in.coszen[col] >= 1.e-10 && cf > 1.e-6 && cf < 1.0 - 1.e-6) atomicAdd(cw.partial, 1);
}

__global__ void sw_cloud_zero_kernel(int *p) { *p = 0; }

// McICA: SWC_BLEND where some g-point's subcolumn has cloud.  A block reads
// the 112 fractions of SWMC_FLAG_PAIRS consecutive (layer, column) pairs as
// one contiguous run and marks the pairs in LDS.
constexpr int SWMC_FLAG_PAIRS = 32;
__global__ __launch_bounds__(256) void sw_mcica_flags_kernel(int ncol, int nlay, SwClouds cl, SwCloudWork cw)
{
    __shared__ int s_fl[SWMC_FLAG_PAIRS];
    const size_t n2 = (size_t)nlay * ncol, p0 = (size_t)blockIdx.x * SWMC_FLAG_PAIRS;
    if (threadIdx.x < SWMC_FLAG_PAIRS) s_fl[threadIdx.x] = 0;
    __syncthreads();
    const size_t np = n2 - p0 < (size_t)SWMC_FLAG_PAIRS ? n2 - p0 : (size_t)SWMC_FLAG_PAIRS;
    const double *f = cl.cldfmc + p0 * 112;
    for (int i = threadIdx.x; i < (int)(np * 112); i += blockDim.x)
        if (f[i] > 0.0) s_fl[i / 112] = SWC_BLEND;
    __syncthreads();
    if (threadIdx.x < np) cw.flags[p0 + threadIdx.x] = s_fl[threadIdx.x];
}

__global__ __launch_bounds__(256) void sw_flags_clear_kernel(size_t n, int *flags)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = 0;
}

// cldprop_sw's table positions for the radii (rrtmg_sw_cldprop.f90:204-208,
// 236-240, 279-283): 1-based lower index and fraction for ice (iceflag 2 and
// 3) and liquid; *nbad counts a radius the reference stops on (the indices
// stay in the tables).
__device__ __forceinline__ void sw_wp_index(int iceflag, double ciwp, double clwp, double radice, double radliq,
                                            int &nbad, int &iindex, double &ifint, int &lindex, double &lfint)
{
#pragma clang fp contract(off)
    iindex = 1;
    lindex = 1;
    ifint = 0.0;
    lfint = 0.0;
    if (ciwp != 0.0) {
        if (iceflag == 1) {
            nbad += radice < 13.0 || radice > 130.0;
        } else {
            const int nr = iceflag == 2 ? 43 : 46;
            nbad += radice < 5.0 || radice > (iceflag == 2 ? 131.0 : 140.0);
            const double factor = (radice - 2.0) / 3.0;
            iindex = (int)factor;
            if (iindex == nr) iindex = nr - 1;
            iindex = iindex < 1 ? 1 : (iindex > nr - 1 ? nr - 1 : iindex);
            ifint = factor - (double)iindex;
        }
    }
    if (clwp != 0.0) {
        nbad += radliq < 2.5 || radliq > 60.0;
        lindex = (int)(radliq - 1.5);
        if (lindex == 0) lindex = 1;
        if (lindex == 58) lindex = 57;
        lindex = lindex < 1 ? 1 : (lindex > 57 ? 57 : lindex);
        lfint = radliq - 1.5 - (double)lindex;
    }
}

// cldprop_sw's final optics of band b (0-based) from the paths
// (rrtmg_sw_cldprop.f90:169-337): ice by iceflag 1 (the abari..fbari fits in
// the radius, the band group by the band's upper wavenumber), 2 or 3 (tables
// interpolated in the radius, 3 with the delta fraction), liquid by liqflag
// 1; each phase delta-scaled by its forward fraction and combined, with the
// reference's floors.  FP64 with contraction off.
__device__ __forceinline__ void sw_wp_band(const double *tab, int iceflag, int b, double ciwp, double clwp,
                                           double radice, int iindex, double ifint, int lindex, double lfint,
                                           int &nbad, double &otau, double &ossa, double &oasm)
{
#pragma clang fp contract(off)
    const double eps = 1.e-06, cldmin = 1.e-20;
    const double *extliq1 = tab, *ssaliq1 = tab + 812, *asyliq1 = tab + 1624;
    const double *extice2 = tab + 2436, *ssaice2 = tab + 3038, *asyice2 = tab + 3640;
    const double *extice3 = tab + 4242, *ssaice3 = tab + 4886, *asyice3 = tab + 5530, *fdlice3 = tab + 6174;
    const double *bar = tab + 6818, *wavenum2 = tab + 6848;
    double extcoice = 0.0, ssacoice = 0.0, gice = 0.0, forwice = 0.0;
    if (ciwp != 0.0) {
        if (iceflag == 1) {
            const double w2 = wavenum2[b];
            const int icx = w2 > 1.43e04 ? 0 : w2 > 7.7e03 ? 1 : w2 > 5.3e03 ? 2 : w2 > 4.0e03 ? 3 : 4;
            extcoice = bar[icx] + bar[5 + icx] / radice;
            ssacoice = 1.0 - bar[10 + icx] - bar[15 + icx] * radice;
            gice = bar[20 + icx] + bar[25 + icx] * radice;
            if (gice >= 1.0) gice = 1.0 - eps;
            forwice = gice * gice;
        } else {
            const int nr = iceflag == 2 ? 43 : 46;
            const double *ex = (iceflag == 2 ? extice2 : extice3) + b * nr;
            const double *ss = (iceflag == 2 ? ssaice2 : ssaice3) + b * nr;
            const double *as = (iceflag == 2 ? asyice2 : asyice3) + b * nr;
            const int i = iindex;
            extcoice = ex[i - 1] + ifint * (ex[i] - ex[i - 1]);
            ssacoice = ss[i - 1] + ifint * (ss[i] - ss[i - 1]);
            gice = as[i - 1] + ifint * (as[i] - as[i - 1]);
            if (iceflag == 2) {
                forwice = gice * gice;
            } else {
                const double *fd = fdlice3 + b * nr;
                const double fdelta = fd[i - 1] + ifint * (fd[i] - fd[i - 1]);
                nbad += fdelta < 0.0 || fdelta > 1.0;
                forwice = fdelta + 0.5 / ssacoice;
                if (forwice > gice) forwice = gice;
            }
        }
        nbad += extcoice < 0.0 || ssacoice > 1.0 || ssacoice < 0.0 || gice > 1.0 || gice < 0.0;
    }
    double extcoliq = 0.0, ssacoliq = 0.0, gliq = 0.0, forwliq = 0.0;
    if (clwp != 0.0) {
        const int i = lindex;
        const double *ex = extliq1 + b * 58, *ss = ssaliq1 + b * 58, *as = asyliq1 + b * 58;
        extcoliq = ex[i - 1] + lfint * (ex[i] - ex[i - 1]);
        ssacoliq = ss[i - 1] + lfint * (ss[i] - ss[i - 1]);
        if (lfint < 0.0 && ssacoliq > 1.0) ssacoliq = ss[i - 1];
        gliq = as[i - 1] + lfint * (as[i] - as[i - 1]);
        forwliq = gliq * gliq;
        nbad += extcoliq < 0.0 || ssacoliq > 1.0 || ssacoliq < 0.0 || gliq > 1.0 || gliq < 0.0;
    }
    const double tauliqorig = clwp * extcoliq, tauiceorig = ciwp * extcoice;
    const double ssaliq = ssacoliq * (1.0 - forwliq) / (1.0 - forwliq * ssacoliq);
    const double tauliq = (1.0 - forwliq * ssacoliq) * tauliqorig;
    const double ssaice = ssacoice * (1.0 - forwice) / (1.0 - forwice * ssacoice);
    const double tauice = (1.0 - forwice * ssacoice) * tauiceorig;
    const double scatliq = ssaliq * tauliq;
    double scatice = ssaice * tauice;
    double t = tauliq + tauice;
    if (t == 0.0) t = cldmin;
    if (scatice == 0.0) scatice = cldmin;
    otau = t;
    ossa = (scatliq + scatice) / t;
    if (iceflag == 3)
        oasm = (1.0 / (scatliq + scatice)) *
               (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * ((gice - forwice) / (1.0 - forwice)));
    else
        oasm = (scatliq * (gliq - forwliq) / (1.0 - forwliq) + scatice * (gice - forwice) / (1.0 - forwice)) /
               (scatliq + scatice);
}

// cldprop_sw with inflag 2 (rrtmg_sw_cldprop.f90:143-338), one thread per
// (layer, column): a layer that passes the cloud test gets each band's
// final optics from its paths; any other depth 0, albedo 1, asymmetry 0.
// inflag 1: no optics (cldprop_sw has no such branch).  Where the
// reference stops, *bad counts the layer.
__global__ __launch_bounds__(256) void sw_cloud_wp_kernel(int ncol, int nlay, SwClouds cl, const double *tab,
                                                          double *otau, double *ossa, double *oasm, int *bad)
{
#pragma clang fp contract(off)
    const size_t ix = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= (size_t)nlay * ncol) return;
    const double cldmin = 1.e-20;
    double *tau = otau + ix * 14, *ssa = ossa + ix * 14, *asy = oasm + ix * 14;
    double tauctot = 0.0;
    for (int b = 0; b < 14; ++b) {
        tau[b] = 0.0;
        ssa[b] = 1.0;
        asy[b] = 0.0;
        tauctot = tauctot + (cl.taucld ? cl.taucld[ix * 14 + b] : 0.0);
    }
    const double ciwp = cl.cicewp ? cl.cicewp[ix] : 0.0, clwp = cl.cliqwp ? cl.cliqwp[ix] : 0.0;
    if (cl.inflag != 2 || !(cl.cldfr[ix] >= cldmin && (ciwp + clwp >= cldmin || tauctot >= cldmin))) return;
    const double radice = cl.reice ? cl.reice[ix] : 0.0, radliq = cl.reliq ? cl.reliq[ix] : 0.0;
    int nbad = 0, iindex, lindex;
    double ifint, lfint;
    sw_wp_index(cl.iceflag, ciwp, clwp, radice, radliq, nbad, iindex, ifint, lindex, lfint);
    for (int b = 0; b < 14; ++b)
        sw_wp_band(tab, cl.iceflag, b, ciwp, clwp, radice, iindex, ifint, lindex, lfint, nbad, tau[b], ssa[b], asy[b]);
    if (nbad && bad) atomicAdd(bad, nbad);
}

// cldprmc_sw with inflag 2 (rrtmg_sw_cldprmc.f90:124-330), one thread per
// subcolumn o = (layer, column, g-point), g fastest: a subcolumn that passes
// its cloud test gets cldprop_sw's optics of its band ngb(g) from its own
// paths and the layer's radii (reicmc, relqmc); any other keeps its input
// depth, albedo and asymmetry.
__global__ __launch_bounds__(256) void sw_mcica_wp_kernel(int ncol, int nlay, SwClouds cl, const double *tab,
                                                          double *otau, double *ossa, double *oasm, int *bad)
{
#pragma clang fp contract(off)
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (size_t)nlay * ncol * 112) return;
    const size_t ix = o / 112;
    const int g = (int)(o - ix * 112);
    const double cldmin = 1.e-20;
    const double cf = cl.cldfmc[o], tin = cl.taucmc[o];
    const double ciwp = cl.ciwpmc ? cl.ciwpmc[o] : 0.0, clwp = cl.clwpmc ? cl.clwpmc[o] : 0.0;
    double t = tin, w = cl.ssacmc[o], a = cl.asmcmc[o];
    int nbad = 0;
    if (cf >= cldmin && (ciwp + clwp >= cldmin || tin >= cldmin)) {
        const int ngc[14] = {6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12};
        int b = 0;
        for (int acc = ngc[0]; g >= acc && b < 13; acc += ngc[++b]) {
        }
        const double radice = cl.reice ? cl.reice[ix] : 0.0, radliq = cl.reliq ? cl.reliq[ix] : 0.0;
        int iindex, lindex;
        double ifint, lfint;
        sw_wp_index(cl.iceflag, ciwp, clwp, radice, radliq, nbad, iindex, ifint, lindex, lfint);
        sw_wp_band(tab, cl.iceflag, b, ciwp, clwp, radice, iindex, ifint, lindex, lfint, nbad, t, w, a);
    }
    otau[o] = t;
    ossa[o] = w;
    oasm[o] = a;
    if (nbad && bad) atomicAdd(bad, nbad);
}

hipError_t launch_sw_cloud_wp(const ColumnInputs &in, const SwClouds &cl, const double *tab, double *tau,
                              double *ssa, double *asm_, int *bad, hipStream_t s)
{
    const size_t n2 = (size_t)in.nlay * in.ncol;
    if (bad) hipLaunchKernelGGL(sw_cloud_zero_kernel, dim3(1), dim3(1), 0, s, bad);
    hipLaunchKernelGGL(sw_cloud_wp_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, in.ncol, in.nlay, cl,
                       tab, tau, ssa, asm_, bad);
    return hipGetLastError();
}

hipError_t launch_sw_mcica_wp(const ColumnInputs &in, const SwClouds &cl, const double *tab, double *tau,
                              double *ssa, double *asm_, int *bad, hipStream_t s)
{
    const size_t n = (size_t)in.nlay * in.ncol * 112;
    if (bad) hipLaunchKernelGGL(sw_cloud_zero_kernel, dim3(1), dim3(1), 0, s, bad);
    hipLaunchKernelGGL(sw_mcica_wp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in.ncol, in.nlay, cl,
                       tab, tau, ssa, asm_, bad);
    return hipGetLastError();
}

// the McICA driver floors the zenith cosine at 1e-10 and computes night
// columns there instead of zeroing them (rrtmg_sw_rad.f90:571-572)
__global__ __launch_bounds__(256) void sw_coszen_floor_kernel(int ncol, const double *__restrict__ c, double *o)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncol) o[i] = c[i] < 1.e-10 ? 1.e-10 : c[i];
}

hipError_t launch_sw_coszen_floor(int ncol, const double *c, double *o, hipStream_t s)
{
    hipLaunchKernelGGL(sw_coszen_floor_kernel, dim3((ncol + 255) / 256), dim3(256), 0, s, ncol, c, o);
    return hipGetLastError();
}

// PIFM two-stream reflectance and transmittance with asymmetry g, kmodts = 2
// (rrtmg_sw_reftra.f90:138-303), FP64; the exp table of the mode (values
// converted to FP64).
template <typename Real>
__device__ void reftra_g(double zto1, double zw, double zg, double prmuz, const Real *__restrict__ exp_tbl,
                         double &pref, double &prefd, double &ptra, double &ptrad)
{
    const double eps = 1.e-08, od_lo = 0.06, zwcrit = 0.9999995;
    auto ex = [&](double x) { return x <= od_lo ? 1.0 - x + 0.5 * x * x : (double)at32(exp_tbl, (unsigned)tbl_index(x)); };
    const double zg3 = 3.0 * zg;
    const double zgamma1 = (8.0 - zw * (5.0 + zg3)) * 0.25;
    const double zgamma2 = 3.0 * (zw * (1.0 - zg)) * 0.25;
    const double zgamma3 = (2.0 - zg3 * prmuz) * 0.25;
    const double zgamma4 = 1.0 - zgamma3;
    const double q = zg / (1.0 - zg);
    const double zwo = zw / (1.0 - (1.0 - zw) * (q * q));
    if (zwo >= zwcrit) {
        const double za1 = zgamma1 * prmuz - zgamma3;
        const double zgt = zgamma1 * zto1;
        const double ze2 = ex(fmin(zto1 / prmuz, 500.0));
        if (ze2 == 1.0) {
            pref = 0.0;
            ptra = 1.0;
            prefd = 0.0;
            ptrad = 1.0;
        } else {
            pref = (zgt - za1 * (1.0 - ze2)) / (1.0 + zgt);
            ptra = 1.0 - pref;
            prefd = zgt / (1.0 + zgt);
            ptrad = 1.0 - prefd;
        }
        return;
    }
    const double za1 = zgamma1 * zgamma4 + zgamma2 * zgamma3;
    const double za2 = zgamma1 * zgamma3 + zgamma2 * zgamma4;
    const double zrk = sqrt(zgamma1 * zgamma1 - zgamma2 * zgamma2);
    const double zrp = zrk * prmuz;
    const double zrp1 = 1.0 + zrp, zrm1 = 1.0 - zrp;
    const double zrk2 = 2.0 * zrk;
    const double zrpp = 1.0 - zrp * zrp;
    const double zrkg = zrk + zgamma1;
    const double zr1 = zrm1 * (za2 + zrk * zgamma3);
    const double zr2 = zrp1 * (za2 - zrk * zgamma3);
    const double zr3 = zrk2 * (zgamma3 - za2 * prmuz);
    const double zr4 = zrpp * zrkg;
    const double zr5 = zrpp * (zrk - zgamma1);
    const double zt1 = zrp1 * (za1 + zrk * zgamma4);
    const double zt2 = zrm1 * (za1 - zrk * zgamma4);
    const double zt3 = zrk2 * (zgamma4 + za1 * prmuz);
    const double zbeta = (zgamma1 - zrk) / zrkg;
    const double zem1 = ex(fmin(zrk * zto1, 500.0));
    const double zem2 = ex(fmin(zto1 / prmuz, 500.0));
    const double zep1 = 1.0 / zem1, zep2 = 1.0 / zem2;
    const double zdenr = zr4 * zep1 + zr5 * zem1;
    const double zdent = zdenr;  // zt4 = zr4, zt5 = zr5
    if (zdenr >= -eps && zdenr <= eps) {
        pref = eps;
        ptra = zem2;
    } else {
        pref = zw * (zr1 * zep1 - zr2 * zem1 - zr3 * zem2) / zdenr;
        ptra = zem2 - zem2 * zw * (zt1 * zep1 - zt2 * zem1 - zt3 * zep2) / zdent;
    }
    const double zemm = zem1 * zem1;
    const double zdend = 1.0 / ((1.0 - zbeta * zemm) * zrkg);
    prefd = zgamma2 * (1.0 - zemm) * zdend;
    ptrad = zrk2 * zem1 * zdend;
}

// layer values of one stream
template <typename Real>
struct SwLayerRT {
    Real r, rd, t, td, dbt;
};

// The cloudy layer values of one (layer, g) blended with the clear ones c
// (rrtmg_sw_spcvrt.f90:459-466, 502-549; no aerosol, so the clear asymmetry
// is 0): the clear depth and albedo plus the delta-scaled cloud optics
// (pt, pw, pg); their reftra where the fraction exceeds 1e-12 (else the
// reference's no-reftra values 0, 0, 1, 1) and direct beam; the blend by
// the fraction in FP64.
template <typename Real>
__device__ __forceinline__ SwLayerRT<Real> sw_cloudy_blend(double ztauc, double zomcc, double pt, double pw, double pg,
                                                           double cf, int fl, double mu0,
                                                           const Real *__restrict__ exp_tbl, SwLayerRT<Real> c)
{
    const double ztauo = ztauc + pt;
    double zomco = ztauc * zomcc + pt * pw;
    const double zgco = (pt * pw * pg) / zomco;
    zomco = zomco / ztauo;
    double ro = 0.0, rdo = 0.0, to = 1.0, tdo = 1.0;
    if (fl & SWC_REFTRA) reftra_g(ztauo, zomco, zgco, mu0, exp_tbl, ro, rdo, to, tdo);
    const double ze1 = ztauo / mu0;
    const double dbto = ze1 <= 0.06 ? 1.0 - ze1 + 0.5 * ze1 * ze1 : (double)at32(exp_tbl, (unsigned)tbl_index(ze1));
    const double zclear = 1.0 - cf;
    SwLayerRT<Real> t;
    t.r = (Real)(zclear * (double)c.r + cf * ro);
    t.rd = (Real)(zclear * (double)c.rd + cf * rdo);
    t.t = (Real)(zclear * (double)c.t + cf * to);
    t.td = (Real)(zclear * (double)c.td + cf * tdo);
    t.dbt = (Real)(zclear * (double)c.dbt + cf * dbto);
    return t;
}

// The clear and total layer values of one (layer, g): the clear optics as
// the clear-sky kernel forms them and, for a layer with cloud, their blend
// with the cloudy ones (cldprop_sw's delta-M scaling of the input optics,
// rrtmg_sw_cldprop.f90:143-156, where the layer passes its cloud test).
// McICA (MC): the lane's subcolumn o = (layer, column, g-point) instead of
// the layer and band; cldprmc_sw scales the optics of a subcolumn that
// passes its cloud test and leaves any other's as input
// (rrtmg_sw_cldprmc.f90:127-143), where cldprop_sw sets depth 0, albedo 1;
// with water paths (scaled) the subcolumn's optics are cldprmc_sw's final ones.
template <int JB, bool MC, typename Real>
__device__ __forceinline__ void sw_cloud_layer(Real taug, Real taur, Real prmu0, const Real *__restrict__ exp_tbl,
                                               int fl, double cf, const SwClouds &cl, size_t ix, size_t o,
                                               SwLayerRT<Real> &c, SwLayerRT<Real> &t)
{
    sw_layer_optics(taug, taur, prmu0, exp_tbl, c.r, c.rd, c.t, c.td, c.dbt);
    t = c;
    if (!(fl & SWC_BLEND)) return;
    double pt = 0.0, pw = 1.0, pg = 0.0;
    if (!MC && cl.scaled) {
        // cldprop's final optics (water paths); 0, 1, 0 where it gave none
        const size_t j = ix * 14 + (JB - 16);
        pt = cl.taucld[j];
        pw = cl.ssacld[j];
        pg = cl.asmcld[j];
    } else if constexpr (MC) {
        pt = cl.taucmc[o];
        pw = cl.ssacmc[o];
        pg = cl.asmcmc[o];
        if ((fl & SWC_OPTICS) && !cl.scaled) {
            const double ffp = cl.fsfcmc[o];
            const double ffp1 = 1.0 - ffp, ffpssa = 1.0 - ffp * pw;
            pw = ffp1 * pw / ffpssa;
            pt = ffpssa * pt;
            pg = (pg - ffp) / ffp1;
        }
    } else if (fl & SWC_OPTICS) {
        const size_t j = ix * 14 + (JB - 16);
        const double tauc = cl.taucld[j], ssac = cl.ssacld[j], asmc = cl.asmcld[j], ffp = cl.fsfcld[j];
        const double ffp1 = 1.0 - ffp, ffpssa = 1.0 - ffp * ssac;
        pw = ffp1 * ssac / ffpssa;
        pt = ffpssa * tauc;
        pg = (asmc - ffp) / ffp1;
    }
    const double ztauc = (double)taur + (double)taug;
    t = sw_cloudy_blend<Real>(ztauc, (double)taur / ztauc, pt, pw, pg, cf, fl, (double)prmu0, exp_tbl, c);
}

// adding-method state of one stream (rrtmg_sw_vrtqdr.f90:103-150)
template <typename Real>
struct SwStream {
    Real prup, prupd;        // bottom-up: reflectances of the atmosphere below the current level
    Real ztdn, prdnd, ptdbt; // top-down
    __device__ __forceinline__ void up(const SwLayerRT<Real> &l)
    {
        const Real zreflect = fdiv((Real)1, (Real)1 - prupd * l.rd);
        const Real pu = l.r + (l.td * ((l.t - l.dbt) * prupd + l.dbt * prup)) * zreflect;
        const Real pud = l.rd + l.td * l.td * prupd * zreflect;
        prup = pu;
        prupd = pud;
    }
    __device__ __forceinline__ void down(const SwLayerRT<Real> &l)
    {
        const Real zreflect = fdiv((Real)1, (Real)1 - l.rd * prdnd);
        const Real tdn = ptdbt * l.t + (l.td * ((ztdn - ptdbt) + ptdbt * l.r * prdnd)) * zreflect;
        const Real dnd = l.rd + l.td * l.td * prdnd * zreflect;
        ptdbt = l.dbt * ptdbt;
        ztdn = tdn;
        prdnd = dnd;
    }
    // level fluxes for the reflectances (pu, pud) of the atmosphere below
    __device__ __forceinline__ void flux(Real pu, Real pud, Real &fu, Real &fd) const
    {
        const Real zreflect = fdiv((Real)1, (Real)1 - prdnd * pud);
        fu = (ptdbt * pu + (ztdn - ptdbt) * pud) * zreflect;
        fd = ptdbt + (ztdn - ptdbt + ptdbt * pu * prdnd) * zreflect;
    }
};

// One band JB of one work unit with clouds.  row: this wave's scratch,
// [4][nlay+1][64] (prup, prupd of the clear stream, then of the total one;
// FP32: then the kept cloudy-layer values);
// s_red: 64 LDS entries for the FP64 g-sum.
template <int JB, bool MC, typename Real>
__device__ __forceinline__ void sw_cloud_band_body(const ColumnInputs &in, const SwWork &w, const SwClouds &cl,
                                                   const SwCloudWork &cw, const SwBand &B, const Real *__restrict__ ktab,
                                                   const Real *__restrict__ exp_tbl, WaveUnit u, int tile, int lane,
                                                   Real *__restrict__ row, Real *s_red)
{
    const int ncol = in.ncol, nlay = in.nlay;
    constexpr int ng = sw_ng(JB);
    SwBand Bc = B;
    Bc.nspa = SW_NSPA[JB - 16];
    Bc.nspb = SW_NSPB[JB - 16];
    const int cl_ = lane / ng, g = lane - cl_ * ng;
    const GSum<Real, ng> gsum(lane, g);
    const int col = tile * TILE_COLS + u.first_col + cl_;
    const bool active = cl_ < u.ncols && col < ncol && in.coszen[min(col, ncol - 1)] >= 1.e-10;
    if (!__any(active)) return;  // night: the finish kernel writes zeros
    const Real *coef = (const Real *)w.coef;
    Tab<Real> T{ktab, ng, g};
    const size_t L1 = (size_t)(nlay + 1) * 64;
    Real *s_cu = row, *s_cud = row + L1, *s_tu = row + 2 * L1, *s_tud = row + 3 * L1;

    Real prmu0 = (Real)1, zincflx = (Real)0, albp = (Real)0, albd = (Real)0;
    int laytrop = 0;
    SwLayer<Real> c;
    if (active) {
        prmu0 = (Real)in.coszen[col];
        laytrop = w.laytrop[col];
        const bool nir = (JB <= 24) || JB == 29;  // rrtmg_sw_rad.nomcica.f90:567-577
        albp = (Real)(nir ? in.aldir[col] : in.asdir[col]);
        albd = (Real)(nir ? in.aldif[col] : in.asdif[col]);
        int lsf = w.laysolfr[(size_t)(JB - 16) * ncol + col];
        sw_load_layer(c, coef, w.ipk, lsf - 1, col, ncol, nlay);
        double sfz = sw_sfluxzen(JB, Bc, T, c, lsf <= laytrop);
        zincflx = (Real)(in.adjflux * sfz * (double)prmu0);
    }
    // both streams' layer values of layer `lay` (1-based, bottom-up).  FP32
    // keeps a cloudy layer's blended values from the bottom-up sweep for the
    // top-down one ([5][nlay][64] after the stream rows; measured at
    // 32768 x 40: 3.47 -> 3.30 ms), FP64 evaluates them again (keeping
    // them, 4.99 -> 5.47 ms)
    constexpr bool keep = sizeof(Real) == 4;
    Real *s_keep = row + 4 * L1;
    const size_t LK = (size_t)nlay * 64;
    auto layer = [&](int lay, bool first, SwLayerRT<Real> &lc, SwLayerRT<Real> &lt) {
        lc = SwLayerRT<Real>{(Real)0, (Real)0, (Real)1, (Real)1, (Real)1};
        lt = lc;
        if (!active) return;
        sw_load_layer(c, coef, w.ipk, lay - 1, col, ncol, nlay);
        Real taug, taur;
        sw_gas(JB, Bc, T, c, lay <= laytrop, taug, taur);
        const size_t ix = (size_t)(lay - 1) * ncol + col, o = ix * 112 + Bc.ngs + g;
        int fl = cw.flags[ix];
        double cf = 0.0;
        if constexpr (MC) {
            // the lane's subcolumn: the tests of spcvmc_sw (:388, 502-549) and
            // cldprmc_sw (rrtmg_sw_cldprmc.f90:127-129) on its own fraction
            if (fl & SWC_BLEND) {
                cf = cl.cldfmc[o];
                const double cwp = (cl.ciwpmc ? cl.ciwpmc[o] : 0.0) + (cl.clwpmc ? cl.clwpmc[o] : 0.0);
                fl = (cf > 0.0 ? SWC_BLEND : 0) | (cf > 1.e-12 ? SWC_REFTRA : 0) |
                     (cf >= 1.e-20 && (cwp >= 1.e-20 || cl.taucmc[o] >= 1.e-20) ? SWC_OPTICS : 0);
            }
        } else {
            cf = (fl & SWC_BLEND) ? cl.cldfr[ix] : 0.0;
        }
        const unsigned k = (unsigned)((lay - 1) * 64 + lane);
        if (keep && !first && (fl & SWC_BLEND)) {
            sw_layer_optics(taug, taur, prmu0, exp_tbl, lc.r, lc.rd, lc.t, lc.td, lc.dbt);
            lt.r = at32(s_keep, k);
            lt.rd = at32(s_keep + LK, k);
            lt.t = at32(s_keep + 2 * LK, k);
            lt.td = at32(s_keep + 3 * LK, k);
            lt.dbt = at32(s_keep + 4 * LK, k);
            return;
        }
        sw_cloud_layer<JB, MC, Real>(taug, taur, prmu0, exp_tbl, fl, cf, cl, ix, o, lc, lt);
        if (keep && (fl & SWC_BLEND)) {
            at32(s_keep, k) = lt.r;
            at32(s_keep + LK, k) = lt.rd;
            at32(s_keep + 2 * LK, k) = lt.t;
            at32(s_keep + 3 * LK, k) = lt.td;
            at32(s_keep + 4 * LK, k) = lt.dbt;
        }
    };

    // bottom-up sweep
    SwStream<Real> sc{albp, albd, (Real)1, (Real)0, (Real)1}, st = sc;
    at32(s_cu, (unsigned)(nlay * 64 + lane)) = albp;
    at32(s_cud, (unsigned)(nlay * 64 + lane)) = albd;
    at32(s_tu, (unsigned)(nlay * 64 + lane)) = albp;
    at32(s_tud, (unsigned)(nlay * 64 + lane)) = albd;
    for (int lay = 1; lay <= nlay; ++lay) {
        const int jk = nlay + 1 - lay;
        SwLayerRT<Real> lc, lt;
        layer(lay, true, lc, lt);
        sc.up(lc);
        st.up(lt);
        const unsigned ix = (unsigned)((jk - 1) * 64 + lane);
        at32(s_cu, ix) = sc.prup;
        at32(s_cud, ix) = sc.prupd;
        at32(s_tu, ix) = st.prup;
        at32(s_tud, ix) = st.prupd;
    }

    // top-down sweep with level fluxes
    const size_t bs = (size_t)(nlay + 1) * ncol;
    Real *bf = (Real *)cw.bflux + (size_t)(JB - 16) * 4 * bs;  // [4][nlay+1][ncol]
    auto store = [&](int k, int lev, Real v) {
        if (active && g == 0) bf[(size_t)k * bs + (size_t)lev * ncol + col] = v;
    };
    for (int jk = 1; jk <= nlay + 1; ++jk) {
        if (jk > 1) {
            SwLayerRT<Real> lc, lt;
            layer(nlay + 2 - jk, false, lc, lt);
            sc.down(lc);
            st.down(lt);
        }
        const unsigned ix = (unsigned)((jk - 1) * 64 + lane);
        Real cu, cd, tu, td;
        sc.flux(at32(s_cu, ix), at32(s_cud, ix), cu, cd);
        st.flux(at32(s_tu, ix), at32(s_tud, ix), tu, td);
        const int lev = nlay + 1 - jk;
        store(0, lev, gsum.sum(active ? zincflx * tu : (Real)0, s_red));
        store(1, lev, gsum.sum(active ? zincflx * td : (Real)0, s_red));
        store(2, lev, gsum.sum(active ? zincflx * cu : (Real)0, s_red));
        store(3, lev, gsum.sum(active ? zincflx * cd : (Real)0, s_red));
    }
}

template <typename Real, bool MC>
__global__ __launch_bounds__(64) void sw_cloud_band_kernel(ColumnInputs in, SwWork w, SwClouds cl, SwCloudWork cw,
                                                           const SwBand *__restrict__ bands, int nwaves, int blk0,
                                                           int nblocks, int exp_off, const Real *__restrict__ tab,
                                                           const WaveUnit *__restrict__ units)
{
    __shared__ Real s_red[64];
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const SwBand B = bands[u.band];
    Real *row = (Real *)cw.rows + (size_t)blk * (sw_cloud_row_bytes(in.nlay, sizeof(Real)) / sizeof(Real));
    const Real *exp_tbl = tab + exp_off;
#define BODY(jb)                                                                                                \
    case jb - 16:                                                                                               \
        sw_cloud_band_body<jb, MC, Real>(in, w, cl, cw, B, tab, exp_tbl, u, tile, (int)threadIdx.x, row, s_red);\
        break;
    switch (u.band) { SW_BANDS(BODY) }
#undef BODY
}

// finish: band sums, net fluxes and heating of both streams
// (rrtmg_sw_rad.nomcica.f90:700-726), one thread per (layer, column); night
// columns zero
template <typename Real>
__global__ __launch_bounds__(256) void sw_cloud_finish_kernel(ColumnInputs in, SwCloudWork cw, SwOutputs out,
                                                              double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    const Real *bflux = (const Real *)cw.bflux;
    const size_t bs = (size_t)(nlay + 1) * ncol;
    const size_t il = (size_t)lay * ncol + col;
    const bool night = in.coszen[col] < 1.e-10;
    double net[2][2];
    for (int j = 0; j < 2; ++j) {
        const int k = lay + j;
        double f[4] = {0.0, 0.0, 0.0, 0.0};
        if (!night)
            for (int b = 0; b < 14; ++b)
                for (int s = 0; s < 4; ++s) f[s] += (double)bflux[((size_t)b * 4 + s) * bs + (size_t)k * ncol + col];
        if (j == 0 || lay == nlay - 1) {
            const size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = f[0];
            out.dflx[ix] = f[1];
            out.uflxc[ix] = f[2];
            out.dflxc[ix] = f[3];
        }
        net[0][j] = f[1] - f[0];
        net[1][j] = f[3] - f[2];
    }
    const double zdpgcp = heatfac / (in.plev[il] - in.plev[il + ncol]);
    const bool top = lay == nlay - 1;
    out.hr[il] = (night || top) ? 0.0 : (net[0][1] - net[0][0]) * zdpgcp;
    out.hrc[il] = (night || top) ? 0.0 : (net[1][1] - net[1][0]) * zdpgcp;
}

hipError_t launch_sw_cloud_flags(const ColumnInputs &in, const SwClouds &cl, const SwCloudWork &cw, hipStream_t s)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    hipLaunchKernelGGL(sw_cloud_zero_kernel, dim3(1), dim3(1), 0, s, cw.partial);
    if (cl.mcica && !cl.cldfmc) {
        // McICA driver with icld = 0: no subcolumns, every layer clear
        hipLaunchKernelGGL(sw_flags_clear_kernel, dim3(nbl), dim3(256), 0, s, (size_t)in.nlay * in.ncol, cw.flags);
    } else if (cl.mcica) {
        const size_t n2 = (size_t)in.nlay * in.ncol;
        hipLaunchKernelGGL(sw_mcica_flags_kernel, dim3((unsigned)((n2 + SWMC_FLAG_PAIRS - 1) / SWMC_FLAG_PAIRS)),
                           dim3(256), 0, s, in.ncol, in.nlay, cl, cw);
    } else {
        hipLaunchKernelGGL(sw_cloud_flags_kernel, dim3(nbl), dim3(256), 0, s, in, cl, cw);
    }
    return hipGetLastError();
}

template <typename Real>
static hipError_t launch_sw_cloud_t(const ColumnInputs &in, const SwClouds &cl, const SwOutputs &out,
                                    const SwWork &w, const SwCloudWork &cw, const SwDevTables &t, int mode,
                                    hipStream_t s, hipEvent_t *ev)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipError_t e = launch_sw_prep(in, w, t, mode, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const int nblocks = ntiles * t.desc.nwaves;
    for (int b0 = 0; b0 < nblocks; b0 += cw.chunk) {
        const int nb_c = std::min(cw.chunk, nblocks - b0);
        if (cl.mcica)
            hipLaunchKernelGGL((sw_cloud_band_kernel<Real, true>), dim3(xcd_grid(nb_c)), dim3(64), 0, s, in, w, cl, cw,
                               t.bands, t.desc.nwaves, b0, nb_c, t.desc.exp_tbl, (const Real *)t.tab, t.units);
        else
            hipLaunchKernelGGL((sw_cloud_band_kernel<Real, false>), dim3(xcd_grid(nb_c)), dim3(64), 0, s, in, w, cl,
                               cw, t.bands, t.desc.nwaves, b0, nb_c, t.desc.exp_tbl, (const Real *)t.tab, t.units);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(sw_cloud_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, cw, out, t.heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_sw_cloud(const ColumnInputs &in, const SwClouds &cl, const SwOutputs &out, const SwWork &w,
                           const SwCloudWork &cw, const SwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev)
{
    return mode == 1 ? launch_sw_cloud_t<float>(in, cl, out, w, cw, t, mode, s, ev)
                     : launch_sw_cloud_t<double>(in, cl, out, w, cw, t, mode, s, ev);
}

}  // namespace rrtmg
