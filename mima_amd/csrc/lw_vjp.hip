// Backward (vector-Jacobian product) of the clear-sky LW path on gfx950:
// given cotangents of uflx, dflx (levels) and hr (layers), the gradients with
// respect to layer temperature, level temperature, water vapour, ozone and
// surface temperature.  No counterpart in the reference, whose only
// derivative is the idrv = 1 surface-Planck term dF/dTs
// (rrtmg_lw_rtrnmr.f90:629-746); the forward being differentiated is
// rrtmg_lw (rrtmg_lw_rad.nomcica.f90:80) as the forward kernels implement it.
//
// Kernels:
//  * prep (one thread per (layer, column)): inatm + setcoef in FP64 dual
//    numbers seeded with (tlay, h2o, o3) of each layer, writing every
//    per-layer coefficient's value (the forward's record) and its nonzero
//    derivatives (a second record, LW_DMASK); level Planck indices; and the per-level
//    radiance cotangents after folding the heating-rate cotangent into the
//    fluxes; then a per-column kernel for laytrop, the column water and
//    dry-air totals, precipitable water and surface temperature.
//  * band (one wave per work unit, lane = (column, g) as in the forward):
//    pass A, downward sweep (values), storing per layer the transmittance
//    in LDS and the upward source and the incoming downward radiance in a
//    per-wave global scratch row (so LDS does not cap occupancy); pass B,
//    upward sweep, replacing the source by (source - incoming upward
//    radiance); pass C, reverse of the upward sweep, top-down, leaving the
//    adjoints of the upward source and of the transmittance (upward part) in
//    place of those two; pass D, reverse of the surface term and of the downward sweep,
//    bottom-up: the layer's optical depth and Planck fraction are recomputed
//    as dual numbers (no g-point-by-level tape is kept), the local derivatives
//    of transmittance and source with respect to optical depth, Planck
//    fraction and Planck integrals are analytic (for table lookups, those of
//    the functions the tables tabulate), and the per-lane gradient is reduced
//    over the band's g-points.
//  * finish (one thread per (layer, column)): sums over bands in FP64 and
//    the precipitable-water chain of the diffusivity angle.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "lw_optics.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

using D3 = Dual<double, 3>;

// value (V = Real) or value and three derivatives (V = Dual<Real, 3>) of one
// field: v, the field in the cell's value record; d, its first derivative in
// the cell's derivative record.  MASK (bit 0: d/dT, 1: d/dq, 2: d/dO3, the
// field's LW_DMASK entry) names the derivatives the field can have; the
// others are known zeros, not stored, so the dual arithmetic on them folds
// away at compile time.
enum : unsigned { DT = 1u, DQ = 2u, DO = 4u };
template <typename V, typename Real, unsigned MASK>
struct DLoad;
template <typename Real, unsigned MASK>
struct DLoad<Real, Real, MASK> {
    __device__ static Real get(const Real *__restrict__ v, const Real *__restrict__) { return v[0]; }
};
template <typename Real, unsigned MASK>
struct DLoad<Dual<Real, 3>, Real, MASK> {
    __device__ static Dual<Real, 3> get(const Real *__restrict__ v, const Real *__restrict__ d)
    {
        Dual<Real, 3> r;
        r.v = v[0];
        r.d[0] = (MASK & DT) ? d[0] : (Real)0;
        r.d[1] = (MASK & DQ) ? d[dmask_count(MASK & DT)] : (Real)0;
        r.d[2] = (MASK & DO) ? d[dmask_count(MASK & (DT | DQ))] : (Real)0;
        return r;
    }
};

// the layer's coefficients from the backward's records (values only when V
// is Real: the derivative record is not read)
template <typename V, typename Real>
__device__ __forceinline__ void lw_load_layer_d(LwLayer<V> &c, const Real *__restrict__ dco,
                                                const Real *__restrict__ dde, const int *__restrict__ ipk, int lay,
                                                int col, int ncol)
{
    const unsigned ix = (unsigned)(lay * ncol + col);
    const Real *vr = (const Real *)__builtin_assume_aligned(dco + (size_t)ix * LWF_STRIDE, 16);
    const Real *dr = (const Real *)__builtin_assume_aligned(dde + (size_t)ix * LW_DSTRIDE, 16);
#define LDD(f, F) c.f = DLoad<V, Real, LW_DMASK[F]>::get(vr + (F), dr + der_offset(LW_DMASK, F))
    LDD(fac00, LWF_FAC00); LDD(fac01, LWF_FAC01); LDD(fac10, LWF_FAC10); LDD(fac11, LWF_FAC11);
    LDD(colh2o, LWF_COLH2O); LDD(colco2, LWF_COLCO2); LDD(colo3, LWF_COLO3);
    LDD(coln2o, LWF_COLN2O); LDD(colco, LWF_COLCO); LDD(colch4, LWF_COLCH4); LDD(colo2, LWF_COLO2);
    LDD(colbrd, LWF_COLBRD); LDD(selffac, LWF_SELFFAC); LDD(selffrac, LWF_SELFFRAC);
    LDD(forfac, LWF_FORFAC); LDD(forfrac, LWF_FORFRAC); LDD(minorfrac, LWF_MINORFRAC);
    LDD(scaleminor, LWF_SCALEMINOR); LDD(scaleminorn2, LWF_SCALEMINORN2);
    LDD(coldry, LWF_COLDRY); LDD(pavel, LWF_PAVEL); LDD(wx1, LWF_WX1); LDD(wx2, LWF_WX2);
    LDD(wx3, LWF_WX3); LDD(wx4, LWF_WX4); LDD(tlayfrac, LWF_TLAYFRAC);
#undef LDD
    const int w = at32(ipk, ix);
    c.jp = unpack_jp(w);
    c.jt = unpack_jt(w);
    c.jt1 = unpack_jt1(w);
    c.inds = unpack_inds(w);
    c.indf = unpack_indf(w);
    c.indm = unpack_indm(w);
    c.indlay = unpack_indlay(w);
}

// ---------------------------------------------------------------------------
// prep: the forward prep's arithmetic (lw.hip, rrtmg_lw_rad.nomcica.f90:729-852,
// rrtmg_lw_setcoef.f90:150-413) on dual numbers, one thread per (layer,
// column): the layer's coefficients with their derivatives, the Planck index
// of the level above it (and of the surface level), and the radiance
// cotangents of its lower level (and of the top level)
template <typename Real>
__global__ __launch_bounds__(256) void lw_vjp_prep_kernel(ColumnInputs in, VjpCot cot, LwVjpWork w,
                                                          const LwConst *__restrict__ cst, double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    Real *dco = (Real *)w.dcoef, *dde = (Real *)w.dder;
    Real *tlevf = (Real *)w.tlev;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    // the value record and the derivative record of the (layer, column)
    auto put = [&](int F, size_t ix, const D3 &x) {
        dco[ix * LWF_STRIDE + F] = (Real)x.v;
        Real *d = dde + ix * LW_DSTRIDE + der_offset(LW_DMASK, F);
        const unsigned m = LW_DMASK[F];
        int k = 0;
        for (int j = 0; j < 3; ++j)
            if (m & (1u << j)) d[k++] = (Real)x.d[j];
    };
    if (lay == 0) {
        double tz0 = in.tlev[col];
        int i0 = clampi((int)(tz0 - 159.0), 1, 180);
        w.ilev[col] = i0;
        tlevf[col] = (Real)(tz0 - 159.0 - (double)i0);
    }
    {
        const size_t ix = (size_t)lay * ncol + col;
        const double pavel = in.play[ix];
        const D3 tavel = D3::seed(in.tlay[ix], 0);
        const D3 h2o = D3::seed(in.h2o[ix], 1);
        const D3 o3 = D3::seed(in.o3[ix], 2);
        const double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        D3 wkl1 = (h2o / (1.0 - h2o)) * amdw;
        const double wkl2 = in.co2[ix], wkl4 = in.n2o[ix], wkl6 = in.ch4[ix], wkl7 = in.o2[ix];
        const D3 wkl3 = o3 * amdo;
        const D3 amm = (1.0 - wkl1) * amd + wkl1 * amw;
        const D3 coldry = ((pz0 - pz1) * 1.e3 * avogad) / (1.e2 * grav * amm * (1.0 + wkl1));
        D3 summol(0.0);
        summol = summol + wkl2;
        summol = summol + wkl3;
        summol = summol + wkl4;
        summol = summol + 0.0;
        summol = summol + wkl6;
        summol = summol + wkl7;
        const D3 wbrodl = coldry * (1.0 - summol);
        wkl1 = coldry * wkl1;
        const D3 wx1 = coldry * in.ccl4[ix] * 1.e-20, wx2 = coldry * in.cfc11[ix] * 1.e-20;
        const D3 wx3 = coldry * in.cfc12[ix] * 1.e-20, wx4 = coldry * in.cfc22[ix] * 1.e-20;

        const int indlay = clampi((int)(tavel.v - 159.0), 1, 180);
        const D3 tlayfrac = tavel - 159.0 - (double)indlay;
        const double tz = in.tlev[ix + ncol];
        const int indlev = clampi((int)(tz - 159.0), 1, 180);
        w.ilev[ix + ncol] = indlev;
        tlevf[ix + ncol] = (Real)(tz - 159.0 - (double)indlev);

        const double plog = log(pavel);
        const int jp = clampi((int)(36.0 - 5 * (plog + 0.04)), 1, 58);
        const double fp = 5.0 * (cst->preflog[jp - 1] - plog);
        const D3 dt0 = (tavel - cst->tref[jp - 1]) / 15.0;
        const int jt = clampi((int)(3.0 + dt0.v), 1, 4);
        const D3 ft = dt0 - (double)(jt - 3);
        const D3 dt1 = (tavel - cst->tref[jp]) / 15.0;
        const int jt1 = clampi((int)(3.0 + dt1.v), 1, 4);
        const D3 ft1 = dt1 - (double)(jt1 - 3);
        const D3 water = wkl1 / coldry;
        const D3 scalefac = pavel * stpfac / tavel;
        const D3 forfac = scalefac / (1. + water);
        D3 forfrac, selffac = water * forfac, selffrac(0.0);
        int indfor, indself = 0;
        if (plog > 4.56) {
            D3 factor = (332.0 - tavel) / 36.0;
            indfor = clampi((int)factor.v, 1, 2);
            forfrac = factor - (double)indfor;
            factor = (tavel - 188.0) / 7.2;
            indself = clampi((int)factor.v - 7, 1, 9);
            selffrac = factor - (double)(indself + 7);
        } else {
            const D3 factor = (tavel - 188.0) / 36.0;
            indfor = 3;
            forfrac = factor - 1.0;
        }
        const D3 scaleminor = pavel / tavel;
        const D3 scaleminorn2 = (pavel / tavel) * (wbrodl / (coldry + wkl1));
        const D3 factor = (tavel - 180.8) / 7.2;
        const int indminor = clampi((int)factor.v, 1, 18);
        const D3 minorfrac = factor - (double)indminor;
        const D3 colh2o = 1.e-20 * wkl1;
        D3 colco2 = 1.e-20 * (coldry * wkl2);
        D3 colo3 = 1.e-20 * (coldry * wkl3);
        D3 coln2o = 1.e-20 * (coldry * wkl4);
        D3 colco(0.0);
        D3 colch4 = 1.e-20 * (coldry * wkl6);
        const D3 colo2 = 1.e-20 * (coldry * wkl7);
        if (colco2.v == 0.0) colco2 = 1.e-32 * coldry;
        if (colo3.v == 0.0) colo3 = 1.e-32 * coldry;
        if (coln2o.v == 0.0) coln2o = 1.e-32 * coldry;
        if (colco.v == 0.0) colco = 1.e-32 * coldry;
        if (colch4.v == 0.0) colch4 = 1.e-32 * coldry;
        const D3 colbrd = 1.e-20 * wbrodl;
        const double compfp = 1. - fp;
        put(LWF_FAC10, ix, compfp * ft);
        put(LWF_FAC00, ix, compfp * (1.0 - ft));
        put(LWF_FAC11, ix, fp * ft1);
        put(LWF_FAC01, ix, fp * (1.0 - ft1));
        put(LWF_COLH2O, ix, colh2o);
        put(LWF_COLCO2, ix, colco2);
        put(LWF_COLO3, ix, colo3);
        put(LWF_COLN2O, ix, coln2o);
        put(LWF_COLCO, ix, colco);
        put(LWF_COLCH4, ix, colch4);
        put(LWF_COLO2, ix, colo2);
        put(LWF_COLBRD, ix, colbrd);
        put(LWF_SELFFAC, ix, colh2o * selffac);
        put(LWF_SELFFRAC, ix, selffrac);
        put(LWF_FORFAC, ix, colh2o * forfac);
        put(LWF_FORFRAC, ix, forfrac);
        put(LWF_MINORFRAC, ix, minorfrac);
        put(LWF_SCALEMINOR, ix, scaleminor);
        put(LWF_SCALEMINORN2, ix, scaleminorn2);
        put(LWF_COLDRY, ix, coldry);
        put(LWF_PAVEL, ix, D3(pavel));
        put(LWF_WX1, ix, wx1);
        put(LWF_WX2, ix, wx2);
        put(LWF_WX3, ix, wx3);
        put(LWF_WX4, ix, wx4);
        put(LWF_TLAYFRAC, ix, tlayfrac);
        w.ipk[ix] = pack_lw(jp, jt, jt1, indself, indfor, indminor, indlay);
    }

    // level radiance cotangents: fluxes' own cotangents plus the heating rate
    // hr(k) = heatfac (fnet(k) - fnet(k+1)) / (p(k) - p(k+1)), fnet = up - dn
    // (rrtmg_lw_rtrnmr.f90:767-772); this thread's level lay, and the top
    Real *cu = (Real *)w.cot, *cd = cu + (size_t)(nlay + 1) * ncol;
    for (int k = lay; k <= (lay == nlay - 1 ? nlay : lay); ++k) {
        const size_t ix = (size_t)k * ncol + col;
        double fbar = 0.0;
        if (cot.hr) {
            if (k < nlay) fbar += heatfac * cot.hr[ix] / (in.plev[ix] - in.plev[ix + ncol]);
            if (k > 0) fbar -= heatfac * cot.hr[ix - ncol] / (in.plev[ix - ncol] - in.plev[ix]);
        }
        cu[ix] = (Real)((cot.uflx ? cot.uflx[ix] : 0.0) + fbar);
        cd[ix] = (Real)((cot.dflx ? cot.dflx[ix] : 0.0) - fbar);
    }
}

// per-column part of the prep (one thread per column): laytrop, the column
// water and dry-air totals (the forward's arithmetic and summation order),
// precipitable water and surface temperature
__global__ __launch_bounds__(256) void lw_vjp_column_kernel(ColumnInputs in, LwVjpWork w)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, avogad = 6.02214199e+23, grav = 9.8066;
    double amttl = 0.0, wvttl = 0.0;
    int laytrop = 0;
    for (int lay = 0; lay < nlay; ++lay) {
        const size_t ix = (size_t)lay * ncol + col;
        const double h = in.h2o[ix], pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double wkl1 = (h / (1.0 - h)) * amdw;
        const double amm = (1.0 - wkl1) * amd + wkl1 * amw;
        const double coldry = ((pz0 - pz1) * 1.e3 * avogad) / (1.e2 * grav * amm * (1.0 + wkl1));
        wkl1 = coldry * wkl1;
        amttl = amttl + coldry + wkl1;
        wvttl = wvttl + wkl1;
        if (log(in.play[ix]) > 4.56) ++laytrop;
    }
    const double wvsh = (amw * wvttl) / (amd * amttl);
    w.laytrop[col] = laytrop;
    w.colscal[col] = wvsh * (1.e3 * in.plev[col]) / (1.e2 * grav);
    w.colscal[(size_t)ncol + col] = in.tsfc[col];
    w.colscal[(size_t)2 * ncol + col] = wvttl;
    w.colscal[(size_t)3 * ncol + col] = amttl;
}

// transmittance and sources of one layer (rrtmg_lw_rtrnmr.f90:591-606) and,
// for the backward, their derivatives with respect to the optical depth.
// Values: the reference's thin-layer series below 0.06 and the tables above,
// or (smooth mode, tables NULL) the tabulated functions themselves at every
// depth; derivatives: always those of the smooth functions
template <typename Real>
struct LwLayerRT {
    Real od, atrans, tf, datr, dtf;
};
template <typename Real>
__device__ __forceinline__ LwLayerRT<Real> lw_layer_rt(Real secdiff, Real taug, const Real *__restrict__ exp_tbl,
                                                       const Real *__restrict__ tfn_tbl, bool derivs)
{
    LwLayerRT<Real> r;
    Real od = secdiff * taug;
    if (od < (Real)0) od = (Real)0;
    r.od = od;
    if (!exp_tbl) {
        const LwRT<Real> x = lw_rt_exact(od, derivs);
        r.atrans = x.atrans;
        r.tf = x.tf;
        r.datr = x.datr;
        r.dtf = x.dtf;
    } else {
        if (od <= (Real)0.06) {
            r.atrans = od - (Real)0.5 * od * od;
            r.tf = (Real)0.166667 * od;
        } else {
            const int itr = tbl_index(od);
            r.atrans = (Real)1 - at32(exp_tbl, (unsigned)itr);
            r.tf = at32(tfn_tbl, (unsigned)itr);
        }
        r.datr = r.dtf = (Real)0;
        if (derivs) {
            // the derivatives of the smooth functions that the series and the
            // tables stand for, 1 - exp(-tau) and the tau transition
            // function, at the same optical depth
            const LwRT<Real> x = lw_rt_exact(od, true);
            r.datr = x.datr;
            r.dtf = x.dtf;
        }
    }
    return r;
}

// One band of one work unit of the backward (passes A-D, described at the top
// of this file).  ktab: the k-tables (the packed global buffer, or the band's
// copy in LDS with rebased offsets in B); tab: the packed global buffer
// (transmittance and Planck tables).  s1: transmittance per layer, [nlay][64]
// (LDS or global); s3: this wave's scratch row, incoming downward radiance
// and upward source, [2][nlay][64]; red: [4][64] LDS entries for FP64 g-sums.
template <int BAND, typename Real>
__device__ __forceinline__ void lw_vjp_band_body(const ColumnInputs &in, const LwVjpWork &w, const LwBand &B,
                                 const Real *__restrict__ ktab, const Real *__restrict__ tab, int exp_off,
                                 int tfn_off, int plk_off, const LwConst *__restrict__ cst, WaveUnit u, int tile,
                                 int lane, Real *__restrict__ s1, Real *__restrict__ s3, Real *red)
{
    constexpr int NG = lw_ng(BAND);
    using V3 = Dual<Real, 3>;
    const int ncol = in.ncol, nlay = in.nlay;
    const int cl = lane / NG, g = lane - cl * NG;
    const int col0 = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col0 < ncol;
    const int col = active ? col0 : 0;
    Real *__restrict__ s2 = s3 + (size_t)nlay * 64;
    // s3: incoming downward radiance per layer, [nlay][64] in this wave's
    // global scratch row (written in pass A, read back in pass D)
    const Real *dco = (const Real *)w.dcoef, *dde = (const Real *)w.dder;
    const Real *tlevf = (const Real *)w.tlev;
    const Real *exp_tbl = exp_off >= 0 ? tab + exp_off : nullptr;
    const Real *tfn_tbl = tfn_off >= 0 ? tab + tfn_off : nullptr;
    const Real *plk = tab + plk_off + (BAND - 1) * 181;
    LTab<Real> T{ktab, NG, g};
    const size_t n2 = (size_t)nlay * ncol, n3 = (size_t)(nlay + 1) * ncol;
    Real *gpart = (Real *)w.gpart + (size_t)(BAND - 1) * 3 * n2;
    Real *glev = (Real *)w.glev + (size_t)(BAND - 1) * n3;
    Real *gsc = (Real *)w.gsc + (size_t)(BAND - 1) * 2 * ncol;

    const int laytrop = w.laytrop[col];
    const double pwvcm = w.colscal[col];
    const double tbound = w.colscal[(size_t)ncol + col];
    double sd = 1.66, dsd = 0.0;
    if (!(BAND == 1 || BAND == 4 || BAND >= 10)) {
        const double e = exp(B.a2 * pwvcm);
        sd = B.a0 + B.a1 * e;
        dsd = B.a1 * B.a2 * e;
        if (sd > 1.80) { sd = 1.80; dsd = 0.0; }
        if (sd < 1.50) { sd = 1.50; dsd = 0.0; }
    }
    const Real secdiff = (Real)sd;
    const Real semiss = (Real)in.emis[(size_t)(BAND - 1) * ncol + col];
    const int indbound = clampi((int)(tbound - 159.0), 1, 180);
    const double tbndfrac = tbound - 159.0 - (double)indbound;
    const Real plankbnd = (Real)((double)semiss * ((double)plk[indbound - 1] +
                                                   tbndfrac * ((double)plk[indbound] - (double)plk[indbound - 1])));
    const Real dplankbnd = semiss * (plk[indbound] - plk[indbound - 1]);
    auto planklev = [&](int lev) {
        const int i = w.ilev[(size_t)lev * ncol + col];
        const Real f = tlevf[(size_t)lev * ncol + col];
        return plk[i - 1] + f * (plk[i] - plk[i - 1]);
    };
    auto plkslope = [&](int lev) {
        const int i = w.ilev[(size_t)lev * ncol + col];
        return plk[i] - plk[i - 1];
    };
    const double *tauaer = in.tauaer;
    auto tau_a = [&](int lev) {
        return tauaer ? (Real)tauaer[((size_t)(BAND - 1) * nlay + (lev - 1)) * ncol + col] : (Real)0;
    };
    // radiance cotangents of this band: flux = fluxfac * sum_b 0.5 delwave_b sum_g rad
    const Real cband = (Real)(2.0 * asin(1.0) * 2.e4 * 0.5 * B.delwave);
    const Real *cu = (const Real *)w.cot, *cd = cu + n3;
    auto ru = [&](int k) { return active ? cu[(size_t)k * ncol + col] * cband : (Real)0; };
    auto rd = [&](int k) { return active ? cd[(size_t)k * ncol + col] * cband : (Real)0; };

    // band fluxes of the recomputed forward, when asked for (as the forward
    // band kernel writes them; uniform branch)
    Real *bf_up = w.bflux ? (Real *)w.bflux + (size_t)(BAND - 1) * 2 * n3 : nullptr;
    Real *bf_dn = bf_up ? bf_up + n3 : nullptr;
    if (bf_dn && active && g == 0) bf_dn[(size_t)nlay * ncol + col] = (Real)0;

    // pass A: downward sweep
    Real radld = (Real)0, fracs1 = (Real)0;
    Real plev_up = planklev(nlay);
    {
        LwLayer<Real> c;
        for (int lev = nlay; lev >= 1; --lev) {
            lw_load_layer_d<Real, Real>(c, dco, dde, w.ipk, lev - 1, col, ncol);
            Real taug, plfrac;
            lw_gas<BAND, Real>(B, T, cst, c, lev <= laytrop, taug, plfrac);
            taug = taug + tau_a(lev);
            const Real plev_dn = planklev(lev - 1);
            const Real blay = plk[c.indlay - 1] + c.tlayfrac * (plk[c.indlay] - plk[c.indlay - 1]);
            const LwLayerRT<Real> rt = lw_layer_rt(secdiff, taug, exp_tbl, tfn_tbl, false);
            const Real bbd = plfrac * (blay + rt.tf * (plev_dn - blay));
            const Real bbu = plfrac * (blay + rt.tf * (plev_up - blay));
            at32(s1, (unsigned)((lev - 1) * 64 + lane)) = rt.atrans;
            at32(s2, (unsigned)((lev - 1) * 64 + lane)) = bbu;
            at32(s3, (unsigned)((lev - 1) * 64 + lane)) = radld;
            radld = radld + (bbd - radld) * rt.atrans;
            if (bf_dn) {
                const Real sd = seg_sum(active ? radld : (Real)0, lane, g, NG, red);
                if (active && g == 0) bf_dn[(size_t)(lev - 1) * ncol + col] = sd;
            }
            plev_up = plev_dn;
            if (lev == 1) fracs1 = plfrac;
        }
    }
    // surface and pass B: upward sweep; keep source minus incoming radiance
    Real radlu = fracs1 * plankbnd + ((Real)1 - semiss) * radld;
    if (bf_up) {
        const Real su = seg_sum(active ? radlu : (Real)0, lane, g, NG, red);
        if (active && g == 0) bf_up[col] = su;
    }
    // the per-layer states are read one level ahead
    Real s1n = at32(s1, (unsigned)lane), s2n = at32(s2, (unsigned)lane);
    for (int lev = 1; lev <= nlay; ++lev) {
        const Real atr = s1n;
        const Real dsrc = s2n - radlu;
        if (lev < nlay) {
            s1n = at32(s1, (unsigned)(lev * 64 + lane));
            s2n = at32(s2, (unsigned)(lev * 64 + lane));
        }
        at32(s2, (unsigned)((lev - 1) * 64 + lane)) = dsrc;
        radlu = radlu + dsrc * atr;
        if (bf_up) {
            const Real su = seg_sum(active ? radlu : (Real)0, lane, g, NG, red);
            if (active && g == 0) bf_up[(size_t)lev * ncol + col] = su;
        }
    }
    // pass C: reverse of the upward sweep (top-down)
    Real a_ru = ru(nlay);
    s1n = at32(s1, (unsigned)((nlay - 1) * 64 + lane));
    s2n = at32(s2, (unsigned)((nlay - 1) * 64 + lane));
    for (int lev = nlay; lev >= 1; --lev) {
        const unsigned k = (unsigned)((lev - 1) * 64 + lane);
        const Real atr = s1n, dsrc = s2n;
        if (lev > 1) {
            s1n = at32(s1, (unsigned)((lev - 2) * 64 + lane));
            s2n = at32(s2, (unsigned)((lev - 2) * 64 + lane));
        }
        at32(s1, k) = a_ru * atr;   // adjoint of the upward source of layer lev
        at32(s2, k) = a_ru * dsrc;  // adjoint of its transmittance, upward part
        a_ru = a_ru * ((Real)1 - atr) + ru(lev - 1);
    }
    // reverse of the surface term
    const Real a_fracs1 = a_ru * plankbnd;
    Real a_tb = a_ru * fracs1 * dplankbnd;
    Real a_rd = a_ru * ((Real)1 - semiss) + rd(0);
    Real a_sec = (Real)0, pend_up = (Real)0;
    // pass D: reverse of the downward sweep (bottom-up) with the layer's
    // optical depth and Planck fraction recomputed as dual numbers
    for (int lev = 1; lev <= nlay; ++lev) {
        LwLayer<V3> c;
        lw_load_layer_d<V3, Real>(c, dco, dde, w.ipk, lev - 1, col, ncol);
        V3 tg3, pf3;
        lw_gas<BAND, Real, V3>(B, T, cst, c, lev <= laytrop, tg3, pf3);
        const Real tg = tg3.v + tau_a(lev), pf = pf3.v;
        const Real plev_dn = planklev(lev - 1), plev_upv = planklev(lev);
        const Real slope = plk[c.indlay] - plk[c.indlay - 1];
        const Real blay = plk[c.indlay - 1] + c.tlayfrac.v * slope;
        const LwLayerRT<Real> rt = lw_layer_rt(secdiff, tg, exp_tbl, tfn_tbl, true);
        const Real dpdn = plev_dn - blay, dpup = plev_upv - blay;
        const Real bbd = pf * (blay + rt.tf * dpdn);
        const size_t k = (size_t)(lev - 1) * 64 + lane;
        const Real a_bbd = a_rd * rt.atrans;
        const Real a_atr = a_rd * (bbd - at32(s3, (unsigned)k)) + at32(s2, (unsigned)k);
        const Real a_bbu = at32(s1, (unsigned)k);
        a_rd = a_rd * ((Real)1 - rt.atrans) + (lev < nlay ? rd(lev) : (Real)0);
        Real a_pf = a_bbd * (blay + rt.tf * dpdn) + a_bbu * (blay + rt.tf * dpup);
        if (lev == 1) a_pf += a_fracs1;
        const Real a_tf = pf * (a_bbd * dpdn + a_bbu * dpup);
        const Real a_blay = pf * ((Real)1 - rt.tf) * (a_bbd + a_bbu);
        const Real a_pdn = a_bbd * pf * rt.tf, a_pup = a_bbu * pf * rt.tf;
        const Real a_od = rt.od > (Real)0 ? a_atr * rt.datr + a_tf * rt.dtf : (Real)0;
        const Real a_tg = a_od * secdiff;
        a_sec += a_od * tg;
        Real v[4];
        v[0] = a_tg * tg3.d[0] + a_pf * pf3.d[0] + a_blay * slope;
        v[1] = a_tg * tg3.d[1] + a_pf * pf3.d[1];
        v[2] = a_tg * tg3.d[2] + a_pf * pf3.d[2];
        v[3] = (pend_up + a_pdn) * plkslope(lev - 1);  // level lev-1 is complete
        pend_up = a_pup;
        if (!active) v[0] = v[1] = v[2] = v[3] = (Real)0;
        seg_sum_n<4>(v, lane, g, NG, red);
        if (active && g == 0) {
            const size_t ix = (size_t)(lev - 1) * ncol + col;
            gpart[ix] = v[0];
            gpart[n2 + ix] = v[1];
            gpart[2 * n2 + ix] = v[2];
            glev[ix] = v[3];
        }
    }
    {
        Real v[3];
        v[0] = active ? pend_up * plkslope(nlay) : (Real)0;
        v[1] = active ? a_tb : (Real)0;
        v[2] = active ? a_sec * (Real)dsd : (Real)0;
        seg_sum_n<3>(v, lane, g, NG, red);
        if (active && g == 0) {
            glev[(size_t)nlay * ncol + col] = v[0];
            gsc[col] = v[1];
            gsc[ncol + col] = v[2];
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 3 : 1)))
void lw_vjp_band_kernel(ColumnInputs in, LwVjpWork w, const LwBand *__restrict__ bands,
                                                         int nwaves, int blk0, int nblocks, int exp_off, int tfn_off,
                                                         int plk_off, const Real *__restrict__ tab,
                                                         const LwConst *__restrict__ cst,
                                                         const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const LwBand B = bands[u.band];
    Real *s3 = (Real *)w.dstate + (size_t)blk * 2 * in.nlay * 64;
    Real *s1 = smem, *red = smem + (size_t)in.nlay * 64;
#define BODY(b)                                                                                          \
    case b - 1:                                                                                          \
        lw_vjp_band_body<b, Real>(in, w, B, tab, tab, exp_off, tfn_off, plk_off, cst, u, tile, (int)threadIdx.x, \
                                  s1, s3, red);                                                          \
        break;
    switch (u.band) {
        BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
        BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
    }
#undef BODY
}

// FP32 backward band kernel with the band's k-tables in LDS, scheduled as the
// forward's lw_band_lds_kernel (persistent workgroups of LWV_WGW waves that
// claim work items from a global counter and copy a band's k-tables into LDS
// before its first item).  All three per-layer states of a wave live in its
// global scratch row, [3][nlay][64].
template <typename Real>
__global__ __launch_bounds__(64 * LWV_WGW) void lw_vjp_band_lds_kernel(ColumnInputs in, LwVjpWork w,
                                                                       const LwBand *__restrict__ bands,
                                                                       LwLdsPlan plan, int nwaves, int nitems,
                                                                       int exp_off, int tfn_off, int plk_off,
                                                                       const Real *__restrict__ tab,
                                                                       const LwConst *__restrict__ cst,
                                                                       const WaveUnit *__restrict__ units)
{
    static_assert(sizeof(Real) == 4, "FP32 only: the FP64 g-sum synchronises the workgroup");
    extern __shared__ unsigned char smem_raw[];
    Real *ktab = (Real *)smem_raw;
    const int nlay = in.nlay;
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    Real *s3 = (Real *)w.dstate + ((size_t)blockIdx.x * LWV_WGW + wave) * 3 * nlay * 64;
    Real *s1 = s3 + (size_t)2 * nlay * 64;
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
            for (int i = threadIdx.x; i < n; i += 64 * LWV_WGW) ktab[i] = tab[t0 + i];
            __syncthreads();
            loaded = b;
        }
        const int k = (r - plan.ubeg[b]) * LWV_WGW + wave;
        const int tile = group * LWV_WGW + k / plan.upt[b];
        if (tile >= ntiles) continue;
        const WaveUnit u = units[plan.ubeg[b] + k % plan.upt[b]];
        const LwBand B = bands[b];
#define BODY(bb)                                                                                              \
    case bb - 1:                                                                                              \
        lw_vjp_band_body<bb, Real>(in, w, B, ktab, tab, exp_off, tfn_off, plk_off, cst, u, tile, lane, s1, s3,  \
                                   nullptr);                                                                  \
        break;
        switch (b) {
            BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
            BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
        }
#undef BODY
    }
}

// finish: band sums and the precipitable-water chain of the diffusivity
// angle (pwvcm depends on the water vapour of every layer), one thread per
// (layer, column); the thread of layer 0 also writes the surface gradient and
// that of the top layer the top level's
template <typename Real>
__global__ __launch_bounds__(256) void lw_vjp_finish_kernel(ColumnInputs in, LwVjpWork w, LwGrad out)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    const size_t n2 = (size_t)nlay * ncol, n3 = (size_t)(nlay + 1) * ncol;
    const Real *gpart = (const Real *)w.gpart, *glev = (const Real *)w.glev, *gsc = (const Real *)w.gsc;
    double a_pw = 0.0, gts = 0.0;
    for (int b = 0; b < 16; ++b) {
        gts += (double)gsc[(size_t)b * 2 * ncol + col];
        a_pw += (double)gsc[(size_t)b * 2 * ncol + ncol + col];
    }
    if (lay == 0) out.tsfc[col] = gts;
    for (int k = lay; k <= (lay == nlay - 1 ? nlay : lay); ++k) {
        double s = 0.0;
        for (int b = 0; b < 16; ++b) s += (double)glev[(size_t)b * n3 + (size_t)k * ncol + col];
        out.tlev[(size_t)k * ncol + col] = s;
    }
    // pwvcm = (amw wvttl / (amd amttl)) 1e3 plev(0) / (1e2 grav), wvttl = sum
    // coldry wkl1, amttl = sum coldry (1 + wkl1) (rrtmg_lw_rad.nomcica.f90:845-852)
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, avogad = 6.02214199e+23, grav = 9.8066;
    using D1 = Dual<double, 1>;
    const double wvttl = w.colscal[(size_t)2 * ncol + col], amttl = w.colscal[(size_t)3 * ncol + col];
    const double cpw = (amw / amd) * (1.e3 * in.plev[col]) / (1.e2 * grav);
    const size_t ix = (size_t)lay * ncol + col;
    const D1 h = D1::seed(in.h2o[ix], 0);
    const double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
    const D1 wkl1 = (h / (1.0 - h)) * amdw;
    const D1 amm = (1.0 - wkl1) * amd + wkl1 * amw;
    const D1 coldry = ((pz0 - pz1) * 1.e3 * avogad) / (1.e2 * grav * amm * (1.0 + wkl1));
    const D1 wv = coldry * wkl1, am = coldry + coldry * wkl1;
    const double dpw = cpw * (wv.d[0] * amttl - wvttl * am.d[0]) / (amttl * amttl);
    double gt = 0.0, gq = 0.0, go = 0.0;
    for (int b = 0; b < 16; ++b) {
        const Real *gp = gpart + (size_t)b * 3 * n2;
        gt += (double)gp[ix];
        gq += (double)gp[n2 + ix];
        go += (double)gp[2 * n2 + ix];
    }
    out.tlay[ix] = gt;
    out.h2o[ix] = gq + a_pw * dpw;
    out.o3[ix] = go;
}

template <typename Real>
static hipError_t launch_lw_vjp_t(const ColumnInputs &in, const VjpCot &cot, const LwGrad &out, const LwVjpWork &w,
                                  const LwDevTables &t, bool smooth, hipStream_t s, hipEvent_t *ev,
                                  const LwOutputs *fwd, int mode)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(lw_vjp_prep_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, cot, w, t.cst, t.heatfac);
    hipLaunchKernelGGL(lw_vjp_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = lw_vjp_lds_bytes(in.nlay, sizeof(Real));
    const int nblocks = ntiles * t.desc.nwaves;
    if (ev) (void)hipEventRecord(ev[1], s);
    bool done = false;
    if constexpr (sizeof(Real) == 4) {
        if (t.lds && !smooth) {
            const int nitems = (ntiles + LWV_WGW - 1) / LWV_WGW * t.desc.nwaves;
            hipLaunchKernelGGL(reset_items_kernel, dim3(1), dim3(1), 0, s, w.items);
            hipLaunchKernelGGL(lw_vjp_band_lds_kernel<Real>, dim3(lw_vjp_lds_grid(t.plan, in.ncol, t.desc.nwaves)),
                               dim3(64 * LWV_WGW), (size_t)t.plan.lds_bytes, s, in, w, t.bands_lds, t.plan,
                               t.desc.nwaves, nitems, t.desc.exp_tbl, t.desc.tfn_tbl, t.desc.totplnk,
                               (const Real *)t.tab, t.cst, t.units);
            done = true;
        }
    }
    // otherwise launches of at most w.chunk wave units, each reusing the scratch rows
    for (int b0 = 0; !done && b0 < nblocks; b0 += w.chunk) {
        const int nb_c = std::min(w.chunk, nblocks - b0);
        hipLaunchKernelGGL(lw_vjp_band_kernel<Real>, dim3(xcd_grid(nb_c)), dim3(64), shmem, s, in, w, t.bands,
                           t.desc.nwaves, b0, nb_c, smooth ? -1 : t.desc.exp_tbl, smooth ? -1 : t.desc.tfn_tbl,
                           t.desc.totplnk, (const Real *)t.tab, t.cst, t.units);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(lw_vjp_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out);
    if (fwd) {
        const hipError_t e = launch_lw_finish(in, *fwd, w.bflux, t, mode, s);
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_lw_vjp(const ColumnInputs &in, const VjpCot &cot, const LwGrad &out, const LwVjpWork &w,
                         const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev, const LwOutputs *fwd)
{
    if (fwd && !w.bflux) return hipErrorInvalidValue;
    return mode == 1 ? launch_lw_vjp_t<float>(in, cot, out, w, t, false, s, ev, fwd, mode)
                     : launch_lw_vjp_t<double>(in, cot, out, w, t, mode == 2, s, ev, fwd, mode);
}

}  // namespace rrtmg
