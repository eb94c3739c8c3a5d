// SW optics of one (layer, g-point): per-layer coefficient record, table
// reader, the 14 band formulas of taumol_sw (rrtmg_sw_taumol.f90), the solar
// source at the reference layer, reftra_sw and the direct-beam transmittance.
// Shared by the forward band kernel and the backward (vector-Jacobian)
// kernel.  `Real` is the table type; `V`/`F` the arithmetic type: Real in the
// forward, dual numbers in the backward.
#pragma once
#include <hip/hip_runtime.h>

#include "dual.hpp"
#include "rrtmg_device.hpp"
#include "rrtmg_types.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// per-layer coefficients of one column, loaded by every lane of that column
template <typename Real>
struct SwLayer {
    Real fac00, fac01, fac10, fac11;
    Real colh2o, colco2, colo3, colch4, colo2, colmol;
    Real selffac, selffrac, forfac, forfrac;
    int jp, jt, jt1, inds, indf;
};

template <typename Real>
__device__ __forceinline__ void sw_load_layer(SwLayer<Real> &c, const Real *__restrict__ coef,
                                              const int *__restrict__ ipk, int lay, int col, int ncol, int nlay)
{
    const unsigned ix = (unsigned)(lay * ncol + col);
    // the (layer, column) record, [lay][col][SWF_STRIDE], 16-byte aligned
    const Real *rec = (const Real *)__builtin_assume_aligned(coef + (size_t)ix * SWF_STRIDE, 16);
    (void)nlay;
#define LD(f, F) c.f = rec[F]
    LD(fac00, SWF_FAC00); LD(fac01, SWF_FAC01); LD(fac10, SWF_FAC10); LD(fac11, SWF_FAC11);
    LD(colh2o, SWF_COLH2O); LD(colco2, SWF_COLCO2); LD(colo3, SWF_COLO3); LD(colch4, SWF_COLCH4);
    LD(colo2, SWF_COLO2); LD(colmol, SWF_COLMOL);
    LD(selffac, SWF_SELFFAC); LD(selffrac, SWF_SELFFRAC); LD(forfac, SWF_FORFAC); LD(forfrac, SWF_FORFRAC);
#undef LD
    int w = at32(ipk, ix);
    c.jp = unpack_jp(w);
    c.jt = unpack_jt(w);
    c.jt1 = unpack_jt1(w);
    c.inds = unpack_inds(w);
    c.indf = unpack_indf(w);
}

// Per-column part of the SW prep for column col (after the per-layer part,
// whose packed indices ipk hold jp): laytrop and the solar-source reference
// layer of each band, laysolfr[b * ncol + col].  Bands searching the lower
// atmosphere (18-25) take min(lay+1, laytrop) at the last jp crossing of
// layreffr, default laytrop; bands searching aloft (16, 17, 27, 28, 29) take
// the last crossing above laytrop, default nlayers; band 26 uses laytrop
// (e.g. rrtmg_sw_taumol.f90:292-295, 391-394, 494-496).  One pass over the
// layers tests every band's crossing.
__device__ __forceinline__ void sw_column_refs(const double *__restrict__ play, const int *__restrict__ ipk,
                                               int ncol, int nlay, int col, int *__restrict__ laytrop_out,
                                               int *__restrict__ laysolfr)
{
    int laytrop = 0;
    for (int lay = 0; lay < nlay; ++lay)
        if (log(play[(size_t)lay * ncol + col]) > 4.56) ++laytrop;
    *laytrop_out = laytrop;
    constexpr int lowref[14] = {0, 0, 6, 3, 3, 8, 2, 6, 1, 2, 0, 0, 0, 0};
    constexpr int upref[14] = {18, 30, 0, 0, 0, 0, 0, 0, 0, 0, 0, 32, 58, 49};
    int lsf[14];
#pragma unroll
    for (int b = 0; b < 14; ++b) lsf[b] = upref[b] > 0 ? nlay : laytrop;
    const int uplo = max(laytrop + 1, 2);
    int jp2 = 0, jp1 = unpack_jp(ipk[col]);  // jp of layers lay-2 and lay-1
    for (int lay = 1; lay <= nlay; ++lay) {
        const int jpc = lay < nlay ? unpack_jp(ipk[(size_t)lay * ncol + col]) : jp1;  // jp(lay), jp(nlay-1) on top
#pragma unroll
        for (int b = 0; b < 14; ++b) {
            if (upref[b] > 0) {
                if (lay >= uplo && jp2 < upref[b] && jp1 >= upref[b]) lsf[b] = lay;
            } else if (lowref[b] > 0) {
                if (lay <= laytrop && jp1 < lowref[b] && jpc >= lowref[b]) lsf[b] = min(lay + 1, laytrop);
            }
        }
        jp2 = jp1;
        jp1 = jpc;
    }
#pragma unroll
    for (int b = 0; b < 14; ++b) laysolfr[(size_t)b * ncol + col] = max(lsf[b], 1);
}

// X(jb) for every SW band, for switches over compile-time band bodies
#define SW_BANDS(X) X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29)

// Table reader for one lane: row r (1-based, as the reference's ind) of a
// [rows][ng] table at `off`.
template <typename Real>
struct Tab {
    const Real *__restrict__ t;
    int ng, g;
    // a 32-bit unsigned byte offset from the uniform base: one load with an
    // SGPR base and a VGPR offset, no 64-bit address arithmetic per gather
    __device__ __forceinline__ Real operator()(int off, int row) const
    {
        const unsigned byte = (unsigned)(off + (row - 1) * ng + g) * (unsigned)sizeof(Real);
        return *(const Real *)((const char *)t + byte);
    }
};

// binary-species parameter of a band (e.g. rrtmg_sw_taumol.f90:276-283)
template <typename R>
__device__ __forceinline__ void sw_eta(R cola, R colb, R strrat, R mult, R &speccomb, int &js, R &fs)
{
    speccomb = cola + strrat * colb;
    R specparm = fdiv(cola, speccomb);
    const R oneminus = (R)(1.0 - 1.e-6);
    if (specparm >= oneminus) specparm = oneminus;
    R specmult = mult * specparm;
    js = 1 + (int)specmult;
    fs = specmult - (R)(int)specmult;
}

// 8-point (eta, T, p) interpolation; s1 = jt stride (9 lower, 5 upper)
template <typename Real, typename V>
__device__ __forceinline__ V sw_maj8(const Tab<Real> &T, int off, int ind0, int ind1, V fs, const SwLayer<V> &c, int s1)
{
    V fac000 = ((Real)1 - fs) * c.fac00, fac010 = ((Real)1 - fs) * c.fac10;
    V fac100 = fs * c.fac00, fac110 = fs * c.fac10;
    V fac001 = ((Real)1 - fs) * c.fac01, fac011 = ((Real)1 - fs) * c.fac11;
    V fac101 = fs * c.fac01, fac111 = fs * c.fac11;
    return fac000 * T(off, ind0) + fac100 * T(off, ind0 + 1) + fac010 * T(off, ind0 + s1) +
           fac110 * T(off, ind0 + s1 + 1) + fac001 * T(off, ind1) + fac101 * T(off, ind1 + 1) +
           fac011 * T(off, ind1 + s1) + fac111 * T(off, ind1 + s1 + 1);
}

template <typename Real, typename V>
__device__ __forceinline__ V sw_maj4(const Tab<Real> &T, int off, int ind0, int ind1, const SwLayer<V> &c)
{
    return c.fac00 * T(off, ind0) + c.fac10 * T(off, ind0 + 1) + c.fac01 * T(off, ind1) + c.fac11 * T(off, ind1 + 1);
}

template <typename Real, typename V>
__device__ __forceinline__ V sw_cont(const Tab<Real> &T, const SwBand &B, const SwLayer<V> &c, bool self)
{
    V f = c.forfac * (T(B.forref, c.indf) + c.forfrac * (T(B.forref, c.indf + 1) - T(B.forref, c.indf)));
    if (!self) return f;
    V s = c.selffac * (T(B.selfref, c.inds) + c.selffrac * (T(B.selfref, c.inds + 1) - T(B.selfref, c.inds)));
    return s + f;
}

// Gas and Rayleigh optical depth of one (layer, g) for band jb (16..29),
// rrtmg_sw_taumol.f90:243-1536.
template <typename Real, typename V = Real>
__device__ __forceinline__ void sw_gas(int jb, const SwBand &B, const Tab<Real> &T, const SwLayer<V> &c, bool lower,
                                       V &taug, V &taur)
{
    V speccomb, fs;
    int js;
    int ind0, ind1;
    if (lower) {
        ind0 = ((c.jp - 1) * 5 + (c.jt - 1)) * B.nspa;
        ind1 = (c.jp * 5 + (c.jt1 - 1)) * B.nspa;
    } else {
        ind0 = ((c.jp - 13) * 5 + (c.jt - 1)) * B.nspb;
        ind1 = ((c.jp - 12) * 5 + (c.jt1 - 1)) * B.nspb;
    }
    Real rayl = (Real)B.rayl;
    switch (jb) {
    case 16:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colch4, (Real)252.131, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            taug = c.colch4 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        }
        taur = c.colmol * rayl;
        break;
    case 17:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colco2, (Real)0.364641, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            sw_eta<V>(c.colh2o, c.colco2, (Real)0.364641, (Real)4, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absb, ind0 + js, ind1 + js, fs, c, 5) + c.colh2o * sw_cont(T, B, c, false);
        }
        taur = c.colmol * rayl;
        break;
    case 18:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colch4, (Real)38.9589, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            taug = c.colch4 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        }
        taur = c.colmol * rayl;
        break;
    case 19:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colco2, (Real)5.49281, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            taug = c.colco2 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        }
        taur = c.colmol * rayl;
        break;
    case 20:
        if (lower)
            taug = c.colh2o * (sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + sw_cont(T, B, c, true)) +
                   c.colch4 * T(B.absch4, 1);
        else
            taug = c.colh2o * (sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c) + sw_cont(T, B, c, false)) +
                   c.colch4 * T(B.absch4, 1);
        taur = c.colmol * rayl;
        break;
    case 21:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colco2, (Real)0.0045321, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            sw_eta<V>(c.colh2o, c.colco2, (Real)0.0045321, (Real)4, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absb, ind0 + js, ind1 + js, fs, c, 5) + c.colh2o * sw_cont(T, B, c, false);
        }
        taur = c.colmol * rayl;
        break;
    case 22: {
        V o2cont = fdiv((Real)4.35e-4 * c.colo2, (Real)(350.0 * 2.0));
        if (lower) {
            sw_eta<V>(c.colh2o, c.colo2, (Real)(1.6 * 0.022708), (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true) +
                   o2cont;
        } else {
            taug = c.colo2 * (Real)1.6 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c) + o2cont;
        }
        taur = c.colmol * rayl;
        break;
    }
    case 23:
        taug = lower ? V(c.colh2o * ((Real)1.029 * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + sw_cont(T, B, c, true)))
                     : V((Real)0);
        taur = c.colmol * T(B.raylg, 1);
        break;
    case 24:
        if (lower) {
            sw_eta<V>(c.colh2o, c.colo2, (Real)0.124692, (Real)8, speccomb, js, fs);
            taur = c.colmol * (T(B.rayla, js) + fs * (T(B.rayla, js + 1) - T(B.rayla, js)));
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colo3 * T(B.abso3a, 1) +
                   c.colh2o * sw_cont(T, B, c, true);
        } else {
            taur = c.colmol * T(B.raylb, 1);
            taug = c.colo2 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c) + c.colo3 * T(B.abso3b, 1);
        }
        break;
    case 25:
        taur = c.colmol * T(B.raylg, 1);
        taug = lower ? V(c.colh2o * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + c.colo3 * T(B.abso3a, 1))
                     : V(c.colo3 * T(B.abso3b, 1));
        break;
    case 26:
        taug = (Real)0;
        taur = c.colmol * T(B.raylg, 1);
        break;
    case 27:
        taur = c.colmol * T(B.raylg, 1);
        taug = lower ? V(c.colo3 * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c))
                     : V(c.colo3 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c));
        break;
    case 28:
        if (lower) {
            sw_eta<V>(c.colo3, c.colo2, (Real)6.67029e-07, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9);
        } else {
            sw_eta<V>(c.colo3, c.colo2, (Real)6.67029e-07, (Real)4, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absb, ind0 + js, ind1 + js, fs, c, 5);
        }
        taur = c.colmol * rayl;
        break;
    default:  // 29
        if (lower)
            taug = c.colh2o * (sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + sw_cont(T, B, c, true)) +
                   c.colco2 * T(B.absco2, 1);
        else
            taug = c.colco2 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c) + c.colh2o * T(B.absh2o, 1);
        taur = c.colmol * rayl;
        break;
    }
}

// Solar source of g-point g at the band's reference layer (sfluxzen),
// evaluated in FP64 from that layer's column amounts.
template <typename Real>
__device__ double sw_sfluxzen(int jb, const SwBand &B, const Tab<Real> &T, const SwLayer<Real> &c, bool lower)
{
    double sp, fs;
    int js;
    auto flux_eta = [&](double cola, double colb, double strrat, double mult) {
        sw_eta<double>(cola, colb, strrat, mult, sp, js, fs);
        double a = (double)T(B.sfluxref, js), b = (double)T(B.sfluxref, js + 1);
        return a + fs * (b - a);
    };
    switch (jb) {
    case 17: return flux_eta(c.colh2o, c.colco2, 0.364641, 4.0);
    case 18: return flux_eta(c.colh2o, c.colch4, 38.9589, 8.0);
    case 19: return flux_eta(c.colh2o, c.colco2, 5.49281, 8.0);
    case 21: return flux_eta(c.colh2o, c.colco2, 0.0045321, 8.0);
    case 22: return flux_eta(c.colh2o, c.colo2, 1.6 * 0.022708, 8.0);
    case 24: return flux_eta(c.colh2o, c.colo2, 0.124692, 8.0);
    case 27: return (50.15 / 48.37) * (double)T(B.sfluxref, 1);
    case 28: return flux_eta(c.colo3, c.colo2, 6.67029e-07, 4.0);
    default: return (double)T(B.sfluxref, 1);
    }
    (void)lower;
}

// The same solar source in arithmetic type V (the backward's dual numbers,
// for its derivative with respect to the reference layer's state).
template <typename Real, typename V>
__device__ V sw_sfluxzen_v(int jb, const SwBand &B, const Tab<Real> &T, const SwLayer<V> &c)
{
    V sp, fs;
    int js;
    auto flux_eta = [&](V cola, V colb, Real strrat, Real mult) {
        sw_eta<V>(cola, colb, V(strrat), V(mult), sp, js, fs);
        Real a = T(B.sfluxref, js), b = T(B.sfluxref, js + 1);
        return a + fs * (b - a);
    };
    switch (jb) {
    case 17: return flux_eta(c.colh2o, c.colco2, (Real)0.364641, (Real)4);
    case 18: return flux_eta(c.colh2o, c.colch4, (Real)38.9589, (Real)8);
    case 19: return flux_eta(c.colh2o, c.colco2, (Real)5.49281, (Real)8);
    case 21: return flux_eta(c.colh2o, c.colco2, (Real)0.0045321, (Real)8);
    case 22: return flux_eta(c.colh2o, c.colo2, (Real)(1.6 * 0.022708), (Real)8);
    case 24: return flux_eta(c.colh2o, c.colo2, (Real)0.124692, (Real)8);
    case 27: return V((Real)(50.15 / 48.37) * T(B.sfluxref, 1));
    case 28: return flux_eta(c.colo3, c.colo2, (Real)6.67029e-07, (Real)4);
    default: return V(T(B.sfluxref, 1));
    }
}

// exp(-x) from the lookup table at int(1e4 x/(bpade + x) + 0.5), or, when tbl
// is NULL, the function the table holds: exp(-x) floored at 1e-20 as the
// table's entries are (rrtmg_sw_init.f90:94-104; the floor also keeps 1/E
// finite in FP32).  The smooth FP64 mode and the backward's derivatives use
// the latter, the smooth mode also without the small-argument series
// 1 - x + x^2/2 that the table path uses below 0.06.  For a dual argument the
// derivative is -E, zero on the floor.
template <typename F, typename Real>
__device__ __forceinline__ F tbl_exp(F x, const Real *__restrict__ tbl)
{
    return tbl ? (F)at32(tbl, (unsigned)tbl_index(x)) : fmax(exp(-x), (F)1.e-20);
}
template <typename F, int N, typename Real>
__device__ __forceinline__ Dual<F, N> tbl_exp(const Dual<F, N> &x, const Real *__restrict__ tbl)
{
    Dual<F, N> r;
    const F e = tbl ? (F)at32(tbl, (unsigned)tbl_index(x.v)) : exp(-x.v);
    r.v = tbl ? e : fmax(e, (F)1.e-20);
    const F s = e > (F)1.e-20 ? -e : (F)0;
    for (int i = 0; i < N; ++i) r.d[i] = s * x.d[i];
    return r;
}

// The FP32 two-stream eigenvalue for asymmetry 0: with g = 0, gamma1 - gamma2
// = 2 (1 - w) exactly, so k = sqrt(gamma1^2 - gamma2^2) = sqrt(2 (1 - w)
// (gamma1 + gamma2)), formed from zomw = 1 - w without the cancellation of
// the difference of squares as w -> 1 (rrtmg_sw_reftra.f90:211 has the
// general form).  One function for reftra and the island test below, so the
// two evaluate it identically and share it.
template <typename F>
__device__ __forceinline__ F eigen_k32(F zw, F zomw)
{
    const F zgamma1 = ((F)8 - zw * (F)5) * (F)0.25, zgamma2 = (F)3 * zw * (F)0.25;
    return sqrt((F)2 * zomw * (zgamma1 + zgamma2));
}

// PIFM two-stream layer reflectance/transmittance, kmodts = 2, asymmetry 0
// (rrtmg_sw_reftra.f90:127-303).  F is the arithmetic type.  zomw = 1 - zw
// formed without cancellation (gas over total optical depth); the FP32
// arithmetic uses it for the eigenvalue, see below.  pdbt, when given,
// receives the direct-beam transmittance exp(-min(zto1/prmuz, 500)) that
// reftra looks up on the way.
template <typename F, typename Real, typename S>
__device__ __forceinline__ void reftra(F zto1, F zw, F zomw, S prmuz, const Real *__restrict__ exp_tbl, F &pref,
                                       F &prefd, F &ptra, F &ptrad, F *pdbt = nullptr)
{
    const F eps = (F)1.e-08, od_lo = (F)0.06, zwcrit = (F)0.9999995;
    const F zg = (F)0, zg3 = (F)3 * zg;
    F zgamma1 = ((F)8 - zw * ((F)5 + zg3)) * (F)0.25;
    F zgamma2 = (F)3 * (zw * ((F)1 - zg)) * (F)0.25;
    F zgamma3 = ((F)2 - zg3 * prmuz) * (F)0.25;
    F zgamma4 = (F)1 - zgamma3;
    F zwo = zw;  // zw / (1 - (1 - zw) * (g/(1-g))^2) with g = 0
    if (zwo >= zwcrit) {
        F za = zgamma1 * prmuz;
        F za1 = za - zgamma3;
        F zgt = zgamma1 * zto1;
        F ze1 = fmin(fdiv(zto1, (F)prmuz), (F)500);
        F ze2 = (exp_tbl && ze1 <= od_lo) ? F((S)1 - ze1 + (S)0.5 * ze1 * ze1) : tbl_exp(ze1, exp_tbl);
        if (pdbt) *pdbt = ze2;
        if (ze2 == (F)1) {
            pref = (F)0;
            ptra = (F)1;
            prefd = (F)0;
            ptrad = (F)1;
        } else {
            pref = fdiv(zgt - za1 * ((F)1 - ze2), (F)1 + zgt);
            ptra = (F)1 - pref;
            prefd = fdiv(zgt, (F)1 + zgt);
            ptrad = (F)1 - prefd;
        }
    } else {
        F za1 = zgamma1 * zgamma4 + zgamma2 * zgamma3;
        F za2 = zgamma1 * zgamma3 + zgamma2 * zgamma4;
        // k = sqrt(g1^2 - g2^2); with g = 0, g1 - g2 = 2 (1 - w) exactly, so
        // in FP32 k = sqrt(2 (1 - w) (g1 + g2)), free of the cancellation of
        // the difference of squares as w -> 1; FP64 keeps the reference's form
        F zrk;
        if (sizeof(val(zto1)) == 4)
            zrk = eigen_k32(zw, zomw);
        else
            zrk = sqrt(zgamma1 * zgamma1 - zgamma2 * zgamma2);
        F zrp = zrk * prmuz;
        F zrp1 = (F)1 + zrp, zrm1 = (F)1 - zrp;
        F zrk2 = (F)2 * zrk;
        F zrpp = (F)1 - zrp * zrp;
        F zrkg = zrk + zgamma1;
        F zr1 = zrm1 * (za2 + zrk * zgamma3);
        F zr2 = zrp1 * (za2 - zrk * zgamma3);
        F zr3 = zrk2 * (zgamma3 - za2 * prmuz);
        F zr4 = zrpp * zrkg;
        F zr5 = zrpp * (zrk - zgamma1);
        F zt1 = zrp1 * (za1 + zrk * zgamma4);
        F zt2 = zrm1 * (za1 - zrk * zgamma4);
        F zt3 = zrk2 * (zgamma4 + za1 * prmuz);
        F ze1 = fmin(zrk * zto1, (F)500);
        F ze2 = fmin(fdiv(zto1, (F)prmuz), (F)500);
        F zem1 = (exp_tbl && ze1 <= od_lo) ? F((S)1 - ze1 + (S)0.5 * ze1 * ze1) : tbl_exp(ze1, exp_tbl);
        F zem2 = (exp_tbl && ze2 <= od_lo) ? F((S)1 - ze2 + (S)0.5 * ze2 * ze2) : tbl_exp(ze2, exp_tbl);
        if (pdbt) *pdbt = zem2;
        F zemm = zem1 * zem1;
        if (sizeof(val(zto1)) == 4) {
            // FP32: numerators and denominator multiplied by zem1, which
            // leaves no 1/zem1 and, with zem2 zep2 = 1, no 1/zem2:
            // zdenr zem1 = zr4 + zr5 zem1^2, and the zdenr test scales with it
            const F den = zr4 + zr5 * zemm;
            const F lim = eps * zem1;
            if (den >= -lim && den <= lim) {
                pref = eps;
                ptra = zem2;
            } else {
                const F rden = fdiv((F)1, den);
                pref = zw * (zr1 - zr2 * zemm - zr3 * (zem2 * zem1)) * rden;
                ptra = zem2 - zw * (zem2 * (zt1 - zt2 * zemm) - zt3 * zem1) * rden;
            }
        } else {
            F zep1 = fdiv((F)1, zem1);
            F zep2 = fdiv((F)1, zem2);
            F zdenr = zr4 * zep1 + zr5 * zem1;
            F zdent = zr4 * zep1 + zr5 * zem1;  // zt4 = zr4, zt5 = zr5
            if (zdenr >= -eps && zdenr <= eps) {
                pref = eps;
                ptra = zem2;
            } else {
                pref = fdiv(zw * (zr1 * zep1 - zr2 * zem1 - zr3 * zem2), zdenr);
                ptra = zem2 - fdiv(zem2 * zw * (zt1 * zep1 - zt2 * zem1 - zt3 * zep2), zdent);
            }
        }
        // (1 - zbeta zemm) zrkg with zbeta = (zgamma1 - zrk) / zrkg; FP32 forms
        // it without the quotient
        F zdend;
        if (sizeof(val(zto1)) == 4) {
            zdend = fdiv((F)1, zrkg - (zgamma1 - zrk) * zemm);
        } else {
            F zbeta = fdiv(zgamma1 - zrk, zrkg);
            zdend = fdiv((F)1, ((F)1 - zbeta * zemm) * zrkg);
        }
        prefd = zgamma2 * ((F)1 - zemm) * zdend;
        ptrad = zrk2 * zem1 * zdend;
    }
}

// reftra in FP64 for an FP32 layer (the islands below): r, rd, t, td as
// plain FP32 values or as FP32 dual numbers.  Out of line on purpose: the
// islands are rare, and inlined, their FP64 temporaries would set the
// register count of the whole FP32 band kernel (forward: 107 VGPRs and 4
// waves per SIMD, against 80 and 6 with the call).
template <typename V>
struct LayerOptics {
    V r, rd, t, td;
};
template <typename V, typename V64>
__device__ __forceinline__ LayerOptics<V> reftra_fp64(V taug, V taur, float prmu0, const float *__restrict__ exp_tbl)
{
    V64 R, RD, T, TD;
    const V64 t64 = cast<V64>(taur) + cast<V64>(taug);
    reftra<V64, float, double>(t64, cast<V64>(taur) / t64, cast<V64>(taug) / t64, (double)prmu0, exp_tbl, R, RD, T,
                               TD);
    return {cast<V>(R), cast<V>(RD), cast<V>(T), cast<V>(TD)};
}
[[maybe_unused]] static __device__ __attribute__((noinline)) LayerOptics<float>
reftra_island(float taug, float taur, float prmu0, const float *__restrict__ exp_tbl)
{
    return reftra_fp64<float, double>(taug, taur, prmu0, exp_tbl);
}
[[maybe_unused]] static __device__ __attribute__((noinline)) LayerOptics<Dual<float, 2>>
reftra_island(Dual<float, 2> taug, Dual<float, 2> taur, float prmu0, const float *__restrict__ exp_tbl)
{
    return reftra_fp64<Dual<float, 2>, Dual<double, 2>>(taug, taur, prmu0, exp_tbl);
}

// optics of one (layer, g): total optical depth, reftra and direct-beam
// transmittance (rrtmg_sw_spcvrt.f90:375-550, clear sky, no aerosol).  V is
// the arithmetic type in the band kernel's precision (plain, or dual numbers
// in the backward).
template <typename Real, typename V = Real>
__device__ __forceinline__ void sw_layer_optics(V taug, V taur, Real prmu0, const Real *__restrict__ exp_tbl, V &r, V &rd,
                                                V &t, V &td, V &dbt)
{
    V tau = taur + taug;
    V w = fdiv(taur, tau);
    V omw = fdiv(taug, tau);
    bool island = false;
    if (sizeof(Real) == 4) {
        // FP64 islands of the mixed path: the removable 0/0 of the direct-beam
        // solution at k mu0 = 1, where both the numerator and zdenr of reftra
        // vanish.  Both are O(d), d = |1 - k mu0|, and carry absolute errors
        // of the FP32 epsilon, so FP32 values are good to ~6e-8/d relative:
        // the island covers d < 0.002 (relative error above 3e-5 outside it).
        // The derivatives (dual V, the backward) divide once more by d and
        // keep d < 0.02.  Near-conservative scattering needs no island for the
        // values (the eigenvalue is formed from 1 - w, see reftra) but does
        // for the derivatives: with w > 0.999 the derivative with respect to
        // the gas optical depth is a (1 - w)-sized difference of O(1) terms,
        // which FP32 resolves to a few percent only
        Real wv = val(w), ov = val(omw);
        Real zrp = eigen_k32(wv, ov) * prmu0;
        const bool dual = sizeof(V) != sizeof(Real);
        island = fabs((Real)1 - zrp) < (dual ? (Real)0.02 : (Real)0.002) || (dual && wv > (Real)0.999);
    }
    // direct-beam transmittance (rrtmg_sw_spcvrt.f90:536-546): the lookup of
    // exp(-tau/mu0) that reftra makes for zem2 (rrtmg_sw_reftra.f90:237-270)
    // with the same operands, so it is taken from there.  Where reftra clamps
    // tau/mu0 at 500, both table entries are the 1e-20 floor
    // (rrtmg_sw_init.f90:96-104; past 46 in the smooth function).
    if constexpr (sizeof(Real) == 4) {
        if (island) {
            const LayerOptics<V> o = reftra_island(taug, taur, prmu0, exp_tbl);
            r = o.r;
            rd = o.rd;
            t = o.t;
            td = o.td;
            V ze1 = fdiv(tau, V(prmu0));
            dbt = (exp_tbl && ze1 <= (Real)0.06) ? V((Real)1 - ze1 + (Real)0.5 * ze1 * ze1) : tbl_exp(ze1, exp_tbl);
        } else {
            reftra<V, Real, Real>(tau, w, omw, prmu0, exp_tbl, r, rd, t, td, &dbt);
        }
    } else {
        reftra<V, Real, Real>(tau, w, omw, prmu0, exp_tbl, r, rd, t, td, &dbt);
    }
}

}  // namespace rrtmg
