// LW gas optics of one (layer, g-point): per-layer coefficient record, table
// reader and the 16 band-specialised taumol formulas (rrtmg_lw_taumol.f90).
// Shared by the forward band kernel and the backward (vector-Jacobian)
// kernel.  `Real` is the table type; `V` is the arithmetic type of the layer
// coefficients and results: Real in the forward, a dual number carrying
// derivatives with respect to layer temperature, water vapour and ozone in
// the backward.
#pragma once
#include <hip/hip_runtime.h>

#include "dual.hpp"
#include "rrtmg_device.hpp"
#include "rrtmg_types.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
template <typename Real>
struct LwLayer {
    Real fac00, fac01, fac10, fac11;
    Real colh2o, colco2, colo3, coln2o, colco, colch4, colo2, colbrd;
    Real selffac, selffrac, forfac, forfrac, minorfrac, scaleminor, scaleminorn2, coldry, pavel;
    Real wx1, wx2, wx3, wx4, tlayfrac;
    int jp, jt, jt1, inds, indf, indm, indlay;
};

template <typename Real>
__device__ __forceinline__ void lw_load_layer(LwLayer<Real> &c, const Real *__restrict__ coef,
                                              const int *__restrict__ ipk, int lay, int col, int ncol, int nlay)
{
    const unsigned ix = (unsigned)(lay * ncol + col);
    // the (layer, column) record, [lay][col][LWF_STRIDE], 16-byte aligned
    const Real *rec = (const Real *)__builtin_assume_aligned(coef + (size_t)ix * LWF_STRIDE, 16);
    (void)nlay;
#define LD(f, F) c.f = rec[F]
    LD(fac00, LWF_FAC00); LD(fac01, LWF_FAC01); LD(fac10, LWF_FAC10); LD(fac11, LWF_FAC11);
    LD(colh2o, LWF_COLH2O); LD(colco2, LWF_COLCO2); LD(colo3, LWF_COLO3); LD(coln2o, LWF_COLN2O);
    LD(colco, LWF_COLCO); LD(colch4, LWF_COLCH4); LD(colo2, LWF_COLO2); LD(colbrd, LWF_COLBRD);
    LD(selffac, LWF_SELFFAC); LD(selffrac, LWF_SELFFRAC); LD(forfac, LWF_FORFAC); LD(forfrac, LWF_FORFRAC);
    LD(minorfrac, LWF_MINORFRAC); LD(scaleminor, LWF_SCALEMINOR); LD(scaleminorn2, LWF_SCALEMINORN2);
    LD(coldry, LWF_COLDRY); LD(pavel, LWF_PAVEL); LD(wx1, LWF_WX1); LD(wx2, LWF_WX2); LD(wx3, LWF_WX3);
    LD(wx4, LWF_WX4); LD(tlayfrac, LWF_TLAYFRAC);
#undef LD
    int w = at32(ipk, ix);
    c.jp = unpack_jp(w);
    c.jt = unpack_jt(w);
    c.jt1 = unpack_jt1(w);
    c.inds = unpack_inds(w);
    c.indf = unpack_indf(w);
    c.indm = unpack_indm(w);
    c.indlay = unpack_indlay(w);
}

template <typename Real>
struct LTab {
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

template <typename R>
struct Eta {
    R speccomb, specparm, fs;
    int js;
};

// binary-species parameter (e.g. rrtmg_lw_taumol.f90:496-509)
template <typename V, typename S>
__device__ __forceinline__ Eta<V> lw_eta(V cola, V colb, S rat, S mult)
{
    Eta<V> e;
    e.speccomb = cola + rat * colb;
    e.specparm = fdiv(cola, e.speccomb);
    const S oneminus = (S)(1.0 - 1.e-6);
    if (e.specparm >= oneminus) e.specparm = V(oneminus);
    V specmult = mult * e.specparm;
    e.js = 1 + (int)specmult;
    e.fs = specmult - (S)(int)specmult;
    return e;
}

// lower-atmosphere major term with the 3-point eta scheme at the edges
// (rrtmg_lw_taumol.f90:550-668): facA/facB = fac00/fac10 or fac01/fac11
template <typename Real, typename V>
__device__ __forceinline__ V lw_major3(const LTab<Real> &T, int off, const Eta<V> &e, int ind, V facA, V facB)
{
    if (e.specparm < (Real)0.125) {
        V p = e.fs - (Real)1, p2 = p * p, p4 = p2 * p2;
        V fk0 = p4, fk1 = (Real)1 - p - (Real)2 * p4, fk2 = p + p4;
        return e.speccomb * (fk0 * facA * T(off, ind) + fk1 * facA * T(off, ind + 1) + fk2 * facA * T(off, ind + 2) +
                             fk0 * facB * T(off, ind + 9) + fk1 * facB * T(off, ind + 10) +
                             fk2 * facB * T(off, ind + 11));
    } else if (e.specparm > (Real)0.875) {
        V p = -e.fs, p2 = p * p, p4 = p2 * p2;
        V fk0 = p4, fk1 = (Real)1 - p - (Real)2 * p4, fk2 = p + p4;
        return e.speccomb * (fk2 * facA * T(off, ind - 1) + fk1 * facA * T(off, ind) + fk0 * facA * T(off, ind + 1) +
                             fk2 * facB * T(off, ind + 8) + fk1 * facB * T(off, ind + 9) +
                             fk0 * facB * T(off, ind + 10));
    }
    return e.speccomb * (((Real)1 - e.fs) * facA * T(off, ind) + e.fs * facA * T(off, ind + 1) +
                         ((Real)1 - e.fs) * facB * T(off, ind + 9) + e.fs * facB * T(off, ind + 10));
}

// upper-atmosphere binary-species term (e.g. rrtmg_lw_taumol.f90:740-749)
template <typename Real, typename V>
__device__ __forceinline__ V lw_major_b(const LTab<Real> &T, int off, const Eta<V> &e, int ind, V facA, V facB)
{
    return e.speccomb * (((Real)1 - e.fs) * facA * T(off, ind) + e.fs * facA * T(off, ind + 1) +
                         ((Real)1 - e.fs) * facB * T(off, ind + 5) + e.fs * facB * T(off, ind + 6));
}

template <typename Real, typename V>
__device__ __forceinline__ V lw_four(const LTab<Real> &T, int off, int ind0, int ind1, const LwLayer<V> &c)
{
    return c.fac00 * T(off, ind0) + c.fac10 * T(off, ind0 + 1) + c.fac01 * T(off, ind1) + c.fac11 * T(off, ind1 + 1);
}

template <typename Real, typename V>
__device__ __forceinline__ V lw_self(const LTab<Real> &T, const LwBand &B, const LwLayer<V> &c)
{
    return c.selffac * (T(B.selfref, c.inds) + c.selffrac * (T(B.selfref, c.inds + 1) - T(B.selfref, c.inds)));
}
template <typename Real, typename V>
__device__ __forceinline__ V lw_for(const LTab<Real> &T, const LwBand &B, const LwLayer<V> &c)
{
    return c.forfac * (T(B.forref, c.indf) + c.forfrac * (T(B.forref, c.indf + 1) - T(B.forref, c.indf)));
}
// minor absorber interpolated in temperature, k(19, ng)
template <typename Real, typename V>
__device__ __forceinline__ V lw_minor1(const LTab<Real> &T, int off, const LwLayer<V> &c)
{
    return T(off, c.indm) + c.minorfrac * (T(off, c.indm + 1) - T(off, c.indm));
}
// minor absorber interpolated in eta and temperature, k(neta, 19, ng)
template <typename Real, typename V>
__device__ __forceinline__ V lw_minor2(const LTab<Real> &T, int off, int neta, const Eta<V> &m, const LwLayer<V> &c)
{
    int r1 = m.js + neta * (c.indm - 1), r2 = r1 + neta;
    V m1 = T(off, r1) + m.fs * (T(off, r1 + 1) - T(off, r1));
    V m2 = T(off, r2) + m.fs * (T(off, r2 + 1) - T(off, r2));
    return m1 + c.minorfrac * (m2 - m1);
}
template <typename Real, typename V>
__device__ __forceinline__ V lw_pfrac(const LTab<Real> &T, int off, const Eta<V> &e)
{
    return T(off, e.js) + e.fs * (T(off, e.js + 1) - T(off, e.js));
}
// column-amount adjustment of minor CO2/N2O (e.g. rrtmg_lw_taumol.f90:1314-1320)
template <typename V, typename S>
__device__ __forceinline__ V lw_adjcol(V col, V coldry, S chiref, S thresh, S base, S ex)
{
    V rat = fdiv((S)1.e20 * fdiv(col, coldry), chiref);
    if (rat > thresh) return (base + pow(rat - base, ex)) * chiref * coldry * (S)1.e-20;
    return col;
}

// Transmittance 1 - exp(-tau) and the tau transition function
// tfn(tau) = 1 - 2 (1/tau - e/(1 - e)), e = exp(-tau) (what the 10000-point
// tables of rrtmg_lw_init.f90:106-123 tabulate), evaluated directly for the
// smooth FP64 mode, with their derivatives.  Below tau = 0.1 tfn and its
// derivative use the series tau/6 - tau^3/360 + tau^5/15120 (from
// 1/(e^tau - 1) = 1/tau - 1/2 + tau/12 - tau^3/720 + tau^5/30240 - ...),
// whose omitted term is below 1e-11 there.
template <typename Real>
struct LwRT {
    Real atrans, tf, datr, dtf;
};
template <typename Real>
__device__ __forceinline__ LwRT<Real> lw_rt_exact(Real od, bool derivs)
{
    LwRT<Real> r;
    const Real e = exp(-od);
    r.atrans = -expm1(-od);
    r.datr = e;
    if (od < (Real)0.1) {
        const Real o2 = od * od;
        r.tf = od * ((Real)1 / (Real)6 - o2 * ((Real)1 / (Real)360 - o2 / (Real)15120));
        r.dtf = (Real)1 / (Real)6 - o2 * ((Real)1 / (Real)120 - o2 / (Real)3024);
    } else {
        const Real ome = -expm1(-od);
        const Real iod = fdiv((Real)1, od), iome = fdiv((Real)1, ome);
        r.tf = (Real)1 - (Real)2 * (iod - e * iome);
        r.dtf = derivs ? (Real)2 * iod * iod - (Real)2 * e * iome * iome : (Real)0;
    }
    return r;
}

// a LwConst table in the band kernel's precision: the FP64 values or their
// FP32 copies (so the FP32 kernels gather floats and convert nothing)
template <typename V>
__device__ __forceinline__ auto lw_cst(const double *d, const float *f)
{
    if constexpr (sizeof(V) == 4)
        return f;
    else
        return d;
}
template <int NUM, int DEN, int LEV>
constexpr int lw_chi_ratio_k()
{
    constexpr int k = lw_chi_ratio_index(NUM, DEN, LEV);
    static_assert(k >= 0, "chi_mls ratio missing from LW_CHI_RATIOS");
    return k;
}
// chi_mls(i, j); the rat_* table NAME at reference level e; the fixed-level
// quotient chi_mls(a, j) / chi_mls(b, j), computed once at init
#define CHI(i, j) ((Real)at32(lw_cst<V>(cst->chi_mls, cst->chi_mls_f), (unsigned)(((j)-1) * 7 + ((i)-1))))
#define RAT(NAME, e) ((Real)at32(lw_cst<V>(cst->NAME, cst->NAME##_f), (unsigned)(e)))
#define CHIQ(a, b, j) ((Real)lw_cst<V>(cst->chi_ratio, cst->chi_ratio_f)[lw_chi_ratio_k<a, b, j>()])

// Optical depth and Planck fraction of one (layer, g) of band BAND,
// rrtmg_lw_taumol.f90:280-3147.
// CONST_NSP: the species-parameter counts as compile-time constants (the
// forward band kernel; the backward's kernels keep the loaded values, as
// constants they ran 2% slower).
template <int BAND, typename Real, typename V = Real, bool CONST_NSP = false>
__device__ __forceinline__ void lw_gas(const LwBand &B, const LTab<Real> &T, const LwConst *__restrict__ cst,
                                       const LwLayer<V> &c, bool lower, V &taug, V &fracs)
{
    const int g1 = T.g + 1;
    int ind0, ind1;
    if (lower) {
        ind0 = ((c.jp - 1) * 5 + (c.jt - 1)) * (CONST_NSP ? LW_NSPA[BAND - 1] : B.nspa);
        ind1 = (c.jp * 5 + (c.jt1 - 1)) * (CONST_NSP ? LW_NSPA[BAND - 1] : B.nspa);
    } else {
        ind0 = ((c.jp - 13) * 5 + (c.jt - 1)) * (CONST_NSP ? LW_NSPB[BAND - 1] : B.nspb);
        ind1 = ((c.jp - 12) * 5 + (c.jt1 - 1)) * (CONST_NSP ? LW_NSPB[BAND - 1] : B.nspb);
    }
    const int jp = c.jp;
    if constexpr (BAND == 1) {
        if (lower) {
            V corradj = c.pavel < (Real)250 ? V((Real)1 - fdiv((Real)0.15 * ((Real)250 - c.pavel), (Real)154.4)) : V((Real)1);
            V taun2 = c.colbrd * c.scaleminorn2 * lw_minor1(T, B.ka_mn2, c);
            taug = corradj * (c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) +
                              lw_for(T, B, c) + taun2);
            fracs = T(B.fracrefa, 1);
        } else {
            V corradj = (Real)1 - (Real)0.15 * fdiv(c.pavel, (Real)95.6);
            V taun2 = c.colbrd * c.scaleminorn2 * lw_minor1(T, B.kb_mn2, c);
            taug = corradj * (c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c) + taun2);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 2) {
        if (lower) {
            V corradj = (Real)1 - fdiv((Real).05 * (c.pavel - (Real)100), (Real)900);
            taug = corradj * (c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c));
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 3) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 3), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 9), (Real)8);
            V adj = lw_adjcol(c.coln2o, c.coldry, CHI(4, jp + 1), (Real)1.5, (Real)0.5, (Real)0.65);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mn2o, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<V> e = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp), (Real)4);
            Eta<V> e1 = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp + 1), (Real)4);
            Eta<V> em = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 13), (Real)4);
            Eta<V> ep = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 13), (Real)4);
            V adj = lw_adjcol(c.coln2o, c.coldry, CHI(4, jp + 1), (Real)1.5, (Real)0.5, (Real)0.65);
            taug = lw_major_b(T, B.absb, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major_b(T, B.absb, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.kb_mn2o, 5, em, c);
            fracs = lw_pfrac(T, B.fracrefb, ep);
        }
    } else if constexpr (BAND == 4) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp + 1), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 11), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<V> e = lw_eta(c.colo3, c.colco2, RAT(rat_o3co2, jp), (Real)4);
            Eta<V> e1 = lw_eta(c.colo3, c.colco2, RAT(rat_o3co2, jp + 1), (Real)4);
            Eta<V> ep = lw_eta(c.colo3, c.colco2, CHIQ(3, 2, 13), (Real)4);
            taug = lw_major_b(T, B.absb, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major_b(T, B.absb, e1, ind1 + e1.js, c.fac01, c.fac11);
            // empirical g-point adjustments aloft (rrtmg_lw_taumol.f90:1011-1017)
            if (g1 >= 8) {
                Real m = g1 == 8 ? (Real)0.92 : g1 == 9 ? (Real)0.88 : g1 == 10 ? (Real)1.07 : g1 == 11 ? (Real)1.1
                       : g1 == 12 ? (Real)0.99 : g1 == 13 ? (Real)0.88 : (Real)0.943;
                taug = taug * m;
            }
            fracs = lw_pfrac(T, B.fracrefb, ep);
        }
    } else if constexpr (BAND == 5) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 7), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 5), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   lw_minor2(T, B.ka_mo3, 9, em, c) * c.colo3 + c.wx1 * T(B.ccl4, 1);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<V> e = lw_eta(c.colo3, c.colco2, RAT(rat_o3co2, jp), (Real)4);
            Eta<V> e1 = lw_eta(c.colo3, c.colco2, RAT(rat_o3co2, jp + 1), (Real)4);
            Eta<V> ep = lw_eta(c.colo3, c.colco2, CHIQ(3, 2, 43), (Real)4);
            taug = lw_major_b(T, B.absb, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major_b(T, B.absb, e1, ind1 + e1.js, c.fac01, c.fac11) + c.wx1 * T(B.ccl4, 1);
            fracs = lw_pfrac(T, B.fracrefb, ep);
        }
    } else if constexpr (BAND == 6) {
        if (lower) {
            V adj = lw_adjcol(c.colco2, c.coldry, CHI(2, jp + 1), (Real)3, (Real)2, (Real)0.77);
            taug = c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor1(T, B.ka_mco2, c) + c.wx2 * T(B.cfc11adj, 1) + c.wx3 * T(B.cfc12, 1);
        } else {
            taug = (Real)0 + c.wx2 * T(B.cfc11adj, 1) + c.wx3 * T(B.cfc12, 1);
        }
        fracs = T(B.fracrefa, 1);
    } else if constexpr (BAND == 7) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colo3, RAT(rat_h2oo3, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colo3, RAT(rat_h2oo3, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.colh2o, c.colo3, CHIQ(1, 3, 3), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colo3, CHIQ(1, 3, 3), (Real)8);
            V adj = lw_adjcol(c.colco2, c.coldry, CHI(2, jp + 1), (Real)3, (Real)3, (Real)0.79);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mco2, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            V adj = lw_adjcol(c.colco2, c.coldry, CHI(2, jp + 1), (Real)3, (Real)2, (Real)0.79);
            taug = c.colo3 * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + adj * lw_minor1(T, B.kb_mco2, c);
            // rrtmg_lw_taumol.f90:1646-1651
            if (g1 >= 6 && g1 <= 11) {
                Real m = g1 == 6 ? (Real)0.92 : g1 == 7 ? (Real)0.88 : g1 == 8 ? (Real)1.07 : g1 == 9 ? (Real)1.1
                       : g1 == 10 ? (Real)0.99 : (Real)0.855;
                taug = taug * m;
            }
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 8) {
        V adj = lw_adjcol(c.colco2, c.coldry, CHI(2, jp + 1), (Real)3, (Real)2, (Real)0.65);
        if (lower) {
            taug = c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor1(T, B.ka_mco2, c) + c.colo3 * lw_minor1(T, B.ka_mo3, c) +
                   c.coln2o * lw_minor1(T, B.ka_mn2o, c) + c.wx3 * T(B.cfc12, 1) + c.wx4 * T(B.cfc22adj, 1);
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colo3 * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + adj * lw_minor1(T, B.kb_mco2, c) +
                   c.coln2o * lw_minor1(T, B.kb_mn2o, c) + c.wx3 * T(B.cfc12, 1) + c.wx4 * T(B.cfc22adj, 1);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 9) {
        V adj = lw_adjcol(c.coln2o, c.coldry, CHI(4, jp + 1), (Real)1.5, (Real)0.5, (Real)0.65);
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colch4, RAT(rat_h2och4, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colch4, RAT(rat_h2och4, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.colh2o, c.colch4, CHIQ(1, 6, 3), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colch4, CHIQ(1, 6, 9), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mn2o, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = c.colch4 * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + adj * lw_minor1(T, B.kb_mn2o, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 10) {
        if (lower) {
            taug = c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 11) {
        V scaleo2 = c.colo2 * c.scaleminor;
        if (lower) {
            taug = c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c) +
                   scaleo2 * lw_minor1(T, B.ka_mo2, c);
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c) +
                   scaleo2 * lw_minor1(T, B.kb_mo2, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 12) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colco2, RAT(rat_h2oco2, jp + 1), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colco2, CHIQ(1, 2, 10), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = (Real)0;
            fracs = (Real)0;
        }
    } else if constexpr (BAND == 13) {
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.coln2o, RAT(rat_h2on2o, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.coln2o, RAT(rat_h2on2o, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.colh2o, c.coln2o, CHIQ(1, 4, 1), (Real)8);
            Eta<V> emco = lw_eta(c.colh2o, c.coln2o, CHIQ(1, 4, 3), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.coln2o, CHIQ(1, 4, 5), (Real)8);
            V adj = lw_adjcol(c.colco2, c.coldry, (Real)3.55e-4, (Real)3, (Real)2, (Real)0.68);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mco2, 9, em, c) + c.colco * lw_minor2(T, B.ka_mco, 9, emco, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = c.colo3 * lw_minor1(T, B.kb_mo3, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 14) {
        if (lower) {
            taug = c.colco2 * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colco2 * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 15) {
        if (lower) {
            Eta<V> e = lw_eta(c.coln2o, c.colco2, RAT(rat_n2oco2, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.coln2o, c.colco2, RAT(rat_n2oco2, jp + 1), (Real)8);
            Eta<V> em = lw_eta(c.coln2o, c.colco2, CHIQ(4, 2, 1), (Real)8);
            Eta<V> ep = lw_eta(c.coln2o, c.colco2, CHIQ(4, 2, 1), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   c.colbrd * c.scaleminor * lw_minor2(T, B.ka_mn2, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = (Real)0;
            fracs = (Real)0;
        }
    } else {  // 16; aloft nspb(16) = 0, so ind0 = ind1 = 1 (rrtmg_lw_taumol.f90:3135-3136)
        if (lower) {
            Eta<V> e = lw_eta(c.colh2o, c.colch4, RAT(rat_h2och4, jp), (Real)8);
            Eta<V> e1 = lw_eta(c.colh2o, c.colch4, RAT(rat_h2och4, jp + 1), (Real)8);
            Eta<V> ep = lw_eta(c.colh2o, c.colch4, CHIQ(1, 6, 6), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = c.colch4 * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c);
            fracs = T(B.fracrefb, 1);
        }
    }
}
#undef CHI
#undef RAT
#undef CHIQ

}  // namespace rrtmg
