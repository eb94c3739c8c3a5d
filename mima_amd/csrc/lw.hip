// Longwave RRTMG on gfx950: prep (inatm + setcoef), band kernel (taumol +
// rtrnmr per g-point) and finish (band sums, fluxfac, heating).  Reference:
// rrtmg_lw/gcm_model/src/ of MiMA's RRTMG_LW (file:line cited per function).
//
// Band kernel: one wave per work unit (band b, up to 64/ng columns of a
// tile), lane = (column, g).  Each band is a compile-time instantiation, so a
// wave runs straight-line band code.  Per lane: a downward sweep computes the
// layer optical depth and Planck fraction, the transmittance and source
// (table or thin-layer series) and the downward radiance, keeping
// (atrans, bbugas) per layer in LDS; the surface term and an upward sweep
// follow.  Radiances are summed over the band's g-points of a column
// (segmented sum over ng adjacent lanes) into per-band level fluxes.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// prep: one thread per column, FP64 (rrtmg_lw_rad.nomcica.f90:729-852 and
// rrtmg_lw_setcoef.f90:150-413)
template <typename Real>
__global__ __launch_bounds__(256) void lw_prep_kernel(ColumnInputs in, LwWork w, const LwConst *__restrict__ cst)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    Real *coef = (Real *)w.coef;
    Real *tlevf = (Real *)w.tlev;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    const size_t fs = (size_t)nlay * ncol;
    double amttl = 0.0, wvttl = 0.0;
    int laytrop = 0;
    // level 0 Planck index (setcoef: indlev0, t0frac)
    {
        double tz0 = in.tlev[col];
        int i0 = clampi((int)(tz0 - 159.0), 1, 180);
        w.ilev[col] = i0;
        tlevf[col] = (Real)(tz0 - 159.0 - (double)i0);
    }
    for (int lay = 0; lay < nlay; ++lay) {
        const size_t ix = (size_t)lay * ncol + col;
        double pavel = in.play[ix], tavel = in.tlay[ix];
        double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double h2o = in.h2o[ix];
        double wkl1 = (h2o / (1.0 - h2o)) * amdw;
        double wkl2 = in.co2[ix], wkl3 = in.o3[ix] * amdo, wkl4 = in.n2o[ix], wkl6 = in.ch4[ix], wkl7 = in.o2[ix];
        double amm = (1.0 - wkl1) * amd + wkl1 * amw;
        double coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1.0 + wkl1));
        double summol = 0.0;
        summol = summol + wkl2;
        summol = summol + wkl3;
        summol = summol + wkl4;
        summol = summol + 0.0;
        summol = summol + wkl6;
        summol = summol + wkl7;
        double wbrodl = coldry * (1.0 - summol);
        wkl1 = coldry * wkl1;
        amttl = amttl + coldry + wkl1;
        wvttl = wvttl + wkl1;
        double wx1 = coldry * in.ccl4[ix] * 1.e-20, wx2 = coldry * in.cfc11[ix] * 1.e-20;
        double wx3 = coldry * in.cfc12[ix] * 1.e-20, wx4 = coldry * in.cfc22[ix] * 1.e-20;

        // Planck interpolation indices of the layer and of its top level
        int indlay = clampi((int)(tavel - 159.0), 1, 180);
        double tlayfrac = tavel - 159.0 - (double)indlay;
        double tz = in.tlev[ix + ncol];
        int indlev = clampi((int)(tz - 159.0), 1, 180);
        w.ilev[ix + ncol] = indlev;
        tlevf[ix + ncol] = (Real)(tz - 159.0 - (double)indlev);

        double plog = log(pavel);
        int jp = clampi((int)(36.0 - 5 * (plog + 0.04)), 1, 58);
        double fp = 5.0 * (cst->preflog[jp - 1] - plog);
        double dt0 = (tavel - cst->tref[jp - 1]) / 15.0;
        int jt = clampi((int)(3.0 + dt0), 1, 4);
        double ft = dt0 - (double)(jt - 3);
        double dt1 = (tavel - cst->tref[jp]) / 15.0;
        int jt1 = clampi((int)(3.0 + dt1), 1, 4);
        double ft1 = dt1 - (double)(jt1 - 3);
        double water = wkl1 / coldry;
        double scalefac = pavel * stpfac / tavel;
        double forfac = scalefac / (1. + water), forfrac, selffac = water * forfac, selffrac = 0.0;
        int indfor, indself = 0;
        if (plog > 4.56) {
            ++laytrop;
            double factor = (332.0 - tavel) / 36.0;
            indfor = clampi((int)factor, 1, 2);
            forfrac = factor - (double)indfor;
            factor = (tavel - 188.0) / 7.2;
            indself = clampi((int)factor - 7, 1, 9);
            selffrac = factor - (double)(indself + 7);
        } else {
            double factor = (tavel - 188.0) / 36.0;
            indfor = 3;
            forfrac = factor - 1.0;
        }
        double scaleminor = pavel / tavel;
        double scaleminorn2 = (pavel / tavel) * (wbrodl / (coldry + wkl1));
        double factor = (tavel - 180.8) / 7.2;
        int indminor = clampi((int)factor, 1, 18);
        double minorfrac = factor - (double)indminor;
        double colh2o = 1.e-20 * wkl1;
        double colco2 = 1.e-20 * (coldry * wkl2);
        double colo3 = 1.e-20 * (coldry * wkl3);
        double coln2o = 1.e-20 * (coldry * wkl4);
        double colco = 0.0;
        double colch4 = 1.e-20 * (coldry * wkl6);
        double colo2 = 1.e-20 * (coldry * wkl7);
        if (colco2 == 0.0) colco2 = 1.e-32 * coldry;
        if (colo3 == 0.0) colo3 = 1.e-32 * coldry;
        if (coln2o == 0.0) coln2o = 1.e-32 * coldry;
        if (colco == 0.0) colco = 1.e-32 * coldry;
        if (colch4 == 0.0) colch4 = 1.e-32 * coldry;
        double colbrd = 1.e-20 * wbrodl;
        double compfp = 1. - fp;
        coef[LWF_FAC10 * fs + ix] = (Real)(compfp * ft);
        coef[LWF_FAC00 * fs + ix] = (Real)(compfp * (1.0 - ft));
        coef[LWF_FAC11 * fs + ix] = (Real)(fp * ft1);
        coef[LWF_FAC01 * fs + ix] = (Real)(fp * (1.0 - ft1));
        coef[LWF_COLH2O * fs + ix] = (Real)colh2o;
        coef[LWF_COLCO2 * fs + ix] = (Real)colco2;
        coef[LWF_COLO3 * fs + ix] = (Real)colo3;
        coef[LWF_COLN2O * fs + ix] = (Real)coln2o;
        coef[LWF_COLCO * fs + ix] = (Real)colco;
        coef[LWF_COLCH4 * fs + ix] = (Real)colch4;
        coef[LWF_COLO2 * fs + ix] = (Real)colo2;
        coef[LWF_COLBRD * fs + ix] = (Real)colbrd;
        coef[LWF_SELFFAC * fs + ix] = (Real)(colh2o * selffac);
        coef[LWF_SELFFRAC * fs + ix] = (Real)selffrac;
        coef[LWF_FORFAC * fs + ix] = (Real)(colh2o * forfac);
        coef[LWF_FORFRAC * fs + ix] = (Real)forfrac;
        coef[LWF_MINORFRAC * fs + ix] = (Real)minorfrac;
        coef[LWF_SCALEMINOR * fs + ix] = (Real)scaleminor;
        coef[LWF_SCALEMINORN2 * fs + ix] = (Real)scaleminorn2;
        coef[LWF_COLDRY * fs + ix] = (Real)coldry;
        coef[LWF_PAVEL * fs + ix] = (Real)pavel;
        coef[LWF_WX1 * fs + ix] = (Real)wx1;
        coef[LWF_WX2 * fs + ix] = (Real)wx2;
        coef[LWF_WX3 * fs + ix] = (Real)wx3;
        coef[LWF_WX4 * fs + ix] = (Real)wx4;
        coef[LWF_TLAYFRAC * fs + ix] = (Real)tlayfrac;
        w.ipk[ix] = pack_lw(jp, jt, jt1, indself, indfor, indminor, indlay);
    }
    double wvsh = (amw * wvttl) / (amd * amttl);
    double pwvcm = wvsh * (1.e3 * in.plev[col]) / (1.e2 * grav);
    w.laytrop[col] = laytrop;
    w.colscal[col] = pwvcm;
    w.colscal[(size_t)ncol + col] = in.tsfc[col];
}

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
    const size_t ix = (size_t)lay * ncol + col, fs = (size_t)nlay * ncol;
#define LD(f, F) c.f = coef[F * fs + ix]
    LD(fac00, LWF_FAC00); LD(fac01, LWF_FAC01); LD(fac10, LWF_FAC10); LD(fac11, LWF_FAC11);
    LD(colh2o, LWF_COLH2O); LD(colco2, LWF_COLCO2); LD(colo3, LWF_COLO3); LD(coln2o, LWF_COLN2O);
    LD(colco, LWF_COLCO); LD(colch4, LWF_COLCH4); LD(colo2, LWF_COLO2); LD(colbrd, LWF_COLBRD);
    LD(selffac, LWF_SELFFAC); LD(selffrac, LWF_SELFFRAC); LD(forfac, LWF_FORFAC); LD(forfrac, LWF_FORFRAC);
    LD(minorfrac, LWF_MINORFRAC); LD(scaleminor, LWF_SCALEMINOR); LD(scaleminorn2, LWF_SCALEMINORN2);
    LD(coldry, LWF_COLDRY); LD(pavel, LWF_PAVEL); LD(wx1, LWF_WX1); LD(wx2, LWF_WX2); LD(wx3, LWF_WX3);
    LD(wx4, LWF_WX4); LD(tlayfrac, LWF_TLAYFRAC);
#undef LD
    int w = ipk[ix];
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
    __device__ __forceinline__ Real operator()(int off, int row) const { return t[off + (row - 1) * ng + g]; }
};

template <typename R>
struct Eta {
    R speccomb, specparm, fs;
    int js;
};

// binary-species parameter (e.g. rrtmg_lw_taumol.f90:496-509)
template <typename R>
__device__ __forceinline__ Eta<R> lw_eta(R cola, R colb, R rat, R mult)
{
    Eta<R> e;
    e.speccomb = cola + rat * colb;
    e.specparm = cola / e.speccomb;
    const R oneminus = (R)(1.0 - 1.e-6);
    if (e.specparm >= oneminus) e.specparm = oneminus;
    R specmult = mult * e.specparm;
    e.js = 1 + (int)specmult;
    e.fs = specmult - (R)(int)specmult;
    return e;
}

// lower-atmosphere major term with the 3-point eta scheme at the edges
// (rrtmg_lw_taumol.f90:550-668): facA/facB = fac00/fac10 or fac01/fac11
template <typename Real>
__device__ __forceinline__ Real lw_major3(const LTab<Real> &T, int off, const Eta<Real> &e, int ind, Real facA, Real facB)
{
    if (e.specparm < (Real)0.125) {
        Real p = e.fs - (Real)1, p2 = p * p, p4 = p2 * p2;
        Real fk0 = p4, fk1 = (Real)1 - p - (Real)2 * p4, fk2 = p + p4;
        return e.speccomb * (fk0 * facA * T(off, ind) + fk1 * facA * T(off, ind + 1) + fk2 * facA * T(off, ind + 2) +
                             fk0 * facB * T(off, ind + 9) + fk1 * facB * T(off, ind + 10) +
                             fk2 * facB * T(off, ind + 11));
    } else if (e.specparm > (Real)0.875) {
        Real p = -e.fs, p2 = p * p, p4 = p2 * p2;
        Real fk0 = p4, fk1 = (Real)1 - p - (Real)2 * p4, fk2 = p + p4;
        return e.speccomb * (fk2 * facA * T(off, ind - 1) + fk1 * facA * T(off, ind) + fk0 * facA * T(off, ind + 1) +
                             fk2 * facB * T(off, ind + 8) + fk1 * facB * T(off, ind + 9) +
                             fk0 * facB * T(off, ind + 10));
    }
    return e.speccomb * (((Real)1 - e.fs) * facA * T(off, ind) + e.fs * facA * T(off, ind + 1) +
                         ((Real)1 - e.fs) * facB * T(off, ind + 9) + e.fs * facB * T(off, ind + 10));
}

// upper-atmosphere binary-species term (e.g. rrtmg_lw_taumol.f90:740-749)
template <typename Real>
__device__ __forceinline__ Real lw_major_b(const LTab<Real> &T, int off, const Eta<Real> &e, int ind, Real facA,
                                           Real facB)
{
    return e.speccomb * (((Real)1 - e.fs) * facA * T(off, ind) + e.fs * facA * T(off, ind + 1) +
                         ((Real)1 - e.fs) * facB * T(off, ind + 5) + e.fs * facB * T(off, ind + 6));
}

template <typename Real>
__device__ __forceinline__ Real lw_four(const LTab<Real> &T, int off, int ind0, int ind1, const LwLayer<Real> &c)
{
    return c.fac00 * T(off, ind0) + c.fac10 * T(off, ind0 + 1) + c.fac01 * T(off, ind1) + c.fac11 * T(off, ind1 + 1);
}

template <typename Real>
__device__ __forceinline__ Real lw_self(const LTab<Real> &T, const LwBand &B, const LwLayer<Real> &c)
{
    return c.selffac * (T(B.selfref, c.inds) + c.selffrac * (T(B.selfref, c.inds + 1) - T(B.selfref, c.inds)));
}
template <typename Real>
__device__ __forceinline__ Real lw_for(const LTab<Real> &T, const LwBand &B, const LwLayer<Real> &c)
{
    return c.forfac * (T(B.forref, c.indf) + c.forfrac * (T(B.forref, c.indf + 1) - T(B.forref, c.indf)));
}
// minor absorber interpolated in temperature, k(19, ng)
template <typename Real>
__device__ __forceinline__ Real lw_minor1(const LTab<Real> &T, int off, const LwLayer<Real> &c)
{
    return T(off, c.indm) + c.minorfrac * (T(off, c.indm + 1) - T(off, c.indm));
}
// minor absorber interpolated in eta and temperature, k(neta, 19, ng)
template <typename Real>
__device__ __forceinline__ Real lw_minor2(const LTab<Real> &T, int off, int neta, const Eta<Real> &m,
                                          const LwLayer<Real> &c)
{
    int r1 = m.js + neta * (c.indm - 1), r2 = r1 + neta;
    Real m1 = T(off, r1) + m.fs * (T(off, r1 + 1) - T(off, r1));
    Real m2 = T(off, r2) + m.fs * (T(off, r2 + 1) - T(off, r2));
    return m1 + c.minorfrac * (m2 - m1);
}
template <typename Real>
__device__ __forceinline__ Real lw_pfrac(const LTab<Real> &T, int off, const Eta<Real> &e)
{
    return T(off, e.js) + e.fs * (T(off, e.js + 1) - T(off, e.js));
}
// column-amount adjustment of minor CO2/N2O (e.g. rrtmg_lw_taumol.f90:1314-1320)
template <typename Real>
__device__ __forceinline__ Real lw_adjcol(Real col, Real coldry, Real chiref, Real thresh, Real base, Real ex)
{
    Real rat = (Real)1.e20 * (col / coldry) / chiref;
    if (rat > thresh) return (base + pow(rat - base, ex)) * chiref * coldry * (Real)1.e-20;
    return col;
}

#define CHI(i, j) cst->chi_mls[((j)-1) * 7 + ((i)-1)]

// Optical depth and Planck fraction of one (layer, g) of band BAND,
// rrtmg_lw_taumol.f90:280-3147.
template <int BAND, typename Real>
__device__ __forceinline__ void lw_gas(const LwBand &B, const LTab<Real> &T, const LwConst *__restrict__ cst,
                                       const LwLayer<Real> &c, bool lower, Real &taug, Real &fracs)
{
    const int g1 = T.g + 1;
    int ind0, ind1;
    if (lower) {
        ind0 = ((c.jp - 1) * 5 + (c.jt - 1)) * B.nspa;
        ind1 = (c.jp * 5 + (c.jt1 - 1)) * B.nspa;
    } else {
        ind0 = ((c.jp - 13) * 5 + (c.jt - 1)) * B.nspb;
        ind1 = ((c.jp - 12) * 5 + (c.jt1 - 1)) * B.nspb;
    }
    const int jp = c.jp;
    auto R = [](double x) { return (Real)x; };
    if constexpr (BAND == 1) {
        if (lower) {
            Real corradj = c.pavel < (Real)250 ? (Real)1 - (Real)0.15 * ((Real)250 - c.pavel) / (Real)154.4 : (Real)1;
            Real taun2 = c.colbrd * c.scaleminorn2 * lw_minor1(T, B.ka_mn2, c);
            taug = corradj * (c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) +
                              lw_for(T, B, c) + taun2);
            fracs = T(B.fracrefa, 1);
        } else {
            Real corradj = (Real)1 - (Real)0.15 * (c.pavel / (Real)95.6);
            Real taun2 = c.colbrd * c.scaleminorn2 * lw_minor1(T, B.kb_mn2, c);
            taug = corradj * (c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c) + taun2);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 2) {
        if (lower) {
            Real corradj = (Real)1 - (Real).05 * (c.pavel - (Real)100) / (Real)900;
            taug = corradj * (c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c));
            fracs = T(B.fracrefa, 1);
        } else {
            taug = c.colh2o * lw_four(T, B.absb, ind0 + 1, ind1 + 1, c) + lw_for(T, B, c);
            fracs = T(B.fracrefb, 1);
        }
    } else if constexpr (BAND == 3) {
        if (lower) {
            Eta<Real> e = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.colh2o, c.colco2, R(CHI(1, 3) / CHI(2, 3)), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colco2, R(CHI(1, 9) / CHI(2, 9)), (Real)8);
            Real adj = lw_adjcol(c.coln2o, c.coldry, R(CHI(4, jp + 1)), (Real)1.5, (Real)0.5, (Real)0.65);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mn2o, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<Real> e = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp]), (Real)4);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp + 1]), (Real)4);
            Eta<Real> em = lw_eta(c.colh2o, c.colco2, R(CHI(1, 13) / CHI(2, 13)), (Real)4);
            Eta<Real> ep = lw_eta(c.colh2o, c.colco2, R(CHI(1, 13) / CHI(2, 13)), (Real)4);
            Real adj = lw_adjcol(c.coln2o, c.coldry, R(CHI(4, jp + 1)), (Real)1.5, (Real)0.5, (Real)0.65);
            taug = lw_major_b(T, B.absb, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major_b(T, B.absb, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.kb_mn2o, 5, em, c);
            fracs = lw_pfrac(T, B.fracrefb, ep);
        }
    } else if constexpr (BAND == 4) {
        if (lower) {
            Eta<Real> e = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp + 1]), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colco2, R(CHI(1, 11) / CHI(2, 11)), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<Real> e = lw_eta(c.colo3, c.colco2, R(cst->rat_o3co2[jp]), (Real)4);
            Eta<Real> e1 = lw_eta(c.colo3, c.colco2, R(cst->rat_o3co2[jp + 1]), (Real)4);
            Eta<Real> ep = lw_eta(c.colo3, c.colco2, R(CHI(3, 13) / CHI(2, 13)), (Real)4);
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
            Eta<Real> e = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.colh2o, c.colco2, R(CHI(1, 7) / CHI(2, 7)), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colco2, R(CHI(1, 5) / CHI(2, 5)), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   lw_minor2(T, B.ka_mo3, 9, em, c) * c.colo3 + c.wx1 * T(B.ccl4, 1);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Eta<Real> e = lw_eta(c.colo3, c.colco2, R(cst->rat_o3co2[jp]), (Real)4);
            Eta<Real> e1 = lw_eta(c.colo3, c.colco2, R(cst->rat_o3co2[jp + 1]), (Real)4);
            Eta<Real> ep = lw_eta(c.colo3, c.colco2, R(CHI(3, 43) / CHI(2, 43)), (Real)4);
            taug = lw_major_b(T, B.absb, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major_b(T, B.absb, e1, ind1 + e1.js, c.fac01, c.fac11) + c.wx1 * T(B.ccl4, 1);
            fracs = lw_pfrac(T, B.fracrefb, ep);
        }
    } else if constexpr (BAND == 6) {
        if (lower) {
            Real adj = lw_adjcol(c.colco2, c.coldry, R(CHI(2, jp + 1)), (Real)3, (Real)2, (Real)0.77);
            taug = c.colh2o * lw_four(T, B.absa, ind0 + 1, ind1 + 1, c) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor1(T, B.ka_mco2, c) + c.wx2 * T(B.cfc11adj, 1) + c.wx3 * T(B.cfc12, 1);
        } else {
            taug = (Real)0 + c.wx2 * T(B.cfc11adj, 1) + c.wx3 * T(B.cfc12, 1);
        }
        fracs = T(B.fracrefa, 1);
    } else if constexpr (BAND == 7) {
        if (lower) {
            Eta<Real> e = lw_eta(c.colh2o, c.colo3, R(cst->rat_h2oo3[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colo3, R(cst->rat_h2oo3[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.colh2o, c.colo3, R(CHI(1, 3) / CHI(3, 3)), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colo3, R(CHI(1, 3) / CHI(3, 3)), (Real)8);
            Real adj = lw_adjcol(c.colco2, c.coldry, R(CHI(2, jp + 1)), (Real)3, (Real)3, (Real)0.79);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c) +
                   adj * lw_minor2(T, B.ka_mco2, 9, em, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            Real adj = lw_adjcol(c.colco2, c.coldry, R(CHI(2, jp + 1)), (Real)3, (Real)2, (Real)0.79);
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
        Real adj = lw_adjcol(c.colco2, c.coldry, R(CHI(2, jp + 1)), (Real)3, (Real)2, (Real)0.65);
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
        Real adj = lw_adjcol(c.coln2o, c.coldry, R(CHI(4, jp + 1)), (Real)1.5, (Real)0.5, (Real)0.65);
        if (lower) {
            Eta<Real> e = lw_eta(c.colh2o, c.colch4, R(cst->rat_h2och4[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colch4, R(cst->rat_h2och4[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.colh2o, c.colch4, R(CHI(1, 3) / CHI(6, 3)), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colch4, R(CHI(1, 9) / CHI(6, 9)), (Real)8);
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
        Real scaleo2 = c.colo2 * c.scaleminor;
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
            Eta<Real> e = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colco2, R(cst->rat_h2oco2[jp + 1]), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colco2, R(CHI(1, 10) / CHI(2, 10)), (Real)8);
            taug = lw_major3(T, B.absa, e, ind0 + e.js, c.fac00, c.fac10) +
                   lw_major3(T, B.absa, e1, ind1 + e1.js, c.fac01, c.fac11) + lw_self(T, B, c) + lw_for(T, B, c);
            fracs = lw_pfrac(T, B.fracrefa, ep);
        } else {
            taug = (Real)0;
            fracs = (Real)0;
        }
    } else if constexpr (BAND == 13) {
        if (lower) {
            Eta<Real> e = lw_eta(c.colh2o, c.coln2o, R(cst->rat_h2on2o[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.coln2o, R(cst->rat_h2on2o[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.colh2o, c.coln2o, R(CHI(1, 1) / CHI(4, 1)), (Real)8);
            Eta<Real> emco = lw_eta(c.colh2o, c.coln2o, R(CHI(1, 3) / CHI(4, 3)), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.coln2o, R(CHI(1, 5) / CHI(4, 5)), (Real)8);
            Real adj = lw_adjcol(c.colco2, c.coldry, (Real)3.55e-4, (Real)3, (Real)2, (Real)0.68);
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
            Eta<Real> e = lw_eta(c.coln2o, c.colco2, R(cst->rat_n2oco2[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.coln2o, c.colco2, R(cst->rat_n2oco2[jp + 1]), (Real)8);
            Eta<Real> em = lw_eta(c.coln2o, c.colco2, R(CHI(4, 1) / CHI(2, 1)), (Real)8);
            Eta<Real> ep = lw_eta(c.coln2o, c.colco2, R(CHI(4, 1) / CHI(2, 1)), (Real)8);
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
            Eta<Real> e = lw_eta(c.colh2o, c.colch4, R(cst->rat_h2och4[jp]), (Real)8);
            Eta<Real> e1 = lw_eta(c.colh2o, c.colch4, R(cst->rat_h2och4[jp + 1]), (Real)8);
            Eta<Real> ep = lw_eta(c.colh2o, c.colch4, R(CHI(1, 6) / CHI(6, 6)), (Real)8);
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

// ng of each band (parrrtm.f90)
__host__ __device__ constexpr int lw_ng(int band)
{
    return band == 1 ? 10 : band == 2 ? 12 : band == 3 ? 16 : band == 4 ? 14 : band == 5 ? 16 : band == 6 ? 8
         : band == 7 ? 12 : band == 8 ? 8 : band == 9 ? 12 : band == 10 ? 6 : band == 11 ? 8 : band == 12 ? 8
         : band == 13 ? 4 : 2;
}

// One band of one work unit (rrtmg_lw_rtrnmr.f90:483-716, clear sky)
template <int BAND, typename Real>
__device__ void lw_band_body(const ColumnInputs &in, const LwWork &w, const LwBand &B, const Real *__restrict__ tab,
                             int exp_off, int tfn_off, int plk_off, const LwConst *__restrict__ cst, WaveUnit u,
                             int tile, Real *smem)
{
    constexpr int NG = lw_ng(BAND);
    const int ncol = in.ncol, nlay = in.nlay;
    const int lane = threadIdx.x;
    const int cl = lane / NG, g = lane - cl * NG;
    const int col0 = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col0 < ncol;
    const int col = active ? col0 : 0;
    Real *s_atr = smem;                            // [nlay][64]
    Real *s_bbu = smem + (size_t)nlay * 64;        // [nlay][64]
    Real *s_red = smem + (size_t)2 * nlay * 64;    // [64]
    const Real *coef = (const Real *)w.coef;
    const Real *tlevf = (const Real *)w.tlev;
    const Real *exp_tbl = tab + exp_off, *tfn_tbl = tab + tfn_off;
    const Real *plk = tab + plk_off + (BAND - 1) * 181;  // totplnk(:, BAND)
    LTab<Real> T{tab, NG, g};
    Real *bf_up = (Real *)w.bflux + (size_t)(BAND - 1) * 2 * (nlay + 1) * ncol;
    Real *bf_dn = bf_up + (size_t)(nlay + 1) * ncol;

    const int laytrop = w.laytrop[col];
    const double pwvcm = w.colscal[col];
    const double tbound = w.colscal[(size_t)ncol + col];
    // diffusivity secant (rrtmg_lw_rtrnmr.f90:272-280)
    double sd = 1.66;
    if (!(BAND == 1 || BAND == 4 || BAND >= 10)) {
        sd = B.a0 + B.a1 * exp(B.a2 * pwvcm);
        if (sd > 1.80) sd = 1.80;
        if (sd < 1.50) sd = 1.50;
    }
    const Real secdiff = (Real)sd;
    const Real semiss = (Real)in.emis[(size_t)(BAND - 1) * ncol + col];
    // surface Planck (rrtmg_lw_setcoef.f90:190-194), FP64
    int indbound = clampi((int)(tbound - 159.0), 1, 180);
    double tbndfrac = tbound - 159.0 - (double)indbound;
    const Real plankbnd =
        (Real)((double)semiss * ((double)plk[indbound - 1] + tbndfrac * ((double)plk[indbound] - (double)plk[indbound - 1])));
    auto planklev = [&](int lev) {
        int i = w.ilev[(size_t)lev * ncol + col];
        Real f = tlevf[(size_t)lev * ncol + col];
        return plk[i - 1] + f * (plk[i] - plk[i - 1]);
    };
    const double *tauaer = in.tauaer;

    // downward sweep
    Real radld = (Real)0, fracs1 = (Real)0;
    Real plev_up = planklev(nlay);
    if (active && g == 0) bf_dn[(size_t)nlay * ncol + col] = (Real)0;  // drad(nlayers) stays 0
    LwLayer<Real> c;
    for (int lev = nlay; lev >= 1; --lev) {
        lw_load_layer(c, coef, w.ipk, lev - 1, col, ncol, nlay);
        Real taug, plfrac;
        lw_gas<BAND, Real>(B, T, cst, c, lev <= laytrop, taug, plfrac);
        if (tauaer) taug = taug + (Real)tauaer[((size_t)(BAND - 1) * nlay + (lev - 1)) * ncol + col];
        Real plev_dn = planklev(lev - 1);
        Real blay = plk[c.indlay - 1] + c.tlayfrac * (plk[c.indlay] - plk[c.indlay - 1]);
        Real dplankup = plev_up - blay, dplankdn = plev_dn - blay;
        Real odepth = secdiff * taug;
        if (odepth < (Real)0) odepth = (Real)0;
        Real atrans, bbd, bbugas;
        if (odepth <= (Real)0.06) {
            atrans = odepth - (Real)0.5 * odepth * odepth;
            odepth = (Real)0.166667 * odepth;
            bbd = plfrac * (blay + dplankdn * odepth);
            bbugas = plfrac * (blay + dplankup * odepth);
        } else {
            int itr = tbl_index(odepth);
            atrans = (Real)1 - exp_tbl[itr];
            Real tausfac = tfn_tbl[itr];
            bbd = plfrac * (blay + tausfac * dplankdn);
            bbugas = plfrac * (blay + tausfac * dplankup);
        }
        radld = radld + (bbd - radld) * atrans;
        s_atr[(size_t)(lev - 1) * 64 + lane] = atrans;
        s_bbu[(size_t)(lev - 1) * 64 + lane] = bbugas;
        Real s = seg_sum(active ? radld : (Real)0, lane, NG, s_red);
        if (active && g == 0) bf_dn[(size_t)(lev - 1) * ncol + col] = s;
        plev_up = plev_dn;
        if (lev == 1) fracs1 = plfrac;
    }
    // surface emission and reflection, upward sweep
    Real radlu = fracs1 * plankbnd + ((Real)1 - semiss) * radld;
    {
        Real s = seg_sum(active ? radlu : (Real)0, lane, NG, s_red);
        if (active && g == 0) bf_up[col] = s;
    }
    for (int lev = 1; lev <= nlay; ++lev) {
        Real atrans = s_atr[(size_t)(lev - 1) * 64 + lane];
        Real bbugas = s_bbu[(size_t)(lev - 1) * 64 + lane];
        radlu = radlu + (bbugas - radlu) * atrans;
        Real s = seg_sum(active ? radlu : (Real)0, lane, NG, s_red);
        if (active && g == 0) bf_up[(size_t)lev * ncol + col] = s;
    }
}

template <typename Real>
__global__ __launch_bounds__(64) void lw_band_kernel(ColumnInputs in, LwWork w, const LwBand *__restrict__ bands,
                                                     int nwaves, int exp_off, int tfn_off, int plk_off,
                                                     const Real *__restrict__ tab, const LwConst *__restrict__ cst,
                                                     const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int tile = blockIdx.x / nwaves;
    const WaveUnit u = units[blockIdx.x % nwaves];
    const LwBand B = bands[u.band];
#define BODY(b)                                                                                          \
    case b - 1:                                                                                          \
        lw_band_body<b, Real>(in, w, B, tab, exp_off, tfn_off, plk_off, cst, u, tile, smem);              \
        break;
    switch (u.band) {
        BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
        BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
    }
#undef BODY
}

// finish: band sums, fluxfac, net flux and heating (rrtmg_lw_rtrnmr.f90:719-777)
template <typename Real>
__global__ __launch_bounds__(256) void lw_finish_kernel(ColumnInputs in, LwWork w, LwOutputs out,
                                                        const LwBand *__restrict__ bands, double heatfac)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    const Real *bflux = (const Real *)w.bflux;
    const size_t bstride = (size_t)(nlay + 1) * ncol;
    const double fluxfac = 2.0 * asin(1.0) * 2.e4;
    double fnet_below = 0.0;
    for (int k = 0; k <= nlay; ++k) {
        double up = 0.0, dn = 0.0;
        for (int b = 0; b < 16; ++b) {
            double dw = bands[b].delwave;
            up = up + ((double)bflux[(size_t)b * 2 * bstride + (size_t)k * ncol + col] * 0.5) * dw;
            dn = dn + ((double)bflux[(size_t)b * 2 * bstride + bstride + (size_t)k * ncol + col] * 0.5) * dw;
        }
        up = up * fluxfac;
        dn = dn * fluxfac;
        size_t ix = (size_t)k * ncol + col;
        out.uflx[ix] = up;
        out.dflx[ix] = dn;
        out.uflxc[ix] = up;
        out.dflxc[ix] = dn;
        double fnet = up - dn;
        if (k > 0) {
            size_t il = (size_t)(k - 1) * ncol + col;
            double h = heatfac * (fnet_below - fnet) / (in.plev[il] - in.plev[il + ncol]);
            out.hr[il] = h;
            out.hrc[il] = h;
        }
        fnet_below = fnet;
    }
}

template <typename Real>
static hipError_t launch_lw_t(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t,
                              double heatfac, hipStream_t s, hipEvent_t *ev)
{
    const int nb = (in.ncol + 255) / 256;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(lw_prep_kernel<Real>, dim3(nb), dim3(256), 0, s, in, w, t.cst);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = sizeof(Real) * ((size_t)2 * in.nlay * 64 + 64);
    if (ev) (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL(lw_band_kernel<Real>, dim3(ntiles * t.desc.nwaves), dim3(64), shmem, s, in, w, t.bands,
                       t.desc.nwaves, t.desc.exp_tbl, t.desc.tfn_tbl, t.desc.totplnk, (const Real *)t.tab, t.cst,
                       t.units);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(lw_finish_kernel<Real>, dim3(nb), dim3(256), 0, s, in, w, out, t.bands, heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_lw(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t, int mixed,
                     hipStream_t s, hipEvent_t *ev)
{
    return mixed ? launch_lw_t<float>(in, out, w, t, t.heatfac, s, ev)
                 : launch_lw_t<double>(in, out, w, t, t.heatfac, s, ev);
}

}  // namespace rrtmg
