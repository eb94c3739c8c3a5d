// Shortwave RRTMG on gfx950: prep (inatm_sw + setcoef_sw), band kernel
// (taumol_sw + spcvrt_sw + reftra_sw + vrtqdr_sw per g-point) and finish
// (band sums, net flux, heating).  Reference: rrtmg_sw/gcm_model/src/ of
// MiMA's RRTMG_SW (file:line cited per function).
//
// Band kernel layout: one wave per work unit (band b, up to 64/ng columns of a
// tile); lane = (column, g).  Each lane runs the adding method for its
// g-point: a bottom-up sweep keeps prup/prupd of every level in LDS, a
// top-down sweep recomputes the layer optics, carries ztdn/prdnd/ptdbt and
// produces level fluxes, which are summed over the band's g-points of a
// column (segmented sum over ng adjacent lanes) into per-band fluxes.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// prep: one thread per column, FP64 (rrtmg_sw_rad.nomcica.f90:933-1047 and
// rrtmg_sw_setcoef.f90:129-284); writes the per-layer coefficients in the
// band kernel's precision and the integer table indices.
template <typename Real>
__global__ __launch_bounds__(256) void sw_prep_kernel(ColumnInputs in, SwWork w, const double *__restrict__ preflog,
                                                      const double *__restrict__ tref)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    Real *coef = (Real *)w.coef;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    int laytrop = 0;
    int jp_store[MAX_NLAY + 1];
    for (int lay = 0; lay < nlay; ++lay) {
        const size_t ix = (size_t)lay * ncol + col;
        double pavel = in.play[ix], tavel = in.tlay[ix];
        double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double h2o = in.h2o[ix];
        double wkl1 = (h2o / (1.0 - h2o)) * amdw;
        double amm = (1.0 - wkl1) * amd + wkl1 * amw;
        double coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1.0 + wkl1));
        double colh2o = 1.e-20 * (coldry * wkl1);
        double colco2 = 1.e-20 * (coldry * in.co2[ix]);
        double colo3 = 1.e-20 * (coldry * (in.o3[ix] * amdo));
        double colch4 = 1.e-20 * (coldry * in.ch4[ix]);
        double colo2 = 1.e-20 * (coldry * in.o2[ix]);
        double colmol = 1.e-20 * coldry + colh2o;
        if (colco2 == 0.0) colco2 = 1.e-32 * coldry;
        if (colch4 == 0.0) colch4 = 1.e-32 * coldry;
        if (colo2 == 0.0) colo2 = 1.e-32 * coldry;

        double plog = log(pavel);
        int jp = clampi((int)(36.0 - 5 * (plog + 0.04)), 1, 58);
        double fp = 5.0 * (preflog[jp - 1] - plog);
        double dt0 = (tavel - tref[jp - 1]) / 15.0;
        int jt = clampi((int)(3.0 + dt0), 1, 4);
        double ft = dt0 - (double)(jt - 3);
        double dt1 = (tavel - tref[jp]) / 15.0;
        int jt1 = clampi((int)(3.0 + dt1), 1, 4);
        double ft1 = dt1 - (double)(jt1 - 3);
        double water = (coldry * wkl1) / coldry;
        double scalefac = pavel * stpfac / tavel;
        double forfac, forfrac, selffac, selffrac;
        int indfor, indself;
        if (plog > 4.56) {
            ++laytrop;
            forfac = scalefac / (1. + water);
            double factor = (332.0 - tavel) / 36.0;
            indfor = clampi((int)factor, 1, 2);
            forfrac = factor - (double)indfor;
            selffac = water * forfac;
            factor = (tavel - 188.0) / 7.2;
            indself = clampi((int)factor - 7, 1, 9);
            selffrac = factor - (double)(indself + 7);
        } else {
            forfac = scalefac / (1. + water);
            double factor = (tavel - 188.0) / 36.0;
            indfor = 3;
            forfrac = factor - 1.0;
            selffac = 0.0;
            selffrac = 0.0;
            indself = 0;
        }
        double compfp = 1.0 - fp;
        const size_t fs = (size_t)nlay * ncol;
        coef[SWF_FAC10 * fs + ix] = (Real)(compfp * ft);
        coef[SWF_FAC00 * fs + ix] = (Real)(compfp * (1.0 - ft));
        coef[SWF_FAC11 * fs + ix] = (Real)(fp * ft1);
        coef[SWF_FAC01 * fs + ix] = (Real)(fp * (1.0 - ft1));
        coef[SWF_COLH2O * fs + ix] = (Real)colh2o;
        coef[SWF_COLCO2 * fs + ix] = (Real)colco2;
        coef[SWF_COLO3 * fs + ix] = (Real)colo3;
        coef[SWF_COLCH4 * fs + ix] = (Real)colch4;
        coef[SWF_COLO2 * fs + ix] = (Real)colo2;
        coef[SWF_COLMOL * fs + ix] = (Real)colmol;
        coef[SWF_SELFFAC * fs + ix] = (Real)selffac;
        coef[SWF_SELFFRAC * fs + ix] = (Real)selffrac;
        coef[SWF_FORFAC * fs + ix] = (Real)forfac;
        coef[SWF_FORFRAC * fs + ix] = (Real)forfrac;
        w.ipk[ix] = pack_sw(jp, jt, jt1, indself, indfor);
        jp_store[lay] = jp;
    }
    w.laytrop[col] = laytrop;
    // reference layers for the solar source (taumol16..29): bands searching
    // the lower atmosphere (18-25) take min(lay+1, laytrop) at the last jp
    // crossing of layreffr, default laytrop; bands searching aloft (16, 17,
    // 27, 28, 29) take the last crossing above laytrop, default nlayers; band
    // 26 uses laytrop.  (e.g. rrtmg_sw_taumol.f90:292-295, 391-394, 494-496)
    const int lowref[14] = {0, 0, 6, 3, 3, 8, 2, 6, 1, 2, 0, 0, 0, 0};
    const int upref[14] = {18, 30, 0, 0, 0, 0, 0, 0, 0, 0, 0, 32, 58, 49};
    for (int b = 0; b < 14; ++b) {
        int lsf;
        if (upref[b] > 0) {
            lsf = nlay;
            for (int lay = max(laytrop + 1, 2); lay <= nlay; ++lay)
                if (jp_store[lay - 2] < upref[b] && jp_store[lay - 1] >= upref[b]) lsf = lay;
        } else {
            lsf = laytrop;
            if (lowref[b] > 0)
                for (int lay = 1; lay <= laytrop; ++lay) {
                    int jpn = lay < nlay ? jp_store[lay] : jp_store[nlay - 1];
                    if (jp_store[lay - 1] < lowref[b] && jpn >= lowref[b]) lsf = min(lay + 1, laytrop);
                }
        }
        w.laysolfr[(size_t)b * ncol + col] = max(lsf, 1);
    }
}

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
    const size_t ix = (size_t)lay * ncol + col, fs = (size_t)nlay * ncol;
    c.fac00 = coef[SWF_FAC00 * fs + ix];
    c.fac01 = coef[SWF_FAC01 * fs + ix];
    c.fac10 = coef[SWF_FAC10 * fs + ix];
    c.fac11 = coef[SWF_FAC11 * fs + ix];
    c.colh2o = coef[SWF_COLH2O * fs + ix];
    c.colco2 = coef[SWF_COLCO2 * fs + ix];
    c.colo3 = coef[SWF_COLO3 * fs + ix];
    c.colch4 = coef[SWF_COLCH4 * fs + ix];
    c.colo2 = coef[SWF_COLO2 * fs + ix];
    c.colmol = coef[SWF_COLMOL * fs + ix];
    c.selffac = coef[SWF_SELFFAC * fs + ix];
    c.selffrac = coef[SWF_SELFFRAC * fs + ix];
    c.forfac = coef[SWF_FORFAC * fs + ix];
    c.forfrac = coef[SWF_FORFRAC * fs + ix];
    int w = ipk[ix];
    c.jp = unpack_jp(w);
    c.jt = unpack_jt(w);
    c.jt1 = unpack_jt1(w);
    c.inds = unpack_inds(w);
    c.indf = unpack_indf(w);
}

// Table reader for one lane: row r (1-based, as the reference's ind) of a
// [rows][ng] table at `off`.
template <typename Real>
struct Tab {
    const Real *__restrict__ t;
    int ng, g;
    __device__ __forceinline__ Real operator()(int off, int row) const { return t[off + (row - 1) * ng + g]; }
};

// binary-species parameter of a band (e.g. rrtmg_sw_taumol.f90:276-283)
template <typename R>
__device__ __forceinline__ void sw_eta(R cola, R colb, R strrat, R mult, R &speccomb, int &js, R &fs)
{
    speccomb = cola + strrat * colb;
    R specparm = cola / speccomb;
    const R oneminus = (R)(1.0 - 1.e-6);
    if (specparm >= oneminus) specparm = oneminus;
    R specmult = mult * specparm;
    js = 1 + (int)specmult;
    fs = specmult - (R)(int)specmult;
}

// 8-point (eta, T, p) interpolation; s1 = jt stride (9 lower, 5 upper)
template <typename Real>
__device__ __forceinline__ Real sw_maj8(const Tab<Real> &T, int off, int ind0, int ind1, Real fs, const SwLayer<Real> &c,
                                        int s1)
{
    Real fac000 = ((Real)1 - fs) * c.fac00, fac010 = ((Real)1 - fs) * c.fac10;
    Real fac100 = fs * c.fac00, fac110 = fs * c.fac10;
    Real fac001 = ((Real)1 - fs) * c.fac01, fac011 = ((Real)1 - fs) * c.fac11;
    Real fac101 = fs * c.fac01, fac111 = fs * c.fac11;
    return fac000 * T(off, ind0) + fac100 * T(off, ind0 + 1) + fac010 * T(off, ind0 + s1) +
           fac110 * T(off, ind0 + s1 + 1) + fac001 * T(off, ind1) + fac101 * T(off, ind1 + 1) +
           fac011 * T(off, ind1 + s1) + fac111 * T(off, ind1 + s1 + 1);
}

template <typename Real>
__device__ __forceinline__ Real sw_maj4(const Tab<Real> &T, int off, int ind0, int ind1, const SwLayer<Real> &c)
{
    return c.fac00 * T(off, ind0) + c.fac10 * T(off, ind0 + 1) + c.fac01 * T(off, ind1) + c.fac11 * T(off, ind1 + 1);
}

template <typename Real>
__device__ __forceinline__ Real sw_cont(const Tab<Real> &T, const SwBand &B, const SwLayer<Real> &c, bool self)
{
    Real f = c.forfac * (T(B.forref, c.indf) + c.forfrac * (T(B.forref, c.indf + 1) - T(B.forref, c.indf)));
    if (!self) return f;
    Real s = c.selffac * (T(B.selfref, c.inds) + c.selffrac * (T(B.selfref, c.inds + 1) - T(B.selfref, c.inds)));
    return s + f;
}

// Gas and Rayleigh optical depth of one (layer, g) for band jb (16..29),
// rrtmg_sw_taumol.f90:243-1536.
template <typename Real>
__device__ __forceinline__ void sw_gas(int jb, const SwBand &B, const Tab<Real> &T, const SwLayer<Real> &c, bool lower,
                                       Real &taug, Real &taur)
{
    const int g1 = T.g + 1;
    (void)g1;
    Real speccomb, fs;
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
            sw_eta<Real>(c.colh2o, c.colch4, (Real)252.131, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            taug = c.colch4 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        }
        taur = c.colmol * rayl;
        break;
    case 17:
        if (lower) {
            sw_eta<Real>(c.colh2o, c.colco2, (Real)0.364641, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            sw_eta<Real>(c.colh2o, c.colco2, (Real)0.364641, (Real)4, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absb, ind0 + js, ind1 + js, fs, c, 5) + c.colh2o * sw_cont(T, B, c, false);
        }
        taur = c.colmol * rayl;
        break;
    case 18:
        if (lower) {
            sw_eta<Real>(c.colh2o, c.colch4, (Real)38.9589, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            taug = c.colch4 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        }
        taur = c.colmol * rayl;
        break;
    case 19:
        if (lower) {
            sw_eta<Real>(c.colh2o, c.colco2, (Real)5.49281, (Real)8, speccomb, js, fs);
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
            sw_eta<Real>(c.colh2o, c.colco2, (Real)0.0045321, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true);
        } else {
            sw_eta<Real>(c.colh2o, c.colco2, (Real)0.0045321, (Real)4, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absb, ind0 + js, ind1 + js, fs, c, 5) + c.colh2o * sw_cont(T, B, c, false);
        }
        taur = c.colmol * rayl;
        break;
    case 22: {
        Real o2cont = (Real)4.35e-4 * c.colo2 / (Real)(350.0 * 2.0);
        if (lower) {
            sw_eta<Real>(c.colh2o, c.colo2, (Real)(1.6 * 0.022708), (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9) + c.colh2o * sw_cont(T, B, c, true) +
                   o2cont;
        } else {
            taug = c.colo2 * (Real)1.6 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c) + o2cont;
        }
        taur = c.colmol * rayl;
        break;
    }
    case 23:
        taug = lower ? c.colh2o * ((Real)1.029 * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + sw_cont(T, B, c, true))
                     : (Real)0;
        taur = c.colmol * T(B.raylg, 1);
        break;
    case 24:
        if (lower) {
            sw_eta<Real>(c.colh2o, c.colo2, (Real)0.124692, (Real)8, speccomb, js, fs);
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
        taug = lower ? c.colh2o * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c) + c.colo3 * T(B.abso3a, 1)
                     : c.colo3 * T(B.abso3b, 1);
        break;
    case 26:
        taug = (Real)0;
        taur = c.colmol * T(B.raylg, 1);
        break;
    case 27:
        taur = c.colmol * T(B.raylg, 1);
        taug = lower ? c.colo3 * sw_maj4(T, B.absa, ind0 + 1, ind1 + 1, c)
                     : c.colo3 * sw_maj4(T, B.absb, ind0 + 1, ind1 + 1, c);
        break;
    case 28:
        if (lower) {
            sw_eta<Real>(c.colo3, c.colo2, (Real)6.67029e-07, (Real)8, speccomb, js, fs);
            taug = speccomb * sw_maj8(T, B.absa, ind0 + js, ind1 + js, fs, c, 9);
        } else {
            sw_eta<Real>(c.colo3, c.colo2, (Real)6.67029e-07, (Real)4, speccomb, js, fs);
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

// PIFM two-stream layer reflectance/transmittance, kmodts = 2, asymmetry 0
// (rrtmg_sw_reftra.f90:127-303).  F is the arithmetic type.
template <typename F, typename Real>
__device__ __forceinline__ void reftra(F zto1, F zw, F prmuz, const Real *__restrict__ exp_tbl, F &pref, F &prefd,
                                       F &ptra, F &ptrad)
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
        F ze1 = fmin(zto1 / prmuz, (F)500);
        F ze2 = ze1 <= od_lo ? (F)1 - ze1 + (F)0.5 * ze1 * ze1 : (F)exp_tbl[tbl_index(ze1)];
        if (ze2 == (F)1) {
            pref = (F)0;
            ptra = (F)1;
            prefd = (F)0;
            ptrad = (F)1;
        } else {
            pref = (zgt - za1 * ((F)1 - ze2)) / ((F)1 + zgt);
            ptra = (F)1 - pref;
            prefd = zgt / ((F)1 + zgt);
            ptrad = (F)1 - prefd;
        }
    } else {
        F za1 = zgamma1 * zgamma4 + zgamma2 * zgamma3;
        F za2 = zgamma1 * zgamma3 + zgamma2 * zgamma4;
        F zrk = sqrt(zgamma1 * zgamma1 - zgamma2 * zgamma2);
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
        F zbeta = (zgamma1 - zrk) / zrkg;
        F ze1 = fmin(zrk * zto1, (F)500);
        F ze2 = fmin(zto1 / prmuz, (F)500);
        F zem1 = ze1 <= od_lo ? (F)1 - ze1 + (F)0.5 * ze1 * ze1 : (F)exp_tbl[tbl_index(ze1)];
        F zep1 = (F)1 / zem1;
        F zem2 = ze2 <= od_lo ? (F)1 - ze2 + (F)0.5 * ze2 * ze2 : (F)exp_tbl[tbl_index(ze2)];
        F zep2 = (F)1 / zem2;
        F zdenr = zr4 * zep1 + zr5 * zem1;
        F zdent = zr4 * zep1 + zr5 * zem1;  // zt4 = zr4, zt5 = zr5
        if (zdenr >= -eps && zdenr <= eps) {
            pref = eps;
            ptra = zem2;
        } else {
            pref = zw * (zr1 * zep1 - zr2 * zem1 - zr3 * zem2) / zdenr;
            ptra = zem2 - zem2 * zw * (zt1 * zep1 - zt2 * zem1 - zt3 * zep2) / zdent;
        }
        F zemm = zem1 * zem1;
        F zdend = (F)1 / (((F)1 - zbeta * zemm) * zrkg);
        prefd = zgamma2 * ((F)1 - zemm) * zdend;
        ptrad = zrk2 * zem1 * zdend;
    }
}

// optics of one (layer, g): total optical depth, reftra and direct-beam
// transmittance (rrtmg_sw_spcvrt.f90:375-550, clear sky, no aerosol)
template <typename Real>
__device__ __forceinline__ void sw_layer_optics(Real taug, Real taur, Real prmu0, const Real *__restrict__ exp_tbl,
                                                Real &r, Real &rd, Real &t, Real &td, Real &dbt)
{
    Real tau = taur + taug;
    Real w = taur / tau;
    bool island = false;
    if (sizeof(Real) == 4) {
        // FP64 islands of the mixed path: near-conservative scattering
        // (cancellation in sqrt(g1^2 - g2^2) and 1 - exp) and the removable
        // 0/0 of the direct-beam solution at k*mu0 = 1, where both the
        // numerator and zdenr of reftra vanish
        Real g1 = ((Real)8 - w * (Real)5) * (Real)0.25, g2 = (Real)3 * w * (Real)0.25;
        Real zrp = sqrt(fmax(g1 * g1 - g2 * g2, (Real)0)) * prmu0;
        island = w > (Real)0.999 || fabs((Real)1 - zrp) < (Real)0.02;
    }
    if (island) {
        double R, RD, T, TD;
        double t64 = (double)taur + (double)taug;
        reftra<double, Real>(t64, (double)taur / t64, (double)prmu0, exp_tbl, R, RD, T, TD);
        r = (Real)R;
        rd = (Real)RD;
        t = (Real)T;
        td = (Real)TD;
    } else {
        reftra<Real, Real>(tau, w, prmu0, exp_tbl, r, rd, t, td);
    }
    Real ze1 = tau / prmu0;
    dbt = ze1 <= (Real)0.06 ? (Real)1 - ze1 + (Real)0.5 * ze1 * ze1 : exp_tbl[tbl_index(ze1)];
}

template <typename Real>
__global__ __launch_bounds__(64) void sw_band_kernel(ColumnInputs in, SwWork w, const SwBand *__restrict__ bands,
                                                     int nwaves, int exp_off, const Real *__restrict__ tab,
                                                     const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int ncol = in.ncol, nlay = in.nlay;
    const int tile = blockIdx.x / nwaves;
    const WaveUnit u = units[blockIdx.x % nwaves];
    const SwBand B = bands[u.band];
    const int jb = u.band + 16;
    const int lane = threadIdx.x;
    const int ng = B.ng;
    const int cl = lane / ng, g = lane - cl * ng;
    const int col = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col < ncol && in.coszen[min(col, ncol - 1)] >= 1.e-10;
    Real *s_prup = smem;                           // [(nlay+1)][64]
    Real *s_prupd = smem + (size_t)(nlay + 1) * 64;
    Real *s_red = smem + (size_t)2 * (nlay + 1) * 64;  // [64]
    const Real *coef = (const Real *)w.coef;
    Real *bflux = (Real *)w.bflux;
    const Real *exp_tbl = tab + exp_off;
    Tab<Real> T{tab, ng, g};

    Real prmu0 = (Real)1, zincflx = (Real)0, albp = (Real)0, albd = (Real)0;
    int laytrop = 0;
    SwLayer<Real> c;
    if (active) {
        prmu0 = (Real)in.coszen[col];
        laytrop = w.laytrop[col];
        const bool nir = (jb <= 24) || jb == 29;  // rrtmg_sw_rad.nomcica.f90:567-577
        albp = (Real)(nir ? in.aldir[col] : in.asdir[col]);
        albd = (Real)(nir ? in.aldif[col] : in.asdif[col]);
        int lsf = w.laysolfr[(size_t)u.band * ncol + col];
        sw_load_layer(c, coef, w.ipk, lsf - 1, col, ncol, nlay);
        double sfz = sw_sfluxzen(jb, B, T, c, lsf <= laytrop);
        zincflx = (Real)(in.adjflux * sfz * (double)prmu0);
    }

    // bottom-up sweep: reflectances of the atmosphere below each level
    Real prup_n = albp, prupd_n = albd;
    s_prup[(size_t)nlay * 64 + lane] = albp;
    s_prupd[(size_t)nlay * 64 + lane] = albd;
    for (int lay = 1; lay <= nlay; ++lay) {
        const int jk = nlay + 1 - lay;  // top-down index of the layer
        Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
        if (active) {
            sw_load_layer(c, coef, w.ipk, lay - 1, col, ncol, nlay);
            Real taug, taur;
            sw_gas(jb, B, T, c, lay <= laytrop, taug, taur);
            sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
        }
        Real zreflect = (Real)1 / ((Real)1 - prupd_n * rd);
        Real prup = r + (td * ((t - dbt) * prupd_n + dbt * prup_n)) * zreflect;
        Real prupd = rd + td * td * prupd_n * zreflect;
        s_prup[(size_t)(jk - 1) * 64 + lane] = prup;
        s_prupd[(size_t)(jk - 1) * 64 + lane] = prupd;
        prup_n = prup;
        prupd_n = prupd;
    }

    // top-down sweep with level fluxes
    Real ztdn = (Real)1, prdnd = (Real)0, ptdbt = (Real)1;
    const size_t lvstride = ncol, bstride = (size_t)(nlay + 1) * ncol;
    Real *bf_up = bflux + (size_t)u.band * 2 * bstride;
    Real *bf_dn = bf_up + bstride;
    for (int jk = 1; jk <= nlay + 1; ++jk) {
        if (jk > 1) {
            const int lay = nlay + 2 - jk;  // layer between levels jk-1 and jk
            Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
            if (active) {
                sw_load_layer(c, coef, w.ipk, lay - 1, col, ncol, nlay);
                Real taug, taur;
                sw_gas(jb, B, T, c, lay <= laytrop, taug, taur);
                sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
            }
            Real zreflect = (Real)1 / ((Real)1 - rd * prdnd);
            Real ztdn_n = ptdbt * t + (td * ((ztdn - ptdbt) + ptdbt * r * prdnd)) * zreflect;
            Real prdnd_n = rd + td * td * prdnd * zreflect;
            ptdbt = dbt * ptdbt;
            ztdn = ztdn_n;
            prdnd = prdnd_n;
        }
        Real prup = s_prup[(size_t)(jk - 1) * 64 + lane];
        Real prupd = s_prupd[(size_t)(jk - 1) * 64 + lane];
        Real zreflect = (Real)1 / ((Real)1 - prdnd * prupd);
        Real pfu = (ptdbt * prup + (ztdn - ptdbt) * prupd) * zreflect;
        Real pfd = ptdbt + (ztdn - ptdbt + ptdbt * prup * prdnd) * zreflect;
        const int lev = nlay + 1 - jk;  // bottom-up level index
        Real su = seg_sum(active ? zincflx * pfu : (Real)0, lane, ng, s_red);
        Real sd = seg_sum(active ? zincflx * pfd : (Real)0, lane, ng, s_red);
        if (active && g == 0) {
            bf_up[(size_t)lev * lvstride + col] = su;
            bf_dn[(size_t)lev * lvstride + col] = sd;
        }
    }
}

// finish: band sums, net flux and heating (rrtmg_sw_rad.nomcica.f90:700-726)
template <typename Real>
__global__ __launch_bounds__(256) void sw_finish_kernel(ColumnInputs in, SwWork w, SwOutputs out, double heatfac)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    const Real *bflux = (const Real *)w.bflux;
    const size_t bstride = (size_t)(nlay + 1) * ncol;
    if (in.coszen[col] < 1.e-10) {
        for (int k = 0; k <= nlay; ++k) {
            size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = 0.0;
            out.dflx[ix] = 0.0;
            out.uflxc[ix] = 0.0;
            out.dflxc[ix] = 0.0;
            if (k < nlay) {
                out.hr[ix] = 0.0;
                out.hrc[ix] = 0.0;
            }
        }
        return;
    }
    double net_below = 0.0;
    for (int k = 0; k <= nlay; ++k) {
        double up = 0.0, dn = 0.0;
        for (int b = 0; b < 14; ++b) {
            up += (double)bflux[(size_t)b * 2 * bstride + (size_t)k * ncol + col];
            dn += (double)bflux[(size_t)b * 2 * bstride + bstride + (size_t)k * ncol + col];
        }
        size_t ix = (size_t)k * ncol + col;
        out.uflx[ix] = up;
        out.dflx[ix] = dn;
        out.uflxc[ix] = up;
        out.dflxc[ix] = dn;
        double net = dn - up;
        if (k > 0) {
            size_t il = (size_t)(k - 1) * ncol + col;
            double pdp = in.plev[il] - in.plev[il + ncol];
            double h = (k == nlay) ? 0.0 : (net - net_below) * (heatfac / pdp);
            out.hr[il] = h;
            out.hrc[il] = h;
        }
        net_below = net;
    }
}

template <typename Real>
static hipError_t launch_sw_t(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t,
                              hipStream_t s, hipEvent_t *ev)
{
    const int nb = (in.ncol + 255) / 256;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(sw_prep_kernel<Real>, dim3(nb), dim3(256), 0, s, in, w, t.preflog, t.tref);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = sizeof(Real) * ((size_t)2 * (in.nlay + 1) * 64 + 64);
    if (ev) (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL(sw_band_kernel<Real>, dim3(ntiles * t.desc.nwaves), dim3(64), shmem, s, in, w, t.bands,
                       t.desc.nwaves, t.desc.exp_tbl, (const Real *)t.tab, t.units);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(sw_finish_kernel<Real>, dim3(nb), dim3(256), 0, s, in, w, out, t.heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_sw(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t, int mixed,
                     hipStream_t s, hipEvent_t *ev)
{
    return mixed ? launch_sw_t<float>(in, out, w, t, s, ev) : launch_sw_t<double>(in, out, w, t, s, ev);
}

}  // namespace rrtmg
