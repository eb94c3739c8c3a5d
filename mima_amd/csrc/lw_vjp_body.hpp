// Backward band body of the clear-sky LW path (passes A-D, lw_vjp.hip) and
// its loaders, shared by the LW backward band kernels and the fused
// LW+SW forward+backward kernel (fused.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lw_optics.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

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

// what the LW backward body reads and writes, for the call below
struct LwUnitArgs {
    int ncol, nlay, exp_off, tfn_off, plk_off;
    gptr<const double> emis, tauaer;  // tauaer NULL: no aerosol
    gptr<const double> colscal;
    gptr<const float> dcoef, dder, tlev, cot, tab;
    gptr<const int> ipk, ilev, laytrop;
    gptr<const LwConst> cst;
    gptr<float> gpart, glev, gsc, bflux;  // bflux NULL: no forward fluxes
};
__device__ __forceinline__ LwUnitArgs lw_unit_args(const ColumnInputs &in, const LwVjpWork &w, int exp_off,
                                                   int tfn_off, int plk_off, const float *tab, const LwConst *cst)
{
    LwUnitArgs a;
    a.ncol = in.ncol;
    a.nlay = in.nlay;
    a.exp_off = exp_off;
    a.tfn_off = tfn_off;
    a.plk_off = plk_off;
    a.emis = (gptr<const double>)in.emis;
    a.tauaer = (gptr<const double>)in.tauaer;
    a.colscal = (gptr<const double>)w.colscal;
    a.dcoef = (gptr<const float>)w.dcoef;
    a.dder = (gptr<const float>)w.dder;
    a.tlev = (gptr<const float>)w.tlev;
    a.cot = (gptr<const float>)w.cot;
    a.tab = (gptr<const float>)tab;
    a.ipk = (gptr<const int>)w.ipk;
    a.ilev = (gptr<const int>)w.ilev;
    a.laytrop = (gptr<const int>)w.laytrop;
    a.cst = (gptr<const LwConst>)cst;
    a.gpart = (gptr<float>)w.gpart;
    a.glev = (gptr<float>)w.glev;
    a.gsc = (gptr<float>)w.gsc;
    a.bflux = (gptr<float>)w.bflux;
    return a;
}

// The FP32 body with its k-tables in LDS as a function of its own, called by
// the LW LDS-table backward kernel and by the fused kernel (the same compiled
// code in both).  s1, s3: the wave's scratch rows.
template <int BAND>
__device__ __noinline__ void lw_vjp_unit(const LwUnitArgs &a, const LwBand &B, lptr<const float> ktab, WaveUnit u,
                                         int tile, int lane, gptr<float> s1, gptr<float> s3)
{
    ColumnInputs in{};
    in.ncol = a.ncol;
    in.nlay = a.nlay;
    in.emis = (const double *)a.emis;
    in.tauaer = (const double *)a.tauaer;
    LwVjpWork w;
    w.colscal = (double *)(const double *)a.colscal;
    w.dcoef = (void *)(const float *)a.dcoef;
    w.dder = (void *)(const float *)a.dder;
    w.tlev = (void *)(const float *)a.tlev;
    w.cot = (void *)(const float *)a.cot;
    w.ipk = (int *)(const int *)a.ipk;
    w.ilev = (int *)(const int *)a.ilev;
    w.laytrop = (int *)(const int *)a.laytrop;
    w.gpart = (void *)(float *)a.gpart;
    w.glev = (void *)(float *)a.glev;
    w.gsc = (void *)(float *)a.gsc;
    w.bflux = (void *)(float *)a.bflux;
    lw_vjp_band_body<BAND, float>(in, w, B, (const float *)ktab, (const float *)a.tab, a.exp_off, a.tfn_off,
                                  a.plk_off, (const LwConst *)a.cst, u, tile, lane, (float *)s1, (float *)s3, nullptr);
}

}  // namespace rrtmg
