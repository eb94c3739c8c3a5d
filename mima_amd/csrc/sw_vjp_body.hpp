// Backward band body of the clear-sky SW path (passes 1-3, sw_vjp.hip) and
// its loaders, shared by the SW backward band kernel and the fused LW+SW
// forward+backward kernel (fused.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "rrtmg_device.hpp"
#include "sw_optics.hpp"

namespace rrtmg {

// value (V = Real) or value and three derivatives (V = Dual<Real, 3>) of one
// field: v, the field in the cell's value record; d, its first derivative in
// the cell's derivative record.  MASK (bit 0: d/dT, 1: d/dq, 2: d/dO3, the
// field's SW_DMASK entry) names the derivatives the field can have; the
// others are known zeros (not stored; folded at compile time)
enum : unsigned { SDT = 1u, SDQ = 2u, SDO = 4u };
template <typename V, typename Real, unsigned MASK>
struct SwDLoad;
template <typename Real, unsigned MASK>
struct SwDLoad<Real, Real, MASK> {
    __device__ static Real get(const Real *__restrict__ v, const Real *__restrict__) { return v[0]; }
};
template <typename Real, unsigned MASK>
struct SwDLoad<Dual<Real, 3>, Real, MASK> {
    __device__ static Dual<Real, 3> get(const Real *__restrict__ v, const Real *__restrict__ d)
    {
        Dual<Real, 3> r;
        r.v = v[0];
        r.d[0] = (MASK & SDT) ? d[0] : (Real)0;
        r.d[1] = (MASK & SDQ) ? d[dmask_count(MASK & SDT)] : (Real)0;
        r.d[2] = (MASK & SDO) ? d[dmask_count(MASK & (SDT | SDQ))] : (Real)0;
        return r;
    }
};

// the layer's coefficients from the backward's records (values only when V
// is Real: the derivative record is not read)
template <typename V, typename Real>
__device__ __forceinline__ void sw_load_layer_d(SwLayer<V> &c, const Real *__restrict__ dco,
                                                const Real *__restrict__ dde, const int *__restrict__ ipk, int lay,
                                                int col, int ncol)
{
    const unsigned ix = (unsigned)(lay * ncol + col);
    const Real *vr = (const Real *)__builtin_assume_aligned(dco + (size_t)ix * SWF_STRIDE, 16);
    const Real *dr = (const Real *)__builtin_assume_aligned(dde + (size_t)ix * SW_DSTRIDE, 16);
#define LDD(f, F) c.f = SwDLoad<V, Real, SW_DMASK[F]>::get(vr + (F), dr + der_offset(SW_DMASK, F))
    LDD(fac00, SWF_FAC00); LDD(fac01, SWF_FAC01); LDD(fac10, SWF_FAC10); LDD(fac11, SWF_FAC11);
    LDD(colh2o, SWF_COLH2O); LDD(colco2, SWF_COLCO2); LDD(colo3, SWF_COLO3);
    LDD(colch4, SWF_COLCH4); LDD(colo2, SWF_COLO2); LDD(colmol, SWF_COLMOL);
    LDD(selffac, SWF_SELFFAC); LDD(selffrac, SWF_SELFFRAC); LDD(forfac, SWF_FORFAC);
    LDD(forfrac, SWF_FORFRAC);
#undef LDD
    const int w = at32(ipk, ix);
    c.jp = unpack_jp(w);
    c.jt = unpack_jt(w);
    c.jt1 = unpack_jt1(w);
    c.inds = unpack_inds(w);
    c.indf = unpack_indf(w);
}

// One band JB of one work unit.  lane: the thread's lane in its wave (0-63);
// row: the wave's row of the global scratch, sw_vjp_state_bytes; smem: the
// wave's LDS, sw_vjp_lds_bytes.  The band's g-point and species-parameter
// counts are compile-time constants, as in the forward band kernel.
template <int JB, typename Real>
__device__ __forceinline__ void sw_vjp_band_body(const ColumnInputs &in, const SwVjpWork &w, const SwBand &B0,
                                                 const WaveUnit u, int tile, int lane, Real *row, int exp_off,
                                                 const Real *__restrict__ tab, Real *smem)
{
    using V3 = Dual<Real, 3>;
    const int ncol = in.ncol, nlay = in.nlay;
    constexpr int jb = JB;
    constexpr int ng = sw_ng(JB);  // B.ng, checked at init
    SwBand B = B0;
    B.nspa = SW_NSPA[JB - 16];
    B.nspb = SW_NSPB[JB - 16];
    const int cl = lane / ng, g = lane - cl * ng;
    const int col0 = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col0 < ncol && in.coszen[min(col0, ncol - 1)] >= 1.e-10;
    // a wave whose columns are all at night writes nothing: the finish
    // kernels set night columns to zero without reading its outputs
    if (!__any(active)) return;
    const int col = active ? col0 : 0;
    const size_t L1 = (size_t)(nlay + 1) * 64;
    const size_t LA = (size_t)nlay * 64;
    Real *s_prup = smem;   // bottom-up sweep's prup, LDS
    Real *red = smem + L1; // [4][64]
    // this wave's row of the global scratch (LDS holds only prup, so that LDS
    // does not cap occupancy below the VGPRs): downward-sweep states, written
    // top-down in pass 2 and read bottom-up in pass 3;
    Real *__restrict__ d_ptdbt = row;
    Real *__restrict__ d_ztdn = d_ptdbt + L1, *__restrict__ d_prdnd = d_ptdbt + 2 * L1;
    // adjoints of each layer's (r, rd, t, td, dbt) from the reverse of the
    // bottom-up sweep (pass 2), [5][nlay][64]: pass 3 adds the reverse of the
    // downward sweep's and projects the sum once onto (T, q, O3);
    Real *__restrict__ d_adj = d_ptdbt + 3 * L1;
    // the bottom-up sweep's prupd (read one level ahead)
    Real *__restrict__ s_prupd = d_adj + 5 * LA;
    const Real *dco = (const Real *)w.dcoef, *dde = (const Real *)w.dder;
    const Real *exp_tbl = exp_off >= 0 ? tab + exp_off : nullptr;
    Tab<Real> T{tab, ng, g};
    const size_t n2 = (size_t)nlay * ncol, n3 = (size_t)(nlay + 1) * ncol;
    Real *gpart = (Real *)w.gpart + (size_t)u.band * 3 * n2;
    const Real *cu = (const Real *)w.cot, *cd = cu + n3;

    Real prmu0 = (Real)1, zincflx = (Real)0, albp = (Real)0, albd = (Real)0;
    int laytrop = 0, lsf = 1;
    if (active) {
        prmu0 = (Real)in.coszen[col];
        laytrop = w.laytrop[col];
        const bool nir = (jb <= 24) || jb == 29;
        albp = (Real)(nir ? in.aldir[col] : in.asdir[col]);
        albd = (Real)(nir ? in.aldif[col] : in.asdif[col]);
        lsf = w.laysolfr[(size_t)u.band * ncol + col];
        SwLayer<Real> c;
        sw_load_layer_d<Real, Real>(c, dco, dde, w.ipk, lsf - 1, col, ncol);
        const double sfz = sw_sfluxzen(jb, B, T, c, lsf <= laytrop);
        zincflx = (Real)(in.adjflux * sfz * (double)prmu0);
    }

    // pass 1: bottom-up sweep (values)
    Real prup_n = albp, prupd_n = albd;
    s_prup[(size_t)nlay * 64 + lane] = albp;
    s_prupd[(size_t)nlay * 64 + lane] = albd;
    for (int lay = 1; lay <= nlay; ++lay) {
        const int jk = nlay + 1 - lay;
        Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
        if (active) {
            SwLayer<Real> c;
            sw_load_layer_d<Real, Real>(c, dco, dde, w.ipk, lay - 1, col, ncol);
            Real taug, taur;
            sw_gas(jb, B, T, c, lay <= laytrop, taug, taur);
            sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
        }
        const Real zreflect = (Real)1 / ((Real)1 - prupd_n * rd);
        const Real prup = r + (td * ((t - dbt) * prupd_n + dbt * prup_n)) * zreflect;
        const Real prupd = rd + td * td * prupd_n * zreflect;
        s_prup[(size_t)(jk - 1) * 64 + lane] = prup;
        at32(s_prupd, (unsigned)((jk - 1) * 64 + lane)) = prupd;
        prup_n = prup;
        prupd_n = prupd;
    }

    // Layer lay's gas and Rayleigh optical depths as dual numbers in (T, q,
    // O3) of that layer, and its two-stream optics as dual numbers in those
    // two optical depths: the optics' adjoint is projected onto (taug, taur)
    // and then onto (T, q, O3), cheaper than carrying three directions
    // through reftra.
    using V2 = Dual<Real, 2>;
    struct LayerD {
        V2 r, rd, t, td, dbt;
        V3 tg, tr;
    };
    auto optics_d = [&](int lay, LayerD &o) {
        SwLayer<V3> c;
        sw_load_layer_d<V3, Real>(c, dco, dde, w.ipk, lay - 1, col, ncol);
        sw_gas<Real, V3>(jb, B, T, c, lay <= laytrop, o.tg, o.tr);
        // exp(-x) itself rather than the table: the backward differentiates
        // the smooth functions the table stands for (the table's neighbouring
        // entries are not consistent to the derivative's accuracy where
        // reftra's k mu0 = 1 quotient is nearly 0/0)
        sw_layer_optics<Real, V2>(V2::seed(o.tg.v, 0), V2::seed(o.tr.v, 1), prmu0, (const Real *)nullptr, o.r, o.rd,
                                  o.t, o.td, o.dbt);
    };
    auto inactive_d = [](LayerD &o) {
        o.r = V2((Real)0);
        o.rd = V2((Real)0);
        o.t = V2((Real)1);
        o.td = V2((Real)1);
        o.dbt = V2((Real)1);
        o.tg = V3((Real)0);
        o.tr = V3((Real)0);
    };
    auto contract = [](Real (&v)[3], Real ar, Real ard, Real at, Real atd, Real adbt, const LayerD &o) {
        Real a_tg = ar * o.r.d[0] + ard * o.rd.d[0] + at * o.t.d[0] + atd * o.td.d[0] + adbt * o.dbt.d[0];
        Real a_tr = ar * o.r.d[1] + ard * o.rd.d[1] + at * o.t.d[1] + atd * o.td.d[1] + adbt * o.dbt.d[1];
        for (int i = 0; i < 3; ++i) v[i] = a_tg * o.tg.d[i] + a_tr * o.tr.d[i];
    };

    // band fluxes of the recomputed forward, when asked for (as the forward
    // band kernel writes them; uniform branch)
    Real *bf_up = w.bflux ? (Real *)w.bflux + (size_t)u.band * 2 * n3 : nullptr;
    Real *bf_dn = bf_up ? bf_up + n3 : nullptr;

    // pass 2: top-down
    Real ztdn = (Real)1, prdnd = (Real)0, ptdbt = (Real)1;
    Real a_pu = (Real)0, a_pud = (Real)0, a_zinc = (Real)0;
    Real pud_n = at32(s_prupd, (unsigned)lane);
    for (int jk = 1; jk <= nlay + 1; ++jk) {
        Real c_pu = (Real)0, c_pud = (Real)0;
        const Real pu = s_prup[(size_t)(jk - 1) * 64 + lane], pud = pud_n;
        if (jk <= nlay) pud_n = at32(s_prupd, (unsigned)(jk * 64 + lane));
        if (jk > 1) {
            const int lay = nlay + 2 - jk;
            Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
            if (active) {
                SwLayer<Real> c;
                sw_load_layer_d<Real, Real>(c, dco, dde, w.ipk, lay - 1, col, ncol);
                Real taug, taur;
                sw_gas(jb, B, T, c, lay <= laytrop, taug, taur);
                sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
            }
            // reverse of the bottom-up step that formed prup/prupd(jk-1) from level jk
            const Real Z = (Real)1 / ((Real)1 - pud * rd);
            const Real A = (t - dbt) * pud + dbt * pu;
            const Real ar = a_pu;
            const Real atd = a_pu * A * Z + a_pud * (Real)2 * td * pud * Z;
            const Real at = a_pu * td * pud * Z;
            const Real adbt = a_pu * td * (pu - pud) * Z;
            const Real ard = a_pu * td * A * pud * Z * Z + a_pud * ((Real)1 + td * td * pud * pud * Z * Z);
            c_pu = a_pu * td * dbt * Z;
            c_pud = a_pu * (td * (t - dbt) * Z + td * A * rd * Z * Z) + a_pud * (td * td * Z + td * td * pud * rd * Z * Z);
            const size_t ka = (size_t)(lay - 1) * 64 + lane;
            d_adj[ka] = ar;
            d_adj[LA + ka] = ard;
            d_adj[2 * LA + ka] = at;
            d_adj[3 * LA + ka] = atd;
            d_adj[4 * LA + ka] = adbt;
            // downward step (values), as the forward
            const Real zreflect = (Real)1 / ((Real)1 - rd * prdnd);
            const Real ztdn_n = ptdbt * t + (td * ((ztdn - ptdbt) + ptdbt * r * prdnd)) * zreflect;
            const Real prdnd_n = rd + td * td * prdnd * zreflect;
            ptdbt = dbt * ptdbt;
            ztdn = ztdn_n;
            prdnd = prdnd_n;
        }
        d_ptdbt[(size_t)(jk - 1) * 64 + lane] = ptdbt;
        d_ztdn[(size_t)(jk - 1) * 64 + lane] = ztdn;
        d_prdnd[(size_t)(jk - 1) * 64 + lane] = prdnd;
        // level fluxes and their local adjoints with respect to prup/prupd(jk)
        const Real zr = (Real)1 / ((Real)1 - prdnd * pud);
        const Real pfu = (ptdbt * pu + (ztdn - ptdbt) * pud) * zr;
        const Real pfd = ptdbt + (ztdn - ptdbt + ptdbt * pu * prdnd) * zr;
        const int lev = nlay + 1 - jk;
        const Real gu = active ? cu[(size_t)lev * ncol + col] : (Real)0;
        const Real gd = active ? cd[(size_t)lev * ncol + col] : (Real)0;
        const Real a_pfu = zincflx * gu, a_pfd = zincflx * gd;
        a_zinc += gu * pfu + gd * pfd;
        if (bf_up) {
            const Real su = seg_sum(active ? zincflx * pfu : (Real)0, lane, g, ng, red);
            const Real sd = seg_sum(active ? zincflx * pfd : (Real)0, lane, g, ng, red);
            if (active && g == 0) {
                bf_up[(size_t)lev * ncol + col] = su;
                bf_dn[(size_t)lev * ncol + col] = sd;
            }
        }
        a_pu = c_pu + a_pfu * ptdbt * zr + a_pfd * ptdbt * prdnd * zr;
        a_pud = c_pud + a_pfu * ((ztdn - ptdbt) * zr + pfu * prdnd * zr) + a_pfd * (pfd - ptdbt) * prdnd * zr;
    }

    // the solar source's dependence on its reference layer
    const Real a_sfz = a_zinc * (Real)(in.adjflux * (double)prmu0);
    V3 sfz3(0);
    if (active) {
        SwLayer<V3> c;
        sw_load_layer_d<V3, Real>(c, dco, dde, w.ipk, lsf - 1, col, ncol);
        sfz3 = sw_sfluxzen_v<Real, V3>(jb, B, T, c);
    }

    // pass 3: bottom-up reverse of the downward sweep
    Real A_b = (Real)0, A_z = (Real)0, A_p = (Real)0;  // adjoints of ptdbt, ztdn, prdnd at level jk
    // downward-sweep states at level jk, carried from the previous iteration
    Real pb_c = d_ptdbt[(size_t)nlay * 64 + lane], pz_c = d_ztdn[(size_t)nlay * 64 + lane],
         pp_c = d_prdnd[(size_t)nlay * 64 + lane], pud_c = at32(s_prupd, (unsigned)(nlay * 64 + lane));
    for (int jk = nlay + 1; jk >= 2; --jk) {
        // level jk-1's states, loaded ahead of their use in the reverse step
        const size_t k0 = (size_t)(jk - 2) * 64 + lane;
        const Real pb = d_ptdbt[k0], pz = d_ztdn[k0], pp = d_prdnd[k0];
        const Real pud_next = at32(s_prupd, (unsigned)k0);
        {
            // local flux adjoints at level jk
            const size_t k = (size_t)(jk - 1) * 64 + lane;
            const Real pu = s_prup[k], pud = pud_c;
            const Real pb = pb_c, pz = pz_c, pp = pp_c;
            const Real zr = (Real)1 / ((Real)1 - pp * pud);
            const Real pfu = (pb * pu + (pz - pb) * pud) * zr;
            const Real pfd = pb + (pz - pb + pb * pu * pp) * zr;
            const int lev = nlay + 1 - jk;
            const Real gu = active ? cu[(size_t)lev * ncol + col] : (Real)0;
            const Real gd = active ? cd[(size_t)lev * ncol + col] : (Real)0;
            const Real a_pfu = zincflx * gu, a_pfd = zincflx * gd;
            A_b += a_pfu * (pu - pud) * zr + a_pfd * ((Real)1 + ((Real)-1 + pu * pp) * zr);
            A_z += a_pfu * pud * zr + a_pfd * zr;
            A_p += a_pfu * pfu * pud * zr + a_pfd * (pb * pu * zr + (pfd - pb) * pud * zr);
        }
        const int lay = nlay + 2 - jk;
        LayerD o;
        if (active) optics_d(lay, o);
        else inactive_d(o);
        const Real r = o.r.v, rd = o.rd.v, t = o.t.v, td = o.td.v, dbt = o.dbt.v;
        pb_c = pb;
        pz_c = pz;
        pp_c = pp;
        pud_c = pud_next;
        const Real Zr = (Real)1 / ((Real)1 - rd * pp);
        const Real Bz = (pz - pb) + pb * r * pp;
        const Real at = A_z * pb;
        const Real atd = A_z * Bz * Zr + A_p * (Real)2 * td * pp * Zr;
        const Real ar = A_z * td * pb * pp * Zr;
        const Real ard = A_z * td * Bz * pp * Zr * Zr + A_p * ((Real)1 + td * td * pp * pp * Zr * Zr);
        const Real adbt = A_b * pb;
        const size_t ka = (size_t)(lay - 1) * 64 + lane;
        const Real ar2 = d_adj[ka], ard2 = d_adj[LA + ka], at2 = d_adj[2 * LA + ka], atd2 = d_adj[3 * LA + ka],
                   adbt2 = d_adj[4 * LA + ka];
        const Real nA_b = A_z * (t + td * ((Real)-1 + r * pp) * Zr) + A_b * dbt;
        const Real nA_z = A_z * td * Zr;
        const Real nA_p = A_z * (td * pb * r * Zr + td * Bz * rd * Zr * Zr) + A_p * (td * td * Zr + td * td * pp * rd * Zr * Zr);
        A_b = nA_b;
        A_z = nA_z;
        A_p = nA_p;
        Real v[3];
        contract(v, ar + ar2, ard + ard2, at + at2, atd + atd2, adbt + adbt2, o);
        if (lay == lsf)
            for (int i = 0; i < 3; ++i) v[i] += a_sfz * sfz3.d[i];
        if (!active) v[0] = v[1] = v[2] = (Real)0;
        seg_sum_n<3>(v, lane, g, ng, red);
        if (active && g == 0) {
            const size_t ix = (size_t)(lay - 1) * ncol + col;
            gpart[ix] = v[0];
            gpart[n2 + ix] = v[1];
            gpart[2 * n2 + ix] = v[2];
        }
    }
}

// what the SW backward body reads and writes, for the call below
template <typename Real>
struct SwUnitArgs {
    int ncol, nlay, exp_off;
    double adjflux;
    gptr<const double> asdir, asdif, aldir, aldif, coszen;
    gptr<const Real> dcoef, dder, cot, tab;
    gptr<const int> ipk, laytrop, laysolfr;
    gptr<Real> gpart, bflux;  // bflux NULL: no forward fluxes
};
template <typename Real>
__device__ __forceinline__ SwUnitArgs<Real> sw_unit_args(const ColumnInputs &in, const SwVjpWork &w, int exp_off,
                                                          const Real *tab)
{
    SwUnitArgs<Real> a;
    a.ncol = in.ncol;
    a.nlay = in.nlay;
    a.exp_off = exp_off;
    a.adjflux = in.adjflux;
    a.asdir = (gptr<const double>)in.asdir;
    a.asdif = (gptr<const double>)in.asdif;
    a.aldir = (gptr<const double>)in.aldir;
    a.aldif = (gptr<const double>)in.aldif;
    a.coszen = (gptr<const double>)in.coszen;
    a.dcoef = (gptr<const Real>)w.dcoef;
    a.dder = (gptr<const Real>)w.dder;
    a.cot = (gptr<const Real>)w.cot;
    a.tab = (gptr<const Real>)tab;
    a.ipk = (gptr<const int>)w.ipk;
    a.laytrop = (gptr<const int>)w.laytrop;
    a.laysolfr = (gptr<const int>)w.laysolfr;
    a.gpart = (gptr<Real>)w.gpart;
    a.bflux = (gptr<Real>)w.bflux;
    return a;
}

// The body as a function of its own, for the fused kernel, where the SW
// body's register allocation then does not weigh on the LW body's.  (The SW
// backward band kernel inlines the body: as a call it ran 9% slower at
// 8192 x 40.  The two compilations may round differently.)  row: the wave's
// scratch row; smem: its LDS.
template <int JB, typename Real>
__device__ __noinline__ void sw_vjp_unit(const SwUnitArgs<Real> &a, const SwBand &B, const WaveUnit u, int tile,
                                         int lane, gptr<Real> row, lptr<Real> smem)
{
    ColumnInputs in{};
    in.ncol = a.ncol;
    in.nlay = a.nlay;
    in.adjflux = a.adjflux;
    in.asdir = (const double *)a.asdir;
    in.asdif = (const double *)a.asdif;
    in.aldir = (const double *)a.aldir;
    in.aldif = (const double *)a.aldif;
    in.coszen = (const double *)a.coszen;
    SwVjpWork w;
    w.dcoef = (void *)(const Real *)a.dcoef;
    w.dder = (void *)(const Real *)a.dder;
    w.cot = (void *)(const Real *)a.cot;
    w.ipk = (int *)(const int *)a.ipk;
    w.laytrop = (int *)(const int *)a.laytrop;
    w.laysolfr = (int *)(const int *)a.laysolfr;
    w.gpart = (void *)(Real *)a.gpart;
    w.bflux = (void *)(Real *)a.bflux;
    sw_vjp_band_body<JB, Real>(in, w, B, u, tile, lane, (Real *)row, a.exp_off, (const Real *)a.tab, (Real *)smem);
}

}  // namespace rrtmg
