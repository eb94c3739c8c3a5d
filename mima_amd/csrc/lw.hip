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
// FP32 runs the same body in lw_band_lds_kernel: persistent 16-wave
// workgroups that copy one band's k-tables (at most 128 KB) into LDS, so the
// ~30 k-table gathers per layer take LDS latency rather than L2 latency
// (64-lane gathers across four columns miss the 32 KB L1); a 16-wave
// workgroup also holds the kernel at 128 VGPRs, four waves per SIMD.  FP64
// (whose bands 3-5 exceed 160 KB of LDS) and RRTMG_HIP_LW_LDS=0 use
// lw_band_kernel, one wave per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "lw_optics.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// prep: FP64 (rrtmg_lw_rad.nomcica.f90:729-852 and rrtmg_lw_setcoef.f90:150-413),
// one thread per (layer, column): the layer's coefficients in the band
// kernel's precision, its integer table indices and the Planck index of the
// level above it (and of the surface level)
template <typename Real>
__global__ __launch_bounds__(256) void lw_prep_kernel(ColumnInputs in, LwWork w, const LwConst *__restrict__ cst)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nlay * ncol, iblk = (size_t)blockIdx.x * blockDim.x;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    // this thread's coefficient record, staged for store_records
    __shared__ __align__(16) Real srec[256 * LWF_STRIDE];
    Real *rec = srec + threadIdx.x * LWF_STRIDE;
    Real *tlevf = (Real *)w.tlev;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    // level 0 Planck index (setcoef: indlev0, t0frac)
    if (i < n && lay == 0) {
        double tz0 = in.tlev[col];
        int i0 = clampi((int)(tz0 - 159.0), 1, 180);
        w.ilev[col] = i0;
        tlevf[col] = (Real)(tz0 - 159.0 - (double)i0);
    }
    if (i < n) {
        const size_t ix = (size_t)lay * ncol + col;
        double pavel = in.play[ix], tavel = in.tlay[ix];
        double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double h2o = in.h2o[ix];
        // MiMA's driver passes specific humidity and an O3 mass mixing
        // ratio (rrtmg_lw_rad.nomcica.f90:775-778); the McICA driver's inatm
        // takes volume mixing ratios (rrtmg_lw_rad.f90:778, 784)
        double wkl1 = in.vmr ? h2o : (h2o / (1.0 - h2o)) * amdw;
        double wkl2 = in.co2[ix], wkl3 = in.vmr ? in.o3[ix] : in.o3[ix] * amdo, wkl4 = in.n2o[ix], wkl6 = in.ch4[ix],
               wkl7 = in.o2[ix];
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
        rec[LWF_FAC10] = (Real)(compfp * ft);
        rec[LWF_FAC00] = (Real)(compfp * (1.0 - ft));
        rec[LWF_FAC11] = (Real)(fp * ft1);
        rec[LWF_FAC01] = (Real)(fp * (1.0 - ft1));
        rec[LWF_COLH2O] = (Real)colh2o;
        rec[LWF_COLCO2] = (Real)colco2;
        rec[LWF_COLO3] = (Real)colo3;
        rec[LWF_COLN2O] = (Real)coln2o;
        rec[LWF_COLCO] = (Real)colco;
        rec[LWF_COLCH4] = (Real)colch4;
        rec[LWF_COLO2] = (Real)colo2;
        rec[LWF_COLBRD] = (Real)colbrd;
        rec[LWF_SELFFAC] = (Real)(colh2o * selffac);
        rec[LWF_SELFFRAC] = (Real)selffrac;
        rec[LWF_FORFAC] = (Real)(colh2o * forfac);
        rec[LWF_FORFRAC] = (Real)forfrac;
        rec[LWF_MINORFRAC] = (Real)minorfrac;
        rec[LWF_SCALEMINOR] = (Real)scaleminor;
        rec[LWF_SCALEMINORN2] = (Real)scaleminorn2;
        rec[LWF_COLDRY] = (Real)coldry;
        rec[LWF_PAVEL] = (Real)pavel;
        rec[LWF_WX1] = (Real)wx1;
        rec[LWF_WX2] = (Real)wx2;
        rec[LWF_WX3] = (Real)wx3;
        rec[LWF_WX4] = (Real)wx4;
        rec[LWF_TLAYFRAC] = (Real)tlayfrac;
        w.ipk[ix] = pack_lw(jp, jt, jt1, indself, indfor, indminor, indlay);
        for (int f = LWF_COUNT; f < LWF_STRIDE; ++f) rec[f] = (Real)0;
    }
    const int nrec = (int)(n - iblk < blockDim.x ? n - iblk : blockDim.x);
    store_records<Real, LWF_STRIDE>((Real *)w.coef + iblk * LWF_STRIDE, srec, nrec);
}

// per-column part of the prep (one thread per column): laytrop and the
// precipitable water from the column's water and dry-air totals, summed in
// the reference's order (rrtmg_lw_rad.nomcica.f90:845-852)
__global__ __launch_bounds__(256) void lw_column_kernel(ColumnInputs in, LwWork w)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = in.ncol, nlay = in.nlay;
    if (col >= ncol) return;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, avogad = 6.02214199e+23, grav = 9.8066;
    double amttl = 0.0, wvttl = 0.0;
    int laytrop = 0;
    for (int lay = 0; lay < nlay; ++lay) {
        const size_t ix = (size_t)lay * ncol + col;
        double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double h2o = in.h2o[ix];
        double wkl1 = in.vmr ? h2o : (h2o / (1.0 - h2o)) * amdw;
        double amm = (1.0 - wkl1) * amd + wkl1 * amw;
        double coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1.0 + wkl1));
        wkl1 = coldry * wkl1;
        amttl = amttl + coldry + wkl1;
        wvttl = wvttl + wkl1;
        if (log(in.play[ix]) > 4.56) ++laytrop;
    }
    double wvsh = (amw * wvttl) / (amd * amttl);
    double pwvcm = wvsh * (1.e3 * in.plev[col]) / (1.e2 * grav);
    w.laytrop[col] = laytrop;
    w.colscal[col] = pwvcm;
    w.colscal[(size_t)ncol + col] = in.tsfc[col];
}

// One band of one work unit (rrtmg_lw_rtrnmr.f90:483-716, clear sky).
// ktab: the k-tables (B's offsets index it; the packed global buffer, or the
// band's copy in LDS with rebased offsets); tab: the packed global buffer
// (transmittance and Planck tables).  s_atr, s_bbu: transmittance and upward
// source per layer, [nlay][64] each (LDS or this wave's global scratch row);
// s_red: 64 LDS entries for the FP64 g-sum.
template <int BAND, typename Real>
__device__ __forceinline__ void lw_band_body(const ColumnInputs &in, const LwWork &w, const LwBand &B, const Real *__restrict__ ktab,
                             const Real *__restrict__ tab, int exp_off, int tfn_off, int plk_off,
                             const LwConst *__restrict__ cst, WaveUnit u, int tile, int lane,
                             Real *__restrict__ s_atr, Real *__restrict__ s_bbu, Real *s_red)
{
    constexpr int NG = lw_ng(BAND);
    const int ncol = in.ncol, nlay = in.nlay;
    const int cl = lane / NG, g = lane - cl * NG;
    const GSum<Real, NG> gsum(lane, g);
    const int col0 = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col0 < ncol;
    const int col = active ? col0 : 0;
    const Real *coef = (const Real *)w.coef;
    const Real *tlevf = (const Real *)w.tlev;
    const Real *exp_tbl = exp_off >= 0 ? tab + exp_off : nullptr;
    const Real *tfn_tbl = tfn_off >= 0 ? tab + tfn_off : nullptr;
    const Real *plk = tab + plk_off + (BAND - 1) * 181;  // totplnk(:, BAND)
    LTab<Real> T{ktab, NG, g};
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
        int i = at32(w.ilev, (unsigned)(lev * ncol + col));
        Real f = at32(tlevf, (unsigned)(lev * ncol + col));
        return at32(plk, (unsigned)(i - 1)) + f * (at32(plk, (unsigned)i) - at32(plk, (unsigned)(i - 1)));
    };
    const double *tauaer = in.tauaer;

    // downward sweep
    Real radld = (Real)0, fracs1 = (Real)0;
    Real plev_up = planklev(nlay);
    if (active && g == 0) at32(bf_dn, (unsigned)(nlay * ncol + col)) = (Real)0;  // drad(nlayers) stays 0
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
        Real atrans, bbd, bbugas;
        if (!exp_tbl) {  // smooth mode: the tabulated functions at every depth
            const LwRT<Real> rt = lw_rt_exact(odepth, false);
            atrans = rt.atrans;
            bbd = plfrac * (blay + rt.tf * dplankdn);
            bbugas = plfrac * (blay + rt.tf * dplankup);
        } else if (odepth <= (Real)0.06) {
            atrans = odepth - (Real)0.5 * odepth * odepth;
            odepth = (Real)0.166667 * odepth;
            bbd = plfrac * (blay + dplankdn * odepth);
            bbugas = plfrac * (blay + dplankup * odepth);
        } else {
            const int itr = tbl_index(odepth);
            atrans = (Real)1 - at32(exp_tbl, (unsigned)itr);
            const Real tausfac = at32(tfn_tbl, (unsigned)itr);
            bbd = plfrac * (blay + tausfac * dplankdn);
            bbugas = plfrac * (blay + tausfac * dplankup);
        }
        radld = radld + (bbd - radld) * atrans;
        at32(s_atr, (unsigned)((lev - 1) * 64 + lane)) = atrans;
        at32(s_bbu, (unsigned)((lev - 1) * 64 + lane)) = bbugas;
        Real s = gsum.sum(active ? radld : (Real)0, s_red);
        if (active && g == 0) at32(bf_dn, (unsigned)((lev - 1) * ncol + col)) = s;
        plev_up = plev_dn;
        if (lev == 1) fracs1 = plfrac;
    }
    // surface emission and reflection, upward sweep
    Real radlu = fracs1 * plankbnd + ((Real)1 - semiss) * radld;
    {
        Real s = gsum.sum(active ? radlu : (Real)0, s_red);
        if (active && g == 0) at32(bf_up, (unsigned)col) = s;
    }
    // the per-layer states are read four levels ahead (rotating registers),
    // so a level's load has three levels of work to complete behind
    auto at = [&](const Real *p, int i) { return i < nlay ? at32(p, (unsigned)(i * 64 + lane)) : (Real)0; };
    Real b0 = at(s_bbu, 0), b1 = at(s_bbu, 1), b2 = at(s_bbu, 2), b3 = at(s_bbu, 3);
    Real a0 = at(s_atr, 0), a1 = at(s_atr, 1), a2 = at(s_atr, 2), a3 = at(s_atr, 3);
    for (int lev = 1; lev <= nlay; ++lev) {
        const Real bbugas = b0, atrans = a0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = at(s_bbu, lev + 3);
        a0 = a1;
        a1 = a2;
        a2 = a3;
        a3 = at(s_atr, lev + 3);
        radlu = radlu + (bbugas - radlu) * atrans;
        Real s = gsum.sum(active ? radlu : (Real)0, s_red);
        if (active && g == 0) at32(bf_up, (unsigned)(lev * ncol + col)) = s;
    }
}

template <typename Real>
__global__ __launch_bounds__(64) void lw_band_kernel(ColumnInputs in, LwWork w, const LwBand *__restrict__ bands,
                                                     int nwaves, int blk0, int nblocks, int exp_off, int tfn_off,
                                                     int plk_off, const Real *__restrict__ tab,
                                                     const LwConst *__restrict__ cst, const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const LwBand B = bands[u.band];
    Real *s_bbu = (Real *)w.dstate + (size_t)blk * in.nlay * 64;
    Real *s_atr = smem, *s_red = smem + (size_t)in.nlay * 64;
#define BODY(b)                                                                                               \
    case b - 1:                                                                                               \
        lw_band_body<b, Real>(in, w, B, tab, tab, exp_off, tfn_off, plk_off, cst, u, tile, (int)threadIdx.x, s_atr, \
                              s_bbu, s_red);                                                                  \
        break;
    switch (u.band) {
        BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
        BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
    }
#undef BODY
}

// FP32 band kernel with the band's k-tables in LDS.  Persistent workgroups of
// LW_WGW waves claim work items from a global counter (next_item);
// item = (group of LW_WGW tiles, unit r of a tile's unit list) covers the
// group's LW_WGW x upt units of the band of unit r, one per wave.  Before an
// item of another band the workgroup copies that band's k-tables (at most
// plan.lds_bytes) into LDS; gathers from them then take LDS latency instead
// of the L2 latency of 64-lane gathers that miss the 32 KB L1.  The
// transmittance and upward source of each wave go to its global scratch row
// (2 x nlay x 64; rows = waves of the grid, so they stay in the last-level
// cache).
template <typename Real>
__global__ __launch_bounds__(64 * LW_WGW) void lw_band_lds_kernel(ColumnInputs in, LwWork w,
                                                                  const LwBand *__restrict__ bands, LwLdsPlan plan,
                                                                  int nwaves, int nitems, int exp_off, int tfn_off,
                                                                  int plk_off, const Real *__restrict__ tab,
                                                                  const LwConst *__restrict__ cst,
                                                                  const WaveUnit *__restrict__ units)
{
    static_assert(sizeof(Real) == 4, "FP32 only: the FP64 g-sum synchronises the workgroup");
    extern __shared__ unsigned char smem_raw[];
    Real *ktab = (Real *)smem_raw;
    const int nlay = in.nlay;
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    Real *s_atr = (Real *)w.dstate + ((size_t)blockIdx.x * LW_WGW + wave) * 2 * nlay * 64;
    Real *s_bbu = s_atr + (size_t)nlay * 64;
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
        lw_band_body<bb, Real>(in, w, B, ktab, tab, exp_off, tfn_off, plk_off, cst, u, tile, lane, s_atr, s_bbu, \
                               nullptr);                                                                      \
        break;
        switch (b) {
            BODY(1) BODY(2) BODY(3) BODY(4) BODY(5) BODY(6) BODY(7) BODY(8)
            BODY(9) BODY(10) BODY(11) BODY(12) BODY(13) BODY(14) BODY(15) BODY(16)
        }
#undef BODY
    }
}

// finish: band sums, fluxfac, net flux and heating (rrtmg_lw_rtrnmr.f90:719-777),
// one thread per (layer, column): the fluxes of the layer's lower level (and
// of the top level) and its heating rate from the net fluxes of both levels
template <typename Real>
__global__ __launch_bounds__(256) void lw_finish_kernel(ColumnInputs in, LwWork w, LwOutputs out,
                                                        const LwBand *__restrict__ bands, double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    const Real *bflux = (const Real *)w.bflux;
    const size_t bstride = (size_t)(nlay + 1) * ncol;
    const double fluxfac = 2.0 * asin(1.0) * 2.e4;
    double fnet[2];
    for (int j = 0; j < 2; ++j) {
        const int k = lay + j;
        double up = 0.0, dn = 0.0;
        for (int b = 0; b < 16; ++b) {
            double dw = bands[b].delwave;
            up = up + ((double)bflux[(size_t)b * 2 * bstride + (size_t)k * ncol + col] * 0.5) * dw;
            dn = dn + ((double)bflux[(size_t)b * 2 * bstride + bstride + (size_t)k * ncol + col] * 0.5) * dw;
        }
        up = up * fluxfac;
        dn = dn * fluxfac;
        if (j == 0 || lay == nlay - 1) {
            size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = up;
            out.dflx[ix] = dn;
            out.uflxc[ix] = up;
            out.dflxc[ix] = dn;
        }
        fnet[j] = up - dn;
    }
    size_t il = (size_t)lay * ncol + col;
    double h = heatfac * (fnet[0] - fnet[1]) / (in.plev[il] - in.plev[il + ncol]);
    out.hr[il] = h;
    out.hrc[il] = h;
}

template <typename Real>
static hipError_t launch_lw_t(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t,
                              double heatfac, bool smooth, hipStream_t s, hipEvent_t *ev)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(lw_prep_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, t.cst);
    hipLaunchKernelGGL(lw_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = lw_fwd_lds_bytes(in.nlay, sizeof(Real));
    if (ev) (void)hipEventRecord(ev[1], s);
    const int nblocks = ntiles * t.desc.nwaves;
    if constexpr (sizeof(Real) == 4) {
        if (t.lds && !smooth) {
            const int nitems = (ntiles + LW_WGW - 1) / LW_WGW * t.desc.nwaves;
            hipLaunchKernelGGL(reset_items_kernel, dim3(1), dim3(1), 0, s, w.items);
            hipLaunchKernelGGL(lw_band_lds_kernel<Real>, dim3(lw_lds_grid(t.plan, in.ncol, t.desc.nwaves)),
                               dim3(64 * LW_WGW), (size_t)t.plan.lds_bytes, s, in, w, t.bands_lds, t.plan,
                               t.desc.nwaves, nitems, t.desc.exp_tbl, t.desc.tfn_tbl, t.desc.totplnk,
                               (const Real *)t.tab, t.cst, t.units);
            if (ev) (void)hipEventRecord(ev[2], s);
            hipLaunchKernelGGL(lw_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out, t.bands, heatfac);
            if (ev) (void)hipEventRecord(ev[3], s);
            return hipGetLastError();
        }
    }
    // launches of at most w.chunk wave units, each reusing the scratch rows
    for (int b0 = 0; b0 < nblocks; b0 += w.chunk) {
        const int nb_c = std::min(w.chunk, nblocks - b0);
        hipLaunchKernelGGL(lw_band_kernel<Real>, dim3(xcd_grid(nb_c)), dim3(64), shmem, s, in, w, t.bands,
                           t.desc.nwaves, b0, nb_c, smooth ? -1 : t.desc.exp_tbl, smooth ? -1 : t.desc.tfn_tbl,
                           t.desc.totplnk, (const Real *)t.tab, t.cst, t.units);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(lw_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out, t.bands, heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_lw_prep(const ColumnInputs &in, const LwWork &w, const LwDevTables &t, int mode, hipStream_t s)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (mode == 1)
        hipLaunchKernelGGL(lw_prep_kernel<float>, dim3(nbl), dim3(256), 0, s, in, w, t.cst);
    else
        hipLaunchKernelGGL(lw_prep_kernel<double>, dim3(nbl), dim3(256), 0, s, in, w, t.cst);
    hipLaunchKernelGGL(lw_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
    return hipGetLastError();
}

hipError_t launch_lw_finish(const ColumnInputs &in, const LwOutputs &out, void *bflux, const LwDevTables &t,
                            int mode, hipStream_t s)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    LwWork w;
    w.bflux = bflux;
    if (mode == 1)
        hipLaunchKernelGGL(lw_finish_kernel<float>, dim3(nbl), dim3(256), 0, s, in, w, out, t.bands, t.heatfac);
    else
        hipLaunchKernelGGL(lw_finish_kernel<double>, dim3(nbl), dim3(256), 0, s, in, w, out, t.bands, t.heatfac);
    return hipGetLastError();
}

hipError_t launch_lw(const ColumnInputs &in, const LwOutputs &out, const LwWork &w, const LwDevTables &t, int mode,
                     hipStream_t s, hipEvent_t *ev)
{
    return mode == 1 ? launch_lw_t<float>(in, out, w, t, t.heatfac, false, s, ev)
                     : launch_lw_t<double>(in, out, w, t, t.heatfac, mode == 2, s, ev);
}

}  // namespace rrtmg
