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
#include "lw_vjp_body.hpp"
#include "rrtmg_device.hpp"

namespace rrtmg {

using D3 = Dual<double, 3>;


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
    const LwUnitArgs ua = lw_unit_args(in, w, exp_off, tfn_off, plk_off, tab, cst);
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
        lw_vjp_unit<bb>(ua, B, (lptr<const float>)ktab, u, tile, lane, (gptr<float>)s1, (gptr<float>)s3);    \
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
static void launch_lw_vjp_prep_t(const ColumnInputs &in, const VjpCot &cot, const LwVjpWork &w,
                                 const LwDevTables &t, hipStream_t s)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    hipLaunchKernelGGL(lw_vjp_prep_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, cot, w, t.cst, t.heatfac);
    hipLaunchKernelGGL(lw_vjp_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
}

template <typename Real>
static void launch_lw_vjp_band_t(const ColumnInputs &in, const LwVjpWork &w, const LwDevTables &t, bool smooth,
                                 hipStream_t s)
{
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = lw_vjp_lds_bytes(in.nlay, sizeof(Real));
    const int nblocks = ntiles * t.desc.nwaves;
    if constexpr (sizeof(Real) == 4) {
        if (t.lds && !smooth) {
            const int nitems = (ntiles + LWV_WGW - 1) / LWV_WGW * t.desc.nwaves;
            hipLaunchKernelGGL(reset_items_kernel, dim3(1), dim3(1), 0, s, w.items);
            hipLaunchKernelGGL(lw_vjp_band_lds_kernel<Real>, dim3(lw_vjp_lds_grid(t.plan, in.ncol, t.desc.nwaves)),
                               dim3(64 * LWV_WGW), (size_t)t.plan.lds_bytes, s, in, w, t.bands_lds, t.plan,
                               t.desc.nwaves, nitems, t.desc.exp_tbl, t.desc.tfn_tbl, t.desc.totplnk,
                               (const Real *)t.tab, t.cst, t.units);
            return;
        }
    }
    // launches of at most w.chunk wave units, each reusing the scratch rows
    for (int b0 = 0; b0 < nblocks; b0 += w.chunk) {
        const int nb_c = std::min(w.chunk, nblocks - b0);
        hipLaunchKernelGGL(lw_vjp_band_kernel<Real>, dim3(xcd_grid(nb_c)), dim3(64), shmem, s, in, w, t.bands,
                           t.desc.nwaves, b0, nb_c, smooth ? -1 : t.desc.exp_tbl, smooth ? -1 : t.desc.tfn_tbl,
                           t.desc.totplnk, (const Real *)t.tab, t.cst, t.units);
    }
}

template <typename Real>
static hipError_t launch_lw_vjp_finish_t(const ColumnInputs &in, const LwGrad &out, const LwVjpWork &w,
                                         const LwDevTables &t, hipStream_t s, const LwOutputs *fwd, int mode)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    hipLaunchKernelGGL(lw_vjp_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out);
    if (fwd) {
        const hipError_t e = launch_lw_finish(in, *fwd, w.bflux, t, mode, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_lw_vjp(const ColumnInputs &in, const VjpCot &cot, const LwGrad &out, const LwVjpWork &w,
                         const LwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev, const LwOutputs *fwd)
{
    if (fwd && !w.bflux) return hipErrorInvalidValue;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipError_t e = launch_lw_vjp_prep(in, cot, w, t, mode, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    e = launch_lw_vjp_band(in, w, t, mode, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], s);
    e = launch_lw_vjp_finish(in, out, w, t, mode, s, fwd);
    if (ev) (void)hipEventRecord(ev[3], s);
    return e;
}

hipError_t launch_lw_vjp_prep(const ColumnInputs &in, const VjpCot &cot, const LwVjpWork &w, const LwDevTables &t,
                              int mode, hipStream_t s)
{
    if (mode == 1) launch_lw_vjp_prep_t<float>(in, cot, w, t, s);
    else launch_lw_vjp_prep_t<double>(in, cot, w, t, s);
    return hipGetLastError();
}

hipError_t launch_lw_vjp_band(const ColumnInputs &in, const LwVjpWork &w, const LwDevTables &t, int mode,
                              hipStream_t s)
{
    if (mode == 1) launch_lw_vjp_band_t<float>(in, w, t, false, s);
    else launch_lw_vjp_band_t<double>(in, w, t, mode == 2, s);
    return hipGetLastError();
}

hipError_t launch_lw_vjp_finish(const ColumnInputs &in, const LwGrad &out, const LwVjpWork &w, const LwDevTables &t,
                                int mode, hipStream_t s, const LwOutputs *fwd)
{
    if (fwd && !w.bflux) return hipErrorInvalidValue;
    return mode == 1 ? launch_lw_vjp_finish_t<float>(in, out, w, t, s, fwd, mode)
                     : launch_lw_vjp_finish_t<double>(in, out, w, t, s, fwd, mode);
}

}  // namespace rrtmg
