// Backward (vector-Jacobian product) of the clear-sky SW path on gfx950:
// given cotangents of swuflx, swdflx (levels) and swhr (layers), the
// gradients with respect to layer temperature, water vapour and ozone.  No
// counterpart in the reference; the forward being differentiated is rrtmg_sw
// (rrtmg_sw_rad.nomcica.f90:78) as the forward kernels implement it.
//
// Kernels:
//  * prep (one thread per (layer, column)): inatm_sw + setcoef_sw in FP64
//    dual numbers seeded with (tlay, h2o, o3) of each layer (a value record
//    and a record of the nonzero derivatives, SW_DMASK, per (layer, column))
//    and the level cotangents after folding the
//    heating-rate cotangent into the fluxes; then a per-column kernel for
//    laytrop and the solar-source reference layers.
//  * band (one wave per work unit, lane = (column, g) as in the forward):
//    pass 1, the adding method's bottom-up sweep (values; reflectances of the
//    atmosphere below each level kept in LDS); pass 2, top-down (values): the
//    downward sweep (its states kept in a per-wave global scratch row), the
//    level fluxes and their local adjoints, and the reverse of the bottom-up
//    sweep, whose adjoints of each layer's (r, rd, t, td, dbt) go to the
//    scratch row; pass 3, bottom-up: the reverse of the downward sweep adds
//    its own layer adjoints, and the sum is projected once onto the layer's
//    (T, q, O3) through its optics recomputed as dual numbers: gas/Rayleigh
//    optical depths in (T, q, O3), two-stream optics in those two optical
//    depths; plus the solar source's dependence on its reference layer.
//    The derivatives are those of the smooth functions behind the exp table
//    (exp(-x), floored at 1e-20) in every precision mode.  Per-lane gradients
//    are reduced over the band's g-points.
//  * finish (one thread per (layer, column)): sums over bands in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "rrtmg_device.hpp"
#include "sw_optics.hpp"
#include "sw_vjp_body.hpp"

namespace rrtmg {

using D3 = Dual<double, 3>;

// ---------------------------------------------------------------------------
// prep: the forward prep's arithmetic (sw.hip, rrtmg_sw_rad.nomcica.f90:933-1047,
// rrtmg_sw_setcoef.f90:129-284) on dual numbers, one thread per (layer,
// column): the layer's coefficients with their derivatives and the flux
// cotangents of its lower level (and of the top level)
template <typename Real>
__global__ __launch_bounds__(256) void sw_vjp_prep_kernel(ColumnInputs in, VjpCot cot, SwVjpWork w,
                                                          const double *__restrict__ preflog,
                                                          const double *__restrict__ tref, double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    Real *dco = (Real *)w.dcoef, *dde = (Real *)w.dder;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    // the value record and the derivative record of the (layer, column)
    auto put = [&](int F, size_t ix, const D3 &x) {
        dco[ix * SWF_STRIDE + F] = (Real)x.v;
        Real *d = dde + ix * SW_DSTRIDE + der_offset(SW_DMASK, F);
        const unsigned m = SW_DMASK[F];
        int k = 0;
        for (int j = 0; j < 3; ++j)
            if (m & (1u << j)) d[k++] = (Real)x.d[j];
    };
    {
        const size_t ix = (size_t)lay * ncol + col;
        const double pavel = in.play[ix];
        const D3 tavel = D3::seed(in.tlay[ix], 0);
        const D3 h2o = D3::seed(in.h2o[ix], 1);
        const D3 o3 = D3::seed(in.o3[ix], 2);
        const double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        const D3 wkl1 = (h2o / (1.0 - h2o)) * amdw;
        const D3 amm = (1.0 - wkl1) * amd + wkl1 * amw;
        const D3 coldry = ((pz0 - pz1) * 1.e3 * avogad) / (1.e2 * grav * amm * (1.0 + wkl1));
        const D3 colh2o = 1.e-20 * (coldry * wkl1);
        D3 colco2 = 1.e-20 * (coldry * in.co2[ix]);
        const D3 colo3 = 1.e-20 * (coldry * (o3 * amdo));
        D3 colch4 = 1.e-20 * (coldry * in.ch4[ix]);
        D3 colo2 = 1.e-20 * (coldry * in.o2[ix]);
        const D3 colmol = 1.e-20 * coldry + colh2o;
        if (colco2.v == 0.0) colco2 = 1.e-32 * coldry;
        if (colch4.v == 0.0) colch4 = 1.e-32 * coldry;
        if (colo2.v == 0.0) colo2 = 1.e-32 * coldry;

        const double plog = log(pavel);
        const int jp = clampi((int)(36.0 - 5 * (plog + 0.04)), 1, 58);
        const double fp = 5.0 * (preflog[jp - 1] - plog);
        const D3 dt0 = (tavel - tref[jp - 1]) / 15.0;
        const int jt = clampi((int)(3.0 + dt0.v), 1, 4);
        const D3 ft = dt0 - (double)(jt - 3);
        const D3 dt1 = (tavel - tref[jp]) / 15.0;
        const int jt1 = clampi((int)(3.0 + dt1.v), 1, 4);
        const D3 ft1 = dt1 - (double)(jt1 - 3);
        const D3 water = (coldry * wkl1) / coldry;
        const D3 scalefac = pavel * stpfac / tavel;
        D3 forfac, forfrac, selffac, selffrac;
        int indfor, indself;
        if (plog > 4.56) {
            forfac = scalefac / (1. + water);
            D3 factor = (332.0 - tavel) / 36.0;
            indfor = clampi((int)factor.v, 1, 2);
            forfrac = factor - (double)indfor;
            selffac = water * forfac;
            factor = (tavel - 188.0) / 7.2;
            indself = clampi((int)factor.v - 7, 1, 9);
            selffrac = factor - (double)(indself + 7);
        } else {
            forfac = scalefac / (1. + water);
            const D3 factor = (tavel - 188.0) / 36.0;
            indfor = 3;
            forfrac = factor - 1.0;
            selffac = D3(0.0);
            selffrac = D3(0.0);
            indself = 0;
        }
        const double compfp = 1.0 - fp;
        put(SWF_FAC10, ix, compfp * ft);
        put(SWF_FAC00, ix, compfp * (1.0 - ft));
        put(SWF_FAC11, ix, fp * ft1);
        put(SWF_FAC01, ix, fp * (1.0 - ft1));
        put(SWF_COLH2O, ix, colh2o);
        put(SWF_COLCO2, ix, colco2);
        put(SWF_COLO3, ix, colo3);
        put(SWF_COLCH4, ix, colch4);
        put(SWF_COLO2, ix, colo2);
        put(SWF_COLMOL, ix, colmol);
        put(SWF_SELFFAC, ix, selffac);
        put(SWF_SELFFRAC, ix, selffrac);
        put(SWF_FORFAC, ix, forfac);
        put(SWF_FORFRAC, ix, forfrac);
        w.ipk[ix] = pack_sw(jp, jt, jt1, indself, indfor);
    }
    // level cotangents: swhr(k) = (net(k+1) - net(k)) heatfac / (p(k) - p(k+1)),
    // net = dn - up, top layer zero (rrtmg_sw_rad.nomcica.f90:713-726); this
    // thread's level lay, and the top
    Real *cu = (Real *)w.cot, *cd = cu + (size_t)(nlay + 1) * ncol;
    for (int k = lay; k <= (lay == nlay - 1 ? nlay : lay); ++k) {
        const size_t ix = (size_t)k * ncol + col;
        double nbar = 0.0;
        if (cot.hr) {
            if (k >= 1 && k <= nlay - 1) nbar += cot.hr[ix - ncol] * heatfac / (in.plev[ix - ncol] - in.plev[ix]);
            if (k <= nlay - 2) nbar -= cot.hr[ix] * heatfac / (in.plev[ix] - in.plev[ix + ncol]);
        }
        cu[ix] = (Real)((cot.uflx ? cot.uflx[ix] : 0.0) - nbar);
        cd[ix] = (Real)((cot.dflx ? cot.dflx[ix] : 0.0) + nbar);
    }
}

// per-column part of the prep (one thread per column, after the per-layer
// part, whose packed indices hold jp): laytrop and the solar-source reference
// layers, as the forward prep
__global__ __launch_bounds__(256) void sw_vjp_column_kernel(ColumnInputs in, SwVjpWork w)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= in.ncol) return;
    sw_column_refs(in.play, w.ipk, in.ncol, in.nlay, col, w.laytrop + col, w.laysolfr);
}

template <typename Real>
__global__ __launch_bounds__(64) void sw_vjp_band_kernel(ColumnInputs in, SwVjpWork w,
                                                         const SwBand *__restrict__ bands, int nwaves, int blk0,
                                                         int nblocks, int exp_off, const Real *__restrict__ tab,
                                                         const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const SwBand B = bands[u.band];
    Real *row = (Real *)w.dstate + (size_t)blk * (sw_vjp_state_bytes(in.nlay, sizeof(Real)) / sizeof(Real));
#define BODY(jb)                                                                                      \
    case jb - 16:                                                                                     \
        sw_vjp_band_body<jb, Real>(in, w, B, u, tile, (int)threadIdx.x, row, exp_off, tab, smem);     \
        break;
    switch (u.band) { SW_BANDS(BODY) }
#undef BODY
}

// finish: band sums in FP64, one thread per (layer, column)
template <typename Real>
__global__ __launch_bounds__(256) void sw_vjp_finish_kernel(ColumnInputs in, SwVjpWork w, SwGrad out)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t n2 = (size_t)nlay * ncol;
    const size_t ix = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= n2) return;
    const int col = (int)(ix % ncol);
    const Real *gpart = (const Real *)w.gpart;
    double gt = 0.0, gq = 0.0, go = 0.0;
    if (in.coszen[col] >= 1.e-10)
        for (int b = 0; b < 14; ++b) {
            const Real *gp = gpart + (size_t)b * 3 * n2;
            gt += (double)gp[ix];
            gq += (double)gp[n2 + ix];
            go += (double)gp[2 * n2 + ix];
        }
    out.tlay[ix] = gt;
    out.h2o[ix] = gq;
    out.o3[ix] = go;
}

template <typename Real>
static void launch_sw_vjp_prep_t(const ColumnInputs &in, const VjpCot &cot, const SwVjpWork &w,
                                 const SwDevTables &t, hipStream_t s)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    hipLaunchKernelGGL(sw_vjp_prep_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, cot, w, t.preflog, t.tref,
                       t.heatfac);
    hipLaunchKernelGGL(sw_vjp_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
}

template <typename Real>
static void launch_sw_vjp_band_t(const ColumnInputs &in, const SwVjpWork &w, const SwDevTables &t, bool smooth,
                                 hipStream_t s)
{
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = sw_vjp_lds_bytes(in.nlay, sizeof(Real));
    const int nblocks = ntiles * t.desc.nwaves;
    // launches of at most w.chunk wave units, each reusing the scratch rows
    for (int b0 = 0; b0 < nblocks; b0 += w.chunk) {
        const int nb_c = std::min(w.chunk, nblocks - b0);
        hipLaunchKernelGGL(sw_vjp_band_kernel<Real>, dim3(xcd_grid(nb_c)), dim3(64), shmem, s, in, w, t.bands,
                           t.desc.nwaves, b0, nb_c, smooth ? -1 : t.desc.exp_tbl, (const Real *)t.tab, t.units);
    }
}

template <typename Real>
static hipError_t launch_sw_vjp_finish_t(const ColumnInputs &in, const SwGrad &out, const SwVjpWork &w,
                                         const SwDevTables &t, hipStream_t s, const SwOutputs *fwd, int mode)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    hipLaunchKernelGGL(sw_vjp_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out);
    if (fwd) {
        const hipError_t e = launch_sw_finish(in, *fwd, w.bflux, t, mode, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_sw_vjp(const ColumnInputs &in, const VjpCot &cot, const SwGrad &out, const SwVjpWork &w,
                         const SwDevTables &t, int mode, hipStream_t s, hipEvent_t *ev, const SwOutputs *fwd)
{
    if (fwd && !w.bflux) return hipErrorInvalidValue;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipError_t e = launch_sw_vjp_prep(in, cot, w, t, mode, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    e = launch_sw_vjp_band(in, w, t, mode, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], s);
    e = launch_sw_vjp_finish(in, out, w, t, mode, s, fwd);
    if (ev) (void)hipEventRecord(ev[3], s);
    return e;
}

hipError_t launch_sw_vjp_prep(const ColumnInputs &in, const VjpCot &cot, const SwVjpWork &w, const SwDevTables &t,
                              int mode, hipStream_t s)
{
    if (mode == 1) launch_sw_vjp_prep_t<float>(in, cot, w, t, s);
    else launch_sw_vjp_prep_t<double>(in, cot, w, t, s);
    return hipGetLastError();
}

hipError_t launch_sw_vjp_band(const ColumnInputs &in, const SwVjpWork &w, const SwDevTables &t, int mode,
                              hipStream_t s)
{
    if (mode == 1) launch_sw_vjp_band_t<float>(in, w, t, false, s);
    else launch_sw_vjp_band_t<double>(in, w, t, mode == 2, s);
    return hipGetLastError();
}

hipError_t launch_sw_vjp_finish(const ColumnInputs &in, const SwGrad &out, const SwVjpWork &w, const SwDevTables &t,
                                int mode, hipStream_t s, const SwOutputs *fwd)
{
    if (fwd && !w.bflux) return hipErrorInvalidValue;
    return mode == 1 ? launch_sw_vjp_finish_t<float>(in, out, w, t, s, fwd, mode)
                     : launch_sw_vjp_finish_t<double>(in, out, w, t, s, fwd, mode);
}

}  // namespace rrtmg
