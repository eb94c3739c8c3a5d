// Shortwave RRTMG on gfx950: prep (inatm_sw + setcoef_sw), band kernel
// (taumol_sw + spcvrt_sw + reftra_sw + vrtqdr_sw per g-point) and finish
// (band sums, net flux, heating).  Reference: rrtmg_sw/gcm_model/src/ of
// MiMA's RRTMG_SW (file:line cited per function).
//
// Band kernel layout: one wave per work unit (band b, up to 64/ng columns of a
// tile); lane = (column, g); each band is a compile-time instantiation of the
// band body (straight-line band code, registers sized per band).  Each lane
// runs the adding method for its
// g-point: a bottom-up sweep keeps prup/prupd of every level (in a per-wave
// global scratch row), a
// top-down sweep recomputes the layer optics, carries ztdn/prdnd/ptdbt and
// produces level fluxes, which are summed over the band's g-points of a
// column (segmented sum over ng adjacent lanes) into per-band fluxes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "rrtmg_device.hpp"
#include "sw_optics.hpp"

namespace rrtmg {

// ---------------------------------------------------------------------------
// prep: FP64 (rrtmg_sw_rad.nomcica.f90:933-1047 and rrtmg_sw_setcoef.f90:129-284),
// one thread per (layer, column): the layer's coefficients in the band
// kernel's precision and its integer table indices.
template <typename Real>
__global__ __launch_bounds__(256) void sw_prep_kernel(ColumnInputs in, SwWork w, const double *__restrict__ preflog,
                                                      const double *__restrict__ tref)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nlay * ncol, iblk = (size_t)blockIdx.x * blockDim.x;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    // this thread's coefficient record, staged for store_records
    __shared__ __align__(16) Real srec[256 * SWF_STRIDE];
    Real *rec = srec + threadIdx.x * SWF_STRIDE;
    const double amd = 28.9660, amw = 18.0160, amdw = 1.607793, amdo = 0.603428;
    const double avogad = 6.02214199e+23, grav = 9.8066;
    const double stpfac = 296.0 / 1013.0;
    if (i < n) {
        const size_t ix = (size_t)lay * ncol + col;
        double pavel = in.play[ix], tavel = in.tlay[ix];
        double pz0 = in.plev[ix], pz1 = in.plev[ix + ncol];
        double h2o = in.h2o[ix];
        // in.vmr: h2o and O3 as volume mixing ratios (the McICA driver's
        // inatm, rrtmg_sw_rad.f90:1015, 1021), else MiMA's specific
        // humidity and O3 mass mixing ratio (rrtmg_sw_rad.nomcica.f90:1004, 1007)
        double wkl1 = in.vmr ? h2o : (h2o / (1.0 - h2o)) * amdw;
        double amm = (1.0 - wkl1) * amd + wkl1 * amw;
        double coldry = (pz0 - pz1) * 1.e3 * avogad / (1.e2 * grav * amm * (1.0 + wkl1));
        double colh2o = 1.e-20 * (coldry * wkl1);
        double colco2 = 1.e-20 * (coldry * in.co2[ix]);
        double colo3 = 1.e-20 * (coldry * (in.vmr ? in.o3[ix] : in.o3[ix] * amdo));
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
        rec[SWF_FAC10] = (Real)(compfp * ft);
        rec[SWF_FAC00] = (Real)(compfp * (1.0 - ft));
        rec[SWF_FAC11] = (Real)(fp * ft1);
        rec[SWF_FAC01] = (Real)(fp * (1.0 - ft1));
        rec[SWF_COLH2O] = (Real)colh2o;
        rec[SWF_COLCO2] = (Real)colco2;
        rec[SWF_COLO3] = (Real)colo3;
        rec[SWF_COLCH4] = (Real)colch4;
        rec[SWF_COLO2] = (Real)colo2;
        rec[SWF_COLMOL] = (Real)colmol;
        rec[SWF_SELFFAC] = (Real)selffac;
        rec[SWF_SELFFRAC] = (Real)selffrac;
        rec[SWF_FORFAC] = (Real)forfac;
        rec[SWF_FORFRAC] = (Real)forfrac;
        w.ipk[ix] = pack_sw(jp, jt, jt1, indself, indfor);
        for (int f = SWF_COUNT; f < SWF_STRIDE; ++f) rec[f] = (Real)0;
    }
    const int nrec = (int)(n - iblk < blockDim.x ? n - iblk : blockDim.x);
    store_records<Real, SWF_STRIDE>((Real *)w.coef + iblk * SWF_STRIDE, srec, nrec);
}

// per-column part of the prep (one thread per column, after the per-layer
// part, whose packed indices hold jp): laytrop and the solar-source
// reference layers
__global__ __launch_bounds__(256) void sw_column_kernel(ColumnInputs in, SwWork w)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= in.ncol) return;
    sw_column_refs(in.play, w.ipk, in.ncol, in.nlay, col, w.laytrop + col, w.laysolfr);
}

// One band JB of one work unit (rrtmg_sw_spcvrt.f90:302-600 with
// rrtmg_sw_vrtqdr.f90:28-152, clear sky).  ktab: the k-tables (B's offsets index
// it: the packed global buffer, or the band's copy in LDS with rebased
// offsets); exp_tbl: the exp table (global) or NULL for the smooth function.
// s_prup, s_prupd: reflectances of the atmosphere below each level,
// [(nlay+1)][64] each (LDS or this wave's global scratch row); s_red: 64 LDS
// entries for the FP64 g-sum.
template <int JB, typename Real>
__device__ __forceinline__ void sw_band_body(const ColumnInputs &in, const SwWork &w, const SwBand &B, const Real *__restrict__ ktab,
                             const Real *__restrict__ exp_tbl, WaveUnit u, int tile, int lane,
                             Real *__restrict__ s_prup, Real *__restrict__ s_prupd, Real *s_red)
{
    const int ncol = in.ncol, nlay = in.nlay;
    constexpr int ng = sw_ng(JB);  // B.ng, checked at init
    // the band record with its species-parameter counts as constants
    SwBand Bc = B;
    Bc.nspa = SW_NSPA[JB - 16];
    Bc.nspb = SW_NSPB[JB - 16];
    const int cl = lane / ng, g = lane - cl * ng;
    const GSum<Real, ng> gsum(lane, g);
    const int col = tile * TILE_COLS + u.first_col + cl;
    const bool active = cl < u.ncols && col < ncol && in.coszen[min(col, ncol - 1)] >= 1.e-10;
    // a wave whose columns are all at night writes nothing: the finish
    // kernel sets night columns to zero without reading the band fluxes
    if (!__any(active)) return;
    const Real *coef = (const Real *)w.coef;
    Real *bflux = (Real *)w.bflux;
    Tab<Real> T{ktab, ng, g};

    Real prmu0 = (Real)1, zincflx = (Real)0, albp = (Real)0, albd = (Real)0;
    int laytrop = 0;
    SwLayer<Real> c;
    if (active) {
        prmu0 = (Real)in.coszen[col];
        laytrop = w.laytrop[col];
        const bool nir = (JB <= 24) || JB == 29;  // rrtmg_sw_rad.nomcica.f90:567-577
        albp = (Real)(nir ? in.aldir[col] : in.asdir[col]);
        albd = (Real)(nir ? in.aldif[col] : in.asdif[col]);
        int lsf = w.laysolfr[(size_t)(JB - 16) * ncol + col];
        sw_load_layer(c, coef, w.ipk, lsf - 1, col, ncol, nlay);
        double sfz = sw_sfluxzen(JB, Bc, T, c, lsf <= laytrop);
        zincflx = (Real)(in.adjflux * sfz * (double)prmu0);
    }

    // bottom-up sweep: reflectances of the atmosphere below each level
    Real prup_n = albp, prupd_n = albd;
    at32(s_prup, (unsigned)(nlay * 64 + lane)) = albp;
    at32(s_prupd, (unsigned)(nlay * 64 + lane)) = albd;
    for (int lay = 1; lay <= nlay; ++lay) {
        const int jk = nlay + 1 - lay;  // top-down index of the layer
        Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
        if (active) {
            sw_load_layer(c, coef, w.ipk, lay - 1, col, ncol, nlay);
            Real taug, taur;
            sw_gas(JB, Bc, T, c, lay <= laytrop, taug, taur);
            sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
        }
        Real zreflect = fdiv((Real)1, (Real)1 - prupd_n * rd);
        Real prup = r + (td * ((t - dbt) * prupd_n + dbt * prup_n)) * zreflect;
        Real prupd = rd + td * td * prupd_n * zreflect;
        at32(s_prup, (unsigned)((jk - 1) * 64 + lane)) = prup;
        at32(s_prupd, (unsigned)((jk - 1) * 64 + lane)) = prupd;
        prup_n = prup;
        prupd_n = prupd;
    }

    // top-down sweep with level fluxes
    Real ztdn = (Real)1, prdnd = (Real)0, ptdbt = (Real)1;
    const size_t bstride = (size_t)(nlay + 1) * ncol;
    Real *bf_up = bflux + (size_t)(JB - 16) * 2 * bstride;
    Real *bf_dn = bf_up + bstride;
    Real prupd_next = at32(s_prupd, (unsigned)lane), prup_next = at32(s_prup, (unsigned)lane);
    for (int jk = 1; jk <= nlay + 1; ++jk) {
        if (jk > 1) {
            const int lay = nlay + 2 - jk;  // layer between levels jk-1 and jk
            Real r = 0, rd = 0, t = 1, td = 1, dbt = 1;
            if (active) {
                sw_load_layer(c, coef, w.ipk, lay - 1, col, ncol, nlay);
                Real taug, taur;
                sw_gas(JB, Bc, T, c, lay <= laytrop, taug, taur);
                sw_layer_optics(taug, taur, prmu0, exp_tbl, r, rd, t, td, dbt);
            }
            Real zreflect = fdiv((Real)1, (Real)1 - rd * prdnd);
            Real ztdn_n = ptdbt * t + (td * ((ztdn - ptdbt) + ptdbt * r * prdnd)) * zreflect;
            Real prdnd_n = rd + td * td * prdnd * zreflect;
            ptdbt = dbt * ptdbt;
            ztdn = ztdn_n;
            prdnd = prdnd_n;
        }
        // the reflectances of the level below are read one level ahead
        const Real prup = prup_next, prupd = prupd_next;
        if (jk <= nlay) {
            prupd_next = at32(s_prupd, (unsigned)(jk * 64 + lane));
            prup_next = at32(s_prup, (unsigned)(jk * 64 + lane));
        }
        Real zreflect = fdiv((Real)1, (Real)1 - prdnd * prupd);
        Real pfu = (ptdbt * prup + (ztdn - ptdbt) * prupd) * zreflect;
        Real pfd = ptdbt + (ztdn - ptdbt + ptdbt * prup * prdnd) * zreflect;
        const int lev = nlay + 1 - jk;  // bottom-up level index
        Real su = gsum.sum(active ? zincflx * pfu : (Real)0, s_red);
        Real sd = gsum.sum(active ? zincflx * pfd : (Real)0, s_red);
        if (active && g == 0) {
            at32(bf_up, (unsigned)(lev * ncol + col)) = su;
            at32(bf_dn, (unsigned)(lev * ncol + col)) = sd;
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 6 : 1))) void sw_band_kernel(ColumnInputs in, SwWork w, const SwBand *__restrict__ bands,
                                                     int nwaves, int blk0, int nblocks, int exp_off,
                                                     const Real *__restrict__ tab, const WaveUnit *__restrict__ units)
{
    extern __shared__ unsigned char smem_raw[];
    Real *smem = (Real *)smem_raw;
    const int nlay = in.nlay;
    const int blk = xcd_block();
    if (blk >= nblocks) return;
    const int tile = (blk0 + blk) / nwaves;
    const WaveUnit u = units[(blk0 + blk) % nwaves];
    const SwBand B = bands[u.band];
    Real *s_red = smem;  // [64], LDS (FP64 g-sums)
    // prupd and prup of every level in this wave's global scratch row,
    // [2][(nlay+1)][64] (in LDS they would cap occupancy below the VGPRs)
    Real *s_prupd = (Real *)w.dstate + (size_t)blk * 2 * (nlay + 1) * 64;
    Real *s_prup = s_prupd + (size_t)(nlay + 1) * 64;
    const Real *exp_tbl = exp_off >= 0 ? tab + exp_off : nullptr;
#define BODY(jb)                                                                                              \
    case jb - 16:                                                                                             \
        sw_band_body<jb, Real>(in, w, B, tab, exp_tbl, u, tile, (int)threadIdx.x, s_prup, s_prupd, s_red);      \
        break;
    switch (u.band) { SW_BANDS(BODY) }
#undef BODY
}

// finish: band sums, net flux and heating (rrtmg_sw_rad.nomcica.f90:700-726),
// one thread per (layer, column): the fluxes of the layer's lower level (and
// of the top level) and its heating rate, zero in the top layer
template <typename Real>
__global__ __launch_bounds__(256) void sw_finish_kernel(ColumnInputs in, SwWork w, SwOutputs out, double heatfac)
{
    const int ncol = in.ncol, nlay = in.nlay;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nlay * ncol) return;
    const int lay = (int)(i / ncol), col = (int)(i - (size_t)lay * ncol);
    const Real *bflux = (const Real *)w.bflux;
    const size_t bstride = (size_t)(nlay + 1) * ncol;
    const size_t il = (size_t)lay * ncol + col;
    if (in.coszen[col] < 1.e-10) {
        for (int k = lay; k <= (lay == nlay - 1 ? nlay : lay); ++k) {
            size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = 0.0;
            out.dflx[ix] = 0.0;
            out.uflxc[ix] = 0.0;
            out.dflxc[ix] = 0.0;
        }
        out.hr[il] = 0.0;
        out.hrc[il] = 0.0;
        return;
    }
    double net[2];
    for (int j = 0; j < 2; ++j) {
        const int k = lay + j;
        double up = 0.0, dn = 0.0;
        for (int b = 0; b < 14; ++b) {
            up += (double)bflux[(size_t)b * 2 * bstride + (size_t)k * ncol + col];
            dn += (double)bflux[(size_t)b * 2 * bstride + bstride + (size_t)k * ncol + col];
        }
        if (j == 0 || lay == nlay - 1) {
            size_t ix = (size_t)k * ncol + col;
            out.uflx[ix] = up;
            out.dflx[ix] = dn;
            out.uflxc[ix] = up;
            out.dflxc[ix] = dn;
        }
        net[j] = dn - up;
    }
    double pdp = in.plev[il] - in.plev[il + ncol];
    double h = (lay == nlay - 1) ? 0.0 : (net[1] - net[0]) * (heatfac / pdp);
    out.hr[il] = h;
    out.hrc[il] = h;
}

template <typename Real>
static hipError_t launch_sw_t(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t,
                              bool smooth, hipStream_t s, hipEvent_t *ev)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(sw_prep_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, t.preflog, t.tref);
    hipLaunchKernelGGL(sw_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
    const int ntiles = (in.ncol + TILE_COLS - 1) / TILE_COLS;
    const size_t shmem = sw_fwd_lds_bytes(in.nlay, sizeof(Real));
    if (ev) (void)hipEventRecord(ev[1], s);
    const int nblocks = ntiles * t.desc.nwaves;
    // launches of at most w.chunk wave units, each reusing the scratch rows
    for (int b0 = 0; b0 < nblocks; b0 += w.chunk) {
        const int nb_c = std::min(w.chunk, nblocks - b0);
        hipLaunchKernelGGL(sw_band_kernel<Real>, dim3(xcd_grid(nb_c)), dim3(64), shmem, s, in, w, t.bands,
                           t.desc.nwaves, b0, nb_c, smooth ? -1 : t.desc.exp_tbl, (const Real *)t.tab, t.units);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(sw_finish_kernel<Real>, dim3(nbl), dim3(256), 0, s, in, w, out, t.heatfac);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

hipError_t launch_sw_prep(const ColumnInputs &in, const SwWork &w, const SwDevTables &t, int mode, hipStream_t s)
{
    const int nb = (in.ncol + 255) / 256;
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    if (mode == 1)
        hipLaunchKernelGGL(sw_prep_kernel<float>, dim3(nbl), dim3(256), 0, s, in, w, t.preflog, t.tref);
    else
        hipLaunchKernelGGL(sw_prep_kernel<double>, dim3(nbl), dim3(256), 0, s, in, w, t.preflog, t.tref);
    hipLaunchKernelGGL(sw_column_kernel, dim3(nb), dim3(256), 0, s, in, w);
    return hipGetLastError();
}

hipError_t launch_sw_finish(const ColumnInputs &in, const SwOutputs &out, void *bflux, const SwDevTables &t,
                            int mode, hipStream_t s)
{
    const int nbl = (int)(((size_t)in.nlay * in.ncol + 255) / 256);
    SwWork w;
    w.bflux = bflux;
    if (mode == 1)
        hipLaunchKernelGGL(sw_finish_kernel<float>, dim3(nbl), dim3(256), 0, s, in, w, out, t.heatfac);
    else
        hipLaunchKernelGGL(sw_finish_kernel<double>, dim3(nbl), dim3(256), 0, s, in, w, out, t.heatfac);
    return hipGetLastError();
}

hipError_t launch_sw(const ColumnInputs &in, const SwOutputs &out, const SwWork &w, const SwDevTables &t, int mode,
                     hipStream_t s, hipEvent_t *ev)
{
    return mode == 1 ? launch_sw_t<float>(in, out, w, t, false, s, ev)
                     : launch_sw_t<double>(in, out, w, t, mode == 2, s, ev);
}

}  // namespace rrtmg
