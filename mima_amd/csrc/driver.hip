// MiMA's radiation driver around rrtmg_sw/rrtmg_lw, on the device:
// compute_zenith (RR/astro.f90:59-248), interp_temp
// (RR/rrtm_radiation.f90:422-461) and run_rrtmg's pre- and post-processing
// (RR/rrtm_radiation.f90:581-677 and :715-826), so a model whose state is
// resident on the GPU hands it to the radiation kernels and takes the
// tendencies back without host copies.  RR = src/atmos_param/rrtm_radiation.
//
// GCM arrays are FMS physics-window arrays (nlon, nlat, nlev) in Fortran
// order (longitude fastest), level 1 at the top; RRTMG arrays are
// (ncol, nlay) column-major, level 0 at the surface, ncol = nlon*nlat/lonstep.
// Every kernel is a one-pass gather/scatter (HBM-bound): one thread per
// output element, consecutive threads on consecutive longitudes (GCM side)
// or columns (RRTMG side).  Sums over longitude (zonal means) run in index
// order, one thread per (latitude, level).
//
// FP contraction is off in this file: the reshapes, clamps, interpolation
// weights and sums are the reference's sequence of IEEE operations, so the
// results are bitwise those of the same arithmetic on the CPU.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace rrtmg {

// ---------------------------------------------------------------------------
// compute_zenith

__global__ __launch_bounds__(256) void zenith_kernel(int n, const double *__restrict__ lat,
                                                     const double *__restrict__ lon, ZenithScalars z,
                                                     double *__restrict__ cosz)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double PI = 3.141592653589793, twopi = 2 * PI;
    double r = fmod(z.radsec + lon[p], 2 * PI);                      // Fortran MODULO (:125)
    if (r != 0.0 && r < 0.0) r += 2 * PI;
    double time_pi = r - PI;
    if (time_pi >= PI) time_pi = time_pi - twopi;
    if (time_pi < -PI) time_pi = time_pi + twopi;
    double lat_h = lat[p];
    if (lat_h == 0.5 * PI) lat_h = lat[p] - 1.0E-05;
    if (lat_h == -0.5 * PI) lat_h = lat[p] + 1.0E-05;
    const double cos_h = -tan(lat_h) * z.dec_tan;
    double h = 0.0;
    if (cos_h <= -1.0) h = PI;
    if (cos_h >= 1.0) h = 0.0;
    if (cos_h > -1.0 && cos_h < 1.0) h = acos(cos_h);
    const double aa = sin(lat[p]) * z.dec_sin;
    const double bb = cos(lat[p]) * z.dec_cos;
    double c;
    const int dt = z.dt;
    if (dt > 0 && dt < 86400.) {
        // the reference's eight cases, in its order: a later case overrides an
        // earlier one, and case 6 combines with the value before it
        const double tt = time_pi + z.dt_pi;
        const double st = sin(time_pi), stt = sin(tt), sh = sin(h);
        c = 0.0;
        if (time_pi < -h && tt < -h) c = 0.0;
        if ((tt + h) != 0.0 && time_pi < -h && fabs(tt) <= h) c = aa + bb * (stt + sh) / (tt + h);
        if (time_pi < -h && h != 0.0 && h < tt) c = aa + bb * (sh + sh) / (h + h);
        if (fabs(time_pi) <= h && fabs(tt) <= h) c = aa + bb * (stt - st) / (tt - time_pi);
        if ((h - time_pi) != 0.0 && fabs(time_pi) <= h && h < tt) c = aa + bb * (sh - st) / (h - time_pi);
        if (twopi - h < tt && (tt + h - twopi) != 0.0 && time_pi <= h)
            c = (c * (h - time_pi) + (aa * (tt + h - twopi) + bb * (stt + sh))) / ((h - time_pi) + (tt + h - twopi));
        if (h < time_pi && twopi - h >= tt) c = 0.0;
        if (h < time_pi && twopi - h < tt) c = aa + bb * (stt + sh) / (tt + h - twopi);
        double fracday = 0.0;
        if (time_pi < -h && tt < -h) fracday = 0.0;
        if (time_pi < -h && fabs(tt) <= h) fracday = (tt + h) / dt;
        if (time_pi < -h && h < tt) fracday = (h + h) / dt;
        if (fabs(time_pi) <= h && fabs(tt) <= h) fracday = (tt - time_pi) / dt;
        if (fabs(time_pi) <= h && h < tt) fracday = (h - time_pi) / dt;
        if (h < time_pi) fracday = 0.0;
        if (twopi - h < tt) fracday = fracday + (tt + h - twopi) / dt;
        c = c * fracday / z.radpersec;
    } else if (dt >= 86400.) {
        c = (aa * h + bb * sin(h)) / PI;
    } else {
        c = fabs(time_pi) <= h ? aa + bb * cos(time_pi) : 0.0;
    }
    cosz[p] = c > 0.0 ? c : 0.0;
}

// ---------------------------------------------------------------------------
// interp_temp: one thread per (point, interface)

__global__ __launch_bounds__(256) void interp_temp_kernel(long nh, int nlev, const double *__restrict__ z_full,
                                                          const double *__restrict__ z_half,
                                                          const double *__restrict__ t_surf,
                                                          const double *__restrict__ t, double *__restrict__ t_half)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh * (nlev + 1)) return;
    const int k = (int)(i / nh);
    const long p = i - (long)k * nh;
    double v;
    if (k == 0) {
        v = 0.5 * (3 * t[p] - t[nh + p]);
    } else if (k == nlev) {
        v = t_surf[p];
    } else {
        const double dzk2 = 1. / (z_full[(k - 1) * nh + p] - z_full[k * nh + p]);
        const double dzk = (z_half[k * nh + p] - z_full[k * nh + p]) * dzk2;
        const double dzk1 = (z_full[(k - 1) * nh + p] - z_half[k * nh + p]) * dzk2;
        v = t[k * nh + p] * dzk1 + t[(k - 1) * nh + p] * dzk;
    }
    t_half[i] = v;
}

// ---------------------------------------------------------------------------
// run_rrtmg pre-processing

// zonal mean of q per (latitude, level), sum in longitude order (:618-622)
__global__ __launch_bounds__(256) void zm_tracer_kernel(DriverDims d, const double *__restrict__ q,
                                                        double *__restrict__ zq)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nlat * d.nlev) return;
    const double *row = q + (long)i * d.nlon;  // (k, j) row: i = k * nlat + j
    double s = 0.0;
    for (int x = 0; x < d.nlon; ++x) s += row[x];
    zq[i] = s / d.nlon;
}

// zonal sum of the precipitation with the tropics (|lat| < precip_lat)
// zeroed (:603-606), and the top-interface flag: set when the top half-level
// pressure of some column handed to RRTM is not positive (:654-656)
__global__ __launch_bounds__(256) void pre_column_kernel(DriverDims d, double precip_lat, int zm_precip,
                                                         const double *__restrict__ lat,
                                                         const double *__restrict__ precip,
                                                         const double *__restrict__ p_half, double *__restrict__ zp,
                                                         int *__restrict__ flag)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.nlat) return;
    const long r = (long)j * d.nlon;
    if (zm_precip) {
        double s = 0.0;
        for (int x = 0; x < d.nlon; ++x) {
            const double pr = fabs(lat[r + x]) < precip_lat * 3.14159265 / 180. ? 0. : precip[r + x];
            s += pr;
        }
        zp[j] = s;
    }
    bool bad = false;
    for (int x = 0; x < d.nlon; x += d.lonstep) bad |= p_half[r + x] * 0.01 <= 0.;
    if (bad) atomicOr(flag, 1);
}

__global__ void flag_clear_kernel(int *flag) { *flag = 0; }

// one thread per (RRTM level kr in 0..nlev, column c): layer fields when kr <
// nlev, level fields always, column fields at kr == 0
__global__ __launch_bounds__(256) void pre_kernel(DriverDims d, DriverPre a)
{
    const long ncol = (long)(d.nlon / d.lonstep) * d.nlat;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol * (d.nlev + 1)) return;
    const int kr = (int)(i / ncol);
    const long c = i - (long)kr * ncol;
    const int si2 = d.nlon / d.lonstep;
    const int j = (int)(c / si2), i2 = (int)(c - (long)j * si2);
    const long nh = (long)d.nlon * d.nlat;
    const long p = (long)j * d.nlon + (long)i2 * d.lonstep;
    const DriverConfigD &cf = a.cf;
    if (kr < d.nlev) {
        const int k = d.nlev - 1 - kr;
        const long g = (long)k * nh + p;
        double qq = cf.do_zm_tracers ? a.zq[(long)k * d.nlat + j] : a.q[g];
        if (cf.do_fixed_water && fabs(a.lat[p]) <= cf.fixed_water_lat && a.p_full[g] <= cf.fixed_water_pres * 100.)
            qq = cf.fixed_water;
        a.play[i] = a.p_full[g] * 0.01;
        double tt = a.t[g];
        tt = tt > cf.temp_lower_limit ? tt : cf.temp_lower_limit;
        tt = tt < cf.temp_upper_limit ? tt : cf.temp_upper_limit;
        a.tlay[i] = tt;
        a.h2o[i] = qq > cf.h2o_lower_limit ? qq : cf.h2o_lower_limit;
        if (cf.do_read_ozone) {
            const double o = a.o3f[g] * cf.scale_ozone;
            a.o3[i] = o > 0.0 ? o : 0.0;
        } else {
            a.o3[i] = cf.o3_val;
        }
        a.co2[i] = cf.co2ppmv * 1.e-6;
    }
    {
        const int k = d.nlev - kr;
        const long g = (long)k * nh + p;
        double ph = a.p_half[g] * 0.01;
        if (kr == d.nlev && *a.flag) ph = a.p_full[p] * 0.01 * 0.5;
        a.plev[i] = ph;
        double th = a.t_half[g];
        th = th > cf.temp_lower_limit ? th : cf.temp_lower_limit;
        th = th < cf.temp_upper_limit ? th : cf.temp_upper_limit;
        a.tlev[i] = th;
    }
    if (kr == 0) {
        double alb = a.albedo[p];
        if (cf.do_precip_albedo && cf.num_precip > 0)
            alb = a.albedo[p] + (cf.precip_albedo - a.albedo[p]) * a.zp[j] / d.nlon / cf.num_precip;
        a.albedo_rr[c] = alb;
        a.tsfc[c] = a.t_surf[p];
        a.coszen_rr[c] = a.coszen[p];
    }
}

// ---------------------------------------------------------------------------
// run_rrtmg post-processing

// zonal means of the RRTM heating (K/s) per (latitude, level) and of the
// surface fluxes per latitude (:766-767, :800-801)
__global__ __launch_bounds__(256) void zm_rad_kernel(DriverDims d, DriverPost a, double *__restrict__ zm,
                                                     double *__restrict__ zf)
{
    const int si2 = d.nlon / d.lonstep;
    const long ncol = (long)si2 * d.nlat;
    const double daypersec = 1. / 86400;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int den = si2 > 1 ? si2 : 1;
    if (i < d.nlat * d.nlev) {
        const int k = i / d.nlat, j = i - k * d.nlat;
        const long r = (long)(d.nlev - 1 - k) * ncol + (long)j * si2;
        double s = 0.0;
        for (int x = 0; x < si2; ++x) s += a.swhr[r + x] * daypersec + a.hr[r + x] * daypersec;
        zm[i] = s / den;
    } else if (i < d.nlat * d.nlev + d.nlat) {
        const int j = i - d.nlat * d.nlev;
        const long r = (long)j * si2;
        double s1 = 0.0, s2 = 0.0;
        for (int x = 0; x < si2; ++x) {
            s1 += a.swdflx[r + x] - a.swuflx[r + x];
            s2 += a.dflx[r + x];
        }
        zf[2 * j] = s1 / den;
        zf[2 * j + 1] = s2 / den;
    }
}

// one thread per GCM (level k, latitude j, longitude x): the tendency and
// the per-level diagnostics; the surface and top-of-atmosphere fields at k == 0
__global__ __launch_bounds__(256) void post_kernel(DriverDims d, DriverPost a, const double *__restrict__ zm,
                                                   const double *__restrict__ zf)
{
    const long nh = (long)d.nlon * d.nlat;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nh * d.nlev) return;
    const int k = (int)(g / nh);
    const long p = g - (long)k * nh;
    const int j = (int)(p / d.nlon), x = (int)(p - (long)j * d.nlon);
    const int si2 = d.nlon / d.lonstep;
    const long ncol = (long)si2 * d.nlat;
    const int i = x / d.lonstep, ij = x - i * d.lonstep;
    const int i1 = i + 1 < si2 ? i + 1 : 0;
    const double daypersec = 1. / 86400;
    const double di = ij * (1. / d.lonstep);
    const long r = (long)(d.nlev - 1 - k) * ncol + (long)j * si2;
    const double sw0 = a.swhr[r + i] * daypersec, sw1 = a.swhr[r + i1] * daypersec;
    const double lw0 = a.hr[r + i] * daypersec, lw1 = a.hr[r + i1] * daypersec;
    const double v = a.do_zm_rad ? zm[(long)k * d.nlat + j] : +di * (sw1 + lw1) + (1. - di) * (sw0 + lw0);
    a.tdt[g] = a.tdt[g] + v;
    if (a.tdt_rad) a.tdt_rad[g] = v;
    if (a.tdt_sw) a.tdt_sw[g] = di * sw1 + (1. - di) * sw0;
    if (a.tdt_lw) a.tdt_lw[g] = di * lw1 + (1. - di) * lw0;
    if (k == 0) {
        const long c0 = (long)j * si2 + i, c1 = (long)j * si2 + i1, top = (long)d.nlev * ncol;
        if (a.olr) a.olr[p] = di * a.uflx[top + c1] + (1. - di) * a.uflx[top + c0];
        if (a.isr)
            a.isr[p] = di * (a.swdflx[top + c1] - a.swuflx[top + c1]) +
                       (1. - di) * (a.swdflx[top + c0] - a.swuflx[top + c0]);
        if (a.flux_sw) {
            if (a.do_zm_rad) {
                a.flux_sw[p] = zf[2 * j];
                a.flux_lw[p] = zf[2 * j + 1];
            } else {
                a.flux_sw[p] = di * (a.swdflx[c1] - a.swuflx[c1]) + (1. - di) * (a.swdflx[c0] - a.swuflx[c0]);
                a.flux_lw[p] = di * a.dflx[c1] + (1. - di) * a.dflx[c0];
            }
        }
    }
}

static unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_zenith(int n, const double *lat, const double *lon, const ZenithScalars &z, double *cosz,
                         hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(zenith_kernel, dim3(blocks(n)), dim3(256), 0, s, n, lat, lon, z, cosz);
    return hipGetLastError();
}

hipError_t launch_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                              const double *t_surf, const double *t, double *t_half, hipStream_t s)
{
    const long nh = (long)nlon * nlat;
    hipLaunchKernelGGL(interp_temp_kernel, dim3(blocks(nh * (nlev + 1))), dim3(256), 0, s, nh, nlev, z_full, z_half,
                       t_surf, t, t_half);
    return hipGetLastError();
}

hipError_t launch_rrtmg_pre(const DriverDims &d, DriverPre a, double *work, hipStream_t s)
{
    // work: [nlat * nlev] zonal-mean q, [nlat] zonal precipitation sums, flag
    a.zq = work;
    a.zp = work + (long)d.nlat * d.nlev;
    a.flag = (int *)(a.zp + d.nlat);
    hipLaunchKernelGGL(flag_clear_kernel, dim3(1), dim3(1), 0, s, a.flag);
    if (a.cf.do_zm_tracers)
        hipLaunchKernelGGL(zm_tracer_kernel, dim3(blocks((long)d.nlat * d.nlev)), dim3(256), 0, s, d, a.q, a.zq);
    const int zm_precip = a.cf.do_precip_albedo && a.cf.num_precip > 0;
    hipLaunchKernelGGL(pre_column_kernel, dim3(blocks(d.nlat)), dim3(256), 0, s, d, a.cf.precip_lat, zm_precip, a.lat,
                       a.precip, a.p_half, a.zp, a.flag);
    const long ncol = (long)(d.nlon / d.lonstep) * d.nlat;
    hipLaunchKernelGGL(pre_kernel, dim3(blocks(ncol * (d.nlev + 1))), dim3(256), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_rrtmg_post(const DriverDims &d, const DriverPost &a, double *work, hipStream_t s)
{
    // work: [nlat * nlev] zonal-mean heating, [2 nlat] zonal-mean surface fluxes
    double *zm = work, *zf = work + (long)d.nlat * d.nlev;
    if (a.do_zm_rad)
        hipLaunchKernelGGL(zm_rad_kernel, dim3(blocks((long)d.nlat * (d.nlev + 1))), dim3(256), 0, s, d, a, zm, zf);
    const long nh = (long)d.nlon * d.nlat;
    hipLaunchKernelGGL(post_kernel, dim3(blocks(nh * d.nlev)), dim3(256), 0, s, d, a, zm, zf);
    return hipGetLastError();
}

}  // namespace rrtmg
