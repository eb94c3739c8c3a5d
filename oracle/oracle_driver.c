/* MiMA's radiation driver around rrtmg_sw/rrtmg_lw (TEST INFRASTRUCTURE ONLY;
 * see oracle.h): compute_zenith, interp_temp and run_rrtmg's pre- and
 * post-processing, restated in plain C and FP64 (MiMA builds with -r8, so
 * every `real` of the reference is FP64).  Paths are relative to
 * src/atmos_param/rrtm_radiation/ (RR/ below).
 *
 * GCM arrays are FMS physics-window arrays (nlon, nlat, nlev) in Fortran
 * order (longitude fastest), level 1 at the top of the atmosphere; RRTMG
 * arrays are (ncol, nlay) column-major with level 0 at the surface. */
#include <math.h>
#include <string.h>

#include "oracle.h"

static const double PI = 3.141592653589793; /* 4 atan(1), shared/fms/fms.f90:412 */

/* Fortran MODULO for reals (sign of the divisor), as gfortran and flang
 * compute it: fmod, then one period added when the signs differ. */
static double fmodulo(double a, double p)
{
    double r = fmod(a, p);
    if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
    return r;
}

static int imodulo(int a, int p)
{
    int r = a % p;
    if (r != 0 && ((r < 0) != (p < 0))) r += p;
    return r;
}

/* RR/astro.f90:59-248 (compute_zenith), for n points of lat/lon (radians);
 * model time as get_time returns it (seconds of the day, days). */
int oracle_compute_zenith(int n, const double *lat, const double *lon, int seconds, int days, int daysperyear,
                          double equinox_day, double obliq, int dt, double *cosz, int *dyofyr)
{
    if (n < 0 || daysperyear <= 0) return -1;
    const double deg2rad = PI / 180.;                                   /* :100 */
    const double twopi = 2 * PI;
    const double radpersec = 2 * PI / 86400.;                           /* :114 */
    const double radsec = seconds * radpersec;                          /* :119-121 */
    const double dt_pi = dt * radpersec;
    int d = days - (int)(equinox_day * daysperyear);                    /* :133 */
    *dyofyr = imodulo(d, daysperyear);                                  /* :134 */
    const double radperday = 2 * PI / daysperyear;
    const double radday = *dyofyr * radperday;                          /* :136 */
    const double dec_sin = sin(obliq * deg2rad) * sin(radday);          /* :138 */
    const double dec = asin(dec_sin);
    const double dec_cos = cos(dec);
    const double dec_tan = tan(dec);                                    /* :143 */
    const double eps = 1.0E-05;
    for (int p = 0; p < n; ++p) {
        double time_pi = fmodulo(radsec + lon[p], 2 * PI) - PI;         /* :125 */
        if (time_pi >= PI) time_pi = time_pi - twopi;                   /* :127 */
        if (time_pi < -PI) time_pi = time_pi + twopi;
        double lat_h = lat[p];                                          /* :144-146 */
        if (lat_h == 0.5 * PI) lat_h = lat[p] - eps;
        if (lat_h == -0.5 * PI) lat_h = lat[p] + eps;
        const double cos_h = -tan(lat_h) * dec_tan;                     /* :147 */
        double h = 0.0;
        if (cos_h <= -1.0) h = PI;                                      /* :148-151 */
        if (cos_h >= 1.0) h = 0.0;
        if (cos_h > -1.0 && cos_h < 1.0) h = acos(cos_h);
        const double aa = sin(lat[p]) * dec_sin;                        /* :155-156 */
        const double bb = cos(lat[p]) * dec_cos;
        double c;
        if (dt > 0 && dt < 86400.) {                                    /* :159-233 */
            const double tt = time_pi + dt_pi;
            const double st = sin(time_pi), stt = sin(tt), sh = sin(h);
            c = 0.0;
            if (time_pi < -h && tt < -h) c = 0.0;                                          /* case 1 */
            if ((tt + h) != 0.0 && time_pi < -h && fabs(tt) <= h) c = aa + bb * (stt + sh) / (tt + h); /* 2 */
            if (time_pi < -h && h != 0.0 && h < tt) c = aa + bb * (sh + sh) / (h + h);     /* 3 */
            if (fabs(time_pi) <= h && fabs(tt) <= h) c = aa + bb * (stt - st) / (tt - time_pi); /* 4 */
            if ((h - time_pi) != 0.0 && fabs(time_pi) <= h && h < tt)
                c = aa + bb * (sh - st) / (h - time_pi);                                    /* 5 */
            if (twopi - h < tt && (tt + h - twopi) != 0.0 && time_pi <= h)                 /* 6 */
                c = (c * (h - time_pi) + (aa * (tt + h - twopi) + bb * (stt + sh))) /
                    ((h - time_pi) + (tt + h - twopi));
            if (h < time_pi && twopi - h >= tt) c = 0.0;                                    /* 7 */
            if (h < time_pi && twopi - h < tt) c = aa + bb * (stt + sh) / (tt + h - twopi); /* 8 */
            double fracday = 0.0;                                                            /* :224-230 */
            if (time_pi < -h && tt < -h) fracday = 0.0;
            if (time_pi < -h && fabs(tt) <= h) fracday = (tt + h) / dt;
            if (time_pi < -h && h < tt) fracday = (h + h) / dt;
            if (fabs(time_pi) <= h && fabs(tt) <= h) fracday = (tt - time_pi) / dt;
            if (fabs(time_pi) <= h && h < tt) fracday = (h - time_pi) / dt;
            if (h < time_pi) fracday = 0.0;
            if (twopi - h < tt) fracday = fracday + (tt + h - twopi) / dt;
            c = c * fracday / radpersec;                                                     /* :235 */
        } else if (dt >= 86400.) {
            c = (aa * h + bb * sin(h)) / PI;                                                 /* :240 */
        } else {
            c = fabs(time_pi) <= h ? aa + bb * cos(time_pi) : 0.0;                           /* :245-250 */
        }
        cosz[p] = c > 0.0 ? c : 0.0;                                                         /* :255 */
    }
    return 0;
}

/* RR/rrtm_radiation.f90:422-461 (interp_temp) */
int oracle_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                       const double *t_surf, const double *t, double *t_half)
{
    const long nh = (long)nlon * nlat;
    if (nlev < 2) return -1;
    for (int k = 1; k < nlev; ++k)                                      /* Fortran k = 2..kend */
        for (long p = 0; p < nh; ++p) {
            const double dzk2 = 1. / (z_full[(k - 1) * nh + p] - z_full[k * nh + p]);
            const double dzk = (z_half[k * nh + p] - z_full[k * nh + p]) * dzk2;
            const double dzk1 = (z_full[(k - 1) * nh + p] - z_half[k * nh + p]) * dzk2;
            t_half[k * nh + p] = t[k * nh + p] * dzk1 + t[(k - 1) * nh + p] * dzk;
        }
    for (long p = 0; p < nh; ++p) {
        t_half[p] = 0.5 * (3 * t[p] - t[nh + p]);                       /* :449 */
        t_half[(long)nlev * nh + p] = t_surf[p];                        /* :456 */
    }
    return 0;
}

/* RR/rrtm_radiation.f90:581-677: the state run_rrtmg hands to rrtmg_sw and
 * rrtmg_lw, after compute_zenith.  precip and o3f may be NULL when their
 * options are off. */
int oracle_rrtmg_pre(const OracleDriverConfig *cf, const double *lat, const double *p_full, const double *p_half,
                     const double *t, const double *t_half, const double *q, const double *o3f,
                     const double *albedo, const double *precip, const double *t_surf, const double *coszen,
                     double *play, double *plev, double *tlay, double *tlev, double *h2ovmr, double *o3vmr,
                     double *co2vmr, double *albedo_rr, double *tsfc, double *coszen_rr)
{
    const int si = cf->nlon, sj = cf->nlat, sk = cf->nlev, ls = cf->lonstep;
    if (si <= 0 || sj <= 0 || sk < 2 || ls <= 0 || si % ls) return -1;
    const long nh = (long)si * sj;
    const int si2 = si / ls;
    const long ncol = (long)si2 * sj;
    /* zonal means (sum over longitude in index order) */
    double zq[4096], zp[4096];
    if ((long)sj * sk > 4096 || sj > 4096) return -2;
    if (cf->do_zm_tracers)                                               /* :618-622 */
        for (int k = 0; k < sk; ++k)
            for (int j = 0; j < sj; ++j) {
                double s = 0.0;
                for (int i = 0; i < si; ++i) s += q[k * nh + (long)j * si + i];
                zq[k * sj + j] = s / si;
            }
    if (cf->do_precip_albedo && cf->num_precip > 0)                     /* :603-609 */
        for (int j = 0; j < sj; ++j) {
            double s = 0.0;
            for (int i = 0; i < si; ++i) {
                const long p = (long)j * si + i;
                const double pr = fabs(lat[p]) < cf->precip_lat * 3.14159265 / 180. ? 0. : precip[p];
                s += pr;
            }
            zp[j] = s;
        }
    /* top interface: minval over the columns handed to RRTM (:654-656) */
    int fix_top = 0;
    for (int j = 0; j < sj; ++j)
        for (int i = 0; i < si; i += ls)
            if (p_half[(long)j * si + i] * 0.01 <= 0.) fix_top = 1;
    for (int j = 0; j < sj; ++j)
        for (int i2 = 0; i2 < si2; ++i2) {
            const int i = i2 * ls;
            const long p = (long)j * si + i, c = (long)j * si2 + i2;
            for (int kr = 0; kr < sk; ++kr) {                            /* RRTM layer kr = GCM level sk-1-kr */
                const int k = sk - 1 - kr;
                const long g = (long)k * nh + p;
                double qq = cf->do_zm_tracers ? zq[k * sj + j] : q[g];
                if (cf->do_fixed_water && fabs(lat[p]) <= cf->fixed_water_lat &&   /* :632-640 */
                    p_full[g] <= cf->fixed_water_pres * 100.)
                    qq = cf->fixed_water;
                play[kr * ncol + c] = p_full[g] * 0.01;                 /* :652 */
                double tt = t[g];
                tt = tt > cf->temp_lower_limit ? tt : cf->temp_lower_limit;   /* :672-674 */
                tt = tt < cf->temp_upper_limit ? tt : cf->temp_upper_limit;
                tlay[kr * ncol + c] = tt;
                h2ovmr[kr * ncol + c] = qq > cf->h2o_lower_limit ? qq : cf->h2o_lower_limit;   /* :671 */
                if (cf->do_read_ozone) {                                 /* :601-602, :660 */
                    const double o = o3f[g] * cf->scale_ozone;
                    o3vmr[kr * ncol + c] = o > 0.0 ? o : 0.0;
                } else {
                    o3vmr[kr * ncol + c] = cf->o3_val;                   /* rrtm_radiation_init */
                }
                co2vmr[kr * ncol + c] = cf->co2ppmv * 1.e-6;
            }
            for (int kr = 0; kr <= sk; ++kr) {
                const int k = sk - kr;
                const long g = (long)k * nh + p;
                double ph = p_half[g] * 0.01;                            /* :653 */
                if (kr == sk && fix_top) ph = p_full[p] * 0.01 * 0.5;    /* pfull(:,sk)*0.5 */
                plev[kr * ncol + c] = ph;
                double th = t_half[g];
                th = th > cf->temp_lower_limit ? th : cf->temp_lower_limit;  /* :675-676 */
                th = th < cf->temp_upper_limit ? th : cf->temp_upper_limit;
                tlev[kr * ncol + c] = th;
            }
            double alb = albedo[p];
            if (cf->do_precip_albedo && cf->num_precip > 0)
                alb = albedo[p] + (cf->precip_albedo - albedo[p]) * zp[j] / si / cf->num_precip;
            albedo_rr[c] = alb;                                          /* :663 */
            tsfc[c] = t_surf[p];                                         /* :664 */
            coszen_rr[c] = coszen[p];                                    /* :662 */
        }
    return 0;
}

/* RR/rrtm_radiation.f90:715-752 and :756-826: the heating-rate tendency and
 * surface fluxes run_rrtmg forms from the RRTM outputs, interpolated back
 * from every lonstep-th longitude (closing toroidally) or zonally averaged.
 * Optional outputs may be NULL. */
int oracle_rrtmg_post(const OracleDriverConfig *cf, const double *swhr, const double *hr, const double *swuflx,
                      const double *swdflx, const double *uflx, const double *dflx, double *tdt, double *tdt_rad,
                      double *flux_sw, double *flux_lw, double *tdt_sw, double *tdt_lw, double *olr, double *isr)
{
    const int si = cf->nlon, sj = cf->nlat, sk = cf->nlev, ls = cf->lonstep;
    if (si <= 0 || sj <= 0 || sk < 2 || ls <= 0 || si % ls) return -1;
    const long nh = (long)si * sj;
    const int si2 = si / ls;
    const long ncol = (long)si2 * sj;
    const double daypersec = 1. / 86400;                                 /* rrtm_vars */
    const double dlon = 1. / ls;
#define SW(i, j, k) (swhr[(long)(sk - 1 - (k)) * ncol + (long)(j) * si2 + (i)] * daypersec)   /* :715 */
#define LWH(i, j, k) (hr[(long)(sk - 1 - (k)) * ncol + (long)(j) * si2 + (i)] * daypersec)    /* :751 */
    double zm[4096], zf[2 * 4096];
    if ((long)sj * sk > 4096 || sj > 4096) return -2;
    if (cf->do_zm_rad) {
        for (int k = 0; k < sk; ++k)
            for (int j = 0; j < sj; ++j) {
                double s = 0.0;
                for (int i = 0; i < si2; ++i) s += SW(i, j, k) + LWH(i, j, k);
                zm[k * sj + j] = s / (si2 > 1 ? si2 : 1);
            }
        for (int j = 0; j < sj; ++j) {
            double s1 = 0.0, s2 = 0.0;
            for (int i = 0; i < si2; ++i) {
                const long c = (long)j * si2 + i;
                s1 += swdflx[c] - swuflx[c];
                s2 += dflx[c];
            }
            zf[2 * j] = s1 / (si2 > 1 ? si2 : 1);
            zf[2 * j + 1] = s2 / (si2 > 1 ? si2 : 1);
        }
    }
    for (int i = 0; i < si2; ++i) {
        const int i1 = i + 1 < si2 ? i + 1 : 0;                          /* :761-762 */
        for (int ij = 0; ij < ls; ++ij) {
            const double di = ij * dlon;                                 /* :764 */
            const int ij1 = i * ls + ij;
            for (int j = 0; j < sj; ++j) {
                for (int k = 0; k < sk; ++k) {
                    double v;
                    if (cf->do_zm_rad)
                        v = zm[k * sj + j];
                    else
                        v = +di * (SW(i1, j, k) + LWH(i1, j, k)) + (1. - di) * (SW(i, j, k) + LWH(i, j, k));
                    const long g = (long)k * nh + (long)j * si + ij1;
                    tdt[g] = tdt[g] + v;                                 /* :786 */
                    if (tdt_rad) tdt_rad[g] = v;                         /* :788 */
                    if (tdt_sw) tdt_sw[g] = di * SW(i1, j, k) + (1. - di) * SW(i, j, k);
                    if (tdt_lw) tdt_lw[g] = di * LWH(i1, j, k) + (1. - di) * LWH(i, j, k);
                }
                const long h = (long)j * si + ij1;
                const long c0 = (long)j * si2 + i, c1 = (long)j * si2 + i1;
                if (olr) olr[h] = di * uflx[(long)sk * ncol + c1] + (1. - di) * uflx[(long)sk * ncol + c0];
                if (isr)
                    isr[h] = di * (swdflx[(long)sk * ncol + c1] - swuflx[(long)sk * ncol + c1]) +
                             (1. - di) * (swdflx[(long)sk * ncol + c0] - swuflx[(long)sk * ncol + c0]);
                if (flux_sw) {                                           /* :796-808 */
                    if (cf->do_zm_rad) {
                        flux_sw[h] = zf[2 * j];
                        flux_lw[h] = zf[2 * j + 1];
                    } else {
                        flux_sw[h] = di * (swdflx[c1] - swuflx[c1]) + (1. - di) * (swdflx[c0] - swuflx[c0]);
                        flux_lw[h] = di * dflx[c1] + (1. - di) * dflx[c0];
                    }
                }
            }
        }
    }
#undef SW
#undef LWH
    return 0;
}
