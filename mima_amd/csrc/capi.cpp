// C ABI of the library (declared in include/rrtmg_hip.h): per-device table
// state, workspaces, argument validation, host/device pointer modes and
// error reporting.  Reference entry points replaced: rrtmg_sw_ini, rrtmg_lw_ini
// (rrtmg_*_init.f90:28), rrtmg_sw (rrtmg_sw_rad.nomcica.f90:78), rrtmg_lw
// (rrtmg_lw_rad.nomcica.f90:80).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rrtmg_hip.h"
#include "capi_internal.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "tables.hpp"

using namespace rrtmg;

namespace {

thread_local std::string t_err;
extern bool g_timing;

int fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

struct Buffer {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap) return 0;
        if (p) {
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        if (std::getenv("RRTMG_HIP_FAIL_ALLOC")) return RRTMG_HIP_ENOMEM;  // injected failure for tests
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return RRTMG_HIP_ENOMEM;
        }
        cap = bytes;
        return 0;
    }
};

struct SideTables {
    bool ready = false;
    void *tab64 = nullptr, *tab32 = nullptr;
    void *bands = nullptr;
    WaveUnit *units = nullptr;
    double *preflog = nullptr, *tref = nullptr;  // SW
    LwConst *cst = nullptr;                      // LW
    LwBand *bands_lds = nullptr;                 // LW: offsets relative to each band's first k-table element
    LwLdsPlan plan{};                            // LW: schedule of the LDS-table kernel
    bool lds = false;                            // LW: that kernel fits this device
    double *cld = nullptr;                       // lwcldpr's / swcldpr's coefficients (LW_CLD_TAB / SW_CLD_TAB), when the file is there
    double cpdair = 0.0;
};

// The workspaces of the calls on one stream.  Calls on one stream run in
// stream order and may share buffers; calls on different streams get
// different buffers, so they can overlap on the device.
// a backward call of one side whose band and finish phases wait for
// rrtmg_hip_fused_end
struct PendingLw {
    bool on = false, fwd = false;
    int mode = 0;
    ColumnInputs in{};
    LwGrad g{};
    LwVjpWork w;
    LwDevTables t;
    LwOutputs fo{};
};
struct PendingSw {
    bool on = false, fwd = false;
    int mode = 0;
    ColumnInputs in{};
    SwGrad g{};
    SwVjpWork w;
    SwDevTables t;
    SwOutputs fo{};
};
struct FusedSession {
    bool open = false;
    PendingLw lw;
    PendingSw sw;
};

struct StreamWork {
    Buffer sw_work, lw_work, sw_stage, lw_stage;
    Buffer sw_vjp_work, lw_vjp_work, vjp_stage, mlp_work, mlp_pack, gemm_ws, flag, drv, fused_rows;
    FusedSession fused;
};

struct DeviceState {
    SwHost sw;
    LwHost lw;
    SideTables swd, lwd;
    std::map<hipStream_t, StreamWork> ws;
    StreamWork &work(hipStream_t st) { return ws[st]; }
    int sw_last_ncol = 0, sw_last_nlay = 0, lw_last_ncol = 0, lw_last_nlay = 0;
    uint32_t *xs_jump = nullptr;  // the KISS 3-shift register's squarings (McICA subcolumns), on first use
    SwWork sw_w;
    LwWork lw_w;
    // timing events per call kind: 0 SW, 1 LW, 2 SW backward, 3 LW backward
    hipEvent_t ev[4][4] = {};
    bool ev_ok = false, ev_recorded[4] = {false, false, false, false};
    std::mutex mu;
    hipEvent_t *events(int side)
    {
        if (!g_timing) return nullptr;
        if (!ev_ok) {
            for (auto &row : ev)
                for (auto &e : row)
                    if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev_ok = true;
        }
        ev_recorded[side] = true;
        return ev[side];
    }
};

std::mutex g_mu;
bool g_timing = false;
bool g_checks = false;
std::map<int, DeviceState *> g_dev;
int g_prec = RRTMG_HIP_PREC_FP64;
// 1 when the last rrtmg_hip_fused_end ran the fused band kernel
static int g_fused_last = 0;

DeviceState *state()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev.find(dev);
    if (it != g_dev.end()) return it->second;
    DeviceState *s = new DeviceState();
    g_dev[dev] = s;
    return s;
}

std::string default_table_dir()
{
    const char *env = std::getenv("RRTMG_HIP_TABLES");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((void *)&default_table_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t k = p.find_last_of('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/tables";
    }
    return "tables";
}

template <typename T>
int upload(T **dst, const void *src, size_t bytes)
{
    if (hipMalloc((void **)dst, bytes) != hipSuccess) return RRTMG_HIP_ENOMEM;
    if (hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return RRTMG_HIP_EDEVICE;
    return 0;
}

int upload_packed(const std::vector<double> &buf, void **d64, void **d32)
{
    std::vector<float> f(buf.size());
    for (size_t i = 0; i < buf.size(); ++i) f[i] = (float)buf[i];
    int rc = upload((double **)d64, buf.data(), buf.size() * sizeof(double));
    if (rc) return rc;
    return upload((float **)d32, f.data(), f.size() * sizeof(float));
}

bool bad(double x) { return !std::isfinite(x); }

int check_host_array(const double *a, size_t n, const char *name)
{
    if (!a) return fail(RRTMG_HIP_EINVAL, std::string("null array argument ") + name);
    for (size_t i = 0; i < n; ++i)
        if (bad(a[i])) return fail(RRTMG_HIP_EINVAL, std::string("non-finite value in ") + name);
    return 0;
}

int check_common(int ncol, int nlay, const double *plev, int ptr_kind)
{
    if (ncol <= 0) return fail(RRTMG_HIP_EINVAL, "ncol must be positive");
    if (nlay < 2 || nlay > RRTMG_HIP_MAX_NLAY) return fail(RRTMG_HIP_EINVAL, "nlay out of range [2, 200]");
    // the band kernels address each per-level field with 32-bit byte offsets
    if ((size_t)(nlay + 1) * (size_t)ncol * sizeof(double) >= ((size_t)1 << 32))
        return fail(RRTMG_HIP_ENOTSUP, "ncol x (nlay+1) too large for one call; split the columns");
    if (ptr_kind != RRTMG_HIP_HOST && ptr_kind != RRTMG_HIP_DEVICE) return fail(RRTMG_HIP_EINVAL, "bad ptr_kind");
    if (ptr_kind == RRTMG_HIP_HOST && plev) {
        for (int i = 0; i < ncol; ++i)
            for (int k = 0; k < nlay; ++k)
                if (!(plev[(size_t)k * ncol + i] > plev[(size_t)(k + 1) * ncol + i]))
                    return fail(RRTMG_HIP_EINVAL, "plev must decrease strictly upward");
        for (int i = 0; i < ncol; ++i)
            if (!(plev[(size_t)nlay * ncol + i] > 0.0)) return fail(RRTMG_HIP_EINVAL, "top plev must be positive");
    }
    return 0;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Device-pointer inputs checked on the device when rrtmg_hip_set_checks is
// on: the host-pointer path's checks (finite inputs, plev decreasing upward
// and positive at the top).  The call waits for the verdict on its stream.
struct NamedArray {
    const char *name;
    const double *p;
    size_t n;
};
int device_checks(int *flag, hipStream_t st, const std::vector<NamedArray> &a, const double *plev, int ncol,
                  int nlay)
{
    std::vector<DeviceArray> d;
    for (const NamedArray &x : a) d.push_back({x.p, x.n});
    int h = 0;
    hipError_t e = launch_input_checks(d.data(), (int)d.size(), plev, ncol, nlay, flag, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("input checks: ") + hipGetErrorString(e));
    if (h == INT_MAX) return 0;
    if (h == (int)a.size()) return fail(RRTMG_HIP_EINVAL, "plev must decrease strictly upward and be positive at the top");
    return fail(RRTMG_HIP_EINVAL, std::string("non-finite value in ") + a[h].name);
}

// The band kernels keep two per-level arrays per lane in LDS; refuse a layer
// count whose footprint exceeds the device's per-workgroup limit.
int device_checks(StreamWork &W, hipStream_t st, const std::vector<NamedArray> &a, const double *plev, int ncol,
                  int nlay)
{
    int rc;
    if ((rc = W.flag.ensure(256))) return fail(rc, "device allocation for the input checks failed");
    return device_checks((int *)W.flag.p, st, a, plev, ncol, nlay);
}

int check_lds(size_t bytes)
{
    static int limit = -1;
    if (limit < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 65536;
        limit = v;
    }
    if (bytes > (size_t)limit)
        return fail(RRTMG_HIP_ENOTSUP, "nlay too large for the LDS of one workgroup in this precision (" +
                                           std::to_string(bytes) + " > " + std::to_string(limit) + " bytes)");
    return 0;
}

// adjes * scon / 1368.22, or Iqbal's earth_sun factor for dyofyr > 0
// (rrtmg_sw_rad.nomcica.f90:734-758, 968-972)
double sw_adjflux(double adjes, int dyofyr, double scon)
{
    double adjflx = adjes;
    if (dyofyr > 0) {
        const double pi = 2.0 * std::asin(1.0);
        double gamma = 2.0 * pi * (dyofyr - 1) / 365.0;
        adjflx = 1.000110 + .034221 * std::cos(gamma) + .001289 * std::sin(gamma) + .000719 * std::cos(2.0 * gamma) +
                 .000077 * std::sin(2.0 * gamma);
    }
    return adjflx * (scon / 1.36822e+03);
}

// Copies host arrays into a device staging buffer on one stream (host-pointer
// calls); device-pointer calls pass the caller's pointers through.
struct Stager {
    char *p;
    bool host;
    hipStream_t st;
    int ck = 0;
    const double *in(const double *a, size_t n)
    {
        if (!host || !a) return a;
        double *d = (double *)p;
        ck |= hipMemcpyAsync(d, a, n * sizeof(double), hipMemcpyHostToDevice, st);
        p += align_up(n * sizeof(double));
        return d;
    }
    double *out(double *a, size_t n)
    {
        if (!host) return a;
        double *d = (double *)p;
        p += align_up(n * sizeof(double));
        return d;
    }
    void back(double *h, const double *d, size_t n)
    {
        if (host) ck |= hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, st);
    }
};

}  // namespace

extern "C" {

const char *rrtmg_hip_last_error(void) { return t_err.c_str(); }

int rrtmg_hip_set_precision(int mode)
{
    if (mode != RRTMG_HIP_PREC_FP64 && mode != RRTMG_HIP_PREC_MIXED && mode != RRTMG_HIP_PREC_FP64_SMOOTH)
        return fail(RRTMG_HIP_EINVAL, "bad precision");
    int prev = g_prec;
    g_prec = mode;
    return prev;
}

// Device attributes of the LDS-table kernels: compute units and the LDS of
// one workgroup.
static int lds_device(int *ncu, int *lds)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return RRTMG_HIP_EDEVICE;
    *ncu = std::max(*ncu, 1);
    return 0;
}

int rrtmg_hip_sw_init(double cpdair, const char *table_dir)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!(cpdair > 0.0) || bad(cpdair)) return fail(RRTMG_HIP_EINVAL, "cpdair must be positive");
    if (s->swd.ready) {
        s->sw.heatfac = 9.8066 * 8.6400e4 / (cpdair * 1.e2);
        s->swd.cpdair = cpdair;
        (void)noalloc_attach();
        return 0;
    }
    std::string dir = table_dir ? table_dir : default_table_dir();
    std::string err;
    if (!build_sw(dir + "/rrtmg_sw.tbl", cpdair, s->sw, err)) return fail(RRTMG_HIP_ENOENT, err);
    int rc = upload_packed(s->sw.buf, &s->swd.tab64, &s->swd.tab32);
    if (!rc) rc = upload((SwBand **)&s->swd.bands, s->sw.tab.band, sizeof s->sw.tab.band);
    if (!rc) rc = upload(&s->swd.units, s->sw.units.data(), s->sw.units.size() * sizeof(WaveUnit));
    if (!rc) rc = upload(&s->swd.preflog, s->sw.preflog.data(), 59 * sizeof(double));
    if (!rc) rc = upload(&s->swd.tref, s->sw.tref.data(), 59 * sizeof(double));
    {
        // optional: the water-path cloud optics' coefficients (inflgsw 2)
        TableFile ct;
        std::string cerr;
        if (!rc && ct.load(dir + "/rrtmg_sw_cld.tbl", cerr)) {
            const char *nm[] = {"swcld/extliq1", "swcld/ssaliq1", "swcld/asyliq1", "swcld/extice2", "swcld/ssaice2",
                                "swcld/asyice2", "swcld/extice3", "swcld/ssaice3", "swcld/asyice3", "swcld/fdlice3",
                                "swcld/abari",   "swcld/bbari",   "swcld/cbari",   "swcld/dbari",   "swcld/ebari",
                                "swcld/fbari",   "swcld/wavenum2"};
            const size_t len[] = {812, 812, 812, 602, 602, 602, 644, 644, 644, 644, 5, 5, 5, 5, 5, 5, 14};
            std::vector<double> pk;
            bool ok = true;
            for (int i = 0; i < 17; ++i) {
                std::vector<double> v = ct.find(nm[i]) ? ct.doubles(nm[i]) : std::vector<double>();
                ok = ok && v.size() == len[i];
                pk.insert(pk.end(), v.begin(), v.end());
            }
            if (ok && pk.size() == (size_t)SW_CLD_TAB) rc = upload(&s->swd.cld, pk.data(), pk.size() * sizeof(double));
        }
    }
    if (rc) return fail(rc, "uploading SW tables failed");
    s->swd.cpdair = cpdair;
    s->swd.ready = true;
    // the no-allocation entry point's scratch address on this device; a
    // failure is reported by that entry point (EDEVICE), not here
    (void)noalloc_attach();
    return 0;
}

// The LDS-table schedule of the FP32 LW band kernel: each band's k-tables
// are one contiguous range of the packed buffer (tables.cpp packs them band
// by band, before the transmittance tables); every k-table offset of the band
// must lie in it, and is rebased to its start for the LDS copy.
static bool make_lw_plan(const LwHost &h, LwLdsPlan &p, LwBand *rebased, std::string &err)
{
    for (int b = 0; b < 16; ++b) {
        const LwBand &d = h.tab.band[b];
        const int offs[] = {d.absa,   d.absb,   d.selfref, d.forref,  d.fracrefa, d.fracrefb, d.ka_mn2,
                            d.kb_mn2, d.ka_mn2o, d.kb_mn2o, d.ka_mo3, d.kb_mo3,   d.ka_mco2,  d.kb_mco2,
                            d.ka_mco, d.ka_mo2, d.kb_mo2,  d.ccl4,   d.cfc11adj, d.cfc12,   d.cfc22adj};
        int lo = 1 << 30;
        for (int o : offs)
            if (o >= 0) lo = std::min(lo, o);
        p.tab0[b] = lo;
    }
    p.tab0[16] = h.tab.exp_tbl;
    p.lds_bytes = 0;
    for (int b = 0; b < 16; ++b) {
        const int lo = p.tab0[b], hi = p.tab0[b + 1];
        if (!(lo < hi)) {
            err = "LW k-tables are not packed band by band";
            return false;
        }
        p.lds_bytes = std::max(p.lds_bytes, (int)sizeof(float) * (hi - lo));
        LwBand r = h.tab.band[b];
        int *fields[] = {&r.absa,   &r.absb,   &r.selfref, &r.forref,  &r.fracrefa, &r.fracrefb, &r.ka_mn2,
                         &r.kb_mn2, &r.ka_mn2o, &r.kb_mn2o, &r.ka_mo3, &r.kb_mo3,   &r.ka_mco2,  &r.kb_mco2,
                         &r.ka_mco, &r.ka_mo2, &r.kb_mo2,  &r.ccl4,   &r.cfc11adj, &r.cfc12,   &r.cfc22adj};
        for (int *f : fields) {
            if (*f < 0) continue;
            if (*f < lo || *f >= hi) {
                err = "LW k-table offset outside its band's range";
                return false;
            }
            *f -= lo;
        }
        rebased[b] = r;
        p.ubeg[b] = -1;
        p.upt[b] = 0;
    }
    for (int i = 0; i < (int)h.units.size(); ++i) {
        const int b = h.units[i].band;
        if (p.ubeg[b] < 0) p.ubeg[b] = i;
        if (p.ubeg[b] + p.upt[b] != i) {
            err = "LW wave units of a band are not contiguous";
            return false;
        }
        ++p.upt[b];
    }
    for (int b = 0; b < 16; ++b)
        if (p.upt[b] < 1) {
            err = "LW band without wave units";
            return false;
        }
    return true;
}

int rrtmg_hip_lw_init(double cpdair, const char *table_dir)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!(cpdair > 0.0) || bad(cpdair)) return fail(RRTMG_HIP_EINVAL, "cpdair must be positive");
    if (s->lwd.ready) {
        s->lw.cst.heatfac = 9.8066 * 8.6400e4 / (cpdair * 1.e2);
        if (hipMemcpy(s->lwd.cst, &s->lw.cst, sizeof(LwConst), hipMemcpyHostToDevice) != hipSuccess)
            return fail(RRTMG_HIP_EDEVICE, "updating heatfac failed");
        s->lwd.cpdair = cpdair;
        return 0;
    }
    std::string dir = table_dir ? table_dir : default_table_dir();
    std::string err;
    if (!build_lw(dir + "/rrtmg_lw.tbl", cpdair, s->lw, err)) return fail(RRTMG_HIP_ENOENT, err);
    int rc = upload_packed(s->lw.buf, &s->lwd.tab64, &s->lwd.tab32);
    if (!rc) rc = upload((LwBand **)&s->lwd.bands, s->lw.tab.band, sizeof s->lw.tab.band);
    if (!rc) rc = upload(&s->lwd.units, s->lw.units.data(), s->lw.units.size() * sizeof(WaveUnit));
    if (!rc) rc = upload(&s->lwd.cst, &s->lw.cst, sizeof(LwConst));
    LwBand rebased[16];
    if (!rc && !make_lw_plan(s->lw, s->lwd.plan, rebased, err)) return fail(RRTMG_HIP_ENOENT, err);
    if (!rc) rc = upload(&s->lwd.bands_lds, rebased, sizeof rebased);
    {
        // optional: the water-path cloud optics' coefficients (inflglw 1, 2)
        TableFile ct;
        std::string cerr;
        if (!rc && ct.load(dir + "/rrtmg_lw_cld.tbl", cerr)) {
            std::vector<double> pk;
            const char *nm[] = {"lwcld/abscld1", "lwcld/absice0", "lwcld/absice1", "lwcld/absice2",
                                "lwcld/absice3", "lwcld/absliq0", "lwcld/absliq1"};
            const size_t len[] = {1, 2, 10, 688, 736, 1, 928};
            bool ok = true;
            for (int i = 0; i < 7; ++i) {
                std::vector<double> v = ct.find(nm[i]) ? ct.doubles(nm[i]) : std::vector<double>();
                ok = ok && v.size() == len[i];
                pk.insert(pk.end(), v.begin(), v.end());
            }
            if (ok && pk.size() == (size_t)LW_CLD_TAB) rc = upload(&s->lwd.cld, pk.data(), pk.size() * sizeof(double));
        }
    }
    if (rc) return fail(rc, "uploading LW tables failed");
    {
        int ncu = 0, lds = 0;
        if (lds_device(&ncu, &lds)) return fail(RRTMG_HIP_EDEVICE, "device attribute query failed");
        s->lwd.plan.ncu = ncu;
        const char *env = std::getenv("RRTMG_HIP_LW_LDS");
        s->lwd.lds = s->lwd.plan.lds_bytes <= lds && !(env && env[0] == '0');
    }
    s->lwd.cpdair = cpdair;
    s->lwd.ready = true;
    return 0;
}

}  // extern "C"

namespace rrtmg {

// rrtmg_hip_sw, with its workspace from the stream's growing buffer (fx NULL)
// or carved from the caller's fixed device region fx (device pointers only;
// nothing is allocated: the band kernel's scratch rows are reused over as
// many launches as the region requires, which leaves every result bitwise
// unchanged because wave units are independent).
int sw_forward(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev, const double *tlay,
               const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
               const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
               const double *asdir, const double *asdif, const double *aldir, const double *aldif,
               const double *coszen, double adjes, int dyofyr, double scon, double *swuflx, double *swdflx,
               double *swhr, double *swuflxc, double *swdflxc, double *swhrc, int ptr_kind, void *stream,
               const FixedWorkspace *fx, const SwClouds *cl)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->swd.ready) return fail(RRTMG_HIP_ENOINIT, "rrtmg_hip_sw called before rrtmg_hip_sw_init");
    if (fx && !fx->p)
        return fail(RRTMG_HIP_EDEVICE, "the no-allocation scratch array has no device address on this device");
    if (fx && ptr_kind != RRTMG_HIP_DEVICE)
        return fail(RRTMG_HIP_EINVAL, "the no-allocation entry point takes device pointers only");
    if (icld != 0 && !cl) return fail(RRTMG_HIP_ENOTSUP, "icld != 0: rrtmg_hip_sw is clear sky; clouds take rrtmg_hip_sw_cloud");
    if (iaer != 0) return fail(RRTMG_HIP_ENOTSUP, "iaer != 0: SW aerosol optics are not implemented");
    if (cl && fx) return fail(RRTMG_HIP_EINVAL, "the no-allocation entry point is clear sky");
    if (cl && g_prec == RRTMG_HIP_PREC_FP64_SMOOTH)
        return fail(RRTMG_HIP_ENOTSUP, "the cloudy path has no smooth (FP64_SMOOTH) mode");
    if (cl && cl->inflag == 2 && !s->swd.cld)
        return fail(RRTMG_HIP_ENOTSUP, "inflgsw = 2 needs the cloud optics table rrtmg_sw_cld.tbl beside "
                                       "rrtmg_sw.tbl at rrtmg_hip_sw_init");
    int rc = check_common(ncol, nlay, plev, ptr_kind);
    if (rc) return rc;
    const size_t n2 = (size_t)ncol * nlay, n3 = (size_t)ncol * (nlay + 1);
    const double *ins2[] = {play, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr};
    const char *nm2[] = {"play", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr"};
    const double *ins3[] = {plev, tlev};
    const double *ins1[] = {tsfc, asdir, asdif, aldir, aldif, coszen};
    double *outs[] = {swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc};
    for (double *o : outs)
        if (!o) return fail(RRTMG_HIP_EINVAL, "null output array");
    if (bad(adjes) || bad(scon)) return fail(RRTMG_HIP_EINVAL, "non-finite adjes/scon");
    if (ptr_kind == RRTMG_HIP_HOST) {
        for (int i = 0; i < 8; ++i)
            if ((rc = check_host_array(ins2[i], n2, nm2[i]))) return rc;
        if ((rc = check_host_array(plev, n3, "plev")) || (rc = check_host_array(tlev, n3, "tlev"))) return rc;
        for (int i = 0; i < 6; ++i)
            if ((rc = check_host_array(ins1[i], ncol, "surface/zenith input"))) return rc;
    } else {
        for (auto p : ins2)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        for (auto p : ins3)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        for (auto p : ins1)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
    }
    hipStream_t st = (hipStream_t)stream;
    const int mixed = g_prec == RRTMG_HIP_PREC_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    if ((rc = check_lds(sw_fwd_lds_bytes(nlay, rs)))) return rc;

    // workspace: [input-check flag (fixed region only)] coefficient records,
    // packed indices, laytrop, laysolfr, band fluxes, scratch rows
    size_t off = 0;
    const size_t o_coef = off; off = align_up(off + rs * SWF_STRIDE * n2);
    const size_t o_ipk = off; off = align_up(off + sizeof(int) * n2);
    const size_t o_lt = off; off = align_up(off + sizeof(int) * ncol);
    const size_t o_lsf = off; off = align_up(off + sizeof(int) * 14 * ncol);
    const size_t o_bf = off; off = align_up(off + rs * 14 * 2 * n3);
    const size_t o_ds = off;
    const int units = (ncol + TILE_COLS - 1) / TILE_COLS * s->sw.tab.nwaves;
    const size_t unit_bytes = cl ? sw_cloud_row_bytes(nlay, rs) : sw_fwd_state_bytes(nlay, rs);
    int chunk;
    char *wb;
    int *flag = nullptr;
    StreamWork *W = nullptr;
    if (fx) {
        const size_t avail = fx->cap > 256 ? fx->cap - 256 : 0;
        const size_t rows = avail > o_ds ? (avail - o_ds) / unit_bytes : 0;
        if (rows < (size_t)std::min(units, 8))
            return fail(RRTMG_HIP_EINVAL, "ncol x nlay too large for the no-allocation entry point (needs more than " +
                                              std::to_string(o_ds + 8 * unit_bytes + 256) + " bytes of its " +
                                              std::to_string(fx->cap) + "-byte scratch array); call rrtmg_hip_sw");
        chunk = (int)std::min((size_t)units, rows);
        flag = (int *)fx->p;
        wb = (char *)fx->p + 256;
    } else {
        W = &s->work(st);
        chunk = vjp_chunk(units, unit_bytes);
    }
    if (ptr_kind == RRTMG_HIP_DEVICE && g_checks) {
        std::vector<NamedArray> a;
        for (int i = 0; i < 8; ++i) a.push_back({nm2[i], ins2[i], n2});
        a.push_back({"plev", plev, n3});
        a.push_back({"tlev", tlev, n3});
        const char *nm1[] = {"tsfc", "asdir", "asdif", "aldir", "aldif", "coszen"};
        for (int i = 0; i < 6; ++i) a.push_back({nm1[i], ins1[i], (size_t)ncol});
        if (cl && cl->mcica) {
            const double *opt[] = {cl->cldfmc, cl->taucmc, cl->ssacmc, cl->asmcmc, cl->fsfcmc, cl->ciwpmc, cl->clwpmc};
            const char *nmo[] = {"cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl"};
            for (int i = 0; i < 7; ++i)
                if (opt[i]) a.push_back({nmo[i], opt[i], 112 * n2});
            if (cl->inflag == 2 && cl->reice) a.push_back({"reicmcl", cl->reice, n2});
            if (cl->inflag == 2 && cl->reliq) a.push_back({"relqmcl", cl->reliq, n2});
        } else if (cl) {
            a.push_back({"cldfr", cl->cldfr, n2});
            const double *opt[] = {cl->taucld, cl->ssacld, cl->asmcld, cl->fsfcld};
            const char *nmo[] = {"taucld", "ssacld", "asmcld", "fsfcld"};
            for (int i = 0; i < 4; ++i)
                if (opt[i]) a.push_back({nmo[i], opt[i], 14 * n2});
            if (cl->inflag == 2 && cl->reice) a.push_back({"reice", cl->reice, n2});
            if (cl->inflag == 2 && cl->reliq) a.push_back({"reliq", cl->reliq, n2});
            if (cl->cicewp) a.push_back({"cicewp", cl->cicewp, n2});
            if (cl->cliqwp) a.push_back({"cliqwp", cl->cliqwp, n2});
        }
        if ((rc = flag ? device_checks(flag, st, a, plev, ncol, nlay) : device_checks(*W, st, a, plev, ncol, nlay)))
            return rc;
    }
    int ck = 0;
    off = align_up(o_ds + (size_t)chunk * unit_bytes);
    // cloud path: flags, four-stream band fluxes, partial-cloud count
    const size_t o_cfl = off; if (cl) off = align_up(off + sizeof(int) * n2);
    const size_t o_cbf = off; if (cl) off = align_up(off + rs * 14 * 4 * n3);
    const size_t o_cpc = off; if (cl) off = align_up(off + sizeof(int));
    const bool mcica = cl && cl->mcica;
    const size_t o_csz = off; if (mcica) off = align_up(off + sizeof(double) * ncol);
    const bool swp = cl && !cl->mcica && cl->inflag != 0;
    const size_t o_wpo = off; if (swp) off = align_up(off + sizeof(double) * 3 * 14 * n2);
    const size_t o_wpb = off; if (swp) off = align_up(off + sizeof(int));
    const bool mwp = mcica && cl->inflag == 2;
    const size_t o_mwo = off; if (mwp) off = align_up(off + sizeof(double) * 3 * 112 * n2);
    const size_t o_mwb = off; if (mwp) off = align_up(off + sizeof(int));
    if (W) {
        if ((rc = W->sw_work.ensure(off))) return fail(rc, "device workspace allocation failed");
        wb = (char *)W->sw_work.p;
    }
    SwWork w;
    w.coef = wb + o_coef;
    w.ipk = (int *)(wb + o_ipk);
    w.laytrop = (int *)(wb + o_lt);
    w.laysolfr = (int *)(wb + o_lsf);
    w.bflux = wb + o_bf;
    w.dstate = wb + o_ds;
    w.chunk = chunk;
    SwCloudWork cw;
    SwClouds cld{};
    if (cl) {
        cw.flags = (int *)(wb + o_cfl);
        cw.bflux = wb + o_cbf;
        cw.rows = wb + o_ds;
        cw.chunk = chunk;
        cw.partial = (int *)(wb + o_cpc);
        cld = *cl;
    }

    ColumnInputs in{};
    in.ncol = ncol;
    in.nlay = nlay;
    in.adjflux = sw_adjflux(adjes, dyofyr, scon);
    SwOutputs out{};
    double *stage = nullptr;
    if (ptr_kind == RRTMG_HIP_HOST) {
        size_t ncl = 0;
        if (mcica) {
            const double *sc[] = {cl->cldfmc, cl->taucmc, cl->ssacmc, cl->asmcmc, cl->fsfcmc, cl->ciwpmc, cl->clwpmc};
            for (const double *q : sc) ncl += q ? 112 * n2 : 0;
            ncl += (mwp && cl->reice ? n2 : 0) + (mwp && cl->reliq ? n2 : 0);
        } else if (cl) {
            const double *sc[] = {cl->taucld, cl->ssacld, cl->asmcld, cl->fsfcld};
            ncl = n2 + (cl->cicewp ? n2 : 0) + (cl->cliqwp ? n2 : 0) + (swp && cl->reice ? n2 : 0) +
                  (swp && cl->reliq ? n2 : 0);
            for (const double *q : sc) ncl += q ? 14 * n2 : 0;
        }
        const size_t need = sizeof(double) * (8 * n2 + 2 * n3 + 6 * (size_t)ncol + ncl + 4 * n3 + 2 * n2);
        if ((rc = W->sw_stage.ensure(need))) return fail(rc, "device staging allocation failed");
        stage = (double *)W->sw_stage.p;
        double *p = stage;
        const double **dst2[] = {&in.play, &in.tlay, &in.h2o, &in.o3, &in.co2, &in.ch4, &in.n2o, &in.o2};
        for (int i = 0; i < 8; ++i) {
            ck |= hipMemcpyAsync(p, ins2[i], n2 * sizeof(double), hipMemcpyHostToDevice, st);
            *dst2[i] = p;
            p += n2;
        }
        ck |= hipMemcpyAsync(p, plev, n3 * sizeof(double), hipMemcpyHostToDevice, st);
        in.plev = p;
        p += n3;
        ck |= hipMemcpyAsync(p, tlev, n3 * sizeof(double), hipMemcpyHostToDevice, st);
        in.tlev = p;
        p += n3;
        const double **dst1[] = {&in.tsfc, &in.asdir, &in.asdif, &in.aldir, &in.aldif, &in.coszen};
        for (int i = 0; i < 6; ++i) {
            ck |= hipMemcpyAsync(p, ins1[i], ncol * sizeof(double), hipMemcpyHostToDevice, st);
            *dst1[i] = p;
            p += ncol;
        }
        if (mcica) {
            const double *src[] = {cl->cldfmc, cl->taucmc, cl->ssacmc, cl->asmcmc, cl->fsfcmc, cl->ciwpmc, cl->clwpmc};
            const double **dst[] = {&cld.cldfmc, &cld.taucmc, &cld.ssacmc, &cld.asmcmc, &cld.fsfcmc, &cld.ciwpmc,
                                    &cld.clwpmc};
            for (int i = 0; i < 7; ++i) {
                if (!src[i]) continue;
                ck |= hipMemcpyAsync(p, src[i], 112 * n2 * sizeof(double), hipMemcpyHostToDevice, st);
                *dst[i] = p;
                p += 112 * n2;
            }
            const double *rsrc[] = {mwp ? cl->reice : nullptr, mwp ? cl->reliq : nullptr};
            const double **rdst[] = {&cld.reice, &cld.reliq};
            for (int i = 0; i < 2; ++i) {
                if (!rsrc[i]) continue;
                ck |= hipMemcpyAsync(p, rsrc[i], n2 * sizeof(double), hipMemcpyHostToDevice, st);
                *rdst[i] = p;
                p += n2;
            }
        } else if (cl) {
            const double *src[] = {cl->cldfr, cl->taucld, cl->ssacld, cl->asmcld, cl->fsfcld, cl->cicewp, cl->cliqwp,
                                   swp ? cl->reice : nullptr, swp ? cl->reliq : nullptr};
            const double **dst[] = {&cld.cldfr, &cld.taucld, &cld.ssacld, &cld.asmcld, &cld.fsfcld, &cld.cicewp,
                                    &cld.cliqwp, &cld.reice, &cld.reliq};
            const size_t len[] = {n2, 14 * n2, 14 * n2, 14 * n2, 14 * n2, n2, n2, n2, n2};
            for (int i = 0; i < 9; ++i) {
                if (!src[i]) continue;
                ck |= hipMemcpyAsync(p, src[i], len[i] * sizeof(double), hipMemcpyHostToDevice, st);
                *dst[i] = p;
                p += len[i];
            }
        }
        out.uflx = p; p += n3;
        out.dflx = p; p += n3;
        out.uflxc = p; p += n3;
        out.dflxc = p; p += n3;
        out.hr = p; p += n2;
        out.hrc = p; p += n2;
    } else {
        in.play = play; in.plev = plev; in.tlay = tlay; in.tlev = tlev; in.tsfc = tsfc;
        in.h2o = h2ovmr; in.o3 = o3vmr; in.co2 = co2vmr; in.ch4 = ch4vmr; in.n2o = n2ovmr; in.o2 = o2vmr;
        in.asdir = asdir; in.asdif = asdif; in.aldir = aldir; in.aldif = aldif; in.coszen = coszen;
        out.uflx = swuflx; out.dflx = swdflx; out.hr = swhr; out.uflxc = swuflxc; out.dflxc = swdflxc; out.hrc = swhrc;
    }
    SwDevTables t;
    t.desc = s->sw.tab;
    t.bands = (const SwBand *)s->swd.bands;
    t.tab = mixed ? s->swd.tab32 : s->swd.tab64;
    t.preflog = s->swd.preflog;
    t.tref = s->swd.tref;
    t.units = s->swd.units;
    t.heatfac = s->sw.heatfac;
    if (ck) return fail(RRTMG_HIP_EDEVICE, "SW: host-to-device copy failed");
    hipError_t e = hipSuccess;
    if (mcica) {
        // the McICA driver's inatm reads h2o and O3 as volume mixing ratios
        // and floors the zenith cosine at 1e-10 (rrtmg_sw_rad.f90:571-572,
        // 1015, 1021)
        in.vmr = 1;
        e = launch_sw_coszen_floor(ncol, in.coszen, (double *)(wb + o_csz), st);
        in.coszen = (const double *)(wb + o_csz);
        if (e == hipSuccess && mwp) {
            // water paths: cldprmc_sw's final optics per subcolumn, then the
            // McICA transfer on them
            double *o3b = (double *)(wb + o_mwo);
            int *mbad = (int *)(wb + o_mwb);
            e = launch_sw_mcica_wp(in, cld, s->swd.cld, o3b, o3b + 112 * n2, o3b + 224 * n2, mbad, st);
            cld.taucmc = o3b;
            cld.ssacmc = o3b + 112 * n2;
            cld.asmcmc = o3b + 224 * n2;
            cld.fsfcmc = nullptr;
            cld.scaled = 1;
            if (e == hipSuccess && (ptr_kind == RRTMG_HIP_HOST || g_checks)) {
                int nb = 0;
                e = hipMemcpyAsync(&nb, mbad, sizeof(int), hipMemcpyDeviceToHost, st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e == hipSuccess && nb > 0)
                    return fail(RRTMG_HIP_EINVAL, "cloud water-path optics outside their parameterisation's range "
                                                  "(an effective radius, or optics out of [0, 1]) in a cloudy "
                                                  "subcolumn (the reference stops)");
            }
        }
        if (e == hipSuccess) e = launch_sw_cloud_flags(in, cld, cw, st);
        if (e == hipSuccess) e = launch_sw_cloud(in, cld, out, w, cw, t, g_prec, st, s->events(0));
    } else if (cl) {
        int *wpbad = nullptr;
        if (swp) {
            // water paths: cldprop_sw's final optics, then the cloudy path on them
            double *o3b = (double *)(wb + o_wpo);
            wpbad = (int *)(wb + o_wpb);
            e = launch_sw_cloud_wp(in, cld, s->swd.cld, o3b, o3b + 14 * n2, o3b + 28 * n2, wpbad, st);
            cld.taucld = o3b;
            cld.ssacld = o3b + 14 * n2;
            cld.asmcld = o3b + 28 * n2;
            cld.fsfcld = nullptr;
            cld.scaled = 1;
            if (e == hipSuccess && (ptr_kind == RRTMG_HIP_HOST || g_checks)) {
                int nb = 0;
                e = hipMemcpyAsync(&nb, wpbad, sizeof(int), hipMemcpyDeviceToHost, st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e == hipSuccess && nb > 0)
                    return fail(RRTMG_HIP_EINVAL, "cloud water-path optics outside their parameterisation's range "
                                                  "(an effective radius, or optics out of [0, 1]) in a cloudy layer "
                                                  "(the reference stops)");
            }
        }
        if (e == hipSuccess) e = launch_sw_cloud_flags(in, cld, cw, st);
        if (e == hipSuccess && ptr_kind == RRTMG_HIP_DEVICE && g_checks) {
            // the nomcica driver stops on partial cloud in a day column
            // (rrtmg_sw_rad.nomcica.f90:535-539); the host-pointer path
            // checked before staging
            int np = 0;
            e = hipMemcpyAsync(&np, cw.partial, sizeof(int), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e == hipSuccess && np > 0)
                return fail(RRTMG_HIP_EINVAL, "partial cloud (fraction in (1e-6, 1 - 1e-6)) in a day column: the "
                                              "SW path without McICA takes overcast or clear layers only");
        }
        if (e == hipSuccess) e = launch_sw_cloud(in, cld, out, w, cw, t, g_prec, st, s->events(0));
    } else {
        e = launch_sw(in, out, w, t, g_prec, st, s->events(0));
    }
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("SW kernel launch: ") + hipGetErrorString(e));
    if (!cl) s->sw_w = w;
    s->sw_last_ncol = ncol;
    s->sw_last_nlay = nlay;
    if (ptr_kind == RRTMG_HIP_HOST) {
        ck |= hipMemcpyAsync(swuflx, out.uflx, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(swdflx, out.dflx, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(swuflxc, out.uflxc, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(swdflxc, out.dflxc, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(swhr, out.hr, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(swhrc, out.hrc, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
        e = hipStreamSynchronize(st);
        if (e != hipSuccess || ck) return fail(RRTMG_HIP_EDEVICE, std::string("SW: ") + hipGetErrorString(e));
    }
    return 0;
}

}  // namespace rrtmg

extern "C" {

int rrtmg_hip_sw(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev, const double *tlay,
                 const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                 const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                 const double *asdir, const double *asdif, const double *aldir, const double *aldif,
                 const double *coszen, double adjes, int dyofyr, double scon, double *swuflx, double *swdflx,
                 double *swhr, double *swuflxc, double *swdflxc, double *swhrc, int ptr_kind, void *stream)
{
    return sw_forward(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx, swdflx, swhr, swuflxc,
                      swdflxc, swhrc, ptr_kind, stream, nullptr);
}

// rrtmg_hip_lw, rrtmg_hip_lw_cloud and rrtmg_hip_lw_mcica: cl NULL or
// cl->icld == 0 runs the clear-sky kernels, otherwise the cloud kernels (the
// cloud arrays already checked by the caller).  vmr: water vapour and ozone
// arrive as volume mixing ratios (the McICA driver's inatm) rather than as
// MiMA's mass mixing ratios.
static int lw_forward(int ncol, int nlay, int icld, int idrv, const double *play, const double *plev,
                      const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                      const double *o3vmr, const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                      const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                      const double *cfc22vmr, const double *ccl4vmr, const double *emis, const double *tauaer,
                      const LwClouds *cl, double *uflx, double *dflx, double *hr, double *uflxc, double *dflxc,
                      double *hrc, int ptr_kind, void *stream, bool vmr = false)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->lwd.ready) return fail(RRTMG_HIP_ENOINIT, "rrtmg_hip_lw called before rrtmg_hip_lw_init");
    const bool cloudy = cl && cl->icld != 0;
    if (icld != 0 && !cloudy) return fail(RRTMG_HIP_ENOTSUP, "icld != 0: rrtmg_hip_lw is clear sky; clouds take rrtmg_hip_lw_cloud");
    if (cloudy && g_prec == RRTMG_HIP_PREC_FP64_SMOOTH)
        return fail(RRTMG_HIP_ENOTSUP, "the cloudy path has no smooth (FP64_SMOOTH) mode");
    const bool wp = cloudy && cl->icld != LW_MCICA && cl->inflag != 0;
    const bool mwp = cloudy && cl->icld == LW_MCICA && cl->inflag == 2;
    if ((wp || mwp) && !s->lwd.cld)
        return fail(RRTMG_HIP_ENOTSUP, "inflglw != 0 needs the cloud optics table rrtmg_lw_cld.tbl beside "
                                       "rrtmg_lw.tbl at rrtmg_hip_lw_init");
    if (idrv != 0) return fail(RRTMG_HIP_ENOTSUP, "idrv != 0: dF/dTs output is not implemented");
    int rc = check_common(ncol, nlay, plev, ptr_kind);
    if (rc) return rc;
    const size_t n2 = (size_t)ncol * nlay, n3 = (size_t)ncol * (nlay + 1);
    const double *ins2[] = {play, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr};
    const char *nm2[] = {"play", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
                         "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr"};
    double *outs[] = {uflx, dflx, hr, uflxc, dflxc, hrc};
    for (double *o : outs)
        if (!o) return fail(RRTMG_HIP_EINVAL, "null output array");
    if (ptr_kind == RRTMG_HIP_HOST) {
        for (int i = 0; i < 12; ++i)
            if ((rc = check_host_array(ins2[i], n2, nm2[i]))) return rc;
        if ((rc = check_host_array(plev, n3, "plev")) || (rc = check_host_array(tlev, n3, "tlev")) ||
            (rc = check_host_array(tsfc, ncol, "tsfc")) || (rc = check_host_array(emis, (size_t)ncol * 16, "emis")))
            return rc;
        if (tauaer && (rc = check_host_array(tauaer, n2 * 16, "tauaer"))) return rc;
    } else {
        for (auto p : ins2)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        if (!plev || !tlev || !tsfc || !emis) return fail(RRTMG_HIP_EINVAL, "null input array");
    }
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    if (ptr_kind == RRTMG_HIP_DEVICE && g_checks) {
        std::vector<NamedArray> a;
        for (int i = 0; i < 12; ++i) a.push_back({nm2[i], ins2[i], n2});
        a.push_back({"plev", plev, n3});
        a.push_back({"tlev", tlev, n3});
        a.push_back({"tsfc", tsfc, (size_t)ncol});
        a.push_back({"emis", emis, (size_t)ncol * 16});
        if (tauaer) a.push_back({"tauaer", tauaer, n2 * 16});
        if (cloudy && cl->icld == LW_MCICA) {
            a.push_back({"cldfmcl", cl->cldfmc, n2 * 140});
            a.push_back({"taucmcl", cl->taucmc, n2 * 140});
            if (mwp && cl->ciwpmc) a.push_back({"ciwpmcl", cl->ciwpmc, n2 * 140});
            if (mwp && cl->clwpmc) a.push_back({"clwpmcl", cl->clwpmc, n2 * 140});
            if (mwp && cl->reice) a.push_back({"reicmcl", cl->reice, n2});
            if (mwp && cl->reliq) a.push_back({"relqmcl", cl->reliq, n2});
        } else if (cloudy) {
            a.push_back({"cldfr", cl->cldfr, n2});
            if (cl->taucld) a.push_back({"taucld", cl->taucld, n2 * 16});
            if (wp && cl->reice) a.push_back({"reice", cl->reice, n2});
            if (wp && cl->reliq) a.push_back({"reliq", cl->reliq, n2});
            if (cl->cicewp) a.push_back({"cicewp", cl->cicewp, n2});
            if (cl->cliqwp) a.push_back({"cliqwp", cl->cliqwp, n2});
        }
        if ((rc = device_checks(W, st, a, plev, ncol, nlay))) return rc;
    }
    const int mixed = g_prec == RRTMG_HIP_PREC_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    if ((rc = check_lds(lw_fwd_lds_bytes(nlay, rs)))) return rc;
    int ck = 0;
    size_t off = 0;
    const size_t o_coef = off; off = align_up(off + rs * LWF_STRIDE * n2);
    const size_t o_tlev = off; off = align_up(off + rs * n3);
    const size_t o_ipk = off; off = align_up(off + sizeof(int) * n2);
    const size_t o_ilev = off; off = align_up(off + sizeof(int) * n3);
    const size_t o_lt = off; off = align_up(off + sizeof(int) * ncol);
    const size_t o_cs = off; off = align_up(off + sizeof(double) * 3 * ncol);
    const size_t o_bf = off; off = align_up(off + rs * 16 * 2 * n3);
    const int chunk = vjp_chunk((ncol + TILE_COLS - 1) / TILE_COLS * s->lw.tab.nwaves, lw_fwd_state_bytes(nlay, rs));
    const bool lds = mixed && s->lwd.lds;
    const size_t ds_bytes = lds ? (size_t)lw_lds_grid(s->lwd.plan, ncol, s->lw.tab.nwaves) * LW_WGW *
                                      lw_lds_state_bytes(nlay)
                                : (size_t)chunk * lw_fwd_state_bytes(nlay, rs);
    const size_t o_ds = off; off = align_up(off + ds_bytes);
    const size_t o_it = off; off = align_up(off + sizeof(int));
    // cloud path: flags, overlap records, four-stream band fluxes, scratch rows
    const int cchunk = vjp_chunk((ncol + TILE_COLS - 1) / TILE_COLS * s->lw.tab.nwaves, lw_cloud_row_bytes(nlay, rs));
    const size_t o_cfl = off; if (cloudy) off = align_up(off + sizeof(int) * n2);
    const size_t o_cfa = off; if (cloudy) off = align_up(off + rs * LWC_STRIDE * n2);
    const size_t o_cbf = off; if (cloudy) off = align_up(off + rs * 16 * 4 * n3);
    const size_t crows = lds ? (size_t)lw_lds_grid(s->lwd.plan, ncol, s->lw.tab.nwaves) * LW_WGW : (size_t)cchunk;
    const size_t o_crw = off; if (cloudy) off = align_up(off + crows * lw_cloud_row_bytes(nlay, rs));
    const size_t o_cod = off; if (wp || mwp) off = align_up(off + sizeof(double) * (mwp ? 140 : 16) * n2);
    const size_t o_cbd = off; if (wp || mwp) off = align_up(off + sizeof(int));
    if ((rc = W.lw_work.ensure(off))) return fail(rc, "device workspace allocation failed");
    char *wb = (char *)W.lw_work.p;
    LwWork w;
    w.coef = wb + o_coef;
    w.tlev = wb + o_tlev;
    w.ipk = (int *)(wb + o_ipk);
    w.ilev = (int *)(wb + o_ilev);
    w.laytrop = (int *)(wb + o_lt);
    w.colscal = (double *)(wb + o_cs);
    w.bflux = wb + o_bf;
    w.dstate = wb + o_ds;
    w.chunk = chunk;
    w.items = (int *)(wb + o_it);
    LwCloudWork cw;
    LwClouds cld{};
    if (cloudy) {
        cw.flags = (int *)(wb + o_cfl);
        cw.fac = wb + o_cfa;
        cw.bflux = wb + o_cbf;
        cw.rows = wb + o_crw;
        cw.chunk = cchunk;
        cld = *cl;
        if (wp || mwp) {
            cw.od = (double *)(wb + o_cod);
            cw.bad = (int *)(wb + o_cbd);
            cw.cldtab = s->lwd.cld;
        }
    }

    ColumnInputs in{};
    in.ncol = ncol;
    in.nlay = nlay;
    in.vmr = vmr ? 1 : 0;
    LwOutputs out{};
    if (ptr_kind == RRTMG_HIP_HOST) {
        const size_t ncl = !cloudy ? 0
                           : cl->icld == LW_MCICA ? 280 * n2 + (mwp && cl->ciwpmc ? 140 * n2 : 0) +
                                                        (mwp && cl->clwpmc ? 140 * n2 : 0) +
                                                        (mwp && cl->reice ? n2 : 0) + (mwp && cl->reliq ? n2 : 0)
                                                  : n2 + (cl->taucld ? 16 * n2 : 0) + (cl->cicewp ? n2 : 0) +
                                                        (cl->cliqwp ? n2 : 0) + (wp && cl->reice ? n2 : 0) +
                                                        (wp && cl->reliq ? n2 : 0);
        const size_t need = sizeof(double) * (12 * n2 + 2 * n3 + (size_t)ncol * 17 + (tauaer ? 16 * n2 : 0) +
                                              ncl + 4 * n3 + 2 * n2);
        if ((rc = W.lw_stage.ensure(need))) return fail(rc, "device staging allocation failed");
        double *p = (double *)W.lw_stage.p;
        const double **dst2[] = {&in.play, &in.tlay, &in.h2o, &in.o3, &in.co2, &in.ch4,
                                 &in.n2o, &in.o2, &in.cfc11, &in.cfc12, &in.cfc22, &in.ccl4};
        for (int i = 0; i < 12; ++i) {
            ck |= hipMemcpyAsync(p, ins2[i], n2 * sizeof(double), hipMemcpyHostToDevice, st);
            *dst2[i] = p;
            p += n2;
        }
        ck |= hipMemcpyAsync(p, plev, n3 * sizeof(double), hipMemcpyHostToDevice, st);
        in.plev = p;
        p += n3;
        ck |= hipMemcpyAsync(p, tlev, n3 * sizeof(double), hipMemcpyHostToDevice, st);
        in.tlev = p;
        p += n3;
        ck |= hipMemcpyAsync(p, tsfc, ncol * sizeof(double), hipMemcpyHostToDevice, st);
        in.tsfc = p;
        p += ncol;
        ck |= hipMemcpyAsync(p, emis, (size_t)ncol * 16 * sizeof(double), hipMemcpyHostToDevice, st);
        in.emis = p;
        p += (size_t)ncol * 16;
        if (tauaer) {
            ck |= hipMemcpyAsync(p, tauaer, n2 * 16 * sizeof(double), hipMemcpyHostToDevice, st);
            in.tauaer = p;
            p += n2 * 16;
        }
        if (cloudy) {
            const bool mc = cl->icld == LW_MCICA;
            const double *src[] = {mc ? cl->cldfmc : cl->cldfr, mc ? cl->taucmc : cl->taucld,
                                   mc ? nullptr : cl->cicewp, mc ? nullptr : cl->cliqwp,
                                   wp || mwp ? cl->reice : nullptr, wp || mwp ? cl->reliq : nullptr,
                                   mwp ? cl->ciwpmc : nullptr, mwp ? cl->clwpmc : nullptr};
            const double **dst[] = {mc ? &cld.cldfmc : &cld.cldfr, mc ? &cld.taucmc : &cld.taucld, &cld.cicewp,
                                    &cld.cliqwp, &cld.reice, &cld.reliq, &cld.ciwpmc, &cld.clwpmc};
            const size_t len[] = {mc ? 140 * n2 : n2, mc ? 140 * n2 : 16 * n2, n2, n2, n2, n2, 140 * n2, 140 * n2};
            for (int i = 0; i < 8; ++i) {
                if (!src[i]) continue;
                ck |= hipMemcpyAsync(p, src[i], len[i] * sizeof(double), hipMemcpyHostToDevice, st);
                *dst[i] = p;
                p += len[i];
            }
        }
        out.uflx = p; p += n3;
        out.dflx = p; p += n3;
        out.uflxc = p; p += n3;
        out.dflxc = p; p += n3;
        out.hr = p; p += n2;
        out.hrc = p; p += n2;
    } else {
        in.play = play; in.plev = plev; in.tlay = tlay; in.tlev = tlev; in.tsfc = tsfc;
        in.h2o = h2ovmr; in.o3 = o3vmr; in.co2 = co2vmr; in.ch4 = ch4vmr; in.n2o = n2ovmr; in.o2 = o2vmr;
        in.cfc11 = cfc11vmr; in.cfc12 = cfc12vmr; in.cfc22 = cfc22vmr; in.ccl4 = ccl4vmr;
        in.emis = emis; in.tauaer = tauaer;
        out.uflx = uflx; out.dflx = dflx; out.hr = hr; out.uflxc = uflxc; out.dflxc = dflxc; out.hrc = hrc;
    }
    LwDevTables t;
    t.desc = s->lw.tab;
    t.bands = (const LwBand *)s->lwd.bands;
    t.tab = mixed ? s->lwd.tab32 : s->lwd.tab64;
    t.cst = s->lwd.cst;
    t.units = s->lwd.units;
    t.heatfac = s->lw.cst.heatfac;
    t.bands_lds = s->lwd.bands_lds;
    t.plan = s->lwd.plan;
    t.lds = lds;
    if (ck) return fail(RRTMG_HIP_EDEVICE, "LW: host-to-device copy failed");
    hipError_t e = cloudy ? launch_lw_cloud(in, cld, out, w, cw, t, g_prec, st, s->events(1))
                          : launch_lw(in, out, w, t, g_prec, st, s->events(1));
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("LW kernel launch: ") + hipGetErrorString(e));
    if ((wp || mwp) && ptr_kind == RRTMG_HIP_DEVICE && g_checks) {
        // the radii cldprop stops on (rrtmg_lw_cldprop.f90:184-259); the
        // host-pointer path checked before staging
        int nb = 0;
        e = hipMemcpyAsync(&nb, cw.bad, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("LW: ") + hipGetErrorString(e));
        if (nb)
            return fail(RRTMG_HIP_EINVAL, "an ice or liquid effective radius outside the range of its "
                                          "parameterisation in a cloudy layer (the reference stops); outputs "
                                          "written from clamped radii");
    }
    if (!cloudy) s->lw_w = w;
    s->lw_last_ncol = ncol;
    s->lw_last_nlay = nlay;
    if (ptr_kind == RRTMG_HIP_HOST) {
        ck |= hipMemcpyAsync(uflx, out.uflx, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(dflx, out.dflx, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(uflxc, out.uflxc, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(dflxc, out.dflxc, n3 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(hr, out.hr, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
        ck |= hipMemcpyAsync(hrc, out.hrc, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
        e = hipStreamSynchronize(st);
        if (e != hipSuccess || ck) return fail(RRTMG_HIP_EDEVICE, std::string("LW: ") + hipGetErrorString(e));
    }
    return 0;
}

int rrtmg_hip_sw_cloud(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev,
                       const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                       const double *o3vmr, const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr, const double *asdir, const double *asdif, const double *aldir,
                       const double *aldif, const double *coszen, double adjes, int dyofyr, double scon,
                       int inflgsw, int iceflgsw, int liqflgsw, const double *cldfr, const double *taucld,
                       const double *ssacld, const double *asmcld, const double *fsfcld, const double *cicewp,
                       const double *cliqwp, const double *reice, const double *reliq, const double *tauaer,
                       const double *ssaaer, const double *asmaer, const double *ecaer, double *swuflx,
                       double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc, int ptr_kind,
                       void *stream)
{
    (void)tauaer;
    (void)ssaaer;
    (void)asmaer;
    (void)ecaer;
    // rrtmg_sw_rad.nomcica.f90:471: icld outside 0..3 is taken as 2
    if (icld < 0 || icld > 3) icld = 2;
    if (icld == 0)
        return sw_forward(ncol, nlay, 0, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                          o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx, swdflx, swhr,
                          swuflxc, swdflxc, swhrc, ptr_kind, stream, nullptr);
    if (inflgsw < 0 || inflgsw > 2) return fail(RRTMG_HIP_EINVAL, "inflgsw must be 0, 1 or 2");
    // cldprop_sw reads ice optics for iceflgsw 1-3 and liquid for liqflgsw
    // 1 only (rrtmg_sw_cldprop.f90:169-287)
    if (inflgsw == 2 && (iceflgsw < 1 || iceflgsw > 3 || liqflgsw != 1))
        return fail(RRTMG_HIP_EINVAL, "iceflgsw must be 1..3 and liqflgsw 1 with inflgsw = 2");
    if (ncol > 0 && nlay >= 2 && nlay <= RRTMG_HIP_MAX_NLAY) {
        const size_t n2 = (size_t)ncol * nlay;
        int rc;
        if (ptr_kind == RRTMG_HIP_HOST) {
            const double *a[] = {taucld, ssacld, asmcld, fsfcld};
            const char *nm[] = {"taucld", "ssacld", "asmcld", "fsfcld"};
            if ((rc = check_host_array(cldfr, n2, "cldfr"))) return rc;
            for (int i = 0; i < 4; ++i)
                if ((inflgsw == 0 || (i == 0 && a[i])) && (rc = check_host_array(a[i], 14 * n2, nm[i]))) return rc;
            if (inflgsw == 2 && reice && (rc = check_host_array(reice, n2, "reice"))) return rc;
            if (inflgsw == 2 && reliq && (rc = check_host_array(reliq, n2, "reliq"))) return rc;
            if (cicewp && (rc = check_host_array(cicewp, n2, "cicewp"))) return rc;
            if (cliqwp && (rc = check_host_array(cliqwp, n2, "cliqwp"))) return rc;
            if ((rc = check_host_array(coszen, ncol, "coszen"))) return rc;
            // rrtmg_sw_rad.nomcica.f90:501, 535-539: day columns take
            // overcast or clear layers only
            for (size_t i = 0; i < n2; ++i) {
                const double f = cldfr[i];
                if (coszen[i % ncol] >= 1.e-10 && f > 1.e-6 && f < 1.0 - 1.e-6)
                    return fail(RRTMG_HIP_EINVAL, "partial cloud (fraction in (1e-6, 1 - 1e-6)) in a day column: the "
                                                  "SW path without McICA takes overcast or clear layers only");
            }
        } else if (!cldfr || (inflgsw == 0 && (!taucld || !ssacld || !asmcld || !fsfcld))) {
            return fail(RRTMG_HIP_EINVAL, "null cloud array");
        }
    }
    SwClouds cl{cldfr, taucld, inflgsw == 0 ? ssacld : nullptr,
                inflgsw == 0 ? asmcld : nullptr, inflgsw == 0 ? fsfcld : nullptr, cicewp, cliqwp};
    cl.inflag = inflgsw;
    cl.iceflag = iceflgsw;
    cl.liqflag = liqflgsw;
    cl.reice = reice;
    cl.reliq = reliq;
    return sw_forward(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx, swdflx, swhr, swuflxc,
                      swdflxc, swhrc, ptr_kind, stream, nullptr, &cl);
}

int rrtmg_hip_sw_mcica(int ncol, int nlay, int icld, int iaer, const double *play, const double *plev,
                       const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                       const double *o3vmr, const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr, const double *asdir, const double *asdif, const double *aldir,
                       const double *aldif, const double *coszen, double adjes, int dyofyr, double scon,
                       int inflgsw, int iceflgsw, int liqflgsw, const double *cldfmcl, const double *taucmcl,
                       const double *ssacmcl, const double *asmcmcl, const double *fsfcmcl, const double *ciwpmcl,
                       const double *clwpmcl, const double *reicmcl, const double *relqmcl, const double *tauaer,
                       const double *ssaaer, const double *asmaer, const double *ecaer, double *swuflx,
                       double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc, int ptr_kind,
                       void *stream)
{
    (void)tauaer;
    (void)ssaaer;
    (void)asmaer;
    (void)ecaer;
    // rrtmg_sw_rad.f90:500: icld outside 0..3 is taken as 2
    if (icld < 0 || icld > 3) icld = 2;
    SwClouds cl{};
    cl.mcica = 1;
    // icld = 0: no subcolumns (rrtmg_sw_rad.f90:591-597, 954)
    if (icld != 0) {
        // rrtmg_sw_cldprmc.f90:146-147: inflag 1 stops under McICA
        if (inflgsw != 0 && inflgsw != 2)
            return fail(RRTMG_HIP_EINVAL, "inflgsw must be 0 or 2 with McICA (the reference stops on 1)");
        if (inflgsw == 2 && (iceflgsw < 1 || iceflgsw > 3))
            return fail(RRTMG_HIP_EINVAL, "iceflgsw must be 1, 2 or 3 with inflgsw = 2");
        if (inflgsw == 2 && liqflgsw != 1)
            return fail(RRTMG_HIP_EINVAL, "liqflgsw must be 1 with inflgsw = 2");
        // inflgsw 2 reads fsfcmcl nowhere (cldprmc_sw replaces the optics it scales)
        const double *arr[] = {cldfmcl, taucmcl, ssacmcl, asmcmcl, inflgsw == 2 ? taucmcl : fsfcmcl};
        const char *nm[] = {"cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl"};
        if (ncol > 0 && nlay >= 2 && nlay <= RRTMG_HIP_MAX_NLAY) {
            const size_t n = (size_t)ncol * nlay * 112;
            int rc;
            if (ptr_kind == RRTMG_HIP_HOST) {
                for (int i = 0; i < 5; ++i)
                    if ((rc = check_host_array(arr[i], n, nm[i]))) return rc;
                if (ciwpmcl && (rc = check_host_array(ciwpmcl, n, "ciwpmcl"))) return rc;
                if (clwpmcl && (rc = check_host_array(clwpmcl, n, "clwpmcl"))) return rc;
                if (inflgsw == 2 && reicmcl && (rc = check_host_array(reicmcl, n / 112, "reicmcl"))) return rc;
                if (inflgsw == 2 && relqmcl && (rc = check_host_array(relqmcl, n / 112, "relqmcl"))) return rc;
            } else {
                for (int i = 0; i < 5; ++i)
                    if (!arr[i]) return fail(RRTMG_HIP_EINVAL, "null cloud array");
            }
        }
        cl.cldfmc = cldfmcl;
        cl.taucmc = taucmcl;
        cl.ssacmc = ssacmcl;
        cl.asmcmc = asmcmcl;
        cl.fsfcmc = fsfcmcl;
        cl.ciwpmc = ciwpmcl;
        cl.clwpmc = clwpmcl;
        cl.inflag = inflgsw;
        cl.iceflag = iceflgsw;
        cl.liqflag = liqflgsw;
        if (inflgsw == 2) {
            cl.reice = reicmcl;
            cl.reliq = relqmcl;
        }
    }
    return sw_forward(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx, swdflx, swhr, swuflxc,
                      swdflxc, swhrc, ptr_kind, stream, nullptr, &cl);
}

int rrtmg_hip_lw(int ncol, int nlay, int icld, int idrv, const double *play, const double *plev, const double *tlay,
                 const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                 const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                 const double *cfc11vmr, const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                 const double *emis, const double *tauaer, double *uflx, double *dflx, double *hr, double *uflxc,
                 double *dflxc, double *hrc, int ptr_kind, void *stream)
{
    return lw_forward(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, nullptr, uflx, dflx, hr, uflxc,
                      dflxc, hrc, ptr_kind, stream);
}

int rrtmg_hip_lw_cloud(int ncol, int nlay, int icld, int idrv, const double *play, const double *plev,
                       const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                       const double *o3vmr, const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                       const double *cfc22vmr, const double *ccl4vmr, const double *emis, int inflglw,
                       int iceflglw, int liqflglw, const double *cldfr, const double *taucld,
                       const double *cicewp, const double *cliqwp, const double *reice, const double *reliq,
                       const double *tauaer, double *uflx, double *dflx, double *hr, double *uflxc, double *dflxc,
                       double *hrc, int ptr_kind, void *stream)
{
    // rrtmg_lw_rad.nomcica.f90:437: icld outside 0..3 selects maximum-random overlap
    if (icld < 0 || icld > 3) icld = 2;
    if (icld == 0)
        return lw_forward(ncol, nlay, 0, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                          n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, nullptr, uflx, dflx,
                          hr, uflxc, dflxc, hrc, ptr_kind, stream);
    if (inflglw < 0 || inflglw > 2) return fail(RRTMG_HIP_EINVAL, "inflglw must be 0, 1 or 2");
    if (inflglw == 2 && (iceflglw < 0 || iceflglw > 3 || liqflglw < 0 || liqflglw > 1))
        return fail(RRTMG_HIP_EINVAL, "iceflglw must be 0..3 and liqflglw 0..1 with inflglw = 2");
    if (ncol > 0 && nlay >= 2 && nlay <= RRTMG_HIP_MAX_NLAY) {
        const size_t n2 = (size_t)ncol * nlay;
        int rc;
        if (ptr_kind == RRTMG_HIP_HOST) {
            if ((rc = check_host_array(cldfr, n2, "cldfr"))) return rc;
            if ((inflglw == 0 || taucld) && (rc = check_host_array(taucld, n2 * 16, "taucld"))) return rc;
            if (cicewp && (rc = check_host_array(cicewp, n2, "cicewp"))) return rc;
            if (cliqwp && (rc = check_host_array(cliqwp, n2, "cliqwp"))) return rc;
            if (inflglw == 2 && reice && (rc = check_host_array(reice, n2, "reice"))) return rc;
            if (inflglw == 2 && reliq && (rc = check_host_array(reliq, n2, "reliq"))) return rc;
            if (inflglw == 2) {
                // the radii cldprop stops on (rrtmg_lw_cldprop.f90:184-259),
                // in the layers that pass its cloud test
                for (size_t i = 0; i < n2; ++i) {
                    const double ci = cicewp ? cicewp[i] : 0.0, cw = cliqwp ? cliqwp[i] : 0.0;
                    double tt = 0.0;
                    for (int b = 0; b < 16 && taucld; ++b) tt = tt + taucld[i * 16 + b];
                    if (!(cldfr[i] >= 1.e-20 && (ci + cw >= 1.e-20 || tt >= 1.e-20))) continue;
                    const double ri = reice ? reice[i] : 0.0, rl = reliq ? reliq[i] : 0.0;
                    const bool bi = ci != 0.0 && (iceflglw == 0   ? ri < 10.0
                                                  : iceflglw == 1 ? ri < 13.0 || ri > 130.0
                                                  : iceflglw == 2 ? ri < 5.0 || ri > 131.0
                                                                  : ri < 5.0 || ri > 140.0);
                    const bool bl = cw != 0.0 && liqflglw == 1 && (rl < 2.5 || rl > 60.0);
                    if (bi || bl)
                        return fail(RRTMG_HIP_EINVAL, "an ice or liquid effective radius outside the range of its "
                                                      "parameterisation in a cloudy layer (the reference stops)");
                }
            }
        } else if (!cldfr || (inflglw == 0 && !taucld)) {
            return fail(RRTMG_HIP_EINVAL, "null cloud array");
        }
    }
    LwClouds cl{icld == 1 ? 1 : 2, cldfr, taucld, cicewp, cliqwp};
    cl.inflag = inflglw;
    cl.iceflag = iceflglw;
    cl.liqflag = liqflglw;
    cl.reice = reice;
    cl.reliq = reliq;
    return lw_forward(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, &cl, uflx, dflx, hr, uflxc, dflxc,
                      hrc, ptr_kind, stream);
}

int rrtmg_hip_lw_mcica(int ncol, int nlay, int icld, int idrv, const double *play, const double *plev,
                       const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                       const double *o3vmr, const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr, const double *cfc11vmr, const double *cfc12vmr,
                       const double *cfc22vmr, const double *ccl4vmr, const double *emis, int inflglw,
                       int iceflglw, int liqflglw, const double *cldfmcl, const double *taucmcl,
                       const double *ciwpmcl, const double *clwpmcl, const double *reicmcl,
                       const double *relqmcl, const double *tauaer, double *uflx, double *dflx, double *hr,
                       double *uflxc, double *dflxc, double *hrc, int ptr_kind, void *stream)
{
    // rrtmg_lw_rad.f90:451: icld outside 0..3 is taken as 2
    if (icld < 0 || icld > 3) icld = 2;
    // icld = 0: inatm leaves the subcolumns at zero (rrtmg_lw_rad.f90:743-744, 874)
    if (icld == 0)
        return lw_forward(ncol, nlay, 0, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                          n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, nullptr, uflx, dflx,
                          hr, uflxc, dflxc, hrc, ptr_kind, stream, true);
    // rrtmg_lw_cldprmc.f90:156-158: inflag 1 stops with McICA
    if (inflglw != 0 && inflglw != 2)
        return fail(RRTMG_HIP_EINVAL, "inflglw must be 0 or 2 with McICA (the reference stops on 1)");
    if (inflglw == 2 && (iceflglw < 0 || iceflglw > 3 || liqflglw < 0 || liqflglw > 1))
        return fail(RRTMG_HIP_EINVAL, "iceflglw must be 0..3 and liqflglw 0..1 with inflglw = 2");
    if (ncol > 0 && nlay >= 2 && nlay <= RRTMG_HIP_MAX_NLAY) {
        const size_t n2 = (size_t)ncol * nlay, n = n2 * 140;
        int rc;
        if (ptr_kind == RRTMG_HIP_HOST) {
            if ((rc = check_host_array(cldfmcl, n, "cldfmcl")) || (rc = check_host_array(taucmcl, n, "taucmcl")))
                return rc;
            if (inflglw == 2) {
                if ((ciwpmcl && (rc = check_host_array(ciwpmcl, n, "ciwpmcl"))) ||
                    (clwpmcl && (rc = check_host_array(clwpmcl, n, "clwpmcl"))) ||
                    (reicmcl && (rc = check_host_array(reicmcl, n2, "reicmcl"))) ||
                    (relqmcl && (rc = check_host_array(relqmcl, n2, "relqmcl"))))
                    return rc;
                // the radii cldprmc stops on (rrtmg_lw_cldprmc.f90:163-218),
                // in the subcolumns that pass its cloud test
                for (size_t t = 0; t < n; ++t) {
                    const double ci = ciwpmcl ? ciwpmcl[t] : 0.0, cw = clwpmcl ? clwpmcl[t] : 0.0;
                    if (!(cldfmcl[t] >= 1.e-20 && (ci + cw >= 1.e-20 || taucmcl[t] >= 1.e-20))) continue;
                    const size_t ix = t / 140;
                    const double ri = reicmcl ? reicmcl[ix] : 0.0, rl = relqmcl ? relqmcl[ix] : 0.0;
                    const bool bi = ci != 0.0 && (iceflglw == 0   ? ri < 10.0
                                                  : iceflglw == 1 ? ri < 13.0 || ri > 130.0
                                                  : iceflglw == 2 ? ri < 5.0 || ri > 131.0
                                                                  : ri < 5.0 || ri > 140.0);
                    const bool bl = cw != 0.0 && liqflglw == 1 && (rl < 2.5 || rl > 60.0);
                    if (bi || bl)
                        return fail(RRTMG_HIP_EINVAL, "an ice or liquid effective radius outside the range of its "
                                                      "parameterisation in a cloudy subcolumn (the reference stops)");
                }
            }
        } else if (!cldfmcl || !taucmcl) {
            return fail(RRTMG_HIP_EINVAL, "null cloud array");
        }
    }
    LwClouds cl{};
    cl.icld = LW_MCICA;
    cl.cldfmc = cldfmcl;
    cl.taucmc = taucmcl;
    if (inflglw == 2) {
        cl.inflag = 2;
        cl.iceflag = iceflglw;
        cl.liqflag = liqflglw;
        cl.ciwpmc = ciwpmcl;
        cl.clwpmc = clwpmcl;
        cl.reice = reicmcl;
        cl.reliq = relqmcl;
    }
    return lw_forward(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                      o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, &cl, uflx, dflx, hr, uflxc, dflxc,
                      hrc, ptr_kind, stream, true);
}

// rrtmg_hip_mcica_subcol_lw (nsub 140) and rrtmg_hip_mcica_subcol_sw (nsub
// 112, with the single-scattering optics)
static int mcica_subcol(const char *who, int nsub, int ncol, int nlay, int icld, int permuteseed, int irng,
                        const double *play, const double *cldfrac, const double *ciwp, const double *clwp,
                        const double *rei, const double *rel, const double *tauc, const double *ssac,
                        const double *asmc, const double *fsfc, double *cldfmcl, double *ciwpmcl, double *clwpmcl,
                        double *reicmcl, double *relqmcl, double *taucmcl, double *ssacmcl, double *asmcmcl,
                        double *fsfcmcl, int ptr_kind, void *stream)
{
    const bool sw = nsub == 112;
    const int nband = sw ? 14 : 16;
    // mcica_subcol_gen_lw.f90:119-123 (SW :115-119): clear sky returns at
    // once, icld outside 0..3 stops
    if (icld == 0) return 0;
    if (icld < 0 || icld > 3) return fail(RRTMG_HIP_EINVAL, "invalid icld (0..3)");
    // mcica_subcol_gen_lw.f90:305: any irng != 0 selects the Mersenne Twister
    const bool mt = irng != 0;
    if (ncol <= 0) return fail(RRTMG_HIP_EINVAL, "ncol must be positive");
    // the KISS seeds read the four lowest layers (mcica_subcol_gen_lw.f90:330-333)
    if (nlay < (mt ? 2 : 4) || nlay > RRTMG_HIP_MAX_NLAY)
        return fail(RRTMG_HIP_EINVAL, mt ? "nlay out of range [2, 200]" : "nlay out of range [4, 200] with KISS");
    if (ptr_kind != RRTMG_HIP_HOST && ptr_kind != RRTMG_HIP_DEVICE) return fail(RRTMG_HIP_EINVAL, "bad ptr_kind");
    if (!play || !cldfrac || !tauc || !cldfmcl || !taucmcl) return fail(RRTMG_HIP_EINVAL, "null array argument");
    if (sw && (!ssac || !asmc || !fsfc || !ssacmcl || !asmcmcl || !fsfcmcl))
        return fail(RRTMG_HIP_EINVAL, "null array argument");
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->xs_jump) {
        std::vector<uint32_t> xs(KISS_XS_JUMPS * 32);
        kiss_xorshift_jumps(xs.data());
        if (upload(&s->xs_jump, xs.data(), xs.size() * sizeof(uint32_t)))
            return fail(RRTMG_HIP_EDEVICE, std::string(who) + ": uploading the generator's jump tables failed");
    }
    const size_t n2 = (size_t)ncol * nlay, ns = n2 * nsub, nb = n2 * nband;
    const double *ins[] = {play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc};
    const char *nm[] = {"play", "cldfrac", "ciwp", "clwp", "rei", "rel", "tauc", "ssac", "asmc", "fsfc"};
    const size_t len[] = {n2, n2, n2, n2, n2, n2, nb, nb, nb, nb};
    const int nin = sw ? 10 : 7;
    double *outs[] = {cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl};
    const size_t olen[] = {ns, ns, ns, n2, n2, ns, ns, ns, ns};
    const int nout = sw ? 9 : 6;
    int rc;
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    if ((rc = W.flag.ensure(256))) return fail(rc, "device allocation for the input checks failed");
    // the jump constants, then (Mersenne Twister) the call's stream
    const size_t jump_bytes = align_up(140 * sizeof(McicaJump)) +
                              (mt ? align_up(sizeof(uint32_t) * mcica_mt_count(ncol, nlay, icld, nsub)) : 0);
    McicaSubcolArgs a{};
    a.ncol = ncol;
    a.nlay = nlay;
    a.icld = icld;
    a.permuteseed = permuteseed;
    a.nsub = nsub;
    int ck = 0;
    const double **din[] = {&a.play, &a.cldfrac, &a.ciwp, &a.clwp, &a.rei, &a.rel, &a.tauc, &a.ssac, &a.asmc, &a.fsfc};
    double **dout[] = {&a.cldfmcl, &a.ciwpmcl, &a.clwpmcl, &a.reicmcl, &a.relqmcl, &a.taucmcl,
                       &a.ssacmcl, &a.asmcmcl, &a.fsfcmcl};
    if (ptr_kind == RRTMG_HIP_HOST) {
        for (int i = 0; i < nin; ++i)
            if (ins[i] && (rc = check_host_array(ins[i], len[i], nm[i]))) return rc;
        for (int i = 0; i < ncol && !mt; ++i)
            if (play[i] * 1.e2 < play[(size_t)ncol + i] * 1.e2)
                return fail(RRTMG_HIP_EINVAL, "play must not rise from the first layer to the second (the KISS "
                                              "seeds read the four lowest layers)");
        size_t need = 0;
        for (int i = 0; i < nin; ++i) need += ins[i] ? len[i] : 0;
        for (int i = 0; i < nout; ++i) need += outs[i] ? olen[i] : 0;
        if ((rc = W.lw_stage.ensure(jump_bytes + need * sizeof(double))))
            return fail(rc, "device staging allocation failed");
        double *p = (double *)((char *)W.lw_stage.p + jump_bytes);
        for (int i = 0; i < nin; ++i) {
            if (!ins[i]) continue;
            ck |= hipMemcpyAsync(p, ins[i], len[i] * sizeof(double), hipMemcpyHostToDevice, st);
            *din[i] = p;
            p += len[i];
        }
        for (int i = 0; i < nout; ++i) {
            if (!outs[i]) continue;
            *dout[i] = p;
            p += olen[i];
        }
    } else {
        if ((rc = W.lw_stage.ensure(jump_bytes))) return fail(rc, "device scratch allocation failed");
        if (g_checks) {
            std::vector<NamedArray> v;
            for (int i = 0; i < nin; ++i)
                if (ins[i]) v.push_back({nm[i], ins[i], len[i]});
            if ((rc = device_checks(W, st, v, nullptr, ncol, nlay))) return rc;
            a.bad = (int *)W.flag.p + 16;
        }
        for (int i = 0; i < nin; ++i) *din[i] = ins[i];
        for (int i = 0; i < nout; ++i) *dout[i] = outs[i];
    }
    if (ck) return fail(RRTMG_HIP_EDEVICE, std::string(who) + ": host-to-device copy failed");
    if (mt) a.mt = (uint32_t *)((char *)W.lw_stage.p + align_up(140 * sizeof(McicaJump)));
    hipError_t e = launch_mcica_subcol(a, s->xs_jump, (McicaJump *)W.lw_stage.p, st);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string(who) + " launch: " + hipGetErrorString(e));
    if (ptr_kind == RRTMG_HIP_HOST) {
        for (int i = 0; i < nout; ++i)
            if (outs[i]) ck |= hipMemcpyAsync(outs[i], *dout[i], olen[i] * sizeof(double), hipMemcpyDeviceToHost, st);
        e = hipStreamSynchronize(st);
        if (e != hipSuccess || ck) return fail(RRTMG_HIP_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
    } else if (a.bad) {
        int nbad = 0;
        e = hipMemcpyAsync(&nbad, a.bad, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
        if (nbad)
            return fail(RRTMG_HIP_EINVAL, "play must not rise from the first layer to the second (the KISS seeds "
                                          "read the four lowest layers); those columns are left unwritten");
    }
    return 0;
}

int rrtmg_hip_mcica_subcol_lw(int iplon, int ncol, int nlay, int icld, int permuteseed, int irng,
                              const double *play, const double *cldfrac, const double *ciwp, const double *clwp,
                              const double *rei, const double *rel, const double *tauc, double *cldfmcl,
                              double *ciwpmcl, double *clwpmcl, double *reicmcl, double *relqmcl,
                              double *taucmcl, int ptr_kind, void *stream)
{
    (void)iplon;
    return mcica_subcol("McICA LW subcolumns", 140, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp,
                        rei, rel, tauc, nullptr, nullptr, nullptr, cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl,
                        taucmcl, nullptr, nullptr, nullptr, ptr_kind, stream);
}

int rrtmg_hip_mcica_subcol_sw(int iplon, int ncol, int nlay, int icld, int permuteseed, int irng,
                              const double *play, const double *cldfrac, const double *ciwp, const double *clwp,
                              const double *rei, const double *rel, const double *tauc, const double *ssac,
                              const double *asmc, const double *fsfc, double *cldfmcl, double *ciwpmcl,
                              double *clwpmcl, double *reicmcl, double *relqmcl, double *taucmcl, double *ssacmcl,
                              double *asmcmcl, double *fsfcmcl, int ptr_kind, void *stream)
{
    (void)iplon;
    return mcica_subcol("McICA SW subcolumns", 112, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp,
                        rei, rel, tauc, ssac, asmc, fsfc, cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl,
                        ssacmcl, asmcmcl, fsfcmcl, ptr_kind, stream);
}

// backward of rrtmg_sw; with swuflx/swdflx/swhr given, also the forward's
// outputs, formed from the forward that the backward band kernel recomputes
static int sw_vjp_impl(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                       const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *asdir, const double *asdif, const double *aldir, const double *aldif,
                       const double *coszen, double adjes, int dyofyr, double scon, const double *swuflx_bar,
                       const double *swdflx_bar, const double *swhr_bar, double *tlay_bar, double *h2ovmr_bar,
                       double *o3vmr_bar, double *swuflx, double *swdflx, double *swhr, int ptr_kind, void *stream)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->swd.ready) return fail(RRTMG_HIP_ENOINIT, "rrtmg_hip_sw_vjp called before rrtmg_hip_sw_init");
    int rc = check_common(ncol, nlay, plev, ptr_kind);
    if (rc) return rc;
    const size_t n2 = (size_t)ncol * nlay, n3 = (size_t)ncol * (nlay + 1);
    const double *ins2[] = {play, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr};
    const char *nm2[] = {"play", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr"};
    const double *ins1[] = {tsfc, asdir, asdif, aldir, aldif, coszen};
    double *outs[] = {tlay_bar, h2ovmr_bar, o3vmr_bar};
    for (double *o : outs)
        if (!o) return fail(RRTMG_HIP_EINVAL, "null gradient array");
    const bool fwd = swuflx || swdflx || swhr;
    if (fwd && !(swuflx && swdflx && swhr)) return fail(RRTMG_HIP_EINVAL, "null forward output array");
    if (bad(adjes) || bad(scon)) return fail(RRTMG_HIP_EINVAL, "non-finite adjes/scon");
    const bool host = ptr_kind == RRTMG_HIP_HOST;
    if (host) {
        for (int i = 0; i < 8; ++i)
            if ((rc = check_host_array(ins2[i], n2, nm2[i]))) return rc;
        if ((rc = check_host_array(plev, n3, "plev")) || (rc = check_host_array(tlev, n3, "tlev"))) return rc;
        for (int i = 0; i < 6; ++i)
            if ((rc = check_host_array(ins1[i], ncol, "surface/zenith input"))) return rc;
        if ((swuflx_bar && (rc = check_host_array(swuflx_bar, n3, "swuflx_bar"))) ||
            (swdflx_bar && (rc = check_host_array(swdflx_bar, n3, "swdflx_bar"))) ||
            (swhr_bar && (rc = check_host_array(swhr_bar, n2, "swhr_bar"))))
            return rc;
    } else {
        for (auto p : ins2)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        for (auto p : ins1)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        if (!plev || !tlev) return fail(RRTMG_HIP_EINVAL, "null input array");
    }
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    const bool fused = W.fused.open;
    if (fused && host) return fail(RRTMG_HIP_EINVAL, "calls inside a fused session take device pointers");
    if (fused && W.fused.sw.on) return fail(RRTMG_HIP_EINVAL, "one SW backward call per fused session");
    if (!host && g_checks) {
        std::vector<NamedArray> a;
        for (int i = 0; i < 8; ++i) a.push_back({nm2[i], ins2[i], n2});
        a.push_back({"plev", plev, n3});
        a.push_back({"tlev", tlev, n3});
        const char *nm1[] = {"tsfc", "asdir", "asdif", "aldir", "aldif", "coszen"};
        for (int i = 0; i < 6; ++i) a.push_back({nm1[i], ins1[i], (size_t)ncol});
        if (swuflx_bar) a.push_back({"swuflx_bar", swuflx_bar, n3});
        if (swdflx_bar) a.push_back({"swdflx_bar", swdflx_bar, n3});
        if (swhr_bar) a.push_back({"swhr_bar", swhr_bar, n2});
        if ((rc = device_checks(W, st, a, plev, ncol, nlay))) return rc;
    }
    const int mixed = g_prec == RRTMG_HIP_PREC_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    if ((rc = check_lds(sw_vjp_lds_bytes(nlay, rs)))) return rc;
    size_t off = 0;
    const size_t o_dc = off; off = align_up(off + rs * SWF_STRIDE * n2);
    const size_t o_dd = off; off = align_up(off + rs * SW_DSTRIDE * n2);
    const size_t o_ipk = off; off = align_up(off + sizeof(int) * n2);
    const size_t o_lt = off; off = align_up(off + sizeof(int) * ncol);
    const size_t o_lsf = off; off = align_up(off + sizeof(int) * 14 * ncol);
    const size_t o_cot = off; off = align_up(off + rs * 2 * n3);
    const size_t o_gp = off; off = align_up(off + rs * 14 * 3 * n2);
    const int chunk = vjp_chunk((ncol + TILE_COLS - 1) / TILE_COLS * s->sw.tab.nwaves, sw_vjp_state_bytes(nlay, rs));
    const size_t o_ds = off; off = align_up(off + (size_t)chunk * sw_vjp_state_bytes(nlay, rs));
    const size_t o_bf = off; off = align_up(off + (fwd ? rs * 14 * 2 * n3 : 0));
    if ((rc = W.sw_vjp_work.ensure(off))) return fail(rc, "device workspace allocation failed");
    char *wb = (char *)W.sw_vjp_work.p;
    SwVjpWork w;
    w.dcoef = wb + o_dc;
    w.dder = wb + o_dd;
    w.ipk = (int *)(wb + o_ipk);
    w.laytrop = (int *)(wb + o_lt);
    w.laysolfr = (int *)(wb + o_lsf);
    w.cot = wb + o_cot;
    w.gpart = wb + o_gp;
    w.dstate = wb + o_ds;
    w.chunk = chunk;
    w.bflux = fwd ? wb + o_bf : nullptr;
    if (host) {
        const size_t need = 256 * 27 + sizeof(double) * (12 * n2 + 4 * n3 + (size_t)ncol * 6 + (fwd ? n2 + 2 * n3 : 0));
        if ((rc = W.vjp_stage.ensure(need))) return fail(rc, "device staging allocation failed");
    }
    Stager sg{host ? (char *)W.vjp_stage.p : nullptr, host, st};
    ColumnInputs in{};
    in.ncol = ncol;
    in.nlay = nlay;
    in.play = sg.in(play, n2); in.plev = sg.in(plev, n3); in.tlay = sg.in(tlay, n2); in.tlev = sg.in(tlev, n3);
    in.tsfc = sg.in(tsfc, ncol); in.h2o = sg.in(h2ovmr, n2); in.o3 = sg.in(o3vmr, n2); in.co2 = sg.in(co2vmr, n2);
    in.ch4 = sg.in(ch4vmr, n2); in.n2o = sg.in(n2ovmr, n2); in.o2 = sg.in(o2vmr, n2);
    in.asdir = sg.in(asdir, ncol); in.asdif = sg.in(asdif, ncol); in.aldir = sg.in(aldir, ncol);
    in.aldif = sg.in(aldif, ncol); in.coszen = sg.in(coszen, ncol);
    in.adjflux = sw_adjflux(adjes, dyofyr, scon);
    VjpCot cot{sg.in(swuflx_bar, n3), sg.in(swdflx_bar, n3), sg.in(swhr_bar, n2)};
    SwGrad g{sg.out(tlay_bar, n2), sg.out(h2ovmr_bar, n2), sg.out(o3vmr_bar, n2)};
    SwOutputs fo{};
    if (fwd) {  // clear sky: the clear-sky outputs are the same arrays
        fo.uflx = fo.uflxc = sg.out(swuflx, n3);
        fo.dflx = fo.dflxc = sg.out(swdflx, n3);
        fo.hr = fo.hrc = sg.out(swhr, n2);
    }
    if (sg.ck) return fail(RRTMG_HIP_EDEVICE, "SW backward: host-to-device copy failed");
    SwDevTables t;
    t.desc = s->sw.tab;
    t.bands = (const SwBand *)s->swd.bands;
    t.tab = mixed ? s->swd.tab32 : s->swd.tab64;
    t.preflog = s->swd.preflog;
    t.tref = s->swd.tref;
    t.units = s->swd.units;
    t.heatfac = s->sw.heatfac;
    if (fused) {
        const hipError_t e = launch_sw_vjp_prep(in, cot, w, t, g_prec, st);
        if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("SW backward launch: ") + hipGetErrorString(e));
        PendingSw &P = W.fused.sw;
        P.on = true;
        P.fwd = fwd;
        P.mode = g_prec;
        P.in = in;
        P.g = g;
        P.w = w;
        P.t = t;
        P.fo = fo;
        return 0;
    }
    hipError_t e = launch_sw_vjp(in, cot, g, w, t, g_prec, st, s->events(2), fwd ? &fo : nullptr);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("SW backward launch: ") + hipGetErrorString(e));
    if (host) {
        sg.back(tlay_bar, g.tlay, n2);
        sg.back(h2ovmr_bar, g.h2o, n2);
        sg.back(o3vmr_bar, g.o3, n2);
        if (fwd) {
            sg.back(swuflx, fo.uflx, n3);
            sg.back(swdflx, fo.dflx, n3);
            sg.back(swhr, fo.hr, n2);
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess || sg.ck) return fail(RRTMG_HIP_EDEVICE, std::string("SW backward: ") + hipGetErrorString(e));
    }
    return 0;
}

// backward of rrtmg_lw; with uflx/dflx/hr given, also the forward's outputs
static int lw_vjp_impl(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                       const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                       const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                       const double *cfc11vmr, const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                       const double *emis, const double *tauaer, const double *uflx_bar, const double *dflx_bar,
                       const double *hr_bar, double *tlay_bar, double *tlev_bar, double *h2ovmr_bar,
                       double *o3vmr_bar, double *tsfc_bar, double *uflx, double *dflx, double *hr, int ptr_kind,
                       void *stream)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->lwd.ready) return fail(RRTMG_HIP_ENOINIT, "rrtmg_hip_lw_vjp called before rrtmg_hip_lw_init");
    int rc = check_common(ncol, nlay, plev, ptr_kind);
    if (rc) return rc;
    const size_t n2 = (size_t)ncol * nlay, n3 = (size_t)ncol * (nlay + 1);
    const double *ins2[] = {play, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr};
    const char *nm2[] = {"play", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
                         "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr"};
    double *outs[] = {tlay_bar, tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar};
    for (double *o : outs)
        if (!o) return fail(RRTMG_HIP_EINVAL, "null gradient array");
    const bool fwd = uflx || dflx || hr;
    if (fwd && !(uflx && dflx && hr)) return fail(RRTMG_HIP_EINVAL, "null forward output array");
    const bool host = ptr_kind == RRTMG_HIP_HOST;
    if (host) {
        for (int i = 0; i < 12; ++i)
            if ((rc = check_host_array(ins2[i], n2, nm2[i]))) return rc;
        if ((rc = check_host_array(plev, n3, "plev")) || (rc = check_host_array(tlev, n3, "tlev")) ||
            (rc = check_host_array(tsfc, ncol, "tsfc")) || (rc = check_host_array(emis, (size_t)ncol * 16, "emis")))
            return rc;
        if (tauaer && (rc = check_host_array(tauaer, n2 * 16, "tauaer"))) return rc;
        if ((uflx_bar && (rc = check_host_array(uflx_bar, n3, "uflx_bar"))) ||
            (dflx_bar && (rc = check_host_array(dflx_bar, n3, "dflx_bar"))) ||
            (hr_bar && (rc = check_host_array(hr_bar, n2, "hr_bar"))))
            return rc;
    } else {
        for (auto p : ins2)
            if (!p) return fail(RRTMG_HIP_EINVAL, "null input array");
        if (!plev || !tlev || !tsfc || !emis) return fail(RRTMG_HIP_EINVAL, "null input array");
    }
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    const bool fused = W.fused.open;
    if (fused && host) return fail(RRTMG_HIP_EINVAL, "calls inside a fused session take device pointers");
    if (fused && W.fused.lw.on) return fail(RRTMG_HIP_EINVAL, "one LW backward call per fused session");
    if (!host && g_checks) {
        std::vector<NamedArray> a;
        for (int i = 0; i < 12; ++i) a.push_back({nm2[i], ins2[i], n2});
        a.push_back({"plev", plev, n3});
        a.push_back({"tlev", tlev, n3});
        a.push_back({"tsfc", tsfc, (size_t)ncol});
        a.push_back({"emis", emis, (size_t)ncol * 16});
        if (tauaer) a.push_back({"tauaer", tauaer, n2 * 16});
        if (uflx_bar) a.push_back({"uflx_bar", uflx_bar, n3});
        if (dflx_bar) a.push_back({"dflx_bar", dflx_bar, n3});
        if (hr_bar) a.push_back({"hr_bar", hr_bar, n2});
        if ((rc = device_checks(W, st, a, plev, ncol, nlay))) return rc;
    }
    const int mixed = g_prec == RRTMG_HIP_PREC_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    if ((rc = check_lds(lw_vjp_lds_bytes(nlay, rs)))) return rc;
    size_t off = 0;
    const size_t o_dc = off; off = align_up(off + rs * LWF_STRIDE * n2);
    const size_t o_dd = off; off = align_up(off + rs * LW_DSTRIDE * n2);
    const size_t o_tlev = off; off = align_up(off + rs * n3);
    const size_t o_ipk = off; off = align_up(off + sizeof(int) * n2);
    const size_t o_ilev = off; off = align_up(off + sizeof(int) * n3);
    const size_t o_lt = off; off = align_up(off + sizeof(int) * ncol);
    const size_t o_cs = off; off = align_up(off + sizeof(double) * 4 * ncol);
    const size_t o_cot = off; off = align_up(off + rs * 2 * n3);
    const size_t o_gp = off; off = align_up(off + rs * 16 * 3 * n2);
    const size_t o_gl = off; off = align_up(off + rs * 16 * n3);
    const size_t o_gs = off; off = align_up(off + rs * 16 * 2 * ncol);
    const int chunk = vjp_chunk((ncol + TILE_COLS - 1) / TILE_COLS * s->lw.tab.nwaves, lw_vjp_state_bytes(nlay, rs));
    const bool lds = mixed && s->lwd.lds;
    const size_t ds_bytes = lds ? (size_t)lw_vjp_lds_grid(s->lwd.plan, ncol, s->lw.tab.nwaves) * LWV_WGW *
                                      lw_vjp_lds_state_bytes(nlay)
                                : (size_t)chunk * lw_vjp_state_bytes(nlay, rs);
    const size_t o_ds = off; off = align_up(off + ds_bytes);
    const size_t o_it = off; off = align_up(off + sizeof(int));
    const size_t o_bf = off; off = align_up(off + (fwd ? rs * 16 * 2 * n3 : 0));
    if ((rc = W.lw_vjp_work.ensure(off))) return fail(rc, "device workspace allocation failed");
    char *wb = (char *)W.lw_vjp_work.p;
    LwVjpWork w;
    w.dcoef = wb + o_dc;
    w.dder = wb + o_dd;
    w.tlev = wb + o_tlev;
    w.ipk = (int *)(wb + o_ipk);
    w.ilev = (int *)(wb + o_ilev);
    w.laytrop = (int *)(wb + o_lt);
    w.colscal = (double *)(wb + o_cs);
    w.cot = wb + o_cot;
    w.gpart = wb + o_gp;
    w.glev = wb + o_gl;
    w.gsc = wb + o_gs;
    w.dstate = wb + o_ds;
    w.chunk = chunk;
    w.items = (int *)(wb + o_it);
    w.bflux = fwd ? wb + o_bf : nullptr;
    if (host) {
        const size_t need = 256 * 27 + sizeof(double) * (16 * n2 + 5 * n3 + (size_t)ncol * 18 + (tauaer ? 16 * n2 : 0) +
                                                         (fwd ? n2 + 2 * n3 : 0));
        if ((rc = W.vjp_stage.ensure(need))) return fail(rc, "device staging allocation failed");
    }
    Stager sg{host ? (char *)W.vjp_stage.p : nullptr, host, st};
    ColumnInputs in{};
    in.ncol = ncol;
    in.nlay = nlay;
    in.play = sg.in(play, n2); in.plev = sg.in(plev, n3); in.tlay = sg.in(tlay, n2); in.tlev = sg.in(tlev, n3);
    in.tsfc = sg.in(tsfc, ncol); in.h2o = sg.in(h2ovmr, n2); in.o3 = sg.in(o3vmr, n2); in.co2 = sg.in(co2vmr, n2);
    in.ch4 = sg.in(ch4vmr, n2); in.n2o = sg.in(n2ovmr, n2); in.o2 = sg.in(o2vmr, n2);
    in.cfc11 = sg.in(cfc11vmr, n2); in.cfc12 = sg.in(cfc12vmr, n2); in.cfc22 = sg.in(cfc22vmr, n2);
    in.ccl4 = sg.in(ccl4vmr, n2); in.emis = sg.in(emis, (size_t)ncol * 16); in.tauaer = sg.in(tauaer, n2 * 16);
    VjpCot cot{sg.in(uflx_bar, n3), sg.in(dflx_bar, n3), sg.in(hr_bar, n2)};
    LwGrad g{sg.out(tlay_bar, n2), sg.out(tlev_bar, n3), sg.out(h2ovmr_bar, n2), sg.out(o3vmr_bar, n2),
             sg.out(tsfc_bar, ncol)};
    LwOutputs fo{};
    if (fwd) {  // clear sky: the clear-sky outputs are the same arrays
        fo.uflx = fo.uflxc = sg.out(uflx, n3);
        fo.dflx = fo.dflxc = sg.out(dflx, n3);
        fo.hr = fo.hrc = sg.out(hr, n2);
    }
    if (sg.ck) return fail(RRTMG_HIP_EDEVICE, "LW backward: host-to-device copy failed");
    LwDevTables t;
    t.desc = s->lw.tab;
    t.bands = (const LwBand *)s->lwd.bands;
    t.tab = mixed ? s->lwd.tab32 : s->lwd.tab64;
    t.cst = s->lwd.cst;
    t.units = s->lwd.units;
    t.heatfac = s->lw.cst.heatfac;
    t.bands_lds = s->lwd.bands_lds;
    t.plan = s->lwd.plan;
    t.lds = lds;
    if (fused) {
        const hipError_t e = launch_lw_vjp_prep(in, cot, w, t, g_prec, st);
        if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("LW backward launch: ") + hipGetErrorString(e));
        PendingLw &P = W.fused.lw;
        P.on = true;
        P.fwd = fwd;
        P.mode = g_prec;
        P.in = in;
        P.g = g;
        P.w = w;
        P.t = t;
        P.fo = fo;
        return 0;
    }
    hipError_t e = launch_lw_vjp(in, cot, g, w, t, g_prec, st, s->events(3), fwd ? &fo : nullptr);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("LW backward launch: ") + hipGetErrorString(e));
    if (host) {
        sg.back(tlay_bar, g.tlay, n2);
        sg.back(tlev_bar, g.tlev, n3);
        sg.back(h2ovmr_bar, g.h2o, n2);
        sg.back(o3vmr_bar, g.o3, n2);
        sg.back(tsfc_bar, g.tsfc, ncol);
        if (fwd) {
            sg.back(uflx, fo.uflx, n3);
            sg.back(dflx, fo.dflx, n3);
            sg.back(hr, fo.hr, n2);
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess || sg.ck) return fail(RRTMG_HIP_EDEVICE, std::string("LW backward: ") + hipGetErrorString(e));
    }
    return 0;
}

int rrtmg_hip_sw_vjp(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                     const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                     const double *asdir, const double *asdif, const double *aldir, const double *aldif,
                     const double *coszen, double adjes, int dyofyr, double scon, const double *swuflx_bar,
                     const double *swdflx_bar, const double *swhr_bar, double *tlay_bar, double *h2ovmr_bar,
                     double *o3vmr_bar, int ptr_kind, void *stream)
{
    return sw_vjp_impl(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, asdir,
                       asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx_bar, swdflx_bar, swhr_bar, tlay_bar,
                       h2ovmr_bar, o3vmr_bar, nullptr, nullptr, nullptr, ptr_kind, stream);
}

int rrtmg_hip_sw_fwd_vjp(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                         const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *asdir, const double *asdif, const double *aldir, const double *aldif,
                         const double *coszen, double adjes, int dyofyr, double scon, const double *swuflx_bar,
                         const double *swdflx_bar, const double *swhr_bar, double *swuflx, double *swdflx,
                         double *swhr, double *tlay_bar, double *h2ovmr_bar, double *o3vmr_bar, int ptr_kind,
                         void *stream)
{
    if (!swuflx || !swdflx || !swhr) return fail(RRTMG_HIP_EINVAL, "null forward output array");
    return sw_vjp_impl(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr, asdir,
                       asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx_bar, swdflx_bar, swhr_bar, tlay_bar,
                       h2ovmr_bar, o3vmr_bar, swuflx, swdflx, swhr, ptr_kind, stream);
}

int rrtmg_hip_lw_vjp(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                     const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                     const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                     const double *cfc11vmr, const double *cfc12vmr, const double *cfc22vmr, const double *ccl4vmr,
                     const double *emis, const double *tauaer, const double *uflx_bar, const double *dflx_bar,
                     const double *hr_bar, double *tlay_bar, double *tlev_bar, double *h2ovmr_bar,
                     double *o3vmr_bar, double *tsfc_bar, int ptr_kind, void *stream)
{
    return lw_vjp_impl(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                       cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, uflx_bar, dflx_bar, hr_bar, tlay_bar,
                       tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar, nullptr, nullptr, nullptr, ptr_kind, stream);
}

int rrtmg_hip_lw_fwd_vjp(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc, const double *h2ovmr, const double *o3vmr,
                         const double *co2vmr, const double *ch4vmr, const double *n2ovmr, const double *o2vmr,
                         const double *cfc11vmr, const double *cfc12vmr, const double *cfc22vmr,
                         const double *ccl4vmr, const double *emis, const double *tauaer, const double *uflx_bar,
                         const double *dflx_bar, const double *hr_bar, double *uflx, double *dflx, double *hr,
                         double *tlay_bar, double *tlev_bar, double *h2ovmr_bar, double *o3vmr_bar,
                         double *tsfc_bar, int ptr_kind, void *stream)
{
    if (!uflx || !dflx || !hr) return fail(RRTMG_HIP_EINVAL, "null forward output array");
    return lw_vjp_impl(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                       cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, uflx_bar, dflx_bar, hr_bar, tlay_bar,
                       tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar, uflx, dflx, hr, ptr_kind, stream);
}

int rrtmg_hip_fused_begin(void *stream)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    StreamWork &W = s->work((hipStream_t)stream);
    if (W.fused.open) return fail(RRTMG_HIP_EINVAL, "a fused session is already open on this stream");
    W.fused.open = true;
    W.fused.lw.on = W.fused.sw.on = false;
    return 0;
}

int rrtmg_hip_fused_end(void *stream)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    if (!W.fused.open) return fail(RRTMG_HIP_EINVAL, "rrtmg_hip_fused_end without rrtmg_hip_fused_begin");
    W.fused.open = false;
    const PendingLw L = W.fused.lw;
    const PendingSw S = W.fused.sw;
    W.fused.lw.on = W.fused.sw.on = false;
    bool fuse = L.on && S.on && L.mode == RRTMG_HIP_PREC_MIXED && S.mode == RRTMG_HIP_PREC_MIXED && L.t.lds;
    if (fuse) {
        int dev = 0, lim = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            lim = 65536;
        fuse = fused_lds_bytes(L.t.plan, S.in.nlay) <= (size_t)lim;
    }
    hipError_t e = hipSuccess;
    g_fused_last = fuse ? 1 : 0;
    if (fuse) {
        FusedBand a;
        a.lin = L.in;
        a.lw = L.w;
        a.lbands = L.t.bands_lds;
        a.plan = L.t.plan;
        a.lnwaves = L.t.desc.nwaves;
        const int ltiles = (L.in.ncol + TILE_COLS - 1) / TILE_COLS;
        a.litems = (ltiles + FUSED_WGW - 1) / FUSED_WGW * a.lnwaves;
        a.exp_off = L.t.desc.exp_tbl;
        a.tfn_off = L.t.desc.tfn_tbl;
        a.plk_off = L.t.desc.totplnk;
        a.ltab = (const float *)L.t.tab;
        a.cst = L.t.cst;
        a.lunits = L.t.units;
        a.sin = S.in;
        a.sw = S.w;
        a.sbands = S.t.bands;
        a.snwaves = S.t.desc.nwaves;
        a.snunits = (S.in.ncol + TILE_COLS - 1) / TILE_COLS * a.snwaves;
        a.sexp_off = S.t.desc.exp_tbl;
        a.stab = (const float *)S.t.tab;
        a.sunits = S.t.units;
        a.row_floats = fused_row_floats(L.in.nlay, S.in.nlay);
        a.items = L.w.items;
        a.nitems = a.litems + (a.snunits + FUSED_WGW - 1) / FUSED_WGW;
        const size_t rows = (size_t)fused_grid(a.plan, a.nitems) * FUSED_WGW;
        int rc = W.fused_rows.ensure(rows * a.row_floats * sizeof(float));
        if (rc) return fail(rc, "device workspace allocation failed");
        a.rows = (float *)W.fused_rows.p;
        e = launch_fused_band(a, st);
    } else {
        if (L.on) e = launch_lw_vjp_band(L.in, L.w, L.t, L.mode, st);
        if (e == hipSuccess && S.on) e = launch_sw_vjp_band(S.in, S.w, S.t, S.mode, st);
    }
    if (e == hipSuccess && L.on) e = launch_lw_vjp_finish(L.in, L.g, L.w, L.t, L.mode, st, L.fwd ? &L.fo : nullptr);
    if (e == hipSuccess && S.on) e = launch_sw_vjp_finish(S.in, S.g, S.w, S.t, S.mode, st, S.fwd ? &S.fo : nullptr);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("fused band launch: ") + hipGetErrorString(e));
    return 0;
}

int rrtmg_hip_fused_last_kernel(void) { return g_fused_last; }

// ---------------------------------------------------------------------------
// MLP surrogate kernels (device pointers, enqueued on `stream`)

int rrtmg_hip_gemm_bf16(int M, int N, int K, const void *A, int lda, int transA, const void *B, int ldb, int transB,
                        const float *bias, int relu, float *out_pre, float *out_f32, void *out_bf16, int ldc,
                        int accumulate, void *stream)
{
    if (M <= 0 || N <= 0 || K <= 0) return fail(RRTMG_HIP_EINVAL, "gemm: sizes must be positive");
    if (!A || !B) return fail(RRTMG_HIP_EINVAL, "gemm: null operand");
    if (!out_pre && !out_f32 && !out_bf16) return fail(RRTMG_HIP_EINVAL, "gemm: no output");
    if (lda < (transA ? M : K) || ldb < (transB ? N : K) || ldc < N)
        return fail(RRTMG_HIP_EINVAL, "gemm: leading dimension smaller than the matrix");
    GemmEpilogue ep;
    ep.bias = bias;
    ep.relu = relu;
    ep.out_pre = out_pre;
    ep.out_f32 = out_f32;
    ep.out_bf16 = (uint16_t *)out_bf16;
    ep.ldc = ldc;
    ep.accumulate = accumulate;
    // A^T B with few output tiles: split K over the device's split-K
    // workspace (falls back to the direct kernel if it cannot be allocated)
    int ks = gemm_splitk_parts(M, N, K, transA, transB, lda, ldb, A, B);
    float *part = nullptr;
    if (ks) {
        DeviceState *s = state();
        if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
        std::lock_guard<std::mutex> lk(s->mu);
        StreamWork &W = s->work((hipStream_t)stream);
        if (W.gemm_ws.ensure((size_t)ks * M * N * sizeof(float)) == 0) part = (float *)W.gemm_ws.p;
        else ks = 0;
    }
    hipError_t e = launch_gemm_bf16(M, N, K, (const uint16_t *)A, lda, transA, (const uint16_t *)B, ldb, transB, ep,
                                    (hipStream_t)stream, part, ks);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("gemm launch: ") + hipGetErrorString(e));
}

int rrtmg_hip_mlp_relu_backward(long n, const float *dy, const float *pre, void *dz_bf16, float *dz_f32, void *stream)
{
    if (n <= 0 || !dy || !pre || !dz_bf16) return fail(RRTMG_HIP_EINVAL, "relu_backward: bad arguments");
    hipError_t e = launch_relu_backward((size_t)n, dy, pre, (uint16_t *)dz_bf16, dz_f32, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_mlp_col_sum(int M, int N, const float *x, int ldx, float *out, void *stream)
{
    if (M <= 0 || N <= 0 || ldx < N || !x || !out) return fail(RRTMG_HIP_EINVAL, "col_sum: bad arguments");
    hipError_t e = launch_col_sum(M, N, x, ldx, out, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_mlp_mse(int M, int N, const float *y, int ldy, const float *t, int ldt, float *dy, void *dy_bf16,
                      int ldd, float *loss_partial, int nblocks, void *stream)
{
    if (M <= 0 || N <= 0 || !y || !t || !loss_partial || nblocks <= 0 || ldy < N || ldt < N ||
        ((dy || dy_bf16) && ldd < N))
        return fail(RRTMG_HIP_EINVAL, "mse: bad arguments");
    hipError_t e = launch_mse(M, N, y, ldy, t, ldt, dy, (uint16_t *)dy_bf16, ldd, loss_partial, nblocks,
                              (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_mlp_adam(long n, float *w, float *m, float *v, const float *g, void *w_bf16, float lr, float beta1,
                       float beta2, float eps, float weight_decay, int step, void *stream)
{
    if (n <= 0 || !w || !m || !v || !g || step < 1) return fail(RRTMG_HIP_EINVAL, "adam: bad arguments");
    AdamParams p{lr, beta1, beta2, eps, weight_decay, (float)(1.0 / (1.0 - std::pow((double)beta1, step))),
                 (float)(1.0 / (1.0 - std::pow((double)beta2, step)))};
    hipError_t e = launch_adam((size_t)n, w, m, v, g, (uint16_t *)w_bf16, p, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_mlp_features(int ncol, int nlay, const double *tlay, const double *h2ovmr, const double *o3vmr,
                           const double *tsfc, const double *coszen, const float *mean, const float *istd,
                           const float *lo, const float *hi, int ldx, void *x_bf16, float *xraw, int *out_of_envelope,
                           void *stream)
{
    if (ncol <= 0 || nlay < 2 || ldx < 3 * nlay + 2 || ldx % 2 || ldx > 240 || !tlay || !h2ovmr || !o3vmr || !tsfc ||
        !coszen || !x_bf16 ||
        (!mean) != (!istd) || (!lo) != (!hi))
        return fail(RRTMG_HIP_EINVAL, "mlp_features: bad arguments");
    hipError_t e = launch_mlp_features(ncol, nlay, tlay, h2ovmr, o3vmr, tsfc, coszen, mean, istd, lo, hi, ldx,
                                       (uint16_t *)x_bf16, xraw, out_of_envelope, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_mlp_targets(int ncol, int nlay, const double *hr_lw, const double *hr_sw, const float *mean,
                          const float *istd, int ldt, float *t, void *stream)
{
    if (ncol <= 0 || nlay < 2 || ldt < 2 * nlay || ldt > 240 || !hr_lw || !hr_sw || !t || (!mean) != (!istd))
        return fail(RRTMG_HIP_EINVAL, "mlp_targets: bad arguments");
    hipError_t e = launch_mlp_targets(ncol, nlay, hr_lw, hr_sw, mean, istd, ldt, t, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, hipGetErrorString(e));
}

int rrtmg_hip_surrogate(int ncol, int nlay, const double *tlay, const double *h2ovmr, const double *o3vmr,
                        const double *tsfc, const double *coszen, const RrtmgHipMlp *net, double *hr_lw,
                        double *hr_sw, int *out_of_envelope, void *stream)
{
    if (ncol <= 0 || !net || !hr_lw || !hr_sw || !tlay || !h2ovmr || !o3vmr || !tsfc || !coszen)
        return fail(RRTMG_HIP_EINVAL, "surrogate: bad arguments");
    if (nlay != net->nlay || net->nin != 3 * nlay + 2 || net->nout != 2 * nlay || net->nlayers < 1 ||
        net->nlayers > RRTMG_HIP_MLP_MAX_LAYERS)
        return fail(RRTMG_HIP_EINVAL, "surrogate: network shape does not match nlay");
    if (net->ldx % 2 || net->ldx > 240 || nlay > 120)  // the layout kernels' 64-column LDS tiles
        return fail(RRTMG_HIP_EINVAL, "surrogate: padded feature count must be even and at most 240, nlay at most 120");
    int width[RRTMG_HIP_MLP_MAX_LAYERS + 1];
    width[0] = net->ldx;
    int wmax = net->ldx;
    for (int l = 0; l < net->nlayers; ++l) {
        width[l + 1] = net->width[l];
        if (!net->w[l] || !net->b[l] || net->width[l] <= 0) return fail(RRTMG_HIP_EINVAL, "surrogate: missing layer");
        wmax = std::max(wmax, net->width[l]);
    }
    if (width[net->nlayers] < net->nout) return fail(RRTMG_HIP_EINVAL, "surrogate: output layer too narrow");
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    int rc;
    if (!std::getenv("RRTMG_HIP_MLP_UNFUSED") && mlp_fused_fits(net->ldx, net->nlayers, net->width, net->nout)) {
        // one kernel: features, every layer and the decode, activations in
        // LDS; the weights first repacked in the order its MFMAs read them
        size_t wbytes = 0;
        for (int l = 0; l < net->nlayers; ++l) wbytes += align_up((size_t)width[l] * width[l + 1] * sizeof(uint16_t));
        if ((rc = W.mlp_pack.ensure(wbytes))) return fail(rc, "surrogate workspace allocation failed");
        const void *pw[RRTMG_HIP_MLP_MAX_LAYERS];
        hipError_t e = hipSuccess;
        char *p = (char *)W.mlp_pack.p;
        for (int l = 0; l < net->nlayers && e == hipSuccess; ++l) {
            e = launch_mlp_pack(width[l + 1], width[l], (const uint16_t *)net->w[l], (uint16_t *)p, st);
            pw[l] = p;
            p += align_up((size_t)width[l] * width[l + 1] * sizeof(uint16_t));
        }
        if (e == hipSuccess)
            e = launch_mlp_infer_fused(ncol, nlay, tlay, h2ovmr, o3vmr, tsfc, coszen, net->ldx, net->nlayers,
                                       net->width, pw, net->b, net->in_mean, net->in_istd, net->in_lo, net->in_hi,
                                       net->out_mean, net->out_std, hr_lw, hr_sw, out_of_envelope, st);
        return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("surrogate: ") + hipGetErrorString(e));
    }
    const size_t act = align_up((size_t)ncol * wmax * sizeof(uint16_t));
    const size_t outb = align_up((size_t)ncol * width[net->nlayers] * sizeof(float));
    if ((rc = W.mlp_work.ensure(2 * act + outb))) return fail(rc, "surrogate workspace allocation failed");
    char *wb = (char *)W.mlp_work.p;
    uint16_t *buf[2] = {(uint16_t *)wb, (uint16_t *)(wb + act)};
    float *y = (float *)(wb + 2 * act);
    hipError_t e = launch_mlp_features(ncol, nlay, tlay, h2ovmr, o3vmr, tsfc, coszen, net->in_mean, net->in_istd,
                                       net->in_lo, net->in_hi, net->ldx, buf[0], nullptr, out_of_envelope, st);
    int cur = 0;
    for (int l = 0; l < net->nlayers && e == hipSuccess; ++l) {
        GemmEpilogue ep;
        ep.bias = net->b[l];
        const bool last = l == net->nlayers - 1;
        ep.relu = !last;
        ep.ldc = width[l + 1];
        if (last) ep.out_f32 = y;
        else ep.out_bf16 = buf[1 - cur];
        e = launch_gemm_bf16(ncol, width[l + 1], width[l], buf[cur], width[l], 0, (const uint16_t *)net->w[l], width[l],
                             0, ep, st);
        cur = 1 - cur;
    }
    if (e == hipSuccess)
        e = launch_mlp_decode(ncol, nlay, y, width[net->nlayers], net->out_mean, net->out_std, hr_lw, hr_sw, st);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("surrogate: ") + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// MiMA's radiation driver around the RRTM calls (kernels in driver.hip)

int rrtmg_hip_compute_zenith(int n, const double *lat, const double *lon, int seconds, int days, int daysperyear,
                             double equinox_day, double obliq, int dt, double *cosz, int *dyofyr, void *stream)
{
    if (n < 0 || !dyofyr || (n > 0 && (!lat || !lon || !cosz)))
        return fail(RRTMG_HIP_EINVAL, "compute_zenith: bad arguments");
    if (daysperyear <= 0 || bad(equinox_day) || bad(obliq))
        return fail(RRTMG_HIP_EINVAL, "compute_zenith: daysperyear must be positive, equinox_day and obliq finite");
    // the per-call scalars, as RR/astro.f90:100-143 forms them
    const double PI = 3.141592653589793;  // 4 atan(1), shared/fms/fms.f90:412
    ZenithScalars z;
    const double deg2rad = PI / 180.;
    z.radpersec = 2 * PI / 86400.;
    z.radsec = seconds * z.radpersec;
    z.dt_pi = dt * z.radpersec;
    z.dt = dt;
    const int d = days - (int)(equinox_day * daysperyear);
    int dy = d % daysperyear;
    if (dy < 0) dy += daysperyear;  // Fortran MODULO
    const double radperday = 2 * PI / daysperyear;
    const double radday = dy * radperday;
    z.dec_sin = std::sin(obliq * deg2rad) * std::sin(radday);
    const double dec = std::asin(z.dec_sin);
    z.dec_cos = std::cos(dec);
    z.dec_tan = std::tan(dec);
    hipError_t e = launch_zenith(n, lat, lon, z, cosz, (hipStream_t)stream);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("compute_zenith: ") + hipGetErrorString(e));
    *dyofyr = dy;
    return 0;
}

int rrtmg_hip_interp_temp(int nlon, int nlat, int nlev, const double *z_full, const double *z_half,
                          const double *t_surf, const double *t, double *t_half, void *stream)
{
    if (nlon <= 0 || nlat <= 0 || nlev < 2 || !z_full || !z_half || !t_surf || !t || !t_half)
        return fail(RRTMG_HIP_EINVAL, "interp_temp: bad arguments (nlev must be at least 2)");
    if ((double)nlon * nlat * (nlev + 1) >= 2147483647.0)
        return fail(RRTMG_HIP_ENOTSUP, "interp_temp: window too large for one call");
    hipError_t e = launch_interp_temp(nlon, nlat, nlev, z_full, z_half, t_surf, t, t_half, (hipStream_t)stream);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("interp_temp: ") + hipGetErrorString(e));
}

static int check_driver_cfg(const RrtmgHipDriverConfig *c)
{
    if (!c) return fail(RRTMG_HIP_EINVAL, "driver: null configuration");
    if (c->nlon <= 0 || c->nlat <= 0 || c->nlev < 2 || c->lonstep <= 0 || c->nlon % c->lonstep)
        return fail(RRTMG_HIP_EINVAL, "driver: nlon, nlat positive, nlev >= 2, lonstep positive and dividing nlon");
    if (c->num_precip < 0) return fail(RRTMG_HIP_EINVAL, "driver: num_precip must not be negative");
    const double v[] = {c->fixed_water, c->fixed_water_pres, c->fixed_water_lat, c->precip_albedo, c->precip_lat,
                        c->scale_ozone,  c->o3_val,          c->co2ppmv,          c->h2o_lower_limit,
                        c->temp_lower_limit, c->temp_upper_limit};
    for (double x : v)
        if (bad(x)) return fail(RRTMG_HIP_EINVAL, "driver: non-finite option");
    if ((double)c->nlon * c->nlat * (c->nlev + 1) >= 2147483647.0)
        return fail(RRTMG_HIP_ENOTSUP, "driver: window too large for one call");
    return 0;
}

int rrtmg_hip_rrtmg_pre(const RrtmgHipDriverConfig *cfg, const double *lat, const double *p_full,
                        const double *p_half, const double *t, const double *t_half, const double *q,
                        const double *o3f, const double *albedo, const double *precip, const double *t_surf,
                        const double *coszen, double *play, double *plev, double *tlay, double *tlev,
                        double *h2ovmr, double *o3vmr, double *co2vmr, double *albedo_rr, double *tsfc,
                        double *coszen_rr, void *stream)
{
    int rc = check_driver_cfg(cfg);
    if (rc) return rc;
    const double *need[] = {lat, p_full, p_half, t, t_half, q, albedo, t_surf, coszen};
    double *outs[] = {play, plev, tlay, tlev, h2ovmr, o3vmr, co2vmr, albedo_rr, tsfc, coszen_rr};
    for (const double *p : need)
        if (!p) return fail(RRTMG_HIP_EINVAL, "rrtmg_pre: null input array");
    for (double *p : outs)
        if (!p) return fail(RRTMG_HIP_EINVAL, "rrtmg_pre: null output array");
    if (cfg->do_read_ozone && !o3f) return fail(RRTMG_HIP_EINVAL, "rrtmg_pre: do_read_ozone needs o3f");
    if (cfg->do_precip_albedo && cfg->num_precip > 0 && !precip)
        return fail(RRTMG_HIP_EINVAL, "rrtmg_pre: do_precip_albedo needs precip");
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    if ((rc = W.drv.ensure(sizeof(double) * driver_work_doubles(cfg->nlat, cfg->nlev))))
        return fail(rc, "rrtmg_pre: device workspace allocation failed");
    DriverDims d{cfg->nlon, cfg->nlat, cfg->nlev, cfg->lonstep};
    DriverPre a;
    a.cf = *cfg;
    a.lat = lat; a.p_full = p_full; a.p_half = p_half; a.t = t; a.t_half = t_half; a.q = q; a.o3f = o3f;
    a.albedo = albedo; a.precip = precip; a.t_surf = t_surf; a.coszen = coszen;
    a.play = play; a.plev = plev; a.tlay = tlay; a.tlev = tlev; a.h2o = h2ovmr; a.o3 = o3vmr; a.co2 = co2vmr;
    a.albedo_rr = albedo_rr; a.tsfc = tsfc; a.coszen_rr = coszen_rr;
    hipError_t e = launch_rrtmg_pre(d, a, (double *)W.drv.p, st);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("rrtmg_pre: ") + hipGetErrorString(e));
}

int rrtmg_hip_rrtmg_post(const RrtmgHipDriverConfig *cfg, const double *swhr, const double *hr,
                         const double *swuflx, const double *swdflx, const double *uflx, const double *dflx,
                         double *tdt, double *tdt_rad, double *flux_sw, double *flux_lw, double *tdt_sw,
                         double *tdt_lw, double *olr, double *isr, void *stream)
{
    int rc = check_driver_cfg(cfg);
    if (rc) return rc;
    if (!swhr || !hr || !swuflx || !swdflx || !uflx || !dflx || !tdt)
        return fail(RRTMG_HIP_EINVAL, "rrtmg_post: null array");
    if (!flux_sw != !flux_lw) return fail(RRTMG_HIP_EINVAL, "rrtmg_post: flux_sw and flux_lw go together");
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    hipStream_t st = (hipStream_t)stream;
    StreamWork &W = s->work(st);
    if ((rc = W.drv.ensure(sizeof(double) * driver_work_doubles(cfg->nlat, cfg->nlev))))
        return fail(rc, "rrtmg_post: device workspace allocation failed");
    DriverDims d{cfg->nlon, cfg->nlat, cfg->nlev, cfg->lonstep};
    DriverPost a{cfg->do_zm_rad, swhr, hr, swuflx, swdflx, uflx, dflx, tdt, tdt_rad, flux_sw, flux_lw,
                 tdt_sw, tdt_lw, olr, isr};
    hipError_t e = launch_rrtmg_post(d, a, (double *)W.drv.p, st);
    return e == hipSuccess ? 0 : fail(RRTMG_HIP_EDEVICE, std::string("rrtmg_post: ") + hipGetErrorString(e));
}

int rrtmg_hip_set_checks(int on)
{
    int prev = g_checks ? 1 : 0;
    g_checks = on != 0;
    return prev;
}

int rrtmg_hip_set_timing(int on)
{
    int prev = g_timing ? 1 : 0;
    g_timing = on != 0;
    return prev;
}

int rrtmg_hip_last_kernel_ms(int lw, float *ms)
{
    DeviceState *s = state();
    if (!s || !ms) return fail(RRTMG_HIP_EINVAL, "no device or null output");
    std::lock_guard<std::mutex> lk(s->mu);
    if (lw < 0 || lw > 3) return fail(RRTMG_HIP_EINVAL, "call kind must be 0..3");
    const int side = lw;
    if (!s->ev_ok || !s->ev_recorded[side]) return fail(RRTMG_HIP_ENOINIT, "no timed call recorded");
    if (hipEventSynchronize(s->ev[side][3]) != hipSuccess) return fail(RRTMG_HIP_EDEVICE, "event sync failed");
    for (int k = 0; k < 3; ++k)
        if (hipEventElapsedTime(&ms[k], s->ev[side][k], s->ev[side][k + 1]) != hipSuccess)
            return fail(RRTMG_HIP_EDEVICE, "event timing failed");
    return 0;
}

int rrtmg_hip_debug_indices(int lw, int *jp, int *jt, int *jt1, int *laytrop)
{
    DeviceState *s = state();
    if (!s) return fail(RRTMG_HIP_EDEVICE, "no HIP device");
    std::lock_guard<std::mutex> lk(s->mu);
    int ncol = lw ? s->lw_last_ncol : s->sw_last_ncol, nlay = lw ? s->lw_last_nlay : s->sw_last_nlay;
    if (!ncol) return fail(RRTMG_HIP_ENOINIT, "no previous call");
    const int *ipk = lw ? s->lw_w.ipk : s->sw_w.ipk;
    const int *lt = lw ? s->lw_w.laytrop : s->sw_w.laytrop;
    std::vector<int> h((size_t)ncol * nlay);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(h.data(), ipk, h.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(RRTMG_HIP_EDEVICE, "copy of indices failed");
    for (size_t i = 0; i < h.size(); ++i) {
        if (jp) jp[i] = unpack_jp(h[i]);
        if (jt) jt[i] = unpack_jt(h[i]);
        if (jt1) jt1[i] = unpack_jt1(h[i]);
    }
    if (laytrop && hipMemcpy(laytrop, lt, ncol * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(RRTMG_HIP_EDEVICE, "copy of laytrop failed");
    return 0;
}

static SwHost g_host_sw;
static LwHost g_host_lw;
static bool g_host_sw_ok = false, g_host_lw_ok = false;

int rrtmg_hip_debug_host_tables(int lw, const char *table_dir, double cpdair)
{
    std::lock_guard<std::mutex> lk(g_mu);
    std::string dir = table_dir ? table_dir : default_table_dir();
    std::string err;
    (lw ? g_host_lw_ok : g_host_sw_ok) = false;  // a failed load leaves no half-built tables readable
    bool ok = lw ? build_lw(dir + "/rrtmg_lw.tbl", cpdair, g_host_lw, err)
                 : build_sw(dir + "/rrtmg_sw.tbl", cpdair, g_host_sw, err);
    if (!ok) return fail(RRTMG_HIP_ENOENT, err);
    (lw ? g_host_lw_ok : g_host_sw_ok) = true;
    return 0;
}

int rrtmg_hip_debug_smooth(int n, const double *x, double *out)
{
    if (n < 0 || (n > 0 && (!x || !out))) return fail(RRTMG_HIP_EINVAL, "bad arguments");
    if (n == 0) return 0;
    double *dx = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&dx, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, sizeof(double) * 6 * n);
    if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_smooth_functions(n, dx, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * 6 * n, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(RRTMG_HIP_EDEVICE, std::string("smooth functions: ") + hipGetErrorString(e));
    return 0;
}

long rrtmg_hip_debug_table(const char *name, double *out, long n)
{
    if (!name) return -1;
    const bool lw = name[0] == 'l';
    if ((lw && g_host_lw_ok) || (!lw && g_host_sw_ok)) {
        const auto &m = lw ? g_host_lw.named : g_host_sw.named;
        auto it = m.find(name);
        if (it == m.end()) return -1;
        long cnt = (long)it->second.size();
        if (out) std::memcpy(out, it->second.data(), sizeof(double) * (size_t)(n < cnt ? n : cnt));
        return cnt;
    }
    DeviceState *s = state();
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    const auto &m = lw ? s->lw.named : s->sw.named;
    auto it = m.find(name);
    if (it == m.end()) return -1;
    long cnt = (long)it->second.size();
    if (out) std::memcpy(out, it->second.data(), sizeof(double) * (size_t)(n < cnt ? n : cnt));
    return cnt;
}

}  // extern "C"
