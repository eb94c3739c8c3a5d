// Host-side table preparation: read the RRTMGTB1 files, reduce the original
// 16-point g-intervals of every band to the model's g-points and pack the
// reduced tables into one buffer per side with the g-point index fastest.
//
// Reduction rule (the cmbgb* routines, rrtmg_sw_init.f90:473-1516 and
// rrtmg_lw_init.f90:366-2015): an absorption-type array (major, continuum,
// minor, cross-section and Rayleigh coefficients) becomes
//     reduced(..., igc) = sum over its original g-points ipr of orig(..., ipr) * rwgt(ipr)
// and a source/fraction array (sfluxref, fracref) the plain sum; rwgt is
// wt(ig) / (sum of wt over the original g-points of the reduced point)
// (rrtmg_sw_init.f90:106-130, rrtmg_lw_init.f90:126-154).
#include "tables.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "table_shapes.hpp"

namespace rrtmg {

bool TableFile::load(const std::string &path, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open table file " + path;
        return false;
    }
    raw_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (raw_.size() < 16 || std::memcmp(raw_.data(), "RRTMGTB1", 8) != 0) {
        err = path + " is not an RRTMGTB1 table file";
        return false;
    }
    uint32_t n;
    std::memcpy(&n, raw_.data() + 8, 4);
    std::memcpy(&flags_, raw_.data() + 12, 4);
    const size_t esz = 48 + 4 + 4 + 16 + 16 + 8 + 8;
    if (16 + n * esz > raw_.size()) {
        err = path + ": truncated header";
        return false;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const unsigned char *e = raw_.data() + 16 + i * esz;
        Entry en;
        char name[49] = {0};
        std::memcpy(name, e, 48);
        std::memcpy(&en.dtype, e + 48, 4);
        std::memcpy(&en.ndim, e + 52, 4);
        std::memcpy(en.lb, e + 56, 16);
        std::memcpy(en.ext, e + 72, 16);
        std::memcpy(&en.off, e + 88, 8);
        std::memcpy(&en.nbytes, e + 96, 8);
        if (en.off + en.nbytes > raw_.size()) {
            err = path + ": entry " + name + " out of range";
            return false;
        }
        entries_[name] = en;
    }
    return true;
}

const TableFile::Entry *TableFile::find(const std::string &name) const
{
    auto it = entries_.find(name);
    return it == entries_.end() ? nullptr : &it->second;
}

std::vector<double> TableFile::doubles(const std::string &name) const
{
    const Entry *e = find(name);
    if (!e || e->dtype != 0) return {};
    std::vector<double> v(e->nbytes / 8);
    std::memcpy(v.data(), raw_.data() + e->off, e->nbytes);
    return v;
}

std::vector<int> TableFile::ints(const std::string &name) const
{
    const Entry *e = find(name);
    if (!e || e->dtype != 1) return {};
    std::vector<int> v(e->nbytes / 4);
    std::memcpy(v.data(), raw_.data() + e->off, e->nbytes);
    return v;
}

// Checks that a table file holds every band array with the shape the band
// kernels index it with (table_shapes.hpp) and reduction bookkeeping that
// make_rwgt and reduce can use: ngs the running sum of ngc, each band's ngn
// counts summing to its 16 original g-points, and ngm numbering every
// original g-point by the reduced point those counts assign it (the
// reference's lwcmbdat/swcmbdat literals satisfy all three,
// rrtmg_lw_init.f90:284-363, rrtmg_sw_init.f90:263-367).
static bool check_layout(const TableFile &t, const TableShape *shapes, int nshapes, int nb,
                         const std::vector<int> &ngc, const std::vector<int> &ngs, const std::vector<int> &ngn,
                         const std::vector<int> &ngm, const char *side, std::string &err)
{
    for (int i = 0; i < nshapes; ++i) {
        const TableShape &sh = shapes[i];
        const TableFile::Entry *e = t.find(sh.name);
        bool ok = e && e->dtype == 0 && (int)e->ndim == sh.ndim;
        uint64_t n = 1;
        for (int d = 0; ok && d < 4; ++d) {
            ok = e->ext[d] == sh.ext[d];
            n *= (uint64_t)sh.ext[d];
        }
        if (!ok || e->nbytes != 8 * n) {
            err = std::string("table array ") + sh.name + " is missing or does not have the shape the band kernels use";
            return false;
        }
    }
    if ((int)ngs.size() != nb || (int)ngm.size() != 16 * nb) {
        err = std::string(side) + " table file: ngs/ngm have the wrong length";
        return false;
    }
    int first = 0;
    for (int b = 0; b < nb; ++b) {
        if (ngs[b] != first + ngc[b]) {
            err = std::string(side) + " table file: ngs is not the running sum of ngc";
            return false;
        }
        int g = 0;
        for (int r = 0; r < ngc[b]; ++r) {
            const int k = ngn[first + r];
            if (k < 1 || g + k > 16) {
                err = std::string(side) + " table file: a band's ngn counts do not cover its 16 g-points";
                return false;
            }
            for (int j = 0; j < k; ++j, ++g)
                if (ngm[16 * b + g] != r + 1) {
                    err = std::string(side) + " table file: ngm disagrees with ngn";
                    return false;
                }
        }
        if (g != 16) {
            err = std::string(side) + " table file: a band's ngn counts do not cover its 16 g-points";
            return false;
        }
        first += ngc[b];
    }
    return true;
}

// rwgt for nb bands of 16 original g-points
static std::vector<double> make_rwgt(int nb, const std::vector<int> &ngc, const std::vector<int> &ngn,
                                     const std::vector<int> &ngm, const std::vector<double> &wt)
{
    std::vector<double> rwgt(nb * 16);
    int grp = 0;
    for (int b = 0; b < nb; ++b) {
        if (ngc[b] < 16) {
            double wtsm[16];
            int ipr = 0;
            for (int igc = 0; igc < ngc[b]; ++igc, ++grp) {
                double s = 0.0;
                for (int k = 0; k < ngn[grp]; ++k) s = s + wt[ipr++];
                wtsm[igc] = s;
            }
            for (int ig = 0; ig < 16; ++ig) rwgt[b * 16 + ig] = wt[ig] / wtsm[ngm[b * 16 + ig] - 1];
        } else {
            for (int ig = 0; ig < 16; ++ig) rwgt[b * 16 + ig] = 1.0;
            grp += 16;
        }
    }
    return rwgt;
}

// Reduce one original array.  `gfirst`: the 16-point g index is the first
// dimension (sfluxref, fracref, Rayleigh by eta) rather than the last.
// Returns the reduced array in the same (Fortran) order and its row count
// (number of elements per g-point).
static std::vector<double> reduce(const TableFile &t, const std::string &name, bool gfirst, bool weighted,
                                  int ngc, const int *ngn_band, const double *rw, int *rows)
{
    const TableFile::Entry *e = t.find(name);
    std::vector<double> x = t.doubles(name);
    if (!e || x.empty()) {
        if (rows) *rows = 0;
        return {};
    }
    size_t total = x.size();
    size_t other = total / 16;
    std::vector<double> out(other * ngc);
    for (size_t o = 0; o < other; ++o) {
        int ipr = 0;
        for (int igc = 0; igc < ngc; ++igc) {
            double s = 0.0;
            for (int k = 0; k < ngn_band[igc]; ++k, ++ipr) {
                double v = gfirst ? x[o * 16 + ipr] : x[(size_t)ipr * other + o];
                s = weighted ? s + v * rw[ipr] : s + v;
            }
            if (gfirst)
                out[o * ngc + igc] = s;
            else
                out[(size_t)igc * other + o] = s;
        }
    }
    if (rows) *rows = (int)other;
    return out;
}

// Append `red` (Fortran order) to the packed buffer with g fastest; returns the offset.
static int pack(std::vector<double> &buf, const std::vector<double> &red, int rows, int ng, bool gfirst)
{
    if (red.empty()) return -1;
    int off = (int)buf.size();
    buf.resize(buf.size() + (size_t)rows * ng);
    for (int r = 0; r < rows; ++r)
        for (int g = 0; g < ng; ++g)
            buf[off + (size_t)r * ng + g] = gfirst ? red[(size_t)r * ng + g] : red[(size_t)g * rows + r];
    return off;
}

static void exp_table(std::vector<double> &buf, int *off_exp, int *off_tfn)
{
    // rrtmg_lw_init.f90:106-123 / rrtmg_sw_init.f90:94-104
    const double bpade = 1.0 / 0.278, expeps = 1.e-20;
    *off_exp = (int)buf.size();
    buf.resize(buf.size() + NTBL + 1);
    double *ex = &buf[*off_exp];
    ex[0] = 1.0;
    ex[NTBL] = expeps;
    std::vector<double> tau(NTBL + 1);
    tau[0] = 0.0;
    tau[NTBL] = 1.e10;
    for (int i = 1; i < NTBL; ++i) {
        double tfn = (double)i / (double)NTBL;
        tau[i] = bpade * tfn / (1.0 - tfn);
        ex[i] = std::exp(-tau[i]);
        if (ex[i] <= expeps) ex[i] = expeps;
    }
    if (off_tfn) {
        *off_tfn = (int)buf.size();
        buf.resize(buf.size() + NTBL + 1);
        double *tf = &buf[*off_tfn];
        ex = &buf[*off_exp];
        tf[0] = 0.0;
        tf[NTBL] = 1.0;
        for (int i = 1; i < NTBL; ++i)
            tf[i] = tau[i] < 0.06 ? tau[i] / 6.0 : 1.0 - 2.0 * ((1.0 / tau[i]) - (ex[i] / (1. - ex[i])));
    }
}

// Wave schedule of one tile: for each band, units of floor(64/ng) columns.
static int build_schedule(int nb, const int *ng, std::vector<WaveUnit> &units)
{
    units.clear();
    for (int b = 0; b < nb; ++b) {
        int cpw = 64 / ng[b];
        for (int c0 = 0; c0 < TILE_COLS; c0 += cpw) {
            WaveUnit u;
            u.band = (int16_t)b;
            u.first_col = (int16_t)c0;
            u.ncols = (int16_t)(c0 + cpw <= TILE_COLS ? cpw : TILE_COLS - c0);
            u.pad = 0;
            units.push_back(u);
        }
    }
    return (int)units.size();
}

bool build_sw(const std::string &path, double cpdair, SwHost &h, std::string &err)
{
    TableFile t;
    if (!t.load(path, err)) return false;
    std::vector<int> ngc = t.ints("sw/ngc"), ngs = t.ints("sw/ngs"), ngn = t.ints("sw/ngn"), ngm = t.ints("sw/ngm");
    std::vector<int> nspa = t.ints("sw/nspa"), nspb = t.ints("sw/nspb");
    std::vector<double> wt = t.doubles("sw/wt");
    if (ngc.size() != 14 || ngn.size() != 112 || ngm.size() != 224 || wt.size() != 16 || nspa.size() != 14 ||
        nspb.size() != 14) {
        err = "SW table file lacks the g-point bookkeeping";
        return false;
    }
    for (int b = 0; b < 14; ++b)
        if (ngc[b] != sw_ng(b + 16) || nspa[b] != SW_NSPA[b] || nspb[b] != SW_NSPB[b]) {
            err = "SW table file's g-point or species-parameter counts differ from the compiled band kernels";
            return false;
        }
    if (!check_layout(t, SW_SHAPES, SW_NSHAPES, 14, ngc, ngs, ngn, ngm, "SW", err)) return false;
    h.pref = t.doubles("sw/pref");
    h.preflog = t.doubles("sw/preflog");
    h.tref = t.doubles("sw/tref");
    if (h.pref.size() != 59 || h.preflog.size() != 59 || h.tref.size() != 59) {
        err = "SW table file lacks the reference atmosphere";
        return false;
    }
    h.heatfac = 9.8066 * 8.6400e4 / (cpdair * 1.e2);  // swdatinit, rrtmg_sw_init.f90:239
    std::vector<double> rwgt = make_rwgt(14, ngc, ngn, ngm, wt);
    h.named["sw/rwgt"] = rwgt;
    std::vector<double> &buf = h.buf;
    buf.clear();
    int first = 0;
    int ngv[14];
    for (int b = 0; b < 14; ++b) {
        int jb = b + 16;
        SwBand &d = h.tab.band[b];
        std::memset(&d, 0xff, sizeof d);  // offsets -1
        d.ng = ngc[b];
        ngv[b] = d.ng;
        d.ngs = b == 0 ? 0 : ngs[b - 1];
        d.nspa = nspa[b];
        d.nspb = nspb[b];
        const int *ngn_b = &ngn[first];
        const double *rw = &rwgt[b * 16];
        char key[64];
        auto red = [&](const char *nm, bool gfirst, bool weighted, int *rows, const char *rname) {
            std::snprintf(key, sizeof key, "sw/b%d/%s", jb, nm);
            std::vector<double> r = reduce(t, key, gfirst, weighted, d.ng, ngn_b, rw, rows);
            if (!r.empty()) {
                std::snprintf(key, sizeof key, "sw/b%d/%s", jb, rname);
                h.named[key] = r;
            }
            return r;
        };
        int rows;
        std::vector<double> r;
        r = red("kao", false, true, &rows, "ka");
        d.absa = pack(buf, r, rows, d.ng, false);
        r = red("kbo", false, true, &rows, "kb");
        d.absb = pack(buf, r, rows, d.ng, false);
        r = red("selfrefo", false, true, &rows, "selfref");
        d.selfref = pack(buf, r, rows, d.ng, false);
        r = red("forrefo", false, true, &rows, "forref");
        d.forref = pack(buf, r, rows, d.ng, false);
        r = red("sfluxrefo", true, false, &rows, "sfluxref");
        d.sfluxref = pack(buf, r, rows, d.ng, true);
        d.nsf = rows;
        r = red("raylo", true, true, &rows, "rayl");
        d.raylg = pack(buf, r, rows, d.ng, true);
        r = red("raylao", true, true, &rows, "rayla");
        d.rayla = pack(buf, r, rows, d.ng, true);
        r = red("raylbo", true, true, &rows, "raylb");
        d.raylb = pack(buf, r, rows, d.ng, true);
        r = red("abso3ao", true, true, &rows, "abso3a");
        d.abso3a = pack(buf, r, rows, d.ng, true);
        r = red("abso3bo", true, true, &rows, "abso3b");
        d.abso3b = pack(buf, r, rows, d.ng, true);
        r = red("absch4o", true, true, &rows, "absch4");
        d.absch4 = pack(buf, r, rows, d.ng, true);
        r = red("absco2o", true, true, &rows, "absco2");
        d.absco2 = pack(buf, r, rows, d.ng, true);
        r = red("absh2oo", true, true, &rows, "absh2o");
        d.absh2o = pack(buf, r, rows, d.ng, true);
        std::snprintf(key, sizeof key, "sw/b%d/rayl", jb);
        std::vector<double> rs = t.doubles(key);
        d.rayl = (rs.size() == 1) ? rs[0] : 0.0;
        first += d.ng;
        if (d.sfluxref < 0) {
            err = "SW band missing sfluxref";
            return false;
        }
    }
    int dummy;
    exp_table(buf, &h.tab.exp_tbl, nullptr);
    (void)dummy;
    h.tab.nwaves = build_schedule(14, ngv, h.units);
    return true;
}

bool build_lw(const std::string &path, double cpdair, LwHost &h, std::string &err)
{
    TableFile t;
    if (!t.load(path, err)) return false;
    h.synthetic = (t.flags() & 1u) != 0;
    std::vector<int> ngc = t.ints("lw/ngc"), ngs = t.ints("lw/ngs"), ngn = t.ints("lw/ngn"), ngm = t.ints("lw/ngm");
    std::vector<int> nspa = t.ints("lw/nspa"), nspb = t.ints("lw/nspb");
    std::vector<double> wt = t.doubles("lw/wt"), delwave = t.doubles("lw/delwave");
    std::vector<double> a0 = t.doubles("lw/a0"), a1 = t.doubles("lw/a1"), a2 = t.doubles("lw/a2");
    std::vector<double> pref = t.doubles("lw/pref"), preflog = t.doubles("lw/preflog"), tref = t.doubles("lw/tref");
    std::vector<double> chi = t.doubles("lw/chi_mls"), totplnk = t.doubles("lw/totplnk");
    if (ngc.size() != 16 || ngn.size() != 140 || ngm.size() != 256 || wt.size() != 16 || a0.size() != 16 ||
        a1.size() != 16 || a2.size() != 16 || delwave.size() != 16 || chi.size() != 7 * 59 ||
        totplnk.size() != 181 * 16 || pref.size() != 59 || preflog.size() != 59 || tref.size() != 59 ||
        nspa.size() != 16 || nspb.size() != 16) {
        err = "LW table file lacks bookkeeping, reference atmosphere or Planck data";
        return false;
    }
    for (int b = 0; b < 16; ++b)
        if (ngc[b] != lw_ng(b + 1) || nspa[b] != LW_NSPA[b] || nspb[b] != LW_NSPB[b]) {
            err = "LW table file's g-point or species-parameter counts differ from the compiled band kernels";
            return false;
        }
    if (!check_layout(t, LW_SHAPES, LW_NSHAPES, 16, ngc, ngs, ngn, ngm, "LW", err)) return false;
    LwConst &c = h.cst;
    std::memcpy(c.pref, pref.data(), sizeof c.pref);
    std::memcpy(c.preflog, preflog.data(), sizeof c.preflog);
    std::memcpy(c.tref, tref.data(), sizeof c.tref);
    std::memcpy(c.chi_mls, chi.data(), sizeof c.chi_mls);
    auto CHI = [&](int i, int j) { return chi[(j - 1) * 7 + (i - 1)]; };
    for (int jp = 1; jp <= 59; ++jp) {
        c.rat_h2oco2[jp] = CHI(1, jp) / CHI(2, jp);
        c.rat_h2oo3[jp] = CHI(1, jp) / CHI(3, jp);
        c.rat_h2on2o[jp] = CHI(1, jp) / CHI(4, jp);
        c.rat_h2och4[jp] = CHI(1, jp) / CHI(6, jp);
        c.rat_n2oco2[jp] = CHI(4, jp) / CHI(2, jp);
        c.rat_o3co2[jp] = CHI(3, jp) / CHI(2, jp);
    }
    c.rat_h2oco2[0] = c.rat_h2oo3[0] = c.rat_h2on2o[0] = c.rat_h2och4[0] = c.rat_n2oco2[0] = c.rat_o3co2[0] = 0.0;
    c.heatfac = 9.8066 * 8.6400e4 / (cpdair * 1.e2);  // lwdatinit, rrtmg_lw_init.f90:279
    for (int k = 0; k < LW_NCHI_RATIO; ++k) {
        const LwChiRatio &r = LW_CHI_RATIOS[k];
        c.chi_ratio[k] = CHI(r.num, r.lev) / CHI(r.den, r.lev);
        c.chi_ratio_f[k] = (float)c.chi_ratio[k];
    }
    for (int i = 0; i < 7 * 59; ++i) c.chi_mls_f[i] = (float)c.chi_mls[i];
    for (int jp = 0; jp < 60; ++jp) {
        c.rat_h2oco2_f[jp] = (float)c.rat_h2oco2[jp];
        c.rat_h2oo3_f[jp] = (float)c.rat_h2oo3[jp];
        c.rat_h2on2o_f[jp] = (float)c.rat_h2on2o[jp];
        c.rat_h2och4_f[jp] = (float)c.rat_h2och4[jp];
        c.rat_n2oco2_f[jp] = (float)c.rat_n2oco2[jp];
        c.rat_o3co2_f[jp] = (float)c.rat_o3co2[jp];
    }
    std::vector<double> rwgt = make_rwgt(16, ngc, ngn, ngm, wt);
    h.named["lw/rwgt"] = rwgt;
    std::vector<double> &buf = h.buf;
    buf.clear();
    int first = 0;
    int ngv[16];
    static const char *minors[][2] = {
        {"kao_mn2", "ka_mn2"}, {"kbo_mn2", "kb_mn2"}, {"kao_mn2o", "ka_mn2o"}, {"kbo_mn2o", "kb_mn2o"},
        {"kao_mo3", "ka_mo3"}, {"kbo_mo3", "kb_mo3"}, {"kao_mco2", "ka_mco2"}, {"kbo_mco2", "kb_mco2"},
        {"kao_mco", "ka_mco"}, {"kao_mo2", "ka_mo2"}, {"kbo_mo2", "kb_mo2"}, {"ccl4o", "ccl4"},
        {"cfc11adjo", "cfc11adj"}, {"cfc12o", "cfc12"}, {"cfc22adjo", "cfc22adj"}};
    for (int b = 0; b < 16; ++b) {
        LwBand &d = h.tab.band[b];
        std::memset(&d, 0xff, sizeof d);
        d.ng = ngc[b];
        ngv[b] = d.ng;
        d.ngs = b == 0 ? 0 : ngs[b - 1];
        d.nspa = nspa[b];
        d.nspb = nspb[b];
        d.delwave = delwave[b];
        d.a0 = a0[b];
        d.a1 = a1[b];
        d.a2 = a2[b];
        const int *ngn_b = &ngn[first];
        const double *rw = &rwgt[b * 16];
        char key[64];
        auto red = [&](const char *nm, bool gfirst, bool weighted, int *rows, const char *rname) {
            std::snprintf(key, sizeof key, "lw/b%02d/%s", b + 1, nm);
            std::vector<double> r = reduce(t, key, gfirst, weighted, d.ng, ngn_b, rw, rows);
            if (!r.empty()) {
                std::snprintf(key, sizeof key, "lw/b%02d/%s", b + 1, rname);
                h.named[key] = r;
            }
            return r;
        };
        int rows;
        std::vector<double> r;
        r = red("kao", false, true, &rows, "ka");
        d.absa = pack(buf, r, rows, d.ng, false);
        r = red("kbo", false, true, &rows, "kb");
        d.absb = pack(buf, r, rows, d.ng, false);
        r = red("selfrefo", false, true, &rows, "selfref");
        d.selfref = pack(buf, r, rows, d.ng, false);
        r = red("forrefo", false, true, &rows, "forref");
        d.forref = pack(buf, r, rows, d.ng, false);
        r = red("fracrefao", true, false, &rows, "fracrefa");
        d.fracrefa = pack(buf, r, rows, d.ng, true);
        d.nfa = rows;
        r = red("fracrefbo", true, false, &rows, "fracrefb");
        d.fracrefb = pack(buf, r, rows, d.ng, true);
        d.nfb = rows;
        int *slots[] = {&d.ka_mn2,  &d.kb_mn2,  &d.ka_mn2o, &d.kb_mn2o, &d.ka_mo3,
                        &d.kb_mo3,  &d.ka_mco2, &d.kb_mco2, &d.ka_mco,  &d.ka_mo2,
                        &d.kb_mo2,  &d.ccl4,    &d.cfc11adj, &d.cfc12,  &d.cfc22adj};
        for (int s = 0; s < 15; ++s) {
            r = red(minors[s][0], false, true, &rows, minors[s][1]);
            *slots[s] = pack(buf, r, rows, d.ng, false);
        }
        first += d.ng;
    }
    exp_table(buf, &h.tab.exp_tbl, &h.tab.tfn_tbl);
    h.tab.totplnk = (int)buf.size();
    buf.insert(buf.end(), totplnk.begin(), totplnk.end());  // already [band][181]
    h.tab.nwaves = build_schedule(16, ngv, h.units);
    return true;
}

}  // namespace rrtmg
