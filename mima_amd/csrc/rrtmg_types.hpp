// Shared host/device data structures of the gfx950 RRTMG library.
//
// Work decomposition (both sides).  Columns are cut into tiles of TILE_COLS
// consecutive columns.  Each one-wave work unit handles one spectral band for
// up to 64/ng columns of a tile, with lane = (column-in-unit, g-point-in-band),
// g fastest.  A band's code therefore runs wave-uniformly, table rows of one
// column are read by ng adjacent lanes (contiguous, g fastest in memory), and
// the per-column g-point flux sums are segmented sums over ng adjacent lanes.
// The wave schedule of one tile (which band, which columns) is the same for
// every tile and is built once at init.
#pragma once
#include <cstdint>

namespace rrtmg {

constexpr int TILE_COLS = 60;
constexpr int NTBL = 10000;
constexpr int MAX_NLAY = 200;
constexpr int MAX_WAVES_PER_TILE = 192;

// One work unit of a tile: band index (0-based within the side), first column
// of the unit within the tile and number of columns it handles.
struct WaveUnit {
    int16_t band;
    int16_t first_col;
    int16_t ncols;
    int16_t pad;
};

// ---------------------------------------------------------------------------
// SW: per-band offsets into one packed table buffer (element offsets; -1 =
// absent).  Every 2-D table is stored row-major with the g-point index
// fastest: t[row * ng + g].
struct SwBand {
    int ng, ngs;         // g-points in band, g-points before this band
    int nspa, nspb;
    int absa, absb;      // rows = 65*nspa / 235*nspb (equivalence of ka/kb)
    int selfref, forref; // 10 / nfor rows
    int sfluxref, nsf;   // nsf rows (1 = no eta dependence)
    int raylg;           // per-g Rayleigh (bands 23, 25, 26, 27)
    int rayla, raylb;    // band 24
    int abso3a, abso3b, absch4, absco2, absh2o;
    double rayl;         // scalar Rayleigh
};

struct SwTables {
    SwBand band[14];
    int exp_tbl;         // 10001 entries
    int nwaves;          // wave units per tile
};

// Per-layer coefficient fields written by the SW prep kernel ([field][lay][col]).
enum SwField {
    SWF_FAC00, SWF_FAC01, SWF_FAC10, SWF_FAC11,
    SWF_COLH2O, SWF_COLCO2, SWF_COLO3, SWF_COLCH4, SWF_COLO2, SWF_COLMOL,
    SWF_SELFFAC, SWF_SELFFRAC, SWF_FORFAC, SWF_FORFRAC,
    SWF_COUNT
};
// the forward prep's coefficient record of one (layer, column), padded to a
// 16-byte multiple (as LWF_STRIDE)
constexpr int SWF_STRIDE = (SWF_COUNT + 3) / 4 * 4;

// ---------------------------------------------------------------------------
// LW
struct LwBand {
    int ng, ngs, nspa, nspb;
    int absa, absb, selfref, forref;
    int fracrefa, nfa, fracrefb, nfb;   // rows = eta points (1 if none)
    int ka_mn2, kb_mn2, ka_mn2o, kb_mn2o, ka_mo3, kb_mo3;
    int ka_mco2, kb_mco2, ka_mco, ka_mo2, kb_mo2;
    int ccl4, cfc11adj, cfc12, cfc22adj;
    double delwave;
    double a0, a1, a2;                  // diffusivity fit (rrtmg_lw_rtrnmr.f90:246-262)
};

struct LwTables {
    LwBand band[16];
    int exp_tbl, tfn_tbl;     // 10001 entries each
    int totplnk;              // [band][181]
    int nwaves;
};

// host-side FP64 constant tables the LW prep and band kernels index by jp
// reduced g-points of each band (parrrtm.f90, parrrsw.f90:53-66): the band
// kernels are compiled for them, and the table loader checks the table file
// against them
constexpr int lw_ng(int band)
{
    return band == 1 ? 10 : band == 2 ? 12 : band == 3 ? 16 : band == 4 ? 14 : band == 5 ? 16 : band == 6 ? 8
         : band == 7 ? 12 : band == 8 ? 8 : band == 9 ? 12 : band == 10 ? 6 : band == 11 ? 8 : band == 12 ? 8
         : band == 13 ? 4 : 2;
}
constexpr int sw_ng(int jb)
{
    return jb == 16 ? 6 : jb == 17 ? 12 : jb == 18 ? 8 : jb == 19 ? 8 : jb == 20 ? 10 : jb == 21 ? 10
         : jb == 22 ? 2 : jb == 23 ? 10 : jb == 24 ? 8 : jb == 25 ? 6 : jb == 26 ? 6 : jb == 27 ? 8
         : jb == 28 ? 6 : 12;
}

// species-parameter counts of the k-tables per band, lower/upper atmosphere
// (rrtmg_lw_init.f90:208-209, rrtmg_sw_init.f90:183-184); checked against
// the table file like the g-point counts
constexpr int LW_NSPA[16] = {1, 1, 9, 9, 9, 1, 9, 1, 9, 1, 1, 9, 9, 1, 9, 9};
constexpr int LW_NSPB[16] = {1, 1, 5, 5, 5, 0, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0};
constexpr int SW_NSPA[14] = {9, 9, 9, 9, 1, 9, 9, 1, 9, 1, 0, 1, 9, 1};
constexpr int SW_NSPB[14] = {1, 5, 1, 1, 1, 5, 1, 0, 1, 0, 0, 1, 5, 1};

// chi_mls ratios at fixed reference levels that LW taumol uses as the eta
// weights of species pairs (e.g. rrtmg_lw_taumol.f90:487-495): numerator
// species, denominator species and level, 1-based as chi_mls(i, j)
struct LwChiRatio {
    int num, den, lev;
};
constexpr LwChiRatio LW_CHI_RATIOS[] = {{1, 2, 3}, {1, 2, 5}, {1, 2, 7},  {1, 2, 9},  {1, 2, 10}, {1, 2, 11},
                                        {1, 2, 13}, {3, 2, 13}, {3, 2, 43}, {1, 3, 3}, {1, 4, 1},  {1, 4, 3},
                                        {1, 4, 5}, {4, 2, 1},  {1, 6, 3},  {1, 6, 6}, {1, 6, 9}};
constexpr int LW_NCHI_RATIO = (int)(sizeof(LW_CHI_RATIOS) / sizeof(LW_CHI_RATIOS[0]));
constexpr int lw_chi_ratio_index(int num, int den, int lev)
{
    for (int k = 0; k < LW_NCHI_RATIO; ++k)
        if (LW_CHI_RATIOS[k].num == num && LW_CHI_RATIOS[k].den == den && LW_CHI_RATIOS[k].lev == lev) return k;
    return -1;
}

struct LwConst {
    double pref[59], preflog[59], tref[59];
    double chi_mls[7 * 59];
    // chi_mls ratios by reference level (rrtmg_lw_setcoef.f90:343-392)
    double rat_h2oco2[60], rat_h2oo3[60], rat_h2on2o[60], rat_h2och4[60], rat_n2oco2[60], rat_o3co2[60];
    double heatfac;
    // the LW_CHI_RATIOS quotients, each one FP64 division
    double chi_ratio[LW_NCHI_RATIO];
    // FP32 copies (each FP64 value rounded once) for the FP32 band kernels
    float chi_mls_f[7 * 59];
    float rat_h2oco2_f[60], rat_h2oo3_f[60], rat_h2on2o_f[60], rat_h2och4_f[60], rat_n2oco2_f[60], rat_o3co2_f[60];
    float chi_ratio_f[LW_NCHI_RATIO];
};

enum LwField {
    LWF_FAC00, LWF_FAC01, LWF_FAC10, LWF_FAC11,
    LWF_COLH2O, LWF_COLCO2, LWF_COLO3, LWF_COLN2O, LWF_COLCO, LWF_COLCH4, LWF_COLO2, LWF_COLBRD,
    LWF_SELFFAC, LWF_SELFFRAC, LWF_FORFAC, LWF_FORFRAC,
    LWF_MINORFRAC, LWF_SCALEMINOR, LWF_SCALEMINORN2, LWF_COLDRY, LWF_PAVEL,
    LWF_WX1, LWF_WX2, LWF_WX3, LWF_WX4, LWF_TLAYFRAC,
    LWF_COUNT
};
// the forward prep's coefficient record of one (layer, column): LWF_COUNT
// fields padded to a 16-byte multiple, so a band kernel lane loads its
// column's fields with a few 16-byte loads
constexpr int LWF_STRIDE = (LWF_COUNT + 3) / 4 * 4;

// The backward's coefficient records: per (layer, column) a value record of
// LWF_STRIDE / SWF_STRIDE fields (the forward's record) and a derivative
// record holding only the derivatives a field can have.  Which of (d/dT,
// d/dq, d/dO3) each field depends on (bits 0, 1, 2; the preps' formulas):
// the interpolation fractions and the temperature scalings on T; column
// amounts on q through the dry-air column (O3's also on O3); the broadening
// column and the N2 minor scaling on q and O3 (they exclude the listed gases,
// the latter also on T); selffac/forfac (stored times the water column) on T
// and q.  The derivative record lists each field's components in field order,
// d/dT before d/dq before d/dO3, padded to a 16-byte multiple.
constexpr unsigned LW_DMASK[LWF_COUNT] = {1, 1, 1, 1, 2, 2, 6, 2, 2, 2, 2, 6, 3,
                                          1, 3, 1, 1, 1, 7, 2, 0, 2, 2, 2, 2, 1};
constexpr unsigned SW_DMASK[SWF_COUNT] = {1, 1, 1, 1, 2, 2, 6, 2, 2, 2, 3, 1, 3, 1};
constexpr int dmask_count(unsigned m) { return (int)(m & 1u) + (int)((m >> 1) & 1u) + (int)((m >> 2) & 1u); }
// offset of field f's first derivative in the derivative record
template <int N>
constexpr int der_offset(const unsigned (&mask)[N], int f)
{
    int o = 0;
    for (int i = 0; i < f; ++i) o += dmask_count(mask[i]);
    return o;
}
constexpr int LW_DSTRIDE = (der_offset(LW_DMASK, LWF_COUNT) + 3) / 4 * 4;
constexpr int SW_DSTRIDE = (der_offset(SW_DMASK, SWF_COUNT) + 3) / 4 * 4;
static_assert(der_offset(LW_DMASK, LWF_COUNT) == 31 && LW_DSTRIDE == 32, "LW derivative record");
static_assert(der_offset(SW_DMASK, SWF_COUNT) == 17 && SW_DSTRIDE == 20, "SW derivative record");

// packed per-layer integer word: jp | jt<<6 | jt1<<9 | indself<<12 | indfor<<16 | indminor<<18 | indlay<<23
__host__ __device__ inline int pack_lw(int jp, int jt, int jt1, int inds, int indf, int indm, int indlay)
{
    return jp | (jt << 6) | (jt1 << 9) | (inds << 12) | (indf << 16) | (indm << 18) | (indlay << 23);
}
// SW word: jp | jt<<6 | jt1<<9 | indself<<12 | indfor<<16
__host__ __device__ inline int pack_sw(int jp, int jt, int jt1, int inds, int indf)
{
    return jp | (jt << 6) | (jt1 << 9) | (inds << 12) | (indf << 16);
}
__host__ __device__ inline int unpack_jp(int w) { return w & 63; }
__host__ __device__ inline int unpack_jt(int w) { return (w >> 6) & 7; }
__host__ __device__ inline int unpack_jt1(int w) { return (w >> 9) & 7; }
__host__ __device__ inline int unpack_inds(int w) { return (w >> 12) & 15; }
__host__ __device__ inline int unpack_indf(int w) { return (w >> 16) & 3; }
__host__ __device__ inline int unpack_indm(int w) { return (w >> 18) & 31; }
__host__ __device__ inline int unpack_indlay(int w) { return (w >> 23) & 255; }

}  // namespace rrtmg
