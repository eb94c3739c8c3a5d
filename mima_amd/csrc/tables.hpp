// Host-side table preparation (see tables.cpp).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "rrtmg_types.hpp"

namespace rrtmg {

class TableFile {
  public:
    struct Entry {
        uint32_t dtype = 0, ndim = 0;
        int32_t lb[4] = {1, 1, 1, 1}, ext[4] = {1, 1, 1, 1};
        uint64_t off = 0, nbytes = 0;
    };
    bool load(const std::string &path, std::string &err);
    const Entry *find(const std::string &name) const;
    std::vector<double> doubles(const std::string &name) const;
    std::vector<int> ints(const std::string &name) const;
    uint32_t flags() const { return flags_; }

  private:
    std::vector<unsigned char> raw_;
    std::map<std::string, Entry> entries_;
    uint32_t flags_ = 0;
};

struct SwHost {
    SwTables tab;
    std::vector<double> buf;       // packed reduced tables + exp table (FP64)
    std::vector<WaveUnit> units;   // wave schedule of one tile
    std::vector<double> pref, preflog, tref;
    double heatfac = 0.0;
    std::map<std::string, std::vector<double>> named;  // reduced tables, Fortran order
};

struct LwHost {
    LwTables tab;
    LwConst cst;
    std::vector<double> buf;
    std::vector<WaveUnit> units;
    bool synthetic = false;
    std::map<std::string, std::vector<double>> named;
};

bool build_sw(const std::string &path, double cpdair, SwHost &h, std::string &err);
bool build_lw(const std::string &path, double cpdair, LwHost &h, std::string &err);

}  // namespace rrtmg
