// Restart and run-data files in the reference's binary formats (host only).
//
// MPCNSRST version 1 (IOModule::WriteRestartBinFile / ReadRestartBinFile,
// src/6_io/IOModule_Bin.cpp:168-420), little-endian:
//   "MPCNSRST" | i32 version = 1 | i32 step | f64 time | i32 nblock | i32 nfield
//   per field: str name | i32 location | i32 ncomp | i32 nghost
//     per block: i32 lo[3] | i32 hi[3] | i32 active | if active:
//       f64 values for i in [lo0,hi0), j in [lo1,hi1), k in [lo2,hi2), m in [0,ncomp)
//   str = i32 length + bytes (no terminator); location codes: cell 0, faces 1-3,
//   edges 4-6, node 7.
// MPCNSRUN version 1 (RunData::WriteBinary / ReadBinary, include/6_io/RunData.h:98-205):
//   "MPCNSRUN" | i32 version = 1 | i32 step | f64 time | f64 dt
//   | i32 nref | f64 residual_ref[nref] | f64 residual_max
//   | i32 nd | (str key, f64 value) * nd | i32 ni | (str key, i32 value) * ni
#pragma once
#include <array>
#include <cstdio>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace merc {
namespace io {

struct FieldMeta {
    std::string name;
    int loc = 0, ncomp = 0, nghost = 0;
};

// one field of a restart file: per block the allocated box (lo[3], hi[3]) and its
// values in file order (empty: the block is written inactive)
struct RestartField {
    FieldMeta meta;
    std::vector<std::array<int, 6>> box;
    std::vector<std::vector<double>> data;
};

void write_restart(const std::string &path, int step, double time, int nblock, const std::vector<RestartField> &fields);

// Reads a restart file whose block count must equal nblock.  For every field,
// want(meta) decides whether its blocks are delivered; check(meta, block, box)
// validates an active block's box before its values are allocated or read (it
// throws to refuse); put(meta, block, box, values) receives the block in file
// order (skipped fields are stepped over, as the reference does).  A block whose
// values do not fit in the rest of the file is an error.  Returns step and time
// through the pointers.
void read_restart(const std::string &path, int nblock, const std::function<bool(const FieldMeta &)> &want,
                  const std::function<void(const FieldMeta &, int, const std::array<int, 6> &)> &check,
                  const std::function<void(const FieldMeta &, int, const std::array<int, 6> &, const double *)> &put, int *step,
                  double *time);

struct RunData {
    int step = 0;
    double time = 0.0, dt = 0.0;
    std::vector<double> residual_ref;
    double residual_max = 0.0;
    std::map<std::string, double> monitor_d;
    std::map<std::string, int> monitor_i;
};

void write_rundata(const std::string &path, const RunData &r);
RunData read_rundata(const std::string &path);

// Tecplot binary (#!TDV112) of ordered zones with every variable at the nodes
// (IOModule::WriteTecplotBinFile in the all-node mode,
// src/6_io/IOModule_Tecplot.cpp:691-1017):
//   "#!TDV112" | i32 1 | i32 0 | str title | i32 nvar | str name * nvar
//   per zone: f32 299 | str zone | i32 -1 | i32 -2 | f64 time | i32 -1 | i32 0
//             | i32 1 | i32 0 * nvar | i32 0 | i32 0 | i32 I | i32 J | i32 K | i32 0
//   f32 357
//   per zone: f32 299 | i32 1 * nvar | i32 0 | i32 0 | i32 -1
//             | (f64 min, f64 max) * nvar | f32 values, variable-major, i fastest
//   str = one i32 per character, then i32 0.
class TecplotWriter {
public:
    struct Zone {
        std::string name;
        int imax = 0, jmax = 0, kmax = 0;
    };
    // writes the header and every zone header
    TecplotWriter(const std::string &path, const std::string &title, const std::vector<std::string> &vars, const std::vector<Zone> &zones,
                  double time);
    ~TecplotWriter();
    // the data section of the next zone: one array per variable, i fastest
    void zone_data(const std::vector<std::vector<double>> &values);
    void close();

private:
    std::FILE *fp_ = nullptr;
    std::string path_;
    std::vector<Zone> zones_;
    size_t nvar_ = 0, next_ = 0;
};

}  // namespace io
}  // namespace merc
