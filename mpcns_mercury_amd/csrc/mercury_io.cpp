// Restart and run-data files in the reference's binary formats (see mercury_io.h).
#include "mercury_io.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace merc {
namespace io {
namespace {

void fail(const std::string &m) { throw std::runtime_error(m); }

void put_i32(std::ofstream &o, int32_t v) { o.write(reinterpret_cast<const char *>(&v), sizeof v); }
void put_f64(std::ofstream &o, double v) { o.write(reinterpret_cast<const char *>(&v), sizeof v); }
void put_str(std::ofstream &o, const std::string &s) {
    put_i32(o, (int32_t)s.size());
    if (!s.empty()) o.write(s.data(), (std::streamsize)s.size());
}
int32_t get_i32(std::ifstream &in, const char *what) {
    int32_t v = 0;
    in.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!in) fail(std::string("truncated file while reading ") + what);
    return v;
}
double get_f64(std::ifstream &in, const char *what) {
    double v = 0.0;
    in.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!in) fail(std::string("truncated file while reading ") + what);
    return v;
}
std::string get_str(std::ifstream &in) {
    const int32_t n = get_i32(in, "a string length");
    if (n < 0) fail("negative string length");
    std::string s((size_t)n, '\0');
    if (n > 0) in.read(&s[0], n);
    if (!in) fail("truncated string");
    return s;
}

}  // namespace

void write_restart(const std::string &path, int step, double time, int nblock, const std::vector<RestartField> &fields) {
    std::ofstream out(path, std::ios::binary);
    if (!out) fail("cannot open " + path + " for writing");
    out.write("MPCNSRST", 8);
    put_i32(out, 1);
    put_i32(out, step);
    put_f64(out, time);
    put_i32(out, nblock);
    put_i32(out, (int32_t)fields.size());
    for (const RestartField &f : fields) {
        put_str(out, f.meta.name);
        put_i32(out, f.meta.loc);
        put_i32(out, f.meta.ncomp);
        put_i32(out, f.meta.nghost);
        for (int b = 0; b < nblock; ++b) {
            for (int c = 0; c < 6; ++c) put_i32(out, f.box[b][c]);
            const bool active = !f.data[b].empty();
            put_i32(out, active ? 1 : 0);
            if (active) out.write(reinterpret_cast<const char *>(f.data[b].data()), (std::streamsize)(f.data[b].size() * sizeof(double)));
        }
    }
    if (!out) fail("write failed: " + path);
}

void read_restart(const std::string &path, int nblock, const std::function<bool(const FieldMeta &)> &want,
                  const std::function<void(const FieldMeta &, int, const std::array<int, 6> &)> &check,
                  const std::function<void(const FieldMeta &, int, const std::array<int, 6> &, const double *)> &put, int *step,
                  double *time) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) fail("cannot open " + path);
    const long long fsize = (long long)in.tellg();
    in.seekg(0, std::ios::beg);
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, "MPCNSRST", 8) != 0) fail("not an MPCNSRST file: " + path);
    const int32_t version = get_i32(in, "the version");
    if (version != 1) fail("unsupported restart version " + std::to_string(version));
    const int32_t st = get_i32(in, "the step");
    const double tm = get_f64(in, "the time");
    const int32_t nb = get_i32(in, "the block count");
    const int32_t nf = get_i32(in, "the field count");
    if (nb != nblock) fail("restart block count " + std::to_string(nb) + " differs from the context's " + std::to_string(nblock));
    std::vector<double> buf;
    for (int32_t f = 0; f < nf; ++f) {
        FieldMeta m;
        m.name = get_str(in);
        m.loc = get_i32(in, "a location");
        m.ncomp = get_i32(in, "a component count");
        m.nghost = get_i32(in, "a ghost depth");
        if (m.ncomp < 0) fail("negative component count for " + m.name);
        const bool take = want(m);
        for (int b = 0; b < nb; ++b) {
            std::array<int, 6> box;
            for (int c = 0; c < 6; ++c) box[c] = get_i32(in, "a block box");
            const long long ni = box[3] - box[0], nj = box[4] - box[1], nk = box[5] - box[2];
            if (ni < 0 || nj < 0 || nk < 0) fail("negative extent for " + m.name);
            const int32_t active = get_i32(in, "an active flag");
            if (!active) continue;
            const long long nval = ni * nj * nk * m.ncomp;
            // a block's values must fit in what is left of the file, and a field
            // the caller takes must match its allocation, before anything is
            // allocated or read
            if (nval > (fsize - (long long)in.tellg()) / (long long)sizeof(double)) fail("truncated data of " + m.name);
            if (take) check(m, b, box);
            if (!take) {
                in.seekg((std::streamoff)(nval * (long long)sizeof(double)), std::ios::cur);
                if (!in) fail("truncated data of " + m.name);
                continue;
            }
            buf.resize((size_t)nval);
            in.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)(nval * (long long)sizeof(double)));
            if (!in) fail("truncated data of " + m.name);
            put(m, b, box, buf.data());
        }
    }
    if (step) *step = st;
    if (time) *time = tm;
}

void write_rundata(const std::string &path, const RunData &r) {
    std::ofstream out(path, std::ios::binary);
    if (!out) fail("cannot open " + path + " for writing");
    out.write("MPCNSRUN", 8);
    put_i32(out, 1);
    put_i32(out, r.step);
    put_f64(out, r.time);
    put_f64(out, r.dt);
    put_i32(out, (int32_t)r.residual_ref.size());
    for (double v : r.residual_ref) put_f64(out, v);
    put_f64(out, r.residual_max);
    put_i32(out, (int32_t)r.monitor_d.size());
    for (const auto &kv : r.monitor_d) {
        put_str(out, kv.first);
        put_f64(out, kv.second);
    }
    put_i32(out, (int32_t)r.monitor_i.size());
    for (const auto &kv : r.monitor_i) {
        put_str(out, kv.first);
        put_i32(out, kv.second);
    }
    if (!out) fail("write failed: " + path);
}

RunData read_rundata(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("cannot open " + path);
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, "MPCNSRUN", 8) != 0) fail("not an MPCNSRUN file: " + path);
    if (get_i32(in, "the version") != 1) fail("unsupported run-data version");
    RunData r;
    r.step = get_i32(in, "the step");
    r.time = get_f64(in, "the time");
    r.dt = get_f64(in, "dt");
    int32_t nref = get_i32(in, "the reference count");
    if (nref < 0) nref = 0;
    r.residual_ref.resize((size_t)nref);
    for (auto &v : r.residual_ref) v = get_f64(in, "a residual reference");
    r.residual_max = get_f64(in, "the residual maximum");
    int32_t nd = get_i32(in, "the real monitor count");
    for (int32_t i = 0; i < nd; ++i) {
        const std::string k = get_str(in);
        r.monitor_d[k] = get_f64(in, "a real monitor");
    }
    int32_t ni = get_i32(in, "the integer monitor count");
    for (int32_t i = 0; i < ni; ++i) {
        const std::string k = get_str(in);
        r.monitor_i[k] = get_i32(in, "an integer monitor");
    }
    return r;
}

namespace {
void tec_i32(std::FILE *fp, int32_t v) { std::fwrite(&v, 4, 1, fp); }
void tec_f32(std::FILE *fp, float v) { std::fwrite(&v, 4, 1, fp); }
void tec_f64(std::FILE *fp, double v) { std::fwrite(&v, 8, 1, fp); }
void tec_str(std::FILE *fp, const std::string &s) {
    for (char c : s) tec_i32(fp, (int32_t)c);
    tec_i32(fp, 0);
}
}  // namespace

TecplotWriter::TecplotWriter(const std::string &path, const std::string &title, const std::vector<std::string> &vars,
                             const std::vector<Zone> &zones, double time)
    : path_(path), zones_(zones), nvar_(vars.size()) {
    fp_ = std::fopen(path.c_str(), "wb");
    if (!fp_) fail("cannot open " + path + " for writing");
    std::fwrite("#!TDV112", 8, 1, fp_);
    tec_i32(fp_, 1);  // byte order: little endian
    tec_i32(fp_, 0);  // full file
    tec_str(fp_, title);
    tec_i32(fp_, (int32_t)vars.size());
    for (const auto &v : vars) tec_str(fp_, v);
    for (const Zone &z : zones_) {
        tec_f32(fp_, 299.0f);
        tec_str(fp_, z.name);
        tec_i32(fp_, -1);  // parent zone
        tec_i32(fp_, -2);  // strand
        tec_f64(fp_, time);
        tec_i32(fp_, -1);  // colour
        tec_i32(fp_, 0);   // ordered
        tec_i32(fp_, 1);   // variable locations given
        for (size_t v = 0; v < nvar_; ++v) tec_i32(fp_, 0);  // all nodal
        tec_i32(fp_, 0);   // raw face neighbours
        tec_i32(fp_, 0);   // miscellaneous face neighbours
        tec_i32(fp_, z.imax);
        tec_i32(fp_, z.jmax);
        tec_i32(fp_, z.kmax);
        tec_i32(fp_, 0);  // auxiliary data
    }
    tec_f32(fp_, 357.0f);  // end of header
}

TecplotWriter::~TecplotWriter() {
    if (fp_) std::fclose(fp_);
}

void TecplotWriter::zone_data(const std::vector<std::vector<double>> &values) {
    if (!fp_) fail("tecplot writer is closed");
    if (next_ >= zones_.size()) fail("more zone data than zones");
    if (values.size() != nvar_) fail("zone data: wrong variable count");
    const Zone &z = zones_[next_++];
    const size_t n = (size_t)z.imax * z.jmax * z.kmax;
    tec_f32(fp_, 299.0f);
    for (size_t v = 0; v < nvar_; ++v) tec_i32(fp_, 1);  // float
    tec_i32(fp_, 0);   // passive variables
    tec_i32(fp_, 0);   // shared variables
    tec_i32(fp_, -1);  // shared connectivity
    for (const auto &a : values) {
        if (a.size() != n) fail("zone data: wrong value count");
        double lo = n ? a[0] : 0.0, hi = lo;
        for (double x : a) {
            if (x < lo) lo = x;
            if (x > hi) hi = x;
        }
        tec_f64(fp_, lo);
        tec_f64(fp_, hi);
    }
    std::vector<float> buf;
    for (const auto &a : values) {
        buf.resize(n);
        for (size_t t = 0; t < n; ++t) buf[t] = (float)a[t];
        std::fwrite(buf.data(), sizeof(float), n, fp_);
    }
}

void TecplotWriter::close() {
    if (!fp_) return;
    if (next_ != zones_.size()) fail("tecplot file closed before every zone was written");
    const bool bad = std::ferror(fp_) != 0;
    std::fclose(fp_);
    fp_ = nullptr;
    if (bad) fail("write failed: " + path_);
}

}  // namespace io
}  // namespace merc
