// A C++ host program written against the drop-in headers the way
// Z4_Mercury/main.cpp:27-99 is written against the reference's classes: Grid,
// Topology, Field, Halo, MercurySolver, then Advance, with DEC diagnostics on
// the device fields.  Driven by tests/test_gpu_cpp_caller.py, which writes the
// case to a binary file, runs this program and compares what it writes back
// with the CPU oracle.
//
//   mercury_caller <in.bin> <out.bin> <nsteps> <integrator>
//
// in.bin (little endian): int32 nb, g; int32 dims[3 nb]; int32 bc[6 nb];
// int32 nbr[6 nb]; int32 transform[54 nb]; float64 owned nodes; float64
// MercuryPhysics (as a struct); int32 ndip; float64 dipoles[7 ndip];
// int32 nfields, then per field: int32 name length, name bytes, int32 ncomp,
// float64 values (per block, per component, the block box).
// out.bin: float64 per step dt, float64 normalised div B, float64
// max|d2 d1 E|, float64 max|d1 d0 (node)| (0 by construction), then U_H, U_Na,
// B_xi, B_eta, B_zeta per block and component.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "MercurySolver.h"
#include "8_dec/DecDiagnostics.h"
#include "8_dec/DecOps.h"
#include "7_metric/SingularEdgeRegistry.h"

namespace {
template <typename T> T get(std::ifstream &in) {
    T v{};
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!in) throw std::runtime_error("truncated input");
    return v;
}
template <typename T> std::vector<T> get_n(std::ifstream &in, size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char *>(v.data()), sizeof(T) * n);
    if (!in) throw std::runtime_error("truncated input");
    return v;
}
long long box_volume(const int *m, int g) { return (long long)(m[0] + 1 + 2 * g) * (m[1] + 1 + 2 * g) * (m[2] + 1 + 2 * g); }
}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s in.bin out.bin nsteps integrator\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        if (!in) throw std::runtime_error("cannot open input");
        const int nb = get<int>(in), g = get<int>(in);
        const std::vector<int> dims = get_n<int>(in, 3 * nb);
        TOPO::Topology topo;
        topo.bc = get_n<int>(in, 6 * nb);
        topo.nbr = get_n<int>(in, 6 * nb);
        topo.transform = get_n<int>(in, 54 * nb);
        const int npatch = get<int>(in);
        topo.rule = get<int>(in);
        topo.patches = get_n<int>(in, 18 * (size_t)npatch);
        size_t nn = 0;
        for (int b = 0; b < nb; ++b) nn += 3 * (size_t)(dims[3 * b] + 1) * (dims[3 * b + 1] + 1) * (dims[3 * b + 2] + 1);
        const std::vector<double> nodes = get_n<double>(in, nn);
        const MercuryPhysics phys = get<MercuryPhysics>(in);
        const int ndip = get<int>(in);
        const std::vector<double> dip = get_n<double>(in, 7 * (size_t)ndip);

        Grid grid;
        grid.SetBlocks(dims, nodes);
        Field field(&grid, topo, phys, g);
        MercurySolver::RegisterFields(&field, g);
        METRIC::SingularEdgeRegistry singular_edges;
        singular_edges.build(topo, field, grid, 0);
        singular_edges.validate_or_abort();
        std::cout << "[SingularEdgeRegistry] physical singular edges=" << singular_edges.size() << "\n";
        Halo halo(&field, &topo);
        MercurySolver::RegisterHaloFields(&field, &halo);
        MercurySolver solver(&grid, &topo, &field, &halo, nullptr, dip);

        const int nf = get<int>(in);
        for (int f = 0; f < nf; ++f) {
            const int len = get<int>(in);
            std::vector<char> name = get_n<char>(in, (size_t)len);
            const std::string fname(name.begin(), name.end());
            const int nc = get<int>(in);
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < nc; ++c) {
                    FieldBlock fb = field.field(fname, b, c);
                    fb.upload(get_n<double>(in, (size_t)box_volume(&dims[3 * b], g)));
                }
        }
        solver.Prepare();
        std::vector<double> dts;
        solver.Advance(std::stoi(argv[3]), std::stoi(argv[4]), &dts);
        halo.sync_group("Ucell");  // idempotent after a step (per-face connections, patches under the position rule)
        const double divb = solver.NormalisedDivB();
        const double dc = DEC::max_abs_div_curl(field, DEC::EdgeFormNames{"E_xi", "E_eta", "E_zeta"});
        // d1 d0 of the node scratch field: the reference's DEC chain identity on solver storage
        {
            FieldBlock n = field.field("node_scalar", 0);
            std::vector<double> v((size_t)n.volume());
            for (size_t q = 0; q < v.size(); ++q) v[q] = (double)((q * 2654435761u) % 1000u) - 500.0;
            n.upload(v);
        }
        const double cg = DEC::max_abs_curl_grad(field, "node_scalar");

        std::ofstream out(argv[2], std::ios::binary);
        out.write(reinterpret_cast<const char *>(dts.data()), sizeof(double) * dts.size());
        out.write(reinterpret_cast<const char *>(&divb), sizeof(double));
        out.write(reinterpret_cast<const char *>(&dc), sizeof(double));
        out.write(reinterpret_cast<const char *>(&cg), sizeof(double));
        const char *state[][2] = {{"U_H", "5"}, {"U_Na", "5"}, {"B_xi", "1"}, {"B_eta", "1"}, {"B_zeta", "1"}};
        for (const auto &s : state)
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < std::atoi(s[1]); ++c) {
                    const std::vector<double> v = field.field(s[0], b, c).download();
                    out.write(reinterpret_cast<const char *>(v.data()), sizeof(double) * v.size());
                }
        if (!out) throw std::runtime_error("cannot write output");
        std::printf("mercury_caller: %zu steps on %d blocks, normalised div B %.3e, max|d2 d1 E| %.3e\n", dts.size(), nb, divb, dc);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mercury_caller: %s\n", e.what());
        return 1;
    }
    return 0;
}
