// Discrete exterior derivatives of the drop-in C++ interface, on the device.
// Replaces DEC::d0_node_to_edge, d1_edge_to_face, d2_face_to_cell and
// faraday_update_face_2form (include/8_dec/DecOps.h:8-68) with the same
// argument order; each call runs on the field context's stream and returns
// without waiting.  The FieldBlock overloads take one component per block
// (node_comp / edge_comp of the reference select it when the views are made).
#pragma once

#include <string>

#include "../3_field/Field.h"

namespace DEC {

// include/8_dec/DecTypes.h
struct EdgeFormNames {
    std::string xi, eta, zeta;
};
struct FaceFormNames {
    std::string xi, eta, zeta;
};

namespace detail {
inline void *stream_of(const Field &f) {
    void *s = nullptr;
    mercury_hip_detail::check(mercury_stream(f.context(), &s), "DEC");
    return s;
}
inline void same_block(const FieldBlock &a, const FieldBlock &b) {
    if (a.block() != b.block() || a.nghost() != b.nghost() || a.cells()[0] != b.cells()[0] || a.cells()[1] != b.cells()[1] ||
        a.cells()[2] != b.cells()[2])
        throw std::runtime_error("DEC: the field blocks belong to different blocks");
}
}  // namespace detail

inline void d0_node_to_edge(FieldBlock &node, FieldBlock &edge_xi, FieldBlock &edge_eta, FieldBlock &edge_zeta, void *stream = nullptr) {
    detail::same_block(node, edge_xi);
    detail::same_block(node, edge_eta);
    detail::same_block(node, edge_zeta);
    double *e[3] = {edge_xi.data(), edge_eta.data(), edge_zeta.data()};
    mercury_hip_detail::check(mercury_dec_d0(node.cells(), node.nghost(), node.data(), e, stream), "DEC::d0_node_to_edge");
}

inline void d1_edge_to_face(FieldBlock &edge_xi, FieldBlock &edge_eta, FieldBlock &edge_zeta, FieldBlock &face_xi,
                            FieldBlock &face_eta, FieldBlock &face_zeta, void *stream = nullptr) {
    for (FieldBlock *f : {&edge_eta, &edge_zeta, &face_xi, &face_eta, &face_zeta}) detail::same_block(edge_xi, *f);
    const double *e[3] = {edge_xi.data(), edge_eta.data(), edge_zeta.data()};
    double *f[3] = {face_xi.data(), face_eta.data(), face_zeta.data()};
    mercury_hip_detail::check(mercury_dec_d1(edge_xi.cells(), edge_xi.nghost(), e, f, stream), "DEC::d1_edge_to_face");
}

inline void d2_face_to_cell(FieldBlock &face_xi, FieldBlock &face_eta, FieldBlock &face_zeta, FieldBlock &cell, void *stream = nullptr) {
    for (FieldBlock *f : {&face_eta, &face_zeta, &cell}) detail::same_block(face_xi, *f);
    const double *f[3] = {face_xi.data(), face_eta.data(), face_zeta.data()};
    mercury_hip_detail::check(mercury_dec_d2(face_xi.cells(), face_xi.nghost(), f, cell.data(), stream), "DEC::d2_face_to_cell");
}

inline void faraday_update_face_2form(FieldBlock &edge_xi, FieldBlock &edge_eta, FieldBlock &edge_zeta, FieldBlock &face_xi,
                                      FieldBlock &face_eta, FieldBlock &face_zeta, double dt, void *stream = nullptr) {
    for (FieldBlock *f : {&edge_eta, &edge_zeta, &face_xi, &face_eta, &face_zeta}) detail::same_block(edge_xi, *f);
    const double *e[3] = {edge_xi.data(), edge_eta.data(), edge_zeta.data()};
    double *f[3] = {face_xi.data(), face_eta.data(), face_zeta.data()};
    mercury_hip_detail::check(mercury_dec_faraday_update(edge_xi.cells(), edge_xi.nghost(), e, f, dt, stream),
                              "DEC::faraday_update_face_2form");
}

// Field overloads: every block, on the context's stream (the reference loops
// over blocks the same way, DecExteriorDerivative.cpp:79-97 and siblings)
inline void d0_node_to_edge(Field &fields, const std::string &node_name, const EdgeFormNames &e, int node_comp = 0, int edge_comp = 0) {
    void *s = detail::stream_of(fields);
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock n = fields.field(node_name, ib, node_comp), x = fields.field(e.xi, ib, edge_comp),
                   y = fields.field(e.eta, ib, edge_comp), z = fields.field(e.zeta, ib, edge_comp);
        d0_node_to_edge(n, x, y, z, s);
    }
}

inline void d1_edge_to_face(Field &fields, const EdgeFormNames &e, const FaceFormNames &f, int edge_comp = 0, int face_comp = 0) {
    void *s = detail::stream_of(fields);
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock ex = fields.field(e.xi, ib, edge_comp), ey = fields.field(e.eta, ib, edge_comp), ez = fields.field(e.zeta, ib, edge_comp);
        FieldBlock fx = fields.field(f.xi, ib, face_comp), fy = fields.field(f.eta, ib, face_comp), fz = fields.field(f.zeta, ib, face_comp);
        d1_edge_to_face(ex, ey, ez, fx, fy, fz, s);
    }
}

inline void d2_face_to_cell(Field &fields, const FaceFormNames &f, const std::string &cell_name, int face_comp = 0, int cell_comp = 0) {
    void *s = detail::stream_of(fields);
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock fx = fields.field(f.xi, ib, face_comp), fy = fields.field(f.eta, ib, face_comp), fz = fields.field(f.zeta, ib, face_comp);
        FieldBlock c = fields.field(cell_name, ib, cell_comp);
        d2_face_to_cell(fx, fy, fz, c, s);
    }
}

inline void faraday_update_face_2form(Field &fields, const EdgeFormNames &e, const FaceFormNames &f, double dt, int edge_comp = 0,
                                      int face_comp = 0) {
    void *s = detail::stream_of(fields);
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock ex = fields.field(e.xi, ib, edge_comp), ey = fields.field(e.eta, ib, edge_comp), ez = fields.field(e.zeta, ib, edge_comp);
        FieldBlock fx = fields.field(f.xi, ib, face_comp), fy = fields.field(f.eta, ib, face_comp), fz = fields.field(f.zeta, ib, face_comp);
        faraday_update_face_2form(ex, ey, ez, fx, fy, fz, dt, s);
    }
}

}  // namespace DEC
