// Chain diagnostics of the drop-in C++ interface, on the device.
// Replaces DEC::max_abs_curl_grad and DEC::max_abs_div_curl
// (include/8_dec/DecDiagnostics.h:8-25); each returns the host value, so it
// waits for its stream.  Uses the HIP runtime API for the one-double result
// (compile with the ROCm include path and link libamdhip64).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>

#include "DecOps.h"

namespace DEC {

namespace detail {
// one device double for a result, read back after the stream drains
struct DeviceScalar {
    double *d = nullptr;
    DeviceScalar() {
        if (hipMalloc((void **)&d, sizeof(double)) != hipSuccess) throw std::runtime_error("DEC: hipMalloc failed");
    }
    ~DeviceScalar() {
        if (d) (void)hipFree(d);
    }
    double read(void *stream) const {
        double h = 0.0;
        if (hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
            hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
            throw std::runtime_error("DEC: result read-back failed");
        return h;
    }
};
}  // namespace detail

inline double max_abs_curl_grad(FieldBlock &node, void *stream = nullptr) {
    detail::DeviceScalar r;
    mercury_hip_detail::check(mercury_dec_max_abs_curl_grad(node.cells(), node.nghost(), node.data(), r.d, stream),
                              "DEC::max_abs_curl_grad");
    return r.read(stream);
}

inline double max_abs_div_curl(FieldBlock &edge_xi, FieldBlock &edge_eta, FieldBlock &edge_zeta, void *stream = nullptr) {
    detail::same_block(edge_xi, edge_eta);
    detail::same_block(edge_xi, edge_zeta);
    detail::DeviceScalar r;
    const double *e[3] = {edge_xi.data(), edge_eta.data(), edge_zeta.data()};
    mercury_hip_detail::check(mercury_dec_max_abs_div_curl(edge_xi.cells(), edge_xi.nghost(), e, r.d, stream), "DEC::max_abs_div_curl");
    return r.read(stream);
}

inline double max_abs_curl_grad(Field &fields, const std::string &node_name, int node_comp = 0) {
    void *s = detail::stream_of(fields);
    double m = 0.0;
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock n = fields.field(node_name, ib, node_comp);
        m = std::max(m, max_abs_curl_grad(n, s));
    }
    return m;
}

inline double max_abs_div_curl(Field &fields, const EdgeFormNames &e, int edge_comp = 0) {
    void *s = detail::stream_of(fields);
    double m = 0.0;
    for (int ib = 0; ib < fields.num_blocks(); ++ib) {
        FieldBlock x = fields.field(e.xi, ib, edge_comp), y = fields.field(e.eta, ib, edge_comp), z = fields.field(e.zeta, ib, edge_comp);
        m = std::max(m, max_abs_div_curl(x, y, z, s));
    }
    return m;
}

}  // namespace DEC
