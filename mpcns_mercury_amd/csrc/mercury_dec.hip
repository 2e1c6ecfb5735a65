// Exterior derivatives of discrete forms on one block held in caller device
// arrays, and the chain diagnostics max|d2 d1| and max|d1 d0|
// (src/8_dec/DecExteriorDerivative.cpp, src/8_dec/DecDiagnostics.cpp).
//
// Every array spans the block box: N_a = m_a + 1 + 2g points per axis for a
// block of m cells, k fastest, point (i,j,k) at ((i+g)*N1 + (j+g))*N2 + (k+g)
// (the layout of every per-block array of this library).  Each operator loops
// over the owned range of its output location, [0, m + 1 - delta) per axis
// (FieldBlock::inner_lo / inner_hi, include/3_field/FieldBlock.h:34-35), and
// reads its inputs at owned index and owned index + 1 only, so ghost layers
// are neither read nor written.
//
// One thread per point of a launch box, k fastest (unit-stride, coalesced
// wave64 accesses), workgroups dealt to the XCDs in contiguous slabs.  The two
// diagnostics reduce to one partial maximum per workgroup (wave64 shuffles,
// then LDS) and fold the partials in a one-workgroup kernel: no atomics, and
// the result equals the reference's sequential std::max loop exactly (every
// partial starts from 0.0 and drops NaN the way std::max(acc, NaN) does).
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "mercury_dec.h"
#include "mercury_device.h"

// Scratch of the no-allocation diagnostics: the partial maxima of up to
// MERC_DEC_SCRATCH_DOUBLES workgroups (1 MiB), one copy per GPU for the life
// of the process.  Host code never passes the array's name to a kernel: it
// uses the device address hipGetSymbolAddress returns for the current device
// (dec_noalloc_scratch below).
static __device__ __align__(256) double scratch[MERC_DEC_SCRATCH_DOUBLES];

namespace {
using merc_dev::dmax;
using merc_dev::xcd_block;

struct DBox {
    int m0, m1, m2;   // cells per axis
    int g;            // ghost layers
    long long sI, sJ; // strides of i and j (k stride 1)
    int n0, n1, n2;   // launch box extents
};

__device__ __forceinline__ bool point(const DBox &b, int &i, int &j, int &k, long long &q) {
    const long long t = (long long)xcd_block() * blockDim.x + threadIdx.x;
    const long long n12 = (long long)b.n1 * b.n2;
    if (t >= (long long)b.n0 * n12) return false;
    i = (int)(t / n12);
    const long long r = t - (long long)i * n12;
    j = (int)(r / b.n2);
    k = (int)(r - (long long)j * b.n2);
    q = (long long)(i + b.g) * b.sI + (long long)(j + b.g) * b.sJ + (k + b.g);
    return true;
}

// d0: node 0-form -> edge 1-form (DecExteriorDerivative.cpp:41-77); launch box = owned nodes
__global__ void __launch_bounds__(256) k_d0(const DBox b, const double *__restrict__ n, double *__restrict__ ex,
                                            double *__restrict__ ey, double *__restrict__ ez) {
    int i, j, k;
    long long q;
    if (!point(b, i, j, k, q)) return;
    const double c = n[q];
    if (i < b.m0) ex[q] = n[q + b.sI] - c;
    if (j < b.m1) ey[q] = n[q + b.sJ] - c;
    if (k < b.m2) ez[q] = n[q + 1] - c;
}

// d1 at one face (d1_xi_at / d1_eta_at / d1_zeta_at, DecExteriorDerivative.cpp:7-39)
__device__ __forceinline__ double d1_xi(const DBox &b, const double *ey, const double *ez, long long q) {
    return (ey[q] - ey[q + 1]) + (ez[q + b.sJ] - ez[q]);
}
__device__ __forceinline__ double d1_eta(const DBox &b, const double *ex, const double *ez, long long q) {
    return (ex[q + 1] - ex[q]) + (ez[q] - ez[q + b.sI]);
}
__device__ __forceinline__ double d1_zeta(const DBox &b, const double *ex, const double *ey, long long q) {
    return (ex[q] - ex[q + b.sJ]) + (ey[q + b.sI] - ey[q]);
}

// d1: edge 1-form -> face 2-form (:99-137), or with update != 0 Faraday's law
// face -= dt * d1(edge) (faraday_update_face_2form, :202-241); launch box = owned nodes
__global__ void __launch_bounds__(256) k_d1(const DBox b, const double *__restrict__ ex, const double *__restrict__ ey,
                                            const double *__restrict__ ez, double *__restrict__ fx, double *__restrict__ fy,
                                            double *__restrict__ fz, int update, double dt) {
    int i, j, k;
    long long q;
    if (!point(b, i, j, k, q)) return;
    if (j < b.m1 && k < b.m2) {
        const double v = d1_xi(b, ey, ez, q);
        fx[q] = update ? fx[q] - dt * v : v;
    }
    if (i < b.m0 && k < b.m2) {
        const double v = d1_eta(b, ex, ez, q);
        fy[q] = update ? fy[q] - dt * v : v;
    }
    if (i < b.m0 && j < b.m1) {
        const double v = d1_zeta(b, ex, ey, q);
        fz[q] = update ? fz[q] - dt * v : v;
    }
}

// d2: face 2-form -> cell 3-form (:164-180); launch box = owned cells
__global__ void __launch_bounds__(256) k_d2(const DBox b, const double *__restrict__ fx, const double *__restrict__ fy,
                                            const double *__restrict__ fz, double *__restrict__ c) {
    int i, j, k;
    long long q;
    if (!point(b, i, j, k, q)) return;
    c[q] = (fx[q + b.sI] - fx[q]) + (fy[q + b.sJ] - fy[q]) + (fz[q + 1] - fz[q]);
}

// maximum of a workgroup's values into partial[blockIdx.x]
__device__ __forceinline__ void group_max(double v, double *__restrict__ partial) {
    __shared__ double lds[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
}

// |d2 d1 E| per owned cell (max_abs_div_curl, DecDiagnostics.cpp:110-135, with
// curl_xi_at / curl_eta_at / curl_zeta_at :37-69); launch box = owned cells
__global__ void __launch_bounds__(256) k_div_curl(const DBox b, const double *__restrict__ ex, const double *__restrict__ ey,
                                                  const double *__restrict__ ez, double *__restrict__ partial) {
    int i, j, k;
    long long q;
    double acc = 0.0;
    if (point(b, i, j, k, q)) {
        const double v = (d1_xi(b, ey, ez, q + b.sI) - d1_xi(b, ey, ez, q)) + (d1_eta(b, ex, ez, q + b.sJ) - d1_eta(b, ex, ez, q)) +
                         (d1_zeta(b, ex, ey, q + 1) - d1_zeta(b, ex, ey, q));
        acc = dmax(acc, fabs(v));
    }
    group_max(acc, partial);
}

// |d1 d0 phi| per owned face of each axis (max_abs_curl_grad, DecDiagnostics.cpp:10-35,71-93);
// launch box = owned nodes
__global__ void __launch_bounds__(256) k_curl_grad(const DBox b, const double *__restrict__ n, double *__restrict__ partial) {
    int i, j, k;
    long long q;
    double acc = 0.0;
    if (point(b, i, j, k, q)) {
        const long long sI = b.sI, sJ = b.sJ;
        if (j < b.m1 && k < b.m2) {
            const double e0 = n[q + sJ] - n[q], e1 = n[q + sJ + 1] - n[q + 1];
            const double z0 = n[q + 1] - n[q], z1 = n[q + sJ + 1] - n[q + sJ];
            acc = dmax(acc, fabs((e0 - e1) + (z1 - z0)));
        }
        if (i < b.m0 && k < b.m2) {
            const double x0 = n[q + sI] - n[q], x1 = n[q + sI + 1] - n[q + 1];
            const double z0 = n[q + 1] - n[q], z1 = n[q + sI + 1] - n[q + sI];
            acc = dmax(acc, fabs((x1 - x0) + (z0 - z1)));
        }
        if (i < b.m0 && j < b.m1) {
            const double x0 = n[q + sI] - n[q], x1 = n[q + sI + sJ] - n[q + sJ];
            const double e0 = n[q + sJ] - n[q], e1 = n[q + sI + sJ] - n[q + sI];
            acc = dmax(acc, fabs((x0 - x1) + (e1 - e0)));
        }
    }
    group_max(acc, partial);
}

// fold ngroups partial maxima into out[0]
__global__ void __launch_bounds__(256) k_fold_max(const double *__restrict__ partial, int ngroups, double *__restrict__ out) {
    double acc = 0.0;
    for (int t = threadIdx.x; t < ngroups; t += 256) acc = dmax(acc, partial[t]);
    __shared__ double lds[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc = dmax(acc, __shfl_xor(acc, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

DBox make_box(const int *dims, int g, int n0, int n1, int n2) {
    DBox b;
    b.m0 = dims[0];
    b.m1 = dims[1];
    b.m2 = dims[2];
    b.g = g;
    const long long N1 = dims[1] + 1 + 2LL * g, N2 = dims[2] + 1 + 2LL * g;
    b.sJ = N2;
    b.sI = N1 * N2;
    b.n0 = n0;
    b.n1 = n1;
    b.n2 = n2;
    return b;
}

long long points(const DBox &b) { return (long long)b.n0 * b.n1 * b.n2; }
unsigned groups_of(long long n) { return (unsigned)((n + 255) / 256); }

#define DEC_HIPCHK(x)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

}  // namespace

namespace merc {

void dec_check_box(const int *dims, int nghost) {
    if (!dims) throw std::runtime_error("DEC: dims is NULL");
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || nghost < 0)
        throw std::runtime_error("DEC: every axis needs at least one cell and nghost >= 0");
    const long long N0 = dims[0] + 1 + 2LL * nghost, N1 = dims[1] + 1 + 2LL * nghost, N2 = dims[2] + 1 + 2LL * nghost;
    if (N0 * N1 * N2 > (1LL << 40)) throw std::runtime_error("DEC: block box too large");
}

void dec_d0(const int *dims, int g, const double *node, double *const *edge, hipStream_t st) {
    const DBox b = make_box(dims, g, dims[0] + 1, dims[1] + 1, dims[2] + 1);
    hipLaunchKernelGGL(k_d0, dim3(groups_of(points(b))), dim3(256), 0, st, b, node, edge[0], edge[1], edge[2]);
    DEC_HIPCHK(hipGetLastError());
}

void dec_d1(const int *dims, int g, const double *const *edge, double *const *face, int update, double dt, hipStream_t st) {
    const DBox b = make_box(dims, g, dims[0] + 1, dims[1] + 1, dims[2] + 1);
    hipLaunchKernelGGL(k_d1, dim3(groups_of(points(b))), dim3(256), 0, st, b, edge[0], edge[1], edge[2], face[0], face[1], face[2],
                       update, dt);
    DEC_HIPCHK(hipGetLastError());
}

void dec_d2(const int *dims, int g, const double *const *face, double *cell, hipStream_t st) {
    const DBox b = make_box(dims, g, dims[0], dims[1], dims[2]);
    hipLaunchKernelGGL(k_d2, dim3(groups_of(points(b))), dim3(256), 0, st, b, face[0], face[1], face[2], cell);
    DEC_HIPCHK(hipGetLastError());
}

long long dec_div_curl_groups(const int *dims) { return groups_of((long long)dims[0] * dims[1] * dims[2]); }
long long dec_curl_grad_groups(const int *dims) { return groups_of((long long)(dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)); }

void dec_div_curl(const int *dims, int g, const double *const *edge, double *partial, double *out, hipStream_t st) {
    const DBox b = make_box(dims, g, dims[0], dims[1], dims[2]);
    const unsigned ng = groups_of(points(b));
    hipLaunchKernelGGL(k_div_curl, dim3(ng), dim3(256), 0, st, b, edge[0], edge[1], edge[2], partial);
    hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(256), 0, st, partial, (int)ng, out);
    DEC_HIPCHK(hipGetLastError());
}

void dec_curl_grad(const int *dims, int g, const double *node, double *partial, double *out, hipStream_t st) {
    const DBox b = make_box(dims, g, dims[0] + 1, dims[1] + 1, dims[2] + 1);
    const unsigned ng = groups_of(points(b));
    hipLaunchKernelGGL(k_curl_grad, dim3(ng), dim3(256), 0, st, b, node, partial);
    hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(256), 0, st, partial, (int)ng, out);
    DEC_HIPCHK(hipGetLastError());
}

// Device address of this GPU's copy of the scratch array, looked up once per
// device with that device current.
double *dec_noalloc_scratch() {
    int dev = -1;
    DEC_HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("DEC: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("DEC: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace merc
