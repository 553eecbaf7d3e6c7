// Caller arrays on the device: FP64 (n,3) rows in the caller's particle order
// to and from the contexts' structure-of-arrays state in `real`, through the
// slot -> caller index map of contexts that keep particles in cell order.
// The conversions are the host path's casts, so both paths give the same
// values.
#include "rrg_internal.hpp"

namespace rrg {
namespace layout {
namespace {

template <class real>
__global__ __launch_bounds__(256) void k_rows_in(int64_t n, const double* __restrict__ rows, real* __restrict__ soa,
                                                 const int* __restrict__ id) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t j = id ? int64_t(id[s]) : s;
#pragma unroll
    for (int c = 0; c < 3; ++c) soa[c * n + s] = (real)rows[3 * j + c];
}

template <class real>
__global__ __launch_bounds__(256) void k_rows_out(int64_t n, const real* __restrict__ soa, double* __restrict__ rows,
                                                  const int* __restrict__ id) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t j = id ? int64_t(id[s]) : s;
#pragma unroll
    for (int c = 0; c < 3; ++c) rows[3 * j + c] = (double)soa[c * n + s];
}

template <class real>
__global__ __launch_bounds__(256) void k_scalar_out(int64_t n, const real* __restrict__ v, double* __restrict__ out,
                                                    const int* __restrict__ id) {
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[id ? int64_t(id[s]) : s] = (double)v[s];
}

unsigned blocks(int64_t n) { return unsigned((n + 255) / 256); }

}  // namespace

template <class real> void rows_in(hipStream_t s, int64_t n, const double* rows, real* soa, const int* id) {
    if (n == 0) return;
    k_rows_in<real><<<blocks(n), 256, 0, s>>>(n, rows, soa, id);
    RRG_LAUNCH_CHECK();
}
template <class real> void rows_out(hipStream_t s, int64_t n, const real* soa, double* rows, const int* id) {
    if (n == 0) return;
    k_rows_out<real><<<blocks(n), 256, 0, s>>>(n, soa, rows, id);
    RRG_LAUNCH_CHECK();
}
template <class real> void scalar_out(hipStream_t s, int64_t n, const real* v, double* out, const int* id) {
    if (n == 0) return;
    k_scalar_out<real><<<blocks(n), 256, 0, s>>>(n, v, out, id);
    RRG_LAUNCH_CHECK();
}

#define RRG_LAYOUT_INST(real)                                                                      \
    template void rows_in<real>(hipStream_t, int64_t, const double*, real*, const int*);          \
    template void rows_out<real>(hipStream_t, int64_t, const real*, double*, const int*);         \
    template void scalar_out<real>(hipStream_t, int64_t, const real*, double*, const int*);

RRG_LAYOUT_INST(float)
RRG_LAYOUT_INST(double)

}  // namespace layout
}  // namespace rrg
