// Direct-sum gravity and velocity-Verlet kernels for gfx950.
//
// Reference: solver/lib/rrgsim_nbody/src/nbody.cu
//   acceleration_kernel_blocked_  nbody.cu:19-65   (pair sum, self pair included)
//   grav_kernel_blocked_          nbody.cu:67-114  (potential, self term included)
//   predict_step_ / correct_step_ nbody.cu:122-181 (predictor/corrector Verlet)
//   calc_conservation             nbody.cu:246-291 (E, P, L definitions)
//
// Design (MI355X): 256-thread workgroups (4 wave64s).  A workgroup evaluates
// one tile: up to 256*IB targets held in registers (IB per lane) against one
// contiguous source range, staged 256 sources at a time through LDS and read
// back as wave-uniform broadcasts.  Two launch forms share the tile code:
//   * grid form: blockIdx = (target tile, source slice) over one system;
//   * queue form: a persistent launch whose workgroups pull (target range,
//     source range, slice) items from a list with an atomic counter -- used
//     for ensembles of independent systems of uneven size.
// Source-range slices write separate partial sums that are added in slice
// order afterwards: deterministic, no floating-point atomics.
//
// The whole file is compiled with -ffp-contract=off; fused multiply-adds are
// written explicitly where they are wanted.
#include "rrg_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace rrg {
namespace nbody {
namespace {

constexpr int TPB = 256;

// ---------------------------------------------------------------- FP64 pair
// RSQ 0: 1/sqrt(r2) with IEEE division and square root (the reference's
// `1. / sqrt(..)`, nbody.cu:50-53); 1: the ROCm device library's FP64 rsqrt.
template <int RSQ> __device__ __forceinline__ double inv_sqrt_f64(double r2) {
    if constexpr (RSQ == 0) {
        return 1.0 / sqrt(r2);
    } else {
        return rsqrt(r2);
    }
}

// One tile: targets [i0, i1) against sources [j0, j1).  UNI: every particle
// has the same softening, so h_i + h_j is the constant `hh`.  POT: the
// potential -sum m_j / r (grav_kernel_blocked_) instead of the acceleration.
// Output o[c*n + i] (c = 0..2) or o[i] for the potential.
template <int IB, int RSQ, bool UNI, bool POT>
__device__ __forceinline__ void tile_f64(const PState<double>& p, double hh, int64_t i0, int64_t i1, int64_t j0,
                                         int64_t j1, double* __restrict__ o, double4* sp, double* sh) {
    const int t = threadIdx.x;
    const int64_t n = p.n;
    const double* __restrict__ X = p.x;
    double xi[IB], yi[IB], zi[IB], hi[IB], ax[IB], ay[IB], az[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(i0 + k * TPB + t, i1 - 1);
        xi[k] = X[i];
        yi[k] = X[n + i];
        zi[k] = X[2 * n + i];
        hi[k] = UNI ? hh : 0.5 * p.s2[i];
        ax[k] = ay[k] = az[k] = 0.0;
    }
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
        double h = 1.0;
        if (j < j1) {
            q = make_double4(X[j], X[n + j], X[2 * n + j], p.m[j]);
            if (!UNI) h = 0.5 * p.s2[j];
        }
        __syncthreads();
        sp[t] = q;
        if (!UNI) sh[t] = h;
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < TPB; ++jj) {
            const double4 s = sp[jj];
            const double hj = UNI ? 0.0 : sh[jj];
#pragma unroll
            for (int k = 0; k < IB; ++k) {
                const double dx = s.x - xi[k];
                const double dy = s.y - yi[k];
                const double dz = s.z - zi[k];
                const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, UNI ? hi[k] : hi[k] + hj)));
                const double inv = inv_sqrt_f64<RSQ>(r2);
                if (POT) {
                    ax[k] = fma(s.w, inv, ax[k]);
                } else {
                    const double c = s.w * (inv * inv * inv);
                    ax[k] = fma(dx, c, ax[k]);
                    ay[k] = fma(dy, c, ay[k]);
                    az[k] = fma(dz, c, az[k]);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = i0 + k * TPB + t;
        if (i < i1) {
            if (POT) {
                o[i] = -ax[k];
            } else {
                o[i] = ax[k];
                o[n + i] = ay[k];
                o[2 * n + i] = az[k];
            }
        }
    }
}

// ------------------------------------------------------------ FP32 kinds
// HILO: positions split x = hi + lo (hi = fp32(x), lo = fp32(x - hi)) and
// d = (hi_j - hi_i) + (lo_j - lo_i) carries ~48 bits of the FP64 separation;
// otherwise d = hi_j - hi_i.  Pair math is FP32 (v_rsq_f32); each 256-source
// tile is summed in FP32 and the tile sums enter a per-target two-sum (Knuth)
// accumulator, converted to FP64 at the end.
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float t) {
    const float u = s + t;
    const float bp = u - s;
    const float e = (s - (u - bp)) + (t - bp);
    s = u;
    c += e;
}

template <int IB, bool HILO, bool UNI, bool POT>
__device__ __forceinline__ void tile_f32(const float4* __restrict__ A, const float4* __restrict__ B, int64_t n,
                                         float hh, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
                                         double* __restrict__ o, float4* sa, float4* sb) {
    constexpr bool NEED_B = HILO || !UNI;
    const int t = threadIdx.x;
    float xh[IB], yh[IB], zh[IB], xl[IB], yl[IB], zl[IB], hs[IB];
    float sx[IB], sy[IB], sz[IB], cx[IB], cy[IB], cz[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(i0 + k * TPB + t, i1 - 1);
        const float4 a = A[i], b = B[i];
        xh[k] = a.x; yh[k] = a.y; zh[k] = a.z;
        xl[k] = b.x; yl[k] = b.y; zl[k] = b.z;
        hs[k] = UNI ? hh : b.w;
        sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
    }
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 qb = make_float4(0.f, 0.f, 0.f, 1.f);
        if (j < j1) {
            qa = A[j];
            if (NEED_B) qb = B[j];
        }
        __syncthreads();
        sa[t] = qa;
        if (NEED_B) sb[t] = qb;
        __syncthreads();
        float tx[IB], ty[IB], tz[IB];
#pragma unroll
        for (int k = 0; k < IB; ++k) tx[k] = ty[k] = tz[k] = 0.f;
#pragma unroll 4
        for (int jj = 0; jj < TPB; ++jj) {
            const float4 a = sa[jj];
            const float4 b = NEED_B ? sb[jj] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < IB; ++k) {
                float dx, dy, dz;
                if (HILO) {
                    dx = (a.x - xh[k]) + (b.x - xl[k]);
                    dy = (a.y - yh[k]) + (b.y - yl[k]);
                    dz = (a.z - zh[k]) + (b.z - zl[k]);
                } else {
                    dx = a.x - xh[k];
                    dy = a.y - yh[k];
                    dz = a.z - zh[k];
                }
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, UNI ? hs[k] : hs[k] + b.w)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                if (POT) {
                    tx[k] = fmaf(a.w, inv, tx[k]);
                } else {
                    const float c = (a.w * inv) * (inv * inv);
                    tx[k] = fmaf(dx, c, tx[k]);
                    ty[k] = fmaf(dy, c, ty[k]);
                    tz[k] = fmaf(dz, c, tz[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < IB; ++k) {
            two_sum_acc(sx[k], cx[k], tx[k]);
            if (!POT) {
                two_sum_acc(sy[k], cy[k], ty[k]);
                two_sum_acc(sz[k], cz[k], tz[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = i0 + k * TPB + t;
        if (i < i1) {
            if (POT) {
                o[i] = -(double(sx[k]) + double(cx[k]));
            } else {
                o[i] = double(sx[k]) + double(cx[k]);
                o[n + i] = double(sy[k]) + double(cy[k]);
                o[2 * n + i] = double(sz[k]) + double(cz[k]);
            }
        }
    }
}

// --------------------------------------------------------------- grid form
template <int IB, int RSQ, bool UNI, bool POT>
__global__ __launch_bounds__(TPB) void k_direct_f64(PState<double> p, int64_t slice_len, double hh,
                                                    double* __restrict__ out) {
    __shared__ double4 sp[TPB];
    __shared__ double sh[UNI ? 1 : TPB];
    const int64_t n = p.n;
    const int64_t i0 = int64_t(blockIdx.x) * TPB * IB;
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    tile_f64<IB, RSQ, UNI, POT>(p, hh, i0, min(n, i0 + TPB * IB), j0, min(n, j0 + slice_len),
                                out + int64_t(blockIdx.y) * (POT ? 1 : 3) * n, sp, sh);
}

template <int IB, bool HILO, bool UNI, bool POT>
__global__ __launch_bounds__(TPB) void k_direct_f32(const float4* __restrict__ A, const float4* __restrict__ B,
                                                    int64_t n, int64_t slice_len, float hh, double* __restrict__ out) {
    __shared__ float4 sa[TPB];
    __shared__ float4 sb[(HILO || !UNI) ? TPB : 1];
    const int64_t i0 = int64_t(blockIdx.x) * TPB * IB;
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    tile_f32<IB, HILO, UNI, POT>(A, B, n, hh, i0, min(n, i0 + TPB * IB), j0, min(n, j0 + slice_len),
                                 out + int64_t(blockIdx.y) * (POT ? 1 : 3) * n, sa, sb);
}

// -------------------------------------------------------------- queue form
// Persistent workgroups pull items until the list is exhausted (every
// workgroup exits when the counter passes nitems).  The last workgroup to
// leave resets both counters, so the next launch finds them at zero.
__device__ __forceinline__ int next_item(int* counters, int nitems, int* s_it) {
    if (threadIdx.x == 0) *s_it = atomicAdd(counters, 1);
    __syncthreads();
    const int it = *s_it;
    __syncthreads();
    if (it >= nitems && threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(counters + 1, 1) == int(gridDim.x) - 1) {
            atomicExch(counters, 0);
            atomicExch(counters + 1, 0);
        }
    }
    return it;
}

template <int IB, int RSQ, bool UNI, bool POT>
__global__ __launch_bounds__(TPB) void k_queue_f64(PState<double> p, const WorkItem* __restrict__ items, int nitems,
                                                   int* counters, double hh, double* __restrict__ out) {
    __shared__ double4 sp[TPB];
    __shared__ double sh[UNI ? 1 : TPB];
    __shared__ int s_it;
    const int64_t n = p.n;
    for (;;) {
        const int it = next_item(counters, nitems, &s_it);
        if (it >= nitems) break;
        const WorkItem w = items[it];
        tile_f64<IB, RSQ, UNI, POT>(p, hh, w.i0, w.i1, w.j0, w.j1, out + int64_t(w.slice) * (POT ? 1 : 3) * n, sp, sh);
    }
}

template <int IB, bool HILO, bool UNI, bool POT>
__global__ __launch_bounds__(TPB) void k_queue_f32(const float4* __restrict__ A, const float4* __restrict__ B,
                                                   int64_t n, const WorkItem* __restrict__ items, int nitems,
                                                   int* counters, float hh, double* __restrict__ out) {
    __shared__ float4 sa[TPB];
    __shared__ float4 sb[(HILO || !UNI) ? TPB : 1];
    __shared__ int s_it;
    for (;;) {
        const int it = next_item(counters, nitems, &s_it);
        if (it >= nitems) break;
        const WorkItem w = items[it];
        tile_f32<IB, HILO, UNI, POT>(A, B, n, hh, w.i0, w.i1, w.j0, w.j1, out + int64_t(w.slice) * (POT ? 1 : 3) * n,
                                     sa, sb);
    }
}

// --------------------------------------------------------- particle kernels
__global__ void k_pack_mixed(PState<double> p, float4* __restrict__ A, float4* __restrict__ B) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    const double x = p.x[i], y = p.x[n + i], z = p.x[2 * n + i];
    const float xh = float(x), yh = float(y), zh = float(z);
    A[i] = make_float4(xh, yh, zh, float(p.m[i]));
    B[i] = make_float4(float(x - double(xh)), float(y - double(yh)), float(z - double(zh)), float(0.5 * p.s2[i]));
}

// predict_step_ (nbody.cu:122-152).  v* = v + dt a in real; the position
// update `r + dt * 0.5 * (v + v*)` is evaluated in double (the 0.5 literal is
// double) and rounded to real on the store.
template <class real, bool PACK>
__global__ void k_predict(PState<real> p, real dt, float4* __restrict__ A, float4* __restrict__ B) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    // every load is issued before the first store: the particle arrays are
    // not declared disjoint, so a load after a store would wait for it
    real v[3], r[3], a[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = p.v[c * n + i];
        r[c] = p.x[c * n + i];
        a[c] = p.a[c * n + i];
    }
    const double m = PACK ? (double)p.m[i] : 0.0, s2 = PACK ? (double)p.s2[i] : 0.0;
    double xn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const real vp = v[c] + dt * a[c];
        p.w[c * n + i] = vp;
        const real rn = (real)((double)r[c] + ((double)dt * 0.5) * (double)(v[c] + vp));
        p.x[c * n + i] = rn;
        xn[c] = (double)rn;
    }
    if constexpr (PACK) {
        const float xh = float(xn[0]), yh = float(xn[1]), zh = float(xn[2]);
        A[i] = make_float4(xh, yh, zh, float(m));
        B[i] = make_float4(float(xn[0] - double(xh)), float(xn[1] - double(yh)), float(xn[2] - double(zh)),
                           float(0.5 * s2));
    }
}

// Slice count of particle i: uniform, or per particle for ensembles.
__device__ __forceinline__ int slices_of(int64_t i, int slices, const unsigned char* __restrict__ ns) {
    return ns ? int(ns[i]) : slices;
}

// correct_step_ (nbody.cu:159-181): v = 0.5 * (v + v* + dt a), after the
// slice partials (if any) are summed into a.
template <class real>
__global__ void k_correct(PState<real> p, real dt, int slices, const unsigned char* __restrict__ ns,
                          const double* __restrict__ part) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    const int si = slices_of(i, slices, ns);
    // loads first (see k_predict)
    real a[3], v[3], vv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (part) {
            double s = part[c * n + i];
            for (int q = 1; q < si; ++q) s += part[(int64_t(q) * 3 + c) * n + i];
            a[c] = (real)s;
        } else {
            a[c] = p.a[c * n + i];
        }
        v[c] = p.v[c * n + i];
        vv[c] = p.w[c * n + i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (part) p.a[c * n + i] = a[c];
        p.v[c * n + i] = (real)(0.5 * (double)(v[c] + vv[c] + dt * a[c]));
    }
}

// k_correct of one step followed by k_predict of the next in one pass over
// the particles (FP64 state): the same values as the two launches, since
// each particle's corrector and predictor touch only that particle.
template <bool PACK>
__global__ void k_correct_predict(PState<double> p, double dt, int slices, const unsigned char* __restrict__ ns,
                                  const double* __restrict__ part, float4* __restrict__ A, float4* __restrict__ B) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    const int si = slices_of(i, slices, ns);
    // loads first (see k_predict)
    double a[3], v[3], vv[3], r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (part) {
            double s = part[c * n + i];
            for (int q = 1; q < si; ++q) s += part[(int64_t(q) * 3 + c) * n + i];
            a[c] = s;
        } else {
            a[c] = p.a[c * n + i];
        }
        v[c] = p.v[c * n + i];
        vv[c] = p.w[c * n + i];
        r[c] = p.x[c * n + i];
    }
    const double m = PACK ? p.m[i] : 0.0, s2 = PACK ? p.s2[i] : 0.0;
    double xn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (part) p.a[c * n + i] = a[c];
        const double vc = 0.5 * (v[c] + vv[c] + dt * a[c]);  // correct_step_
        p.v[c * n + i] = vc;
        const double vp = vc + dt * a[c];  // predict_step_ of the next step
        p.w[c * n + i] = vp;
        const double rn = r[c] + (dt * 0.5) * (vc + vp);
        p.x[c * n + i] = rn;
        xn[c] = rn;
    }
    if constexpr (PACK) {
        const float xh = float(xn[0]), yh = float(xn[1]), zh = float(xn[2]);
        A[i] = make_float4(xh, yh, zh, float(m));
        B[i] = make_float4(float(xn[0] - double(xh)), float(xn[1] - double(yh)), float(xn[2] - double(zh)),
                           float(0.5 * s2));
    }
}

__global__ void k_reduce_slices(int64_t n, int comps, int slices, const unsigned char* __restrict__ ns,
                                const double* __restrict__ part, double* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int si = slices_of(i, slices, ns);
    for (int c = 0; c < comps; ++c) {
        double s = part[c * n + i];
        for (int q = 1; q < si; ++q) s += part[(int64_t(q) * comps + c) * n + i];
        dst[c * n + i] = s;
    }
}

// Wave64 + LDS block sum of 8 FP64 quantities over particles [b, e) with a
// grid stride of `stride` blocks starting at block `blk`.
template <class real>
__device__ __forceinline__ void conservation_block(const PState<real>& p, int64_t b, int64_t e, int64_t blk,
                                                   int64_t stride, double* __restrict__ dst) {
    __shared__ double red[4][8];
    const int64_t n = p.n;
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = b + blk * 256 + threadIdx.x; i < e; i += stride * 256) {
        const double x = p.x[i], y = p.x[n + i], z = p.x[2 * n + i];
        const double vx = p.v[i], vy = p.v[n + i], vz = p.v[2 * n + i];
        const double m = p.m[i];
        q[0] += m * (vx * vx + vy * vy + vz * vz);
        q[1] += m * (double)p.g[i];
        q[2] += m * vx;
        q[3] += m * vy;
        q[4] += m * vz;
        q[5] += m * (vz * y - vy * z);
        q[6] += m * (vx * z - vz * x);
        q[7] += m * (vy * x - vx * y);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        for (int off = 32; off > 0; off >>= 1) q[k] += __shfl_down(q[k], off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) red[wid][k] = q[k];
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        dst[k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_conservation(PState<real> p, double* __restrict__ partial) {
    conservation_block<real>(p, 0, p.n, blockIdx.x, gridDim.x, partial + int64_t(blockIdx.x) * 8);
}

// One block per system of an ensemble: offsets[g] .. offsets[g+1].
__global__ __launch_bounds__(256) void k_conservation_seg(PState<double> p, const int64_t* __restrict__ offsets,
                                                          double* __restrict__ out) {
    const int g = blockIdx.x;
    conservation_block<double>(p, offsets[g], offsets[g + 1], 0, 1, out + int64_t(g) * 8);
}

__global__ void k_fill_i32(int* p, int count, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

inline unsigned grid1d(int64_t n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

// -------------------------------------------------------------- dispatch
// Template arguments are selected from the run-time (kind, ib, uniform,
// potential) tuple through a table of launchers.
struct Sel {
    int rsq;    // FP64: 0 exact 1/sqrt, 1 rsqrt
    bool hilo;  // FP32: hi/lo positions
    bool f32;
};

Sel select(int kind) {
    Sel s;
    s.f32 = is_fp32_kind(kind);
    s.rsq = kind == RRG_FORCE_F64_EXACT ? 0 : 1;
    s.hilo = kind == RRG_FORCE_MIXED;
    return s;
}

template <int IB, int RSQ, bool UNI, bool POT>
void launch_f64(hipStream_t s, const dim3& grid, const PState<double>& p, int64_t slice_len, double hh, double* out,
                const WorkItem* items, int nitems, int* counters) {
    if (items) k_queue_f64<IB, RSQ, UNI, POT><<<grid, TPB, 0, s>>>(p, items, nitems, counters, hh, out);
    else k_direct_f64<IB, RSQ, UNI, POT><<<grid, TPB, 0, s>>>(p, slice_len, hh, out);
}

template <int IB, bool HILO, bool UNI, bool POT>
void launch_f32(hipStream_t s, const dim3& grid, int64_t n, const float4* A, const float4* B, int64_t slice_len,
                float hh, double* out, const WorkItem* items, int nitems, int* counters) {
    if (items) k_queue_f32<IB, HILO, UNI, POT><<<grid, TPB, 0, s>>>(A, B, n, items, nitems, counters, hh, out);
    else k_direct_f32<IB, HILO, UNI, POT><<<grid, TPB, 0, s>>>(A, B, n, slice_len, hh, out);
}

void dispatch(hipStream_t s, const dim3& grid, int ib, const Force& f, bool pot, const PState<double>& p,
              const float4* A, const float4* B, int64_t slice_len, double* out, const WorkItem* items, int nitems,
              int* counters) {
    const Sel sel = select(f.kind);
    const bool uni = f.uniform && !pot;
    bool done = false;
#define RRG_F64(IB, RSQ, UNI, POT)                                                                       \
    if (!done && !sel.f32 && ib == IB && sel.rsq == RSQ && uni == UNI && pot == POT) {                   \
        launch_f64<IB, RSQ, UNI, POT>(s, grid, p, slice_len, f.hh, out, items, nitems, counters);        \
        done = true;                                                                                     \
    }
#define RRG_F64_ALL(IB)                                                                                 \
    RRG_F64(IB, 0, false, false) RRG_F64(IB, 1, false, false) RRG_F64(IB, 0, true, false)               \
    RRG_F64(IB, 1, true, false) RRG_F64(IB, 0, false, true) RRG_F64(IB, 1, false, true)
    RRG_F64_ALL(1) RRG_F64_ALL(2)
#undef RRG_F64_ALL
#undef RRG_F64
#define RRG_F32(IB, HILO, UNI, POT)                                                                      \
    if (!done && sel.f32 && ib == IB && sel.hilo == HILO && uni == UNI && pot == POT) {                  \
        launch_f32<IB, HILO, UNI, POT>(s, grid, p.n, A, B, slice_len, float(f.hh), out, items, nitems,   \
                                       counters);                                                        \
        done = true;                                                                                     \
    }
#define RRG_F32_ALL(IB)                                                                                 \
    RRG_F32(IB, true, false, false) RRG_F32(IB, false, false, false) RRG_F32(IB, true, true, false)     \
    RRG_F32(IB, false, true, false) RRG_F32(IB, true, false, true) RRG_F32(IB, false, false, true)
    RRG_F32_ALL(1) RRG_F32_ALL(2) RRG_F32_ALL(4)
#undef RRG_F32_ALL
#undef RRG_F32
    if (!done) fail(RRG_ERR_INVALID, "direct sum: unsupported launch shape");
    RRG_LAUNCH_CHECK();
}

}  // namespace

bool is_fp32_kind(int kind) { return kind == RRG_FORCE_MIXED || kind == RRG_FORCE_FP32; }

bool valid_ib(int kind, int ib) { return is_fp32_kind(kind) ? (ib == 1 || ib == 2 || ib == 4) : (ib == 1 || ib == 2); }

Plan plan_direct(int64_t n, int force_kind, int target_workgroups, int ib) {
    Plan pl;
    const bool f32 = is_fp32_kind(force_kind);
    if (ib > 0 && valid_ib(force_kind, ib)) {
        pl.ib = ib;
    } else {
        pl.ib = f32 ? 2 : 1;
        if (n >= 65536) pl.ib = f32 ? 4 : 2;
    }
    const int64_t per_tile = int64_t(TPB) * pl.ib;
    pl.tiles = (n + per_tile - 1) / per_tile;
    const int64_t target = target_workgroups > 0 ? target_workgroups : 4096;
    int64_t slices = (target + pl.tiles - 1) / pl.tiles;
    const int64_t max_slices = std::max<int64_t>(1, n / TPB);
    slices = std::min(std::max<int64_t>(slices, 1), std::min<int64_t>(max_slices, 256));
    int64_t len = (n + slices - 1) / slices;
    len = (len + TPB - 1) / TPB * TPB;
    pl.slice_len = std::max<int64_t>(len, TPB);
    pl.slices = int((n + pl.slice_len - 1) / pl.slice_len);
    if (pl.slices < 1) pl.slices = 1;
    return pl;
}

int targets_per_tile(int ib) { return TPB * ib; }

void pack_mixed(hipStream_t s, const PState<double>& p, float4* A, float4* B) {
    k_pack_mixed<<<grid1d(p.n, 256), 256, 0, s>>>(p, A, B);
    RRG_LAUNCH_CHECK();
}

void accel(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
           double* out) {
    dispatch(s, dim3(unsigned(pl.tiles), unsigned(pl.slices)), pl.ib, f, false, p, A, B, pl.slice_len, out, nullptr,
             0, nullptr);
}

void grav(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
          double* out) {
    dispatch(s, dim3(unsigned(pl.tiles), unsigned(pl.slices)), pl.ib, f, true, p, A, B, pl.slice_len, out, nullptr,
             0, nullptr);
}

void queue(hipStream_t s, const QueuePlan& q, const Force& f, bool potential, const PState<double>& p,
           const float4* A, const float4* B, const WorkItem* items, int* counters, double* out) {
    dispatch(s, dim3(unsigned(q.workgroups)), q.ib, f, potential, p, A, B, 0, out, items, q.nitems, counters);
}

void reset_counters(hipStream_t s, int* counters) {
    k_fill_i32<<<1, 64, 0, s>>>(counters, 2, 0);
    RRG_LAUNCH_CHECK();
}

void reduce_slices(hipStream_t s, int64_t n, int comps, int slices, const unsigned char* ns, const double* part,
                   double* dst) {
    k_reduce_slices<<<grid1d(n, 256), 256, 0, s>>>(n, comps, slices, ns, part, dst);
    RRG_LAUNCH_CHECK();
}

template <class real> void predict(hipStream_t s, const PState<real>& p, real dt, float4* A, float4* B) {
    if (A && B) {
        if constexpr (std::is_same_v<real, double>) {
            k_predict<real, true><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, A, B);
        } else {
            fail(RRG_ERR_INVALID, "FP32 direct-sum kinds need FP64 particle state");
        }
    } else {
        k_predict<real, false><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, nullptr, nullptr);
    }
    RRG_LAUNCH_CHECK();
}

void correct_reduce(hipStream_t s, const PState<double>& p, double dt, int slices, const unsigned char* ns,
                    const double* part) {
    k_correct<double><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, slices, ns, part);
    RRG_LAUNCH_CHECK();
}

void correct_predict(hipStream_t s, const PState<double>& p, double dt, int slices, const unsigned char* ns,
                     const double* part, float4* A, float4* B) {
    if (A && B) k_correct_predict<true><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, slices, ns, part, A, B);
    else k_correct_predict<false><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, slices, ns, part, nullptr, nullptr);
    RRG_LAUNCH_CHECK();
}

template <class real> void correct(hipStream_t s, const PState<real>& p, real dt) {
    k_correct<real><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, 1, nullptr, nullptr);
    RRG_LAUNCH_CHECK();
}

int64_t conservation_blocks(int64_t n) { return std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256)); }

template <class real> int64_t conservation(hipStream_t s, const PState<real>& p, double* partial) {
    const int64_t nb = conservation_blocks(p.n);
    k_conservation<real><<<unsigned(nb), 256, 0, s>>>(p, partial);
    RRG_LAUNCH_CHECK();
    return nb;
}

void conservation_segments(hipStream_t s, const PState<double>& p, int k, const int64_t* offsets, double* out) {
    k_conservation_seg<<<unsigned(k), 256, 0, s>>>(p, offsets, out);
    RRG_LAUNCH_CHECK();
}

template void predict<double>(hipStream_t, const PState<double>&, double, float4*, float4*);
template void predict<float>(hipStream_t, const PState<float>&, float, float4*, float4*);
template void correct<double>(hipStream_t, const PState<double>&, double);
template void correct<float>(hipStream_t, const PState<float>&, float);
template int64_t conservation<double>(hipStream_t, const PState<double>&, double*);
template int64_t conservation<float>(hipStream_t, const PState<float>&, double*);

}  // namespace nbody
}  // namespace rrg
