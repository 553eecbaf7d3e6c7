// Direct-sum gravity and velocity-Verlet kernels for gfx950.
//
// Reference: solver/lib/rrgsim_nbody/src/nbody.cu
//   acceleration_kernel_blocked_  nbody.cu:19-65   (pair sum, self pair included)
//   grav_kernel_blocked_          nbody.cu:67-114  (potential, self term included)
//   predict_step_ / correct_step_ nbody.cu:122-181 (predictor/corrector Verlet)
//   calc_conservation             nbody.cu:246-291 (E, P, L definitions)
//
// Design (MI355X): 256-thread workgroups (4 wave64s).  Each workgroup owns a
// tile of 256*IB targets, held in registers (IB targets per lane), and one
// slice of the source range.  Sources are staged 256 at a time through LDS
// and read back as wave-uniform broadcasts.  Small N is spread over the chip
// by splitting the source range into slices whose partial sums are added in
// a fixed order afterwards (deterministic results, no atomics).
//
// The whole file is compiled with -ffp-contract=off; fused multiply-adds are
// written explicitly where they are wanted.
#include "rrg_internal.hpp"

#include <algorithm>

namespace rrg {
namespace nbody {
namespace {

constexpr int TPB = 256;

// ---------------------------------------------------------------- FP64 pair
// rsqrt_kind 0: 1/sqrt(r2) with IEEE division and square root (the
// reference's `1. / sqrt(..)`); 1: the ROCm device library's FP64 rsqrt.
template <int RSQ> __device__ __forceinline__ double inv_sqrt_f64(double r2) {
    if constexpr (RSQ == 0) {
        return 1.0 / sqrt(r2);
    } else {
        return rsqrt(r2);
    }
}

// UNI: every particle has the same softening, so h_i + h_j is the constant
// `hh` and the per-source softening is neither staged nor added.
template <int IB, int RSQ, bool UNI>
__global__ __launch_bounds__(TPB) void k_accel_f64(PState<double> p, int64_t slice_len, double hh,
                                                   double* __restrict__ out) {
    __shared__ double4 sp[TPB];  // (x, y, z, m)
    __shared__ double sh[UNI ? 1 : TPB];  // 0.5 * soft2
    const int t = threadIdx.x;
    const int64_t n = p.n;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    const double* __restrict__ X = p.x;
    double xi[IB], yi[IB], zi[IB], hi[IB], ax[IB], ay[IB], az[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        xi[k] = X[i];
        yi[k] = X[n + i];
        zi[k] = X[2 * n + i];
        hi[k] = UNI ? hh : 0.5 * p.s2[i];
        ax[k] = ay[k] = az[k] = 0.0;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
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
                const double c = s.w * (inv * inv * inv);
                ax[k] = fma(dx, c, ax[k]);
                ay[k] = fma(dy, c, ay[k]);
                az[k] = fma(dz, c, az[k]);
            }
        }
    }
    double* o = out + int64_t(blockIdx.y) * 3 * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) {
            o[i] = ax[k];
            o[n + i] = ay[k];
            o[2 * n + i] = az[k];
        }
    }
}

template <int IB, int RSQ>
__global__ __launch_bounds__(TPB) void k_grav_f64(PState<double> p, int64_t slice_len, double* __restrict__ out) {
    __shared__ double4 sp[TPB];
    __shared__ double sh[TPB];
    const int t = threadIdx.x;
    const int64_t n = p.n;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    const double* __restrict__ X = p.x;
    double xi[IB], yi[IB], zi[IB], hi[IB], g[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        xi[k] = X[i];
        yi[k] = X[n + i];
        zi[k] = X[2 * n + i];
        hi[k] = 0.5 * p.s2[i];
        g[k] = 0.0;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
        double h = 1.0;
        if (j < j1) {
            q = make_double4(X[j], X[n + j], X[2 * n + j], p.m[j]);
            h = 0.5 * p.s2[j];
        }
        __syncthreads();
        sp[t] = q;
        sh[t] = h;
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < TPB; ++jj) {
            const double4 s = sp[jj];
            const double hj = sh[jj];
#pragma unroll
            for (int k = 0; k < IB; ++k) {
                const double dx = s.x - xi[k];
                const double dy = s.y - yi[k];
                const double dz = s.z - zi[k];
                const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, hi[k] + hj)));
                g[k] = fma(s.w, inv_sqrt_f64<RSQ>(r2), g[k]);
            }
        }
    }
    double* o = out + int64_t(blockIdx.y) * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) o[i] = -g[k];
    }
}

// ------------------------------------------------------------ mixed precision
// Positions are split x = hi + lo with hi = fp32(x), lo = fp32(x - hi), so
// d = (hi_j - hi_i) + (lo_j - lo_i) carries ~48 bits of the FP64 separation.
// Pair math is FP32 (v_rsq_f32); each 256-source tile is summed in FP32 and
// the tile sums enter a per-target two-sum (Knuth) accumulator.
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float t) {
    const float u = s + t;
    const float bp = u - s;
    const float e = (s - (u - bp)) + (t - bp);
    s = u;
    c += e;
}

// HILO = false is the plain FP32 variant: d = x_j - x_i on the FP32 values
// (the lo parts are ignored), cheaper and less exact for close pairs; the
// accumulation is compensated either way.
template <int IB, bool HILO, bool UNI>
__global__ __launch_bounds__(TPB) void k_accel_mixed(const float4* __restrict__ A, const float4* __restrict__ B,
                                                     int64_t n, int64_t slice_len, float hh, double* __restrict__ out) {
    constexpr bool NEED_B = HILO || !UNI;
    __shared__ float4 sa[TPB];
    __shared__ float4 sb[NEED_B ? TPB : 1];
    const int t = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    float xh[IB], yh[IB], zh[IB], xl[IB], yl[IB], zl[IB], hs[IB];
    float sx[IB], sy[IB], sz[IB], cx[IB], cy[IB], cz[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        const float4 a = A[i], b = B[i];
        xh[k] = a.x; yh[k] = a.y; zh[k] = a.z;
        xl[k] = b.x; yl[k] = b.y; zl[k] = b.z;
        hs[k] = UNI ? hh : b.w;
        sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
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
                const float c = (a.w * inv) * (inv * inv);
                tx[k] = fmaf(dx, c, tx[k]);
                ty[k] = fmaf(dy, c, ty[k]);
                tz[k] = fmaf(dz, c, tz[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IB; ++k) {
            two_sum_acc(sx[k], cx[k], tx[k]);
            two_sum_acc(sy[k], cy[k], ty[k]);
            two_sum_acc(sz[k], cz[k], tz[k]);
        }
    }
    double* o = out + int64_t(blockIdx.y) * 3 * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) {
            o[i] = double(sx[k]) + double(cx[k]);
            o[n + i] = double(sy[k]) + double(cy[k]);
            o[2 * n + i] = double(sz[k]) + double(cz[k]);
        }
    }
}

template <int IB>
__global__ __launch_bounds__(TPB) void k_grav_mixed(const float4* __restrict__ A, const float4* __restrict__ B,
                                                    int64_t n, int64_t slice_len, double* __restrict__ out) {
    __shared__ float4 sa[TPB];
    __shared__ float4 sb[TPB];
    const int t = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    float xh[IB], yh[IB], zh[IB], xl[IB], yl[IB], zl[IB], hs[IB], sg[IB], cg[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        const float4 a = A[i], b = B[i];
        xh[k] = a.x; yh[k] = a.y; zh[k] = a.z;
        xl[k] = b.x; yl[k] = b.y; zl[k] = b.z; hs[k] = b.w;
        sg[k] = cg[k] = 0.f;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 qb = make_float4(0.f, 0.f, 0.f, 1.f);
        if (j < j1) {
            qa = A[j];
            qb = B[j];
        }
        __syncthreads();
        sa[t] = qa;
        sb[t] = qb;
        __syncthreads();
        float tg[IB];
#pragma unroll
        for (int k = 0; k < IB; ++k) tg[k] = 0.f;
#pragma unroll 4
        for (int jj = 0; jj < TPB; ++jj) {
            const float4 a = sa[jj];
            const float4 b = sb[jj];
#pragma unroll
            for (int k = 0; k < IB; ++k) {
                const float dx = (a.x - xh[k]) + (b.x - xl[k]);
                const float dy = (a.y - yh[k]) + (b.y - yl[k]);
                const float dz = (a.z - zh[k]) + (b.z - zl[k]);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, hs[k] + b.w)));
                tg[k] = fmaf(a.w, __builtin_amdgcn_rsqf(r2), tg[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IB; ++k) two_sum_acc(sg[k], cg[k], tg[k]);
    }
    double* o = out + int64_t(blockIdx.y) * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) o[i] = -(double(sg[k]) + double(cg[k]));
    }
}

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
    double xn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const real v = p.v[c * n + i];
        const real r = p.x[c * n + i];
        const real a = p.a[c * n + i];
        const real vp = v + dt * a;
        p.w[c * n + i] = vp;
        const real rn = (real)((double)r + ((double)dt * 0.5) * (double)(v + vp));
        p.x[c * n + i] = rn;
        xn[c] = (double)rn;
    }
    if constexpr (PACK) {
        const float xh = float(xn[0]), yh = float(xn[1]), zh = float(xn[2]);
        A[i] = make_float4(xh, yh, zh, float(p.m[i]));
        B[i] = make_float4(float(xn[0] - double(xh)), float(xn[1] - double(yh)), float(xn[2] - double(zh)),
                           float(0.5 * (double)p.s2[i]));
    }
}

// correct_step_ (nbody.cu:159-181): v = 0.5 * (v + v* + dt a), after the
// slice partials (if any) are summed into a.
template <class real>
__global__ void k_correct(PState<real> p, real dt, int slices, const double* __restrict__ part) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        real a;
        if (part) {
            double s = part[c * n + i];
            for (int q = 1; q < slices; ++q) s += part[(int64_t(q) * 3 + c) * n + i];
            a = (real)s;
            p.a[c * n + i] = a;
        } else {
            a = p.a[c * n + i];
        }
        const real v = p.v[c * n + i];
        const real vv = p.w[c * n + i];
        p.v[c * n + i] = (real)(0.5 * (double)(v + vv + dt * a));
    }
}

__global__ void k_reduce_slices(int64_t n, int comps, int slices, const double* __restrict__ part,
                                double* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < comps; ++c) {
        double s = part[c * n + i];
        for (int q = 1; q < slices; ++q) s += part[(int64_t(q) * comps + c) * n + i];
        dst[c * n + i] = s;
    }
}

// Wave64 + LDS block sum of 8 FP64 quantities.
template <class real>
__global__ __launch_bounds__(256) void k_conservation(PState<real> p, double* __restrict__ partial) {
    __shared__ double red[4][8];
    const int64_t n = p.n;
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
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
        partial[int64_t(blockIdx.x) * 8 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

inline unsigned grid1d(int64_t n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

}  // namespace

bool is_fp32_kind(int kind) { return kind == RRG_FORCE_MIXED || kind == RRG_FORCE_FP32; }

Plan plan_direct(int64_t n, int force_kind, int target_workgroups) {
    Plan pl;
    const bool f32 = is_fp32_kind(force_kind);
    pl.ib = f32 ? 2 : 1;
    if (n >= 65536) pl.ib = f32 ? 4 : 2;
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

void pack_mixed(hipStream_t s, const PState<double>& p, float4* A, float4* B) {
    k_pack_mixed<<<grid1d(p.n, 256), 256, 0, s>>>(p, A, B);
    RRG_LAUNCH_CHECK();
}

void accel(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
           double* out) {
    const dim3 grid(unsigned(pl.tiles), unsigned(pl.slices));
    const int64_t n = p.n;
    const int rsq = f.kind == RRG_FORCE_F64_EXACT ? 0 : 1;
    bool done = false;
#define RRG_F64(IB, RSQ, UNI)                                                                          \
    if (!done && !is_fp32_kind(f.kind) && pl.ib == IB && rsq == RSQ && f.uniform == UNI) {             \
        k_accel_f64<IB, RSQ, UNI><<<grid, TPB, 0, s>>>(p, pl.slice_len, f.hh, out);                   \
        done = true;                                                                                   \
    }
    RRG_F64(1, 0, false) RRG_F64(2, 0, false) RRG_F64(1, 1, false) RRG_F64(2, 1, false)
    RRG_F64(1, 0, true) RRG_F64(2, 0, true) RRG_F64(1, 1, true) RRG_F64(2, 1, true)
#undef RRG_F64
    const bool hilo = f.kind == RRG_FORCE_MIXED;
#define RRG_F32(IB, HILO, UNI)                                                                         \
    if (!done && is_fp32_kind(f.kind) && pl.ib == IB && hilo == HILO && f.uniform == UNI) {            \
        k_accel_mixed<IB, HILO, UNI><<<grid, TPB, 0, s>>>(A, B, n, pl.slice_len, float(f.hh), out);   \
        done = true;                                                                                   \
    }
    RRG_F32(1, true, false) RRG_F32(2, true, false) RRG_F32(4, true, false)
    RRG_F32(1, true, true) RRG_F32(2, true, true) RRG_F32(4, true, true)
    RRG_F32(1, false, false) RRG_F32(2, false, false) RRG_F32(4, false, false)
    RRG_F32(1, false, true) RRG_F32(2, false, true) RRG_F32(4, false, true)
#undef RRG_F32
    if (!done) fail(RRG_ERR_INVALID, "direct-sum acceleration: unsupported launch shape");
    RRG_LAUNCH_CHECK();
}

void grav(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
          double* out) {
    const dim3 grid(unsigned(pl.tiles), unsigned(pl.slices));
    const int64_t n = p.n;
    if (is_fp32_kind(f.kind)) {
        switch (pl.ib) {
            case 1: k_grav_mixed<1><<<grid, TPB, 0, s>>>(A, B, n, pl.slice_len, out); break;
            case 2: k_grav_mixed<2><<<grid, TPB, 0, s>>>(A, B, n, pl.slice_len, out); break;
            case 4: k_grav_mixed<4><<<grid, TPB, 0, s>>>(A, B, n, pl.slice_len, out); break;
            default: fail(RRG_ERR_INVALID, "direct-sum potential: unsupported launch shape");
        }
    } else {
        const int rsq = f.kind == RRG_FORCE_F64_EXACT ? 0 : 1;
        if (pl.ib == 1 && rsq == 0) k_grav_f64<1, 0><<<grid, TPB, 0, s>>>(p, pl.slice_len, out);
        else if (pl.ib == 2 && rsq == 0) k_grav_f64<2, 0><<<grid, TPB, 0, s>>>(p, pl.slice_len, out);
        else if (pl.ib == 1 && rsq == 1) k_grav_f64<1, 1><<<grid, TPB, 0, s>>>(p, pl.slice_len, out);
        else if (pl.ib == 2 && rsq == 1) k_grav_f64<2, 1><<<grid, TPB, 0, s>>>(p, pl.slice_len, out);
        else fail(RRG_ERR_INVALID, "direct-sum potential: unsupported launch shape");
    }
    RRG_LAUNCH_CHECK();
}

void reduce_slices(hipStream_t s, int64_t n, int comps, int slices, const double* part, double* dst) {
    k_reduce_slices<<<grid1d(n, 256), 256, 0, s>>>(n, comps, slices, part, dst);
    RRG_LAUNCH_CHECK();
}

template <class real> void predict(hipStream_t s, const PState<real>& p, real dt, float4* A, float4* B) {
    if (A && B) {
        if constexpr (std::is_same_v<real, double>) {
            k_predict<real, true><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, A, B);
        } else {
            fail(RRG_ERR_INVALID, "mixed direct sum needs FP64 particle state");
        }
    } else {
        k_predict<real, false><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, nullptr, nullptr);
    }
    RRG_LAUNCH_CHECK();
}

void correct_reduce(hipStream_t s, const PState<double>& p, double dt, int slices, const double* part) {
    k_correct<double><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, slices, part);
    RRG_LAUNCH_CHECK();
}

template <class real> void correct(hipStream_t s, const PState<real>& p, real dt) {
    k_correct<real><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, 1, nullptr);
    RRG_LAUNCH_CHECK();
}

int64_t conservation_blocks(int64_t n) { return std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256)); }

template <class real> int64_t conservation(hipStream_t s, const PState<real>& p, double* partial) {
    const int64_t nb = conservation_blocks(p.n);
    k_conservation<real><<<unsigned(nb), 256, 0, s>>>(p, partial);
    RRG_LAUNCH_CHECK();
    return nb;
}

template void predict<double>(hipStream_t, const PState<double>&, double, float4*, float4*);
template void predict<float>(hipStream_t, const PState<float>&, float, float4*, float4*);
template void correct<double>(hipStream_t, const PState<double>&, double);
template void correct<float>(hipStream_t, const PState<float>&, float);
template int64_t conservation<double>(hipStream_t, const PState<double>&, double*);
template int64_t conservation<float>(hipStream_t, const PState<float>&, double*);

}  // namespace nbody
}  // namespace rrg
