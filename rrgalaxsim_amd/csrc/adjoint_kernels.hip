// Adjoint (reverse-mode) kernels of the direct-sum path.  No counterpart in
// the reference.
//
// Forward step (predict_step_ / correct_step_, nbody.cu:122-181):
//   v* = v + dt a(x);  x' = x + dt (v + v*)/2;  a' = a(x');  v' = (v + v* + dt a')/2
// with a_k(x; m) = sum_j m_j K(x_j - x_k), K(d) = d (|d|^2 + h_kj)^(-3/2).
// Reverse step, given the adjoints (X, V, A) of (x', v', a'):
//   A~ = A + dt V / 2
//   X~ = X + J(x')^T A~           m_bar += (da/dm)(x')^T A~
//   x_bar = X~;  v_bar = V + dt X~;  a_bar = dt V / 2 + dt^2 X~ / 2
// where, with T(d) = r^-3 I - 3 r^-5 d d^T (r^2 = |d|^2 + h):
//   (J^T u)_k     = sum_j T(x_j - x_k) (m_k u_j - m_j u_k)
//   (dm)^T u |_k  = sum_i u_i . K(x_k - x_i) = -sum_j (u_j . d) r^-3,  d = x_j - x_k.
// The pair sums reuse the forward kernels' layout: 256 sources per LDS tile,
// IB targets per lane, source-range slices with partials summed in order.
// Pair math follows the context's force kind (FP64, or FP32 with the same
// hi/lo option and two-sum accumulation).
#include "rrg_internal.hpp"

#include <algorithm>

namespace rrg {
namespace adjoint {
namespace {

constexpr int TPB = 256;

template <int RSQ> __device__ __forceinline__ double inv_sqrt_f64(double r2) {
    if constexpr (RSQ == 0) return 1.0 / sqrt(r2);
    else return rsqrt(r2);
}

__device__ __forceinline__ void two_sum_acc(float& s, float& c, float t) {
    const float u = s + t;
    const float bp = u - s;
    const float e = (s - (u - bp)) + (t - bp);
    s = u;
    c += e;
}

// FP64 VJP.  x: SoA positions (3n), u: SoA adjoint vector (3n).
// out[(slice*4 + c)*n + i]: c = 0..2 J^T u, c = 3 mass gradient.
// With w = m_k u_j - m_j u_k the pair term T(d) w splits into
//   r^-3 w = m_k (r^-3 u_j) - u_k (m_j r^-3)   and
//   -3 r^-5 (d.w) d,  d.w = m_k (d.u_j) - m_j (d.u_k),
// so per pair the kernel accumulates P += r^-3 u_j, S += m_j r^-3 and
// Q += r^-5 (d.w) d, and each target combines m_k P - u_k S - 3 Q once at
// the end (four fewer FP64 operations per pair than forming w; d.u_j also
// gives the mass gradient).
template <int IB, int RSQ>
__global__ __launch_bounds__(TPB) void k_vjp_f64(int64_t n, const double* __restrict__ X, const double* __restrict__ M,
                                                 const double* __restrict__ S2, const double* __restrict__ U,
                                                 int64_t slice_len, double* __restrict__ out) {
    __shared__ double4 sp[TPB];  // x, y, z, m
    __shared__ double4 su[TPB];  // ux, uy, uz, h
    const int t = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    double xk[IB], yk[IB], zk[IB], mk[IB], hk[IB], uxk[IB], uyk[IB], uzk[IB];
    double px[IB], py[IB], pz[IB], sm[IB], qx[IB], qy[IB], qz[IB], gm[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        xk[k] = X[i]; yk[k] = X[n + i]; zk[k] = X[2 * n + i];
        mk[k] = M[i]; hk[k] = 0.5 * S2[i];
        uxk[k] = U[i]; uyk[k] = U[n + i]; uzk[k] = U[2 * n + i];
        px[k] = py[k] = pz[k] = sm[k] = qx[k] = qy[k] = qz[k] = gm[k] = 0.0;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        double4 q = make_double4(0.0, 0.0, 0.0, 0.0), w = make_double4(0.0, 0.0, 0.0, 1.0);
        if (j < j1) {
            q = make_double4(X[j], X[n + j], X[2 * n + j], M[j]);
            w = make_double4(U[j], U[n + j], U[2 * n + j], 0.5 * S2[j]);
        }
        __syncthreads();
        sp[t] = q;
        su[t] = w;
        __syncthreads();
#pragma unroll 2
        for (int jj = 0; jj < TPB; ++jj) {
            const double4 s = sp[jj];
            const double4 u = su[jj];
#pragma unroll
            for (int k = 0; k < IB; ++k) {
                const double dx = s.x - xk[k], dy = s.y - yk[k], dz = s.z - zk[k];
                const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, hk[k] + u.w)));
                const double inv = inv_sqrt_f64<RSQ>(r2);
                const double inv2 = inv * inv;
                const double inv3 = inv * inv2;
                const double inv5 = inv3 * inv2;
                const double du = fma(dz, u.z, fma(dy, u.y, dx * u.x));            // d . u_j
                const double dk = fma(dz, uzk[k], fma(dy, uyk[k], dx * uxk[k]));  // d . u_k
                const double sc = inv5 * fma(mk[k], du, -s.w * dk);                // r^-5 (d . w)
                px[k] = fma(inv3, u.x, px[k]);
                py[k] = fma(inv3, u.y, py[k]);
                pz[k] = fma(inv3, u.z, pz[k]);
                sm[k] = fma(inv3, s.w, sm[k]);
                qx[k] = fma(sc, dx, qx[k]);
                qy[k] = fma(sc, dy, qy[k]);
                qz[k] = fma(sc, dz, qz[k]);
                gm[k] = fma(-du, inv3, gm[k]);
            }
        }
    }
    double* o = out + int64_t(blockIdx.y) * 4 * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) {
            o[i] = fma(mk[k], px[k], fma(-uxk[k], sm[k], -3.0 * qx[k]));
            o[n + i] = fma(mk[k], py[k], fma(-uyk[k], sm[k], -3.0 * qy[k]));
            o[2 * n + i] = fma(mk[k], pz[k], fma(-uzk[k], sm[k], -3.0 * qz[k]));
            o[3 * n + i] = gm[k];
        }
    }
}

// FP32 VJP on packed records A = (x_hi, m), B = (x_lo, h), Uf = (u, 0), with
// the split of k_vjp_f64: per pair P += r^-3 u_j, S += m_j r^-3 and
// Q += r^-5 (d.w) d in FP32 tile sums entering two-sum accumulators, and
// m_k P - u_k S - 3 Q formed in FP64 at the end.
template <int IB, bool HILO>
__global__ __launch_bounds__(TPB) void k_vjp_f32(int64_t n, const float4* __restrict__ A, const float4* __restrict__ B,
                                                 const float4* __restrict__ Uf, int64_t slice_len,
                                                 double* __restrict__ out) {
    constexpr int NC = 8;  // P (3), S, Q (3), mass gradient
    __shared__ float4 sa[TPB];
    __shared__ float4 sb[TPB];
    __shared__ float4 su[TPB];
    const int t = threadIdx.x;
    const int64_t base = int64_t(blockIdx.x) * TPB * IB;
    float xh[IB], yh[IB], zh[IB], xl[IB], yl[IB], zl[IB], mk[IB], hk[IB], uxk[IB], uyk[IB], uzk[IB];
    float sg[NC][IB], cg[NC][IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = min(base + k * TPB + t, n - 1);
        const float4 a = A[i], b = B[i], u = Uf[i];
        xh[k] = a.x; yh[k] = a.y; zh[k] = a.z; mk[k] = a.w;
        xl[k] = b.x; yl[k] = b.y; zl[k] = b.z; hk[k] = b.w;
        uxk[k] = u.x; uyk[k] = u.y; uzk[k] = u.z;
#pragma unroll
        for (int c = 0; c < NC; ++c) sg[c][k] = cg[c][k] = 0.f;
    }
    const int64_t j0 = int64_t(blockIdx.y) * slice_len;
    const int64_t j1 = min(n, j0 + slice_len);
    for (int64_t jt = j0; jt < j1; jt += TPB) {
        const int64_t j = jt + t;
        float4 qa = make_float4(0.f, 0.f, 0.f, 0.f), qb = make_float4(0.f, 0.f, 0.f, 1.f);
        float4 qu = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < j1) {
            qa = A[j];
            qb = B[j];
            qu = Uf[j];
        }
        __syncthreads();
        sa[t] = qa;
        sb[t] = qb;
        su[t] = qu;
        __syncthreads();
        float tg[NC][IB];
#pragma unroll
        for (int k = 0; k < IB; ++k)
#pragma unroll
            for (int c = 0; c < NC; ++c) tg[c][k] = 0.f;
#pragma unroll 2
        for (int jj = 0; jj < TPB; ++jj) {
            const float4 a = sa[jj], b = sb[jj], u = su[jj];
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
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, hk[k] + b.w)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv2 = inv * inv;
                const float inv3 = inv * inv2;
                const float inv5 = inv3 * inv2;
                const float du = fmaf(dz, u.z, fmaf(dy, u.y, dx * u.x));            // d . u_j
                const float dk = fmaf(dz, uzk[k], fmaf(dy, uyk[k], dx * uxk[k]));  // d . u_k
                const float sc = inv5 * fmaf(mk[k], du, -a.w * dk);                // r^-5 (d . w)
                tg[0][k] = fmaf(inv3, u.x, tg[0][k]);
                tg[1][k] = fmaf(inv3, u.y, tg[1][k]);
                tg[2][k] = fmaf(inv3, u.z, tg[2][k]);
                tg[3][k] = fmaf(inv3, a.w, tg[3][k]);
                tg[4][k] = fmaf(sc, dx, tg[4][k]);
                tg[5][k] = fmaf(sc, dy, tg[5][k]);
                tg[6][k] = fmaf(sc, dz, tg[6][k]);
                tg[7][k] = fmaf(-du, inv3, tg[7][k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IB; ++k)
#pragma unroll
            for (int c = 0; c < NC; ++c) two_sum_acc(sg[c][k], cg[c][k], tg[c][k]);
    }
    double* o = out + int64_t(blockIdx.y) * 4 * n;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
        const int64_t i = base + k * TPB + t;
        if (i < n) {
            double v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = double(sg[c][k]) + double(cg[c][k]);
            const double m = double(mk[k]), ux = double(uxk[k]), uy = double(uyk[k]), uz = double(uzk[k]);
            o[i] = fma(m, v[0], fma(-ux, v[3], -3.0 * v[4]));
            o[n + i] = fma(m, v[1], fma(-uy, v[3], -3.0 * v[5]));
            o[2 * n + i] = fma(m, v[2], fma(-uz, v[3], -3.0 * v[6]));
            o[3 * n + i] = v[7];
        }
    }
}

// A~ = A + dt V / 2; for the FP32 kinds also the packed copy of A~.
__global__ void k_adj_pre(int64_t n, double dt, const double* __restrict__ V, const double* __restrict__ Aa,
                          double* __restrict__ At, float4* __restrict__ Uf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        u[c] = Aa[c * n + i] + 0.5 * dt * V[c * n + i];
        At[c * n + i] = u[c];
    }
    if (Uf) Uf[i] = make_float4(float(u[0]), float(u[1]), float(u[2]), 0.f);
}

// Sum the VJP slice partials and advance the adjoint state one step back:
// X~ = X + J^T A~;  m_bar += dm;  V <- V + dt X~;  A <- dt V/2 + dt^2 X~/2;  X <- X~.
// `last` (the closing VJP at x_0): only X and m_bar are updated.
__global__ void k_adj_post(int64_t n, double dt, int slices, const double* __restrict__ part, double* __restrict__ X,
                           double* __restrict__ V, double* __restrict__ Aa, double* __restrict__ gm, int last) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double acc = part[int64_t(c) * n + i];
        for (int q = 1; q < slices; ++q) acc += part[(int64_t(q) * 4 + c) * n + i];
        s[c] = acc;
    }
    gm[i] += s[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double xt = X[c * n + i] + s[c];
        X[c * n + i] = xt;
        if (!last) {
            const double v = V[c * n + i];
            V[c * n + i] = v + dt * xt;
            Aa[c * n + i] = 0.5 * dt * v + 0.5 * dt * dt * xt;
        }
    }
}

// Packed FP32 source records from SoA positions (as k_pack_mixed).
__global__ void k_pack_pos(int64_t n, const double* __restrict__ X, const double* __restrict__ M,
                           const double* __restrict__ S2, float4* __restrict__ A, float4* __restrict__ B) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = X[i], y = X[n + i], z = X[2 * n + i];
    const float xh = float(x), yh = float(y), zh = float(z);
    A[i] = make_float4(xh, yh, zh, float(M[i]));
    B[i] = make_float4(float(x - double(xh)), float(y - double(yh)), float(z - double(zh)), float(0.5 * S2[i]));
}

inline unsigned grid1d(int64_t n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

}  // namespace

void vjp(hipStream_t s, const nbody::Plan& pl, int kind, int64_t n, const double* X, const double* M,
         const double* S2, const double* U, const float4* A, const float4* B, const float4* Uf, double* out) {
    const dim3 grid(unsigned(pl.tiles), unsigned(pl.slices));
    const int ib = pl.ib;
    if (nbody::is_fp32_kind(kind)) {
        const bool hilo = kind == RRG_FORCE_MIXED;
        if (ib == 1 && hilo) k_vjp_f32<1, true><<<grid, TPB, 0, s>>>(n, A, B, Uf, pl.slice_len, out);
        else if (ib == 2 && hilo) k_vjp_f32<2, true><<<grid, TPB, 0, s>>>(n, A, B, Uf, pl.slice_len, out);
        else if (ib == 1 && !hilo) k_vjp_f32<1, false><<<grid, TPB, 0, s>>>(n, A, B, Uf, pl.slice_len, out);
        else if (ib == 2 && !hilo) k_vjp_f32<2, false><<<grid, TPB, 0, s>>>(n, A, B, Uf, pl.slice_len, out);
        else fail(RRG_ERR_INVALID, "adjoint: unsupported launch shape");
    } else {
        const int rsq = kind == RRG_FORCE_F64_EXACT ? 0 : 1;
        if (ib == 1 && rsq == 1) k_vjp_f64<1, 1><<<grid, TPB, 0, s>>>(n, X, M, S2, U, pl.slice_len, out);
        else if (ib == 2 && rsq == 1) k_vjp_f64<2, 1><<<grid, TPB, 0, s>>>(n, X, M, S2, U, pl.slice_len, out);
        else if (ib == 1 && rsq == 0) k_vjp_f64<1, 0><<<grid, TPB, 0, s>>>(n, X, M, S2, U, pl.slice_len, out);
        else if (ib == 2 && rsq == 0) k_vjp_f64<2, 0><<<grid, TPB, 0, s>>>(n, X, M, S2, U, pl.slice_len, out);
        else fail(RRG_ERR_INVALID, "adjoint: unsupported launch shape");
    }
    RRG_LAUNCH_CHECK();
}

void pre(hipStream_t s, int64_t n, double dt, const double* V, const double* Aa, double* At, float4* Uf) {
    k_adj_pre<<<grid1d(n, 256), 256, 0, s>>>(n, dt, V, Aa, At, Uf);
    RRG_LAUNCH_CHECK();
}

void post(hipStream_t s, int64_t n, double dt, int slices, const double* part, double* X, double* V, double* Aa,
          double* gm, bool last) {
    k_adj_post<<<grid1d(n, 256), 256, 0, s>>>(n, dt, slices, part, X, V, Aa, gm, last ? 1 : 0);
    RRG_LAUNCH_CHECK();
}

void pack_pos(hipStream_t s, int64_t n, const double* X, const double* M, const double* S2, float4* A, float4* B) {
    k_pack_pos<<<grid1d(n, 256), 256, 0, s>>>(n, X, M, S2, A, B);
    RRG_LAUNCH_CHECK();
}

}  // namespace adjoint
}  // namespace rrg
