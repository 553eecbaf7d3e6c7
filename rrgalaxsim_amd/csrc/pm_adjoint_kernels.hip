// Reverse sweep of the PM step for gfx950 (no reference counterpart: the
// reference solver has no adjoint).  Forward map, per grid build:
//   x' = x + dt v + dt^2/2 a                      predict_step_ (nbody.cu:135-151)
//   M  = CIC(x', m)                                deposit_cell_mass_CIC_ (rm_p2mesh.cu:176-236)
//   s  = 4 pi k_f (M / dx^3 - rho_mean)            wave.cu:150-153, 183-186
//   phi_{k+1} = B(K_m L phi_k + K_a phi_k + K_o phi_{k-1} + s), k < I   wave.cu:106-193
//   G  = -D phi_I / dx                             calc_acceleration_field_ (rm_p2mesh.cu:238-290)
//   a' = Interp(x', G)                             rm_mesh2p.cu:48-114
//   v' = v + dt/2 (a + a')                         correct_step_ (nbody.cu:171-180)
// B zeroes the faces on zero-Dirichlet grids.  The transposed operators:
//   * Interp^T is a deposit of the particle adjoint with the interpolation
//     weights, plus d(weights)/dx' dotted with G; the deposit is int64 fixed
//     point (resolution 2^-E, n max|abar + h vbar| 2^E < 2^61, E found on the
//     device), so the grid adjoint does not depend on the order of the
//     atomics and the gradients repeat bitwise from run to run;
//   * D^T is evaluated per cell as a gather from the three cells that read it;
//   * the stencil recurrence reverses into the same stencil with no source,
//     mu_{k-1} = B(K_m L mu_k + K_a mu_k + K_o mu_{k+1}), while the source
//     adjoint sbar accumulates sum_k mu_k;
//   * CIC^T is an interpolation of sbar with the CIC weights, plus their
//     derivatives times m (zero on a clamped axis).
// Particle adjoints are FP64 whatever the state precision; grid adjoints are
// kept in the state's precision except the interpolation adjoint, which is
// accumulated in int64 fixed point.
#include "rrg_internal.hpp"
#include "wave_fused.hpp"

#include <cmath>

namespace rrg {
namespace pm {
namespace {

template <class real> __device__ __forceinline__ real clampr(real v, real lo, real hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

__device__ __forceinline__ void xcd_block(int nbx, int nby, int nbz, int& bx, int& by, int& bz) {
    const int nb = nbx * nby * nbz;
    const int b = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    int r = b;
    if ((nb & 7) == 0) {
        const int per = nb >> 3;
        r = (b & 7) * per + (b >> 3);
    }
    bx = r % nbx;
    by = (r / nbx) % nby;
    bz = r / (nbx * nby);
}

// Interpolation adjoint at x' (positions SoA, n per component).  The cell
// choice and weights repeat k_gather's arithmetic in `real`; the products are
// formed in FP64.  at = abar + h vbar is the adjoint of the interpolated
// acceleration (h = dt/2 for a step, 0 for the precalc build).
//
// The deposit of at into Gb merges lanes first: the wave's particles are
// sorted by their 8-cell stencil (a 64-bit key of the clamped corner
// indices) with a shuffle bitonic sort, a segmented inclusive scan sums each
// run of equal stencils, and the run's last lane rounds its 8 x 3 sums to
// multiples of 2^-E and issues them as int64 atomics.  The particle arrays
// are kept in cell order during a recorded run, so dense regions form long
// runs; the merge order within a wave follows the particle order, so it is
// the same in every run.
constexpr unsigned long long kNoStencil = ~0ull;

__device__ __forceinline__ double wave_max_d(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// max over particles and components of |abar + h vbar|, as a bit pattern
// (non-negative doubles order as their bits): one atomic per workgroup.
__global__ __launch_bounds__(256) void k_adj_at_max(int64_t n, double h, const double* __restrict__ Vb,
                                                    const double* __restrict__ Ab, AdjScale* __restrict__ sc) {
    __shared__ double wm[4];
    double mx = 0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        for (int d = 0; d < 3; ++d) mx = fmax(mx, fabs(Ab[d * n + i] + h * Vb[d * n + i]));
    mx = wave_max_d(mx);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        mx = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
        atomicMax(&sc->maxbits, (unsigned long long)__double_as_longlong(mx));
    }
}

// E with n * max * 2^E < 2^61 (the bound of any cell's, and any partial,
// sum); the max is cleared for the next grid build.
__global__ void k_adj_scale(int64_t n, AdjScale* __restrict__ sc) {
    const double mx = __longlong_as_double((long long)sc->maxbits);
    sc->maxbits = 0;
    const double b = mx * (double)n;
    int ex = 0;
    (void)frexp(b > 1e-300 ? b : 1e-300, &ex);  // b in [2^(ex-1), 2^ex)
    sc->e = (b < 1e300) ? min(61 - ex, 1000) : 0;
}

template <class real>
__global__ __launch_bounds__(256) void k_adj_interp(int64_t n, const real* __restrict__ x, Geo<real> g,
                                                    const real* __restrict__ G4, double h, double* __restrict__ Xb,
                                                    const double* __restrict__ Vb, const double* __restrict__ Ab,
                                                    unsigned long long* __restrict__ Gq,
                                                    const AdjScale* __restrict__ sc) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int nx = g.nx;
    const size_t ncell = size_t(nx) * nx * nx;
    const unsigned long long K = 2ull * unsigned(nx);
    unsigned long long key = kNoStencil;
    double at[3] = {0, 0, 0};
    double w[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    if (q < n) {
        int lo[3], hi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const real f = (x[d * n + q] - g.domain_min) / g.dx - (real)0.5f;
            const int i0 = (int)floorf((float)f);
            const real wt = f - (real)i0;
            w[d][0] = (double)((real)1.0f - wt);
            w[d][1] = (double)wt;
            lo[d] = max(0, min(i0, nx - 1));
            hi[d] = max(0, min(i0 + 1, nx - 1));
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) at[d] = Ab[d * n + q] + h * Vb[d * n + q];
        const double inv = 1.0 / (double)g.dx;
        double gx = 0, gy = 0, gz = 0;
#pragma unroll
        for (int k = 0; k <= 1; ++k) {
#pragma unroll
            for (int j = 0; j <= 1; ++j) {
#pragma unroll
                for (int i = 0; i <= 1; ++i) {
                    const size_t c = size_t(i ? hi[0] : lo[0]) + size_t(j ? hi[1] : lo[1]) * nx +
                                     size_t(k ? hi[2] : lo[2]) * nx * nx;
                    const real* f = G4 + 4 * c;
                    const double ga = (double)f[0] * at[0] + (double)f[1] * at[1] + (double)f[2] * at[2];
                    const double wi = w[0][i], wj = w[1][j], wk = w[2][k];
                    gx += (i ? inv : -inv) * wj * wk * ga;
                    gy += wi * (j ? inv : -inv) * wk * ga;
                    gz += wi * wj * (k ? inv : -inv) * ga;
                }
            }
        }
        Xb[q] += gx;
        Xb[n + q] += gy;
        Xb[2 * n + q] += gz;
        // stencil key: per axis 2*lo + (hi != lo)
        key = 0;
#pragma unroll
        for (int d = 2; d >= 0; --d) key = key * K + (unsigned long long)(2 * lo[d] + (hi[d] != lo[d] ? 1 : 0));
    }
    // bitonic sort of (key, lane) over the wave
    int src = lane;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const unsigned long long ok = __shfl_xor(key, stride, 64);
            const int os = __shfl_xor(src, stride, 64);
            const bool up = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            const bool other_smaller = ok < key || (ok == key && os < src);
            if (lower == up ? other_smaller : !other_smaller) {
                key = ok;
                src = os;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        at[d] = __shfl(at[d], src, 64);
        w[d][0] = __shfl(w[d][0], src, 64);
        w[d][1] = __shfl(w[d][1], src, 64);
    }
    const unsigned long long kprev = __shfl_up(key, 1, 64);
    const unsigned long long knext = __shfl_down(key, 1, 64);
    const int head = (lane == 0 || kprev != key) ? 1 : 0;
    const bool tail = (lane == 63 || knext != key) && key != kNoStencil;
    size_t cell[8];
    if (tail) {
        unsigned long long r = key;
        int lo[3], hi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const unsigned c = unsigned(r % K);
            r /= K;
            lo[d] = int(c >> 1);
            hi[d] = lo[d] + int(c & 1);
        }
#pragma unroll
        for (int k = 0; k <= 1; ++k)
#pragma unroll
            for (int j = 0; j <= 1; ++j)
#pragma unroll
                for (int i = 0; i <= 1; ++i)
                    cell[i + 2 * j + 4 * k] = size_t(i ? hi[0] : lo[0]) + size_t(j ? hi[1] : lo[1]) * nx +
                                              size_t(k ? hi[2] : lo[2]) * nx * nx;
    }
    // one component at a time: corner sums, segmented inclusive scan, atomics
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double v[8];
#pragma unroll
        for (int k = 0; k <= 1; ++k)
#pragma unroll
            for (int j = 0; j <= 1; ++j)
#pragma unroll
                for (int i = 0; i <= 1; ++i) v[i + 2 * j + 4 * k] = w[0][i] * w[1][j] * w[2][k] * at[d];
        int f = head;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int fu = __shfl_up(f, off, 64);
            double vu[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) vu[c] = __shfl_up(v[c], off, 64);
            if (lane >= off) {
                if (!f) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) v[c] += vu[c];
                }
                f |= fu;
            }
        }
        if (tail) {
            const double scale = ldexp(1.0, sc->e);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const long long qv = (long long)rint(v[c] * scale);
                if (qv != 0) atomicAdd(Gq + d * ncell + cell[c], (unsigned long long)qv);
            }
        }
    }
}

// mu += B(coef * D^T Gb): each cell gathers the terms of the (up to three)
// cells per axis whose difference read it.  Gb is SoA, three grids.
template <class real, bool PERIODIC>
__global__ __launch_bounds__(256) void k_adj_gradient_t(int NX, double coef, const long long* __restrict__ Gq,
                                                        const AdjScale* __restrict__ sc, real* __restrict__ mu) {
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    const int iz = bz;
    if (ix >= NX || iy >= NX) return;
    if (!PERIODIC && (ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1 || iz == 0 || iz == NX - 1)) return;
    const size_t ncell = size_t(NX) * NX * NX;
    const size_t st[3] = {1, size_t(NX), size_t(NX) * NX};
    const int id[3] = {ix, iy, iz};
    const size_t c = size_t(ix) + iy * st[1] + iz * st[2];
    const double unit = ldexp(1.0, -sc->e);
    double sum = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long* gq = Gq + d * ncell;
        auto gd = [&](size_t k) { return (double)gq[k] * unit; };
        const int i = id[d];
        if (i >= 1) sum += (i - 1 == 0 ? 1.0 : 0.5) * gd(c - st[d]);  // this cell is the +1 neighbour
        if (i == 0) sum -= gd(c);                                      // one-sided faces read the cell itself
        else if (i == NX - 1) sum += gd(c);
        if (i + 1 <= NX - 1) sum -= (i + 1 == NX - 1 ? 1.0 : 0.5) * gd(c + st[d]);  // the -1 neighbour
    }
    mu[c] += (real)(coef * sum);
}

// One transposed stencil iteration (k_wave's z-march with no source), and
// the source adjoint sbar (= cur on the first iteration of a grid build,
// += cur afterwards).  Same z-march depth rule as the forward stencil.
inline int adj_wave_zb(int nx) { return nx <= 96 ? 2 : 4; }

template <class real, bool PERIODIC, bool FIRST, int ZB>
__global__ __launch_bounds__(256) void k_adj_wave(real* __restrict__ pn, const real* __restrict__ p,
                                                  const real* __restrict__ po, real* __restrict__ sbar, int NX,
                                                  WaveK<real> k) {
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    if (ix >= NX || iy >= NX) return;
    const size_t sy = size_t(NX), sz = size_t(NX) * NX;
    auto nb = [&](int i) { return PERIODIC ? (i < 0 ? i + NX : (i >= NX ? i - NX : i)) : max(0, min(i, NX - 1)); };
    const int xm = nb(ix - 1), xp = nb(ix + 1), ym = nb(iy - 1), yp = nb(iy + 1);
    const bool edge_xy = ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1;
    const size_t col = size_t(ix) + iy * sy;
    const int z0 = bz * ZB;
    real below = p[col + nb(z0 - 1) * sz];
    real cen = p[col + size_t(z0) * sz];
#pragma unroll 2
    for (int t = 0; t < ZB; ++t) {
        const int iz = z0 + t;
        if (iz >= NX) break;
        const real above = p[col + nb(iz + 1) * sz];
        const size_t pl = size_t(iz) * sz;
        const size_t c = col + pl;
        real out;
        if (!PERIODIC && (edge_xy || iz == 0 || iz == NX - 1)) {
            out = 0;
        } else {
            const real lap_sum = p[xm + iy * sy + pl] + p[xp + iy * sy + pl] + p[ix + ym * sy + pl] +
                                 p[ix + yp * sy + pl] + below + above;
            out = lap_sum * k.k_main + cen * k.k_actual + po[c] * k.k_old;
        }
        pn[c] = out;
        sbar[c] = FIRST ? cen : sbar[c] + cen;
        below = cen;
        cen = above;
    }
}

// The transposed stencil with per-cell coefficients (non-periodic grids,
// where the absorbing layer varies them): the forward reads its neighbours
// with its own K_MAIN, so the transpose gathers K_MAIN of each neighbour:
//   pn[c] = B(sum_nbrs km[n] p[n] + ka[c] p[c] + ko[c] po[c]),  sbar (+)= p.
// p is zero on the faces, so the clamped neighbour of a face cell is unused.
template <class real, bool FIRST, int ZB>
__global__ __launch_bounds__(256) void k_adj_wave_var(real* __restrict__ pn, const real* __restrict__ p,
                                                      const real* __restrict__ po, real* __restrict__ sbar, int NX,
                                                      const real* __restrict__ km, const real* __restrict__ ka,
                                                      const real* __restrict__ ko) {
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    if (ix >= NX || iy >= NX) return;
    const size_t sy = size_t(NX), sz = size_t(NX) * NX;
    const bool edge_xy = ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1;
    const size_t col = size_t(ix) + iy * sy;
    for (int t = 0; t < ZB; ++t) {
        const int iz = bz * ZB + t;
        if (iz >= NX) break;
        const size_t c = col + size_t(iz) * sz;
        real out = 0;
        if (!(edge_xy || iz == 0 || iz == NX - 1)) {
            const size_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
            real acc = 0;
#pragma unroll
            for (int q = 0; q < 6; ++q) acc += km[nb[q]] * p[nb[q]];
            out = acc + ka[c] * p[c] + ko[c] * po[c];
        }
        pn[c] = out;
        sbar[c] = FIRST ? p[c] : sbar[c] + p[c];
    }
}

// CIC adjoint at x' and the transposed Verlet step.  mbar = CIC^T (alpha
// sbar) feeds the mass gradient (alpha = d source / d cell mass, per cell on
// non-periodic grids); m * d(CIC^T (alpha sbar))/dx' joins the position
// adjoint; then (step builds only)
//   vbar <- vbar + dt xbar,  abar <- dt/2 vbar_old + dt^2/2 xbar.
template <class real, bool CLOSING>
__global__ __launch_bounds__(256) void k_adj_cic(int64_t n, const real* __restrict__ x, const real* __restrict__ m,
                                                 Geo<real> g, const real* __restrict__ sbar, double alpha, double dt,
                                                 double* __restrict__ Xb, double* __restrict__ Vb,
                                                 double* __restrict__ Ab, double* __restrict__ gm,
                                                 const real* __restrict__ alpha_cell) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int NX = g.nx;
    int i0[3], i1[3];
    double w[3][2], dw[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const real raw = (x[d * n + q] - g.domain_min) * g.inv_dx - (real)0.5f;
        const real u = clampr<real>(raw, 0, g.upper);
        const int b = max(0, min((int)floor(u), NX - 1));
        i0[d] = b;
        i1[d] = min(b + 1, NX - 1);
        const real f = u - (real)b;
        w[d][0] = (double)((real)1.f - f);
        w[d][1] = (double)f;
        dw[d] = (raw < (real)0 || raw > g.upper) ? 0.0 : (double)g.inv_dx;
    }
    double mb = 0, dm0 = 0, dm1 = 0, dm2 = 0;
#pragma unroll
    for (int k = 0; k <= 1; ++k) {
#pragma unroll
        for (int j = 0; j <= 1; ++j) {
#pragma unroll
            for (int i = 0; i <= 1; ++i) {
                const size_t c = size_t(i ? i1[0] : i0[0]) + size_t(j ? i1[1] : i0[1]) * NX +
                                 size_t(k ? i1[2] : i0[2]) * NX * NX;
                const double s = alpha_cell ? (double)alpha_cell[c] * (double)sbar[c] : alpha * (double)sbar[c];
                const double wi = w[0][i], wj = w[1][j], wk = w[2][k];
                mb += wi * wj * wk * s;
                dm0 += (i ? dw[0] : -dw[0]) * wj * wk * s;
                dm1 += wi * (j ? dw[1] : -dw[1]) * wk * s;
                dm2 += wi * wj * (k ? dw[2] : -dw[2]) * s;
            }
        }
    }
    gm[q] += mb;
    const double am = (double)m[q];
    const double dm[3] = {dm0, dm1, dm2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double xt = Xb[d * n + q] + am * dm[d];
        Xb[d * n + q] = xt;
        if (!CLOSING) {
            const double vb = Vb[d * n + q];
            Vb[d * n + q] = vb + dt * xt;
            Ab[d * n + q] = 0.5 * dt * vb + 0.5 * dt * dt * xt;
        }
    }
}

inline unsigned grid1d(int64_t n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

}  // namespace

template <class real>
void adj_interp(hipStream_t s, int64_t n, const real* x, const Geo<real>& g, const real* G4, double h, double* Xb,
                const double* Vb, const double* Ab, long long* Gq, AdjScale* sc) {
    if (n == 0) return;
    const unsigned nb = unsigned(std::min<int64_t>(512, (n + 255) / 256));
    k_adj_at_max<<<nb, 256, 0, s>>>(n, h, Vb, Ab, sc);
    RRG_LAUNCH_CHECK();
    k_adj_scale<<<1, 1, 0, s>>>(n, sc);
    RRG_LAUNCH_CHECK();
    k_adj_interp<real><<<grid1d(n, 256), 256, 0, s>>>(n, x, g, G4, h, Xb, Vb, Ab,
                                                      reinterpret_cast<unsigned long long*>(Gq), sc);
    RRG_LAUNCH_CHECK();
}

template <class real>
void adj_gradient_t(hipStream_t s, int nx, real dx, bool periodic, const long long* Gq, const AdjScale* sc, real* mu) {
    const double coef = (double)(real)(-1. / (double)dx);  // the forward's coefficient
    const dim3 grid(unsigned((nx + 63) / 64), unsigned((nx + 3) / 4), unsigned(nx));
    if (periodic) k_adj_gradient_t<real, true><<<grid, dim3(64, 4, 1), 0, s>>>(nx, coef, Gq, sc, mu);
    else k_adj_gradient_t<real, false><<<grid, dim3(64, 4, 1), 0, s>>>(nx, coef, Gq, sc, mu);
    RRG_LAUNCH_CHECK();
}

template <class real>
void adj_wave(hipStream_t s, int nx, bool periodic, const WaveK<real>& k, real* pn, const real* p, const real* po,
              real* sbar, bool first, const real* km, const real* ka, const real* ko) {
    const int zb = adj_wave_zb(nx);
    const dim3 grid(unsigned((nx + 63) / 64), unsigned((nx + 3) / 4), unsigned((nx + zb - 1) / zb));
    const dim3 block(64, 4, 1);
    if (km) {
        auto gv = [&](auto kern) { kern<<<grid, block, 0, s>>>(pn, p, po, sbar, nx, km, ka, ko); };
        if (zb == 2) first ? gv(k_adj_wave_var<real, true, 2>) : gv(k_adj_wave_var<real, false, 2>);
        else first ? gv(k_adj_wave_var<real, true, 4>) : gv(k_adj_wave_var<real, false, 4>);
        RRG_LAUNCH_CHECK();
        return;
    }
    auto go = [&](auto kern) { kern<<<grid, block, 0, s>>>(pn, p, po, sbar, nx, k); };
    if (zb == 2) {
        if (periodic) first ? go(k_adj_wave<real, true, true, 2>) : go(k_adj_wave<real, true, false, 2>);
        else first ? go(k_adj_wave<real, false, true, 2>) : go(k_adj_wave<real, false, false, 2>);
    } else {
        if (periodic) first ? go(k_adj_wave<real, true, true, 4>) : go(k_adj_wave<real, true, false, 4>);
        else first ? go(k_adj_wave<real, false, true, 4>) : go(k_adj_wave<real, false, false, 4>);
    }
    RRG_LAUNCH_CHECK();
}

template <class real> int adj_wave_levels(int nx, bool periodic, int remaining) {
    return fused_levels<real>(nx, periodic ? RRG_BC_PERIODIC : RRG_BC_ZERO_DIRICHLET, remaining, 3);
}

template <class real>
void adj_wave_fused(hipStream_t s, int nx, const WaveK<real>& k, int levels, real* qa, real* qb, const real* p,
                    const real* po, real* sbar, bool first) {
    if (levels < 2 || adj_wave_levels<real>(nx, true, levels) != levels)
        fail(RRG_ERR_INVALID, "adj_wave_fused: no launch shape for this grid and iteration count");
    if (first) launch_fused<real, true, true>(s, nx, levels, qa, qb, p, po, nullptr, sbar, k);
    else launch_fused<real, true, false>(s, nx, levels, qa, qb, p, po, nullptr, sbar, k);
    RRG_LAUNCH_CHECK();
}

template <class real>
void adj_cic(hipStream_t s, int64_t n, const real* x, const real* m, const Geo<real>& g, const real* sbar,
             double alpha, double dt, bool closing, double* Xb, double* Vb, double* Ab, double* gm,
             const real* alpha_cell) {
    if (n == 0) return;
    if (closing) k_adj_cic<real, true><<<grid1d(n, 256), 256, 0, s>>>(n, x, m, g, sbar, alpha, dt, Xb, Vb, Ab, gm,
                                                                      alpha_cell);
    else k_adj_cic<real, false><<<grid1d(n, 256), 256, 0, s>>>(n, x, m, g, sbar, alpha, dt, Xb, Vb, Ab, gm,
                                                                alpha_cell);
    RRG_LAUNCH_CHECK();
}

#define RRG_PMADJ_INST(real)                                                                                     \
    template void adj_interp<real>(hipStream_t, int64_t, const real*, const Geo<real>&, const real*, double,     \
                                   double*, const double*, const double*, long long*, AdjScale*);                \
    template void adj_gradient_t<real>(hipStream_t, int, real, bool, const long long*, const AdjScale*, real*);  \
    template void adj_wave<real>(hipStream_t, int, bool, const WaveK<real>&, real*, const real*, const real*,    \
                                 real*, bool, const real*, const real*, const real*);                            \
    template int adj_wave_levels<real>(int, bool, int);                                                          \
    template void adj_wave_fused<real>(hipStream_t, int, const WaveK<real>&, int, real*, real*, const real*,     \
                                       const real*, real*, bool);                                                \
    template void adj_cic<real>(hipStream_t, int64_t, const real*, const real*, const Geo<real>&, const real*,  \
                                double, double, bool, double*, double*, double*, double*, const real*);

RRG_PMADJ_INST(float)
RRG_PMADJ_INST(double)

}  // namespace pm
}  // namespace rrg
