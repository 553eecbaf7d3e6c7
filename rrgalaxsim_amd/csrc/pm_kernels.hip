// Particle-mesh ("wave" model) kernels for gfx950.
//
// Reference:
//   deposit_cell_mass_CIC_       solver/lib/rrgsim_rect_mesh/src/rm_p2mesh.cu:176-236
//   calc_acceleration_field_     rm_p2mesh.cu:238-290
//   assign_particles_to_cells_   rm_p2mesh.cu:25-66 (NGP cell only)
//   interpolate_acceleration_to_particles_kernel  rm_mesh2p.cu:48-114
//   projectCellPhiOnParticles    rm_mesh2p.cu:29-38
//   wave_diss_iteration          solver/lib/rrgsim_wave/src/wave.cu:106-193
//   reduce_step_error            solver/lib/rrgsim_wave/src/wave_steady_checker.cu:30-85
//
// Every kernel keeps the reference's operation order and its float/double
// promotions, and the file is compiled with -ffp-contract=off, so for a given
// `real` the results are bitwise those of a sequential CPU evaluation.  The
// deposit is the exception by design: its eight products m*w are rounded to
// int64 fixed point (resolution 2^-e) and summed with integer atomics, which
// makes the cell masses independent of particle order.
#include "rrg_internal.hpp"
#include "wave_fused.hpp"

#include <cmath>
#include <cstring>
#include <type_traits>

namespace rrg {
namespace pm {
namespace {

constexpr double kPi = 3.14159265358979;  // co_info_structs.h:6

template <class real> struct Vec4;
template <> struct Vec4<float> { using T = float4; };
template <> struct Vec4<double> { using T = double4; };

template <class real> __device__ __forceinline__ real clampr(real v, real lo, real hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// ------------------------------------------------------------------ deposit
// rint(x) as an integer for |x| < 2^51: x + 1.5 * 2^52 lands in [2^52, 2^53),
// where doubles are the integers, so the addition rounds x to the nearest
// integer (ties to even, as rint in the default rounding mode) and the bit
// patterns of the sum and of 1.5 * 2^52 differ by that integer.
__device__ __forceinline__ long long rint_small(double x) {
    const double magic = 6755399441055744.0;  // 1.5 * 2^52
    return __double_as_longlong(x + magic) - __double_as_longlong(magic);
}

// CIC stencil of deposit_cell_mass_CIC_ at position r: the 8 cells, the 8
// fixed-point contributions rint(m*w * 2^e) and the base cell (i0, j0, k0)
// with the upper corner's steps (0 or 1 per axis; 0 where clamped onto the
// last cell).  Returns the base cell index.  FASTQ: |m| * 2^e < 2^51 for
// every particle, so the rounding can take rint_small.
template <class real, bool FASTQ>
__device__ __forceinline__ unsigned cic_fixed(const real r[3], real m, const Geo<real>& g, double scale, long long q[8],
                                              unsigned c[8], int base[3], int step[3]) {
    const int NX = g.nx;
    real u = (r[0] - g.domain_min) * g.inv_dx - (real)0.5f;
    real v = (r[1] - g.domain_min) * g.inv_dx - (real)0.5f;
    real w = (r[2] - g.domain_min) * g.inv_dx - (real)0.5f;
    u = clampr<real>(u, 0, g.upper);
    v = clampr<real>(v, 0, g.upper);
    w = clampr<real>(w, 0, g.upper);
    // The clamp above already bounds finite u; the index clamp only matters
    // for a NaN coordinate, which must not address outside the grid.
    const int i0 = max(0, min((int)floor(u), NX - 1));
    const int j0 = max(0, min((int)floor(v), NX - 1));
    const int k0 = max(0, min((int)floor(w), NX - 1));
    const int i1 = min(i0 + 1, NX - 1), j1 = min(j0 + 1, NX - 1), k1 = min(k0 + 1, NX - 1);
    base[0] = i0; base[1] = j0; base[2] = k0;
    step[0] = i1 - i0; step[1] = j1 - j0; step[2] = k1 - k0;
    const real fx = u - (real)i0, fy = v - (real)j0, fz = w - (real)k0;
    const real wx0 = (real)1.f - fx, wx1 = fx;
    const real wy0 = (real)1.f - fy, wy1 = fy;
    const real wz0 = (real)1.f - fz, wz1 = fz;
    const unsigned sy = unsigned(NX), sz = unsigned(NX) * unsigned(NX);
    c[0] = i0 + j0 * sy + k0 * sz;
    c[1] = i1 + j0 * sy + k0 * sz;
    c[2] = i0 + j1 * sy + k0 * sz;
    c[3] = i1 + j1 * sy + k0 * sz;
    c[4] = i0 + j0 * sy + k1 * sz;
    c[5] = i1 + j0 * sy + k1 * sz;
    c[6] = i0 + j1 * sy + k1 * sz;
    c[7] = i1 + j1 * sy + k1 * sz;
    const real wt[8] = {wx0 * wy0 * wz0, wx1 * wy0 * wz0, wx0 * wy1 * wz0, wx1 * wy1 * wz0,
                        wx0 * wy0 * wz1, wx1 * wy0 * wz1, wx0 * wy1 * wz1, wx1 * wy1 * wz1};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double x = (double)(m * wt[k]) * scale;
        q[k] = FASTQ ? rint_small(x) : (long long)rint(x);
    }
    return c[0];
}

// Predictor fused with the deposit.  Adjacent lanes of a row of 16 lanes
// whose particles share a base cell (runs of equal keys; the particle arrays
// are kept sorted by cell, so dense regions form long runs) are merged by a
// segmented inclusive scan within the row; the last lane of each run's piece
// in a row then adds its 8 corner sums into a workgroup grid in LDS spanning
// the bounding box of the workgroup's CIC stencils, and the workgroup adds
// each nonzero box cell to the global grid with one atomic.  A run that
// crosses rows leaves one tail per row (at most four LDS atomics on one
// address per instruction); keeping the scan inside rows needs four DPP
// steps that write every lane (bound_ctrl) and stops as soon as every tail
// holds its whole piece, instead of six steps with masked cross-row
// broadcasts.  Particles move less than a cell between the periodic re-sorts,
// so the 256 particles of a workgroup stay in a small box (median 60 cells,
// 99% within 1,711, on the 10^6-particle disk at 128^3, 16 steps after a
// sort) and the LDS atomics absorb the runs that the drift splits; a
// workgroup whose box exceeds kDepositBox cells issues each tail's 8 atomics
// directly.  Integer addition is exact, so the grid does not depend on the
// merging.
constexpr unsigned kNoCell = 0xffffffffu;
constexpr int kDepositBox = 2048;

// v of lane l - n (CTRL row_shr:n = 0x110 + n) or l + n (row_shl:n = 0x100 +
// n) within its row of 16 lanes; lanes whose source lies outside the row read
// 0 (bound_ctrl), so every lane is written and no old value is needed.
template <int CTRL> __device__ __forceinline__ int row_mov(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL> __device__ __forceinline__ long long row_mov64(long long v) {
    return __builtin_amdgcn_update_dpp((long long)0, v, CTRL, 0xf, 0xf, true);
}
// One step of the segmented scan on (q, f), f = "a run head lies in my
// window": the window to the left joins unless a head is already in mine.
// Row starts are heads, so no window reaches across a row.
template <int CTRL> __device__ __forceinline__ void seg_step(long long (&q)[8], int& f) {
    const int fs = row_mov<CTRL>(f);
    long long s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = row_mov64<CTRL>(q[k]);
    if (!f) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] += s[k];
    }
    f |= fs;
}

// Minimum / maximum over the wave by DPP row shifts and row broadcasts, read
// from lane 63.  Lanes without a source read the identity (INT_MAX for min,
// INT_MIN for max), which lets the compiler fold each step into one DPP min /
// max.
template <bool MAX> __device__ __forceinline__ int wave_ext(int v) {
    constexpr int ident = MAX ? int(0x80000000u) : 0x7fffffff;
    auto step = [&](int t) { v = MAX ? max(v, t) : min(v, t); };
    step(__builtin_amdgcn_update_dpp(ident, v, 0x111, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(ident, v, 0x112, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(ident, v, 0x114, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(ident, v, 0x118, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(ident, v, 0x142, 0xa, 0xf, false));
    step(__builtin_amdgcn_update_dpp(ident, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// e / d for 0 <= e < 2^12 and 1 <= d <= 2^12, given rd = 1.f / d: the float
// estimate is within 1e-3 of e / d, so one correction either way makes it
// exact.
__device__ __forceinline__ int small_div(int e, int d, float rd) {
    int q = (int)((float)e * rd);
    const int r = e - __mul24(q, d);
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// The bounding box of a 256-thread workgroup from each wave's (min, max) per
// axis: lane 0 of every wave stores its six values in the wave's own LDS slot
// (no atomics and no initialisation, so one barrier suffices), and after the
// caller's barrier every thread reduces the four slots.
__device__ __forceinline__ void box_slot_store(int (*slot)[6], const int lo[3], const int hi[3]) {
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            slot[w][d] = lo[d];
            slot[w][3 + d] = hi[d];
        }
    }
}
__device__ __forceinline__ void box_slot_reduce(const int (*slot)[6], int b[6]) {
#pragma unroll
    for (int d = 0; d < 6; ++d) b[d] = slot[0][d];
#pragma unroll
    for (int w = 1; w < 4; ++w)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            b[d] = min(b[d], slot[w][d]);
            b[3 + d] = max(b[3 + d], slot[w][3 + d]);
        }
}

template <class real, bool FASTQ>
__global__ __launch_bounds__(256) void k_predict_deposit(PState<real> p, real dt, int do_predict, Geo<real> g,
                                                         double scale, unsigned long long* __restrict__ mfix) {
    __shared__ unsigned long long box[kDepositBox];
    __shared__ int slot[4][6];  // each wave's min i0, j0, k0 and max i1, j1, k1
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t n = p.n;
    const int NX = g.nx;
    unsigned key = kNoCell;  // lanes past n keep it: every thread reaches the barriers below
    long long q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    unsigned c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int base[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, step[3] = {0, 0, 0};
    if (i < n) {
        // every load is issued before the first store: the particle arrays
        // are not declared disjoint, so a load after a store would wait for it
        real r[3], v[3] = {0, 0, 0}, a[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = p.x[d * n + i];
        if (do_predict) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                v[d] = p.v[d * n + i];
                a[d] = p.a[d * n + i];
            }
        }
        const real m = p.m[i];
        if (do_predict) {  // predict_step_ (nbody.cu:135-151)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const real vp = v[d] + dt * a[d];
                p.w[d * n + i] = vp;
                r[d] = (real)((double)r[d] + ((double)dt * 0.5) * (double)(v[d] + vp));
                p.x[d * n + i] = r[d];
            }
        }
        key = cic_fixed<real, FASTQ>(r, m, g, scale, q, c, base, step);
    }
    // the wave's part of the bounding box of the stencils (lanes past n add
    // nothing: base INT_MAX, base + step -1 is below every corner)
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_ext<false>(base[d]);
        hi[d] = wave_ext<true>(i < n ? base[d] + step[d] : -1);
    }
    box_slot_store(slot, lo, hi);
    // run heads and tails within rows of 16 lanes
    const int rl = lane & 15;
    const unsigned kprev = (unsigned)row_mov<0x111>((int)key);
    const unsigned knext = (unsigned)row_mov<0x101>((int)key);
    int f = (rl == 0 || kprev != key) ? 1 : 0;
    const bool tail = (rl == 15 || knext != key);
    // inclusive segmented scan within the row, until every tail holds its run
    // piece (a uniform exit)
    if (__ballot(tail && !f) != 0ull) {
        seg_step<0x111>(q, f);
        if (__ballot(tail && !f) != 0ull) {
            seg_step<0x112>(q, f);
            if (__ballot(tail && !f) != 0ull) {
                seg_step<0x114>(q, f);
                if (__ballot(tail && !f) != 0ull) seg_step<0x118>(q, f);
            }
        }
    }
    __syncthreads();  // the waves' bounds stored
    int bnd[6];
    box_slot_reduce(slot, bnd);
    const int bx = bnd[0], by = bnd[1], bz = bnd[2];
    const int ex = bnd[3] - bx + 1, ey = bnd[4] - by + 1, ez = bnd[5] - bz + 1;
    const bool in_box = bnd[3] >= 0 && ex * ey * ez <= kDepositBox;  // uniform over the workgroup
    const int exy = in_box ? ex * ey : 0, vol = in_box ? exy * ez : 0;
    if (in_box) {
        for (int e = threadIdx.x; e < vol; e += blockDim.x) box[e] = 0ull;
        __syncthreads();  // the box's cells cleared
    }
    if (tail && key != kNoCell) {
        if (in_box) {
            const int di = int(c[1] != c[0]), dj = int(c[2] != c[0]) * ex, dk = int(c[4] != c[0]) * exy;
            const int b0 = (base[0] - bx) + ex * ((base[1] - by) + ey * (base[2] - bz));
#pragma unroll
            for (int k = 0; k < 8; ++k)  // corner k: steps (k & 1, (k >> 1) & 1, k >> 2); adding 0 is harmless
                atomicAdd(&box[b0 + (k & 1) * di + ((k >> 1) & 1) * dj + (k >> 2) * dk], (unsigned long long)q[k]);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (q[k] != 0) atomicAdd(mfix + c[k], (unsigned long long)q[k]);
        }
    }
    if (in_box) {
        __syncthreads();
        const float rxy = 1.f / float(exy), rx = 1.f / float(ex);
        for (int e = threadIdx.x; e < vol; e += blockDim.x) {
            const unsigned long long v = box[e];
            if (v != 0ull) {
                const int lk = small_div(e, exy, rxy), r = e - lk * exy;
                const int lj = small_div(r, ex, rx), li = r - lj * ex;
                atomicAdd(mfix + (unsigned(bx + li) + unsigned(NX) * (unsigned(by + lj) + unsigned(NX) * unsigned(bz + lk))),
                          v);
            }
        }
    }
}

// Sort keys for the spatial ordering of particles: the CIC base cell.
template <class real>
__global__ __launch_bounds__(256) void k_cell_keys(PState<real> p, Geo<real> g, unsigned* __restrict__ keys,
                                                   unsigned* __restrict__ idx) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        real u = (p.x[d * n + i] - g.domain_min) * g.inv_dx - (real)0.5f;
        u = clampr<real>(u, 0, g.upper);
        b[d] = max(0, min((int)floor(u), g.nx - 1));
    }
    keys[i] = unsigned(b[0]) + unsigned(g.nx) * (unsigned(b[1]) + unsigned(g.nx) * unsigned(b[2]));
    idx[i] = unsigned(i);
}

// Gather the particle state through a permutation: dst[i] = src[idx[i]].
template <class real>
__global__ __launch_bounds__(256) void k_permute(PState<real> src, PState<real> dst, const unsigned* __restrict__ idx,
                                                 const int* __restrict__ id_src, int* __restrict__ id_dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= src.n) return;
    const int64_t n = src.n;
    const int64_t j = idx[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dst.x[d * n + i] = src.x[d * n + j];
        dst.v[d * n + i] = src.v[d * n + j];
        dst.a[d * n + i] = src.a[d * n + j];
        dst.w[d * n + i] = src.w[d * n + j];
    }
    dst.m[i] = src.m[j];
    dst.g[i] = src.g[j];
    id_dst[i] = id_src[j];
}

// ------------------------------------------------- non-periodic boundaries
// Division and square root rounded once to `real`: FP32 goes through FP64,
// which rounds to the same float as a correctly rounded FP32 operation
// (53 >= 2*24 + 2 bits), so results match an IEEE host evaluation.
template <class real> __device__ __forceinline__ real div_r(real a, real b) {
    if constexpr (std::is_same_v<real, float>) return (float)((double)a / (double)b);
    else return a / b;
}
template <class real> __device__ __forceinline__ real sqrt_r(real a) {
    if constexpr (std::is_same_v<real, float>) return (float)sqrt((double)a);
    else return sqrt(a);
}

template <class real> __device__ __forceinline__ real coord_of(const WaveModel<real>& w, int i) {  // wave.cu:18-20
    return w.domain_min + (real)i * w.dx;
}

template <class real> __device__ __forceinline__ real layer_diss(const WaveModel<real>& w, real x) {  // wave.cu:27-33
    const real d = fabs(x) - w.x_in;
    return (d > 0) ? w.diss_extra * d * d : real(0);
}

// The scheme coefficients of one cell (wave.cu:155-171); non-periodic grids
// add the layer dissipation of each axis.
template <class real> __device__ __forceinline__ WaveK<real> cell_coef(const WaveModel<real>& w, int ix, int iy, int iz) {
    real diss = w.diss_base;
    diss += layer_diss(w, coord_of(w, ix));
    diss += layer_diss(w, coord_of(w, iy));
    diss += layer_diss(w, coord_of(w, iz));
    const int DIM = 3;
    const real C = w.C, DT = w.DT, DX = w.dx;
    const real Q = diss * C * DT * real(0.5);
    const real W = 1 + Q;
    const real K = div_r<real>(C * C * DT * DT, DX * DX);
    WaveK<real> k;
    k.k_main = div_r<real>(K, W);
    k.k_f = div_r<real>(-C * C * DT * DT, W);
    k.k_actual = div_r<real>(2 * (1 - DIM * K), W);
    k.k_old = div_r<real>(Q - 1, W);
    return k;
}

template <class real> __device__ __forceinline__ real monopole_phi(const WaveModel<real>& w, real x, real y, real z) {
    const real G = 1.;  // wave.cu:36-41
    const real rx = x - w.com[0], ry = y - w.com[1], rz = z - w.com[2];
    const real r = fmax(sqrt_r<real>(rx * rx + ry * ry + rz * rz), w.r_min);
    return div_r<real>(-G * w.mass, r);
}

template <class real>
__device__ __forceinline__ real psi_at(const WaveModel<real>& w, const real* __restrict__ phi, int i, int j, int k) {
    const size_t NX = size_t(w.nx);  // wave.cu:44-49
    return phi[size_t(i) + size_t(j) * NX + size_t(k) * NX * NX] -
           monopole_phi(w, coord_of(w, i), coord_of(w, j), coord_of(w, k));
}

template <class real>
__device__ real radiative_value(const WaveModel<real>& w, const real* __restrict__ phi, int ix, int iy, int iz) {
    const real x = coord_of(w, ix), y = coord_of(w, iy), z = coord_of(w, iz);  // wave.cu:80-104
    const real rx = x - w.com[0], ry = y - w.com[1], rz = z - w.com[2];
    const real r = fmax(sqrt_r<real>(rx * rx + ry * ry + rz * rz), w.r_min);
    const real n3[3] = {div_r<real>(rx, r), div_r<real>(ry, r), div_r<real>(rz, r)};
    const real psi0 = psi_at(w, phi, ix, iy, iz);
    real dpsi_dr = 0;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {  // upwind_axis (wave.cu:53-76), summed in axis order
        const int i0 = axis == 0 ? ix : (axis == 1 ? iy : iz);
        int in = (n3[axis] > 0) ? i0 - 1 : i0 + 1;
        if (in < 0 || in >= w.nx) in = 2 * i0 - in;
        const int jx = axis == 0 ? in : ix, jy = axis == 1 ? in : iy, jz = axis == 2 ? in : iz;
        const real d = div_r<real>(psi0 - psi_at(w, phi, jx, jy, jz), (real)(i0 - in) * w.dx);
        const real term = n3[axis] * d;
        dpsi_dr = axis == 0 ? term : dpsi_dr + term;
    }
    const real psi_new = psi0 - w.C * w.DT * (dpsi_dr + div_r<real>(psi0, r));
    return psi_new + monopole_phi(w, x, y, z);
}

// Fixed point -> cell mass and source term.  rho = m / (DX*DX*DX) - rho_mean
// and f = 4*PI*G*rho in double then rounded to real (wave.cu:150-153); the
// stencil's last term f*K_F is formed here once per deposit.
template <class real>
__device__ __forceinline__ real source_term(real m, real dx3, real rho_mean, int periodic, real k_f) {
    real rho = m / dx3;
    if (periodic) rho -= rho_mean;
    const real f = (real)(4 * kPi * (double)(real)1. * (double)rho);
    return f * k_f;
}

template <class real, bool LAYER>
__global__ __launch_bounds__(256) void k_fixed_to_source(int64_t ncell, unsigned long long* __restrict__ mfix,
                                                         double inv_scale, real dx3, real rho_mean, int periodic,
                                                         real k_f, real* __restrict__ mass, real* __restrict__ src,
                                                         WaveModel<real> w, int* __restrict__ occ,
                                                         real* __restrict__ src_empty) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const long long q = (long long)mfix[c];
    if (q != 0) mfix[c] = 0ull;  // cleared for the next deposit (most cells hold no mass: no store)
    if (!LAYER && occ != nullptr) {
        // one wave = one 64-cell segment (c of lane 0 is a multiple of 64)
        const unsigned long long any = __ballot(q != 0);
        if ((threadIdx.x & 63) == 0) occ[c >> 6] = any != 0ull ? 1 : 0;
        if (c == 0) *src_empty = source_term<real>((real)((double)0ll * inv_scale), dx3, rho_mean, periodic, k_f);
    }
    const real m = (real)((double)q * inv_scale);
    if (mass != nullptr) mass[c] = m;
    real kf = k_f;
    if (LAYER) {
        const int NX = w.nx;
        kf = cell_coef(w, int(c % NX), int((c / NX) % NX), int(c / (int64_t(NX) * NX))).k_f;
    }
    src[c] = source_term<real>(m, dx3, rho_mean, periodic, kf);
}

// k_fixed_to_source with a uniform K_F and four consecutive cells per thread:
// the int64 masses in two 16-byte loads and the four source values in one
// store (ncell a multiple of 4); a 64-cell segment's flag comes from 16 lanes.
template <class real>
__global__ __launch_bounds__(256) void k_fixed_to_source4(int64_t ncell, unsigned long long* __restrict__ mfix,
                                                          double inv_scale, real dx3, real rho_mean, int periodic,
                                                          real k_f, real* __restrict__ mass, real* __restrict__ src,
                                                          int* __restrict__ occ, real* __restrict__ src_empty) {
    const int64_t c = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
    if (c >= ncell) return;  // whole 16-lane segments (ncell is a multiple of 64 where flags are written)
    const ulonglong2 qa = reinterpret_cast<const ulonglong2*>(mfix + c)[0];
    const ulonglong2 qb = reinterpret_cast<const ulonglong2*>(mfix + c)[1];
    const long long q[4] = {(long long)qa.x, (long long)qa.y, (long long)qb.x, (long long)qb.y};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (q[k] != 0) mfix[c + k] = 0ull;  // cleared for the next deposit (the others are already zero)
    if (occ != nullptr) {
        const unsigned long long any = __ballot((q[0] | q[1] | q[2] | q[3]) != 0);
        const int lane = threadIdx.x & 63;
        if ((lane & 15) == 0) occ[c >> 6] = ((any >> lane) & 0xffffull) != 0ull ? 1 : 0;
        if (c == 0) *src_empty = source_term<real>((real)((double)0ll * inv_scale), dx3, rho_mean, periodic, k_f);
    }
    typename Vec4<real>::T out;
    real* o = reinterpret_cast<real*>(&out);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const real m = (real)((double)q[k] * inv_scale);
        if (mass != nullptr) mass[c + k] = m;
        o[k] = source_term<real>(m, dx3, rho_mean, periodic, k_f);
    }
    *reinterpret_cast<typename Vec4<real>::T*>(src + c) = out;
}

template <class real, bool LAYER>
__global__ __launch_bounds__(256) void k_mass_to_source(int64_t ncell, const real* __restrict__ mass, real dx3,
                                                        real rho_mean, int periodic, real k_f, real* __restrict__ src,
                                                        WaveModel<real> w) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    real kf = k_f;
    if (LAYER) {
        const int NX = w.nx;
        kf = cell_coef(w, int(c % NX), int((c / NX) % NX), int(c / (int64_t(NX) * NX))).k_f;
    }
    src[c] = source_term<real>(mass[c], dx3, rho_mean, periodic, kf);
}

// -------------------------------------------------------------- wave stencil
// Thread per cell, workgroup = 64 x 4 cells of one z-plane.  Workgroups are
// renumbered so that each of the 8 XCDs (which receive workgroups round-robin)
// sweeps one contiguous slab of z-planes, keeping the +-z neighbour planes in
// that XCD's L2.
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

// Each thread owns one (x, y) column and marches ZB planes in z, keeping
// phi at z-1, z, z+1 in registers; x/y neighbours come from the cache.  ZB
// trades the z-halo re-reads (ZB+2 planes per ZB) against the workgroup
// count: 2 planes up to 96^3 (enough workgroups to fill 256 CUs), 4 above.
inline int wave_zb(int nx) { return nx <= 96 ? 2 : 4; }

template <class real, bool PERIODIC, bool CHECK, int ZB>
__global__ __launch_bounds__(256) void k_wave(real* __restrict__ pn, const real* __restrict__ p,
                                              const real* __restrict__ po, const real* __restrict__ src, int NX,
                                              WaveK<real> k, typename MaxBits<real>::T* __restrict__ maxbits,
                                              const int* __restrict__ halt) {
    if (halt != nullptr && *halt != 0) return;  // a converged solve's later launches
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    real dmax = 0, amax = 0;
    if (ix < NX && iy < NX) {
        const size_t sy = size_t(NX), sz = size_t(NX) * NX;
        // neighbour indices: wrapped (periodic) or clamped (edge cells are not
        // computed from them, the clamp only keeps the loads in bounds)
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
                out = 0;  // ZeroDirichlet (wave.cu:140-145)
            } else {
                // phi[xm] + phi[xp] + phi[ym] + phi[yp] + phi[zm] + phi[zp] (wave.cu:183-186)
                const real lap_sum = p[xm + iy * sy + pl] + p[xp + iy * sy + pl] + p[ix + ym * sy + pl] +
                                     p[ix + yp * sy + pl] + below + above;
                out = lap_sum * k.k_main + cen * k.k_actual + po[c] * k.k_old + src[c];
            }
            pn[c] = out;
            if (CHECK) {
                dmax = fmax(dmax, fabs(out - cen));
                amax = fmax(amax, fabs(out));
            }
            below = cen;
            cen = above;
        }
    }
    if (CHECK) check_accumulate<real>(dmax, amax, maxbits);
}

// ------------------------------------------- steady check on the device
// solve_wave_equation's SteadyStateChecker (wave.cu:253-287,
// wave_steady_checker.cu:95-143) with its state in a device control block,
// so a solve with checks is a fixed launch sequence that a graph can hold:
// each checking launch (a single iteration, or a fused launch for its last
// level) reduces max|phi_new - phi| and max|phi_new| into the maxbits slots,
// a one-workgroup launch then applies the hold rule; once the
// hold is reached `done` is set and every later stencil launch of the solve
// returns at once, and a last launch moves the two live levels into the
// buffers the full schedule would have left them in.
template <class real>
__global__ __launch_bounds__(kCheckSlots) void k_solve_init(SolveCtl* __restrict__ c,
                                                          typename MaxBits<real>::T* __restrict__ maxbits) {
    const int t = threadIdx.x;
    if (t == 0) {
        c->done = 0;
        c->hold = 0;
        c->iters = 0;
        c->ip = -1;
        c->ic = -1;
    }
    maxbits[t] = 0;
    maxbits[kCheckSlots + t] = 0;
}

// After the checking launch that completes `iters` iterations and leaves
// phi^(n-1), phi^n in buffers (ip, ic): the max over the slots, then
// err = diff / (absu + 1e-30) in real (wave_steady_checker.cu:119-126) and
// SteadyStateChecker::check; the slots are cleared for the next check.
template <class real>
__global__ __launch_bounds__(kCheckSlots) void k_solve_decide(SolveCtl* __restrict__ c,
                                                            typename MaxBits<real>::T* __restrict__ maxbits, real tol,
                                                            int hold_need, int iters, int ip, int ic) {
    using MB = typename MaxBits<real>::T;
    static_assert(sizeof(MB) == sizeof(real), "bit pattern width");
    __shared__ MB sd[kCheckSlots / 64], sa[kCheckSlots / 64];
    if (c->done) return;  // uniform
    const int t = threadIdx.x;
    MB bd = maxbits[t], ba = maxbits[kCheckSlots + t];
    maxbits[t] = 0;
    maxbits[kCheckSlots + t] = 0;
    // non-negative values: the bit patterns order as the values do
    for (int off = 32; off > 0; off >>= 1) {
        const MB od = __shfl_xor(bd, off, 64), oa = __shfl_xor(ba, off, 64);
        bd = od > bd ? od : bd;
        ba = oa > ba ? oa : ba;
    }
    if ((t & 63) == 0) {
        sd[t >> 6] = bd;
        sa[t >> 6] = ba;
    }
    __syncthreads();
    if (t != 0) return;
    for (int w = 1; w < kCheckSlots / 64; ++w) {
        bd = sd[w] > bd ? sd[w] : bd;
        ba = sa[w] > ba ? sa[w] : ba;
    }
    real diff, absu;
    __builtin_memcpy(&diff, &bd, sizeof(real));
    __builtin_memcpy(&absu, &ba, sizeof(real));
    const real EPS = (real)1.E-30;
    const real err = diff / (absu + EPS);
    if ((double)err < (double)tol) ++c->hold;
    else c->hold = 0;
    if (c->hold >= hold_need) {
        c->done = 1;
        c->iters = iters;
        c->ip = ip;
        c->ic = ic;
    }
}

// Moves phi^(n-1), phi^n of a converged solve from buffers (c->ip, c->ic) to
// (ip, ic), where the unconverged schedule ends; each thread reads both
// values of its cells before writing, so overlapping buffer pairs are safe.
template <class real>
__global__ __launch_bounds__(256) void k_solve_fixup(const SolveCtl* __restrict__ c, PhiBufs<real> b, int ip, int ic,
                                                     int64_t ncell) {
    if (!c->done) return;
    const int sp = c->ip, sc = c->ic;
    if (sp == ip && sc == ic) return;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < ncell; i += int64_t(gridDim.x) * 256) {
        const real vp = b.p[sp][i], vc = b.p[sc][i];
        b.p[ip][i] = vp;
        b.p[ic][i] = vc;
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_max_diff(int64_t ncell, const real* __restrict__ a, const real* __restrict__ b,
                                                  typename MaxBits<real>::T* __restrict__ maxbits) {
    real dmax = 0, amax = 0;
    for (int64_t c = int64_t(blockIdx.x) * 256 + threadIdx.x; c < ncell; c += int64_t(gridDim.x) * 256) {
        dmax = fmax(dmax, fabs(a[c] - b[c]));
        amax = fmax(amax, fabs(a[c]));
    }
    check_accumulate<real>(dmax, amax, maxbits);
}

// wave_diss_iteration on a non-periodic grid: per-cell coefficients, face
// values by boundary type BT (RRG_BC_ZERO_DIRICHLET / _MONOPOLE_*).  The
// source term src already carries the cell's own K_F.
template <class real, int BT, bool CHECK, int ZB>
__global__ __launch_bounds__(256) void k_wave_bc(real* __restrict__ pn, const real* __restrict__ p,
                                                 const real* __restrict__ po, const real* __restrict__ src,
                                                 WaveModel<real> w, typename MaxBits<real>::T* __restrict__ maxbits,
                                                 const int* __restrict__ halt) {
    if (halt != nullptr && *halt != 0) return;  // a converged solve's later launches
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int NX = w.nx;
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    real dmax = 0, amax = 0;
    if (ix < NX && iy < NX) {
        const size_t sy = size_t(NX), sz = size_t(NX) * NX;
        const bool edge_xy = ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1;
        const size_t col = size_t(ix) + iy * sy;
        for (int t = 0; t < ZB; ++t) {
            const int iz = bz * ZB + t;
            if (iz >= NX) break;
            const size_t c = col + size_t(iz) * sz;
            const real cen = p[c];
            real out;
            if (edge_xy || iz == 0 || iz == NX - 1) {  // wave.cu:129-147
                if constexpr (BT == RRG_BC_MONOPOLE_DIRICHLET) out = monopole_phi(w, coord_of(w, ix), coord_of(w, iy), coord_of(w, iz));
                else if constexpr (BT == RRG_BC_MONOPOLE_RADIATIVE) out = radiative_value(w, p, ix, iy, iz);
                else out = 0;
            } else {
                const WaveK<real> k = cell_coef(w, ix, iy, iz);
                const real lap_sum = p[c - 1] + p[c + 1] + p[c - sy] + p[c + sy] + p[c - sz] + p[c + sz];
                out = lap_sum * k.k_main + cen * k.k_actual + po[c] * k.k_old + src[c];
            }
            pn[c] = out;
            if (CHECK) {
                dmax = fmax(dmax, fabs(out - cen));
                amax = fmax(amax, fabs(out));
            }
        }
    }
    if (CHECK) check_accumulate<real>(dmax, amax, maxbits);
}

// Per-cell coefficient grids of a non-periodic grid (transposed stencil).
template <class real>
__global__ __launch_bounds__(256) void k_cell_coefs(WaveModel<real> w, real* __restrict__ km, real* __restrict__ ka,
                                                    real* __restrict__ ko, real* __restrict__ alpha) {
    const int NX = w.nx;
    const int64_t ncell = int64_t(NX) * NX * NX;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int ix = int(c % NX), iy = int((c / NX) % NX), iz = int(c / (int64_t(NX) * NX));
    const WaveK<real> k = cell_coef(w, ix, iy, iz);
    km[c] = k.k_main;
    ka[c] = k.k_actual;
    ko[c] = k.k_old;
    alpha[c] = (real)(4 * kPi * (double)k.k_f / ((double)w.dx * (double)w.dx * (double)w.dx));
}

// Centre of mass in fixed point: pass 1 max|m x| (non-negative doubles
// order like their bit patterns), pass 2 the int64 moment and mass sums.
template <class real>
__global__ __launch_bounds__(256) void k_com_max(PState<real> p, unsigned long long* __restrict__ maxbits) {
    const int64_t n = p.n;
    double mx = 0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        for (int d = 0; d < 3; ++d) mx = fmax(mx, fabs((double)p.m[i] * (double)p.x[d * n + i]));
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(mx));
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <class real>
__global__ __launch_bounds__(256) void k_com_sum(PState<real> p, double scale_x, double scale_m,
                                                 unsigned long long* __restrict__ sums) {
    const int64_t n = p.n;
    long long q[4] = {0, 0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double m = (double)p.m[i];
        for (int d = 0; d < 3; ++d) q[d] += (long long)rint(m * (double)p.x[d * n + i] * scale_x);
        q[3] += (long long)rint(m * scale_m);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long v = wave_sum_i64(q[k]);
        if ((threadIdx.x & 63) == 0) atomicAdd(sums + k, (unsigned long long)v);
    }
}

// ----------------------------------------------------------------- gradient
// Each thread marches ZB planes of one (x, y) column, keeping phi at z-1, z
// and z+1 in registers; the x and y neighbours come from the cache.  Every
// cell uses calc_acceleration_field_'s differences (one-sided on the faces).
template <class real, int ZB>
__global__ __launch_bounds__(256) void k_gradient(int NX, real coef, const real* __restrict__ phi,
                                                  typename Vec4<real>::T* __restrict__ acc) {
    int bx, by, bz;
    xcd_block(gridDim.x, gridDim.y, gridDim.z, bx, by, bz);
    const int ix = bx * 64 + threadIdx.x;
    const int iy = by * 4 + threadIdx.y;
    const int z0 = bz * ZB;
    if (ix >= NX || iy >= NX || z0 >= NX) return;
    const size_t sy = size_t(NX), sz = size_t(NX) * NX;
    const size_t col = size_t(ix) + iy * sy;
    real pm = z0 > 0 ? phi[col + (z0 - 1) * sz] : real(0);
    real pc = phi[col + z0 * sz];
#pragma unroll 2
    for (int t = 0; t < ZB; ++t) {
        const int iz = z0 + t;
        if (iz >= NX) break;
        const size_t c = col + iz * sz;
        const real pp = iz + 1 < NX ? phi[c + sz] : real(0);
        real d[3];
        if (ix == 0) d[0] = phi[c + 1] - pc;
        else if (ix == NX - 1) d[0] = pc - phi[c - 1];
        else d[0] = (real)(0.5 * (double)(phi[c + 1] - phi[c - 1]));
        if (iy == 0) d[1] = phi[c + sy] - pc;
        else if (iy == NX - 1) d[1] = pc - phi[c - sy];
        else d[1] = (real)(0.5 * (double)(phi[c + sy] - phi[c - sy]));
        if (iz == 0) d[2] = pp - pc;
        else if (iz == NX - 1) d[2] = pc - pm;
        else d[2] = (real)(0.5 * (double)(pp - pm));
        typename Vec4<real>::T o;
        o.x = d[0] * coef;
        o.y = d[1] * coef;
        o.z = d[2] * coef;
        o.w = 0;
        acc[c] = o;
        pm = pc;
        pc = pp;
    }
}

// ------------------------------------------------------------ gather + kick
template <class real, bool KICK>
__global__ __launch_bounds__(256) void k_gather(PState<real> p, Geo<real> g, const typename Vec4<real>::T* __restrict__ F,
                                                real dt) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    const int64_t n = p.n;
    const int nx = g.nx;
    const real dx = g.dx, domain_min = g.domain_min;
    const real px = p.x[q], py = p.x[n + q], pz = p.x[2 * n + q];
    // the kick's inputs are loaded before the first store (the particle
    // arrays are not declared disjoint, so a later load would wait for it)
    real v[3] = {0, 0, 0}, vv[3] = {0, 0, 0};
    if (KICK) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = p.v[c * n + q];
            vv[c] = p.w[c * n + q];
        }
    }
    const real fx = (px - domain_min) / dx - (real)0.5f;
    const real fy = (py - domain_min) / dx - (real)0.5f;
    const real fz = (pz - domain_min) / dx - (real)0.5f;
    int i0 = (int)floorf((float)fx);
    int j0 = (int)floorf((float)fy);
    int k0 = (int)floorf((float)fz);
    int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
    const real wx = fx - (real)i0, wy = fy - (real)j0, wz = fz - (real)k0;
    i0 = max(0, min(i0, nx - 1));
    i1 = max(0, min(i1, nx - 1));
    j0 = max(0, min(j0, nx - 1));
    j1 = max(0, min(j1, nx - 1));
    k0 = max(0, min(k0, nx - 1));
    k1 = max(0, min(k1, nx - 1));
    real ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int k = 0; k <= 1; k++) {
        const int kk = (k == 0) ? k0 : k1;
        const real wk = (k == 0) ? ((real)1.0f - wz) : wz;
#pragma unroll
        for (int j = 0; j <= 1; j++) {
            const int jj = (j == 0) ? j0 : j1;
            const real wj = (j == 0) ? ((real)1.0f - wy) : wy;
#pragma unroll
            for (int i = 0; i <= 1; i++) {
                const int ii = (i == 0) ? i0 : i1;
                const real wi = (i == 0) ? ((real)1.0f - wx) : wx;
                const real weight = wi * wj * wk;
                const typename Vec4<real>::T f = F[kk * nx * nx + jj * nx + ii];
                ax += weight * f.x;
                ay += weight * f.y;
                az += weight * f.z;
            }
        }
    }
    p.a[q] = ax;
    p.a[n + q] = ay;
    p.a[2 * n + q] = az;
    if (KICK) {  // correct_step_ (nbody.cu:171-180)
        const real a3[3] = {ax, ay, az};
#pragma unroll
        for (int c = 0; c < 3; ++c) p.v[c * n + q] = (real)(0.5 * (double)(v[c] + vv[c] + dt * a3[c]));
    }
}

// -grad phi of cell (ix, iy, iz) as calc_acceleration_field_ computes it
// (rm_p2mesh.cu:238-290, and k_gradient above): one-sided differences on the
// faces, half the central difference inside (through double, as the
// reference's 0.5 literal promotes it), times coef = -1/dx.
template <class real>
__device__ __forceinline__ void cell_grad(const real* __restrict__ phi, int NX, int ix, int iy, int iz, real coef,
                                          real g[3]) {
    const size_t sy = size_t(NX), sz = size_t(NX) * NX;
    const size_t c = size_t(ix) + iy * sy + iz * sz;
    const real pc = phi[c];
    real d[3];
    if (ix == 0) d[0] = phi[c + 1] - pc;
    else if (ix == NX - 1) d[0] = pc - phi[c - 1];
    else d[0] = (real)(0.5 * (double)(phi[c + 1] - phi[c - 1]));
    if (iy == 0) d[1] = phi[c + sy] - pc;
    else if (iy == NX - 1) d[1] = pc - phi[c - sy];
    else d[1] = (real)(0.5 * (double)(phi[c + sy] - phi[c - sy]));
    if (iz == 0) d[2] = phi[c + sz] - pc;
    else if (iz == NX - 1) d[2] = pc - phi[c - sz];
    else d[2] = (real)(0.5 * (double)(phi[c + sz] - phi[c - sz]));
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = d[k] * coef;
}

// The gather of k_gather with -grad phi computed from phi on the way, so a
// step writes no gradient grid.  A workgroup of 256 particles (kept sorted by
// cell) reduces the bounding box of the cells its particles interpolate from;
// if the box holds at most kGatherBox cells, its threads compute each box
// cell's -grad phi once into LDS and every particle reads its eight corners
// from there; otherwise (sparse regions) each particle computes its corners'
// -grad phi itself.  Either way each corner value is cell_grad's, the value
// k_gradient writes, and the interpolation keeps k_gather's order, so the
// result is bitwise that of k_gradient followed by k_gather.
template <class real> constexpr int kGatherBox = 16384 / int(sizeof(typename Vec4<real>::T));

template <class real, bool KICK>
__global__ __launch_bounds__(256) void k_gather_phi(PState<real> p, Geo<real> g, const real* __restrict__ phi,
                                                    real coef, real dt) {
    using V4 = typename Vec4<real>::T;
    __shared__ V4 box[kGatherBox<real>];
    __shared__ int slot[4][6];  // each wave's min and max corner index per axis
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = p.n;
    const int nx = g.nx;
    const bool live = q < n;  // lanes past n reach every barrier and store nothing
    const int64_t qq = live ? q : n - 1;
    const real dx = g.dx, domain_min = g.domain_min;
    const real px = p.x[qq], py = p.x[n + qq], pz = p.x[2 * n + qq];
    // the kick's inputs are loaded before the first store (the particle
    // arrays are not declared disjoint, so a later load would wait for it)
    real v[3] = {0, 0, 0}, vv[3] = {0, 0, 0};
    if (KICK) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = p.v[c * n + qq];
            vv[c] = p.w[c * n + qq];
        }
    }
    // interpolate_acceleration_to_particles_kernel (rm_mesh2p.cu:60-111)
    const real fx = (px - domain_min) / dx - (real)0.5f;
    const real fy = (py - domain_min) / dx - (real)0.5f;
    const real fz = (pz - domain_min) / dx - (real)0.5f;
    int i0 = (int)floorf((float)fx);
    int j0 = (int)floorf((float)fy);
    int k0 = (int)floorf((float)fz);
    int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
    const real wx = fx - (real)i0, wy = fy - (real)j0, wz = fz - (real)k0;
    i0 = max(0, min(i0, nx - 1));
    i1 = max(0, min(i1, nx - 1));
    j0 = max(0, min(j0, nx - 1));
    j1 = max(0, min(j1, nx - 1));
    k0 = max(0, min(k0, nx - 1));
    k1 = max(0, min(k1, nx - 1));
    int bnd[6];
    {
        const int lo[3] = {wave_ext<false>(i0), wave_ext<false>(j0), wave_ext<false>(k0)};
        const int hi[3] = {wave_ext<true>(i1), wave_ext<true>(j1), wave_ext<true>(k1)};
        box_slot_store(slot, lo, hi);
        __syncthreads();  // the waves' bounds stored
        box_slot_reduce(slot, bnd);
    }
    const int bx = bnd[0], by = bnd[1], bz = bnd[2];
    const int ex = bnd[3] - bx + 1, ey = bnd[4] - by + 1, ez = bnd[5] - bz + 1;
    const bool in_box = ex * ey * ez <= kGatherBox<real>;  // uniform over the workgroup (every box cell is a grid cell)
    if (in_box) {
        const int exy = ex * ey, vol = exy * ez;
        const float rxy = 1.f / float(exy), rx = 1.f / float(ex);
        for (int e = threadIdx.x; e < vol; e += blockDim.x) {
            const int lk = small_div(e, exy, rxy), r = e - lk * exy;
            const int lj = small_div(r, ex, rx), li = r - lj * ex;
            real gc[3];
            cell_grad<real>(phi, nx, bx + li, by + lj, bz + lk, coef, gc);
            V4 o;
            o.x = gc[0];
            o.y = gc[1];
            o.z = gc[2];
            o.w = 0;
            box[e] = o;
        }
        __syncthreads();  // box filled
    }
    if (!live) return;
    real ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int k = 0; k <= 1; k++) {
        const int kk = (k == 0) ? k0 : k1;
        const real wk = (k == 0) ? ((real)1.0f - wz) : wz;
#pragma unroll
        for (int j = 0; j <= 1; j++) {
            const int jj = (j == 0) ? j0 : j1;
            const real wj = (j == 0) ? ((real)1.0f - wy) : wy;
#pragma unroll
            for (int i = 0; i <= 1; i++) {
                const int ii = (i == 0) ? i0 : i1;
                const real wi = (i == 0) ? ((real)1.0f - wx) : wx;
                const real weight = wi * wj * wk;
                real f[3];
                if (in_box) {
                    const V4 o = box[(ii - bx) + ex * ((jj - by) + ey * (kk - bz))];
                    f[0] = o.x;
                    f[1] = o.y;
                    f[2] = o.z;
                } else {
                    cell_grad<real>(phi, nx, ii, jj, kk, coef, f);
                }
                ax += weight * f[0];
                ay += weight * f[1];
                az += weight * f[2];
            }
        }
    }
    p.a[q] = ax;
    p.a[n + q] = ay;
    p.a[2 * n + q] = az;
    if (KICK) {  // correct_step_ (nbody.cu:171-180)
        const real a3[3] = {ax, ay, az};
#pragma unroll
        for (int c = 0; c < 3; ++c) p.v[c * n + q] = (real)(0.5 * (double)(v[c] + vv[c] + dt * a3[c]));
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_ngp_phi(PState<real> p, Geo<real> g, const real* __restrict__ phi) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    const int64_t n = p.n;
    const int NX = g.nx;
    const int ix = clampr<int>((int)((p.x[q] - g.domain_min) / g.dx), 0, NX - 1);
    const int iy = clampr<int>((int)((p.x[n + q] - g.domain_min) / g.dx), 0, NX - 1);
    const int iz = clampr<int>((int)((p.x[2 * n + q] - g.domain_min) / g.dx), 0, NX - 1);
    p.g[q] = phi[size_t(ix) + size_t(iy) * NX + size_t(iz) * NX * NX];
}

// NGP binning (assign_particles_to_cells_, rm_p2mesh.cu:25-66): the cell of
// each particle, and the per-cell counts (exact integer atomics).
template <class real>
__global__ __launch_bounds__(256) void k_ngp_keys(PState<real> p, Geo<real> g, unsigned* __restrict__ keys,
                                                  unsigned* __restrict__ idx, int* __restrict__ count) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    const int64_t n = p.n;
    const int NX = g.nx;
    const int ix = clampr<int>((int)((p.x[q] - g.domain_min) / g.dx), 0, NX - 1);
    const int iy = clampr<int>((int)((p.x[n + q] - g.domain_min) / g.dx), 0, NX - 1);
    const int iz = clampr<int>((int)((p.x[2 * n + q] - g.domain_min) / g.dx), 0, NX - 1);
    const unsigned cell = unsigned(ix) + unsigned(NX) * (unsigned(iy) + unsigned(NX) * unsigned(iz));
    keys[q] = cell;
    idx[q] = unsigned(q);
    atomicAdd(count + cell, 1);
}

// Cell averages of the particle potential and acceleration
// (deposit_cell_grav_NGP_ / deposit_cell_acc_NGP_, rm_p2mesh.cu:139-168):
// one thread per cell sums its particles in index order (the stable sort
// keeps that order), so the result is deterministic.
template <class real>
__global__ __launch_bounds__(256) void k_ngp_project(int64_t ncell, const int* __restrict__ count,
                                                     const int* __restrict__ start, const unsigned* __restrict__ order,
                                                     int64_t n, const real* __restrict__ grav,
                                                     const real* __restrict__ acc, real* __restrict__ cell_phi,
                                                     real* __restrict__ cell_acc) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int cnt = count[c];
    real ph = 0, ax = 0, ay = 0, az = 0;
    if (cnt > 0) {
        const real inv_count = (real)(1. / (double)cnt);
        for (int k = start[c]; k < start[c] + cnt; ++k) {
            const int64_t q = order[k];
            ph += grav[q] / (real)cnt;
            ax += acc[q] * inv_count;
            ay += acc[n + q] * inv_count;
            az += acc[2 * n + q] * inv_count;
        }
    }
    cell_phi[c] = ph;
    cell_acc[3 * c + 0] = ax;
    cell_acc[3 * c + 1] = ay;
    cell_acc[3 * c + 2] = az;
}

template <class real>
__global__ __launch_bounds__(256) void k_fixed_to_mass(int64_t ncell, unsigned long long* __restrict__ mfix,
                                                       double inv_scale, real* __restrict__ mass) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const long long q = (long long)mfix[c];
    mfix[c] = 0ull;
    mass[c] = (real)((double)q * inv_scale);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t count, unsigned long long value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

inline unsigned grid1d(int64_t n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

}  // namespace

template <class real> Geo<real> make_geo(int nx, double dx_in, double bc_frac) {
    (void)bc_frac;
    Geo<real> g;
    g.nx = nx;
    g.dx = (real)dx_in;
    const real domain_l = g.dx * nx;             // io_json.cpp:32
    g.domain_min = (real)(-0.5 * domain_l);      // io_json.cpp:39
    g.inv_dx = (real)(1.f / g.dx);               // rm_p2mesh.cu:185
    const real eps = (real)1E-6;                 // rm_p2mesh.cu:196
    g.upper = (real)(nx - 1) - eps;
    return g;
}

template <class real> WaveK<real> wave_coef(real diss, real C, real DT, real DX) {
    const int DIM = 3;  // wave.cu:164-171
    const real Q = diss * C * DT * real(0.5);
    const real W = 1 + Q;
    const real K = C * C * DT * DT / (DX * DX);
    WaveK<real> k;
    k.k_main = K / W;
    k.k_f = -C * C * DT * DT / W;
    k.k_actual = 2 * (1 - DIM * K) / W;
    k.k_old = (Q - 1) / W;
    return k;
}

template <class real>
void predict_deposit(hipStream_t s, const PState<real>& p, real dt, bool predict, const Geo<real>& g, double scale,
                     unsigned long long* mfix, double max_q) {
    if (p.n == 0) return;
    if (max_q < 2251799813685248.0)  // 2^51
        k_predict_deposit<real, true><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, predict ? 1 : 0, g, scale, mfix);
    else
        k_predict_deposit<real, false><<<grid1d(p.n, 256), 256, 0, s>>>(p, dt, predict ? 1 : 0, g, scale, mfix);
    RRG_LAUNCH_CHECK();
}

template <class real>
void cell_keys(hipStream_t s, const PState<real>& p, const Geo<real>& g, unsigned* keys, unsigned* idx) {
    if (p.n == 0) return;
    k_cell_keys<real><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, keys, idx);
    RRG_LAUNCH_CHECK();
}

template <class real>
void permute(hipStream_t s, const PState<real>& src, const PState<real>& dst, const unsigned* idx, const int* id_src,
             int* id_dst) {
    if (src.n == 0) return;
    k_permute<real><<<grid1d(src.n, 256), 256, 0, s>>>(src, dst, idx, id_src, id_dst);
    RRG_LAUNCH_CHECK();
}

template <class real>
WaveModel<real> make_model(int nx, double dx_in, double bc_frac_in, double diss_base, double diss_extra,
                           double wave_speed, double wave_dt, int bc) {
    // GridInfo from params (io_json.cpp:28-40), in `real`
    WaveModel<real> w{};
    w.nx = nx;
    w.bc = bc;
    w.dx = (real)dx_in;
    const real domain_l = w.dx * nx;
    const real bc_frac = (real)bc_frac_in;
    const real sim_l = (real)((1. - bc_frac) * domain_l);
    w.domain_min = (real)(-0.5 * domain_l);
    w.x_in = real(0.5) * sim_l;
    w.diss_base = (real)diss_base;
    w.diss_extra = (real)diss_extra;
    w.C = (real)wave_speed;
    w.DT = (real)wave_dt;
    w.mass = 0;
    w.r_min = real(0.5) * w.dx;
    w.com[0] = w.com[1] = w.com[2] = 0;
    return w;
}

template <class real>
void center_of_mass(hipStream_t s, const PState<real>& p, int e_mass, unsigned long long* scratch, real out[3]) {
    out[0] = out[1] = out[2] = 0;
    if (p.n == 0) return;
    const unsigned nb = unsigned(std::min<int64_t>(1024, (p.n + 255) / 256));
    k_fill_u64<<<1, 64, 0, s>>>(scratch, 5, 0ull);
    k_com_max<real><<<nb, 256, 0, s>>>(p, scratch);
    RRG_LAUNCH_CHECK();
    unsigned long long h[5];
    RRG_HIP(hipMemcpyAsync(h, scratch, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
    double mx;
    std::memcpy(&mx, &h[0], sizeof(mx));
    int ex = 0;
    std::frexp(std::max(mx * (double)p.n, 1e-300), &ex);
    const int E = std::min(61 - ex, 1000);
    k_com_sum<real><<<nb, 256, 0, s>>>(p, std::ldexp(1.0, E), std::ldexp(1.0, e_mass), scratch + 1);
    RRG_LAUNCH_CHECK();
    RRG_HIP(hipMemcpyAsync(h + 1, scratch + 1, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
    const double M = (double)(long long)h[4] * std::ldexp(1.0, -e_mass);
    if (M > 0)
        for (int d = 0; d < 3; ++d) out[d] = (real)((double)(long long)h[1 + d] * std::ldexp(1.0, -E) / M);
}

template <class real>
void fixed_to_source(hipStream_t s, int nx, unsigned long long* mfix, double inv_scale, real dx, real rho_mean,
                     bool periodic, real k_f, real* mass, real* src, const WaveModel<real>* model, int* occ,
                     real* src_empty) {
    const int64_t ncell = int64_t(nx) * nx * nx;
    const real dx3 = dx * dx * dx;
    const WaveModel<real> w = model ? *model : WaveModel<real>{};
    if (occ != nullptr && (model != nullptr || ncell % 64 != 0 || src_empty == nullptr))
        fail(RRG_ERR_INVALID, "fixed_to_source: segment flags need a uniform K_F, nx^3 a multiple of 64 and src_empty");
    if (model) k_fixed_to_source<real, true><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, mfix, inv_scale, dx3, rho_mean,
                                                                               periodic ? 1 : 0, k_f, mass, src, w,
                                                                               nullptr, nullptr);
    else if (ncell % 4 == 0)
        k_fixed_to_source4<real><<<grid1d(ncell / 4, 256), 256, 0, s>>>(ncell, mfix, inv_scale, dx3, rho_mean,
                                                                       periodic ? 1 : 0, k_f, mass, src, occ,
                                                                       src_empty);
    else k_fixed_to_source<real, false><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, mfix, inv_scale, dx3, rho_mean,
                                                                         periodic ? 1 : 0, k_f, mass, src, w, occ,
                                                                         src_empty);
    RRG_LAUNCH_CHECK();
}

template <class real>
void mass_to_source(hipStream_t s, int nx, const real* mass, real dx, real rho_mean, bool periodic, real k_f,
                    real* src, const WaveModel<real>* model) {
    const int64_t ncell = int64_t(nx) * nx * nx;
    const real dx3 = dx * dx * dx;
    const WaveModel<real> w = model ? *model : WaveModel<real>{};
    if (model) k_mass_to_source<real, true><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, mass, dx3, rho_mean,
                                                                             periodic ? 1 : 0, k_f, src, w);
    else k_mass_to_source<real, false><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, mass, dx3, rho_mean,
                                                                        periodic ? 1 : 0, k_f, src, w);
    RRG_LAUNCH_CHECK();
}

template <class real>
void wave_iteration(hipStream_t s, const WaveModel<real>& m, const WaveK<real>& k, real* pn, const real* p,
                    const real* po, const real* src, void* maxbits, const int* halt) {
    using MB = typename MaxBits<real>::T;
    MB* mb = reinterpret_cast<MB*>(maxbits);
    const int nx = m.nx;
    const int zb = wave_zb(nx);
    const dim3 grid(unsigned((nx + 63) / 64), unsigned((nx + 3) / 4), unsigned((nx + zb - 1) / zb));
    const dim3 block(64, 4, 1);
    if (m.bc == RRG_BC_PERIODIC) {
        auto go = [&](auto kern) { kern<<<grid, block, 0, s>>>(pn, p, po, src, nx, k, mb, halt); };
        if (zb == 2) mb ? go(k_wave<real, true, true, 2>) : go(k_wave<real, true, false, 2>);
        else mb ? go(k_wave<real, true, true, 4>) : go(k_wave<real, true, false, 4>);
    } else {
        auto go = [&](auto kern) { kern<<<grid, block, 0, s>>>(pn, p, po, src, m, mb, halt); };
        // face-heavy small grids use the same z depth rule
        if (m.bc == RRG_BC_MONOPOLE_DIRICHLET) {
            if (zb == 2) mb ? go(k_wave_bc<real, RRG_BC_MONOPOLE_DIRICHLET, true, 2>) : go(k_wave_bc<real, RRG_BC_MONOPOLE_DIRICHLET, false, 2>);
            else mb ? go(k_wave_bc<real, RRG_BC_MONOPOLE_DIRICHLET, true, 4>) : go(k_wave_bc<real, RRG_BC_MONOPOLE_DIRICHLET, false, 4>);
        } else if (m.bc == RRG_BC_MONOPOLE_RADIATIVE) {
            if (zb == 2) mb ? go(k_wave_bc<real, RRG_BC_MONOPOLE_RADIATIVE, true, 2>) : go(k_wave_bc<real, RRG_BC_MONOPOLE_RADIATIVE, false, 2>);
            else mb ? go(k_wave_bc<real, RRG_BC_MONOPOLE_RADIATIVE, true, 4>) : go(k_wave_bc<real, RRG_BC_MONOPOLE_RADIATIVE, false, 4>);
        } else {
            if (zb == 2) mb ? go(k_wave_bc<real, RRG_BC_ZERO_DIRICHLET, true, 2>) : go(k_wave_bc<real, RRG_BC_ZERO_DIRICHLET, false, 2>);
            else mb ? go(k_wave_bc<real, RRG_BC_ZERO_DIRICHLET, true, 4>) : go(k_wave_bc<real, RRG_BC_ZERO_DIRICHLET, false, 4>);
        }
    }
    RRG_LAUNCH_CHECK();
}

template <class real> int wave_fused_levels(const WaveModel<real>& m, int remaining) {
    return fused_levels<real>(m.nx, m.bc, remaining);
}

template <class real> bool wave_fused_pairs(const WaveModel<real>& m, int levels) {
    return fused_pairs<real>(m.nx, m.bc, levels);
}

template <class real>
void wave_fused(hipStream_t s, const WaveModel<real>& m, const WaveK<real>& k, int levels, real* qa, real* qb,
                const real* p, const real* po, const real* src, const int* halt, const int* occ,
                const real* src_empty, bool pin, bool pout, void* maxbits) {
    if (levels < 2 || fused_levels<real>(m.nx, m.bc, levels) != levels)
        fail(RRG_ERR_INVALID, "wave_fused: no launch shape for this grid and iteration count");
    if ((occ == nullptr) != (src_empty == nullptr)) fail(RRG_ERR_INVALID, "wave_fused: occ and src_empty go together");
    if ((pin || pout) && (!fused_pairs<real>(m.nx, m.bc, levels) || occ != nullptr))
        fail(RRG_ERR_INVALID, "wave_fused: no pair-grid launch shape for this grid and iteration count");
    if (maxbits != nullptr && (pin || pout)) fail(RRG_ERR_INVALID, "wave_fused: a checking launch uses split grids");
    if (maxbits != nullptr)
        launch_fused<real, false, false, true>(s, m.nx, levels, qa, qb, p, po, src, nullptr, k, halt, occ, src_empty,
                                               false, false, reinterpret_cast<typename MaxBits<real>::T*>(maxbits));
    else
        launch_fused<real, false, false>(s, m.nx, levels, qa, pout ? qa : qb, p, pin ? p : po, src, nullptr, k, halt,
                                         occ, src_empty, pin, pout);
    RRG_LAUNCH_CHECK();
}

template <class real>
void cell_coefficients(hipStream_t s, const WaveModel<real>& m, real* km, real* ka, real* ko, real* alpha) {
    const int64_t ncell = int64_t(m.nx) * m.nx * m.nx;
    k_cell_coefs<real><<<grid1d(ncell, 256), 256, 0, s>>>(m, km, ka, ko, alpha);
    RRG_LAUNCH_CHECK();
}

template <class real> void gradient(hipStream_t s, int nx, real dx, const real* phi, real* acc4) {
    const real coef = (real)(-1. / (double)dx);  // rm_p2mesh.cu:284
    auto* out = reinterpret_cast<typename Vec4<real>::T*>(acc4);
    const dim3 blk(64, 4, 1);
    if (nx >= 128) {
        const dim3 grid(unsigned((nx + 63) / 64), unsigned((nx + 3) / 4), unsigned((nx + 7) / 8));
        k_gradient<real, 8><<<grid, blk, 0, s>>>(nx, coef, phi, out);
    } else {
        const dim3 grid(unsigned((nx + 63) / 64), unsigned((nx + 3) / 4), unsigned((nx + 1) / 2));
        k_gradient<real, 2><<<grid, blk, 0, s>>>(nx, coef, phi, out);
    }
    RRG_LAUNCH_CHECK();
}

template <class real>
void gather(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* acc4, real dt, bool kick) {
    if (p.n == 0) return;
    auto F = reinterpret_cast<const typename Vec4<real>::T*>(acc4);
    if (kick) k_gather<real, true><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, F, dt);
    else k_gather<real, false><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, F, dt);
    RRG_LAUNCH_CHECK();
}

template <class real>
void gather_phi(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* phi, real dt, bool kick) {
    if (p.n == 0) return;
    const real coef = (real)(-1. / (double)g.dx);  // rm_p2mesh.cu:284
    if (kick) k_gather_phi<real, true><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, phi, coef, dt);
    else k_gather_phi<real, false><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, phi, coef, dt);
    RRG_LAUNCH_CHECK();
}

template <class real> void ngp_phi(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* phi) {
    if (p.n == 0) return;
    k_ngp_phi<real><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, phi);
    RRG_LAUNCH_CHECK();
}

template <class real> void max_diff(hipStream_t s, int64_t ncell, const real* a, const real* b, void* maxbits) {
    using MB = typename MaxBits<real>::T;
    const unsigned nb = unsigned(std::min<int64_t>(4096, (ncell + 255) / 256));
    k_max_diff<real><<<nb, 256, 0, s>>>(ncell, a, b, reinterpret_cast<MB*>(maxbits));
    RRG_LAUNCH_CHECK();
}

template <class real> void solve_init(hipStream_t s, SolveCtl* c, void* maxbits) {
    k_solve_init<real><<<1, kCheckSlots, 0, s>>>(c, reinterpret_cast<typename MaxBits<real>::T*>(maxbits));
    RRG_LAUNCH_CHECK();
}

template <class real>
void solve_decide(hipStream_t s, SolveCtl* c, void* maxbits, real tol, int hold_need, int iters, int ip, int ic) {
    k_solve_decide<real><<<1, kCheckSlots, 0, s>>>(c, reinterpret_cast<typename MaxBits<real>::T*>(maxbits), tol,
                                                   hold_need, iters, ip, ic);
    RRG_LAUNCH_CHECK();
}

template <class real>
void solve_fixup(hipStream_t s, const SolveCtl* c, const PhiBufs<real>& b, int ip, int ic, int64_t ncell) {
    const unsigned nb = unsigned(std::min<int64_t>(2048, (ncell + 255) / 256));
    k_solve_fixup<real><<<nb, 256, 0, s>>>(c, b, ip, ic, ncell);
    RRG_LAUNCH_CHECK();
}

void fill_u64(hipStream_t s, unsigned long long* p, size_t count, unsigned long long value) {
    if (count == 0) return;
    const unsigned nb = unsigned(std::min<size_t>(4096, (count + 255) / 256));
    k_fill_u64<<<nb, 256, 0, s>>>(p, count, value);
    RRG_LAUNCH_CHECK();
}

template <class real>
void ngp_bin(hipStream_t s, const PState<real>& p, const Geo<real>& g, NgpBins& b) {
    const int64_t ncell = int64_t(g.nx) * g.nx * g.nx;
    RRG_HIP(hipMemsetAsync(b.count, 0, sizeof(int) * ncell, s));
    if (p.n > 0) {
        k_ngp_keys<real><<<grid1d(p.n, 256), 256, 0, s>>>(p, g, b.keys, b.idx, b.count);
        RRG_LAUNCH_CHECK();
        int end_bit = 1;
        while ((int64_t(1) << end_bit) < ncell) ++end_bit;
        size_t bytes = b.temp_bytes;
        radix_sort_pairs(s, b.temp, bytes, b.keys, b.keys2, b.idx, b.order, p.n, end_bit);
    }
    size_t bytes = b.temp_bytes;
    exclusive_scan_i32(s, b.temp, bytes, b.count, b.start, ncell);
}

template <class real>
void ngp_project(hipStream_t s, int nx, const NgpBins& b, int64_t n, const real* grav, const real* acc, real* cell_phi,
                 real* cell_acc) {
    const int64_t ncell = int64_t(nx) * nx * nx;
    k_ngp_project<real><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, b.count, b.start, b.order, n, grav, acc, cell_phi,
                                                          cell_acc);
    RRG_LAUNCH_CHECK();
}

template <class real>
void fixed_to_mass(hipStream_t s, int nx, unsigned long long* mfix, double inv_scale, real* mass) {
    const int64_t ncell = int64_t(nx) * nx * nx;
    k_fixed_to_mass<real><<<grid1d(ncell, 256), 256, 0, s>>>(ncell, mfix, inv_scale, mass);
    RRG_LAUNCH_CHECK();
}

size_t ngp_temp_bytes(int64_t n, int nx) {
    const int64_t ncell = int64_t(nx) * nx * nx;
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < ncell) ++end_bit;
    size_t a = 0, b = 0;
    radix_sort_pairs(nullptr, nullptr, a, nullptr, nullptr, nullptr, nullptr, std::max<int64_t>(n, 1), end_bit);
    exclusive_scan_i32(nullptr, nullptr, b, nullptr, nullptr, ncell);
    return std::max<size_t>(std::max(a, b), 1);
}

#define RRG_PM_INST(real)                                                                                       \
    template Geo<real> make_geo<real>(int, double, double);                                                     \
    template WaveK<real> wave_coef<real>(real, real, real, real);                                               \
    template void predict_deposit<real>(hipStream_t, const PState<real>&, real, bool, const Geo<real>&, double, \
                                        unsigned long long*, double);                                           \
    template void fixed_to_source<real>(hipStream_t, int, unsigned long long*, double, real, real, bool, real,  \
                                        real*, real*, const WaveModel<real>*, int*, real*);                     \
    template WaveModel<real> make_model<real>(int, double, double, double, double, double, double, int);        \
    template void center_of_mass<real>(hipStream_t, const PState<real>&, int, unsigned long long*, real*);       \
    template void cell_coefficients<real>(hipStream_t, const WaveModel<real>&, real*, real*, real*, real*);      \
    template void cell_keys<real>(hipStream_t, const PState<real>&, const Geo<real>&, unsigned*, unsigned*);     \
    template void permute<real>(hipStream_t, const PState<real>&, const PState<real>&, const unsigned*,          \
                                const int*, int*);                                                              \
    template void mass_to_source<real>(hipStream_t, int, const real*, real, real, bool, real, real*,            \
                                       const WaveModel<real>*);                                                 \
    template void wave_iteration<real>(hipStream_t, const WaveModel<real>&, const WaveK<real>&, real*,          \
                                       const real*, const real*, const real*, void*, const int*);               \
    template int wave_fused_levels<real>(const WaveModel<real>&, int);                                          \
    template bool wave_fused_pairs<real>(const WaveModel<real>&, int);                                          \
    template void wave_fused<real>(hipStream_t, const WaveModel<real>&, const WaveK<real>&, int, real*, real*,  \
                                   const real*, const real*, const real*, const int*, const int*, const real*,  \
                                   bool, bool, void*);                                                          \
    template void solve_init<real>(hipStream_t, SolveCtl*, void*);                                              \
    template void solve_decide<real>(hipStream_t, SolveCtl*, void*, real, int, int, int, int);                  \
    template void solve_fixup<real>(hipStream_t, const SolveCtl*, const PhiBufs<real>&, int, int, int64_t);     \
    template void gradient<real>(hipStream_t, int, real, const real*, real*);                                   \
    template void gather<real>(hipStream_t, const PState<real>&, const Geo<real>&, const real*, real, bool);    \
    template void gather_phi<real>(hipStream_t, const PState<real>&, const Geo<real>&, const real*, real, bool);\
    template void ngp_phi<real>(hipStream_t, const PState<real>&, const Geo<real>&, const real*);               \
    template void max_diff<real>(hipStream_t, int64_t, const real*, const real*, void*);                       \
    template void ngp_bin<real>(hipStream_t, const PState<real>&, const Geo<real>&, NgpBins&);                   \
    template void ngp_project<real>(hipStream_t, int, const NgpBins&, int64_t, const real*, const real*, real*,  \
                                    real*);                                                                     \
    template void fixed_to_mass<real>(hipStream_t, int, unsigned long long*, double, real*);

RRG_PM_INST(float)
RRG_PM_INST(double)

}  // namespace pm
}  // namespace rrg
