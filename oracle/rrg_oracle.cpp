// rrg_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the RRGalaxSim solver's per-step hot path, used as the
// checker for the HIP library.  Only the tests, the smoke check and the CPU
// baseline leg of the benchmark load it; the product never does.
//
// Every routine follows one reference function and keeps its operation order,
// its mixed float/double promotions (the reference's `real` is `float` under
// USE_FP32, but several literals are `double`), and its quirks.  The file is
// compiled with -ffp-contract=off so no multiply-add is fused behind our back.
// All routines are templated on `real` (double = reference default, float =
// reference USE_FP32 build, co_types.h:4-18) and exported with _f64/_f32
// suffixes through a plain C ABI for ctypes.
//
// Reference files are cited as path:line relative to the reference root.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr double kPi = 3.14159265358979;  // co_info_structs.h:6 (PI macro, a double literal)

template <class real> struct V3 { real x, y, z; };

template <class real> inline real cube(real x) { return x * x * x; }  // co_device_utils.cuh:55-58

template <class T> inline T clampv(T v, T lo, T hi) {  // co_device_utils.cuh:43-48
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// ---------------------------------------------------------------------------
// Direct sum: acceleration_kernel_blocked_ (nbody.cu:19-65).
// a_i = sum_j m_j d_ij / sqrt(|d|^2 + 0.5 (s_i + s_j))^3, d = x_j - x_i, j in
// index order 0..N-1, self pair included.  `1. / sqrt(dot + 0.5 * (..))` is
// evaluated in double even for real=float (the literals are double) and the
// quotient is then rounded to real (nbody.cu:50-53).
// ---------------------------------------------------------------------------
template <class real>
void accel(int64_t n, const real* pos, const real* mass, const real* soft2, real* acc) {
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    V3<real>* a = reinterpret_cast<V3<real>*>(acc);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const V3<real> pc = p[i];
        const real sc = soft2[i];
        V3<real> f{0, 0, 0};
        for (int64_t j = 0; j < n; ++j) {
            const real dx = p[j].x - pc.x;
            const real dy = p[j].y - pc.y;
            const real dz = p[j].z - pc.z;
            const real dot = dx * dx + dy * dy + dz * dz;
            const real den = 1. / std::sqrt(dot + 0.5 * (sc + soft2[j]));
            const real k = mass[j] * cube(den);
            f.x += dx * k;
            f.y += dy * k;
            f.z += dz * k;
        }
        a[i] = f;
    }
}

// Subset variant for large N: accelerations of the targets idx[0..m) only.
template <class real>
void accel_subset(int64_t n, const real* pos, const real* mass, const real* soft2,
                  int64_t m, const int64_t* idx, real* acc) {
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    V3<real>* a = reinterpret_cast<V3<real>*>(acc);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < m; ++q) {
        const int64_t i = idx[q];
        const V3<real> pc = p[i];
        const real sc = soft2[i];
        V3<real> f{0, 0, 0};
        for (int64_t j = 0; j < n; ++j) {
            const real dx = p[j].x - pc.x;
            const real dy = p[j].y - pc.y;
            const real dz = p[j].z - pc.z;
            const real dot = dx * dx + dy * dy + dz * dz;
            const real den = 1. / std::sqrt(dot + 0.5 * (sc + soft2[j]));
            const real k = mass[j] * cube(den);
            f.x += dx * k;
            f.y += dy * k;
            f.z += dz * k;
        }
        a[q] = f;
    }
}

// grav_kernel_blocked_ (nbody.cu:67-114): grav_i = -sum_j m_j / sqrt(r^2 + h_ij),
// self term j = i INCLUDED (nbody.cu:108-113).  `grav_sum += m / sqrt(double)`
// accumulates through double and rounds back to real at every j.
template <class real>
void grav(int64_t n, const real* pos, const real* mass, const real* soft2, real* out) {
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const V3<real> pc = p[i];
        const real sc = soft2[i];
        real g = 0.;
        for (int64_t j = 0; j < n; ++j) {
            const real dx = p[j].x - pc.x;
            const real dy = p[j].y - pc.y;
            const real dz = p[j].z - pc.z;
            const real dot = dx * dx + dy * dy + dz * dz;
            g += mass[j] / std::sqrt(dot + 0.5 * (sc + soft2[j]));
        }
        out[i] = -g;
    }
}

// predict_step_ (nbody.cu:122-152): v* = v + dt a;  x += dt * 0.5 * (v + v*)
// (`dt * 0.5` is double, so the position update runs in double before the
// store rounds it to real).
template <class real>
void predict(int64_t n, const real* acc, const real* vel, real* pos, real* vel_pred, real dt) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            const real v = vel[3 * i + c];
            const real r = pos[3 * i + c];
            const real a = acc[3 * i + c];
            const real vp = v + dt * a;
            vel_pred[3 * i + c] = vp;
            pos[3 * i + c] = r + dt * 0.5 * (v + vp);
        }
    }
}

// correct_step_ (nbody.cu:159-181): v = 0.5 * (v + v* + dt a_new)
template <class real>
void correct(int64_t n, const real* acc_new, real* vel, const real* vel_pred, real dt) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            const real v = vel[3 * i + c];
            const real vv = vel_pred[3 * i + c];
            const real aa = acc_new[3 * i + c];
            vel[3 * i + c] = 0.5 * (v + vv + dt * aa);
        }
    }
}

// calc_conservation (nbody.cu:246-291): sequential sums in real.
// out = {Ek, Ep, Px, Py, Pz, Lx, Ly, Lz}
template <class real>
void conservation(int64_t n, const real* pos, const real* vel, const real* mass, const real* gr,
                  real* out) {
    real px = 0, py = 0, pz = 0, lx = 0, ly = 0, lz = 0, ek = 0, ep = 0;
    for (int64_t i = 0; i < n; ++i) {
        const real x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        const real m = mass[i];
        const real vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        px += m * vx;
        py += m * vy;
        pz += m * vz;
        lx += m * (vz * y - vy * z);
        ly += m * (vx * z - vz * x);
        lz += m * (vy * x - vx * y);
        ek += m * (vx * vx + vy * vy + vz * vz);
        ep += m * gr[i];
    }
    ek *= 0.5;
    ep *= 0.5;
    out[0] = ek; out[1] = ep;
    out[2] = px; out[3] = py; out[4] = pz;
    out[5] = lx; out[6] = ly; out[7] = lz;
}

// ---------------------------------------------------------------------------
// Grid geometry (io_json.cpp:28-40) and wave parameters (io_json.cpp:53-99,
// 170-186), kept in `real` like the reference's GridInfo/WaveParams.
// ---------------------------------------------------------------------------
template <class real> struct Grid {
    int nx;
    real dx, domain_l, domain_min, domain_max, sim_l, bc_l;
};

template <class real> Grid<real> make_grid(int nx, real dx, real bc_frac) {
    Grid<real> g;
    g.nx = nx;
    g.dx = dx;
    g.domain_l = g.dx * g.nx;
    g.sim_l = (1. - bc_frac) * g.domain_l;
    g.bc_l = 0.5 * bc_frac * g.domain_l;
    g.domain_max = 0.5 * g.domain_l;
    g.domain_min = -0.5 * g.domain_l;
    return g;
}

inline size_t AT(int x, int y, int z, int NX) {  // co_device_utils.cuh:61
    return size_t(x) + size_t(y) * NX + size_t(z) * NX * NX;
}

// CIC weights of deposit_cell_mass_CIC_ (rm_p2mesh.cu:182-226): cell-centred
// (-0.5 shift), u clamped to [0, nx-1-1e-6], upper index clamped to nx-1.
template <class real> struct CicStencil {
    int i0, j0, k0, i1, j1, k1;
    real w[8];  // order 000,100,010,110,001,101,011,111 (rm_p2mesh.cu:219-226)
};

template <class real>
CicStencil<real> cic_stencil(const V3<real>& pos, const Grid<real>& g) {
    const int NX = g.nx;
    const real inv_dx = 1.f / g.dx;
    real u = (pos.x - g.domain_min) * inv_dx - 0.5f;
    real v = (pos.y - g.domain_min) * inv_dx - 0.5f;
    real w = (pos.z - g.domain_min) * inv_dx - 0.5f;
    const real eps = 1E-6;
    u = clampv<real>(u, 0, (NX - 1) - eps);
    v = clampv<real>(v, 0, (NX - 1) - eps);
    w = clampv<real>(w, 0, (NX - 1) - eps);
    CicStencil<real> s;
    s.i0 = (int)std::floor(u);
    s.j0 = (int)std::floor(v);
    s.k0 = (int)std::floor(w);
    s.i1 = std::min(s.i0 + 1, NX - 1);
    s.j1 = std::min(s.j0 + 1, NX - 1);
    s.k1 = std::min(s.k0 + 1, NX - 1);
    const real fx = u - s.i0, fy = v - s.j0, fz = w - s.k0;
    const real wx0 = 1.f - fx, wx1 = fx;
    const real wy0 = 1.f - fy, wy1 = fy;
    const real wz0 = 1.f - fz, wz1 = fz;
    s.w[0] = wx0 * wy0 * wz0;
    s.w[1] = wx1 * wy0 * wz0;
    s.w[2] = wx0 * wy1 * wz0;
    s.w[3] = wx1 * wy1 * wz0;
    s.w[4] = wx0 * wy0 * wz1;
    s.w[5] = wx1 * wy0 * wz1;
    s.w[6] = wx0 * wy1 * wz1;
    s.w[7] = wx1 * wy1 * wz1;
    return s;
}

template <class real> inline void cic_cells(const CicStencil<real>& s, int NX, size_t* c) {
    c[0] = AT(s.i0, s.j0, s.k0, NX);
    c[1] = AT(s.i1, s.j0, s.k0, NX);
    c[2] = AT(s.i0, s.j1, s.k0, NX);
    c[3] = AT(s.i1, s.j1, s.k0, NX);
    c[4] = AT(s.i0, s.j0, s.k1, NX);
    c[5] = AT(s.i1, s.j0, s.k1, NX);
    c[6] = AT(s.i0, s.j1, s.k1, NX);
    c[7] = AT(s.i1, s.j1, s.k1, NX);
}

// Reference deposit: floating-point adds in particle order (the reference's
// atomicAdd order on the GPU is unspecified; sequential order is one of them).
template <class real>
void cic_deposit_float(int64_t n, const real* pos, const real* mass, int nx, real dx, real bc_frac,
                       real* grid) {
    const Grid<real> g = make_grid<real>(nx, dx, bc_frac);
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    std::fill(grid, grid + size_t(nx) * nx * nx, real(0));
    for (int64_t q = 0; q < n; ++q) {
        const CicStencil<real> s = cic_stencil(p[q], g);
        size_t c[8];
        cic_cells(s, nx, c);
        for (int k = 0; k < 8; ++k) grid[c[k]] += mass[q] * s.w[k];
    }
}

// Fixed-point deposit (the build's order-independent semantics): each of the
// 8 products m*w is rounded once to an integer multiple of 2^-e and summed in
// int64, so the grid is independent of particle order.
template <class real> inline int64_t to_fixed(real v, int e) {
    return (int64_t)std::llrint(std::ldexp((double)v, e));
}

// Integer sums are exact, so the particles are split over the host threads:
// each thread deposits into a private grid and the grids are added cell by
// cell (integer atomics instead when the private grids would pass 1 GiB);
// the result is the same as a sequential deposit.
template <class real>
void cic_deposit_fixed(int64_t n, const real* pos, const real* mass, int nx, real dx, real bc_frac,
                       int e, int64_t* grid) {
    const Grid<real> g = make_grid<real>(nx, dx, bc_frac);
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    const size_t ncell = size_t(nx) * nx * nx;
    const int nt = omp_get_max_threads();
    const bool priv = nt > 1 && size_t(nt) * ncell * sizeof(int64_t) <= (size_t(1) << 30) &&
                      n >= int64_t(ncell) / 8;
    if (!priv) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < int64_t(ncell); ++c) grid[c] = 0;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; ++q) {
            const CicStencil<real> s = cic_stencil(p[q], g);
            size_t c[8];
            cic_cells(s, nx, c);
            for (int k = 0; k < 8; ++k) {
                const int64_t v = to_fixed<real>(mass[q] * s.w[k], e);
#pragma omp atomic
                grid[c[k]] += v;
            }
        }
        return;
    }
    std::vector<int64_t> part(size_t(nt) * ncell);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        int64_t* mine = part.data() + size_t(t) * ncell;
        std::fill(mine, mine + ncell, int64_t(0));
#pragma omp for schedule(static)
        for (int64_t q = 0; q < n; ++q) {
            const CicStencil<real> s = cic_stencil(p[q], g);
            size_t c[8];
            cic_cells(s, nx, c);
            for (int k = 0; k < 8; ++k) mine[c[k]] += to_fixed<real>(mass[q] * s.w[k], e);
        }
#pragma omp for schedule(static)
        for (int64_t c = 0; c < int64_t(ncell); ++c) {
            int64_t v = 0;
            for (int u = 0; u < nt; ++u) v += part[size_t(u) * ncell + size_t(c)];
            grid[c] = v;
        }
    }
}

template <class real> void fixed_to_real(size_t ncell, const int64_t* fx, int e, real* out) {
    for (size_t c = 0; c < ncell; ++c) out[c] = (real)std::ldexp((double)fx[c], -e);
}

// ---------------------------------------------------------------------------
// Damped wave iteration: wave_diss_iteration (wave.cu:106-193), periodic and
// zero-Dirichlet boundary types (co_boundary.cuh:4-9) with a uniform base
// dissipation (no absorbing layer: the live configuration sets
// periodic_keep_layer = false, rrgsim_integration_rect.cuh:65-69).
// ---------------------------------------------------------------------------
template <class real> struct WaveCoef {
    real k_main, k_f, k_actual, k_old;
};

template <class real> WaveCoef<real> wave_coef(real diss, real C, real DT, real DX) {
    const int DIM = 3;  // wave.cu:164-171
    const real Q = diss * C * DT * real(0.5);
    const real W = 1 + Q;
    const real K = C * C * DT * DT / (DX * DX);
    WaveCoef<real> k;
    k.k_main = K / W;
    k.k_f = -C * C * DT * DT / W;
    k.k_actual = 2 * (1 - DIM * K) / W;
    k.k_old = (Q - 1) / W;
    return k;
}

inline int wrapi(int i, int n) { return (i < 0) ? i + n : ((i >= n) ? i - n : i); }  // wave.cu:22-24

// Boundary description (co_boundary.cuh:4-27): the reference's BoundaryType
// values, the monopole mass and centre, r_min, and the periodic mean density.
enum { BC_ZERO_DIRICHLET = 0, BC_PERIODIC = 1, BC_MONOPOLE_DIRICHLET = 2, BC_MONOPOLE_RADIATIVE = 3 };

template <class real> struct WaveBC {
    int type;
    real mass;
    V3<real> com;
    real r_min;
    real rho_mean;
};

// Everything wave_diss_iteration reads from its __constant__ copies
// (grid_info_, wave_params_, bc_params_; co_device_structs.cu:7-13).
template <class real> struct WaveModel {
    Grid<real> g;
    real diss_base, diss_extra, C, DT;
    WaveBC<real> bc;
};

template <class real> inline real coord(const Grid<real>& g, int i) {  // wave.cu:18-20
    return g.domain_min + i * g.dx;
}

template <class real> inline real layer_diss(const WaveModel<real>& w, real x) {  // wave.cu:27-33
    const real x_in = real(0.5) * w.g.sim_l;
    const real d = std::fabs(x) - x_in;
    return (d > 0) ? w.diss_extra * d * d : real(0);
}

template <class real> inline real monopole_phi(const WaveModel<real>& w, real x, real y, real z) {  // wave.cu:36-41
    const real G = 1.;
    const real rx = x - w.bc.com.x, ry = y - w.bc.com.y, rz = z - w.bc.com.z;  // distance(com, p)
    const real dist = std::sqrt(rx * rx + ry * ry + rz * rz);
    const real r = std::fmax(dist, w.bc.r_min);
    return -G * w.bc.mass / r;
}

template <class real> inline real psi_at(const WaveModel<real>& w, const real* phi, int i, int j, int k) {
    const int NX = w.g.nx;  // wave.cu:44-49
    return phi[AT(i, j, k, NX)] - monopole_phi(w, coord(w.g, i), coord(w.g, j), coord(w.g, k));
}

template <class real>
inline real upwind_axis(const WaveModel<real>& w, const real* phi, real psi0, real n_comp, int ix, int iy, int iz,
                        int axis) {  // wave.cu:53-76
    const int NX = w.g.nx;
    const real DX = w.g.dx;
    const int i0 = (axis == 0) ? ix : (axis == 1) ? iy : iz;
    int in = (n_comp > 0) ? i0 - 1 : i0 + 1;
    if (in < 0 || in >= NX) in = 2 * i0 - in;
    int jx = ix, jy = iy, jz = iz;
    if (axis == 0) jx = in;
    else if (axis == 1) jy = in;
    else jz = in;
    const real d = (psi0 - psi_at(w, phi, jx, jy, jz)) / ((i0 - in) * DX);
    return n_comp * d;
}

template <class real> real radiative_boundary(const WaveModel<real>& w, const real* phi, int ix, int iy, int iz) {
    const real C = w.C, DT = w.DT;  // wave.cu:80-104
    const real x = coord(w.g, ix), y = coord(w.g, iy), z = coord(w.g, iz);
    const real rx = x - w.bc.com.x, ry = y - w.bc.com.y, rz = z - w.bc.com.z;
    const real r = std::fmax(std::sqrt(rx * rx + ry * ry + rz * rz), w.bc.r_min);
    const real nx = rx / r, ny = ry / r, nz = rz / r;
    const real psi0 = psi_at(w, phi, ix, iy, iz);
    const real dpsi_dr = upwind_axis(w, phi, psi0, nx, ix, iy, iz, 0) + upwind_axis(w, phi, psi0, ny, ix, iy, iz, 1) +
                         upwind_axis(w, phi, psi0, nz, ix, iy, iz, 2);
    const real psi_new = psi0 - C * DT * (dpsi_dr + psi0 / r);
    return psi_new + monopole_phi(w, x, y, z);
}

// Per-cell scheme coefficients (wave.cu:155-171): the absorbing layer adds
// to the dissipation on every non-periodic grid (periodic_keep_layer is 0).
template <class real> WaveCoef<real> cell_coef(const WaveModel<real>& w, int ix, int iy, int iz) {
    real diss = w.diss_base;
    if (w.bc.type != BC_PERIODIC) {
        diss += layer_diss(w, coord(w.g, ix));
        diss += layer_diss(w, coord(w.g, iy));
        diss += layer_diss(w, coord(w.g, iz));
    }
    return wave_coef<real>(diss, w.C, w.DT, w.g.dx);
}

// wave_diss_iteration (wave.cu:106-193) for every boundary type.
template <class real>
void wave_iteration_model(const WaveModel<real>& w, real* phi_new, const real* phi, const real* phi_old,
                          const real* mass) {
    const int NX = w.g.nx;
    const real DX = w.g.dx;
    const bool periodic = w.bc.type == BC_PERIODIC;
    const real G = 1.;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < NX; ++iz) {
        for (int iy = 0; iy < NX; ++iy) {
            for (int ix = 0; ix < NX; ++ix) {
                const bool on_edge = ix == 0 || ix == NX - 1 || iy == 0 || iy == NX - 1 ||
                                     iz == 0 || iz == NX - 1;
                if (!periodic && on_edge) {  // wave.cu:129-147
                    real v;
                    if (w.bc.type == BC_MONOPOLE_DIRICHLET)
                        v = monopole_phi(w, coord(w.g, ix), coord(w.g, iy), coord(w.g, iz));
                    else if (w.bc.type == BC_MONOPOLE_RADIATIVE)
                        v = radiative_boundary(w, phi, ix, iy, iz);
                    else
                        v = 0.;
                    phi_new[AT(ix, iy, iz, NX)] = v;
                    continue;
                }
                real rho = mass[AT(ix, iy, iz, NX)] / (DX * DX * DX);
                if (periodic) rho -= w.bc.rho_mean;
                const real f = 4 * kPi * G * rho;
                const WaveCoef<real> k = cell_coef(w, ix, iy, iz);
                int xm = ix - 1, xp = ix + 1, ym = iy - 1, yp = iy + 1, zm = iz - 1, zp = iz + 1;
                if (periodic) {
                    xm = wrapi(xm, NX); xp = wrapi(xp, NX);
                    ym = wrapi(ym, NX); yp = wrapi(yp, NX);
                    zm = wrapi(zm, NX); zp = wrapi(zp, NX);
                }
                const real lap_sum = phi[AT(xm, iy, iz, NX)] + phi[AT(xp, iy, iz, NX)] +
                                     phi[AT(ix, ym, iz, NX)] + phi[AT(ix, yp, iz, NX)] +
                                     phi[AT(ix, iy, zm, NX)] + phi[AT(ix, iy, zp, NX)];
                phi_new[AT(ix, iy, iz, NX)] = lap_sum * k.k_main + phi[AT(ix, iy, iz, NX)] * k.k_actual +
                                              phi_old[AT(ix, iy, iz, NX)] * k.k_old + f * k.k_f;
            }
        }
    }
}

template <class real>
WaveModel<real> plain_model(int nx, real dx, real diss, real C, real DT, int periodic, real rho_mean) {
    WaveModel<real> w{};
    w.g = make_grid<real>(nx, dx, real(0));
    w.diss_base = diss;
    w.diss_extra = 0;
    w.C = C;
    w.DT = DT;
    w.bc.type = periodic ? BC_PERIODIC : BC_ZERO_DIRICHLET;
    w.bc.rho_mean = rho_mean;
    return w;
}

// The periodic / zero-Dirichlet iteration with no absorbing layer.
template <class real>
void wave_iteration(int nx, real dx, real diss, real C, real DT, int periodic, real rho_mean,
                    real* phi_new, const real* phi, const real* phi_old, const real* mass) {
    wave_iteration_model(plain_model<real>(nx, dx, diss, C, DT, periodic, rho_mean), phi_new, phi, phi_old, mass);
}

// SteadyStateChecker::calc_max_diff (wave_steady_checker.cu:95-127):
// err = max|u_new - u_old| / (max|u_new| + 1e-30), in real.
template <class real> double steady_err(size_t ncell, const real* u_new, const real* u_old) {
    real diff = 0, absu = 0;
    for (size_t c = 0; c < ncell; ++c) {
        diff = std::max<real>(diff, std::fabs(u_new[c] - u_old[c]));
        absu = std::max<real>(absu, std::fabs(u_new[c]));
    }
    const real EPS = 1.E-30;
    const real err = diff / (absu + EPS);
    return err;
}

// solve_wave_equation (wave.cu:230-297): three-buffer rotation, steady check
// every `check_step` iterations after the rotation, stop after `hold`
// consecutive checks below `tol` (wave_steady_checker.cu:131-143).
// bufs = {prev, curr, next}; on return bufs[1] holds the newest field.  The
// rotation is done by permuting the three pointers, and `perm` reports it.
template <class real>
int solve_wave_model(const WaveModel<real>& w, real* prev, real* curr, real* next, const real* mass, int max_iter,
                     real tol, int check_step, int hold, int* perm) {
    const int nx = w.g.nx;
    real* b[3] = {prev, curr, next};
    int id[3] = {0, 1, 2};
    int hold_count = 0;
    const size_t ncell = size_t(nx) * nx * nx;
    int iter = 0;
    for (iter = 0; iter < max_iter; ++iter) {
        wave_iteration_model<real>(w, b[2], b[1], b[0], mass);
        std::swap(b[0], b[1]); std::swap(id[0], id[1]);
        std::swap(b[1], b[2]); std::swap(id[1], id[2]);
        if (iter && iter % check_step == 0) {
            const double err = steady_err<real>(ncell, b[1], b[0]);
            if (err < tol) ++hold_count; else hold_count = 0;
            if (hold_count >= hold) { ++iter; break; }
        }
    }
    for (int q = 0; q < 3; ++q) perm[q] = id[q];
    return iter;  // number of stencil iterations performed
}

template <class real>
int solve_wave(int nx, real dx, real diss, real C, real DT, int periodic, real rho_mean,
               real* prev, real* curr, real* next, const real* mass, int max_iter, real tol,
               int check_step, int hold, int* perm) {
    return solve_wave_model(plain_model<real>(nx, dx, diss, C, DT, periodic, rho_mean), prev, curr, next, mass,
                            max_iter, tol, check_step, hold, perm);
}

// calc_acceleration_field_ (rm_p2mesh.cu:238-290): a = -grad(phi), central
// differences inside, one-sided on the faces.  `0.5 * (..)` is a double
// product and `-1. / DX` a double quotient, both rounded to real.
template <class real> void gradient(int nx, real dx, const real* phi, real* acc) {
    const int NX = nx;
    const real DX = dx;
    const real coef = -1. / DX;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < NX; ++iz) {
        for (int iy = 0; iy < NX; ++iy) {
            for (int ix = 0; ix < NX; ++ix) {
                const size_t xyz = AT(ix, iy, iz, NX);
                real d[3];
                if (ix == 0) d[0] = phi[AT(ix + 1, iy, iz, NX)] - phi[xyz];
                else if (ix == NX - 1) d[0] = phi[xyz] - phi[AT(ix - 1, iy, iz, NX)];
                else d[0] = 0.5 * (phi[AT(ix + 1, iy, iz, NX)] - phi[AT(ix - 1, iy, iz, NX)]);
                if (iy == 0) d[1] = phi[AT(ix, iy + 1, iz, NX)] - phi[xyz];
                else if (iy == NX - 1) d[1] = phi[xyz] - phi[AT(ix, iy - 1, iz, NX)];
                else d[1] = 0.5 * (phi[AT(ix, iy + 1, iz, NX)] - phi[AT(ix, iy - 1, iz, NX)]);
                if (iz == 0) d[2] = phi[AT(ix, iy, iz + 1, NX)] - phi[xyz];
                else if (iz == NX - 1) d[2] = phi[xyz] - phi[AT(ix, iy, iz - 1, NX)];
                else d[2] = 0.5 * (phi[AT(ix, iy, iz + 1, NX)] - phi[AT(ix, iy, iz - 1, NX)]);
                acc[3 * xyz + 0] = d[0] * coef;
                acc[3 * xyz + 1] = d[1] * coef;
                acc[3 * xyz + 2] = d[2] * coef;
            }
        }
    }
}

// interpolate_acceleration_to_particles_kernel (rm_mesh2p.cu:48-114):
// trilinear, cell-centred; the cell index goes through floorf (a float floor
// even when real = double), the weights come from the unclamped offset and
// only the indices are clamped; corner loop order k, j, i.
template <class real>
void gather(int64_t n, const real* pos, int nx, real dx, real bc_frac, const real* field, real* out) {
    const Grid<real> g = make_grid<real>(nx, dx, bc_frac);
    const real domain_min = g.domain_min;
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    const V3<real>* F = reinterpret_cast<const V3<real>*>(field);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
        const V3<real> ps = p[q];
        const real fx = (ps.x - domain_min) / dx - 0.5f;
        const real fy = (ps.y - domain_min) / dx - 0.5f;
        const real fz = (ps.z - domain_min) / dx - 0.5f;
        int i0 = (int)std::floor((float)fx);
        int j0 = (int)std::floor((float)fy);
        int k0 = (int)std::floor((float)fz);
        int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
        const real wx = fx - i0, wy = fy - j0, wz = fz - k0;
        i0 = std::max(0, std::min(i0, nx - 1));
        i1 = std::max(0, std::min(i1, nx - 1));
        j0 = std::max(0, std::min(j0, nx - 1));
        j1 = std::max(0, std::min(j1, nx - 1));
        k0 = std::max(0, std::min(k0, nx - 1));
        k1 = std::max(0, std::min(k1, nx - 1));
        V3<real> a{0.0f, 0.0f, 0.0f};
        for (int k = 0; k <= 1; k++) {
            const int kk = (k == 0) ? k0 : k1;
            const real wk = (k == 0) ? (1.0f - wz) : wz;
            for (int j = 0; j <= 1; j++) {
                const int jj = (j == 0) ? j0 : j1;
                const real wj = (j == 0) ? (1.0f - wy) : wy;
                for (int i = 0; i <= 1; i++) {
                    const int ii = (i == 0) ? i0 : i1;
                    const real wi = (i == 0) ? (1.0f - wx) : wx;
                    const real weight = wi * wj * wk;
                    const int idx = kk * nx * nx + jj * nx + ii;
                    a.x += weight * F[idx].x;
                    a.y += weight * F[idx].y;
                    a.z += weight * F[idx].z;
                }
            }
        }
        reinterpret_cast<V3<real>*>(out)[q] = a;
    }
}

// NGP cell of assign_particles_to_cells_ (rm_p2mesh.cu:34-53) and the phi
// projection projectCellPhiOnParticles (rm_mesh2p.cu:29-38).
template <class real>
void ngp_phi(int64_t n, const real* pos, int nx, real dx, real bc_frac, const real* phi, real* out) {
    const Grid<real> g = make_grid<real>(nx, dx, bc_frac);
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    for (int64_t q = 0; q < n; ++q) {
        const int ix = clampv<int>((int)((p[q].x - g.domain_min) / g.dx), 0, nx - 1);
        const int iy = clampv<int>((int)((p[q].y - g.domain_min) / g.dx), 0, nx - 1);
        const int iz = clampv<int>((int)((p[q].z - g.domain_min) / g.dx), 0, nx - 1);
        out[q] = phi[AT(ix, iy, iz, nx)];
    }
}

// CFL step of validate_dependent_fields (io_json.cpp:174-178).
// NGP binning of the particles (assign_particles_to_cells_ and the prefix
// sums, rm_p2mesh.cu:25-126) as a counting sort: each particle's cell (the
// truncated (x - min)/dx, clamped), the per-cell counts, their exclusive
// prefix sums and the particle order sorted by cell (ascending particle index
// within a cell).  The reference's prefix kernel never writes its in-block
// scan back, so its start_id holds only block offsets; this is the intended
// exclusive prefix.
template <class real>
void ngp_bin(int64_t n, const real* pos, int nx, real dx, real bc_frac, int* cell_of, int* count, int* start,
             int* order) {
    const Grid<real> g = make_grid<real>(nx, dx, bc_frac);
    const V3<real>* p = reinterpret_cast<const V3<real>*>(pos);
    const size_t ncell = size_t(nx) * nx * nx;
    std::fill(count, count + ncell, 0);
    for (int64_t q = 0; q < n; ++q) {
        const int ix = clampv<int>((int)((p[q].x - g.domain_min) / g.dx), 0, nx - 1);
        const int iy = clampv<int>((int)((p[q].y - g.domain_min) / g.dx), 0, nx - 1);
        const int iz = clampv<int>((int)((p[q].z - g.domain_min) / g.dx), 0, nx - 1);
        cell_of[q] = int(AT(ix, iy, iz, nx));
        ++count[cell_of[q]];
    }
    int run = 0;
    for (size_t c = 0; c < ncell; ++c) {
        start[c] = run;
        run += count[c];
    }
    std::vector<int> fill(start, start + ncell);
    for (int64_t q = 0; q < n; ++q) order[fill[cell_of[q]]++] = int(q);
}

// Cell averages of the particle potential and acceleration
// (deposit_cell_grav_NGP_ and deposit_cell_acc_NGP_, rm_p2mesh.cu:139-168):
// phi_c = sum_p phi_p / count_c and acc_c = sum_p a_p * (1. / count_c), the
// sums taken in particle order (the reference's atomics leave it open).
template <class real>
void ngp_project(int nx, const int* count, const int* start, const int* order, const real* grav, const real* acc,
                 real* cell_phi, real* cell_acc) {
    const size_t ncell = size_t(nx) * nx * nx;
    for (size_t c = 0; c < ncell; ++c) {
        real ph = 0, ax = 0, ay = 0, az = 0;
        if (count[c] > 0) {
            const real inv_count = 1. / count[c];
            for (int k = start[c]; k < start[c] + count[c]; ++k) {
                const int q = order[k];
                ph += grav[q] / count[c];
                ax += acc[3 * q + 0] * inv_count;
                ay += acc[3 * q + 1] * inv_count;
                az += acc[3 * q + 2] * inv_count;
            }
        }
        cell_phi[c] = ph;
        cell_acc[3 * c + 0] = ax;
        cell_acc[3 * c + 1] = ay;
        cell_acc[3 * c + 2] = az;
    }
}

template <class real> real wave_dt(real dx, real c) {
    constexpr real DIM = 3;
    const real dt = 0.5 * dx / (c * std::sqrt(DIM));
    return dt;
}

// set_boundary_periodic (wave.cu:214-228) with the hard-coded total mass 1
// (rrgsim_integration_rect.cuh:65-69).
template <class real> real rho_mean_periodic(real total_mass, int nx, real dx) {
    const real L = nx * dx;
    return total_mass / (L * L * L);
}

// ---------------------------------------------------------------------------
// Whole runs: the N-body strategy (rrgsim_integration_nbody.cuh:26-56) and the
// PM ("wave") strategy (rrgsim_integration_rect.cuh:53-112) with sync after
// precalc (rrgsim_integration.cu:49-51).  State is AoS real3.
// ---------------------------------------------------------------------------
template <class real>
void nbody_run(int64_t n, real* pos, real* vel, const real* mass, const real* soft2, real dt,
               int nsteps, real* acc_out, real* grav_out) {
    std::vector<real> acc(3 * n), vp(3 * n);
    grav<real>(n, pos, mass, soft2, grav_out);  // precalc
    accel<real>(n, pos, mass, soft2, acc.data());
    for (int s = 0; s < nsteps; ++s) {
        predict<real>(n, acc.data(), vel, pos, vp.data(), dt);
        accel<real>(n, pos, mass, soft2, acc.data());
        correct<real>(n, acc.data(), vel, vp.data(), dt);
    }
    if (acc_out) std::memcpy(acc_out, acc.data(), sizeof(real) * 3 * n);
}

// ---------------------------------------------------------------------------
// Reverse-mode gradient of the N-body run above (no reference counterpart):
// the transpose of predict/accel/correct, FP64, all positions kept.  Pinned
// by central finite differences of nbody_run in the tests.
// VJP at x: g_k = sum_j T(x_j - x_k)(m_k u_j - m_j u_k) with
// T(d) = r^-3 I - 3 r^-5 d d^T (r^2 = |d|^2 + h_kj), and the mass gradient
// gm_k = sum_i u_i . K(x_k - x_i), K(d) = d r^-3.
// ---------------------------------------------------------------------------
void vjp(int64_t n, const double* x, const double* m, const double* s2, const double* u, double* g, double* gm) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double gk[3] = {0, 0, 0}, gmk = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == k) continue;
            double d[3], w[3];
            for (int c = 0; c < 3; ++c) {
                d[c] = x[3 * j + c] - x[3 * k + c];
                w[c] = m[k] * u[3 * j + c] - m[j] * u[3 * k + c];
            }
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + 0.5 * (s2[k] + s2[j]);
            const double r = std::sqrt(r2);
            const double inv3 = 1.0 / (r2 * r);
            const double inv5 = inv3 / r2;
            const double dw = d[0] * w[0] + d[1] * w[1] + d[2] * w[2];
            for (int c = 0; c < 3; ++c) gk[c] += inv3 * w[c] - 3.0 * inv5 * dw * d[c];
            // gm_k collects u_j . K(x_k - x_j) = -(u_j . d) r^-3
            gmk -= (u[3 * j] * d[0] + u[3 * j + 1] * d[1] + u[3 * j + 2] * d[2]) * inv3;
        }
        for (int c = 0; c < 3; ++c) g[3 * k + c] = gk[c];
        gm[k] = gmk;
    }
}

void nbody_adjoint(int64_t n, const double* x0, const double* v0, const double* m, const double* s2, double dt,
                   int nsteps, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gmo) {
    std::vector<double> xs(size_t(nsteps + 1) * 3 * n), x(3 * n), v(3 * n), a(3 * n), vp(3 * n);
    std::memcpy(x.data(), x0, sizeof(double) * 3 * n);
    std::memcpy(v.data(), v0, sizeof(double) * 3 * n);
    std::memcpy(xs.data(), x0, sizeof(double) * 3 * n);
    accel<double>(n, x.data(), m, s2, a.data());
    for (int s = 0; s < nsteps; ++s) {
        predict<double>(n, a.data(), v.data(), x.data(), vp.data(), dt);
        accel<double>(n, x.data(), m, s2, a.data());
        correct<double>(n, a.data(), v.data(), vp.data(), dt);
        std::memcpy(xs.data() + size_t(s + 1) * 3 * n, x.data(), sizeof(double) * 3 * n);
    }
    std::vector<double> X(gxT, gxT + 3 * n), V(gvT, gvT + 3 * n), A(3 * n, 0.0), U(3 * n), g(3 * n), gm(n, 0.0),
        gmt(n);
    for (int s = nsteps - 1; s >= 0; --s) {
        // step s: (x_s, v_s) -> (x_{s+1}, v_{s+1}); adjoints (X, V, A) of (x_{s+1}, v_{s+1}, a_{s+1})
        for (int64_t q = 0; q < 3 * n; ++q) U[q] = A[q] + 0.5 * dt * V[q];       // correct_step_
        vjp(n, xs.data() + size_t(s + 1) * 3 * n, m, s2, U.data(), g.data(), gmt.data());  // a_{s+1} = F(x_{s+1})
        for (int64_t i = 0; i < n; ++i) gm[i] += gmt[i];
        for (int64_t q = 0; q < 3 * n; ++q) {
            const double Xt = X[q] + g[q];
            const double vstar = 0.5 * V[q] + 0.5 * dt * Xt;  // adjoint of v*
            const double vbar = 0.5 * V[q] + 0.5 * dt * Xt + vstar;
            A[q] = dt * vstar;                                 // adjoint of a_s
            V[q] = vbar;
            X[q] = Xt;
        }
    }
    vjp(n, x0, m, s2, A.data(), g.data(), gmt.data());  // a_0 = F(x_0)
    for (int64_t q = 0; q < 3 * n; ++q) X[q] += g[q];
    for (int64_t i = 0; i < n; ++i) gm[i] += gmt[i];
    std::memcpy(gx0, X.data(), sizeof(double) * 3 * n);
    std::memcpy(gv0, V.data(), sizeof(double) * 3 * n);
    std::memcpy(gmo, gm.data(), sizeof(double) * n);
}

struct PmConfig {
    int nx;
    double dx, bc_frac;
    double diss_base, wave_speed, tolerance;
    int setup_iterations, iterations, check_step, hold;
    int boundary;   // BoundaryType value (co_boundary.cuh:4-9): 0 zero-Dirichlet, 1 periodic, 2, 3 monopole
    int fixed_e;    // fixed-point exponent (deposit resolution 2^-e)
    double dt;      // particle time step (sim_params.dt_dynamics)
    double diss_extra;
};

// Centre of mass for the monopole boundaries (calc_mass_center,
// nbody.cu:294-439, in an order-independent form): each product m*x (FP64)
// and each mass are rounded once to fixed point and summed in int64.  The
// moment scale 2^E is the largest keeping n * max|m x| below 2^61, the mass
// scale 2^e is the deposit's.  No mass gives the origin, as in the reference.
template <class real>
V3<real> center_of_mass_fixed(int64_t n, const real* pos, const real* mass, int e_mass) {
    double mx = 0;
    for (int64_t q = 0; q < n; ++q)
        for (int d = 0; d < 3; ++d) mx = std::max(mx, std::fabs((double)mass[q] * (double)pos[3 * q + d]));
    int ex = 0;
    std::frexp(std::max(mx * (double)std::max<int64_t>(n, 1), 1e-300), &ex);
    const int E = std::min(61 - ex, 1000);
    int64_t S[3] = {0, 0, 0}, Sm = 0;
    for (int64_t q = 0; q < n; ++q) {
        for (int d = 0; d < 3; ++d) S[d] += (int64_t)std::rint((double)mass[q] * (double)pos[3 * q + d] * std::ldexp(1.0, E));
        Sm += (int64_t)std::rint((double)mass[q] * std::ldexp(1.0, e_mass));
    }
    const double M = (double)Sm * std::ldexp(1.0, -e_mass);
    V3<real> c{0, 0, 0};
    if (M > 0) {
        c.x = (real)((double)S[0] * std::ldexp(1.0, -E) / M);
        c.y = (real)((double)S[1] * std::ldexp(1.0, -E) / M);
        c.z = (real)((double)S[2] * std::ldexp(1.0, -E) / M);
    }
    return c;
}

// What the reverse sweep of a PM run needs per grid build (entry 0 = precalc,
// entry s = step s): the positions the mass was deposited from (after the
// drift), the potential the gradient read, and the stencil iteration count.
template <class real> struct PmTape {
    std::vector<std::vector<real>> x, phi;
    std::vector<int> iters;
};

template <class real>
void pm_run(const PmConfig& cf, int64_t n, real* pos, real* vel, const real* mass, int nsteps,
            real* acc_out, real* phi_out, real* grid_acc_out, real* grav_out, int* iters_out,
            PmTape<real>* tape = nullptr) {
    const int nx = cf.nx;
    const size_t ncell = size_t(nx) * nx * nx;
    const real dx = (real)cf.dx, bcf = (real)cf.bc_frac;
    const real C = (real)cf.wave_speed;
    const real DT = wave_dt<real>(dx, C);
    const real diss = (real)cf.diss_base;
    const real dtp = (real)cf.dt;
    const real rho_mean = rho_mean_periodic<real>(1., nx, dx);
    WaveModel<real> wm{};
    wm.g = make_grid<real>(nx, dx, bcf);
    wm.diss_base = diss;
    wm.diss_extra = (real)cf.diss_extra;
    wm.C = C;
    wm.DT = DT;
    wm.bc.type = cf.boundary;
    wm.bc.rho_mean = rho_mean;
    std::vector<int64_t> mfix(ncell);
    std::vector<real> m(ncell), pa(ncell, 0), pb(ncell, 0), pc(ncell, 0), gacc(3 * ncell);
    std::vector<real> acc(3 * n), vp(3 * n);
    real* b[3] = {pa.data(), pb.data(), pc.data()};  // prev, curr, next (zero-filled)
    auto deposit = [&]() {
        cic_deposit_fixed<real>(n, pos, mass, nx, dx, bcf, cf.fixed_e, mfix.data());
        fixed_to_real<real>(ncell, mfix.data(), cf.fixed_e, m.data());
    };
    auto solve = [&](int iters) {
        int perm[3];
        real* s[3] = {b[0], b[1], b[2]};
        const int done = solve_wave_model<real>(wm, s[0], s[1], s[2], m.data(), iters, (real)cf.tolerance,
                                                cf.check_step, cf.hold, perm);
        real* nb[3] = {s[perm[0]], s[perm[1]], s[perm[2]]};
        b[0] = nb[0]; b[1] = nb[1]; b[2] = nb[2];
        return done;
    };
    int it_idx = 0;
    auto record = [&](int done) {
        if (iters_out) iters_out[it_idx++] = done;
        if (tape) {
            tape->x.emplace_back(pos, pos + 3 * n);
            tape->phi.emplace_back(b[1], b[1] + ncell);
            tape->iters.push_back(done);
        }
    };
    // precalc (rrgsim_integration_rect.cuh:53-83); the monopole boundaries
    // take total mass 1 and the centre of mass at precalc
    deposit();
    if (cf.boundary == BC_MONOPOLE_DIRICHLET || cf.boundary == BC_MONOPOLE_RADIATIVE) {
        wm.bc.mass = 1.;
        wm.bc.com = center_of_mass_fixed<real>(n, pos, mass, cf.fixed_e);
        wm.bc.r_min = real(0.5) * dx;
    }
    int done = solve(cf.setup_iterations);
    record(done);
    gradient<real>(nx, dx, b[1], gacc.data());
    // sync: gather acceleration onto the particles
    gather<real>(n, pos, nx, dx, bcf, gacc.data(), acc.data());
    for (int s = 0; s < nsteps; ++s) {
        predict<real>(n, acc.data(), vel, pos, vp.data(), dtp);
        deposit();
        done = solve(cf.iterations);
        record(done);
        gradient<real>(nx, dx, b[1], gacc.data());
        gather<real>(n, pos, nx, dx, bcf, gacc.data(), acc.data());
        correct<real>(n, acc.data(), vel, vp.data(), dtp);
    }
    if (acc_out) std::memcpy(acc_out, acc.data(), sizeof(real) * 3 * n);
    if (phi_out) {
        std::memcpy(phi_out, b[1], sizeof(real) * ncell);               // curr
        std::memcpy(phi_out + ncell, b[0], sizeof(real) * ncell);       // prev
    }
    if (grid_acc_out) std::memcpy(grid_acc_out, gacc.data(), sizeof(real) * 3 * ncell);
    if (grav_out) ngp_phi<real>(n, pos, nx, dx, bcf, b[1], grav_out);
}

// ---------------------------------------------------------------------------
// Reverse-mode gradient of pm_run (no reference counterpart), FP64.  Pinned by
// central finite differences of pm_run in the tests.  The deposit's fixed-point
// rounding (2^-e) is treated as the identity.  Per grid build s = T .. 0
// (s = 0 is precalc + sync, which builds a_0):
//   a' = Interp(x', G),  G = -D phi_I / dx      -> Gbar += Interp^T abar, xbar += dInterp
//   phibar_I += (-1/dx) D^T Gbar                  (masked on zero-Dirichlet edges)
//   phi_{k+1} = B(K_m L phi_k + K_a phi_k + K_o phi_{k-1} + s) reversed:
//       mu_{k-1} = B(K_m L mu_k + K_a mu_k + K_o mu_{k+1}),  sbar = sum mu_k
//   s = 4 pi k_f (CIC(x', m) / dx^3 - rho_mean)  -> mbar += alpha CIC^T sbar, xbar += alpha m dCIC
//   Verlet (s > 0): xbar, vbar += dt xbar, abar = dt/2 vbar + dt^2/2 xbar,
//   where the gather's adjoint input is abar + dt/2 vbar.
// The potential adjoints (mu_k, mu_{k+1}) carry over between grid builds
// because the stencil sequence continues across steps.
// ---------------------------------------------------------------------------
void gradient_t(int nx, double dx, const double* gbar, double* phibar) {  // transpose of gradient()
    const int NX = nx;
    const double coef = -1. / dx;
    const size_t ncell = size_t(NX) * NX * NX;
    std::fill(phibar, phibar + ncell, 0.0);
    const size_t st[3] = {1, size_t(NX), size_t(NX) * NX};
    for (int iz = 0; iz < NX; ++iz)
        for (int iy = 0; iy < NX; ++iy)
            for (int ix = 0; ix < NX; ++ix) {
                const size_t c = AT(ix, iy, iz, NX);
                const int id[3] = {ix, iy, iz};
                for (int d = 0; d < 3; ++d) {
                    const double g = coef * gbar[3 * c + d];
                    if (id[d] == 0) { phibar[c + st[d]] += g; phibar[c] -= g; }
                    else if (id[d] == NX - 1) { phibar[c] += g; phibar[c - st[d]] -= g; }
                    else { phibar[c + st[d]] += 0.5 * g; phibar[c - st[d]] -= 0.5 * g; }
                }
            }
}

// Interpolation weights of gather() for one particle, with d(weight)/d(pos).
struct InterpW {
    size_t cell[8];
    double w[8], dw[8][3];
};

InterpW interp_weights(const double* p, int nx, double dx, double domain_min) {
    InterpW r;
    double f[3];
    int lo[3], hi[3];
    double wt[3];
    for (int d = 0; d < 3; ++d) {
        f[d] = (p[d] - domain_min) / dx - 0.5f;
        const int i0 = (int)std::floor((float)f[d]);
        wt[d] = f[d] - i0;
        lo[d] = std::max(0, std::min(i0, nx - 1));
        hi[d] = std::max(0, std::min(i0 + 1, nx - 1));
    }
    int q = 0;
    for (int k = 0; k <= 1; ++k)
        for (int j = 0; j <= 1; ++j)
            for (int i = 0; i <= 1; ++i, ++q) {
                const int sel[3] = {i, j, k};
                double w1[3], dw1[3];
                for (int d = 0; d < 3; ++d) {
                    w1[d] = sel[d] ? wt[d] : 1.0 - wt[d];
                    dw1[d] = (sel[d] ? 1.0 : -1.0) / dx;
                }
                r.cell[q] = AT(i ? hi[0] : lo[0], j ? hi[1] : lo[1], k ? hi[2] : lo[2], nx);
                r.w[q] = w1[0] * w1[1] * w1[2];
                r.dw[q][0] = dw1[0] * w1[1] * w1[2];
                r.dw[q][1] = w1[0] * dw1[1] * w1[2];
                r.dw[q][2] = w1[0] * w1[1] * dw1[2];
            }
    return r;
}

// CIC weights of cic_stencil() with their derivatives (zero where clamped).
InterpW cic_weights(const double* p, const Grid<double>& g) {
    const int NX = g.nx;
    const double inv_dx = 1.f / g.dx;
    const double hi_u = (NX - 1) - 1E-6;
    double u[3], du[3];
    int i0[3], i1[3];
    for (int d = 0; d < 3; ++d) {
        const double raw = (p[d] - g.domain_min) * inv_dx - 0.5f;
        u[d] = clampv<double>(raw, 0, hi_u);
        du[d] = (raw < 0 || raw > hi_u) ? 0.0 : inv_dx;
        i0[d] = (int)std::floor(u[d]);
        i1[d] = std::min(i0[d] + 1, NX - 1);
        u[d] -= i0[d];
    }
    InterpW r;
    int q = 0;
    for (int k = 0; k <= 1; ++k)
        for (int j = 0; j <= 1; ++j)
            for (int i = 0; i <= 1; ++i, ++q) {
                const int sel[3] = {i, j, k};
                double w1[3], dw1[3];
                for (int d = 0; d < 3; ++d) {
                    w1[d] = sel[d] ? u[d] : 1.0 - u[d];
                    dw1[d] = (sel[d] ? 1.0 : -1.0) * du[d];
                }
                r.cell[q] = AT(i ? i1[0] : i0[0], j ? i1[1] : i0[1], k ? i1[2] : i0[2], NX);
                r.w[q] = w1[0] * w1[1] * w1[2];
                r.dw[q][0] = dw1[0] * w1[1] * w1[2];
                r.dw[q][1] = w1[0] * dw1[1] * w1[2];
                r.dw[q][2] = w1[0] * w1[1] * dw1[2];
            }
    return r;
}

int pm_adjoint(const PmConfig& cf, int64_t n, const double* x0, const double* v0, const double* mass, int nsteps,
               const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) {
    if (cf.boundary != BC_PERIODIC && cf.boundary != BC_ZERO_DIRICHLET) return -1;  // monopole: not differentiated
    std::vector<double> pos(x0, x0 + 3 * n), vel(v0, v0 + 3 * n);
    PmTape<double> tape;
    pm_run<double>(cf, n, pos.data(), vel.data(), mass, nsteps, nullptr, nullptr, nullptr, nullptr, nullptr, &tape);
    const int nx = cf.nx;
    const size_t ncell = size_t(nx) * nx * nx;
    const double dx = cf.dx, dt = cf.dt;
    const double C = cf.wave_speed;
    const double DT = wave_dt<double>(dx, C);
    const Grid<double> g = make_grid<double>(nx, dx, cf.bc_frac);
    WaveModel<double> wm{};
    wm.g = g;
    wm.diss_base = cf.diss_base;
    wm.diss_extra = cf.diss_extra;
    wm.C = C;
    wm.DT = DT;
    wm.bc.type = cf.boundary;
    // per-cell coefficients (the absorbing layer makes them vary off-periodic)
    std::vector<double> Km(ncell), Ka(ncell), Ko(ncell), alpha(ncell);
    for (int iz = 0; iz < nx; ++iz)
        for (int iy = 0; iy < nx; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const WaveCoef<double> k = cell_coef(wm, ix, iy, iz);
                const size_t c = AT(ix, iy, iz, nx);
                Km[c] = k.k_main;
                Ka[c] = k.k_actual;
                Ko[c] = k.k_old;
                alpha[c] = 4 * kPi * k.k_f / (dx * dx * dx);  // d source / d cell mass
            }
    const bool periodic = cf.boundary == BC_PERIODIC;
    std::vector<double> X(3 * n, 0.0), V(3 * n, 0.0), A(3 * n, 0.0), GM(n, 0.0);
    if (gxT) std::memcpy(X.data(), gxT, sizeof(double) * 3 * n);
    if (gvT) std::memcpy(V.data(), gvT, sizeof(double) * 3 * n);
    std::vector<double> G(3 * ncell), Gb(3 * ncell), phib(ncell), sbar(ncell);
    std::vector<double> mu[3] = {std::vector<double>(ncell, 0.0), std::vector<double>(ncell, 0.0),
                                 std::vector<double>(ncell, 0.0)};
    int ip = 0, ic = 1;  // mu_{k+1}, mu_k
    auto edge = [&](size_t c) {
        const int ix = int(c % nx), iy = int((c / nx) % nx), iz = int(c / (size_t(nx) * nx));
        return ix == 0 || ix == nx - 1 || iy == 0 || iy == nx - 1 || iz == 0 || iz == nx - 1;
    };
    for (int s = nsteps; s >= 0; --s) {
        const double* x = tape.x[s].data();
        gradient<double>(nx, dx, tape.phi[s].data(), G.data());
        const double h = s > 0 ? 0.5 * dt : 0.0;
        std::fill(Gb.begin(), Gb.end(), 0.0);
        for (int64_t q = 0; q < n; ++q) {
            const InterpW W = interp_weights(x + 3 * q, nx, dx, g.domain_min);
            double at[3];
            for (int d = 0; d < 3; ++d) at[d] = A[3 * q + d] + h * V[3 * q + d];
            for (int c = 0; c < 8; ++c) {
                const double* gc = &G[3 * W.cell[c]];
                const double ga = gc[0] * at[0] + gc[1] * at[1] + gc[2] * at[2];
                for (int d = 0; d < 3; ++d) {
                    Gb[3 * W.cell[c] + d] += W.w[c] * at[d];
                    X[3 * q + d] += W.dw[c][d] * ga;
                }
            }
        }
        gradient_t(nx, dx, Gb.data(), phib.data());
        for (size_t c = 0; c < ncell; ++c)
            if (periodic || !edge(c)) mu[ic][c] += phib[c];
        std::fill(sbar.begin(), sbar.end(), 0.0);
        for (int it = 0; it < tape.iters[s]; ++it) {
            for (size_t c = 0; c < ncell; ++c) sbar[c] += mu[ic][c];
            const int in = 3 - ip - ic;
            // transposed iteration: mu_{k-1}[c] = B(sum_n Km[n] mu_k[n] + Ka[c] mu_k[c] + Ko[c] mu_{k+1}[c])
            const double* mk = mu[ic].data();
            for (int iz = 0; iz < nx; ++iz)
                for (int iy = 0; iy < nx; ++iy)
                    for (int ix = 0; ix < nx; ++ix) {
                        const size_t c = AT(ix, iy, iz, nx);
                        if (!periodic && edge(c)) {
                            mu[in][c] = 0.0;
                            continue;
                        }
                        const int xm = wrapi(ix - 1, nx), xp = wrapi(ix + 1, nx), ym = wrapi(iy - 1, nx),
                                  yp = wrapi(iy + 1, nx), zm = wrapi(iz - 1, nx), zp = wrapi(iz + 1, nx);
                        const size_t nb[6] = {AT(xm, iy, iz, nx), AT(xp, iy, iz, nx), AT(ix, ym, iz, nx),
                                              AT(ix, yp, iz, nx), AT(ix, iy, zm, nx), AT(ix, iy, zp, nx)};
                        double acc = 0;
                        for (int q = 0; q < 6; ++q) acc += Km[nb[q]] * mk[nb[q]];
                        mu[in][c] = acc + Ka[c] * mk[c] + Ko[c] * mu[ip][c];
                    }
            ip = ic;
            ic = in;
        }
        for (int64_t q = 0; q < n; ++q) {
            const InterpW W = cic_weights(x + 3 * q, g);
            double mb = 0, dm[3] = {0, 0, 0};
            for (int c = 0; c < 8; ++c) {
                const double sa = alpha[W.cell[c]] * sbar[W.cell[c]];
                mb += W.w[c] * sa;
                for (int d = 0; d < 3; ++d) dm[d] += W.dw[c][d] * sa;
            }
            GM[q] += mb;
            for (int d = 0; d < 3; ++d) X[3 * q + d] += mass[q] * dm[d];
        }
        if (s > 0) {
            for (int64_t i = 0; i < 3 * n; ++i) {
                const double vb = V[i];
                V[i] = vb + dt * X[i];
                A[i] = h * vb + h * dt * X[i];
            }
        }
    }
    std::memcpy(gx0, X.data(), sizeof(double) * 3 * n);
    std::memcpy(gv0, V.data(), sizeof(double) * 3 * n);
    std::memcpy(gm, GM.data(), sizeof(double) * n);
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (ctypes).  Suffix _f64 = real double, _f32 = real float.
// ---------------------------------------------------------------------------
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

// OpenMP threads of the parallel routines (the CPU baseline also times one
// core); the default is the OpenMP runtime's.
ORC_EXPORT void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
ORC_EXPORT int orc_get_threads() { return omp_get_max_threads(); }

#define ORC_INSTANTIATE(SUF, real)                                                                   \
    ORC_EXPORT void orc_accel_##SUF(int64_t n, const real* pos, const real* m, const real* s2,     \
                                    real* acc) { accel<real>(n, pos, m, s2, acc); }                \
    ORC_EXPORT void orc_accel_subset_##SUF(int64_t n, const real* pos, const real* m,              \
                                           const real* s2, int64_t k, const int64_t* idx,          \
                                           real* acc) { accel_subset<real>(n, pos, m, s2, k, idx, acc); } \
    ORC_EXPORT void orc_grav_##SUF(int64_t n, const real* pos, const real* m, const real* s2,      \
                                   real* out) { grav<real>(n, pos, m, s2, out); }                  \
    ORC_EXPORT void orc_predict_##SUF(int64_t n, const real* acc, const real* vel, real* pos,      \
                                      real* vp, real dt) { predict<real>(n, acc, vel, pos, vp, dt); } \
    ORC_EXPORT void orc_correct_##SUF(int64_t n, const real* acc, real* vel, const real* vp,       \
                                      real dt) { correct<real>(n, acc, vel, vp, dt); }             \
    ORC_EXPORT void orc_conservation_##SUF(int64_t n, const real* pos, const real* vel,            \
                                           const real* m, const real* g, real* out) {              \
        conservation<real>(n, pos, vel, m, g, out); }                                              \
    ORC_EXPORT void orc_cic_float_##SUF(int64_t n, const real* pos, const real* m, int nx,         \
                                        real dx, real bcf, real* grid) {                           \
        cic_deposit_float<real>(n, pos, m, nx, dx, bcf, grid); }                                   \
    ORC_EXPORT void orc_cic_fixed_##SUF(int64_t n, const real* pos, const real* m, int nx,         \
                                        real dx, real bcf, int e, int64_t* grid) {                 \
        cic_deposit_fixed<real>(n, pos, m, nx, dx, bcf, e, grid); }                                \
    ORC_EXPORT void orc_fixed_to_real_##SUF(int64_t ncell, const int64_t* g, int e, real* out) {   \
        fixed_to_real<real>((size_t)ncell, g, e, out); }                                           \
    ORC_EXPORT void orc_wave_iteration_##SUF(int nx, real dx, real diss, real c, real dt,          \
                                             int periodic, real rho_mean, real* pn, const real* p, \
                                             const real* po, const real* m) {                      \
        wave_iteration<real>(nx, dx, diss, c, dt, periodic, rho_mean, pn, p, po, m); }             \
    ORC_EXPORT void orc_wave_iteration_bc_##SUF(int nx, real dx, real bc_frac, real diss_base,          \
                                                real diss_extra, real c, real dt, int boundary,            \
                                                real rho_mean, real bc_mass, const real* com, real r_min,  \
                                                real* pn, const real* p, const real* po, const real* m) {  \
        WaveModel<real> w{};                                                                        \
        w.g = make_grid<real>(nx, dx, bc_frac);                                                     \
        w.diss_base = diss_base; w.diss_extra = diss_extra; w.C = c; w.DT = dt;                     \
        w.bc.type = boundary; w.bc.rho_mean = rho_mean; w.bc.mass = bc_mass;                        \
        w.bc.com = V3<real>{com[0], com[1], com[2]}; w.bc.r_min = r_min;                            \
        wave_iteration_model<real>(w, pn, p, po, m); }                                              \
    ORC_EXPORT int orc_solve_wave_##SUF(int nx, real dx, real diss, real c, real dt, int periodic, \
                                        real rho_mean, real* prev, real* curr, real* next,         \
                                        const real* m, int max_iter, real tol, int check_step,     \
                                        int hold, int* perm) {                                     \
        return solve_wave<real>(nx, dx, diss, c, dt, periodic, rho_mean, prev, curr, next, m,      \
                                max_iter, tol, check_step, hold, perm); }                          \
    ORC_EXPORT double orc_steady_err_##SUF(int64_t ncell, const real* un, const real* uo) {        \
        return steady_err<real>((size_t)ncell, un, uo); }                                          \
    ORC_EXPORT void orc_gradient_##SUF(int nx, real dx, const real* phi, real* acc) {              \
        gradient<real>(nx, dx, phi, acc); }                                                        \
    ORC_EXPORT void orc_gather_##SUF(int64_t n, const real* pos, int nx, real dx, real bcf,        \
                                     const real* field, real* out) {                               \
        gather<real>(n, pos, nx, dx, bcf, field, out); }                                           \
    ORC_EXPORT void orc_ngp_phi_##SUF(int64_t n, const real* pos, int nx, real dx, real bcf,       \
                                      const real* phi, real* out) {                                \
        ngp_phi<real>(n, pos, nx, dx, bcf, phi, out); }                                            \
    ORC_EXPORT void orc_ngp_bin_##SUF(int64_t n, const real* pos, int nx, real dx, real bcf,         \
                                      int* cell_of, int* count, int* start, int* order) {          \
        ngp_bin<real>(n, pos, nx, dx, bcf, cell_of, count, start, order); }                        \
    ORC_EXPORT void orc_ngp_project_##SUF(int nx, const int* count, const int* start,               \
                                          const int* order, const real* grav, const real* acc,     \
                                          real* cell_phi, real* cell_acc) {                        \
        ngp_project<real>(nx, count, start, order, grav, acc, cell_phi, cell_acc); }               \
    ORC_EXPORT real orc_wave_dt_##SUF(real dx, real c) { return wave_dt<real>(dx, c); }            \
    ORC_EXPORT real orc_rho_mean_##SUF(real total_mass, int nx, real dx) {                         \
        return rho_mean_periodic<real>(total_mass, nx, dx); }                                      \
    ORC_EXPORT void orc_wave_coef_##SUF(real diss, real c, real dt, real dx, real* out4) {         \
        WaveCoef<real> k = wave_coef<real>(diss, c, dt, dx);                                       \
        out4[0] = k.k_main; out4[1] = k.k_f; out4[2] = k.k_actual; out4[3] = k.k_old; }            \
    ORC_EXPORT void orc_nbody_run_##SUF(int64_t n, real* pos, real* vel, const real* m,            \
                                        const real* s2, real dt, int nsteps, real* acc,            \
                                        real* gr) {                                                \
        nbody_run<real>(n, pos, vel, m, s2, dt, nsteps, acc, gr); }                                \
    ORC_EXPORT void orc_pm_run_##SUF(const PmConfig* cf, int64_t n, real* pos, real* vel,          \
                                     const real* m, int nsteps, real* acc, real* phi2,             \
                                     real* gacc, real* gr, int* iters) {                           \
        pm_run<real>(*cf, n, pos, vel, m, nsteps, acc, phi2, gacc, gr, iters); }

ORC_EXPORT void orc_nbody_adjoint_f64(int64_t n, const double* x0, const double* v0, const double* m,
                                      const double* s2, double dt, int nsteps, const double* gxT,
                                      const double* gvT, double* gx0, double* gv0, double* gm) {
    nbody_adjoint(n, x0, v0, m, s2, dt, nsteps, gxT, gvT, gx0, gv0, gm);
}
ORC_EXPORT int orc_pm_adjoint_f64(const PmConfig* cf, int64_t n, const double* x0, const double* v0,
                                  const double* m, int nsteps, const double* gxT, const double* gvT,
                                  double* gx0, double* gv0, double* gm) {
    return pm_adjoint(*cf, n, x0, v0, m, nsteps, gxT, gvT, gx0, gv0, gm);
}
ORC_EXPORT void orc_center_of_mass_f64(int64_t n, const double* pos, const double* m, int e, double* out) {
    const V3<double> c = center_of_mass_fixed<double>(n, pos, m, e);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
ORC_EXPORT void orc_gradient_t_f64(int nx, double dx, const double* gbar, double* phibar) {
    gradient_t(nx, dx, gbar, phibar);
}
ORC_EXPORT void orc_vjp_f64(int64_t n, const double* x, const double* m, const double* s2, const double* u,
                            double* g, double* gm) {
    vjp(n, x, m, s2, u, g, gm);
}

ORC_INSTANTIATE(f64, double)
ORC_INSTANTIATE(f32, float)
