// Internal declarations shared by the librrg.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "rrg.h"
#include "rrg_error.hpp"

namespace rrg {

#define RRG_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t rrg_e_ = (call);                                                            \
        if (rrg_e_ != hipSuccess)                                                              \
            ::rrg::fail(RRG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(rrg_e_) +   \
                                         " (" __FILE__ ":" + std::to_string(__LINE__) + ")");  \
    } while (0)

#define RRG_LAUNCH_CHECK() RRG_HIP(hipGetLastError())

// Owning device array.  Allocation failures surface as rrg::Error.
template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        if (count) RRG_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    size_t bytes() const { return n * sizeof(T); }
};

// Stores that write through the XCD's L2 to memory: buffer stores with the
// device-scope cache bit (sc1).  A grid that the next launch reads, from any
// XCD, reaches memory either way; written through while the kernel runs,
// its lines need not be flushed from L2 at the kernel boundary, which
// otherwise waits for the whole output there (the four-iteration 128^3 FP32
// stencil launch: 12.5 -> 11.3 us).  Same values, same bits: only the cache
// policy differs.  It pays only where each store instruction writes a span
// that is contiguous across the wave: runs of more than 16 bytes per lane
// take several instructions, each writing part of every cache line, which
// written through cost several times more (FP64 512^3 four-cell runs: 573 ->
// 976 us per iteration), so such runs are stored with the default policy;
// and on 64^3 grids, whose outputs flush quickly, it costs 5-14%.
//
// `base` is wave-uniform (it becomes the buffer resource, held in scalar
// registers) and `bytes` its extent; each lane writes at byte offset `off`
// (offsets at or beyond `bytes` are dropped by the hardware).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t through_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), 0x00020000);  // gfx9 raw
}
constexpr int kPolicyThrough = 16;  // sc1
// N consecutive values at byte offset off (16-byte runs where N * size allows;
// off must then be a multiple of 16, else of 8 or 4).
template <class real, int N>
__device__ __forceinline__ void store_through_run(__amdgpu_buffer_rsrc_t r, unsigned off, const real (&v)[N]) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    constexpr int W = int(sizeof(real) / 4);  // dwords per value
    constexpr int kCachePolicySC1 = N * sizeof(real) <= 16 ? kPolicyThrough : 0;
    unsigned d[N * W];
    __builtin_memcpy(d, v, sizeof(d));
    int i = 0;
#pragma unroll
    for (; i + 4 <= N * W; i += 4) {
        const u4 q = {d[i], d[i + 1], d[i + 2], d[i + 3]};
        __builtin_amdgcn_raw_buffer_store_b128(q, r, int(off + 4u * unsigned(i)), 0, kCachePolicySC1);
    }
    if constexpr ((N * W) % 4 >= 2) {
        const u2 q = {d[i], d[i + 1]};
        __builtin_amdgcn_raw_buffer_store_b64(q, r, int(off + 4u * unsigned(i)), 0, kCachePolicySC1);
        i += 2;
    }
    if constexpr ((N * W) % 2 == 1)
        __builtin_amdgcn_raw_buffer_store_b32(d[N * W - 1], r, int(off + 4u * unsigned(N * W - 1)), 0, kCachePolicySC1);
}
template <class real>
__device__ __forceinline__ void store_through(__amdgpu_buffer_rsrc_t r, unsigned off, real v) {
    const real a[1] = {v};
    store_through_run<real, 1>(r, off, a);
}

// Particle state in structure-of-arrays form: component c of a vector array
// starts at c * n.  `w` is the predicted velocity v* of the predictor step.
template <class real> struct PState {
    real* x;   // 3n
    real* v;   // 3n
    real* a;   // 3n
    real* w;   // 3n
    real* m;   // n
    real* s2;  // n (direct sum only)
    real* g;   // n, particle potential
    int64_t n;
};

// ----------------------------------------------------------------- direct sum
namespace nbody {

// Grid-form launch shape of a direct-sum evaluation: `ib` targets per lane
// (tiles of 256*ib targets), the source range split into `slices` contiguous
// slices of `slice_len` sources; partial results are summed in slice order.
struct Plan {
    int ib = 1;
    int slices = 1;
    int64_t slice_len = 0;
    int64_t tiles = 0;  // workgroups along the target axis
};

// ib <= 0 selects the default for (n, kind); target_workgroups <= 0 the default.
Plan plan_direct(int64_t n, int force_kind, int target_workgroups = 0, int ib = 0);
bool is_fp32_kind(int force_kind);
bool valid_ib(int force_kind, int ib);
int targets_per_tile(int ib);

// Force arithmetic of one evaluation: kind (RRG_FORCE_*) and, when every
// particle has the same softening, hh = h_i + h_j (h = 0.5 * soft2).
struct Force {
    int kind = RRG_FORCE_F64;
    bool uniform = false;
    double hh = 0.0;
};

// One item of the queue form: targets [i0, i1) against sources [j0, j1),
// written to partial slice `slice`.
struct WorkItem {
    int i0, i1, j0, j1, slice, pad0, pad1, pad2;
};

// Queue-form launch: `workgroups` persistent workgroups over `nitems` items.
struct QueuePlan {
    int ib = 1;
    int nitems = 0;
    int workgroups = 0;
};

// Packed FP32 source records for the FP32 kinds:
//   A = (x_hi, y_hi, z_hi, m), B = (x_lo, y_lo, z_lo, 0.5*soft2).
void pack_mixed(hipStream_t s, const PState<double>& p, float4* A, float4* B);

// Acceleration partials: out[(slice*3 + c)*n + i].  A/B are used by the
// FP32 kinds only.
void accel(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
           double* out);
// Potential partials: out[slice*n + i].
void grav(hipStream_t s, const Plan& pl, const Force& f, const PState<double>& p, const float4* A, const float4* B,
          double* out);
// Queue form (acceleration or potential).  counters = 2 ints, zero before
// the first launch; the launch leaves them at zero.
void queue(hipStream_t s, const QueuePlan& q, const Force& f, bool potential, const PState<double>& p,
           const float4* A, const float4* B, const WorkItem* items, int* counters, double* out);
void reset_counters(hipStream_t s, int* counters);

// Sum the slice partials of `comps` components into dst (comps*n); the slice
// count is `slices`, or ns[i] per particle when ns is non-null.
void reduce_slices(hipStream_t s, int64_t n, int comps, int slices, const unsigned char* ns, const double* part,
                   double* dst);

// predict_step_ (nbody.cu:122-152); with A/B non-null also repacks the
// FP32 source records from the new positions.
template <class real>
void predict(hipStream_t s, const PState<real>& p, real dt, float4* A, float4* B);
// correct_step_ (nbody.cu:159-181) on a[] (after summing the slice partials
// from `part` into a[] when part != nullptr).
void correct_reduce(hipStream_t s, const PState<double>& p, double dt, int slices, const unsigned char* ns,
                    const double* part);
// correct_reduce followed by predict (the next step's) in one launch, the
// same values as the two.
void correct_predict(hipStream_t s, const PState<double>& p, double dt, int slices, const unsigned char* ns,
                     const double* part, float4* A, float4* B);
template <class real> void correct(hipStream_t s, const PState<real>& p, real dt);

// Conservation sums (nbody.cu:246-291) as FP64 per-block partials:
// partial[b*8 + q], q = {2 Ek, 2 Ep, Px, Py, Pz, Lx, Ly, Lz}.  Returns the
// number of blocks written (at most conservation_blocks(n)).
int64_t conservation_blocks(int64_t n);
template <class real> int64_t conservation(hipStream_t s, const PState<real>& p, double* partial);
// Same sums per system of an ensemble: out[g*8 + q], offsets on the device.
void conservation_segments(hipStream_t s, const PState<double>& p, int k, const int64_t* offsets, double* out);

}  // namespace nbody

// ----------------------------------------------------------------- adjoint
namespace adjoint {

// J^T u and the mass gradient at positions X (SoA) for the force kind `kind`;
// out[(slice*4 + c)*n + i], c = 0..2 (J^T u), 3 (dL/dm).  FP32 kinds read the
// packed records A/B (positions) and Uf (u as float4); FP64 kinds X and U.
void vjp(hipStream_t s, const nbody::Plan& pl, int kind, int64_t n, const double* X, const double* M,
         const double* S2, const double* U, const float4* A, const float4* B, const float4* Uf, double* out);
// At = Aa + dt V / 2 (and its packed FP32 copy when Uf is non-null).
void pre(hipStream_t s, int64_t n, double dt, const double* V, const double* Aa, double* At, float4* Uf);
// Sum VJP slices; X += J^T A~, gm += dm; unless `last`: V += dt X, Aa = dt V_old/2 + dt^2 X/2.
void post(hipStream_t s, int64_t n, double dt, int slices, const double* part, double* X, double* V, double* Aa,
          double* gm, bool last);
void pack_pos(hipStream_t s, int64_t n, const double* X, const double* M, const double* S2, float4* A, float4* B);

}  // namespace adjoint

// ------------------------------------------------------------ particle-mesh
namespace layout {
// Caller rows (n,3) FP64, in caller order, to and from the state's SoA
// arrays; id (device, may be null = identity) maps slot s to caller index.
template <class real> void rows_in(hipStream_t s, int64_t n, const double* rows, real* soa, const int* id);
template <class real> void rows_out(hipStream_t s, int64_t n, const real* soa, double* rows, const int* id);
template <class real> void scalar_out(hipStream_t s, int64_t n, const real* v, double* out, const int* id);
}  // namespace layout

namespace pm {

// Grid geometry in `real`, derived like GridInfo (io_json.cpp:28-40).
template <class real> struct Geo {
    int nx;
    real dx;
    real domain_min;
    real inv_dx;  // 1.f / dx  (rm_p2mesh.cu:185)
    real upper;   // (nx - 1) - 1e-6 (rm_p2mesh.cu:196-199)
};
template <class real> Geo<real> make_geo(int nx, double dx, double bc_frac);

// Wave-iteration coefficients (wave.cu:164-171), computed on the host in `real`.
template <class real> struct WaveK {
    real k_main, k_f, k_actual, k_old;
};
template <class real> WaveK<real> wave_coef(real diss, real c, real dt, real dx);

// What the stencil reads per cell on a non-periodic grid (wave.cu:18-193):
// the absorbing layer (x_in = sim_l/2, diss_extra) makes the coefficients
// vary per cell, and the face cells take the boundary type's value.
template <class real> struct WaveModel {
    int nx;
    int bc;                 // RRG_BC_* (the reference's BoundaryType values)
    real dx, domain_min;
    real x_in;              // 0.5 * sim_l, sim_l = (1 - bc_frac) * nx * dx
    real diss_base, diss_extra, C, DT;
    real mass, r_min;       // monopole: total mass, guard radius 0.5 dx
    real com[3];            // monopole centre
};
template <class real>
WaveModel<real> make_model(int nx, double dx, double bc_frac, double diss_base, double diss_extra, double wave_speed,
                           double wave_dt, int bc);

// Centre of mass for the monopole boundaries, order-independent: every m*x
// (FP64 product) and m rounded once to fixed point (2^E, 2^e) and summed in
// int64; E keeps n * max|m x| below 2^61.  `scratch` holds 5 int64.
template <class real>
void center_of_mass(hipStream_t s, const PState<real>& p, int e_mass, unsigned long long* scratch, real out[3]);

// predict_step_ fused with deposit_cell_mass_CIC_ at the new position into an
// int64 fixed-point grid (scale = 2^e, exact).  When `predict` is false only
// the deposit runs (precalc).  max_q: an upper bound of |m| * scale over the
// particles (it selects the cheaper exact rounding below 2^51).
template <class real>
void predict_deposit(hipStream_t s, const PState<real>& p, real dt, bool predict, const Geo<real>& g,
                     double scale, unsigned long long* mfix, double max_q);

// Spatial ordering: keys[i] = CIC base cell of particle i, idx[i] = i.
template <class real>
void cell_keys(hipStream_t s, const PState<real>& p, const Geo<real>& g, unsigned* keys, unsigned* idx);
// dst[i] = src[idx[i]] for x, v, a, v*, m, phi_p and the particle ids.
template <class real>
void permute(hipStream_t s, const PState<real>& src, const PState<real>& dst, const unsigned* idx, const int* id_src,
             int* id_dst);
// Key/value radix sort (hipCUB); with temp == nullptr only sets temp_bytes.
void radix_sort_pairs(hipStream_t s, void* temp, size_t& temp_bytes, const unsigned* keys_in, unsigned* keys_out,
                      const unsigned* vals_in, unsigned* vals_out, int64_t n, int end_bit);

// Fixed-point grid -> stencil source term K_F * 4 pi G (m/dx^3 - rho_mean)
// and, when `mass` is non-null, the cell mass (real); clears the fixed-point
// grid for the next deposit.  On a non-periodic grid (`model` non-null) K_F
// is the cell's own.  With a uniform K_F (`model` null) and `occ` non-null
// (nx^3 a multiple of 64) it also writes, per 64-cell segment of the grid,
// occ[c / 64] = 1 if a cell of the segment holds mass and 0 if none does,
// and *src_empty = the source term of an empty cell, which every cell of an
// unmarked segment holds.
template <class real>
void fixed_to_source(hipStream_t s, int nx, unsigned long long* mfix, double inv_scale, real dx,
                     real rho_mean, bool periodic, real k_f, real* mass, real* src,
                     const WaveModel<real>* model = nullptr, int* occ = nullptr, real* src_empty = nullptr);

// Cell mass (real) -> stencil source term, for callers that hold a mass grid.
template <class real>
void mass_to_source(hipStream_t s, int nx, const real* mass, real dx, real rho_mean, bool periodic, real k_f,
                    real* src, const WaveModel<real>* model = nullptr);

// Slots of a checking launch's max-reduction: maxbits holds 2 kCheckSlots
// bit patterns of `real` (max|new - cur| in [0, kCheckSlots), max|new| in
// [kCheckSlots, 2 kCheckSlots)), zero before the launch; the maxima are the
// maxima over the slots.
constexpr int kCheckSlots = 256;

// One wave_diss_iteration over the grid: uniform coefficients `k` on a
// periodic grid, per-cell coefficients and boundary values from `m`
// otherwise.  When `maxbits` is non-null the kernel also accumulates
// max|new - cur| and max|new| as bit patterns into its slots (above).
// `halt` (optional): a device flag; when it holds nonzero at launch time the
// kernel returns without touching any grid.
template <class real>
void wave_iteration(hipStream_t s, const WaveModel<real>& m, const WaveK<real>& k, real* pn, const real* p,
                    const real* po, const real* src, void* maxbits, const int* halt = nullptr);

// `levels` (2 to 4) wave_diss_iterations in one launch, (p = phi^n,
// po = phi^{n-1}) -> (qa = phi^{n+levels-1}, qb = phi^{n+levels}), bitwise
// equal to that many wave_iteration calls; qa and qb must not alias p, po or
// src.  wave_fused_levels(m, r) is the number of iterations the next launch
// should run when r remain: 4, 3 or 2 where a launch shape exists (periodic
// grids of the sizes it lists), else 1 (one wave_iteration).
template <class real> int wave_fused_levels(const WaveModel<real>& m, int remaining);
// Whether launches of `levels` iterations can pass their two levels to the
// next such launch as one interleaved pair grid of 2 nx^3 reals (phi^n and
// phi^(n-1) of each two-cell block side by side): 128^3 FP32, four levels.
template <class real> bool wave_fused_pairs(const WaveModel<real>& m, int levels);
// With `occ` and `src_empty` from fixed_to_source (optional), the source
// rows of unmarked segments are not read from src: their cells hold
// *src_empty.  pin: p is a pair grid (po unused); pout: the two output levels
// go to the pair grid qa (qb unused); either needs wave_fused_pairs.
// maxbits (split grids only): a checking launch, which reduces
// max|phi^(n+levels) - phi^(n+levels-1)| and max|phi^(n+levels)| into the
// slots of maxbits like wave_iteration's check of its one level.
template <class real>
void wave_fused(hipStream_t s, const WaveModel<real>& m, const WaveK<real>& k, int levels, real* qa, real* qb,
                const real* p, const real* po, const real* src, const int* halt = nullptr, const int* occ = nullptr,
                const real* src_empty = nullptr, bool pin = false, bool pout = false, void* maxbits = nullptr);

// Steady-state check of a solve decided on the device (see pm_kernels.hip):
// `done` stops the solve's later stencil launches (pass &ctl->done as their
// halt flag); at convergence the block records the iteration count and the
// buffers holding phi^(n-1), phi^n.
struct SolveCtl {
    int done, hold, iters, ip, ic, pad[3];
};
template <class real> struct PhiBufs {
    real* p[4];
};
// done = hold = 0, maxbits = 0.
template <class real> void solve_init(hipStream_t s, SolveCtl* c, void* maxbits);
// After the checking launch that completes `iters` iterations with
// phi^(n-1), phi^n in buffers (ip, ic): apply the hold rule and clear maxbits.
template <class real>
void solve_decide(hipStream_t s, SolveCtl* c, void* maxbits, real tol, int hold_need, int iters, int ip, int ic);
// Converged: move the two live levels into buffers (ip, ic).
template <class real>
void solve_fixup(hipStream_t s, const SolveCtl* c, const PhiBufs<real>& b, int ip, int ic, int64_t ncell);

// The buffer schedule of `iters` unchecked iterations of solve_wave_equation
// (wave.cu:253-272), (ip, ic) = (phi^{n-1}, phi^n): levels(r) iterations
// per launch while r remain.  single(out, cur, prev) and
// multi(levels, out1, out2, cur, prev) issue the launches; (ip, ic) end on
// the last two levels.  Outputs are grids other than the two inputs, so four
// grids are needed when levels() can exceed 1, three otherwise.
template <class Levels, class Single, class Multi>
void wave_schedule(int iters, Levels&& levels, int& ip, int& ic, Single&& single, Multi&& multi) {
    auto free_after = [&](int from) {
        int f = from;
        while (f == ip || f == ic) ++f;
        return f;
    };
    while (iters > 0) {
        const int L = levels(iters);
        if (L >= 2) {
            const int f1 = free_after(0), f2 = free_after(f1 + 1);
            multi(L, f1, f2, ic, ip);
            ip = f1;
            ic = f2;
        } else {
            const int f = free_after(0);
            single(f, ic, ip);
            ip = ic;
            ic = f;
        }
        iters -= L >= 2 ? L : 1;
    }
}

// Per-cell K_MAIN, K_ACTUAL, K_OLD and d(source)/d(cell mass) = 4 pi K_F / dx^3
// of a non-periodic grid (for the transposed stencil).
template <class real>
void cell_coefficients(hipStream_t s, const WaveModel<real>& m, real* km, real* ka, real* ko, real* alpha);

// calc_acceleration_field_ into a padded (x, y, z, 0) real4 grid.
template <class real> void gradient(hipStream_t s, int nx, real dx, const real* phi, real* acc4);

// interpolate_acceleration_to_particles_kernel into p.a, followed by the
// correct_step_ kick when `kick` is true.
template <class real>
void gather(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* acc4, real dt, bool kick);

// gradient + gather in one launch: -grad phi is computed from phi at the
// cells the particles interpolate from (bitwise the same accelerations), so
// no gradient grid is written.
template <class real>
void gather_phi(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* phi, real dt, bool kick);

// NGP phi projection onto p.g (rm_p2mesh.cu:34-53, rm_mesh2p.cu:29-38).
template <class real> void ngp_phi(hipStream_t s, const PState<real>& p, const Geo<real>& g, const real* phi);

// Standalone max|a-b|, max|a| reduction into the slots of `maxbits` (see kCheckSlots).
template <class real> void max_diff(hipStream_t s, int64_t ncell, const real* a, const real* b, void* maxbits);

// NGP counting sort of the particles by cell (the binning of
// rm_p2mesh.cu:25-126 with the intended exclusive prefix): count[c], start[c]
// and order[] (particle indices sorted by cell, ascending within a cell).
struct NgpBins {
    unsigned *keys = nullptr, *keys2 = nullptr, *idx = nullptr, *order = nullptr;  // n each
    int *count = nullptr, *start = nullptr;                                       // ncell each
    void* temp = nullptr;
    size_t temp_bytes = 0;
};
size_t ngp_temp_bytes(int64_t n, int nx);
void exclusive_scan_i32(hipStream_t s, void* temp, size_t& temp_bytes, const int* in, int* out, int64_t n);
template <class real> void ngp_bin(hipStream_t s, const PState<real>& p, const Geo<real>& g, NgpBins& b);
// Cell averages of the particle potential (grav, n) and acceleration (acc, SoA
// 3n) into cell_phi (ncell) and cell_acc (ncell x 3).
template <class real>
void ngp_project(hipStream_t s, int nx, const NgpBins& b, int64_t n, const real* grav, const real* acc, real* cell_phi,
                 real* cell_acc);
// Fixed-point grid -> cell mass; clears the fixed-point grid.
template <class real> void fixed_to_mass(hipStream_t s, int nx, unsigned long long* mfix, double inv_scale, real* mass);

// Fill kernel (used instead of hipMemsetAsync so the work stays capturable).
void fill_u64(hipStream_t s, unsigned long long* p, size_t count, unsigned long long value);

// Reverse sweep of the PM step (pm_adjoint_kernels.hip).  Particle arrays SoA
// (n per component), particle adjoints FP64; Gq = three int64 fixed-point
// grids (x, y, z) of resolution 2^-sc->e, zero on entry.
// Fixed-point scale of the interpolation adjoint, found on the device; zero
// before the first use (each use leaves `maxbits` zero).
struct AdjScale {
    unsigned long long maxbits;
    int e, pad;
};
// Interpolation adjoint at x': Xb += d(Interp)/dx' . G, Gq += Interp^T (Ab + h Vb).
template <class real>
void adj_interp(hipStream_t s, int64_t n, const real* x, const Geo<real>& g, const real* G4, double h, double* Xb,
                const double* Vb, const double* Ab, long long* Gq, AdjScale* sc);
// mu += B(-1/dx D^T Gq) (B masks the faces of zero-Dirichlet grids).
template <class real>
void adj_gradient_t(hipStream_t s, int nx, real dx, bool periodic, const long long* Gq, const AdjScale* sc, real* mu);
// One transposed stencil iteration pn = B(K_m L p + K_a p + K_o po) with
// sbar = p (first) or sbar += p.  With per-cell coefficient grids (km
// non-null, non-periodic grids) pn = B(L(km p) + ka p + ko po).
template <class real>
void adj_wave(hipStream_t s, int nx, bool periodic, const WaveK<real>& k, real* pn, const real* p, const real* po,
              real* sbar, bool first, const real* km = nullptr, const real* ka = nullptr, const real* ko = nullptr);
// `levels` (2 or 3) periodic transposed iterations in one launch, (p = mu_k,
// po = mu_{k+1}) -> (qa = mu_{k-levels+1}, qb = mu_{k-levels}), with sbar
// accumulated as by that many adj_wave calls (bitwise).  adj_wave_levels is
// wave_fused_levels for the transposed stencil.
template <class real> int adj_wave_levels(int nx, bool periodic, int remaining);
template <class real>
void adj_wave_fused(hipStream_t s, int nx, const WaveK<real>& k, int levels, real* qa, real* qb, const real* p,
                    const real* po, real* sbar, bool first);
// CIC adjoint at x' (gm += alpha CIC^T sbar, Xb += alpha m dCIC^T sbar; alpha
// per cell when `alpha_cell` is non-null) and, unless `closing`, the
// transposed Verlet update of (Xb, Vb, Ab).
template <class real>
void adj_cic(hipStream_t s, int64_t n, const real* x, const real* m, const Geo<real>& g, const real* sbar,
             double alpha, double dt, bool closing, double* Xb, double* Vb, double* Ab, double* gm,
             const real* alpha_cell = nullptr);

}  // namespace pm

}  // namespace rrg
