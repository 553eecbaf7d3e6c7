/*
 * rrg.h -- C ABI of the MI355X (gfx950) RRGalaxSim hot-path library, librrg.so.
 *
 * The library replaces the reference solver's per-step hot path:
 *   - the tiled direct-sum gravity kernel and the velocity-Verlet
 *     predictor/corrector (solver/lib/rrgsim_nbody/src/nbody.cu),
 *   - CIC mass deposition (solver/lib/rrgsim_rect_mesh/src/rm_p2mesh.cu),
 *   - the damped-wave potential relaxation (solver/lib/rrgsim_wave/src/wave.cu),
 *   - the -grad(phi) field and trilinear force gather
 *     (rm_p2mesh.cu, solver/lib/rrgsim_rect_mesh/src/rm_mesh2p.cu),
 * behind the reference's params.json-driven solver (solver/app/src/rrgsim.cu,
 * solver/app/src/rrgsim_integration.cu).
 *
 * Conventions
 *   - Every function returns an int status: RRG_OK (0) or a negative code.  No
 *     C++ exception crosses the ABI; rrg_last_error() gives the message of the
 *     calling thread's last failure.
 *   - Host buffers are owned by the caller and copied in/out.  Vectors are AoS
 *     x,y,z triples (the reference's real3 arrays, co_types.h:4-18).
 *   - A context owns all its device memory and one HIP stream.  A context is
 *     thread-compatible (one host thread at a time); independent contexts may
 *     be used concurrently.  There are no process-wide parameters (the
 *     reference keeps them in __constant__ singletons, co_device_structs.cu:7-13).
 *   - Entry points ending in _dev take device pointers and a hipStream_t
 *     (passed as void*), for callers that already hold device memory.
 */
#ifndef RRG_H
#define RRG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRG_OK 0
#define RRG_ERR_INVALID (-1) /* bad argument or unsupported combination */
#define RRG_ERR_HIP (-2)     /* HIP runtime error (launch, allocation, copy) */
#define RRG_ERR_IO (-3)      /* file missing or malformed */
#define RRG_ERR_STATE (-4)   /* call out of order (e.g. step before precalc) */

/* Precision of particle state and grid: the reference's compile-time `real`
 * switch (co_types.h:4-18) made a run-time choice. */
enum { RRG_REAL_F64 = 0, RRG_REAL_F32 = 1 };

/* Direct-sum force arithmetic (N-body model).
 *   RRG_FORCE_F64       : FP64 pair math with the device library's FP64
 *                         rsqrt (default; the fastest FP64 variant).
 *   RRG_FORCE_MIXED     : FP32 pair math on hi/lo-split positions
 *                         (d = (hi_j - hi_i) + (lo_j - lo_i)) with compensated
 *                         (two-sum) per-particle accumulation, FP64 state.
 *   RRG_FORCE_FP32      : FP32 pair math on FP32-rounded positions with the
 *                         same compensated accumulation; cheaper, larger
 *                         error on close pairs.
 *   RRG_FORCE_F64_EXACT : FP64 pair math with IEEE 1./sqrt, the reference
 *                         kernel's arithmetic (nbody.cu:50-54).
 * The FP32 kinds have no counterpart in the reference. */
enum { RRG_FORCE_F64 = 0, RRG_FORCE_MIXED = 1, RRG_FORCE_FP32 = 2, RRG_FORCE_F64_EXACT = 3 };

/* Wave-equation boundary types (co_boundary.cuh:4-9, same values).  The
 * reference's live configuration is periodic (rrgsim_integration_rect.cuh:65-69).
 * Off-periodic grids carry the absorbing layer (grid bc_frac, wave diss_extra;
 * wave.cu:27-33, 155-161); the monopole types take total mass 1 and the centre
 * of mass at precalc (the reference's disabled call, rrgsim_integration_rect.cuh:57-63):
 * phi = -M/r on the faces (Dirichlet) or the radiation condition for
 * phi + M/r (wave.cu:36-104). */
enum {
    RRG_BC_ZERO_DIRICHLET = 0,
    RRG_BC_PERIODIC = 1,
    RRG_BC_MONOPOLE_DIRICHLET = 2,
    RRG_BC_MONOPOLE_RADIATIVE = 3
};

/* sim_params section (co_sim_params.h:6-13; io_json.cpp:13-21). */
typedef struct {
    double time_max;
    double dt_save;
    double dt_dynamics;
    uint64_t saves_without_dump; /* default: UINT64_MAX */
} rrg_sim_params;

/* grid_params section (co_info_structs.h:16-35; io_json.cpp:28-40). */
typedef struct {
    int32_t nx;
    double dx;
    double bc_frac;
} rrg_grid_params;

/* wave_params section (co_info_structs.h:37-69; io_json.cpp:53-99);
 * `dt` is derived by rrg_params_finalize (io_json.cpp:174-178). */
typedef struct {
    double diss_base;
    double diss_extra;
    double wave_speed;
    double dt;
    int32_t setup_iterations;  /* default 30000 */
    int32_t iterations;        /* default 3000 */
    double tolerance;          /* default 1e-6 */
    int32_t print_step;        /* default 500 */
    int32_t steady_check_step; /* default 30 */
    int32_t steady_hold_steps; /* default 3 */
} rrg_wave_params;

/* ini_params section (io_ini_galaxy_data.h:13-20; io_json.cpp:115-131).
 * An empty path string stands for JSON null / absent. */
typedef struct {
    double mass_star;
    double mass_dark;
    double soft_star;
    double soft_dark;
    char ini_file_star[1024];
    char ini_file_dark[1024];
} rrg_ini_params;

typedef struct {
    rrg_sim_params sim;
    int32_t has_grid;
    rrg_grid_params grid;
    int32_t has_wave;
    rrg_wave_params wave;
    int32_t use_wave_model; /* derived: has_grid && has_wave (io_json.cpp:171-172) */
    /* Options of the optional "mi355x" params.json section (ignored by the
     * reference's reader, so files stay interchangeable). */
    int32_t real_type;      /* RRG_REAL_*; default RRG_REAL_F64 */
    int32_t force_kind;     /* RRG_FORCE_*; default RRG_FORCE_F64 */
    int32_t boundary;       /* RRG_BC_*; default RRG_BC_PERIODIC */
    int32_t fixed_exponent; /* CIC fixed-point resolution 2^-e; 0 = derive from total mass */
    int32_t device;         /* HIP device ordinal; default 0 */
    int32_t use_graphs;     /* 1 (default): replay captured hipGraphs for steps */
    int32_t autotune;       /* 1: time candidate launch shapes at precalc and keep
                               the fastest (cached per process); default 0 */
    int32_t wave_fused;     /* 1 (default): several stencil iterations per
                               launch on periodic 64^3-512^3 grids (up to four
                               at 128^3 and FP32 256^3, three elsewhere, two
                               for FP64 512^3), bitwise the same results;
                               0: one */
    int32_t sort_interval;  /* PM: steps between re-sorts of the particles by
                               cell (results do not depend on it); 0 = 48 */
} rrg_params;

typedef struct {
    double time;
    double Ek, Ep;
    double momentum[3];
    double angular[3];
} rrg_conservation;

typedef struct {
    int64_t n;                 /* particles */
    int64_t steps;             /* completed calc_step calls */
    int32_t last_wave_iterations; /* stencil iterations run by the last wave solve */
    int32_t fixed_exponent;    /* CIC fixed-point exponent in use */
    double time;               /* sum of dt_dynamics over completed steps */
    int32_t systems;           /* independent systems (1, or k of an ensemble) */
    int32_t tile_ib;           /* direct sum: targets per lane */
    int64_t slice_len;         /* direct sum: sources per slice / work item */
    int32_t workgroups;        /* direct sum: workgroups of the force launch */
    int32_t work_items;        /* direct sum, ensemble: items in the work list */
    int32_t tuned;             /* 1 when the launch shape came from the autotuner */
    int32_t tune_candidates;   /* candidates timed by the autotuner */
    int64_t graph_launches;    /* steps replayed from a captured hipGraph */
} rrg_stats;

/* Stages timed by rrg_set_timing / rrg_get_timing. */
enum {
    RRG_STAGE_DEPOSIT = 0,  /* predict + CIC deposit (PM) */
    RRG_STAGE_SOURCE = 1,   /* fixed point -> mass, source term (PM) */
    RRG_STAGE_WAVE = 2,     /* wave_diss_iteration launches (PM) */
    RRG_STAGE_GRADIENT = 3, /* -grad phi (PM; a step forms it inside the gather, so no launches) */
    RRG_STAGE_GATHER = 4,   /* trilinear gather + kick (PM) */
    RRG_STAGE_DIRECT = 5,   /* direct-sum force launches (N-body) */
    RRG_STAGE_PREDICT = 6,  /* predictor (N-body) */
    RRG_STAGE_CORRECT = 7,  /* corrector (N-body) */
    RRG_STAGE_SORT = 8,     /* spatial re-sort of the particles (PM) */
    RRG_STAGE_ADJ_FORCE = 9,     /* adjoint: pair-force VJP (N-body); interpolation adjoint + gradient transpose (PM) */
    RRG_STAGE_ADJ_WAVE = 10,     /* adjoint: transposed stencil iterations (PM) */
    RRG_STAGE_ADJ_PARTICLE = 11, /* adjoint: CIC adjoint and transposed Verlet update */
    RRG_NSTAGES = 12
};

/* GPU time per stage from HIP events recorded on the context stream around
 * each stage's launches (ms), and the number of kernel launches timed
 * (entries past RRG_NSTAGES are zero). */
typedef struct {
    double ms[12];
    int64_t launches[12];
} rrg_timing;

typedef struct rrg_ctx rrg_ctx;

/* ---- library ---------------------------------------------------------- */
const char* rrg_last_error(void);
const char* rrg_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int rrg_device_count(void);

/* ---- params.json / ini (replaces rrgsim_io: io_json.cpp:188-242,
 *      io_particles_csv.cpp:11-134) --------------------------------------- */
/* Fill defaults: sections absent, saves_without_dump = UINT64_MAX, the wave
 * defaults of io_json.cpp:58-98, build options as documented above. */
int rrg_params_default(rrg_params* p);
/* Derived fields (validate_dependent_fields, io_json.cpp:170-186). */
int rrg_params_finalize(rrg_params* p);
/* parse_params_json (io_json.cpp:188-242).  `ini` may be NULL. */
int rrg_params_parse_json(const char* path, rrg_params* p, rrg_ini_params* ini);
/* read_simple_particles_data (io_particles_csv.cpp:89-134).  Paths are
 * resolved against `base_dir` when relative (NULL = process cwd, like the
 * reference).  Call with pos == NULL to get the count in *n; then call with
 * buffers of that size.  Unlike the reference, N need not be a multiple of
 * 256. Masses follow the reference: M_component / N_cumulative
 * (io_particles_csv.cpp:36-47). */
int rrg_ini_read(const rrg_ini_params* ini, const char* base_dir, int64_t* n, double* pos,
                 double* vel, double* mass, double* soft2);

/* ---- solver context (IntegrationSolver, rrgsim_integration.cuh:22-42) --- */
/* use_wave_model selects the PM strategy (RectWaveIntegrationSolver), else
 * the direct-sum strategy (NBodyIntegrationSolver).  soft2 may be NULL in PM
 * mode. */
int rrg_create(const rrg_params* p, int64_t n, const double* pos, const double* vel,
               const double* mass, const double* soft2, rrg_ctx** out);
/* Ensemble of k independent systems of uneven size in one context (no
 * counterpart in the reference): system g holds particles
 * [offsets[g], offsets[g+1]) of the concatenated arrays (offsets[0] = 0,
 * offsets[k] = n).  Direct-sum model only; every force evaluation is one
 * persistent work-queue launch over (system, target tile, source slice)
 * items, and steps replay a captured hipGraph. */
int rrg_create_ensemble(const rrg_params* p, int32_t k, const int64_t* offsets, const double* pos,
                        const double* vel, const double* mass, const double* soft2, rrg_ctx** out);
int rrg_destroy(rrg_ctx* ctx);
/* precalc(): initial potential and acceleration
 * (rrgsim_integration_nbody.cuh:26-30; rrgsim_integration_rect.cuh:53-83). */
int rrg_precalc(rrg_ctx* ctx);
/* nsteps x calc_step() (rrgsim_integration_nbody.cuh:50-56;
 * rrgsim_integration_rect.cuh:99-112), queued on the context stream. */
int rrg_step(rrg_ctx* ctx, int32_t nsteps);
/* sync_particles_and_grid(): particle potential (and, after precalc, the PM
 * acceleration) for output (rrgsim_integration_nbody.cuh:32-48;
 * rrgsim_integration_rect.cuh:85-97).  In N-body mode the potential is
 * recomputed by direct sum (the reference leaves it stale). */
int rrg_sync(rrg_ctx* ctx);
/* Copy state out (any pointer may be NULL): pos, vel, acc as (n,3), grav (n). */
int rrg_get_state(rrg_ctx* ctx, double* pos, double* vel, double* acc, double* grav);
/* Replace positions and velocities (n,3); acceleration is left as is. */
int rrg_set_state(rrg_ctx* ctx, const double* pos, const double* vel);
/* rrg_get_state / rrg_set_state with device arrays (FP64 (n,3) rows in the
 * caller's particle order, device memory of the context's device) instead of
 * host buffers: the transfers stay on the device, ordered after the work
 * queued on `stream` (a hipStream_t, NULL = the null stream), and `stream`
 * waits for them before its later work.  They replace direct access to the
 * device arrays of ParticlesContext_ (co_particles_context.cuh:17-31) by code
 * that shares the GPU, such as a torch caller.  Host pointers are rejected. */
int rrg_get_state_dev(rrg_ctx* ctx, double* pos, double* vel, double* acc, void* stream);
int rrg_set_state_dev(rrg_ctx* ctx, const double* pos, const double* vel, void* stream);
/* Energy, momentum and angular momentum by FP64 device reduction, with the
 * definitions of calc_conservation (nbody.cu:246-291). */
int rrg_get_conservation(rrg_ctx* ctx, rrg_conservation* out);
/* Grids (any pointer may be NULL), x-fastest (co_device_utils.cuh:61): PM
 * contexts give cell mass (nx^3), phi (nx^3) and -grad phi (nx^3,3);
 * direct-sum contexts created with a grid section give the CIC cell mass and
 * the NGP cell averages of the particle potential and acceleration, as the
 * reference's direct-sum strategy projects them at sync
 * (rrgsim_integration_nbody.cuh:32-48, rm_p2mesh.cu:139-168, 453-485). */
int rrg_get_grid(rrg_ctx* ctx, double* mass, double* phi, double* acc);
/* Per-system conservation sums of an ensemble (k entries; k = 1 for a
 * single system gives the same as rrg_get_conservation). */
int rrg_get_conservation_systems(rrg_ctx* ctx, int32_t k, rrg_conservation* out);
int rrg_get_stats(rrg_ctx* ctx, rrg_stats* out);
/* How a context of this real type, grid size and boundary type splits
 * `iterations` unchecked stencil iterations into launches (no counterpart in
 * the reference): launches[0..3] = launches running 1, 2, 3 and 4
 * iterations; `fused` is the wave_fused option. */
int rrg_wave_schedule(int32_t real_type, int32_t nx, int32_t boundary, int32_t iterations, int32_t fused,
                      int32_t* launches);
/* Wait for all work queued on the context stream. */
int rrg_synchronize(rrg_ctx* ctx);
/* Turn per-stage event timing on (1) or off (0); either resets the totals. */
int rrg_set_timing(rrg_ctx* ctx, int32_t enabled);
/* Synchronize the context stream and return the accumulated totals. */
int rrg_get_timing(rrg_ctx* ctx, rrg_timing* out);

/* ---- adjoints of the direct-sum and PM runs (no counterpart in the
 * reference) ----
 * rrg_adjoint_begin marks the current state (after rrg_precalc) as step 0 of
 * a recorded run; the following rrg_step calls store checkpoints every
 * `checkpoint_every` steps on the device ((x, v, a), and for PM also the two
 * potential levels).  rrg_adjoint then returns the gradients of a loss
 * L(x_T, v_T) of the state after the recorded steps with respect to the
 * initial positions, velocities and the masses, given dL/dx_T and dL/dv_T
 * (either may be NULL = zero).  It recomputes each checkpoint segment forward
 * and runs the transposed step (direct sum: reverse Verlet with one
 * vector-Jacobian product of the pair forces; PM: the reversed grid build,
 * stencil and CIC).  Any output pointer may be NULL.  Single-system contexts;
 * PM on periodic and zero-Dirichlet grids. */
int rrg_adjoint_begin(rrg_ctx* ctx, int32_t checkpoint_every);
int rrg_adjoint(rrg_ctx* ctx, const double* gx_T, const double* gv_T, double* gx_0, double* gv_0,
                double* gm);
/* rrg_adjoint with device arrays, ordered on `stream` as rrg_get_state_dev. */
int rrg_adjoint_dev(rrg_ctx* ctx, const double* gx_T, const double* gv_T, double* gx_0, double* gv_0, double* gm,
                    void* stream);

/* ---- stateless operators (one call = one kernel family on device 0) ----
 * Arrays of `real` are double for RRG_REAL_F64 and float for RRG_REAL_F32.
 * They are the reference's L2 free functions without the contexts. */

/* nbody_acceleration (nbody.cu:200-213). force_kind = RRG_FORCE_*.
 * pos (n,3), mass (n), soft2 (n), acc (n,3) -- all double. */
int rrg_op_direct_accel(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                        const double* soft2, double* acc);
/* nbody_grav (nbody.cu:215-228), self term included. */
int rrg_op_direct_grav(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                       const double* soft2, double* grav);
/* deposit_cell_mass_CIC_ (rm_p2mesh.cu:176-236) with int64 fixed-point
 * accumulation: grid (nx^3 int64) = sum of rint(m*w * 2^e). */
int rrg_op_cic_deposit(int32_t real_type, int64_t n, const void* pos, const void* mass, int32_t nx,
                       double dx, double bc_frac, int32_t e, int64_t* grid);
/* `iterations` x wave_diss_iteration (wave.cu:106-193) with the buffer
 * rotation of solve_wave_equation (wave.cu:253-272), no steady check.
 * In: phi_prev, phi_curr, mass (nx^3 real). Out: phi_prev, phi_curr updated. */
int rrg_op_wave_iterate(int32_t real_type, int32_t nx, double dx, double diss_base,
                        double wave_speed, double wave_dt, int32_t boundary, double rho_mean,
                        int32_t iterations, void* phi_prev, void* phi_curr, const void* mass);
/* calc_acceleration_field_ (rm_p2mesh.cu:238-290): acc (nx^3,3 real). */
int rrg_op_gradient(int32_t real_type, int32_t nx, double dx, const void* phi, void* acc);
/* interpolate_acceleration_to_particles_kernel (rm_mesh2p.cu:48-114). */
int rrg_op_gather(int32_t real_type, int64_t n, const void* pos, int32_t nx, double dx,
                  double bc_frac, const void* field, void* acc);
/* NGP binning of particles (assign_particles_to_cells_ and the prefix sums,
 * rm_p2mesh.cu:25-126) as a counting sort: cell_of[n] (AT index of the
 * truncated (x - min)/dx, clamped), count[nx^3], start[nx^3] (the exclusive
 * prefix of count, which the reference's prefix kernel intends but does not
 * write), order[n] (particle indices sorted by cell, ascending within a
 * cell).  Any output may be NULL. */
int rrg_op_ngp_bin(int32_t real_type, int64_t n, const void* pos, int32_t nx, double dx, double bc_frac,
                   int32_t* cell_of, int32_t* count, int32_t* start, int32_t* order);
/* SteadyStateChecker::calc_max_diff (wave_steady_checker.cu:95-127). */
int rrg_op_steady_err(int32_t real_type, int64_t ncell, const void* u_new, const void* u_old,
                      double* err);

/* rrg_op_direct_accel without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipStream capture, tight serving loops, harnesses that forbid hipMalloc).
 * pos (n,3), mass (n), soft2 (n) and acc (n,3) are FP64 device arrays of the
 * current device; the work is queued on `stream` (a hipStream_t, NULL = the
 * null stream) and the call returns without waiting for it.  The result is
 * bitwise that of rrg_op_direct_accel.  The working storage (source records,
 * per-slice partial sums) lives in a fixed 64 MiB device area that the
 * library holds on every device it runs on until the process ends, so n is
 * limited to what fits: every n <= 65535 for every force kind, and every
 * n <= 250000 for the FP64 kinds; a larger n returns RRG_ERR_INVALID with the
 * bytes it would need.  Calls share that area per device: calls whose work overlaps
 * on one device (other streams, other threads) race, so serialise them. */
int rrg_op_direct_accel_noalloc(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                                const double* soft2, double* acc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RRG_H */
