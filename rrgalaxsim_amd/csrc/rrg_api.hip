// librrg.so C ABI: solver contexts (the reference's IntegrationSolver
// strategies) and stateless operators.
//
// Reference:
//   IntegrationSolver                 solver/app/include/rrgsim_integration.cuh:22-42
//   NBodyIntegrationSolver            solver/app/include/rrgsim_integration_nbody.cuh:10-65
//   RectWaveIntegrationSolver         solver/app/include/rrgsim_integration_rect.cuh:13-123
//   solve_wave_equation               solver/lib/rrgsim_wave/src/wave.cu:230-297
//   SteadyStateChecker                solver/lib/rrgsim_wave/src/wave_steady_checker.cu:88-152
//   calc_conservation                 solver/lib/rrgsim_nbody/src/nbody.cu:246-291
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "rrg_internal.hpp"
#include "rrg_io.hpp"

#define RRG_API extern "C" __attribute__((visibility("default")))

namespace rrg {
namespace {

thread_local std::string g_last_error;

template <class F> int guarded(F&& f) {
    try {
        f();
        return RRG_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return RRG_ERR_INVALID;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return RRG_ERR_INVALID;
    }
}

void require(bool cond, const std::string& msg) {
    if (!cond) fail(RRG_ERR_INVALID, msg);
}

// Makes `dev` current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        RRG_HIP(hipGetDevice(&prev));
        if (prev != dev) RRG_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// Pinned host staging buffer (allocated, never registered).
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        RRG_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

// ----------------------------------------------------------------- particles
template <class real> struct Particles {
    int64_t n = 0;
    DBuf<real> X, V, A, W, M, S2, G;

    void alloc(int64_t count) {
        n = count;
        X.alloc(3 * n); V.alloc(3 * n); A.alloc(3 * n); W.alloc(3 * n);
        M.alloc(n); S2.alloc(n); G.alloc(n);
    }
    PState<real> view() const {
        return PState<real>{X.p, V.p, A.p, W.p, M.p, S2.p, G.p, n};
    }
    void swap(Particles& o) {
        std::swap(X, o.X); std::swap(V, o.V); std::swap(A, o.A); std::swap(W, o.W);
        std::swap(M, o.M); std::swap(S2, o.S2); std::swap(G, o.G);
    }
};

// out[id[i]] = in[i] for AoS triples (id: internal slot -> caller's index).
void unpermute_vec(int64_t n, const std::vector<int>& id, const double* in, double* out) {
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[3 * int64_t(id[i]) + c] = in[3 * i + c];
}
void permute_vec_in(int64_t n, const std::vector<int>& id, const double* in, double* out) {
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = in[3 * int64_t(id[i]) + c];
}

template <class real> void aos_to_soa(int64_t n, const double* aos, std::vector<real>& soa) {
    soa.resize(3 * n);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = (real)aos[3 * i + c];
}
template <class real> void soa_to_aos(int64_t n, const std::vector<real>& soa, double* aos) {
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) aos[3 * i + c] = (double)soa[c * n + i];
}
template <class real> void upload_vec(hipStream_t s, int64_t n, const double* aos, real* dst) {
    std::vector<real> tmp;
    aos_to_soa<real>(n, aos, tmp);
    RRG_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(real) * 3 * n, hipMemcpyHostToDevice, s));
    RRG_HIP(hipStreamSynchronize(s));
}
template <class real> void upload_scalar(hipStream_t s, int64_t n, const double* src, real* dst) {
    std::vector<real> tmp(n);
    for (int64_t i = 0; i < n; ++i) tmp[i] = (real)src[i];
    RRG_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(real) * n, hipMemcpyHostToDevice, s));
    RRG_HIP(hipStreamSynchronize(s));
}
template <class real> void download_vec(hipStream_t s, int64_t n, const real* src, double* aos) {
    std::vector<real> tmp(3 * n);
    RRG_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(real) * 3 * n, hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
    soa_to_aos<real>(n, tmp, aos);
}
template <class real> void download_scalar(hipStream_t s, int64_t n, const real* src, double* dst) {
    std::vector<real> tmp(n);
    RRG_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(real) * n, hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) dst[i] = (double)tmp[i];
}
void zero_async(hipStream_t s, void* p, size_t bytes) {
    if (bytes) RRG_HIP(hipMemsetAsync(p, 0, bytes, s));
}

// The maxima of a checking launch's slots (non-negative values, so the bit
// patterns order as the values do): out[0] = max|new - cur|, out[1] = max|new|.
template <class MB> void max_over_slots(const void* slots, MB* out) {
    const MB* v = static_cast<const MB*>(slots);
    out[0] = out[1] = 0;
    for (int q = 0; q < pm::kCheckSlots; ++q) {
        out[0] = std::max(out[0], v[q]);
        out[1] = std::max(out[1], v[pm::kCheckSlots + q]);
    }
}

// Force arithmetic for a direct-sum evaluation; detects uniform softening
// (then h_i + h_j is folded into one constant, formed in the kernel's
// precision exactly as the per-pair sum would be).
nbody::Force make_force(int kind, int64_t n, const double* soft2) {
    nbody::Force f;
    f.kind = kind;
    f.uniform = n > 0;
    for (int64_t i = 1; i < n && f.uniform; ++i) f.uniform = soft2[i] == soft2[0];
    if (f.uniform) {
        if (nbody::is_fp32_kind(kind)) {
            const float h = float(0.5 * soft2[0]);
            f.hh = double(h + h);
        } else {
            const double h = 0.5 * soft2[0];
            f.hh = h + h;
        }
    }
    return f;
}

bool valid_force_kind(int k) {
    return k == RRG_FORCE_F64 || k == RRG_FORCE_MIXED || k == RRG_FORCE_FP32 || k == RRG_FORCE_F64_EXACT;
}

// ------------------------------------------------------------ stage timing
// HIP events recorded on the context stream around each timed stage.  The
// event pairs are resolved (and returned to the pool) by collect().
class StageTimer {
public:
    ~StageTimer() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (const Rec& r : pending_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    void set(bool on) {
        on_ = on;
        reset();
    }
    bool on() const { return on_; }
    // Returns a token for stop(); -1 when timing is off.
    int start(hipStream_t s, int stage) {
        if (!on_) return -1;
        Rec r{stage, get(), get(), 0};
        RRG_HIP(hipEventRecord(r.a, s));
        pending_.push_back(r);
        return int(pending_.size()) - 1;
    }
    void stop(hipStream_t s, int token, int64_t launches) {
        if (token < 0) return;
        Rec& r = pending_[size_t(token)];
        r.launches = launches;
        RRG_HIP(hipEventRecord(r.b, s));
    }
    void collect(hipStream_t s, rrg_timing* out) {
        RRG_HIP(hipStreamSynchronize(s));
        for (const Rec& r : pending_) {
            float ms = 0.f;
            RRG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            ms_[r.stage] += ms;
            launches_[r.stage] += r.launches;
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        pending_.clear();
        std::memset(out, 0, sizeof(*out));
        for (int k = 0; k < RRG_NSTAGES; ++k) {
            out->ms[k] = ms_[k];
            out->launches[k] = launches_[k];
        }
    }

private:
    struct Rec {
        int stage;
        hipEvent_t a, b;
        int64_t launches;
    };
    hipEvent_t get() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        RRG_HIP(hipEventCreate(&e));
        return e;
    }
    void reset() {
        for (const Rec& r : pending_) {
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        pending_.clear();
        for (int k = 0; k < RRG_NSTAGES; ++k) {
            ms_[k] = 0.0;
            launches_[k] = 0;
        }
    }
    bool on_ = false;
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> pool_;
    double ms_[RRG_NSTAGES] = {};
    int64_t launches_[RRG_NSTAGES] = {};
};

// ------------------------------------------------------------------ solvers
struct Solver {
    StageTimer timer;
    virtual ~Solver() = default;
    virtual void precalc() = 0;
    virtual void step(int nsteps) = 0;
    virtual void sync() = 0;
    virtual void get_state(double* pos, double* vel, double* acc, double* grav) = 0;
    virtual void set_state(const double* pos, const double* vel) = 0;
    // The same with device arrays, enqueued on the context stream.
    virtual void get_state_dev(double* pos, double* vel, double* acc) = 0;
    virtual void set_state_dev(const double* pos, const double* vel) = 0;
    virtual void conservation(rrg_conservation* out) = 0;
    virtual void conservation_systems(int k, rrg_conservation* out) = 0;
    virtual void get_grid(double* mass, double* phi, double* acc) = 0;
    virtual void stats(rrg_stats* out) = 0;
    virtual void adjoint_begin(int) { fail(RRG_ERR_STATE, "adjoint: supported by direct-sum contexts only"); }
    virtual void adjoint(const double*, const double*, double*, double*, double*) {
        fail(RRG_ERR_STATE, "adjoint: supported by direct-sum contexts only");
    }
    virtual void adjoint_dev(const double*, const double*, double*, double*, double*) {
        fail(RRG_ERR_STATE, "adjoint: supported by direct-sum contexts only");
    }
};

void fill_conservation(const double* q, double time, rrg_conservation* out) {
    out->time = time;
    out->Ek = 0.5 * q[0];
    out->Ep = 0.5 * q[1];
    for (int c = 0; c < 3; ++c) {
        out->momentum[c] = q[2 + c];
        out->angular[c] = q[5 + c];
    }
}

template <class real> void conservation_common(hipStream_t s, const PState<real>& v, DBuf<double>& part,
                                               PinnedBuf& host, double time, rrg_conservation* out) {
    const int64_t nb = nbody::conservation_blocks(v.n);
    if (part.n < size_t(nb * 8)) part.alloc(nb * 8);
    nbody::conservation<real>(s, v, part.p);
    host.ensure(sizeof(double) * nb * 8);
    RRG_HIP(hipMemcpyAsync(host.p, part.p, sizeof(double) * nb * 8, hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
    const double* h = static_cast<const double*>(host.p);
    double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t b = 0; b < nb; ++b)
        for (int k = 0; k < 8; ++k) q[k] += h[b * 8 + k];
    fill_conservation(q, time, out);
}

// Captured step graphs, keyed by the state that changes the launches'
// arguments (buffer rotation, particle buffers).  A graph holds kernel nodes
// only: the captured work records no memset.
class StepGraphs {
public:
    ~StepGraphs() { clear(); }  // owners synchronise their stream before destroying
    // Waits for `s` (the stream the graphs are launched on) first, so no
    // executable graph is destroyed while a launch of it may be in flight.
    void clear(hipStream_t s) {
        if (!execs_.empty()) RRG_HIP(hipStreamSynchronize(s));
        clear();
    }
    void clear() {
        for (auto& kv : execs_) (void)hipGraphExecDestroy(kv.second);
        execs_.clear();
    }
    template <class F> void run(hipStream_t s, uint64_t key, F&& enqueue) {
        auto it = execs_.find(key);
        if (it == execs_.end()) {
            hipGraph_t g = nullptr;
            RRG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue();
            } catch (...) {
                if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                throw;
            }
            RRG_HIP(hipStreamEndCapture(s, &g));
            hipGraphExec_t e = nullptr;
            const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            RRG_HIP(err);
            it = execs_.emplace(key, e).first;
        }
        RRG_HIP(hipGraphLaunch(it->second, s));
        ++launches_;
    }
    int64_t launches() const { return launches_; }

private:
    std::map<uint64_t, hipGraphExec_t> execs_;
    int64_t launches_ = 0;
};

// Cache of autotuned direct-sum launch shapes, keyed by GPU architecture,
// force kind, softening form and size buckets; kept for the process and in a
// text file (one "key ib len" line per shape) so later processes reuse it:
// $RRG_TUNE_CACHE if set (empty: no file), else
// ${XDG_CACHE_HOME:-$HOME/.cache}/rrgalaxsim_amd/autotune.txt.
struct TunedShape {
    int ib = 0;
    int64_t len = 0;  // grid form: target workgroups; queue form: sources per item
};
std::mutex g_tune_mutex;
std::map<std::string, TunedShape> g_tune_cache;
bool g_tune_file_read = false;

std::string tune_cache_path() {
    if (const char* e = std::getenv("RRG_TUNE_CACHE")) return e;
    std::string base;
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) base = x;
    else if (const char* h = std::getenv("HOME"); h && *h) base = std::string(h) + "/.cache";
    else return "";
    return base + "/rrgalaxsim_amd/autotune.txt";
}

// With g_tune_mutex held.
void tune_file_load() {
    if (g_tune_file_read) return;
    g_tune_file_read = true;
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    std::ifstream f(path);
    std::string key;
    TunedShape t;
    while (f >> key >> t.ib >> t.len)
        if (t.ib > 0 && t.len > 0) g_tune_cache[key] = t;
}

// With g_tune_mutex held; a cache that cannot be written is skipped.
void tune_file_append(const std::string& key, const TunedShape& t) {
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    std::error_code ec;
    std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
    std::ofstream f(path, std::ios::app);
    if (f) f << key << ' ' << t.ib << ' ' << t.len << '\n';
}

// floor(2 log2 v): half-octave size bucket.
int half_octave(double v) { return v > 0 ? int(std::floor(2.0 * std::log2(v))) : -1; }

// Direct-sum strategy (NBodyIntegrationSolver), FP64 particle state, for one
// system (grid-form launches) or an ensemble of independent systems
// (queue-form launches over a work list).
class NBodySolver final : public Solver {
public:
    NBodySolver(const rrg_params& p, hipStream_t s, int64_t n, const double* pos, const double* vel,
                const double* mass, const double* soft2, const std::vector<int64_t>& offsets)
        : prm_(p), s_(s), kind_(p.force_kind), offsets_(offsets) {
        require(soft2 != nullptr, "rrg_create: soft2 is required by the direct-sum model");
        require(valid_force_kind(kind_), "rrg_create: unknown force_kind");
        k_ = int(offsets_.size()) - 1;
        ensemble_ = k_ > 1;
        force_ = make_force(kind_, n, soft2);
        P_.alloc(n);
        upload_vec<double>(s_, n, pos, P_.X.p);
        upload_vec<double>(s_, n, vel, P_.V.p);
        upload_scalar<double>(s_, n, mass, P_.M.p);
        upload_scalar<double>(s_, n, soft2, P_.S2.p);
        zero_async(s_, P_.A.p, P_.A.bytes());
        zero_async(s_, P_.W.p, P_.W.bytes());
        zero_async(s_, P_.G.p, P_.G.bytes());
        d_offsets_.alloc(offsets_.size());
        RRG_HIP(hipMemcpyAsync(d_offsets_.p, offsets_.data(), sizeof(int64_t) * offsets_.size(),
                               hipMemcpyHostToDevice, s_));
        counters_.alloc(2);
        zero_async(s_, counters_.p, counters_.bytes());
        if (nbody::is_fp32_kind(kind_)) {
            A_.alloc(n);
            B_.alloc(n);
        }
        if (p.has_grid) grid_init(n, mass);
        configure(0, 0);
        RRG_HIP(hipStreamSynchronize(s_));
    }

    void precalc() override {  // rrgsim_integration_nbody.cuh:26-30
        if (nbody::is_fp32_kind(kind_)) nbody::pack_mixed(s_, P_.view(), A_.p, B_.p);
        if (prm_.autotune) tune();
        grav();
        accel(true);
        grav_actual_ = true;
        precalc_done_ = true;
    }

    void step(int nsteps) override {  // rrgsim_integration_nbody.cuh:50-56
        if (!precalc_done_) fail(RRG_ERR_STATE, "rrg_step before rrg_precalc");
        // Steps with no host work between them (no stage timing, no adjoint
        // recording) go kStepsPerGraph to a graph launch: at small N a step's
        // kernels take less time than launching them.
        if (prm_.use_graphs && !timer.on() && !recording_) {
            const double dt = prm_.sim.dt_dynamics;
            for (; nsteps >= kStepsPerGraph; nsteps -= kStepsPerGraph) {
                graphs_.run(s_, 4, [&] {
                    // each step's corrector and the next step's predictor
                    // run as one launch inside the graph
                    predict_kernel(dt);
                    for (int q = 0; q < kStepsPerGraph; ++q) {
                        accel(false);
                        if (q + 1 < kStepsPerGraph) correct_predict_kernel(dt);
                        else correct_kernel(dt);
                    }
                });
                steps_ += kStepsPerGraph;
                for (int q = 0; q < kStepsPerGraph; ++q) time_ += dt;
                grav_actual_ = false;
            }
        }
        for (int k = 0; k < nsteps; ++k) {
            enqueue_step(prm_.use_graphs != 0);
            ++steps_;
            time_ += prm_.sim.dt_dynamics;
            if (recording_) {
                ++rec_steps_;
                // positions of the open segment (its steps 1 .. K): the
                // reverse sweep takes the last segment from here instead of
                // recomputing it
                const int64_t slot = rec_steps_ - 1 - ck_every_ * ((rec_steps_ - 1) / ck_every_);
                RRG_HIP(hipMemcpyAsync(tape_.p + size_t(slot) * 3 * P_.n, P_.X.p, sizeof(double) * 3 * P_.n,
                                       hipMemcpyDeviceToDevice, s_));
                if (rec_steps_ % ck_every_ == 0) store_checkpoint();
            }
        }
        if (nsteps > 0) grav_actual_ = false;
    }

    void adjoint_begin(int every) override {
        require(!ensemble_, "rrg_adjoint_begin: not supported for ensembles");
        require(every >= 1, "rrg_adjoint_begin: checkpoint_every must be >= 1");
        if (!precalc_done_) fail(RRG_ERR_STATE, "rrg_adjoint_begin before rrg_precalc");
        ck_every_ = every;
        rec_steps_ = 0;
        ck_.clear();
        tape_.alloc(size_t(every) * 3 * P_.n);
        recording_ = true;
        store_checkpoint();
    }

    void adjoint(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) override {
        adjoint_impl(gxT, gvT, gx0, gv0, gm, false);
    }
    void adjoint_dev(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) override {
        adjoint_impl(gxT, gvT, gx0, gv0, gm, true);
    }

    // Reverse sweep over the recorded steps (see adjoint_kernels.hip); `dev`:
    // the seeds and gradients are device arrays.
    void adjoint_impl(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm, bool dev) {
        if (!recording_) fail(RRG_ERR_STATE, "rrg_adjoint without rrg_adjoint_begin");
        const int64_t n = P_.n;
        const double dt = prm_.sim.dt_dynamics;
        const bool f32 = nbody::is_fp32_kind(kind_);
        adj_alloc();
        if (dev) {
            if (gxT) layout::rows_in<double>(s_, n, gxT, Xb_.p, nullptr);
            else zero_async(s_, Xb_.p, Xb_.bytes());
            if (gvT) layout::rows_in<double>(s_, n, gvT, Vb_.p, nullptr);
            else zero_async(s_, Vb_.p, Vb_.bytes());
        } else {
            std::vector<double> zero(3 * n, 0.0);
            upload_vec<double>(s_, n, gxT ? gxT : zero.data(), Xb_.p);
            upload_vec<double>(s_, n, gvT ? gvT : zero.data(), Vb_.p);
        }
        zero_async(s_, Ab_.p, Ab_.bytes());
        zero_async(s_, gm_.p, gm_.bytes());
        const int T = int(rec_steps_);
        const int K = ck_every_;
        const int nseg = (T + K - 1) / K;
        const size_t v3 = sizeof(double) * 3 * n;
        for (int c = nseg - 1; c >= 0; --c) {
            const int s0 = c * K, len = std::min(T, s0 + K) - s0;
            if (c == nseg - 1) {
                // the last segment's positions were kept by the recorded run;
                // reverse steps n = s0+q, q = len-1 .. 0, each at x_{n+1}
                for (int q = len - 1; q >= 0; --q) reverse_step(tape_.p + size_t(q) * 3 * n, dt, false);
                continue;
            }
            // recompute the segment forward on the scratch state
            const DBuf<double>& ck = ck_[c];
            RRG_HIP(hipMemcpyAsync(R_.X.p, ck.p, v3, hipMemcpyDeviceToDevice, s_));
            RRG_HIP(hipMemcpyAsync(R_.V.p, ck.p + 3 * n, v3, hipMemcpyDeviceToDevice, s_));
            RRG_HIP(hipMemcpyAsync(R_.A.p, ck.p + 6 * n, v3, hipMemcpyDeviceToDevice, s_));
            P_.swap(R_);
            RRG_HIP(hipMemcpyAsync(xs_.p, P_.X.p, v3, hipMemcpyDeviceToDevice, s_));
            if (f32) nbody::pack_mixed(s_, P_.view(), A_.p, B_.p);
            for (int q = 1; q <= len; ++q) {
                enqueue_step(false);
                RRG_HIP(hipMemcpyAsync(xs_.p + size_t(q) * 3 * n, P_.X.p, v3, hipMemcpyDeviceToDevice, s_));
            }
            P_.swap(R_);
            // reverse steps n = s0+q, q = len-1 .. 0, each at x_{n+1}
            for (int q = len - 1; q >= 0; --q) reverse_step(xs_.p + size_t(q + 1) * 3 * n, dt, false);
        }
        // a_0 = F(x_0): the closing product at the initial positions
        reverse_step(ck_[0].p, 0.0, true);
        if (f32) nbody::pack_mixed(s_, P_.view(), A_.p, B_.p);
        if (dev) {
            if (gx0) layout::rows_out<double>(s_, n, Xb_.p, gx0, nullptr);
            if (gv0) layout::rows_out<double>(s_, n, Vb_.p, gv0, nullptr);
            if (gm) layout::scalar_out<double>(s_, n, gm_.p, gm, nullptr);
            return;
        }
        if (gx0) download_vec<double>(s_, n, Xb_.p, gx0);
        if (gv0) download_vec<double>(s_, n, Vb_.p, gv0);
        if (gm) download_scalar<double>(s_, n, gm_.p, gm);
        RRG_HIP(hipStreamSynchronize(s_));
    }

    // The reference returns early without a grid and leaves the potential
    // stale (rrgsim_integration_nbody.cuh:32-35); here it is recomputed.
    void sync() override {  // rrgsim_integration_nbody.cuh:32-48
        if (!grav_actual_) grav();
        grav_actual_ = true;
        if (has_grid_) project_grid();
    }

    void get_state(double* pos, double* vel, double* acc, double* gr) override {
        const int64_t n = P_.n;
        if (pos) download_vec<double>(s_, n, P_.X.p, pos);
        if (vel) download_vec<double>(s_, n, P_.V.p, vel);
        if (acc) download_vec<double>(s_, n, P_.A.p, acc);
        if (gr) download_scalar<double>(s_, n, P_.G.p, gr);
    }

    void set_state(const double* pos, const double* vel) override {
        if (pos) upload_vec<double>(s_, P_.n, pos, P_.X.p);
        if (vel) upload_vec<double>(s_, P_.n, vel, P_.V.p);
        state_changed(pos != nullptr);
    }

    void get_state_dev(double* pos, double* vel, double* acc) override {
        const int64_t n = P_.n;
        if (pos) layout::rows_out<double>(s_, n, P_.X.p, pos, nullptr);
        if (vel) layout::rows_out<double>(s_, n, P_.V.p, vel, nullptr);
        if (acc) layout::rows_out<double>(s_, n, P_.A.p, acc, nullptr);
    }

    void set_state_dev(const double* pos, const double* vel) override {
        if (pos) layout::rows_in<double>(s_, P_.n, pos, P_.X.p, nullptr);
        if (vel) layout::rows_in<double>(s_, P_.n, vel, P_.V.p, nullptr);
        state_changed(pos != nullptr);
    }

    void state_changed(bool moved) {
        if (moved && nbody::is_fp32_kind(kind_)) nbody::pack_mixed(s_, P_.view(), A_.p, B_.p);
        grav_actual_ = false;
        recording_ = false;  // a recorded run no longer describes this state
    }

    void conservation(rrg_conservation* out) override {
        sync();
        conservation_common<double>(s_, P_.view(), cpart_, host_, time_, out);
    }

    void conservation_systems(int k, rrg_conservation* out) override {
        require(k == k_, "rrg_get_conservation_systems: k does not match the context");
        sync();
        if (cpart_.n < size_t(8 * k_)) cpart_.alloc(8 * k_);
        nbody::conservation_segments(s_, P_.view(), k_, d_offsets_.p, cpart_.p);
        host_.ensure(sizeof(double) * 8 * k_);
        RRG_HIP(hipMemcpyAsync(host_.p, cpart_.p, sizeof(double) * 8 * k_, hipMemcpyDeviceToHost, s_));
        RRG_HIP(hipStreamSynchronize(s_));
        const double* h = static_cast<const double*>(host_.p);
        for (int g = 0; g < k_; ++g) fill_conservation(h + 8 * g, time_, out + g);
    }

    // With a grid the direct-sum model projects onto it at every sync
    // (rrgsim_integration_nbody.cuh:32-48): CIC cell mass, and the NGP cell
    // averages of the particle potential and acceleration.
    void get_grid(double* mass, double* phi, double* acc) override {
        if (!has_grid_) fail(RRG_ERR_STATE, "rrg_get_grid: this direct-sum context has no grid");
        sync();
        const size_t nc = size_t(ncell_);
        if (mass) RRG_HIP(hipMemcpyAsync(mass, gmass_.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s_));
        if (phi) RRG_HIP(hipMemcpyAsync(phi, gphi_.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s_));
        if (acc) RRG_HIP(hipMemcpyAsync(acc, gacc_.p, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, s_));
        RRG_HIP(hipStreamSynchronize(s_));
    }

    void stats(rrg_stats* out) override {
        std::memset(out, 0, sizeof(*out));
        out->n = P_.n;
        out->steps = steps_;
        out->time = time_;
        out->systems = k_;
        out->tile_ib = ensemble_ ? qplan_.ib : plan_.ib;
        out->slice_len = ensemble_ ? item_len_ : plan_.slice_len;
        out->workgroups = ensemble_ ? qplan_.workgroups : int(plan_.tiles * plan_.slices);
        out->work_items = ensemble_ ? qplan_.nitems : 0;
        out->tuned = tuned_ ? 1 : 0;
        out->tune_candidates = tune_candidates_;
        out->graph_launches = graphs_.launches();
        out->fixed_exponent = has_grid_ ? grid_e_ : 0;
    }

private:
    // Launch shape.  Grid form: ib and target workgroup count (0 = default).
    // Queue form: ib and sources per item (0 = default).
    void configure(int ib, int64_t len) {
        const int64_t n = P_.n;
        graphs_.clear(s_);
        if (!ensemble_) {
            plan_ = nbody::plan_direct(n, kind_, int(len), ib);
            if (plan_.slices > 1 && part_.n < size_t(plan_.slices) * 3 * n) part_.alloc(size_t(plan_.slices) * 3 * n);
            if (plan_.slices > 1 && gpart_.n < size_t(plan_.slices) * n) gpart_.alloc(size_t(plan_.slices) * n);
            return;
        }
        if (ib <= 0 || !nbody::valid_ib(kind_, ib)) ib = nbody::is_fp32_kind(kind_) ? 2 : 1;
        if (len <= 0) len = 4096;
        const int tile = nbody::targets_per_tile(ib);
        std::vector<nbody::WorkItem> items;
        std::vector<unsigned char> ns(n);
        int smax = 1;
        for (int g = 0; g < k_; ++g) {
            const int64_t o = offsets_[g], ng = offsets_[g + 1] - o;
            if (ng <= 0) continue;
            int64_t sg = std::min<int64_t>(255, std::max<int64_t>(1, (ng + len - 1) / len));
            const int64_t lg = (ng + sg - 1) / sg;
            sg = (ng + lg - 1) / lg;
            smax = std::max<int>(smax, int(sg));
            for (int64_t i = 0; i < ng; ++i) ns[o + i] = (unsigned char)sg;
            for (int64_t t0 = 0; t0 < ng; t0 += tile)
                for (int64_t q = 0; q < sg; ++q) {
                    nbody::WorkItem w{};
                    w.i0 = int(o + t0);
                    w.i1 = int(o + std::min(ng, t0 + tile));
                    w.j0 = int(o + q * lg);
                    w.j1 = int(o + std::min(ng, (q + 1) * lg));
                    w.slice = int(q);
                    items.push_back(w);
                }
        }
        // largest items first, so the queue's tail is made of small items
        std::stable_sort(items.begin(), items.end(), [](const nbody::WorkItem& a, const nbody::WorkItem& b) {
            return int64_t(a.i1 - a.i0) * (a.j1 - a.j0) > int64_t(b.i1 - b.i0) * (b.j1 - b.j0);
        });
        qplan_.ib = ib;
        qplan_.nitems = int(items.size());
        qplan_.workgroups = std::max(1, std::min<int>(qplan_.nitems, 2048));
        item_len_ = len;
        smax_ = smax;
        items_.alloc(items.size());
        RRG_HIP(hipMemcpyAsync(items_.p, items.data(), sizeof(nbody::WorkItem) * items.size(), hipMemcpyHostToDevice, s_));
        ns_.alloc(n);
        RRG_HIP(hipMemcpyAsync(ns_.p, ns.data(), n, hipMemcpyHostToDevice, s_));
        if (part_.n < size_t(smax) * 3 * n) part_.alloc(size_t(smax) * 3 * n);
        if (gpart_.n < size_t(smax) * n) gpart_.alloc(size_t(smax) * n);
        RRG_HIP(hipStreamSynchronize(s_));
    }

    // Autotuner: time every candidate shape on the live data (one warm-up and
    // three timed force launches each, median) and keep the fastest; results
    // are cached per (architecture, kind, softening, size buckets): for one
    // system its N, for an ensemble the number of systems, sum N^2 and the
    // largest N, each in half-octaves.
    std::string tune_key() const {
        hipDeviceProp_t prop{};
        RRG_HIP(hipGetDeviceProperties(&prop, prm_.device));
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        std::string key = arch + "|k" + std::to_string(kind_) + (force_.uniform ? "u" : "v");
        if (!ensemble_) return key + "|n" + std::to_string(half_octave(double(P_.n)));
        double s2 = 0, nmax = 0;
        for (int g = 0; g < k_; ++g) {
            const double ng = double(offsets_[g + 1] - offsets_[g]);
            s2 += ng * ng;
            nmax = std::max(nmax, ng);
        }
        return key + "|e" + std::to_string(k_) + "s" + std::to_string(half_octave(s2)) + "m" +
               std::to_string(half_octave(nmax));
    }

    void tune() {
        const std::string key = tune_key();
        {
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            tune_file_load();
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end() && nbody::valid_ib(kind_, it->second.ib)) {
                configure(it->second.ib, it->second.len);
                tuned_ = true;
                tune_candidates_ = 0;
                return;
            }
        }
        const std::vector<int> ibs = nbody::is_fp32_kind(kind_) ? std::vector<int>{1, 2, 4} : std::vector<int>{1, 2};
        const std::vector<int64_t> lens = {1024, 2048, 4096, 8192};
        hipEvent_t ev[4];
        for (auto& e : ev) RRG_HIP(hipEventCreate(&e));
        TunedShape best;
        float best_ms = 1e30f;
        tune_candidates_ = 0;
        try {
            for (int ib : ibs)
                for (int64_t len : lens) {
                    configure(ib, len);
                    accel(false);
                    RRG_HIP(hipEventRecord(ev[0], s_));
                    for (int r = 0; r < 3; ++r) {
                        accel(false);
                        RRG_HIP(hipEventRecord(ev[r + 1], s_));
                    }
                    RRG_HIP(hipEventSynchronize(ev[3]));
                    float t[3];
                    for (int r = 0; r < 3; ++r) RRG_HIP(hipEventElapsedTime(&t[r], ev[r], ev[r + 1]));
                    std::sort(t, t + 3);
                    ++tune_candidates_;
                    if (t[1] < best_ms) {
                        best_ms = t[1];
                        best.ib = ib;
                        best.len = len;
                    }
                }
        } catch (...) {
            for (auto& e : ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        configure(best.ib, best.len);
        tuned_ = true;
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune_cache[key] = best;
        tune_file_append(key, best);
    }

    void grid_init(int64_t n, const double* mass) {
        require(!ensemble_, "rrg_create_ensemble: an ensemble has no grid");
        const rrg_grid_params& g = prm_.grid;
        require(g.nx >= 2 && g.nx <= 1290, "rrg_create: grid nx must be in [2, 1290]");
        require(g.dx > 0, "rrg_create: grid dx must be positive");
        has_grid_ = true;
        nx_ = g.nx;
        ncell_ = int64_t(nx_) * nx_ * nx_;
        geo_ = pm::make_geo<double>(nx_, g.dx, g.bc_frac);
        double msum = 0;
        for (int64_t i = 0; i < n; ++i) msum += std::fabs(mass[i]);
        int ex = 0;
        std::frexp(std::max(msum, 1e-300), &ex);
        const int e = prm_.fixed_exponent > 0 ? prm_.fixed_exponent : std::min(61 - ex, 1000);
        require(std::ldexp(msum, e) < std::ldexp(1.0, 62), "rrg_create: fixed_exponent too large for total mass");
        grid_e_ = e;
        scale_ = std::ldexp(1.0, e);
        inv_scale_ = std::ldexp(1.0, -e);
        double mmax = 0;
        for (int64_t i = 0; i < n; ++i) mmax = std::max(mmax, std::fabs(mass[i]));
        qmax_ = mmax * scale_;
        mfix_.alloc(ncell_);
        gmass_.alloc(ncell_);
        gphi_.alloc(ncell_);
        gacc_.alloc(3 * ncell_);
        bkeys_.alloc(n); bkeys2_.alloc(n); bidx_.alloc(n); border_.alloc(n);
        bcount_.alloc(ncell_); bstart_.alloc(ncell_);
        btemp_.alloc(pm::ngp_temp_bytes(n, nx_));
        zero_async(s_, mfix_.p, mfix_.bytes());
    }

    // convert_particles_to_grid (CIC mass, rm_p2mesh.cu:398-451) and the NGP
    // projections of potential and acceleration (rm_p2mesh.cu:453-485).
    void project_grid() {
        pm::predict_deposit<double>(s_, P_.view(), 0.0, false, geo_, scale_, mfix_.p, qmax_);
        pm::fixed_to_mass<double>(s_, nx_, mfix_.p, inv_scale_, gmass_.p);
        pm::NgpBins b;
        b.keys = bkeys_.p; b.keys2 = bkeys2_.p; b.idx = bidx_.p; b.order = border_.p;
        b.count = bcount_.p; b.start = bstart_.p; b.temp = btemp_.p; b.temp_bytes = btemp_.n;
        pm::ngp_bin<double>(s_, P_.view(), geo_, b);
        pm::ngp_project<double>(s_, nx_, b, P_.n, P_.G.p, P_.A.p, gphi_.p, gacc_.p);
    }

    void store_checkpoint() {
        const int64_t n = P_.n;
        ck_.emplace_back(9 * n);
        DBuf<double>& ck = ck_.back();
        const size_t v3 = sizeof(double) * 3 * n;
        RRG_HIP(hipMemcpyAsync(ck.p, P_.X.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 3 * n, P_.V.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 6 * n, P_.A.p, v3, hipMemcpyDeviceToDevice, s_));
    }

    void adj_alloc() {
        const int64_t n = P_.n;
        if (R_.n != n) {
            R_.alloc(n);
            RRG_HIP(hipMemcpyAsync(R_.M.p, P_.M.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s_));
            RRG_HIP(hipMemcpyAsync(R_.S2.p, P_.S2.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s_));
            zero_async(s_, R_.W.p, R_.W.bytes());
            zero_async(s_, R_.G.p, R_.G.bytes());
            Xb_.alloc(3 * n); Vb_.alloc(3 * n); Ab_.alloc(3 * n); At_.alloc(3 * n); gm_.alloc(n);
            vplan_ = nbody::plan_direct(n, kind_, 0, std::min(plan_.ib, 2));
            vpart_.alloc(size_t(vplan_.slices) * 4 * n);
            if (nbody::is_fp32_kind(kind_)) {
                Uf_.alloc(n);
                Ap_.alloc(n);
                Bp_.alloc(n);
            }
        }
        if (xs_.n < size_t(ck_every_ + 1) * 3 * n) xs_.alloc(size_t(ck_every_ + 1) * 3 * n);
    }

    // One reverse Verlet step at positions x (the state after the forward
    // step); `last` is the closing product for a_0 = F(x_0).
    void reverse_step(const double* x, double dt, bool last) {
        const int64_t n = P_.n;
        const bool f32 = nbody::is_fp32_kind(kind_);
        int t = timer.start(s_, RRG_STAGE_ADJ_PARTICLE);
        adjoint::pre(s_, n, dt, Vb_.p, Ab_.p, At_.p, f32 ? Uf_.p : nullptr);
        if (f32) adjoint::pack_pos(s_, n, x, P_.M.p, P_.S2.p, Ap_.p, Bp_.p);
        timer.stop(s_, t, f32 ? 2 : 1);
        t = timer.start(s_, RRG_STAGE_ADJ_FORCE);
        adjoint::vjp(s_, vplan_, kind_, n, x, P_.M.p, P_.S2.p, At_.p, Ap_.p, Bp_.p, Uf_.p, vpart_.p);
        timer.stop(s_, t, 1);
        t = timer.start(s_, RRG_STAGE_ADJ_PARTICLE);
        adjoint::post(s_, n, prm_.sim.dt_dynamics, vplan_.slices, vpart_.p, Xb_.p, Vb_.p, Ab_.p, gm_.p, last);
        timer.stop(s_, t, 1);
    }

    // One step as three segments (predictor, force, corrector); with `graph`
    // each segment is replayed from its captured graph, and stage timing
    // events sit between the segment launches.
    template <class F> void segment(int stage, bool graph, uint64_t key, F&& f) {
        const int t = timer.start(s_, stage);
        if (graph) graphs_.run(s_, key, f);
        else f();
        timer.stop(s_, t, 1);
    }

    static constexpr int kStepsPerGraph = 8;

    void predict_kernel(double dt) { nbody::predict<double>(s_, P_.view(), dt, A_.p, B_.p); }
    void correct_kernel(double dt) {
        if (ensemble_) nbody::correct_reduce(s_, P_.view(), dt, 0, ns_.p, part_.p);
        else nbody::correct_reduce(s_, P_.view(), dt, plan_.slices, nullptr, plan_.slices > 1 ? part_.p : nullptr);
    }
    void correct_predict_kernel(double dt) {
        if (ensemble_) nbody::correct_predict(s_, P_.view(), dt, 0, ns_.p, part_.p, A_.p, B_.p);
        else
            nbody::correct_predict(s_, P_.view(), dt, plan_.slices, nullptr, plan_.slices > 1 ? part_.p : nullptr,
                                   A_.p, B_.p);
    }
    // The kernels of one step: predictor, force, corrector.
    void enqueue_kernels(double dt) {
        predict_kernel(dt);
        accel(false);
        correct_kernel(dt);
    }

    void enqueue_step(bool graph) {
        const double dt = prm_.sim.dt_dynamics;
        auto predict = [&] { predict_kernel(dt); };
        auto force = [&] { accel(false); };
        auto correct = [&] { correct_kernel(dt); };
        if (graph && !timer.on()) {
            // without stage timing the whole step is one graph launch (small N
            // is bound by the host's launch rate)
            graphs_.run(s_, 3, [&] { enqueue_kernels(dt); });
            return;
        }
        segment(RRG_STAGE_PREDICT, graph, 0, predict);
        segment(RRG_STAGE_DIRECT, graph, 1, force);
        segment(RRG_STAGE_CORRECT, graph, 2, correct);
    }

    void accel(bool reduce_now) {
        const int64_t n = P_.n;
        if (ensemble_) {
            nbody::queue(s_, qplan_, force_, false, P_.view(), A_.p, B_.p, items_.p, counters_.p, part_.p);
            if (reduce_now) nbody::reduce_slices(s_, n, 3, 0, ns_.p, part_.p, P_.A.p);
            return;
        }
        double* out = plan_.slices > 1 ? part_.p : P_.A.p;
        nbody::accel(s_, plan_, force_, P_.view(), A_.p, B_.p, out);
        if (reduce_now && plan_.slices > 1) nbody::reduce_slices(s_, n, 3, plan_.slices, nullptr, part_.p, P_.A.p);
    }

    void grav() {
        const int64_t n = P_.n;
        if (ensemble_) {
            nbody::queue(s_, qplan_, force_, true, P_.view(), A_.p, B_.p, items_.p, counters_.p, gpart_.p);
            nbody::reduce_slices(s_, n, 1, 0, ns_.p, gpart_.p, P_.G.p);
            return;
        }
        double* out = plan_.slices > 1 ? gpart_.p : P_.G.p;
        nbody::grav(s_, plan_, force_, P_.view(), A_.p, B_.p, out);
        if (plan_.slices > 1) nbody::reduce_slices(s_, n, 1, plan_.slices, nullptr, gpart_.p, P_.G.p);
    }

    rrg_params prm_;
    hipStream_t s_;
    int kind_;
    nbody::Force force_;
    std::vector<int64_t> offsets_;
    int k_ = 1;
    bool ensemble_ = false;
    Particles<double> P_;
    DBuf<int64_t> d_offsets_;
    nbody::Plan plan_;
    nbody::QueuePlan qplan_;
    int64_t item_len_ = 0;
    int smax_ = 1;
    DBuf<nbody::WorkItem> items_;
    DBuf<unsigned char> ns_;
    DBuf<int> counters_;
    DBuf<double> part_, gpart_, cpart_;
    DBuf<float4> A_, B_;
    PinnedBuf host_;
    StepGraphs graphs_;
    bool precalc_done_ = false;
    bool grav_actual_ = false;
    bool tuned_ = false;
    int tune_candidates_ = 0;
    int64_t steps_ = 0;
    double time_ = 0.0;
    // adjoint
    bool recording_ = false;
    int ck_every_ = 1;
    int64_t rec_steps_ = 0;
    std::vector<DBuf<double>> ck_;  // (x, v, a) every ck_every_ steps
    Particles<double> R_;           // scratch state for segment recomputation
    DBuf<double> xs_;               // positions of one segment
    DBuf<double> tape_;             // positions after steps 1 .. K of the recorded run's open segment
    DBuf<double> Xb_, Vb_, Ab_, At_, gm_, vpart_;
    DBuf<float4> Uf_, Ap_, Bp_;
    nbody::Plan vplan_;
    // optional grid (direct-sum model with a grid section)
    bool has_grid_ = false;
    int nx_ = 0;
    int64_t ncell_ = 0;
    pm::Geo<double> geo_{};
    double scale_ = 0, inv_scale_ = 0, qmax_ = 0;  // qmax_: max |m| * scale_
    int grid_e_ = 0;
    DBuf<unsigned long long> mfix_;
    DBuf<double> gmass_, gphi_, gacc_;
    DBuf<unsigned> bkeys_, bkeys2_, bidx_, border_;
    DBuf<int> bcount_, bstart_;
    DBuf<char> btemp_;
};

// Particle-mesh strategy (RectWaveIntegrationSolver), state in `real`.
template <class real> class WaveSolver final : public Solver {
public:
    WaveSolver(const rrg_params& p, hipStream_t s, int64_t n, const double* pos, const double* vel, const double* mass)
        : prm_(p), s_(s) {
        const rrg_grid_params& g = p.grid;
        require(g.nx >= 2 && g.nx <= 1290, "rrg_create: grid nx must be in [2, 1290]");
        require(g.dx > 0, "rrg_create: grid dx must be positive");
        require(p.wave.iterations >= 0 && p.wave.setup_iterations >= 0, "rrg_create: negative wave iterations");
        require(p.wave.steady_check_step > 0, "rrg_create: steady_check_step must be positive");
        require(p.sort_interval >= 0, "rrg_create: sort_interval must be non-negative");
        sort_interval_ = p.sort_interval > 0 ? p.sort_interval : kSortInterval;
        require(p.boundary >= RRG_BC_ZERO_DIRICHLET && p.boundary <= RRG_BC_MONOPOLE_RADIATIVE,
                "rrg_create: unknown boundary");
        require(g.bc_frac >= 0 && g.bc_frac < 1, "rrg_create: grid bc_frac must be in [0, 1)");
        nx_ = g.nx;
        ncell_ = int64_t(nx_) * nx_ * nx_;
        periodic_ = p.boundary == RRG_BC_PERIODIC;
        geo_ = pm::make_geo<real>(nx_, g.dx, g.bc_frac);
        const real dxr = (real)g.dx;
        const real C = (real)p.wave.wave_speed;
        const real DT = (real)p.wave.dt;
        k_ = pm::wave_coef<real>((real)p.wave.diss_base, C, DT, dxr);
        model_ = pm::make_model<real>(nx_, g.dx, g.bc_frac, p.wave.diss_base, p.wave.diss_extra, p.wave.wave_speed,
                                      p.wave.dt, p.boundary);
        // set_boundary_periodic(1., ...) (rrgsim_integration_rect.cuh:65-69; wave.cu:214-228)
        {
            const real L = nx_ * dxr;
            rho_mean_ = (real)1. / (L * L * L);
        }
        tol_ = (real)p.wave.tolerance;
        // several stencil iterations per launch where a launch shape exists;
        // such a launch writes two grids other than its inputs, hence a fourth
        fused_ = p.wave_fused != 0 && pm::wave_fused_levels<real>(model_, 3) > 1;
        nbuf_ = fused_ ? 4 : 3;
        // Fixed-point exponent: the largest e with sum|m| * 2^e below 2^62.
        double msum = 0;
        for (int64_t i = 0; i < n; ++i) msum += std::fabs(mass[i]);
        if (p.fixed_exponent > 0) {
            e_ = p.fixed_exponent;
        } else {
            int ex = 0;
            std::frexp(std::max(msum, 1e-300), &ex);  // msum in [2^(ex-1), 2^ex)
            e_ = 61 - ex;
            if (e_ > 1000) e_ = 1000;
        }
        require(std::ldexp(std::max(msum, 0.0), e_) < std::ldexp(1.0, 62), "rrg_create: fixed_exponent too large for total mass");
        scale_ = std::ldexp(1.0, e_);
        inv_scale_ = std::ldexp(1.0, -e_);
        {
            double mmax = 0;  // the state's masses are `real`: bound them as stored
            for (int64_t i = 0; i < n; ++i) mmax = std::max(mmax, std::fabs((double)(real)mass[i]));
            qmax_ = mmax * scale_;
        }

        P_.alloc(n);
        upload_vec<real>(s_, n, pos, P_.X.p);
        upload_vec<real>(s_, n, vel, P_.V.p);
        upload_scalar<real>(s_, n, mass, P_.M.p);
        // spatial ordering buffers (particles are kept sorted by CIC cell)
        x_first_ = P_.X.p;
        P2_.alloc(n);
        id_.alloc(n);
        id2_.alloc(n);
        keys_.alloc(n); keys2_.alloc(n); idx_.alloc(n); idx2_.alloc(n);
        {
            std::vector<int> ident(n);
            for (int64_t i = 0; i < n; ++i) ident[i] = int(i);
            RRG_HIP(hipMemcpyAsync(id_.p, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice, s_));
            RRG_HIP(hipStreamSynchronize(s_));
        }
        end_bit_ = 1;
        while ((int64_t(1) << end_bit_) < ncell_) ++end_bit_;
        pm::radix_sort_pairs(s_, nullptr, sort_bytes_, keys_.p, keys2_.p, idx_.p, idx2_.p, n, end_bit_);
        sort_temp_.alloc(std::max<size_t>(sort_bytes_, 1));
        zero_async(s_, P_.S2.p, P_.S2.bytes());
        zero_async(s_, P_.A.p, P_.A.bytes());
        zero_async(s_, P_.W.p, P_.W.bytes());
        zero_async(s_, P_.G.p, P_.G.bytes());
        mfix_.alloc(ncell_);
        mass_.alloc(ncell_);
        src_.alloc(ncell_);
        // Source segment flags for the fused stencil (periodic grids): measured
        // faster where the stencil is bound by its memory traffic (FP64 from
        // 128^3, FP32 from 256^3: 4-14% per iteration on the disk, which
        // leaves ~90% of the rows without mass), slower on 64^3 and on FP32
        // 128^3, where it is bound by its instruction issue.
        if (fused_ && (nx_ >= 256 || (sizeof(real) == 8 && nx_ >= 128))) {
            occ_.alloc(size_t(ncell_ / 64));
            src_empty_.alloc(1);
        }
        for (int i = 0; i < nbuf_; ++i) phi_[i].alloc(ncell_);
        if (fused_ && pm::wave_fused_pairs<real>(model_, 4))
            for (auto& q : pair_) q.alloc(2 * size_t(ncell_));
        gacc_.alloc(4 * ncell_);
        maxbits_.alloc(2 * pm::kCheckSlots);
        ctl_.alloc(1);
        zero_async(s_, ctl_.p, ctl_.bytes());
        zero_async(s_, mfix_.p, mfix_.bytes());
        zero_async(s_, mass_.p, mass_.bytes());
        zero_async(s_, src_.p, src_.bytes());
        for (int i = 0; i < nbuf_; ++i) zero_async(s_, phi_[i].p, phi_[i].bytes());  // initial potential: zero
        zero_async(s_, gacc_.p, gacc_.bytes());
        RRG_HIP(hipStreamSynchronize(s_));
    }

    void precalc() override {  // rrgsim_integration_rect.cuh:53-83
        graphs_.clear(s_);  // captured segments hold the boundary model by value (the centre of mass may change)
        sort_particles();
        deposit(false);
        if (model_.bc == RRG_BC_MONOPOLE_DIRICHLET || model_.bc == RRG_BC_MONOPOLE_RADIATIVE) {
            // set_boundary_monopole(1., calc_mass_center(...), ...) (rrgsim_integration_rect.cuh:57-63)
            if (com_scratch_.n == 0) com_scratch_.alloc(5);
            pm::center_of_mass<real>(s_, P_.view(), e_, com_scratch_.p, model_.com);
            model_.mass = (real)1.;
        }
        int launches = 0;
        last_iters_ = solve(prm_.wave.setup_iterations, &launches);
        grav_actual_ = false;
        acc_actual_ = false;
        precalc_done_ = true;
        fresh_ = true;
        recording_ = false;
    }

    void step(int nsteps) override {  // rrgsim_integration_rect.cuh:99-112
        if (!precalc_done_) fail(RRG_ERR_STATE, "rrg_step before rrg_precalc");
        const real dt = (real)prm_.sim.dt_dynamics;
        // Steps are replayed from captured graphs (one set per buffer
        // arrangement); a steady-state check is then decided on the device.
        // A run recorded for the adjoint with checks keeps the host's
        // decisions, which give it each step's iteration count.
        const bool graph = prm_.use_graphs && !(tol_ > (real)0 && recording_);
        for (int k = 0; k < nsteps; ++k) {
            // a recorded run keeps one particle order so its tapes line up
            if (++since_sort_ >= sort_interval_ && !recording_) sort_particles();
            enqueue_step(dt, graph);
            ++steps_;
            time_ += prm_.sim.dt_dynamics;
            if (recording_) {
                rec_iters_.push_back(last_iters_);
                ++rec_steps_;
                if (rec_steps_ % ck_every_ == 0) store_checkpoint(ck_.emplace_back());
            }
        }
        if (nsteps > 0) {
            acc_actual_ = true;
            grav_actual_ = false;
            fresh_ = false;
        }
    }

    // Records the run from the state right after precalc: the particles are
    // sorted once and stay in that order, a_0 is interpolated from the
    // precalc grid (the integrate loop's sync after precalc), and
    // (x, v, a, phi_prev, phi_curr) is kept every `every` steps.
    void adjoint_begin(int every) override {
        require(every >= 1, "rrg_adjoint_begin: checkpoint_every must be >= 1");
        if (!precalc_done_) fail(RRG_ERR_STATE, "rrg_adjoint_begin before rrg_precalc");
        if (!fresh_) fail(RRG_ERR_STATE, "rrg_adjoint_begin: a PM context records from the state right after rrg_precalc");
        require(periodic_ || model_.bc == RRG_BC_ZERO_DIRICHLET,
                "rrg_adjoint_begin: the PM adjoint covers periodic and zero-Dirichlet boundaries");
        sort_particles();
        sync();
        ck_every_ = every;
        rec_steps_ = 0;
        ck_.clear();
        rec_iters_.assign(1, last_iters_);
        store_checkpoint(ck_.emplace_back());
        recording_ = true;
    }

    // Reverse sweep: per checkpoint segment (last first) the steps are
    // recomputed on the live buffers, keeping x' and phi_I of each step, and
    // then reversed; the precalc build closes the sweep.  The live state is
    // restored afterwards.
    void adjoint(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) override {
        adjoint_impl(gxT, gvT, gx0, gv0, gm, false);
    }
    void adjoint_dev(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm) override {
        adjoint_impl(gxT, gvT, gx0, gv0, gm, true);
    }

    void adjoint_impl(const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm, bool dev) {
        if (!recording_) fail(RRG_ERR_STATE, "rrg_adjoint without rrg_adjoint_begin");
        const int64_t n = P_.n;
        adj_alloc();
        std::vector<int> id;
        if (dev) {
            if (gxT) layout::rows_in<double>(s_, n, gxT, Xb_.p, id_.p);
            else zero_async(s_, Xb_.p, Xb_.bytes());
            if (gvT) layout::rows_in<double>(s_, n, gvT, Vb_.p, id_.p);
            else zero_async(s_, Vb_.p, Vb_.bytes());
        } else {
            id = ids();
            std::vector<double> tmp(3 * n, 0.0), zero(3 * n, 0.0);
            if (gxT) permute_vec_in(n, id, gxT, tmp.data());
            upload_vec<double>(s_, n, gxT ? tmp.data() : zero.data(), Xb_.p);
            if (gvT) permute_vec_in(n, id, gvT, tmp.data());
            upload_vec<double>(s_, n, gvT ? tmp.data() : zero.data(), Vb_.p);
        }
        zero_async(s_, Ab_.p, Ab_.bytes());
        zero_async(s_, gm_.p, gm_.bytes());
        for (int i = 0; i < nbuf_; ++i) zero_async(s_, mu_[i].p, mu_[i].bytes());
        aip_ = 0;
        aic_ = 1;
        // keep the live state
        const int ip0 = ip_, ic0 = ic_, it0 = last_iters_;
        store_checkpoint(backup_);
        const real dt = (real)prm_.sim.dt_dynamics;
        const bool graph = prm_.use_graphs && !(tol_ > (real)0);
        const int T = int(rec_steps_), K = ck_every_;
        const int nseg = (T + K - 1) / K;
        const size_t v3 = sizeof(real) * 3 * n, vg = sizeof(real) * ncell_;
        for (int c = nseg - 1; c >= 0; --c) {
            const int s0 = c * K, len = std::min(T, s0 + K) - s0;
            load_checkpoint(ck_[c]);
            for (int q = 1; q <= len; ++q) {
                enqueue_step(dt, graph, rec_iters_[s0 + q]);
                RRG_HIP(hipMemcpyAsync(xs_.p + size_t(q) * 3 * n, P_.X.p, v3, hipMemcpyDeviceToDevice, s_));
                RRG_HIP(hipMemcpyAsync(phis_.p + size_t(q) * ncell_, cur(), vg, hipMemcpyDeviceToDevice, s_));
            }
            for (int q = len; q >= 1; --q)
                reverse_build(xs_.p + size_t(q) * 3 * n, phis_.p + size_t(q) * ncell_, rec_iters_[s0 + q], false);
        }
        // precalc build at x_0 (a_0 interpolated from it, phi from the setup solve)
        reverse_build(ck_[0].p, ck_[0].p + 9 * n + ncell_, rec_iters_[0], true);
        ip_ = ip0;
        ic_ = ic0;
        load_checkpoint(backup_);
        last_iters_ = it0;
        if (dev) {
            if (gx0) layout::rows_out<double>(s_, n, Xb_.p, gx0, id_.p);
            if (gv0) layout::rows_out<double>(s_, n, Vb_.p, gv0, id_.p);
            if (gm) layout::scalar_out<double>(s_, n, gm_.p, gm, id_.p);
            return;
        }
        std::vector<double> tmp(3 * n);
        if (gx0) { download_vec<double>(s_, n, Xb_.p, tmp.data()); unpermute_vec(n, id, tmp.data(), gx0); }
        if (gv0) { download_vec<double>(s_, n, Vb_.p, tmp.data()); unpermute_vec(n, id, tmp.data(), gv0); }
        if (gm) {
            download_scalar<double>(s_, n, gm_.p, tmp.data());
            for (int64_t i = 0; i < n; ++i) gm[id[i]] = tmp[i];
        }
        RRG_HIP(hipStreamSynchronize(s_));
    }

    void sync() override {  // rrgsim_integration_rect.cuh:85-97
        if (!grav_actual_) pm::ngp_phi<real>(s_, P_.view(), geo_, cur());
        if (!acc_actual_) pm::gather_phi<real>(s_, P_.view(), geo_, cur(), (real)0, false);
        grav_actual_ = acc_actual_ = true;
    }

    // The device arrays hold the particles in cell order; id_ maps each slot
    // back to the caller's index.
    void get_state(double* pos, double* vel, double* acc, double* gr) override {
        const int64_t n = P_.n;
        const std::vector<int> id = ids();
        std::vector<double> tmp(3 * n);
        if (pos) { download_vec<real>(s_, n, P_.X.p, tmp.data()); unpermute_vec(n, id, tmp.data(), pos); }
        if (vel) { download_vec<real>(s_, n, P_.V.p, tmp.data()); unpermute_vec(n, id, tmp.data(), vel); }
        if (acc) { download_vec<real>(s_, n, P_.A.p, tmp.data()); unpermute_vec(n, id, tmp.data(), acc); }
        if (gr) {
            download_scalar<real>(s_, n, P_.G.p, tmp.data());
            for (int64_t i = 0; i < n; ++i) gr[id[i]] = tmp[i];
        }
    }

    void set_state(const double* pos, const double* vel) override {
        const int64_t n = P_.n;
        const std::vector<int> id = ids();
        std::vector<double> tmp(3 * n);
        if (pos) { permute_vec_in(n, id, pos, tmp.data()); upload_vec<real>(s_, n, tmp.data(), P_.X.p); }
        if (vel) { permute_vec_in(n, id, vel, tmp.data()); upload_vec<real>(s_, n, tmp.data(), P_.V.p); }
        grav_actual_ = acc_actual_ = false;
        fresh_ = recording_ = false;
    }

    void get_state_dev(double* pos, double* vel, double* acc) override {
        const int64_t n = P_.n;
        if (pos) layout::rows_out<real>(s_, n, P_.X.p, pos, id_.p);
        if (vel) layout::rows_out<real>(s_, n, P_.V.p, vel, id_.p);
        if (acc) layout::rows_out<real>(s_, n, P_.A.p, acc, id_.p);
    }

    void set_state_dev(const double* pos, const double* vel) override {
        const int64_t n = P_.n;
        if (pos) layout::rows_in<real>(s_, n, pos, P_.X.p, id_.p);
        if (vel) layout::rows_in<real>(s_, n, vel, P_.V.p, id_.p);
        grav_actual_ = acc_actual_ = false;
        fresh_ = recording_ = false;
    }

    void conservation(rrg_conservation* out) override {
        sync();
        conservation_common<real>(s_, P_.view(), cpart_, host_, time_, out);
    }

    void conservation_systems(int k, rrg_conservation* out) override {
        require(k == 1, "rrg_get_conservation_systems: a PM context holds one system");
        conservation(out);
    }

    // The steps keep no cell-mass grid (only its source term): the cell mass
    // is deposited again from the current positions, which are those of the
    // last deposit (the kick does not move particles; set_state would).
    void get_grid(double* mass, double* phi, double* acc) override {
        if (mass) {
            pm::predict_deposit<real>(s_, P_.view(), (real)0, false, geo_, scale_, mfix_.p, qmax_);
            pm::fixed_to_mass<real>(s_, nx_, mfix_.p, inv_scale_, mass_.p);  // leaves mfix_ zero
            download_scalar<real>(s_, ncell_, mass_.p, mass);
        }
        if (phi) download_scalar<real>(s_, ncell_, cur(), phi);
        if (acc) {
            // steps keep no gradient grid: -grad phi of the current potential
            pm::gradient<real>(s_, nx_, geo_.dx, cur(), gacc_.p);
            std::vector<real> tmp(4 * ncell_);
            RRG_HIP(hipMemcpyAsync(tmp.data(), gacc_.p, sizeof(real) * 4 * ncell_, hipMemcpyDeviceToHost, s_));
            RRG_HIP(hipStreamSynchronize(s_));
            for (int64_t c = 0; c < ncell_; ++c)
                for (int q = 0; q < 3; ++q) acc[3 * c + q] = (double)tmp[4 * c + q];
        }
    }

    void stats(rrg_stats* out) override {
        std::memset(out, 0, sizeof(*out));
        out->n = P_.n;
        out->steps = steps_;
        out->time = time_;
        out->last_wave_iterations = solved_iterations();
        out->fixed_exponent = e_;
        out->systems = 1;
        out->graph_launches = graphs_.launches();
    }

private:
    // Steps between spatial re-sorts.  The deposit's LDS box absorbs the
    // drift: on the 10^6-particle disk at 128^3 its time stays flat for ~48
    // steps, so a re-sort (~150 us) every 48 steps costs least overall.
    static constexpr int kSortInterval = 48;

    template <class F> void segment(int stage, int64_t launches, bool graph, uint64_t key, F&& f) {
        const int t = timer.start(s_, stage);
        if (graph) graphs_.run(s_, key, f);
        else f();
        timer.stop(s_, t, launches);
    }

    uint64_t graph_base() const {
        return (uint64_t(ip_) * 4 + uint64_t(ic_)) * 16 + (P_.X.p == x_first_ ? 0 : 8);
    }

    // One PM step as segments: predictor + deposit, source, the wave
    // iterations, gradient, gather + kick.  With `graph` every segment is
    // replayed from its captured graph and stage timing events sit between
    // segment launches.
    void enqueue_step(real dt, bool graph, int fixed_iters = -1) {
        const uint64_t b0 = graph_base();
        const bool dev_check = graph && tol_ > (real)0;  // steady check decided on the device
        const int iters = prm_.wave.iterations;
        // the solve's launches from arrangement (ip, ic); `launch` false: the rotation only
        auto solve_launches = [&](int& ip, int& ic, bool launch) {
            return dev_check ? wave_solve_device(iters, ip, ic, launch) : wave_launches(iters, ip, ic, nullptr, launch);
        };
        if (graph && !timer.on()) {
            // without stage timing the whole step is one graph launch
            const int ip0 = ip_, ic0 = ic_;
            solve_launches(ip_, ic_, false);
            graphs_.run(s_, b0 + 5, [&] {
                pm::predict_deposit<real>(s_, P_.view(), dt, true, geo_, scale_, mfix_.p, qmax_);
                pm::fixed_to_source<real>(s_, nx_, mfix_.p, inv_scale_, geo_.dx, rho_mean_, periodic_, k_.k_f,
                                          nullptr, src_.p, layered(), occ_.p, src_empty_.p);
                int ip = ip0, ic = ic0;  // from the arrangement before the rotation
                solve_launches(ip, ic, true);
                pm::gather_phi<real>(s_, P_.view(), geo_, cur(), dt, true);
            });
            last_iters_ = iters;
            iters_on_device_ = dev_check;
            return;
        }
        segment(RRG_STAGE_DEPOSIT, 1, graph, b0 + 0, [&] {
            pm::predict_deposit<real>(s_, P_.view(), dt, true, geo_, scale_, mfix_.p, qmax_);
        });
        segment(RRG_STAGE_SOURCE, 1, graph, b0 + 1, [&] {
            pm::fixed_to_source<real>(s_, nx_, mfix_.p, inv_scale_, geo_.dx, rho_mean_, periodic_, k_.k_f, nullptr,
                                      src_.p, layered(), occ_.p, src_empty_.p);
        });
        iters_on_device_ = false;
        if (graph) {
            const int ip0 = ip_, ic0 = ic_;
            const int launches = solve_launches(ip_, ic_, false);
            segment(RRG_STAGE_WAVE, launches, true, b0 + 2, [&] {
                int ip = ip0, ic = ic0;  // captured from the arrangement before the rotation
                solve_launches(ip, ic, true);
            });
            last_iters_ = iters;
            iters_on_device_ = dev_check;
        } else if (fixed_iters >= 0) {  // a recorded count, no checks (recomputation)
            const int t = timer.start(s_, RRG_STAGE_WAVE);
            const int launches = wave_launches(fixed_iters, ip_, ic_);
            last_iters_ = fixed_iters;
            timer.stop(s_, t, launches);
        } else {
            const int t = timer.start(s_, RRG_STAGE_WAVE);
            int launches = 0;
            last_iters_ = solve(iters, &launches);
            timer.stop(s_, t, launches);
        }
        // -grad phi is computed inside the gather (no gradient grid is written)
        const uint64_t b1 = graph_base();
        segment(RRG_STAGE_GATHER, 1, graph, b1 + 4, [&] { pm::gather_phi<real>(s_, P_.view(), geo_, cur(), dt, true); });
    }

    // iterations the next launch runs when `remaining` are left
    int levels(int remaining) const { return fused_ ? pm::wave_fused_levels<real>(model_, remaining) : 1; }

    // `iters` stencil iterations without a check; returns the launches.
    // `halt`: the device stop flag of a solve checked on the device;
    // `launch` false: only the buffer rotation the launches would leave.
    // Consecutive fused launches with a pair-grid shape hand their two levels
    // over in pair_[0] / pair_[1] (alternately); the stretch's first launch
    // reads and its last writes the split grids of the rotation.  `check`
    // (the maxbits pair): the last launch is a checking one, on split grids.
    int wave_launches(int iters, int& ip, int& ic, const int* halt = nullptr, bool launch = true,
                      void* check = nullptr) {
        int launches = 0;
        std::vector<int> seq;  // iterations of each launch
        for (int r = iters; r > 0; r -= seq.back()) seq.push_back(levels(r));
        auto paired = [&](size_t i) {  // launches i and i+1 exchange a pair grid
            return pair_[0].p != nullptr && i + 1 < seq.size() && seq[i] == seq[i + 1] &&
                   pm::wave_fused_pairs<real>(model_, seq[i]) && !(check != nullptr && i + 2 == seq.size());
        };
        pm::wave_schedule(
            iters, [&](int r) { return levels(r); }, ip, ic,
            [&](int out, int cur, int prev) {
                if (launch)
                    pm::wave_iteration<real>(s_, model_, k_, phi_[out].p, phi_[cur].p, phi_[prev].p, src_.p,
                                             size_t(launches) + 1 == seq.size() ? check : nullptr, halt);
                ++launches;
            },
            [&](int lv, int o1, int o2, int cur, int prev) {
                if (launch) {
                    const size_t i = size_t(launches);
                    const bool pin = i > 0 && paired(i - 1), pout = paired(i);
                    real* qa = pout ? pair_[i & 1].p : phi_[o1].p;
                    const real* p = pin ? pair_[(i - 1) & 1].p : phi_[cur].p;
                    pm::wave_fused<real>(s_, model_, k_, lv, qa, phi_[o2].p, p, phi_[prev].p, src_.p, halt,
                                         occ_.p, src_empty_.p, pin, pout, i + 1 == seq.size() ? check : nullptr);
                }
                ++launches;
            });
        return launches;
    }

    // solve_wave_equation (wave.cu:230-297) as a fixed launch sequence with
    // the steady check decided on the device: each stretch between checks and
    // the checking iteration after it (iter > 0, iter % steady_check_step ==
    // 0, after the rotation) run as in wave_launches, the checking iteration
    // being the last level of the stretch's last launch, which reduces its
    // maxima (a fused launch checks its last level against the level before,
    // a single launch its one level against its input: the same values);
    // the hold rule follows on the device.  After convergence every launch
    // returns at once and a last launch leaves phi^(n-1), phi^n where the
    // full schedule ends (ip, ic).  Returns the launches; `launch` false:
    // only the rotation.
    int wave_solve_device(int iters, int& ip, int& ic, bool launch) {
        const int cs = prm_.wave.steady_check_step;
        const int* halt = reinterpret_cast<const int*>(reinterpret_cast<const char*>(ctl_.p) +
                                                       offsetof(pm::SolveCtl, done));
        int launches = 0, iter = 0;
        if (launch) pm::solve_init<real>(s_, ctl_.p, maxbits_.p);
        ++launches;
        while (iter < iters) {
            int c = ((iter + cs - 1) / cs) * cs;  // the next check follows iteration c
            if (c == 0) c = cs;
            if (c >= iters) {  // no check before the limit
                launches += wave_launches(iters - iter, ip, ic, halt, launch);
                break;
            }
            launches += wave_launches(c - iter + 1, ip, ic, halt, launch, maxbits_.p);
            iter = c + 1;
            if (launch) pm::solve_decide<real>(s_, ctl_.p, maxbits_.p, tol_, prm_.wave.steady_hold_steps, iter, ip, ic);
            ++launches;
        }
        if (launch) {
            pm::PhiBufs<real> b{};
            for (int q = 0; q < nbuf_; ++q) b.p[q] = phi_[q].p;
            pm::solve_fixup<real>(s_, ctl_.p, b, ip, ic, ncell_);
        }
        return launches + 1;
    }

    // Stencil iterations run by the last solve (a device-checked one is read
    // back from its control block).
    int solved_iterations() {
        if (!iters_on_device_) return last_iters_;
        pm::SolveCtl c{};
        RRG_HIP(hipMemcpyAsync(&c, ctl_.p, sizeof(c), hipMemcpyDeviceToHost, s_));
        RRG_HIP(hipStreamSynchronize(s_));
        return c.done ? c.iters : last_iters_;
    }

    real* cur() { return phi_[ic_].p; }

    // the per-cell source coefficient applies off-periodic (absorbing layer)
    const pm::WaveModel<real>* layered() const { return periodic_ ? nullptr : &model_; }

    std::vector<int> ids() {
        std::vector<int> id(P_.n);
        RRG_HIP(hipMemcpyAsync(id.data(), id_.p, sizeof(int) * P_.n, hipMemcpyDeviceToHost, s_));
        RRG_HIP(hipStreamSynchronize(s_));
        return id;
    }

    // (x, v, a, phi_prev, phi_curr) of the live state into / out of `ck`.
    void store_checkpoint(DBuf<real>& ck) {
        const int64_t n = P_.n;
        if (ck.n != size_t(9 * n + 2 * ncell_)) ck.alloc(9 * n + 2 * ncell_);
        const size_t v3 = sizeof(real) * 3 * n, vg = sizeof(real) * ncell_;
        RRG_HIP(hipMemcpyAsync(ck.p, P_.X.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 3 * n, P_.V.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 6 * n, P_.A.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 9 * n, phi_[ip_].p, vg, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(ck.p + 9 * n + ncell_, phi_[ic_].p, vg, hipMemcpyDeviceToDevice, s_));
    }

    void load_checkpoint(const DBuf<real>& ck) {
        const int64_t n = P_.n;
        const size_t v3 = sizeof(real) * 3 * n, vg = sizeof(real) * ncell_;
        RRG_HIP(hipMemcpyAsync(P_.X.p, ck.p, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(P_.V.p, ck.p + 3 * n, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(P_.A.p, ck.p + 6 * n, v3, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(phi_[ip_].p, ck.p + 9 * n, vg, hipMemcpyDeviceToDevice, s_));
        RRG_HIP(hipMemcpyAsync(phi_[ic_].p, ck.p + 9 * n + ncell_, vg, hipMemcpyDeviceToDevice, s_));
    }

    void adj_alloc() {
        const int64_t n = P_.n;
        if (Xb_.n != size_t(3 * n)) {
            Xb_.alloc(3 * n); Vb_.alloc(3 * n); Ab_.alloc(3 * n); gm_.alloc(n);
            Gb_.alloc(3 * ncell_);
            ascale_.alloc(1);
            zero_async(s_, ascale_.p, ascale_.bytes());
            for (int i = 0; i < nbuf_; ++i) mu_[i].alloc(ncell_);
            sbar_.alloc(ncell_);
            gtmp_.alloc(4 * ncell_);
            zero_async(s_, Gb_.p, Gb_.bytes());
            if (!periodic_) {
                for (auto& b : kc_) b.alloc(ncell_);
                pm::cell_coefficients<real>(s_, model_, kc_[0].p, kc_[1].p, kc_[2].p, kc_[3].p);
            }
        }
        const size_t K1 = size_t(ck_every_) + 1;
        if (xs_.n < K1 * 3 * n) xs_.alloc(K1 * 3 * n);
        if (phis_.n < K1 * ncell_) phis_.alloc(K1 * ncell_);
    }

    // Reverse of one grid build (a step, or with `closing` the precalc build)
    // given the positions x' it deposited from and the potential phi_I its
    // gradient read.  Gb is zero on entry and left zero.
    void reverse_build(const real* x, const real* phi, int iters, bool closing) {
        const int64_t n = P_.n;
        const double dt = prm_.sim.dt_dynamics;
        int t = timer.start(s_, RRG_STAGE_ADJ_FORCE);
        pm::gradient<real>(s_, nx_, geo_.dx, phi, gtmp_.p);
        pm::adj_interp<real>(s_, n, x, geo_, gtmp_.p, closing ? 0.0 : 0.5 * dt, Xb_.p, Vb_.p, Ab_.p, Gb_.p,
                             ascale_.p);
        pm::adj_gradient_t<real>(s_, nx_, geo_.dx, periodic_, Gb_.p, ascale_.p, mu_[aic_].p);
        pm::fill_u64(s_, reinterpret_cast<unsigned long long*>(Gb_.p), 3 * size_t(ncell_), 0ull);
        timer.stop(s_, t, 4);
        if (iters == 0) {
            zero_async(s_, sbar_.p, sbar_.bytes());
        } else {
            // the transposed iterations, several per launch like the forward's
            auto levels = [&](int r) { return fused_ ? pm::adj_wave_levels<real>(nx_, periodic_, r) : 1; };
            auto run = [&](int& ip, int& ic, bool launch) {
                int launches = 0;
                pm::wave_schedule(
                    iters, levels, ip, ic,
                    [&](int out, int cur, int prev) {
                        if (launch)
                            pm::adj_wave<real>(s_, nx_, periodic_, k_, mu_[out].p, mu_[cur].p, mu_[prev].p, sbar_.p,
                                               launches == 0, periodic_ ? nullptr : kc_[0].p, kc_[1].p, kc_[2].p);
                        ++launches;
                    },
                    [&](int lv, int o1, int o2, int cur, int prev) {
                        if (launch)
                            pm::adj_wave_fused<real>(s_, nx_, k_, lv, mu_[o1].p, mu_[o2].p, mu_[cur].p, mu_[prev].p,
                                                     sbar_.p, launches == 0);
                        ++launches;
                    });
                return launches;
            };
            t = timer.start(s_, RRG_STAGE_ADJ_WAVE);
            int launches = 0;
            if (prm_.use_graphs) {
                const uint64_t key = (uint64_t(1) << 40) | (uint64_t(iters) << 4) | uint64_t(aip_ * 4 + aic_);
                const int ip0 = aip_, ic0 = aic_;
                launches = run(aip_, aic_, false);  // the rotation the launches leave
                graphs_.run(s_, key, [&] {
                    int ip = ip0, ic = ic0;
                    run(ip, ic, true);
                });
            } else {
                launches = run(aip_, aic_, true);
            }
            timer.stop(s_, t, launches);
        }
        const real dx3 = geo_.dx * geo_.dx * geo_.dx;
        const double alpha = 4 * 3.14159265358979 * (double)k_.k_f / (double)dx3;  // d source / d cell mass
        t = timer.start(s_, RRG_STAGE_ADJ_PARTICLE);
        pm::adj_cic<real>(s_, n, x, P_.M.p, geo_, sbar_.p, alpha, dt, closing, Xb_.p, Vb_.p, Ab_.p, gm_.p,
                          periodic_ ? nullptr : kc_[3].p);
        timer.stop(s_, t, 1);
    }

    // Reorder the particle arrays by CIC base cell (radix sort + gather), so
    // the deposit's wave-level run merging sees long runs of equal cells.
    void sort_particles() {
        int t = timer.start(s_, RRG_STAGE_SORT);
        pm::cell_keys<real>(s_, P_.view(), geo_, keys_.p, idx_.p);
        size_t bytes = sort_bytes_;
        pm::radix_sort_pairs(s_, sort_temp_.p, bytes, keys_.p, keys2_.p, idx_.p, idx2_.p, P_.n, end_bit_);
        pm::permute<real>(s_, P_.view(), P2_.view(), idx2_.p, id_.p, id2_.p);
        P_.swap(P2_);
        std::swap(id_, id2_);
        timer.stop(s_, t, 3);
        since_sort_ = 0;
    }

    void deposit(bool predict) {
        int t = timer.start(s_, RRG_STAGE_DEPOSIT);
        pm::predict_deposit<real>(s_, P_.view(), (real)prm_.sim.dt_dynamics, predict, geo_, scale_, mfix_.p, qmax_);
        timer.stop(s_, t, 1);
        t = timer.start(s_, RRG_STAGE_SOURCE);
        pm::fixed_to_source<real>(s_, nx_, mfix_.p, inv_scale_, geo_.dx, rho_mean_, periodic_, k_.k_f, nullptr, src_.p,
                                  layered(), occ_.p, src_empty_.p);
        timer.stop(s_, t, 1);
    }

    // solve_wave_equation (wave.cu:230-297) with the SteadyStateChecker
    // (wave_steady_checker.cu:95-143).  A tolerance <= 0 can never be met
    // (the error is >= 0), so the checks are skipped and exactly max_iter
    // iterations run.
    int solve(int max_iter, int* launches) {
        using MB = std::conditional_t<std::is_same_v<real, float>, unsigned int, unsigned long long>;
        iters_on_device_ = false;
        const bool checking = tol_ > (real)0;
        const int check_step = prm_.wave.steady_check_step;
        int hold_count = 0;
        int iter = 0;
        *launches = 0;
        while (iter < max_iter) {
            // the iterations before the next check (iter > 0, iter % check_step == 0) run unchecked
            int c = max_iter;
            if (checking) {
                c = ((iter + check_step - 1) / check_step) * check_step;
                if (c == 0) c = check_step;
            }
            if (!checking || c >= max_iter) {  // no check before the limit
                *launches += wave_launches(max_iter - iter, ip_, ic_);
                iter = max_iter;
                break;
            }
            // the stretch and the checking iteration after it, the check
            // reduced by the launch that completes that iteration
            pm::fill_u64(s_, maxbits_.p, 2 * pm::kCheckSlots, 0ull);
            *launches += wave_launches(c - iter + 1, ip_, ic_, nullptr, true, maxbits_.p);
            iter = c + 1;
            host_.ensure(sizeof(MB) * 2 * pm::kCheckSlots);
            RRG_HIP(hipMemcpyAsync(host_.p, maxbits_.p, sizeof(MB) * 2 * pm::kCheckSlots, hipMemcpyDeviceToHost, s_));
            RRG_HIP(hipStreamSynchronize(s_));
            MB bits[2];
            max_over_slots<MB>(host_.p, bits);
            real diff, absu;
            std::memcpy(&diff, &bits[0], sizeof(real));
            std::memcpy(&absu, &bits[1], sizeof(real));
            const real EPS = (real)1.E-30;
            const real err = diff / (absu + EPS);
            if ((double)err < (double)tol_) ++hold_count;
            else hold_count = 0;
            if (hold_count >= prm_.wave.steady_hold_steps) break;
        }
        return iter;
    }

    rrg_params prm_;
    hipStream_t s_;
    int nx_ = 0;
    int64_t ncell_ = 0;
    bool periodic_ = true;
    pm::Geo<real> geo_{};
    pm::WaveK<real> k_{};
    real rho_mean_ = 0, tol_ = 0;
    pm::WaveModel<real> model_{};
    DBuf<unsigned long long> com_scratch_;
    int e_ = 61;
    double scale_ = 0, inv_scale_ = 0, qmax_ = 0;  // qmax_: max |m| * scale_
    Particles<real> P_, P2_;
    DBuf<int> id_, id2_;
    DBuf<unsigned> keys_, keys2_, idx_, idx2_;
    DBuf<char> sort_temp_;
    size_t sort_bytes_ = 0;
    int end_bit_ = 32;
    int since_sort_ = 0;
    int sort_interval_ = kSortInterval;
    real* x_first_ = nullptr;  // P_.X at creation (the sort swaps buffers)
    StepGraphs graphs_;
    DBuf<unsigned long long> mfix_;
    DBuf<real> mass_, src_, phi_[4], gacc_;
    // (fused stencil) 64-cell segments of the source that hold mass, and the
    // source value of an empty cell; the stencil skips reading the others
    DBuf<int> occ_;
    DBuf<real> src_empty_;
    DBuf<real> pair_[2];  // pair grids between consecutive fused launches (wave_fused_pairs shapes)
    bool fused_ = false;  // several iterations per launch
    int nbuf_ = 3;       // potential grids in the rotation
    DBuf<unsigned long long> maxbits_;
    DBuf<pm::SolveCtl> ctl_;       // steady check decided on the device
    bool iters_on_device_ = false;  // the last solve's iteration count is in ctl_
    DBuf<double> cpart_;
    PinnedBuf host_;
    int ip_ = 0, ic_ = 1;  // indices of phi^{n-1} (prev) and phi^n (curr)
    bool precalc_done_ = false, grav_actual_ = false, acc_actual_ = false;
    bool fresh_ = false;  // state is the one precalc left (no step, no set_state since)
    int last_iters_ = 0;
    int64_t steps_ = 0;
    double time_ = 0.0;
    // adjoint
    bool recording_ = false;
    int ck_every_ = 1;
    int64_t rec_steps_ = 0;
    std::vector<int> rec_iters_;        // stencil iterations per grid build (0 = precalc)
    std::vector<DBuf<real>> ck_;        // (x, v, a, phi_prev, phi_curr) every ck_every_ steps
    DBuf<real> backup_, xs_, phis_;     // live state; x' and phi_I of one segment
    DBuf<double> Xb_, Vb_, Ab_, gm_;
    DBuf<long long> Gb_;             // interpolation adjoint, int64 fixed point
    DBuf<pm::AdjScale> ascale_;      // its scale
    DBuf<real> mu_[4], sbar_, gtmp_;
    DBuf<real> kc_[4];                  // off-periodic K_MAIN, K_ACTUAL, K_OLD, d source / d mass per cell
    int aip_ = 0, aic_ = 1;             // potential adjoints mu_{k+1}, mu_k
};

}  // namespace
}  // namespace rrg

struct rrg_ctx {
    rrg_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<rrg::Solver> solver;
};

using namespace rrg;

namespace {

void check_ctx(rrg_ctx* ctx) {
    if (!ctx || !ctx->solver) fail(RRG_ERR_INVALID, "null or destroyed context");
}

void check_finite(int64_t n, const double* a, const char* what) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) fail(RRG_ERR_INVALID, std::string(what) + " holds a non-finite value");
}

}  // namespace

RRG_API const char* rrg_last_error(void) { return g_last_error.c_str(); }

RRG_API const char* rrg_version(void) { return "rrg-mi355x 0.1 (gfx950)"; }

RRG_API int rrg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

RRG_API int rrg_params_default(rrg_params* p) {
    return guarded([&] {
        require(p != nullptr, "rrg_params_default: null params");
        io::params_default(p);
    });
}

RRG_API int rrg_params_finalize(rrg_params* p) {
    return guarded([&] {
        require(p != nullptr, "rrg_params_finalize: null params");
        io::params_finalize(p);
    });
}

RRG_API int rrg_params_parse_json(const char* path, rrg_params* p, rrg_ini_params* ini) {
    return guarded([&] {
        require(path && p, "rrg_params_parse_json: null argument");
        io::params_parse_json(path, p, ini);
    });
}

RRG_API int rrg_ini_read(const rrg_ini_params* ini, const char* base_dir, int64_t* n, double* pos, double* vel,
                         double* mass, double* soft2) {
    return guarded([&] {
        require(ini && n, "rrg_ini_read: null argument");
        io::ini_read(ini, base_dir, n, pos, vel, mass, soft2);
    });
}

RRG_API int rrg_create(const rrg_params* p, int64_t n, const double* pos, const double* vel, const double* mass,
                       const double* soft2, rrg_ctx** out) {
    return guarded([&] {
        require(p && out, "rrg_create: null argument");
        *out = nullptr;
        require(n > 0, "rrg_create: n must be positive");
        require(pos && vel && mass, "rrg_create: pos, vel and mass are required");
        require(n <= (int64_t(1) << 31) - 1, "rrg_create: n exceeds 2^31-1");
        check_finite(3 * n, pos, "pos");
        check_finite(3 * n, vel, "vel");
        check_finite(n, mass, "mass");
        if (soft2) check_finite(n, soft2, "soft2");
        int ndev = 0;
        RRG_HIP(hipGetDeviceCount(&ndev));
        require(p->device >= 0 && p->device < ndev, "rrg_create: invalid device ordinal");
        require(p->real_type == RRG_REAL_F64 || p->real_type == RRG_REAL_F32, "rrg_create: unknown real_type");
        require(std::isfinite(p->sim.dt_dynamics), "rrg_create: sim.dt_dynamics is not set");
        auto ctx = std::make_unique<rrg_ctx>();
        ctx->p = *p;
        io::params_finalize(&ctx->p);
        ctx->device = p->device;
        DeviceGuard dg(ctx->device);
        RRG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        try {
            if (ctx->p.use_wave_model) {
                require(std::isfinite(ctx->p.wave.dt) && ctx->p.wave.dt > 0, "rrg_create: wave dt is not positive");
                if (ctx->p.real_type == RRG_REAL_F32)
                    ctx->solver = std::make_unique<WaveSolver<float>>(ctx->p, ctx->stream, n, pos, vel, mass);
                else
                    ctx->solver = std::make_unique<WaveSolver<double>>(ctx->p, ctx->stream, n, pos, vel, mass);
            } else {
                require(ctx->p.real_type == RRG_REAL_F64,
                        "rrg_create: the direct-sum model keeps FP64 state (use force_kind for FP32 pair math)");
                ctx->solver = std::make_unique<NBodySolver>(ctx->p, ctx->stream, n, pos, vel, mass, soft2,
                                                            std::vector<int64_t>{0, n});
            }
        } catch (...) {
            ctx->solver.reset();
            (void)hipStreamDestroy(ctx->stream);
            throw;
        }
        *out = ctx.release();
    });
}

RRG_API int rrg_create_ensemble(const rrg_params* p, int32_t k, const int64_t* offsets, const double* pos,
                                const double* vel, const double* mass, const double* soft2, rrg_ctx** out) {
    return guarded([&] {
        require(p && out && offsets, "rrg_create_ensemble: null argument");
        *out = nullptr;
        require(k >= 1, "rrg_create_ensemble: k must be positive");
        require(offsets[0] == 0, "rrg_create_ensemble: offsets[0] must be 0");
        for (int g = 0; g < k; ++g)
            require(offsets[g + 1] > offsets[g], "rrg_create_ensemble: offsets must be strictly increasing");
        const int64_t n = offsets[k];
        require(n <= (int64_t(1) << 31) - 1, "rrg_create_ensemble: n exceeds 2^31-1");
        require(pos && vel && mass && soft2, "rrg_create_ensemble: pos, vel, mass and soft2 are required");
        check_finite(3 * n, pos, "pos");
        check_finite(3 * n, vel, "vel");
        check_finite(n, mass, "mass");
        check_finite(n, soft2, "soft2");
        int ndev = 0;
        RRG_HIP(hipGetDeviceCount(&ndev));
        require(p->device >= 0 && p->device < ndev, "rrg_create_ensemble: invalid device ordinal");
        require(p->real_type == RRG_REAL_F64, "rrg_create_ensemble: the direct-sum model keeps FP64 state");
        require(std::isfinite(p->sim.dt_dynamics), "rrg_create_ensemble: sim.dt_dynamics is not set");
        auto ctx = std::make_unique<rrg_ctx>();
        ctx->p = *p;
        io::params_finalize(&ctx->p);
        ctx->p.use_wave_model = 0;
        ctx->device = p->device;
        DeviceGuard dg(ctx->device);
        RRG_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        try {
            ctx->solver = std::make_unique<NBodySolver>(ctx->p, ctx->stream, n, pos, vel, mass, soft2,
                                                        std::vector<int64_t>(offsets, offsets + k + 1));
        } catch (...) {
            ctx->solver.reset();
            (void)hipStreamDestroy(ctx->stream);
            throw;
        }
        *out = ctx.release();
    });
}

RRG_API int rrg_destroy(rrg_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        {
            DeviceGuard dg(ctx->device);
            if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
            ctx->solver.reset();
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        }
        delete ctx;
    });
}

#define RRG_CTX_CALL(ctx, body)              \
    return guarded([&] {                     \
        check_ctx(ctx);                      \
        DeviceGuard dg((ctx)->device);       \
        body;                                \
    })

RRG_API int rrg_precalc(rrg_ctx* ctx) { RRG_CTX_CALL(ctx, ctx->solver->precalc()); }

RRG_API int rrg_step(rrg_ctx* ctx, int32_t nsteps) {
    RRG_CTX_CALL(ctx, {
        require(nsteps >= 0, "rrg_step: nsteps must be non-negative");
        ctx->solver->step(nsteps);
    });
}

RRG_API int rrg_sync(rrg_ctx* ctx) { RRG_CTX_CALL(ctx, ctx->solver->sync()); }

RRG_API int rrg_get_state(rrg_ctx* ctx, double* pos, double* vel, double* acc, double* grav) {
    RRG_CTX_CALL(ctx, ctx->solver->get_state(pos, vel, acc, grav));
}

RRG_API int rrg_set_state(rrg_ctx* ctx, const double* pos, const double* vel) {
    RRG_CTX_CALL(ctx, ctx->solver->set_state(pos, vel));
}

namespace {

// Device arrays of the caller: each must be device memory of the context's
// device; the work is ordered after `stream`'s queued work and `stream` waits
// for it before its own later work.
void require_device_array(const void* p, int device, const char* what) {
    if (!p) return;
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(RRG_ERR_INVALID, std::string(what) + ": not device memory");
    }
    if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)
        fail(RRG_ERR_INVALID, std::string(what) + ": not device memory");
    if (a.device != device) fail(RRG_ERR_INVALID, std::string(what) + ": memory of another device");
}

class StreamJoin {
public:
    StreamJoin(hipStream_t caller, hipStream_t own) : caller_(caller), own_(own) {
        RRG_HIP(hipEventCreateWithFlags(&e_, hipEventDisableTiming));
        RRG_HIP(hipEventRecord(e_, caller_));
        RRG_HIP(hipStreamWaitEvent(own_, e_, 0));
    }
    ~StreamJoin() { (void)hipEventDestroy(e_); }
    void done() {
        RRG_HIP(hipEventRecord(e_, own_));
        RRG_HIP(hipStreamWaitEvent(caller_, e_, 0));
    }

private:
    hipStream_t caller_, own_;
    hipEvent_t e_ = nullptr;
};

}  // namespace

RRG_API int rrg_get_state_dev(rrg_ctx* ctx, double* pos, double* vel, double* acc, void* stream) {
    RRG_CTX_CALL(ctx, {
        for (const void* q : {(const void*)pos, (const void*)vel, (const void*)acc})
            require_device_array(q, ctx->device, "rrg_get_state_dev");
        StreamJoin j(static_cast<hipStream_t>(stream), ctx->stream);
        ctx->solver->get_state_dev(pos, vel, acc);
        j.done();
    });
}

RRG_API int rrg_set_state_dev(rrg_ctx* ctx, const double* pos, const double* vel, void* stream) {
    RRG_CTX_CALL(ctx, {
        require_device_array(pos, ctx->device, "rrg_set_state_dev");
        require_device_array(vel, ctx->device, "rrg_set_state_dev");
        StreamJoin j(static_cast<hipStream_t>(stream), ctx->stream);
        ctx->solver->set_state_dev(pos, vel);
        j.done();
    });
}

RRG_API int rrg_adjoint_dev(rrg_ctx* ctx, const double* gx_T, const double* gv_T, double* gx_0, double* gv_0,
                            double* gm, void* stream) {
    RRG_CTX_CALL(ctx, {
        for (const void* q : {(const void*)gx_T, (const void*)gv_T, (const void*)gx_0, (const void*)gv_0, (const void*)gm})
            require_device_array(q, ctx->device, "rrg_adjoint_dev");
        StreamJoin j(static_cast<hipStream_t>(stream), ctx->stream);
        ctx->solver->adjoint_dev(gx_T, gv_T, gx_0, gv_0, gm);
        j.done();
    });
}

RRG_API int rrg_wave_schedule(int32_t real_type, int32_t nx, int32_t boundary, int32_t iterations, int32_t fused,
                              int32_t* launches) {
    return guarded([&] {
        require(launches != nullptr && nx >= 1 && iterations >= 0, "rrg_wave_schedule: bad arguments");
        require(real_type == RRG_REAL_F32 || real_type == RRG_REAL_F64, "rrg_wave_schedule: bad real_type");
        launches[0] = launches[1] = launches[2] = launches[3] = 0;
        auto levels = [&](int r) {
            if (!fused) return 1;
            return real_type == RRG_REAL_F32 ? pm::wave_fused_levels<float>(pm::make_model<float>(nx, 1, 0, 0, 0, 1, 1,
                                                                                                  boundary), r)
                                             : pm::wave_fused_levels<double>(pm::make_model<double>(nx, 1, 0, 0, 0, 1,
                                                                                                    1, boundary), r);
        };
        int ip = 0, ic = 1;
        pm::wave_schedule(
            iterations, levels, ip, ic, [&](int, int, int) { ++launches[0]; },
            [&](int lv, int, int, int, int) { ++launches[lv - 1]; });
    });
}

RRG_API int rrg_get_conservation(rrg_ctx* ctx, rrg_conservation* out) {
    RRG_CTX_CALL(ctx, {
        require(out != nullptr, "rrg_get_conservation: null output");
        ctx->solver->conservation(out);
    });
}

RRG_API int rrg_adjoint_begin(rrg_ctx* ctx, int32_t checkpoint_every) {
    RRG_CTX_CALL(ctx, ctx->solver->adjoint_begin(checkpoint_every));
}

RRG_API int rrg_adjoint(rrg_ctx* ctx, const double* gx_T, const double* gv_T, double* gx_0, double* gv_0,
                        double* gm) {
    RRG_CTX_CALL(ctx, ctx->solver->adjoint(gx_T, gv_T, gx_0, gv_0, gm));
}

RRG_API int rrg_get_conservation_systems(rrg_ctx* ctx, int32_t k, rrg_conservation* out) {
    RRG_CTX_CALL(ctx, {
        require(out != nullptr && k >= 1, "rrg_get_conservation_systems: bad arguments");
        ctx->solver->conservation_systems(k, out);
    });
}

RRG_API int rrg_get_grid(rrg_ctx* ctx, double* mass, double* phi, double* acc) {
    RRG_CTX_CALL(ctx, ctx->solver->get_grid(mass, phi, acc));
}

RRG_API int rrg_get_stats(rrg_ctx* ctx, rrg_stats* out) {
    RRG_CTX_CALL(ctx, {
        require(out != nullptr, "rrg_get_stats: null output");
        ctx->solver->stats(out);
    });
}

RRG_API int rrg_synchronize(rrg_ctx* ctx) { RRG_CTX_CALL(ctx, RRG_HIP(hipStreamSynchronize(ctx->stream))); }

RRG_API int rrg_set_timing(rrg_ctx* ctx, int32_t enabled) {
    RRG_CTX_CALL(ctx, {
        RRG_HIP(hipStreamSynchronize(ctx->stream));
        ctx->solver->timer.set(enabled != 0);
    });
}

RRG_API int rrg_get_timing(rrg_ctx* ctx, rrg_timing* out) {
    RRG_CTX_CALL(ctx, {
        require(out != nullptr, "rrg_get_timing: null output");
        ctx->solver->timer.collect(ctx->stream, out);
    });
}

// ------------------------------------------------------------------ operators
namespace {

struct OpStream {
    hipStream_t s = nullptr;
    DeviceGuard dg{0};
    OpStream() { RRG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~OpStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

size_t real_size(int32_t real_type) {
    if (real_type == RRG_REAL_F64) return sizeof(double);
    if (real_type == RRG_REAL_F32) return sizeof(float);
    fail(RRG_ERR_INVALID, "unknown real_type");
}

template <class T> void h2d(hipStream_t s, T* dst, const void* src, size_t bytes) {
    RRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}
template <class T> void d2h(hipStream_t s, void* dst, const T* src, size_t bytes) {
    RRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    RRG_HIP(hipStreamSynchronize(s));
}

void direct_op(int32_t kind, int64_t n, const double* pos, const double* mass, const double* soft2, double* out,
               bool potential) {
    require(n > 0 && pos && mass && soft2 && out, "direct-sum operator: bad arguments");
    require(valid_force_kind(kind), "direct-sum operator: unknown force_kind");
    OpStream os;
    Particles<double> P;
    P.alloc(n);
    upload_vec<double>(os.s, n, pos, P.X.p);
    upload_scalar<double>(os.s, n, mass, P.M.p);
    upload_scalar<double>(os.s, n, soft2, P.S2.p);
    nbody::Plan pl = nbody::plan_direct(n, kind);
    const int comps = potential ? 1 : 3;
    DBuf<double> part(size_t(pl.slices) * comps * n);
    DBuf<double> res(size_t(comps) * n);
    DBuf<float4> A, B;
    const nbody::Force f = make_force(kind, n, soft2);
    if (nbody::is_fp32_kind(kind)) {
        A.alloc(n);
        B.alloc(n);
        nbody::pack_mixed(os.s, P.view(), A.p, B.p);
    }
    if (potential) nbody::grav(os.s, pl, f, P.view(), A.p, B.p, part.p);
    else nbody::accel(os.s, pl, f, P.view(), A.p, B.p, part.p);
    nbody::reduce_slices(os.s, n, comps, pl.slices, nullptr, part.p, res.p);
    if (potential) {
        d2h(os.s, out, res.p, sizeof(double) * n);
    } else {
        std::vector<double> soa(3 * n);
        d2h(os.s, soa.data(), res.p, sizeof(double) * 3 * n);
        soa_to_aos<double>(n, soa, out);
    }
}

template <class real>
void cic_op(int64_t n, const void* pos, const void* mass, int32_t nx, double dx, double bc_frac, int32_t e,
            int64_t* grid) {
    OpStream os;
    const int64_t ncell = int64_t(nx) * nx * nx;
    Particles<real> P;
    P.alloc(n);
    std::vector<real> soa(3 * n);
    const real* pr = static_cast<const real*>(pos);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = pr[3 * i + c];
    h2d(os.s, P.X.p, soa.data(), sizeof(real) * 3 * n);
    h2d(os.s, P.M.p, mass, sizeof(real) * n);
    DBuf<unsigned long long> g(ncell);
    pm::fill_u64(os.s, g.p, ncell, 0ull);
    const pm::Geo<real> geo = pm::make_geo<real>(nx, dx, bc_frac);
    double mmax = 0;
    for (int64_t i = 0; i < n; ++i) mmax = std::max(mmax, std::fabs((double)static_cast<const real*>(mass)[i]));
    pm::predict_deposit<real>(os.s, P.view(), (real)0, false, geo, std::ldexp(1.0, e), g.p, mmax * std::ldexp(1.0, e));
    d2h(os.s, grid, g.p, sizeof(int64_t) * ncell);
}

template <class real>
void wave_op(int32_t nx, double dx, double diss, double c, double wdt, int32_t boundary, double rho_mean,
             int32_t iters, void* prev, void* curr, const void* mass) {
    OpStream os;
    const int64_t ncell = int64_t(nx) * nx * nx;
    DBuf<real> b[4], m(ncell), src(ncell);
    for (auto& x : b) x.alloc(ncell);
    h2d(os.s, b[0].p, prev, sizeof(real) * ncell);
    h2d(os.s, b[1].p, curr, sizeof(real) * ncell);
    h2d(os.s, m.p, mass, sizeof(real) * ncell);
    const bool periodic = boundary == RRG_BC_PERIODIC;
    const pm::WaveK<real> k = pm::wave_coef<real>((real)diss, (real)c, (real)wdt, (real)dx);
    const pm::WaveModel<real> wm = pm::make_model<real>(nx, dx, 0.0, diss, 0.0, c, wdt, boundary);
    pm::mass_to_source<real>(os.s, nx, m.p, (real)dx, (real)rho_mean, periodic, k.k_f, src.p, periodic ? nullptr : &wm);
    int ip = 0, ic = 1;
    pm::wave_schedule(
        iters, [&](int r) { return pm::wave_fused_levels<real>(wm, r); }, ip, ic,
        [&](int out, int cur, int prv) {
            pm::wave_iteration<real>(os.s, wm, k, b[out].p, b[cur].p, b[prv].p, src.p, nullptr);
        },
        [&](int lv, int o1, int o2, int cur, int prv) {
            pm::wave_fused<real>(os.s, wm, k, lv, b[o1].p, b[o2].p, b[cur].p, b[prv].p, src.p);
        });
    d2h(os.s, prev, b[ip].p, sizeof(real) * ncell);
    d2h(os.s, curr, b[ic].p, sizeof(real) * ncell);
}

template <class real> void gradient_op(int32_t nx, double dx, const void* phi, void* acc) {
    OpStream os;
    const int64_t ncell = int64_t(nx) * nx * nx;
    DBuf<real> p(ncell), a(4 * ncell);
    h2d(os.s, p.p, phi, sizeof(real) * ncell);
    pm::gradient<real>(os.s, nx, (real)dx, p.p, a.p);
    std::vector<real> tmp(4 * ncell);
    d2h(os.s, tmp.data(), a.p, sizeof(real) * 4 * ncell);
    real* out = static_cast<real*>(acc);
    for (int64_t c = 0; c < ncell; ++c)
        for (int q = 0; q < 3; ++q) out[3 * c + q] = tmp[4 * c + q];
}

template <class real>
void gather_op(int64_t n, const void* pos, int32_t nx, double dx, double bc_frac, const void* field, void* acc) {
    OpStream os;
    const int64_t ncell = int64_t(nx) * nx * nx;
    Particles<real> P;
    P.alloc(n);
    std::vector<real> soa(3 * n);
    const real* pr = static_cast<const real*>(pos);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = pr[3 * i + c];
    h2d(os.s, P.X.p, soa.data(), sizeof(real) * 3 * n);
    std::vector<real> f4(4 * ncell);
    const real* fr = static_cast<const real*>(field);
    for (int64_t c = 0; c < ncell; ++c) {
        for (int q = 0; q < 3; ++q) f4[4 * c + q] = fr[3 * c + q];
        f4[4 * c + 3] = 0;
    }
    DBuf<real> F(4 * ncell);
    h2d(os.s, F.p, f4.data(), sizeof(real) * 4 * ncell);
    const pm::Geo<real> geo = pm::make_geo<real>(nx, dx, bc_frac);
    pm::gather<real>(os.s, P.view(), geo, F.p, (real)0, false);
    d2h(os.s, soa.data(), P.A.p, sizeof(real) * 3 * n);
    real* out = static_cast<real*>(acc);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = soa[c * n + i];
}

template <class real>
void ngp_bin_op(int64_t n, const void* pos, int32_t nx, double dx, double bc_frac, int32_t* cell_of, int32_t* count,
                int32_t* start, int32_t* order) {
    OpStream os;
    const int64_t ncell = int64_t(nx) * nx * nx;
    Particles<real> P;
    P.alloc(n);
    std::vector<real> soa(3 * n);
    const real* pr = static_cast<const real*>(pos);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = pr[3 * i + c];
    if (n) h2d(os.s, P.X.p, soa.data(), sizeof(real) * 3 * n);
    DBuf<unsigned> keys(std::max<int64_t>(n, 1)), keys2(std::max<int64_t>(n, 1)), idx(std::max<int64_t>(n, 1)),
        ord(std::max<int64_t>(n, 1));
    DBuf<int> cnt(ncell), st(ncell);
    DBuf<char> temp(pm::ngp_temp_bytes(n, nx));
    pm::NgpBins b;
    b.keys = keys.p; b.keys2 = keys2.p; b.idx = idx.p; b.order = ord.p;
    b.count = cnt.p; b.start = st.p; b.temp = temp.p; b.temp_bytes = temp.n;
    const pm::Geo<real> geo = pm::make_geo<real>(nx, dx, bc_frac);
    pm::ngp_bin<real>(os.s, P.view(), geo, b);
    if (cell_of && n) d2h(os.s, cell_of, keys.p, sizeof(int32_t) * n);
    if (count) d2h(os.s, count, cnt.p, sizeof(int32_t) * ncell);
    if (start) d2h(os.s, start, st.p, sizeof(int32_t) * ncell);
    if (order && n) d2h(os.s, order, ord.p, sizeof(int32_t) * n);
}

template <class real> void steady_op(int64_t ncell, const void* un, const void* uo, double* err) {
    using MB = std::conditional_t<std::is_same_v<real, float>, unsigned int, unsigned long long>;
    OpStream os;
    DBuf<real> a(ncell), b(ncell);
    DBuf<unsigned long long> mb(2 * pm::kCheckSlots);
    h2d(os.s, a.p, un, sizeof(real) * ncell);
    h2d(os.s, b.p, uo, sizeof(real) * ncell);
    pm::fill_u64(os.s, mb.p, 2 * pm::kCheckSlots, 0ull);
    pm::max_diff<real>(os.s, ncell, a.p, b.p, mb.p);
    std::vector<MB> slots(2 * pm::kCheckSlots);
    d2h(os.s, slots.data(), reinterpret_cast<MB*>(mb.p), sizeof(MB) * slots.size());
    MB bits[2];
    max_over_slots<MB>(slots.data(), bits);
    real diff, absu;
    std::memcpy(&diff, &bits[0], sizeof(real));
    std::memcpy(&absu, &bits[1], sizeof(real));
    const real EPS = (real)1.E-30;
    const real e = diff / (absu + EPS);
    *err = (double)e;
}

}  // namespace

RRG_API int rrg_op_direct_accel(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                                const double* soft2, double* acc) {
    return guarded([&] { direct_op(force_kind, n, pos, mass, soft2, acc, false); });
}

RRG_API int rrg_op_direct_grav(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                               const double* soft2, double* grav) {
    return guarded([&] { direct_op(force_kind, n, pos, mass, soft2, grav, true); });
}

RRG_API int rrg_op_cic_deposit(int32_t real_type, int64_t n, const void* pos, const void* mass, int32_t nx, double dx,
                               double bc_frac, int32_t e, int64_t* grid) {
    return guarded([&] {
        require(n > 0 && pos && mass && grid && nx >= 2 && dx > 0 && e >= 0 && e < 1000, "rrg_op_cic_deposit: bad arguments");
        real_size(real_type);
        if (real_type == RRG_REAL_F64) cic_op<double>(n, pos, mass, nx, dx, bc_frac, e, grid);
        else cic_op<float>(n, pos, mass, nx, dx, bc_frac, e, grid);
    });
}

RRG_API int rrg_op_wave_iterate(int32_t real_type, int32_t nx, double dx, double diss_base, double wave_speed,
                                double wave_dt, int32_t boundary, double rho_mean, int32_t iterations, void* phi_prev,
                                void* phi_curr, const void* mass) {
    return guarded([&] {
        require(nx >= 2 && dx > 0 && iterations >= 0 && phi_prev && phi_curr && mass, "rrg_op_wave_iterate: bad arguments");
        require(boundary == RRG_BC_PERIODIC || boundary == RRG_BC_ZERO_DIRICHLET, "rrg_op_wave_iterate: unknown boundary");
        real_size(real_type);
        if (real_type == RRG_REAL_F64)
            wave_op<double>(nx, dx, diss_base, wave_speed, wave_dt, boundary, rho_mean, iterations, phi_prev, phi_curr, mass);
        else
            wave_op<float>(nx, dx, diss_base, wave_speed, wave_dt, boundary, rho_mean, iterations, phi_prev, phi_curr, mass);
    });
}

RRG_API int rrg_op_gradient(int32_t real_type, int32_t nx, double dx, const void* phi, void* acc) {
    return guarded([&] {
        require(nx >= 2 && dx > 0 && phi && acc, "rrg_op_gradient: bad arguments");
        real_size(real_type);
        if (real_type == RRG_REAL_F64) gradient_op<double>(nx, dx, phi, acc);
        else gradient_op<float>(nx, dx, phi, acc);
    });
}

RRG_API int rrg_op_gather(int32_t real_type, int64_t n, const void* pos, int32_t nx, double dx, double bc_frac,
                          const void* field, void* acc) {
    return guarded([&] {
        require(n > 0 && pos && field && acc && nx >= 2 && dx > 0, "rrg_op_gather: bad arguments");
        real_size(real_type);
        if (real_type == RRG_REAL_F64) gather_op<double>(n, pos, nx, dx, bc_frac, field, acc);
        else gather_op<float>(n, pos, nx, dx, bc_frac, field, acc);
    });
}

RRG_API int rrg_op_ngp_bin(int32_t real_type, int64_t n, const void* pos, int32_t nx, double dx, double bc_frac,
                           int32_t* cell_of, int32_t* count, int32_t* start, int32_t* order) {
    return guarded([&] {
        require(n >= 0 && n <= (int64_t(1) << 31) - 1 && (n == 0 || pos) && nx >= 1 && nx <= 1290 && dx > 0,
                "rrg_op_ngp_bin: bad arguments");
        real_size(real_type);
        if (real_type == RRG_REAL_F64) ngp_bin_op<double>(n, pos, nx, dx, bc_frac, cell_of, count, start, order);
        else ngp_bin_op<float>(n, pos, nx, dx, bc_frac, cell_of, count, start, order);
    });
}

RRG_API int rrg_op_steady_err(int32_t real_type, int64_t ncell, const void* u_new, const void* u_old, double* err) {
    return guarded([&] {
        require(ncell > 0 && u_new && u_old && err, "rrg_op_steady_err: bad arguments");
        real_size(real_type);
        if (real_type == RRG_REAL_F64) steady_op<double>(ncell, u_new, u_old, err);
        else steady_op<float>(ncell, u_new, u_old, err);
    });
}

// ------------------------------------------------ allocation-free operator
// rrg_op_direct_accel for callers that cannot allocate device memory (code
// under stream capture, serving loops): device arrays in and out, the
// caller's stream, and a fixed scratch area that lives in a file-scope device
// array for the life of the process.  The array's name in host code is a
// host-side stand-in; its device address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that address is handed to kernels.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // per device
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on device `dev` (the current device).
unsigned char* scratch_address(int dev) {
    require(dev >= 0 && dev < kMaxScratchDevices, "rrg_op_direct_accel_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        RRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "rrg_op_direct_accel_noalloc: no device address for the scratch area");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// The evaluation of direct_op with the same launch plan and kernels, so the
// result is bitwise the same.  Softening is always taken per pair:
// h_i + h_j with h = 0.5 * soft2 in the kernel's precision rounds exactly as
// the uniform-softening constant h + h does, so no host read of soft2 is
// needed.  Every slice partial is written before the reduction reads it, so
// nothing in the scratch area has to start at zero.
void direct_noalloc(int32_t kind, int64_t n, const double* pos, const double* mass, const double* soft2, double* acc,
                    hipStream_t s) {
    require(n > 0 && pos && mass && soft2 && acc, "rrg_op_direct_accel_noalloc: bad arguments");
    require(valid_force_kind(kind), "rrg_op_direct_accel_noalloc: unknown force_kind");
    int dev = 0;
    RRG_HIP(hipGetDevice(&dev));
    require_device_array(pos, dev, "rrg_op_direct_accel_noalloc: pos");
    require_device_array(mass, dev, "rrg_op_direct_accel_noalloc: mass");
    require_device_array(soft2, dev, "rrg_op_direct_accel_noalloc: soft2");
    require_device_array(acc, dev, "rrg_op_direct_accel_noalloc: acc");
    const nbody::Plan pl = nbody::plan_direct(n, kind);
    const bool f32 = nbody::is_fp32_kind(kind);
    const size_t v3 = round256(sizeof(double) * 3 * size_t(n));
    const size_t rec = f32 ? round256(sizeof(float4) * size_t(n)) : 0;
    const size_t part = round256(sizeof(double) * 3 * size_t(n) * size_t(pl.slices));
    const size_t need = 2 * v3 + 2 * rec + part;
    if (need > kNoAllocScratchBytes)
        fail(RRG_ERR_INVALID, "rrg_op_direct_accel_noalloc: n = " + std::to_string(n) + " needs " +
                                  std::to_string(need) + " bytes of scratch, more than the " +
                                  std::to_string(kNoAllocScratchBytes) + " reserved per device");
    unsigned char* base = scratch_address(dev);
    double* X = reinterpret_cast<double*>(base);
    double* R = reinterpret_cast<double*>(base + v3);
    double* P = reinterpret_cast<double*>(base + 2 * v3);
    float4* A = f32 ? reinterpret_cast<float4*>(base + 2 * v3 + part) : nullptr;
    float4* B = f32 ? reinterpret_cast<float4*>(base + 2 * v3 + part + rec) : nullptr;
    const PState<double> p{X, nullptr, nullptr, nullptr, const_cast<double*>(mass), const_cast<double*>(soft2),
                           nullptr, n};
    nbody::Force f;
    f.kind = kind;
    f.uniform = false;
    layout::rows_in<double>(s, n, pos, X, nullptr);
    if (f32) nbody::pack_mixed(s, p, A, B);
    nbody::accel(s, pl, f, p, A, B, P);
    nbody::reduce_slices(s, n, 3, pl.slices, nullptr, P, R);
    layout::rows_out<double>(s, n, R, acc, nullptr);
}

}  // namespace

RRG_API int rrg_op_direct_accel_noalloc(int32_t force_kind, int64_t n, const double* pos, const double* mass,
                                        const double* soft2, double* acc, void* stream) {
    return guarded([&] { direct_noalloc(force_kind, n, pos, mass, soft2, acc, static_cast<hipStream_t>(stream)); });
}
