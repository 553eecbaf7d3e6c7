// rrgsim <work_dir> -- the params.json-driven solver executable, written over
// the librrg.so C ABI (include/rrg.h).
//
// Reference: solver/app/src/rrgsim.cu:12-62 (main), rrgsim_integration.cu:19-145
// (integrate: precalc, sync, save cadence), rrgsim_conservation.cu:9-98 and
// io_handler.cpp:71-105 (data/delta_*.csv tables), rrgsim_conservation.cu:280-424
// (check_bounds, data/sample_particle.csv and the grid axis cuts
// data/{a_grid,grav_grid,mass_grid}/), io_particles_bin.cu:9-37
// (dump/S_{iter:>5d}.bin), rrgsim_log.cpp:27-74 (logs/rrgsim_<time>.log).
//
// Deliberate differences from the reference, all on error paths or quirks:
//   - a failure exits with status 2 (the reference logs it and exits 0);
//   - the direct-sum model is selected when either grid_params or
//     wave_params is absent (the reference throws on the missing optional);
//   - N need not be a multiple of 256;
//   - the particle potential used for Ep is recomputed at every save.
#include <cmath>
#include <algorithm>
#include <charconv>
#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "rrg.h"

namespace fs = std::filesystem;

namespace {

std::FILE* g_log = nullptr;

void log_line(const char* level, const char* fmt, ...) {
    char msg[2048];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    std::time_t now = std::time(nullptr);
    std::tm tm_now{};
    localtime_r(&now, &tm_now);
    char ts[32];
    std::strftime(ts, sizeof(ts), "%H:%M:%S", &tm_now);
    std::printf("[%s] [%s] %s\n", ts, level, msg);
    if (g_log) {
        char tsl[40];
        std::strftime(tsl, sizeof(tsl), "%Y-%m-%d %H:%M:%S", &tm_now);
        std::fprintf(g_log, "[%s] [%s] %s\n", tsl, level, msg);
        std::fflush(g_log);
    }
}

struct Fail {
    std::string what;
};

void check(int rc, const char* call) {
    if (rc != RRG_OK) throw Fail{std::string(call) + " failed (" + std::to_string(rc) + "): " + rrg_last_error()};
}

// IOHandler::append_table: header on first use, rows "{:.{prec}f}, "; a
// table may live in a subdirectory of data/.
class Tables {
public:
    explicit Tables(fs::path dir) : dir_(std::move(dir)) {}
    struct Item {
        const char* tag;
        double value;
        int precision;
    };
    void append(const std::string& name, std::initializer_list<Item> items, const char* subdir = nullptr) {
        const fs::path dir = subdir ? dir_ / subdir : dir_;
        fs::create_directories(dir);
        const fs::path path = dir / (name + ".csv");
        std::ofstream f;
        if (filled_[name]) {
            f.open(path, std::ios::app);
        } else {
            f.open(path);
            size_t k = 0;
            for (const Item& it : items) f << it.tag << (++k < items.size() ? ", " : "\n");
            filled_[name] = true;
        }
        size_t k = 0;
        char buf[128];
        for (const Item& it : items) {
            std::snprintf(buf, sizeof(buf), "%.*f", it.precision, it.value);
            f << buf << (++k < items.size() ? ", " : "\n");
        }
    }

private:
    fs::path dir_;
    std::map<std::string, bool> filled_;
};

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

void print_conservation(const rrg_conservation& c) {
    log_line("info", "------------------------------------------");
    log_line("info", "Conservation time: %.10g", c.time);
    log_line("info", "-- momentum: %.15f", norm3(c.momentum));
    log_line("info", "-- angular momentum: %.15f", norm3(c.angular));
    log_line("info", "-- energy: %.15f (kinetic %.15f, potential %.15f)", c.Ek + c.Ep, c.Ek, c.Ep);
}

// rrgsim_conservation.cu:9-98 and :213-278.
void save_conservation(Tables& t, const rrg_conservation& c0, const rrg_conservation& c) {
    const double ang_z_rel = c.angular[2] / c0.angular[2] - 1.;
    t.append("delta_ang_z", {{"t", c.time, 6}, {"ang_z", c.angular[2], 20}, {"ang_z_rel_signed", ang_z_rel, 20},
                             {"ang_z_rel", std::fabs(ang_z_rel), 20}});
    const double an = norm3(c.angular), an0 = norm3(c0.angular);
    const double an_rel = an / an0 - 1.;
    t.append("delta_ang_norm", {{"t", c.time, 6}, {"ang_norm", an, 20}, {"ang_norm_rel_signed", an_rel, 20},
                                {"ang_norm_rel", std::fabs(an_rel), 20}});
    const double pn = norm3(c.momentum), pn0 = norm3(c0.momentum);
    const double p_rel = pn / pn0 - 1.;
    t.append("delta_momentum", {{"t", c.time, 20}, {"momentum_norm", pn, 20}, {"momentum_diff", pn - pn0, 20},
                                {"momentum_rel_signed", p_rel, 20}, {"momentum_rel", std::fabs(p_rel), 20}});
    const double E = c.Ek + c.Ep, E0 = c0.Ek + c0.Ep;
    const double e_rel = E / E0 - 1.;
    t.append("delta_energy", {{"t", c.time, 6}, {"E", E, 15}, {"E_rel_signed", e_rel, 15}, {"E_rel", std::fabs(e_rel), 15},
                              {"Ek", c.Ek, 15}, {"Ep", c.Ep, 15}});
    log_line("info", "-- energy total: diff %.6e ; rel %.6e", E - E0, e_rel);
    if (std::fabs(e_rel) > 0.15) log_line("warning", "-- energy total drifted by more than 15%%");
}

// Shortest round-trip text of a value of the run's precision (what fmt's
// "{}" prints for a float or a double).
template <class real> std::string shortest(real v) {
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v);
    return std::string(buf, r.ptr);
}

// The reference's GridInfo for a grid section (io_json.cpp:28-40), in `real`.
template <class real> struct GridGeom {
    int nx;
    real dx, domain_min, domain_max;
};
template <class real> GridGeom<real> grid_geom(const rrg_grid_params& g) {
    GridGeom<real> o;
    o.nx = g.nx;
    o.dx = (real)g.dx;
    const real domain_l = o.dx * (real)g.nx;
    o.domain_max = (real)(0.5 * domain_l);
    o.domain_min = (real)(-0.5 * domain_l);
    return o;
}

// print_sample_particle (rrgsim_conservation.cu:334-359): particle 0.
void save_sample_particle(Tables& t, const double* x, const double* v, const double* a, double grav) {
    t.append("sample_particle", {{"i", 0.0, 8}, {"x", x[0], 8}, {"y", x[1], 8}, {"z", x[2], 8}, {"vx", v[0], 8},
                                 {"vy", v[1], 8}, {"vz", v[2], 8}, {"grav", grav, 8}, {"ax", a[0], 8},
                                 {"ay", a[1], 8}, {"az", a[2], 8}});
}

// check_bounds (rrgsim_conservation.cu:280-332): the particles' extent, and an
// error line per axis and side that leaves the domain.
template <class real> void check_bounds(int64_t n, const std::vector<double>& pos, const GridGeom<real>& g) {
    log_line("debug", "Check bounds");
    if (n == 0) {
        log_line("debug", "\t No particles_pos passed");
        return;
    }
    real lo[3], hi[3];
    for (int c = 0; c < 3; ++c) lo[c] = hi[c] = (real)pos[c];
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const real v = (real)pos[3 * i + c];
            hi[c] = std::max(hi[c], v);
            lo[c] = std::min(lo[c], v);
        }
    const char* axis[3] = {"X", "Y", "Z"};
    for (int c = 0; c < 3; ++c)
        log_line("debug", "\t %s: [%s .. %s]", axis[c], shortest(lo[c]).c_str(), shortest(hi[c]).c_str());
    for (int c = 0; c < 3; ++c)
        if (lo[c] < g.domain_min)
            log_line("error", "%s min bounds violation! Particle %s: %s; domain_min: %s", axis[c], axis[c],
                     shortest(lo[c]).c_str(), shortest(g.domain_min).c_str());
    for (int c = 0; c < 3; ++c)
        if (hi[c] > g.domain_max)
            log_line("error", "%s max bounds violation! Particle %s: %s; domain_max: %s", axis[c], axis[c],
                     shortest(hi[c]).c_str(), shortest(g.domain_max).c_str());
}

// print_acc_projection / print_grav_projection / print_mass_projection
// (rrgsim_conservation.cu:360-424): cuts through the grid centre, x-fastest
// cell index AT(x, y, z); the coordinate column is domain_min + dx * i.
template <class real>
void save_grid_cuts(Tables& t, const GridGeom<real>& g, real time, const std::vector<double>& mass,
                    const std::vector<double>& phi, const std::vector<double>& acc) {
    const size_t NX = size_t(g.nx), h = NX / 2;
    auto at = [NX](size_t x, size_t y, size_t z) { return x + NX * (y + NX * z); };
    char name[96];
    std::snprintf(name, sizeof(name), "a_grid_center_%.8f", (double)time);
    for (size_t i = 0; i < NX; ++i) {
        const double x = (double)(real)(g.domain_min + g.dx * (real)i);
        const size_t c = at(i, h, h);
        t.append(name, {{"x", x, 8}, {"ax", acc[3 * c], 8}, {"ay", acc[3 * c + 1], 8}, {"az", acc[3 * c + 2], 8}},
                 "a_grid");
    }
    std::snprintf(name, sizeof(name), "grav_axis_center_%.10f", (double)time);
    for (size_t i = 0; i < NX; ++i) {
        const double r = (double)(real)(g.domain_min + g.dx * (real)i);
        t.append(name, {{"r", r, 8}, {"gx", phi[at(i, h, h)], 8}, {"gy", phi[at(h, i, h)], 8},
                        {"gz", phi[at(h, h, i)], 8}}, "grav_grid");
    }
    std::snprintf(name, sizeof(name), "mass_axis_center_%.10f", (double)time);
    for (size_t i = 0; i < NX; ++i) {
        const double r = (double)(real)(g.domain_min + g.dx * (real)i);
        t.append(name, {{"r", r, 8}, {"mx", mass[at(i, h, h)], 8}, {"my", mass[at(h, i, h)], 8},
                        {"mz", mass[at(h, h, i)], 8}}, "mass_grid");
    }
}

// write_particles_dump (io_particles_bin.cu:9-37): size_t N, double t,
// real3 pos[N], real3 vel[N] in the run's precision.
void write_dump(const fs::path& dir, int64_t iter, double time, int64_t n, const std::vector<double>& pos,
                const std::vector<double>& vel, bool fp32) {
    fs::create_directories(dir);
    char name[64];
    std::snprintf(name, sizeof(name), "S_%5lld.bin", (long long)iter);
    std::ofstream f(dir / name, std::ios::binary);
    const size_t N = size_t(n);
    f.write(reinterpret_cast<const char*>(&N), sizeof(N));
    f.write(reinterpret_cast<const char*>(&time), sizeof(time));
    if (fp32) {
        std::vector<float> tmp(pos.begin(), pos.end());
        f.write(reinterpret_cast<const char*>(tmp.data()), sizeof(float) * tmp.size());
        tmp.assign(vel.begin(), vel.end());
        f.write(reinterpret_cast<const char*>(tmp.data()), sizeof(float) * tmp.size());
    } else {
        f.write(reinterpret_cast<const char*>(pos.data()), sizeof(double) * pos.size());
        f.write(reinterpret_cast<const char*>(vel.data()), sizeof(double) * vel.size());
    }
    log_line("info", "Write particles dump. Time: %.5f", time);
}

template <class real> int64_t run_loop(rrg_ctx* ctx, const rrg_params& p, int64_t n, const fs::path& work_dir) {
    Tables tables(work_dir / "data");
    check(rrg_precalc(ctx), "rrg_precalc");
    check(rrg_sync(ctx), "rrg_sync");
    rrg_conservation c0;
    check(rrg_get_conservation(ctx, &c0), "rrg_get_conservation");
    print_conservation(c0);

    // integrate (rrgsim_integration.cu:79-138); time accumulates in `real`.
    int64_t iteration = 0;
    real time = 0.;
    real next_save = time + (real)p.sim.dt_save;
    uint64_t saves_without_dump = 0;
    std::vector<double> pos(3 * n), vel(3 * n), acc(3 * n), grav(n);
    const GridGeom<real> geom = grid_geom<real>(p.grid);
    const size_t ncell = p.has_grid ? size_t(p.grid.nx) * p.grid.nx * p.grid.nx : 0;
    std::vector<double> gmass(ncell), gphi(ncell), gacc(3 * ncell);
    const auto t_start = std::chrono::steady_clock::now();
    while (time < (real)p.sim.time_max) {
        check(rrg_step(ctx, 1), "rrg_step");
        ++iteration;
        time += (real)p.sim.dt_dynamics;
        if (time >= next_save) {
            log_line("info", "Simulation time: %.5f / %.5f (iter: %lld)", (double)time, p.sim.time_max,
                     (long long)iteration);
            check(rrg_sync(ctx), "rrg_sync");
            check(rrg_get_state(ctx, pos.data(), vel.data(), acc.data(), grav.data()), "rrg_get_state");
            if (n > 0) save_sample_particle(tables, pos.data(), vel.data(), acc.data(), grav[0]);
            rrg_conservation c;
            check(rrg_get_conservation(ctx, &c), "rrg_get_conservation");
            c.time = (double)time;
            save_conservation(tables, c0, c);
            if (p.has_grid) {
                check(rrg_get_grid(ctx, gmass.data(), gphi.data(), gacc.data()), "rrg_get_grid");
                check_bounds<real>(n, pos, geom);
                save_grid_cuts<real>(tables, geom, time, gmass, gphi, gacc);
            }
            next_save += (real)p.sim.dt_save;
            ++saves_without_dump;
            if (saves_without_dump < p.sim.saves_without_dump) {
                write_dump(work_dir / "dump", iteration, (double)time, n, pos, vel, p.real_type == RRG_REAL_F32);
            }
        }
    }
    check(rrg_synchronize(ctx), "rrg_synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    log_line("info", "integration completed: %lld steps in %.3f s", (long long)iteration, secs);
    return iteration;
}

}  // namespace

int main(int argc, const char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: rrgsim <work_dir>\n");
        return 1;
    }
    const fs::path work_dir = argv[1];
    {
        std::error_code ec;
        fs::create_directories(work_dir / "logs", ec);
        std::time_t now = std::time(nullptr);
        std::tm tm_now{};
        localtime_r(&now, &tm_now);
        char buf[80];
        std::strftime(buf, sizeof(buf), "%Y-%m-%d_%H-%M-%S", &tm_now);
        const fs::path log_path = work_dir / "logs" / (std::string("rrgsim_") + buf + ".log");
        g_log = std::fopen(log_path.c_str(), "w");
        std::printf("Ready to log: %s\n", log_path.c_str());
        fs::remove_all(work_dir / "data", ec);  // IOHandler::setup (io_handler.cpp:12-15)
    }
    rrg_ctx* ctx = nullptr;
    int status = 0;
    try {
        log_line("info", "%s; %d HIP device(s)", rrg_version(), rrg_device_count());
        rrg_params p;
        rrg_ini_params ini;
        check(rrg_params_parse_json((work_dir / "params.json").c_str(), &p, &ini), "rrg_params_parse_json");
        log_line("info", "Params json parsed successfully (%s model)", p.use_wave_model ? "wave" : "direct-sum");
        int64_t n = 0;
        check(rrg_ini_read(&ini, nullptr, &n, nullptr, nullptr, nullptr, nullptr), "rrg_ini_read");
        std::vector<double> pos(3 * n), vel(3 * n), mass(n), soft2(n);
        check(rrg_ini_read(&ini, nullptr, &n, pos.data(), vel.data(), mass.data(), soft2.data()), "rrg_ini_read");
        log_line("info", "%lld particles loaded", (long long)n);
        check(rrg_create(&p, n, pos.data(), vel.data(), mass.data(), soft2.data(), &ctx), "rrg_create");
        if (p.real_type == RRG_REAL_F32) run_loop<float>(ctx, p, n, work_dir);
        else run_loop<double>(ctx, p, n, work_dir);
    } catch (const Fail& f) {
        log_line("error", "Error: %s", f.what.c_str());
        status = 2;
    } catch (const std::exception& e) {
        log_line("error", "Exception: %s", e.what());
        status = 2;
    }
    if (ctx) rrg_destroy(ctx);
    if (g_log) std::fclose(g_log);
    return status;
}
