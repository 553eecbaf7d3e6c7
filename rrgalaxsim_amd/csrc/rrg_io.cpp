// params.json and ini-file readers (host only).
//
// Reference: solver/lib/rrgsim_io/src/io_json.cpp (schema, defaults, derived
// fields) and solver/lib/rrgsim_io/src/io_particles_csv.cpp (ini text files).
// The reference uses nlohmann::json and csv-parser; neither is available
// here, so this file carries a small JSON reader of its own (objects, arrays,
// strings, numbers, true/false/null).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "rrg_io.hpp"

namespace rrg {
namespace io {

// ------------------------------------------------------------------- JSON
struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::map<std::string, JValue> obj;

    const JValue* find(const std::string& key) const {
        if (kind != Object) return nullptr;
        auto it = obj.find(key);
        return it == obj.end() ? nullptr : &it->second;
    }
};

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;
    explicit Parser(const std::string& text) : s(text) {}

    [[noreturn]] void error(const std::string& what) {
        fail(RRG_ERR_IO, "JSON parse error at offset " + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { ++i; return true; }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) error(std::string("expected '") + c + "'");
    }
    JValue value() {
        ws();
        if (i >= s.size()) error("unexpected end of input");
        const char c = s[i];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') { JValue v; v.kind = JValue::String; v.str = string(); return v; }
        if (s.compare(i, 4, "true") == 0) { i += 4; JValue v; v.kind = JValue::Bool; v.b = true; return v; }
        if (s.compare(i, 5, "false") == 0) { i += 5; JValue v; v.kind = JValue::Bool; v.b = false; return v; }
        if (s.compare(i, 4, "null") == 0) { i += 4; return JValue{}; }
        return number();
    }
    JValue object() {
        expect('{');
        JValue v;
        v.kind = JValue::Object;
        if (eat('}')) return v;
        do {
            ws();
            const std::string key = string();
            expect(':');
            v.obj[key] = value();
        } while (eat(','));
        expect('}');
        return v;
    }
    JValue array() {
        expect('[');
        JValue v;
        v.kind = JValue::Array;
        if (eat(']')) return v;
        do { v.arr.push_back(value()); } while (eat(','));
        expect(']');
        return v;
    }
    std::string string() {
        ws();
        if (i >= s.size() || s[i] != '"') error("expected string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) error("bad escape");
                const char e = s[i++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (i + 4 > s.size()) error("bad \\u escape");
                        const unsigned cp = unsigned(std::strtoul(s.substr(i, 4).c_str(), nullptr, 16));
                        i += 4;
                        if (cp < 0x80) out += char(cp);
                        else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                        else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) error("unterminated string");
        ++i;
        return out;
    }
    JValue number() {
        const char* begin = s.c_str() + i;
        char* end = nullptr;
        const double d = std::strtod(begin, &end);
        if (end == begin) error("unexpected character");
        i += size_t(end - begin);
        JValue v;
        v.kind = JValue::Number;
        v.num = d;
        return v;
    }
};

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(RRG_ERR_IO, "Expected params json: '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// nlohmann-style `j.at(key).get<T>()`: throws on absence or type mismatch.
const JValue& at(const JValue& j, const char* section, const char* key) {
    const JValue* v = j.find(key);
    if (!v) fail(RRG_ERR_IO, std::string("Unexpected error on params json parsing: key '") + key +
                                 "' not found in '" + section + "'");
    return *v;
}
double num(const JValue& v, const char* key) {
    if (v.kind != JValue::Number) fail(RRG_ERR_IO, std::string("params json: '") + key + "' must be a number");
    return v.num;
}
double req_num(const JValue& j, const char* section, const char* key) { return num(at(j, section, key), key); }
template <class T> void opt_num(const JValue& j, const char* key, T& out) {
    if (const JValue* v = j.find(key)) out = T(num(*v, key));
}

void set_path(char* dst, size_t cap, const JValue* v, const char* key) {
    dst[0] = '\0';
    if (!v || v->kind == JValue::Null) return;
    if (v->kind != JValue::String) fail(RRG_ERR_IO, std::string("params json: '") + key + "' must be a string or null");
    if (v->str.size() + 1 > cap) fail(RRG_ERR_IO, std::string("params json: '") + key + "' is too long");
    std::memcpy(dst, v->str.c_str(), v->str.size() + 1);
}

}  // namespace

JValue parse_json_text(const std::string& text) {
    Parser p(text);
    JValue v = p.value();
    p.ws();
    if (p.i != text.size()) p.error("trailing characters");
    return v;
}

void params_default(rrg_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->sim.time_max = NAN;
    p->sim.dt_save = NAN;
    p->sim.dt_dynamics = NAN;
    p->sim.saves_without_dump = UINT64_MAX;  // co_sim_params.h:11
    p->has_grid = 0;
    p->has_wave = 0;
    // wave defaults (io_json.cpp:58-98)
    p->wave.setup_iterations = 30000;
    p->wave.iterations = 3000;
    p->wave.tolerance = 1e-6;
    p->wave.print_step = 500;
    p->wave.steady_check_step = 30;
    p->wave.steady_hold_steps = 3;
    p->real_type = RRG_REAL_F64;
    p->force_kind = RRG_FORCE_F64;
    p->boundary = RRG_BC_PERIODIC;
    p->fixed_exponent = 0;
    p->device = 0;
    p->use_graphs = 1;
    p->autotune = 0;
    p->wave_fused = 1;
    p->sort_interval = 0;
}

// validate_dependent_fields (io_json.cpp:170-186).  The CFL step is formed
// in the configured `real`: dt = 0.5 * dx / (c * sqrt(3)).
void params_finalize(rrg_params* p) {
    p->use_wave_model = (p->has_grid && p->has_wave) ? 1 : 0;
    if (p->use_wave_model) {
        if (p->real_type == RRG_REAL_F32) {
            const float dx = float(p->grid.dx), c = float(p->wave.wave_speed);
            const float DIM = 3;
            p->wave.dt = double(float(0.5 * dx / (c * std::sqrt(DIM))));
        } else {
            const double dx = p->grid.dx, c = p->wave.wave_speed;
            const double DIM = 3;
            p->wave.dt = 0.5 * dx / (c * std::sqrt(DIM));
        }
    }
}

void params_parse_json(const std::string& path, rrg_params* p, rrg_ini_params* ini) {
    const std::string text = read_file(path);
    JValue j;
    try {
        j = parse_json_text(text);
    } catch (const Error& e) {
        fail(RRG_ERR_IO, std::string("Unexpected error on params json parsing: ") + e.what());
    }
    if (j.kind != JValue::Object) fail(RRG_ERR_IO, "params json: top level must be an object");
    params_default(p);

    // grid_params (optional; io_json.cpp:28-40)
    if (const JValue* g = j.find("grid_params")) {
        p->has_grid = 1;
        p->grid.nx = int32_t(req_num(*g, "grid_params", "nx"));
        p->grid.dx = req_num(*g, "grid_params", "dx");
        p->grid.bc_frac = req_num(*g, "grid_params", "bc_frac");
    }
    // sim_params (mandatory; io_json.cpp:13-21)
    const JValue* s = j.find("sim_params");
    if (!s) fail(RRG_ERR_IO, "No sim_params section");
    p->sim.time_max = req_num(*s, "sim_params", "time_max");
    p->sim.dt_save = req_num(*s, "sim_params", "dt_save");
    p->sim.dt_dynamics = req_num(*s, "sim_params", "dt_dynamics");
    if (const JValue* v = s->find("saves_without_dump")) p->sim.saves_without_dump = uint64_t(num(*v, "saves_without_dump"));
    // ini_params (mandatory; io_json.cpp:115-131)
    const JValue* in = j.find("ini_params");
    if (!in) fail(RRG_ERR_IO, "No ini_params section");
    rrg_ini_params tmp;
    rrg_ini_params* ip = ini ? ini : &tmp;
    std::memset(ip, 0, sizeof(*ip));
    ip->mass_star = req_num(*in, "ini_params", "mass_star");
    ip->mass_dark = req_num(*in, "ini_params", "mass_dark");
    ip->soft_star = req_num(*in, "ini_params", "soft_star");
    ip->soft_dark = req_num(*in, "ini_params", "soft_dark");
    set_path(ip->ini_file_star, sizeof(ip->ini_file_star), in->find("ini_file_star"), "ini_file_star");
    set_path(ip->ini_file_dark, sizeof(ip->ini_file_dark), in->find("ini_file_dark"), "ini_file_dark");
    // wave_params (optional; io_json.cpp:53-99)
    if (const JValue* w = j.find("wave_params")) {
        p->has_wave = 1;
        p->wave.diss_base = req_num(*w, "wave_params", "diss_base");
        p->wave.diss_extra = req_num(*w, "wave_params", "diss_extra");
        p->wave.wave_speed = req_num(*w, "wave_params", "wave_speed");
        opt_num(*w, "setup_iterations", p->wave.setup_iterations);
        opt_num(*w, "iterations", p->wave.iterations);
        opt_num(*w, "tolerance", p->wave.tolerance);
        opt_num(*w, "print_step", p->wave.print_step);
        opt_num(*w, "steady_check_step", p->wave.steady_check_step);
        opt_num(*w, "steady_hold_steps", p->wave.steady_hold_steps);
    }
    // Build options: optional "mi355x" section, unknown to the reference reader.
    if (const JValue* m = j.find("mi355x")) {
        if (const JValue* v = m->find("real")) {
            if (v->kind != JValue::String || (v->str != "fp64" && v->str != "fp32"))
                fail(RRG_ERR_IO, "mi355x.real must be \"fp64\" or \"fp32\"");
            p->real_type = v->str == "fp32" ? RRG_REAL_F32 : RRG_REAL_F64;
        }
        if (const JValue* v = m->find("force")) {
            const std::map<std::string, int> kinds = {{"fp64", RRG_FORCE_F64}, {"mixed", RRG_FORCE_MIXED},
                                                      {"fp32", RRG_FORCE_FP32}, {"fp64_exact", RRG_FORCE_F64_EXACT}};
            if (v->kind != JValue::String || !kinds.count(v->str))
                fail(RRG_ERR_IO, "mi355x.force must be \"fp64\", \"mixed\", \"fp32\" or \"fp64_exact\"");
            p->force_kind = kinds.at(v->str);
        }
        if (const JValue* v = m->find("boundary")) {
            static const std::map<std::string, int32_t> types = {
                {"periodic", RRG_BC_PERIODIC}, {"zero_dirichlet", RRG_BC_ZERO_DIRICHLET},
                {"monopole_dirichlet", RRG_BC_MONOPOLE_DIRICHLET}, {"monopole_radiative", RRG_BC_MONOPOLE_RADIATIVE}};
            if (v->kind != JValue::String || !types.count(v->str))
                fail(RRG_ERR_IO, "mi355x.boundary must be \"periodic\", \"zero_dirichlet\", \"monopole_dirichlet\" "
                                 "or \"monopole_radiative\"");
            p->boundary = types.at(v->str);
        }
        opt_num(*m, "fixed_exponent", p->fixed_exponent);
        opt_num(*m, "device", p->device);
        if (const JValue* v = m->find("graphs")) {
            if (v->kind != JValue::Bool) fail(RRG_ERR_IO, "mi355x.graphs must be true or false");
            p->use_graphs = v->b ? 1 : 0;
        }
        if (const JValue* v = m->find("autotune")) {
            if (v->kind != JValue::Bool) fail(RRG_ERR_IO, "mi355x.autotune must be true or false");
            p->autotune = v->b ? 1 : 0;
        }
        if (const JValue* v = m->find("wave_fused")) {
            if (v->kind != JValue::Bool) fail(RRG_ERR_IO, "mi355x.wave_fused must be true or false");
            p->wave_fused = v->b ? 1 : 0;
        }
        if (const JValue* v = m->find("sort_interval")) {
            if (v->kind != JValue::Number || v->num < 0 || v->num > 1e9 || v->num != (double)(int32_t)v->num)
                fail(RRG_ERR_IO, "mi355x.sort_interval must be a non-negative integer");
            p->sort_interval = (int32_t)v->num;
        }
    }
    params_finalize(p);
}

// --------------------------------------------------------------- ini files
namespace {

std::string resolve(const char* base_dir, const char* path) {
    std::string p(path);
    for (char& c : p)
        if (c == '\\') c = '/';  // the sample params.json was written on Windows ("ini\\S.txt")
    if (base_dir && *base_dir && !p.empty() && p[0] != '/') return std::string(base_dir) + "/" + p;
    return p;
}

struct Component {
    std::string path;
    double mass, soft;
};

// read_particles_data_component (io_particles_csv.cpp:11-87): header "N t",
// then N rows "x y z vx vy vz".  m = M_component / N_cumulative and
// soft2 = soft^2 (io_particles_csv.cpp:46-47).
int64_t read_component(const Component& c, int64_t prev_n, int64_t cap, double* pos, double* vel, double* mass,
                       double* soft2) {
    std::FILE* f = std::fopen(c.path.c_str(), "r");
    if (!f) fail(RRG_ERR_IO, "Can't read simple particles data: cannot open '" + c.path + "'");
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> guard(f, &std::fclose);
    double hn = 0, ht = 0;
    if (std::fscanf(f, "%lf %lf", &hn, &ht) != 2) fail(RRG_ERR_IO, "Failed to read ini header");
    const int64_t ini_n = int64_t(hn);
    const int64_t total = prev_n + ini_n;
    const double pm = c.mass / double(total);
    const double ps2 = c.soft * c.soft;
    int64_t rows = 0;
    double r[6];
    while (true) {
        int got = std::fscanf(f, "%lf %lf %lf %lf %lf %lf", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5]);
        if (got == EOF || got == 0) break;
        if (got != 6) fail(RRG_ERR_IO, "ini row with fewer than 6 columns in '" + c.path + "'");
        const int64_t k = prev_n + rows;
        if (pos && k < cap) {
            for (int q = 0; q < 3; ++q) {
                pos[3 * k + q] = r[q];
                vel[3 * k + q] = r[3 + q];
            }
            mass[k] = pm;
            soft2[k] = ps2;
        }
        ++rows;
    }
    if (rows != ini_n)
        fail(RRG_ERR_IO, "Rows header/file mismatch (" + std::to_string(total) + "/" + std::to_string(prev_n + rows) + ")");
    return ini_n;
}

}  // namespace

void ini_read(const rrg_ini_params* ini, const char* base_dir, int64_t* n, double* pos, double* vel, double* mass,
              double* soft2) {
    std::vector<Component> comps;
    if (ini->ini_file_star[0]) comps.push_back({resolve(base_dir, ini->ini_file_star), ini->mass_star, ini->soft_star});
    if (ini->ini_file_dark[0]) comps.push_back({resolve(base_dir, ini->ini_file_dark), ini->mass_dark, ini->soft_dark});
    const int64_t cap = pos ? *n : 0;
    if (pos && (!vel || !mass || !soft2)) fail(RRG_ERR_INVALID, "rrg_ini_read: vel, mass and soft2 are required with pos");
    int64_t total = 0;
    for (const Component& c : comps) total += read_component(c, total, cap, pos, vel, mass, soft2);
    if (pos && total != cap)
        fail(RRG_ERR_INVALID, "rrg_ini_read: buffers sized for " + std::to_string(cap) + " particles, files hold " +
                                  std::to_string(total));
    *n = total;
}

}  // namespace io
}  // namespace rrg
