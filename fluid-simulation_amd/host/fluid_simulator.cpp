#include "fluid_simulator.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace wcsph_host {

// ---------------------------------------------------------------- math
mat4 mat4::identity() {
    mat4 r{};
    for (int i = 0; i < 4; ++i) r.m[5 * i] = 1.f;
    return r;
}

mat4 mat4::operator*(const mat4& o) const {
    mat4 r{};
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += (double)m[4 * k + row] * (double)o.m[4 * c + k];
            r.m[4 * c + row] = (float)acc;
        }
    return r;
}

vec3 mat4::apply(vec3 p, float w) const {
    const float in[4] = {p.x, p.y, p.z, w};
    float out[3];
    for (int row = 0; row < 3; ++row) {
        double acc = 0.0;
        for (int k = 0; k < 4; ++k) acc += (double)m[4 * k + row] * in[k];
        out[row] = (float)acc;
    }
    return {out[0], out[1], out[2]};
}

mat4 mat4::inverse() const {
    // Gauss-Jordan with partial pivoting in double
    double a[4][8];
    for (int row = 0; row < 4; ++row)
        for (int c = 0; c < 4; ++c) {
            a[row][c] = m[4 * c + row];
            a[row][4 + c] = row == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int row = c + 1; row < 4; ++row)
            if (std::fabs(a[row][c]) > std::fabs(a[piv][c])) piv = row;
        if (std::fabs(a[piv][c]) < 1e-300) throw std::runtime_error("singular model matrix");
        if (piv != c)
            for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int row = 0; row < 4; ++row) {
            if (row == c) continue;
            const double f = a[row][c];
            for (int k = 0; k < 8; ++k) a[row][k] -= f * a[c][k];
        }
    }
    mat4 r{};
    for (int row = 0; row < 4; ++row)
        for (int c = 0; c < 4; ++c) r.m[4 * c + row] = (float)a[row][4 + c];
    return r;
}

mat4 translate(const mat4& m, vec3 t) {
    mat4 tr = mat4::identity();
    tr.m[12] = t.x;
    tr.m[13] = t.y;
    tr.m[14] = t.z;
    return m * tr;
}

mat4 rotate(const mat4& m, float angle, vec3 axis) {
    // Rodrigues: R = c I + (1 - c) a a^T + s [a]x, a the unit axis
    const double len = std::sqrt((double)axis.x * axis.x + (double)axis.y * axis.y + (double)axis.z * axis.z);
    const double a[3] = {axis.x / len, axis.y / len, axis.z / len};
    const double c = std::cos((double)angle), s = std::sin((double)angle);
    const double k[3][3] = {{0.0, -a[2], a[1]}, {a[2], 0.0, -a[0]}, {-a[1], a[0], 0.0}};
    mat4 r = mat4::identity();
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col)
            r.m[4 * col + row] = (float)((row == col ? c : 0.0) + (1.0 - c) * a[row] * a[col] + s * k[row][col]);
    return m * r;
}

mat4 scale(const mat4& m, vec3 s) {
    mat4 sc = mat4::identity();
    sc.m[0] = s.x;
    sc.m[5] = s.y;
    sc.m[10] = s.z;
    return m * sc;
}

wcsph_box BoundingBox::to_abi() const {
    wcsph_box b{};
    const float lo[3] = {min.x, min.y, min.z}, hi[3] = {max.x, max.y, max.z};
    for (int a = 0; a < 3; ++a) {
        b.min[a] = lo[a];
        b.max[a] = hi[a];
    }
    std::memcpy(b.model, model.m, sizeof(b.model));
    std::memcpy(b.model_inv, model_inv.m, sizeof(b.model_inv));
    return b;
}

Box::Box(vec3 half_size) : half_(half_size) {
    bb_.min = {-half_size.x, -half_size.y, -half_size.z};
    bb_.max = half_size;
}

void Box::set_model(const mat4& m) {
    bb_.model = m;
    bb_.model_inv = m.inverse();
}

std::vector<float> Box::triangles() const {
    const float prm[3] = {half_.x, half_.y, half_.z};
    uint32_t nt = 0;
    if (wcsph_mesh_procedural(WCSPH_MESH_CUBE, prm, 0, bb_.model.m, nullptr, 0, &nt) != WCSPH_OK)
        throw std::runtime_error("box mesh: bad half size");
    std::vector<float> tri(9 * (size_t)nt);
    wcsph_mesh_procedural(WCSPH_MESH_CUBE, prm, 0, bb_.model.m, tri.data(), nt, &nt);
    return tri;
}

std::vector<float> load_obj_triangles(const std::string& path, const mat4& model) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open mesh file: " + path);
    std::vector<vec3> v;
    std::vector<float> tri;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string tag;
        is >> tag;
        if (tag == "v") {
            vec3 p;
            is >> p.x >> p.y >> p.z;
            if (!is) throw std::runtime_error("bad vertex record in " + path);
            v.push_back(p);
        } else if (tag == "f") {
            std::vector<size_t> idx;
            std::string tok;
            while (is >> tok) {
                const long k = std::stol(tok.substr(0, tok.find('/')));
                const long res = k < 0 ? (long)v.size() + k : k - 1;
                if (res < 0 || res >= (long)v.size()) throw std::runtime_error("bad face index in " + path);
                idx.push_back((size_t)res);
            }
            for (size_t k = 1; k + 1 < idx.size(); ++k)
                for (size_t q : {idx[0], idx[k], idx[k + 1]}) {
                    const vec3 w = model.apply(v[q]);
                    tri.insert(tri.end(), {w.x, w.y, w.z});
                }
        }
    }
    return tri;
}

// ---------------------------------------------------------------- simulators
FluidSimulator::FluidSimulator(const opts_t& opts, float r)
    : initial_fluid{opts.initial_fluid},
      fluid_particles{opts.initial_fluid.empty() ? opts.grid_dims.x * opts.grid_dims.y * opts.grid_dims.z
                                                 : (unsigned)(opts.initial_fluid.size() / 4)},
      bounding_box{opts.bounding_box},
      grid_dims{opts.grid_dims},
      origin{opts.origin},
      particle_radius{r} {
    if (initial_fluid.empty() && (grid_dims.x < 2 || grid_dims.y < 2 || grid_dims.z < 2))
        throw std::invalid_argument("grid_dims must be > 1");
    if (initial_fluid.size() % 4) throw std::invalid_argument("initial_fluid holds float4 per particle");
    init_positions();
    // boundary particles: voxel centres of every collision object at
    // 0.75 * 2r (src/simulation/SPH/boundary.cpp:8-35)
    const float voxel = 2.f * particle_radius * 0.75f;
    for (const auto& tri : opts.collision_objects) {
        const uint32_t nt = (uint32_t)(tri.size() / 9);
        uint32_t cnt = 0;
        if (wcsph_sample_mesh(tri.data(), nt, voxel, nullptr, 0, &cnt) != WCSPH_OK)
            throw std::runtime_error("collision object: bad mesh");
        const size_t off = positions.size();
        positions.resize(off + 4 * (size_t)cnt);
        wcsph_sample_mesh(tri.data(), nt, voxel, positions.data() + off, cnt, &cnt);
    }
    if (opts.extra_boundary.size() % 4) throw std::invalid_argument("extra_boundary holds float4 per particle");
    positions.insert(positions.end(), opts.extra_boundary.begin(), opts.extra_boundary.end());
    total_particles = (unsigned)(positions.size() / 4);
    boundary_particles = total_particles - fluid_particles;
}

// x-major lattice at spacing 2r centred on origin (fluid_simulator.cpp:24-54).
void FluidSimulator::init_positions() {
    if (!initial_fluid.empty()) {
        if (positions.size() < initial_fluid.size()) positions.resize(initial_fluid.size());
        std::copy(initial_fluid.begin(), initial_fluid.end(), positions.begin());
        return;
    }
    const float sp = 2.f * particle_radius;
    const vec3 start{origin.x - (float)(grid_dims.x - 1) * particle_radius,
                     origin.y - (float)(grid_dims.y - 1) * particle_radius,
                     origin.z - (float)(grid_dims.z - 1) * particle_radius};
    if (positions.size() < 4 * (size_t)fluid_particles) positions.resize(4 * (size_t)fluid_particles);
    size_t i = 0;
    for (unsigned x = 0; x < grid_dims.x; ++x)
        for (unsigned y = 0; y < grid_dims.y; ++y)
            for (unsigned z = 0; z < grid_dims.z; ++z, ++i) {
                positions[4 * i + 0] = start.x + (float)x * sp;
                positions[4 * i + 1] = start.y + (float)y * sp;
                positions[4 * i + 2] = start.z + (float)z * sp;
                positions[4 * i + 3] = 1.f;
            }
}

std::span<const float> FluidSimulator::get_position_data() {
    refresh_positions();
    return {positions.data(), positions.size()};
}

void FluidSimulator::reset() { init_positions(); }

SPHSimulator::SPHSimulator(const opts_t& opts, const wcsph_params& params)
    : FluidSimulator(opts, params.particle_radius) {
    wcsph_params p = params;
    if (opts.external_force.size() > WCSPH_MAX_FORCE_RULES)
        throw std::invalid_argument("at most WCSPH_MAX_FORCE_RULES external force rules");
    p.n_force_rules = (uint32_t)opts.external_force.size();
    for (size_t k = 0; k < opts.external_force.size(); ++k) p.force_rules[k] = opts.external_force[k];
    wcsph_scene sc{};
    sc.n_fluid = fluid_particles;
    sc.n_boundary = boundary_particles;
    sc.fluid_pos = positions.data();
    sc.boundary_pos = boundary_particles ? positions.data() + 4 * (size_t)fluid_particles : nullptr;
    sc.box = bounding_box.to_abi();
    if (opts.domain) {
        const vec3 lo = opts.domain->first, hi = opts.domain->second;
        const float l[3] = {lo.x, lo.y, lo.z}, u[3] = {hi.x, hi.y, hi.z};
        for (int a = 0; a < 3; ++a) {
            sc.domain_min[a] = l[a];
            sc.domain_max[a] = u[a];
        }
    }
    if (wcsph_create(&p, &sc, &h_) != WCSPH_OK) throw std::runtime_error(wcsph_last_error(nullptr));
    stale_ = false;
}

SPHSimulator::~SPHSimulator() { wcsph_destroy(h_); }

void SPHSimulator::check(int rc, const char* what) const {
    if (rc != WCSPH_OK) throw std::runtime_error(std::string(what) + ": " + wcsph_last_error(h_));
}

void SPHSimulator::update(float delta) {
    check(wcsph_step(h_, delta, 1, &stats_), "wcsph_step");
    stale_ = true;
}

void SPHSimulator::reset() {
    FluidSimulator::reset();
    check(wcsph_set_state(h_, positions.data(), nullptr), "wcsph_set_state");
    stale_ = false;
}

void SPHSimulator::refresh_positions() {
    if (!stale_) return;
    check(wcsph_get_state(h_, positions.data(), nullptr, nullptr, nullptr, nullptr), "wcsph_get_state");
    stale_ = false;
}

void SPHSimulator::set_frozen_config(const wcsph_variant& v) { check(wcsph_set_variant(h_, &v), "wcsph_set_variant"); }

wcsph_variant SPHSimulator::variant() const {
    wcsph_variant v{};
    wcsph_get_variant(h_, &v);
    return v;
}

std::pair<double, double> SPHSimulator::compute_state_metrics() {
    double v = 0.0, ke = 0.0;
    check(wcsph_state_metrics(h_, &v, &ke), "wcsph_state_metrics");
    return {v, ke};
}

void SPHSimulator::set_box(const BoundingBox& box) {
    const wcsph_box b = box.to_abi();
    check(wcsph_set_box(h_, 0, &b), "wcsph_set_box");
    bounding_box = box;
}

std::string SPHSimulator::save_snapshot(const std::string& path, const std::string& app_name) {
    std::vector<float> pos(4 * (size_t)fluid_particles), vel(4 * (size_t)fluid_particles);
    if (wcsph_get_state(h_, pos.data(), vel.data(), nullptr, nullptr, nullptr) != WCSPH_OK)
        return wcsph_last_error(h_);
    if (wcsph_snapshot_write(path.c_str(), app_name.c_str(), fluid_particles, pos.data(), vel.data()) != WCSPH_OK)
        return wcsph_snapshot_error();
    return {};
}

std::string SPHSimulator::load_snapshot(const std::string& path, const std::string& app_name) {
    std::vector<float> pos(4 * (size_t)fluid_particles), vel(4 * (size_t)fluid_particles);
    if (wcsph_snapshot_read(path.c_str(), app_name.c_str(), fluid_particles, pos.data(), vel.data()) != WCSPH_OK)
        return wcsph_snapshot_error();
    if (wcsph_set_state(h_, pos.data(), vel.data()) != WCSPH_OK) return wcsph_last_error(h_);
    stale_ = true;
    return {};
}

std::vector<std::pair<wcsph_variant, double>> SPHSimulator::autotune(float dt_hint, unsigned steps_per_trial,
                                                                     const std::vector<wcsph_variant>& candidates) {
    std::vector<wcsph_variant> cands = candidates;
    if (cands.empty()) {
        // block size x builder (x its workgroup count) at cell size h, and the
        // per-particle builder at the finer cell sizes 2/3 h and h/2, from the
        // current variant
        wcsph_variant base = variant();
        base.cell_size_mult = 1.f;
        for (uint32_t bs : {128u, 256u, 512u}) {
            wcsph_variant v = base;
            v.block_size = bs;
            v.neighbor_mode = WCSPH_NEIGHBORS_PARTICLE;
            cands.push_back(v);
        }
        for (uint32_t g : {4096u, 8192u}) {
            wcsph_variant v = base;
            v.neighbor_mode = WCSPH_NEIGHBORS_CELL;
            v.neighbor_groups = g;
            cands.push_back(v);
        }
        for (uint32_t mode : {(uint32_t)WCSPH_NEIGHBORS_CELL_LOOP, (uint32_t)WCSPH_NEIGHBORS_BRICK}) {
            wcsph_variant v = base;
            v.neighbor_mode = mode;
            cands.push_back(v);
        }
        for (float m : {2.f / 3.f, 0.5f}) {
            wcsph_variant v = base;
            v.neighbor_mode = WCSPH_NEIGHBORS_PARTICLE;
            v.block_size = 256;
            v.cell_size_mult = m;
            cands.push_back(v);
        }
    }
    std::vector<double> ms(cands.size(), 0.0);
    wcsph_variant best{};
    check(wcsph_autotune(h_, dt_hint, steps_per_trial, cands.data(), (uint32_t)cands.size(), ms.data(), &best),
          "wcsph_autotune");
    std::vector<std::pair<wcsph_variant, double>> out;
    for (size_t k = 0; k < cands.size(); ++k) out.push_back({cands[k], ms[k]});
    stale_ = true;
    return out;
}

// ---------------------------------------------------------------- KTT configs
// Why a KTT parameter of the reference's step (step_tuner.cuh:80-121) has no
// counterpart here.
std::string ktt_ignored_reason(const std::string& name) {
    auto ends = [&](const char* suf) {
        const size_t n = std::strlen(suf);
        return name.size() >= n && name.compare(name.size() - n, n, suf) == 0;
    };
    if (ends("_u_n")) return "loop unrolling is fixed at compile time";
    if (ends("_shared") || ends("_cache_log2")) return "the LDS forms are chosen by neighbor_mode";
    if (name.find("TABLE_SIZE") != std::string::npos) return "the cell grid is dense: no hash table";
    if (ends("_block")) return "one block size serves every particle pass (the force pass's is taken)";
    if (name == "EXTERNAL_FORCE") return "the external force is the scene's force rules";
    if (name == "KERNEL_DIR") return "no run-time compilation";
    return "no counterpart";
}

wcsph_variant variant_from_ktt_config(const std::string& path, const wcsph_variant& base,
                                      std::vector<std::string>* ignored) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open config file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string txt = ss.str();
    // Innermost {...} objects, skipping braces inside strings; each one is
    // a configuration entry {"Name", "Value", "ValueType"}.
    std::vector<std::pair<size_t, size_t>> objects;
    std::vector<size_t> open;
    std::vector<bool> has_child;
    bool in_str = false;
    for (size_t i = 0; i < txt.size(); ++i) {
        const char c = txt[i];
        if (in_str) {
            if (c == '\\') ++i;
            else if (c == '"') in_str = false;
            continue;
        }
        if (c == '"') {
            in_str = true;
        } else if (c == '{') {
            if (!has_child.empty()) has_child.back() = true;
            open.push_back(i);
            has_child.push_back(false);
        } else if (c == '}' && !open.empty()) {
            if (!has_child.back()) objects.push_back({open.back(), i});
            open.pop_back();
            has_child.pop_back();
        }
    }
    // value of "key" in one flat object: a string (escapes kept) or a bare token
    auto field = [&](size_t b, size_t e, const std::string& key, std::string& out) {
        bool s2 = false;
        for (size_t i = b; i < e; ++i) {
            if (txt[i] != '"') continue;
            size_t j = i + 1;
            while (j < e && txt[j] != '"') j += txt[j] == '\\' ? 2 : 1;
            const std::string tok = txt.substr(i + 1, j - i - 1);
            i = j;
            if (s2) { s2 = false; continue; }  // a string value, not a key
            size_t k = j + 1;
            while (k < e && std::isspace((unsigned char)txt[k])) ++k;
            if (k >= e || txt[k] != ':') continue;
            ++k;
            while (k < e && std::isspace((unsigned char)txt[k])) ++k;
            if (tok != key) {
                if (k < e && txt[k] == '"') s2 = true;
                continue;
            }
            if (k < e && txt[k] == '"') {
                size_t q = k + 1;
                while (q < e && txt[q] != '"') q += txt[q] == '\\' ? 2 : 1;
                out = txt.substr(k + 1, q - k - 1);
            } else {
                size_t q = k;
                while (q < e && txt[q] != ',' && txt[q] != '}' && !std::isspace((unsigned char)txt[q])) ++q;
                out = txt.substr(k, q - k);
            }
            return true;
        }
        return false;
    };
    if (objects.empty()) throw std::runtime_error("Config JSON must be an array or contain a 'Configuration' array");
    auto number = [](const std::string& name, const std::string& value, auto parse) {
        try {
            size_t used = 0;
            const auto x = parse(value, &used);
            if (used != value.size()) throw std::invalid_argument(value);
            return x;
        } catch (const std::exception&) {
            throw std::runtime_error("Config value of " + name + " is not a number: '" + value + "'");
        }
    };
    auto to_long = [](const std::string& t, size_t* u) { return std::stol(t, u); };
    auto to_float = [](const std::string& t, size_t* u) { return std::stof(t, u); };
    wcsph_variant v = base;
    int nlist_off = 0;  // SPH passes configured without lists
    for (const auto& [b, e] : objects) {
        std::string name, value, type;
        if (!field(b, e, "Name", name) || !field(b, e, "Value", value) || !field(b, e, "ValueType", type))
            throw std::runtime_error("Config entry missing Name, Value, or ValueType");
        if (name == "compute_pressure_accel_n_normal_block") {
            const long bs = number(name, value, to_long);
            v.block_size = (uint32_t)std::clamp<long>((bs + 63) / 64 * 64, 64, 1024);
        } else if (name == "CELL_SIZE_MULT") {
            v.cell_size_mult = number(name, value, to_float);
        } else if (name == "compute_rho_p_nlist" || name == "compute_pressure_accel_n_normal_nlist" ||
                   name == "compute_non_pressure_accel_nlist") {
            if (number(name, value, to_long) == 0) ++nlist_off;
        } else if (ignored) {
            ignored->push_back(name + ": " + ktt_ignored_reason(name));
        }
    }
    if (v.cell_size_mult != 0.f && v.cell_size_mult != 1.f) v.neighbor_mode = WCSPH_NEIGHBORS_PARTICLE;
    else if (nlist_off == 3) v.neighbor_mode = WCSPH_NEIGHBORS_CELL_LOOP;
    return v;
}

}  // namespace wcsph_host
