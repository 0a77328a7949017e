#include "apps.h"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <memory>
#include <stdexcept>

namespace wcsph_host {

namespace {

constexpr float kPiF = 3.14159265358979323846f;

wcsph_params preset(int which) {
    wcsph_params p{};
    wcsph_default_params(which, &p);
    return p;
}

wcsph_force_rule rule(uint32_t region, uint32_t kind, float y_below, float radius2, vec3 value, float magnitude) {
    wcsph_force_rule r{};
    r.region = region;
    r.kind = kind;
    r.y_below = y_below;
    r.radius2 = radius2;
    r.value[0] = value.x;
    r.value[1] = value.y;
    r.value[2] = value.z;
    r.magnitude = magnitude;
    return r;
}

void add_dragon(AppScene& a, const std::optional<std::string>& mesh, const mat4& model) {
    if (!mesh) {
        std::cerr << a.name << ": no dragon mesh given (--mesh stanford-dragon.obj); running without the dragon\n";
        return;
    }
    a.opts.collision_objects.push_back(load_obj_triangles(*mesh, model));
}

// Reference apps: opts{origin, grid_dims, bounding_box, collision_objects,
// external_force} exactly as their setup_scene builds them.
AppScene reference_app(const std::string& name, const std::optional<std::string>& mesh) {
    AppScene a{name, preset(WCSPH_PRESET_REFERENCE), {{0.f, 0.f, 0.f}, 30u, {}, {}, {}, {}, {}, {}}, {}};
    const mat4 I = mat4::identity();
    if (name == "dragon_collision") {  // app/dragon_collision/application.cu
        Box box({2.5f, 0.7f, 0.7f});
        a.opts.origin = {-1.5f, 0.f, 0.f};
        a.opts.bounding_box = box.bounding_box();
        a.opts.collision_objects.push_back(box.triangles());
        add_dragon(a, mesh, rotate(translate(I, {1.f, -0.45f, 0.f}), -kPiF * 0.5f, {1.f, 0.f, 0.f}));
    } else if (name == "surface_tension") {  // app/surface_tension/application.cu
        Box box({0.8f, 0.8f, 0.8f});
        a.opts.bounding_box = box.bounding_box();
        a.opts.external_force.push_back(rule(WCSPH_REGION_ALL, WCSPH_FORCE_CONSTANT, 0.f, 0.f, {0.f, 9.81f, 0.f}, 0.f));
    } else if (name == "tilting_box") {  // app/tilting_box/application.cu
        const vec3 half{2.3f, 0.7f, 0.6f};
        Box box(half);
        a.opts.bounding_box = box.bounding_box();
        // the box turns about z: the grid covers its bounding sphere
        const float rad = std::sqrt(half.x * half.x + half.y * half.y + half.z * half.z) + 0.1f;
        a.opts.domain = std::make_pair(vec3{-rad, -rad, -rad}, vec3{rad, rad, rad});
        auto update_time = std::make_shared<float>(0.f);
        a.update_objects = [update_time, half](SPHSimulator& sim, float delta) {
            *update_time += std::min(delta, 1.f / 30.f);
            Box b(half);
            b.set_model(rotate(mat4::identity(), std::sin(*update_time / 5.f) / 5.f, {0.f, 0.f, 1.f}));
            sim.set_box(b.bounding_box());
        };
    } else if (name == "fountain") {  // app/fountain/application.cu
        Box box({2.5f, 1.f, 0.82f});
        box.set_model(translate(I, {0.f, 0.3f, 0.f}));
        a.opts.origin = {-1.5f, 0.25f, 0.f};
        a.opts.grid_dims = 40u;
        a.opts.bounding_box = box.bounding_box();
        a.opts.collision_objects.push_back(box.triangles());
        add_dragon(a, mesh, rotate(translate(I, {1.f, -0.45f, 0.f}), -kPiF * 0.5f, {1.f, 0.f, 0.f}));
        a.opts.external_force.push_back(
            rule(WCSPH_REGION_XZ_DISC_Y_BELOW, WCSPH_FORCE_CONSTANT, 0.f, 0.2f, {0.f, 25.f, 0.f}, 0.f));
        a.opts.external_force.push_back(rule(WCSPH_REGION_Y_BELOW, WCSPH_FORCE_TOWARD_Y_AXIS, -0.3f, 0.f, {}, 2.f));
    } else if (name == "vortex") {  // app/vortex/application.cu
        const float S = 1.5f;
        Box box({1.f * S, 0.7f * S, 1.f * S});
        a.opts.origin = {-0.25f * S, 0.f, 0.f};
        a.opts.grid_dims = 50u;
        a.opts.bounding_box = box.bounding_box();
        mat4 dm = scale(I, {S, S, S});
        dm = translate(dm, {0.75f, -0.45f, 0.f});
        dm = rotate(dm, -kPiF * 0.5f, {0.f, 1.f, 0.f});
        dm = rotate(dm, -kPiF * 0.5f, {1.f, 0.f, 0.f});
        add_dragon(a, mesh, dm);
        a.opts.collision_objects.push_back(box.triangles());
        a.opts.external_force.push_back(rule(WCSPH_REGION_Y_BELOW, WCSPH_FORCE_SWIRL_Y_AXIS, -0.5f, 0.f, {}, 2.f));
    } else if (name == "BIG") {  // app/BIG/application.cu
        const float S = 2.25f;
        Box box({2.5f * S, 0.7f * S, 0.7f * S});
        a.opts.origin = {-1.5f * S, 0.f, 0.f};
        a.opts.grid_dims = 75u;
        a.opts.bounding_box = box.bounding_box();
        mat4 dm = scale(I, {S, S, S});
        dm = translate(dm, {1.f, -0.45f, 0.f});
        dm = rotate(dm, -kPiF * 0.5f, {1.f, 0.f, 0.f});
        add_dragon(a, mesh, dm);
        a.opts.collision_objects.push_back(box.triangles());
    } else {
        throw std::invalid_argument("Unknown app: " + name);
    }
    return a;
}

// Benchmark dam breaks: the seeded jittered lattice and 3 wall shells of
// wcsph_make_dam_break, the box's min corner at the origin.
AppScene dam_break(const std::string& name, uint32_t nx, uint32_t ny, uint32_t nz, vec3 box_max, bool adaptive) {
    AppScene a{name, preset(WCSPH_PRESET_BENCHMARK), {{0.f, 0.f, 0.f}, 2u, {}, {}, {}, {}, {}, {}}, {}};
    a.params.adaptive_dt = adaptive ? 1u : 0u;
    const float r = a.params.particle_radius;
    const uint32_t dims[3] = {nx, ny, nz};
    const float lmin[3] = {0.f + r, 0.f + r, 0.f + r};
    const float bmin[3] = {0.f, 0.f, 0.f}, bmax[3] = {box_max.x, box_max.y, box_max.z};
    const float jitter = 0.1f * r;
    uint32_t nf = 0, nb = 0;
    if (wcsph_make_dam_break(r, dims, lmin, jitter, 12345, bmin, bmax, 3, nullptr, &nf, nullptr, &nb) != WCSPH_OK)
        throw std::runtime_error("bad dam-break description");
    a.opts.initial_fluid.resize(4 * (size_t)nf);
    a.opts.extra_boundary.resize(4 * (size_t)nb);
    wcsph_make_dam_break(r, dims, lmin, jitter, 12345, bmin, bmax, 3, a.opts.initial_fluid.data(), &nf,
                         a.opts.extra_boundary.data(), &nb);
    a.opts.bounding_box.min = {0.f, 0.f, 0.f};
    a.opts.bounding_box.max = box_max;
    return a;
}

std::vector<float> sampled(int kind, const float prm[3], uint32_t res, const mat4& model, float voxel) {
    uint32_t nt = 0, np = 0;
    if (wcsph_mesh_procedural(kind, prm, res, model.m, nullptr, 0, &nt) != WCSPH_OK)
        throw std::runtime_error("bad mesh");
    std::vector<float> tri(9 * (size_t)nt);
    wcsph_mesh_procedural(kind, prm, res, model.m, tri.data(), nt, &nt);
    wcsph_sample_mesh(tri.data(), nt, voxel, nullptr, 0, &np);
    std::vector<float> pts(4 * (size_t)np);
    wcsph_sample_mesh(tri.data(), nt, voxel, pts.data(), np, &np);
    return pts;
}

// Column collapse onto a sphere + torus obstacle (8.3M fluid particles):
// a 2 x 1 x 1 m block with 10% of its particles in denser clusters.
AppScene obstacle_collapse() {
    AppScene a{"obstacle_collapse_8m", preset(WCSPH_PRESET_BENCHMARK), {{0.f, 0.f, 0.f}, 2u, {}, {}, {}, {}, {}, {}},
               {}};
    a.params.adaptive_dt = 1;
    const float r = 0.003125f;
    a.params.particle_radius = r;
    const float bmin[3] = {-2.1f, 0.f, -0.5f}, bmax[3] = {-0.1f, 1.f, 0.5f};
    const float jitter = 0.1f * r, cspacing = 1.8f * r;
    uint32_t nf = 0;
    if (wcsph_make_clustered_block(r, bmin, bmax, 0.1f, 0.1f, cspacing, jitter, 4242, nullptr, &nf) != WCSPH_OK)
        throw std::runtime_error("bad clustered block");
    a.opts.initial_fluid.resize(4 * (size_t)nf);
    wcsph_make_clustered_block(r, bmin, bmax, 0.1f, 0.1f, cspacing, jitter, 4242, a.opts.initial_fluid.data(), &nf);
    const float voxel = 1.0f * 2.0f * r;
    const float sph[3] = {0.5f, 0.f, 0.f}, tor[3] = {0.8f, 0.2f, 0.f};
    const mat4 I = mat4::identity();
    auto s = sampled(WCSPH_MESH_SPHERE, sph, 96, translate(I, {1.1f, 0.5f, 0.f}), voxel);
    auto t = sampled(WCSPH_MESH_TORUS, tor, 96, translate(I, {1.1f, 0.2f, 0.f}), voxel);
    a.opts.extra_boundary = s;
    a.opts.extra_boundary.insert(a.opts.extra_boundary.end(), t.begin(), t.end());
    a.opts.bounding_box.min = {-2.2f, 0.f, -1.2f};
    a.opts.bounding_box.max = {2.2f, 2.f, 1.2f};
    return a;
}

}  // namespace

std::vector<std::string> app_names() {
    return {"dam_break_32k", "dam_break_1m", "obstacle_collapse_8m", "dragon_collision",
            "surface_tension", "tilting_box", "fountain", "vortex", "BIG"};
}

AppScene make_app(const std::string& name, const std::optional<std::string>& mesh) {
    if (name == "dam_break_32k") return dam_break(name, 32, 32, 32, {1.6f, 1.f, 1.f}, false);
    if (name == "dam_break_1m") return dam_break(name, 128, 128, 64, {4.f, 2.6f, 2.f}, true);
    if (name == "obstacle_collapse_8m") return obstacle_collapse();
    return reference_app(name, mesh);
}

}  // namespace wcsph_host
