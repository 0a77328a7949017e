// C++ host over the libwcsph C ABI, keeping the reference's simulator
// interface: FluidSimulator (src/simulation/fluid_simulator.h:6-54) with
// update / reset / get_position_data / get_fluid_particles, and the
// extras of CUDASPHSimulator (src/cuda/SPH/sph.cuh:14-86):
// compute_state_metrics, set_frozen_config, was_scheduled_step, and the
// snapshot access the reference grants Snapshot/Application as friends.
// No OpenGL: positions are copied to host memory on request.
#pragma once

#include <cstdint>
#include <optional>
#include <span>
#include <string>
#include <utility>
#include <vector>

#include "wcsph.h"

namespace wcsph_host {

struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

// Column-major 4x4 like glm::mat4: m[4 * column + row].
struct mat4 {
    float m[16];
    static mat4 identity();
    mat4 operator*(const mat4& o) const;
    vec3 apply(vec3 p, float w = 1.f) const;
    mat4 inverse() const;
};
// glm::translate / glm::rotate / glm::scale: M * T, M * R, M * S.
mat4 translate(const mat4& m, vec3 t);
mat4 rotate(const mat4& m, float angle, vec3 axis);
mat4 scale(const mat4& m, vec3 s);

// BoundingBox (src/cuda/util.cuh:9-19): [min, max] in the box frame,
// model maps it to world.
struct BoundingBox {
    vec3 min, max;
    mat4 model = mat4::identity();
    mat4 model_inv = mat4::identity();
    wcsph_box to_abi() const;
};

// The fluid box: a collision object whose surface is procedural::cube
// (src/render/box.cpp:7-29, src/render/geometry.cpp:381-).
class Box {
public:
    explicit Box(vec3 half_size);
    void set_model(const mat4& m);
    const BoundingBox& bounding_box() const { return bb_; }
    std::vector<float> triangles() const;  // world space, 9 floats per triangle

private:
    vec3 half_;
    BoundingBox bb_;
};

// Triangle soup of a mesh file (Wavefront OBJ: v and f records, polygons
// fanned), transformed by model; replaces Geometry::from_file +
// Geometry::load_triangles (src/render/geometry.cpp:185-207).
std::vector<float> load_obj_triangles(const std::string& path, const mat4& model);

class FluidSimulator {
public:
    struct grid_dims_t {
        grid_dims_t(unsigned x, unsigned y, unsigned z) : x(x), y(y), z(z) {}
        grid_dims_t(unsigned a) : grid_dims_t{a, a, a} {}
        unsigned x, y, z;
    };

    // fluid_simulator.h:22-28; collision objects are world-space triangle
    // soups, the external force is a list of rules (wcsph_force_rule) in
    // place of the reference's CUDA lambda text, and `domain` is the world
    // region the box may later move through (wcsph_scene.domain_*).
    struct opts_t {
        vec3 origin;
        grid_dims_t grid_dims;
        BoundingBox bounding_box;
        std::vector<std::vector<float>> collision_objects;
        std::vector<wcsph_force_rule> external_force;
        std::optional<std::pair<vec3, vec3>> domain;
        // Explicit particles instead of the centred lattice / in addition to
        // the sampled collision objects (float4 each; the benchmark scenes).
        std::vector<float> initial_fluid;
        std::vector<float> extra_boundary;
    };

    FluidSimulator(const opts_t& opts, float particle_radius);
    virtual ~FluidSimulator() = default;

    virtual void update(float delta) = 0;

    // float4 per particle: fluid (current state) then boundary.
    std::span<const float> get_position_data();

    virtual void reset();

    unsigned get_fluid_particles() const { return fluid_particles; }
    unsigned get_boundary_particles() const { return boundary_particles; }

protected:
    virtual void refresh_positions() {}
    void init_positions();

    std::vector<float> positions;
    std::vector<float> initial_fluid;  // explicit initial fluid (empty: the lattice)
    unsigned total_particles = 0, fluid_particles = 0, boundary_particles = 0;
    BoundingBox bounding_box;
    grid_dims_t grid_dims;
    vec3 origin;
    float particle_radius;
};

// The WCSPH step on the GPU through libwcsph.
class SPHSimulator final : public FluidSimulator {
public:
    SPHSimulator(const opts_t& opts, const wcsph_params& params);
    ~SPHSimulator() override;
    SPHSimulator(const SPHSimulator&) = delete;
    SPHSimulator& operator=(const SPHSimulator&) = delete;

    void update(float delta) override;
    void reset() override;

    void set_frozen_config(const wcsph_variant& v);
    wcsph_variant variant() const;
    bool was_scheduled_step() const { return false; }  // tuning never runs inside a step
    std::pair<double, double> compute_state_metrics();  // mean |v|, kinetic energy (FP64)
    void set_box(const BoundingBox& box);
    const wcsph_stats& last_stats() const { return stats_; }

    // Snapshot::save / Snapshot::load (src/cuda/SPH/snapshot.cu:10-106):
    // "" on success, else the message.
    std::string save_snapshot(const std::string& path, const std::string& app_name);
    std::string load_snapshot(const std::string& path, const std::string& app_name);

    // One-shot project autotuner; returns (variant, ms per step) per
    // candidate (empty candidates = the library's default space).
    std::vector<std::pair<wcsph_variant, double>> autotune(float dt_hint, unsigned steps_per_trial,
                                                           const std::vector<wcsph_variant>& candidates = {});

    wcsph_t* handle() { return h_; }

private:
    void refresh_positions() override;
    void check(int rc, const char* what) const;

    wcsph_t* h_ = nullptr;
    wcsph_stats stats_{};
    bool stale_ = true;
};

// Parses a reference KTT frozen-configuration JSON (measurements/configs/
// *.json; config_loader.cpp:8-54) into a variant: the force pass's block
// size (compute_pressure_accel_n_normal_block) rounded up to a whole
// wavefront; neighbour lists are always built.  Throws on a bad file.
// Variant from a reference KTT frozen configuration (measurements/configs/*.json):
// CELL_SIZE_MULT -> cell_size_mult (per-particle builder when it is not 1),
// compute_pressure_accel_n_normal_block -> block_size (a whole wavefront),
// all three SPH passes with *_nlist = 0 -> the cell-loop form.  `ignored`
// receives "name: reason" for every parameter without a counterpart.
wcsph_variant variant_from_ktt_config(const std::string& path, const wcsph_variant& base,
                                      std::vector<std::string>* ignored = nullptr);
std::string ktt_ignored_reason(const std::string& name);

}  // namespace wcsph_host
