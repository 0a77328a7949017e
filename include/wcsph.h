/*
 * wcsph.h — C ABI of the MI355X (gfx950) WCSPH solver step.
 *
 * One handle owns one scene: fluid particles, static boundary particles
 * (Akinci), a box with collision response, and every device buffer the
 * step needs.  All entry points return an int status (WCSPH_OK == 0) and
 * never throw; wcsph_last_error() gives the message of the last failure on
 * a handle (or of the last failed wcsph_create when the handle is NULL).
 *
 * Plain pointers and sizes only.  Host pointers unless the name says
 * "device".  Positions and velocities are float4 (x, y, z, w) per particle,
 * the reference's own layout (src/cuda/SPH/particle_data.cuh:217-221);
 * w = 1 for positions and 0 for velocities.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef WCSPH_H
#define WCSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
enum {
    WCSPH_OK = 0,
    WCSPH_ERR_INVALID_ARGUMENT = 1,
    WCSPH_ERR_HIP = 2,                /* a HIP runtime call failed */
    WCSPH_ERR_NEIGHBOR_OVERFLOW = 3,  /* a particle had more neighbours than the list holds */
    WCSPH_ERR_DOMAIN = 4,             /* a particle left the cell grid of the scene */
    WCSPH_ERR_NONFINITE = 5,          /* a position or velocity became NaN/Inf */
    WCSPH_ERR_NO_DEVICE = 6,
    WCSPH_ERR_OUT_OF_MEMORY = 7
};

/* Bits of wcsph_stats.flags (sticky within one wcsph_step call).
 * LISTS_GROWN is a note, not an error: a step's neighbour lists overflowed,
 * wcsph_sync grew the capacities (wcsph_get_variant shows the new ones) and
 * ran the lost steps again. */
enum {
    WCSPH_FLAG_FLUID_OVERFLOW = 1u << 0,
    WCSPH_FLAG_BOUNDARY_OVERFLOW = 1u << 1,
    WCSPH_FLAG_DOMAIN = 1u << 2,
    WCSPH_FLAG_NONFINITE = 1u << 3,
    WCSPH_FLAG_LISTS_GROWN = 1u << 4,
    /* the sort's cell counts did not add up to the particle count: a defect
     * of this library, reported as WCSPH_ERR_HIP (the cell ranges are kept
     * inside the particle arrays, so no access leaves them) */
    WCSPH_FLAG_INTERNAL = 1u << 5
};

/* Physics presets (wcsph_default_params). */
enum {
    /* The physics of the dam-break benchmarks: r = 0.01 m, cubic-spline
     * kernel with h = 4r for density and every gradient, Tait k = 1000 Pa,
     * gamma = 7, Monaghan artificial viscosity alpha = 0.01, no surface
     * tension, one time step for all accelerations. */
    WCSPH_PRESET_BENCHMARK = 0,
    /* The reference's compiled-in constants: r = 0.02 m, k = 0.1,
     * spiky pressure gradient with its h^-9 normalisation, Laplacian
     * viscosity nu = 0.001, surface tension 0.13 with normals, and the
     * separate non-pressure time step min(dt_hint, 0.015).
     * (src/cuda/tuning/kernels/common.cuh:18-23, compute_rho_p.cu:10-12,
     *  compute_non_pressure_accel.cu:14-16, sph.cuh:19-32, sph.cu:39-40) */
    WCSPH_PRESET_REFERENCE = 1
};

enum { WCSPH_GRAD_CUBIC = 0, WCSPH_GRAD_SPIKY_REF = 1 };
/* Neighbour search forms: lists built by one wavefront per occupied cell
 * (candidates loaded once per cell) or by one lane per particle walking its
 * 27 cells; no lists at all (CELL_LOOP): the density and force passes each
 * walk the cells, one workgroup per 4^3-cell brick with the brick's halo
 * staged in LDS; or brick lists (BRICK): the brick search records each
 * neighbour's LDS slot, and the density and force passes read their
 * neighbours' data from LDS.  All forms give identical results. */
enum { WCSPH_NEIGHBORS_CELL = 0, WCSPH_NEIGHBORS_PARTICLE = 1, WCSPH_NEIGHBORS_CELL_LOOP = 2,
       WCSPH_NEIGHBORS_BRICK = 3 };
enum { WCSPH_VISC_ARTIFICIAL = 0, WCSPH_VISC_LAPLACIAN = 1 };

/*
 * Position-dependent external acceleration, added to gravity in the
 * non-pressure force.  Replaces the EXTERNAL_FORCE lambda the reference
 * compiles into its kernel (compute_non_pressure_accel.cu:9-11, 109, 142;
 * set per scene through FluidSimulator::opts_t::external_force,
 * fluid_simulator.h:27, step_tuner.cuh:81-86).  The rules are tried in
 * order at the particle's position at the start of the step; the first
 * whose region holds it gives the acceleration, none gives zero.  The
 * reference's scene lambdas map onto them:
 *   fountain (app/fountain/application.cu:41-46):
 *     {XZ_DISC_Y_BELOW, r2 0.2, y 0: CONSTANT (0, 25, 0)},
 *     {Y_BELOW -0.3: TOWARD_Y_AXIS 2}
 *   vortex (app/vortex/application.cu:45): {Y_BELOW -0.5: SWIRL_Y_AXIS 2}
 *   surface_tension (app/surface_tension/application.cu):
 *     {ALL: CONSTANT (0, 9.81, 0)}
 * TOWARD_Y_AXIS is magnitude * normalize(-x, 0, -z), SWIRL_Y_AXIS is
 * magnitude * normalize(z, 0, -x); both are zero on the y axis itself.
 */
enum { WCSPH_REGION_ALL = 0, WCSPH_REGION_Y_BELOW = 1, WCSPH_REGION_XZ_DISC_Y_BELOW = 2 };
enum { WCSPH_FORCE_CONSTANT = 0, WCSPH_FORCE_TOWARD_Y_AXIS = 1, WCSPH_FORCE_SWIRL_Y_AXIS = 2 };
#define WCSPH_MAX_FORCE_RULES 4
typedef struct wcsph_force_rule {
    uint32_t region;     /* WCSPH_REGION_*                                   */
    uint32_t kind;       /* WCSPH_FORCE_*                                    */
    float y_below;       /* Y_BELOW, XZ_DISC_Y_BELOW: y < y_below            */
    float radius2;       /* XZ_DISC_Y_BELOW: x*x + z*z < radius2             */
    float value[3];      /* CONSTANT: the acceleration                       */
    float magnitude;     /* TOWARD_Y_AXIS, SWIRL_Y_AXIS: its length          */
} wcsph_force_rule;

/*
 * Physical and numerical parameters.  One runtime struct replaces the
 * constants the reference duplicates between host and kernels
 * (sph.cuh:19-32, fluid_simulator.h:11-12, common.cuh:18-23,
 *  compute_rho_p.cu:10-12, compute_non_pressure_accel.cu:14-16,
 *  update_positions.cu:10-12).
 * Derived: spacing s = 2r, support radius h = 2s, cell size = h,
 * particle mass m = rest_density * (s*s*s*0.8).
 */
typedef struct wcsph_params {
    float particle_radius;     /* r                                  */
    float rest_density;        /* rho0                               */
    float stiffness;           /* Tait k                             */
    float exponent;            /* Tait gamma                         */
    float max_density_ratio;   /* upper clamp of rho/rho0 in the EOS */
    float viscosity;           /* alpha (artificial) or nu (Laplacian) */
    float surface_tension;     /* gamma_st; 0 disables the normal pass */
    float gravity[3];
    float cfl_factor;          /* lambda of dt <= lambda*s/vmax      */
    float dt_min, dt_max;      /* clamp of dt_hint in adaptive mode  */
    float dt_np_max;           /* non-pressure step cap (reference preset) */
    float collision_buffer;    /* 0.75: wall test at x -/+ 0.75r      */
    float elasticity;          /* 0.5: v_n *= -elasticity on contact */
    uint32_t pressure_gradient;/* WCSPH_GRAD_*                       */
    uint32_t viscosity_model;  /* WCSPH_VISC_*                       */
    uint32_t adaptive_dt;      /* 1: CFL dt (sph.cu:115-130); 0: dt = dt_hint */
    uint32_t split_np_dt;      /* 1: v += dt_np*a_np + dt*a_p (update_positions.cu:64) */
    uint32_t n_force_rules;    /* external force: rules in use (0..WCSPH_MAX_FORCE_RULES) */
    wcsph_force_rule force_rules[WCSPH_MAX_FORCE_RULES];
} wcsph_params;

/* Box the fluid is kept in: [min, max] in the box frame; model maps the
 * box frame to world, model_inv back (column-major 4x4, like glm).
 * Replaces BoundingBox / BoundingBoxGPU (src/cuda/util.cuh:9-19,
 * src/render/box.cpp:25-29). */
typedef struct wcsph_box {
    float min[3];
    float max[3];
    float model[16];
    float model_inv[16];
} wcsph_box;

/* Scene handed to wcsph_create.  Replaces FluidSimulator::opts_t and the
 * lattice / boundary-particle set-up of its constructor
 * (src/simulation/fluid_simulator.h:22-28, fluid_simulator.cpp:10-18). */
typedef struct wcsph_scene {
    uint32_t n_fluid;
    uint32_t n_boundary;
    const float* fluid_pos;     /* 4*n_fluid                                */
    const float* fluid_vel;     /* 4*n_fluid, or NULL for zero velocities   */
    const float* boundary_pos;  /* 4*n_boundary (w ignored), may be NULL if 0 */
    wcsph_box box;
    /* Optional world-space region the cell grid must also cover (used when
     * wcsph_set_box will later move the box, e.g. a tilting box).  Ignored
     * unless domain_max > domain_min on every axis.  The grid always covers
     * the box, the boundary and the initial fluid. */
    float domain_min[3];
    float domain_max[3];
} wcsph_scene;

/* Kernel-variant choice (what the reference's KTT parameters selected,
 * src/cuda/tuning/step_tuner.cuh:80-121).  Default: block 256, capacities
 * 128 / 64, graph replay, the per-particle list builder and cell size h;
 * handles of at most 131,072 fluid particles start with the
 * cell-cooperative builder at n/8 workgroups (1024..65536) instead.
 *
 * cell_size_mult is the reference's CELL_SIZE_MULT (step_tuner.cuh:88,
 * particle_data.cuh:116, common.cuh:37): the fluid particles are sorted and
 * searched in cells of size fp32(h) * fp32(mult), walking R = ceil(h / cell)
 * cells per axis (R = 2 for 0.5 and 2/3, 1 for 1, 1.5 and 2).  Any value in
 * [0.5, 2] is accepted; 0 means 1.  It changes the particle order and the
 * order of each neighbour list (the reference's order for that multiplier)
 * and so the rounding of the sums, not the neighbour sets.  The boundary
 * stays in cells of size h.  A multiplier other than 1 needs the
 * per-particle builder; setting it rebuilds the fluid cell grid. */
typedef struct wcsph_variant {
    uint32_t block_size;         /* threads per workgroup of the particle passes: 64..1024 */
    uint32_t fluid_capacity;     /* neighbour-list slots per particle (fluid), multiple of 4    */
    uint32_t boundary_capacity;  /* neighbour-list slots per particle (boundary), multiple of 4 */
    uint32_t use_graph;          /* 1: replay the step as a hipGraph             */
    uint32_t neighbor_mode;      /* WCSPH_NEIGHBORS_*                             */
    uint32_t neighbor_groups;    /* workgroups of the cell-cooperative builder    */
    float cell_size_mult;        /* fluid cell size / h, in [0.5, 2] (0: 1)       */
} wcsph_variant;

/* Returned by wcsph_step / wcsph_sync; flags and maxima cover the steps
 * since the previous sync. */
typedef struct wcsph_stats {
    double sim_time;             /* sum of the dt actually used (all calls) */
    double last_dt;              /* dt of the last step                     */
    uint64_t steps;              /* steps taken over the handle's life      */
    uint32_t flags;              /* WCSPH_FLAG_*                             */
    uint32_t max_fluid_neighbors;
    uint32_t max_boundary_neighbors;
    uint32_t reserved;
} wcsph_stats;

typedef struct wcsph wcsph_t;

/* Fill params with a preset.  No reference counterpart as a function; the
 * values are the constants cited at WCSPH_PRESET_*. */
int wcsph_default_params(int preset, wcsph_params* out);

/* Identity box [min, max].  Replaces Box::Box + bounding_box()
 * (src/render/box.cpp:7-17). */
int wcsph_box_aabb(const float min[3], const float max[3], wcsph_box* out);

/*
 * Dam-break scene generator.  Replaces FluidSimulator::init_positions
 * (src/simulation/fluid_simulator.cpp:24-54: x-major lattice at spacing 2r,
 * w = 1) and the box's boundary particles (fluid_simulator.cpp:56-62),
 * adding the seeded jitter and explicit wall layers of the benchmarks.
 * Lattice site (ix,iy,iz) sits at lattice_min + s*(ix,iy,iz) plus a
 * uniform jitter in [-jitter, +jitter) per axis (counter-based hash of
 * seed and index).  Walls: `layers` shells of boundary particles at spacing s
 * outside the box, at distances r + s*l from each face.
 * Call with fluid_pos/boundary_pos NULL to get the counts first.
 * WCSPH_ERR_INVALID_ARGUMENT for non-finite inputs, more than 64 layers,
 * more than 2^21 sites along a box edge, or more than 2^32 - 1 particles.
 */
int wcsph_make_dam_break(float particle_radius,
                         const uint32_t dims[3], const float lattice_min[3],
                         float jitter, uint64_t seed,
                         const float box_min[3], const float box_max[3],
                         uint32_t layers,
                         float* fluid_pos, uint32_t* n_fluid,
                         float* boundary_pos, uint32_t* n_boundary);

/*
 * Fluid block with denser clusters (the obstacle-collapse workload).  No
 * reference counterpart: it extends init_positions
 * (fluid_simulator.cpp:24-54) with skewed neighbour counts.  Sites of the
 * main lattice sit at block_min + r + s*(i, j, k), s = 2r, x-major (x outer,
 * z inner) like the reference lattice, up to block_max - r.  Balls of
 * radius cluster_radius (centres drawn uniformly with a counter-based
 * generator of seed, at least cluster_radius + s inside the block, at
 * least 2*(cluster_radius + s) apart) replace the main sites closer than
 * cluster_radius + 0.9 s to their centre with a lattice of spacing
 * cluster_spacing centred on the ball; balls are added until the cluster
 * particles make up at least cluster_fraction of the total.  Every site
 * then gets a uniform jitter in [-jitter, +jitter) per axis.  Call with
 * fluid_pos NULL for the count.  cluster_radius / cluster_spacing is at
 * most 64.
 */
int wcsph_make_clustered_block(float particle_radius, const float block_min[3], const float block_max[3],
                               float cluster_fraction, float cluster_radius, float cluster_spacing,
                               float jitter, uint64_t seed, float* fluid_pos, uint32_t* n_fluid);

/*
 * Procedural triangle meshes, triangle soup of 9 floats per triangle,
 * transformed by `model` (column-major 4x4 like glm; NULL = identity) as
 * Geometry::load_triangles does (src/render/geometry.cpp:185-207).
 *   WCSPH_MESH_CUBE:  params = half size (x, y, z); 12 triangles, the
 *                     faces of procedural::cube (geometry.cpp:381-).
 *   WCSPH_MESH_SPHERE: params[0] = radius; `resolution` latitude bands and
 *                     2*resolution longitudes (UV sphere).
 *   WCSPH_MESH_TORUS: params[0] = major radius R, params[1] = tube radius;
 *                     2*resolution segments around the axis (y) and
 *                     `resolution` around the tube.
 * Call with vertices NULL for the triangle count; capacity is in triangles.
 * No reference counterpart for the sphere and torus.
 */
enum { WCSPH_MESH_CUBE = 0, WCSPH_MESH_SPHERE = 1, WCSPH_MESH_TORUS = 2 };
int wcsph_mesh_procedural(int kind, const float params[3], uint32_t resolution, const float model[16],
                          float* vertices, uint32_t capacity, uint32_t* n_triangles);

/*
 * Boundary particles on a triangle mesh: the centres of the cubic voxels of
 * edge voxel_size that intersect at least one triangle (separating-axis
 * triangle/box test, touching counts), on a voxel grid whose origin is the
 * mesh's minimum corner minus voxel_size/2.  Output is 4 floats per
 * particle (w = 1) in ascending voxel index order (x outer, z inner).
 * Replaces generate_boundary_particles (src/simulation/SPH/boundary.cpp:8-35),
 * which takes Open3D's VoxelGrid::CreateFromTriangleMesh at voxel size
 * 0.75 * 2r; the reference's order is that of Open3D's hash map and so is
 * unspecified.  Call with out_pos NULL for the count; a capacity below the
 * count returns WCSPH_ERR_INVALID_ARGUMENT with *n_out set to the count.
 * A voxel grid over the mesh of more than 2^31 voxels (2^21 along an axis)
 * is refused with WCSPH_ERR_INVALID_ARGUMENT.
 */
int wcsph_sample_mesh(const float* triangle_vertices, uint32_t n_triangles, float voxel_size,
                      float* out_pos, uint32_t capacity, uint32_t* n_out);

/*
 * Create a solver on the current HIP device.  Boundary particles are
 * Morton-sorted and their masses Psi computed once here.  Replaces the
 * CUDASPHSimulator constructor + init_positions/init_boundary
 * (src/cuda/SPH/sph.cu:9-29, 84-103, 66-71).
 */
int wcsph_create(const wcsph_params* params, const wcsph_scene* scene,
                 wcsph_t** out);

/*
 * Batched form: n_scenes independent scenes (each with its own box,
 * boundary particles, cell grid and CFL dt) stepped together, one launch
 * per kernel covering all scenes.  Fluid particles, boundary particles and
 * states of all scenes are concatenated scene-major in every array the
 * other entry points take or return (wcsph_scene_range gives a scene's
 * slice).  No reference counterpart: the reference steps one scene per
 * process (sph.cu:31-60).  wcsph_create is wcsph_create_batch with one scene.
 */
int wcsph_create_batch(const wcsph_params* params, const wcsph_scene* scenes, uint32_t n_scenes,
                       wcsph_t** out);
uint32_t wcsph_scene_count(const wcsph_t* h);
int wcsph_scene_range(const wcsph_t* h, uint32_t scene, uint32_t* first, uint32_t* count,
                      uint32_t* boundary_first, uint32_t* boundary_count);
/* Per-scene sim_time / last_dt / steps (flags and maxima are per handle). */
int wcsph_scene_stats(wcsph_t* h, uint32_t scene, wcsph_stats* stats);
int wcsph_scene_metrics(wcsph_t* h, uint32_t scene, double* mean_speed, double* kinetic_energy);

int wcsph_destroy(wcsph_t* h);

/* Last error text of h (or of the last failed create when h is NULL). */
const char* wcsph_last_error(const wcsph_t* h);

/* Replaces StepTuner's tuned parameters (step_tuner.cuh:80-121, 129-142).
 * Waits for the handle's stream.  Checks every field first
 * (WCSPH_ERR_INVALID_ARGUMENT, nothing changed); then reallocates the lists
 * for new capacities and rebuilds the fluid cell grid for a new
 * cell_size_mult.  If an allocation fails the handle keeps its grid and
 * variant, except that capacities whose lists were already reallocated
 * are kept.  A captured step graph is dropped. */
int wcsph_set_variant(wcsph_t* h, const wcsph_variant* v);
int wcsph_get_variant(const wcsph_t* h, wcsph_variant* v);

/*
 * Per-scene list builders of a batch: modes[s] (WCSPH_NEIGHBORS_CELL or
 * WCSPH_NEIGHBORS_PARTICLE, one per scene of the handle) picks the builder
 * of scene s, in place of the variant's neighbor_mode, for the list forms;
 * each step then runs each builder over its own scenes (the lists, and so
 * every result, are the same whichever builds them).  modes == NULL goes
 * back to the variant's builder for every scene.  A variant with the
 * cell-loop or brick-list form still runs that form for all scenes.  Every
 * scene must use the per-particle builder while cell_size_mult is not 1.
 * No reference counterpart: the reference steps one scene.  It is what the
 * size-bucket autotuner of wcsph/autotune.py applies.  Waits for the
 * handle's stream; drops a captured step graph.
 */
int wcsph_set_scene_neighbor_modes(wcsph_t* h, const uint32_t* modes);
/* modes[s] = the builder of scene s (the variant's neighbor_mode when no
 * per-scene builders are set). */
int wcsph_get_scene_neighbor_modes(const wcsph_t* h, uint32_t* modes);

/*
 * Advance n_steps steps.  Each step: dt from the CFL rule on the current
 * velocities (or dt_hint when adaptive_dt = 0), Morton sort of the fluid,
 * cell ranges, neighbour lists, density + Tait EOS, forces + symplectic
 * Euler + box collisions.  Replaces FluidSimulator::update(float) /
 * CUDASPHSimulator::update (src/simulation/fluid_simulator.h:33,
 * src/cuda/SPH/sph.cu:31-60).  Synchronises the handle's stream once at
 * the end to report stats (may be NULL).
 */
int wcsph_step(wcsph_t* h, float dt_hint, uint32_t n_steps, wcsph_stats* stats);

/* Same, without the final synchronisation or stats readback: the steps are
 * only enqueued on the handle's stream.  wcsph_sync waits for them and
 * reports the flags and neighbour maxima of every step since the previous
 * sync (wcsph_step is wcsph_step_async followed by wcsph_sync).
 *
 * Neighbour-list overflow.  The reference writes past its 96-entry lists
 * unchecked (compute_neighbor_list.cu:40-43, step_tuner.cuh:21).  Here a
 * step whose lists overflow stores no state, and every later step up to the
 * sync is skipped whole, so the state stays the one the failed step started
 * from.  wcsph_sync then grows the list capacities past the largest count
 * seen (with 1/8 headroom, up to 4096 per particle), restores the time
 * state and runs the lost steps again with the dt hints they were enqueued
 * with; the result equals a run started with the larger capacities.  The
 * stats then carry WCSPH_FLAG_LISTS_GROWN; WCSPH_ERR_NEIGHBOR_OVERFLOW is
 * returned only when the lists cannot grow far enough (the state then
 * stays at the start of the failed step). */
int wcsph_step_async(wcsph_t* h, float dt_hint, uint32_t n_steps);
int wcsph_sync(wcsph_t* h, wcsph_stats* stats);

uint32_t wcsph_fluid_count(const wcsph_t* h);    /* get_fluid_particles (fluid_simulator.h:39) */
uint32_t wcsph_boundary_count(const wcsph_t* h);

/*
 * Copy the fluid state in the handle's current (last sorted) order.  Any
 * pointer may be NULL.  ids[k] is the creation index of the particle now
 * at k.  rho/p are the values of the last density pass.
 * Replaces get_position_data (fluid_simulator.h:35) and the friend access
 * of Snapshot::save to ParticleData (snapshot.cu:10-35).
 */
int wcsph_get_state(wcsph_t* h, float* pos, float* vel, float* rho, float* p,
                    uint32_t* ids);

/* Overwrite fluid positions/velocities (4*n_fluid each, vel may be NULL for
 * zero); ids restart at 0..n-1.  Replaces Snapshot::load's writes
 * (snapshot.cu:81-106) and FluidSimulator::reset (fluid_simulator.cpp:64-66). */
int wcsph_set_state(wcsph_t* h, const float* pos, const float* vel);

/*
 * Replace the box of one scene (bounds and/or model) for the following
 * steps.  Replaces Box::set_model (src/render/box.cpp:25-29), which the
 * tilting-box scene calls every frame (app/tilting_box/application.cu,
 * TiltingBoxApp::update_objects).  The world-space AABB of the new box must
 * lie inside the scene's cell grid (size the grid with
 * wcsph_scene.domain_min/domain_max at creation), else WCSPH_ERR_DOMAIN and
 * the box is unchanged.  Boundary particles stay where they are.
 * Synchronises the handle's stream.
 */
int wcsph_set_box(wcsph_t* h, uint32_t scene, const wcsph_box* box);
int wcsph_get_box(const wcsph_t* h, uint32_t scene, wcsph_box* box);

/* Parity hooks.  perm[k] = index, in the order before the last step's sort,
 * of the particle placed at k (ParticleData::indices after sort_by_key,
 * particle_data.cuh:113-133). */
int wcsph_get_permutation(wcsph_t* h, uint32_t* perm);

/* Positions of the last step in sorted order, before integration (what the
 * neighbour search saw; positions_dst after ParticleData::sort). */
int wcsph_get_sorted_positions(wcsph_t* h, float* pos);

/*
 * Neighbour lists of the last step in CSR form, in the reference's
 * enumeration order (cells dx, dy, dz from -1 to 1, then ascending index;
 * common.cuh:126-151; compute_neighbor_list.cu:10-44).  counts[n_fluid];
 * list receives sum(counts) entries if list_len is large enough
 * (list may be NULL to get counts only; *total gets the sum).
 * which = 0: fluid neighbours (sorted fluid indices);
 * which = 1: boundary neighbours (sorted boundary indices), the set the
 * reference visits on the fly (common.cuh:153-178).
 */
int wcsph_get_neighbors(wcsph_t* h, int which, uint32_t* counts,
                        uint32_t* list, uint64_t list_len, uint64_t* total);

/* Sorted boundary positions with w = Psi (boundary mass), 4*n_boundary.
 * Replaces reading positions[fluid_n..] + boundary_mass
 * (sph.cu:66-71, 84-96; compute_boundary_mass.cu:10-30). */
int wcsph_get_boundary(wcsph_t* h, float* pos_psi);

/* FP64 diagnostics over the fluid of scene 0: mean |v| and kinetic energy
 * sum(0.5 m |v|^2).  Replaces compute_state_metrics (sph.cu:180-197). */
int wcsph_state_metrics(wcsph_t* h, double* mean_speed, double* kinetic_energy);

/*
 * Per-kernel timing with HIP events on the handle's stream.  While enabled,
 * steps are launched eagerly with an event between consecutive kernels;
 * wcsph_sync folds the elapsed times into per-kernel totals.  No reference
 * counterpart as a function: it replaces the per-kernel Duration KTT writes
 * into its result JSON (src/cuda/tuning/tuner.cpp:21-29).
 */
enum {
    WCSPH_K_TIME_STEP = 0,  /* a launch of its own only in steps without fluid; else inside WCSPH_K_SCAN */
    WCSPH_K_CELL_COUNT = 1,
    WCSPH_K_SCAN = 2,
    WCSPH_K_SCATTER = 3,
    WCSPH_K_PERMUTE = 4,
    WCSPH_K_NEIGHBORS = 5,
    WCSPH_K_DENSITY = 6,
    WCSPH_K_NORMALS = 7,
    WCSPH_K_FORCE = 8,
    WCSPH_NUM_KERNELS = 9
};
int wcsph_set_profiling(wcsph_t* h, int enable);
/* ms_total[WCSPH_NUM_KERNELS], launches[WCSPH_NUM_KERNELS]; reset != 0 zeroes
 * the totals after reading. */
int wcsph_kernel_times(wcsph_t* h, double* ms_total, uint64_t* launches, int reset);
const char* wcsph_kernel_name(int id);

/*
 * Differentiable runs (hand-written adjoint of the step).  No reference
 * counterpart.  Covers the benchmark physics (WCSPH_PRESET_BENCHMARK) with a
 * fixed dt (adaptive_dt = 0), one scene, an identity box.  The neighbour
 * sets, the sort and the collision decisions of each step are held fixed.
 *
 * wcsph_grad_forward restarts from the creation positions with velocities
 * vel0 (4 per particle, creation order; NULL = zero), runs n_steps steps of
 * dt, keeps a checkpoint every `checkpoint_every` steps, and writes the
 * final positions in creation order (4 per particle).
 * wcsph_grad_backward takes dL/d(final positions) (creation order) and
 * returns dL/d(vel0) and dL/d(viscosity alpha); it recomputes each segment
 * from its checkpoint with a tape and runs the adjoint kernels backwards.
 * Pointers may be host or device memory.  Afterwards the handle's state is
 * that of an intermediate step: set it again before stepping on.
 * wcsph_set_viscosity sets alpha (artificial viscosity) for later steps.
 */
int wcsph_set_viscosity(wcsph_t* h, float alpha);
int wcsph_grad_forward(wcsph_t* h, const float* vel0, float dt, uint32_t n_steps,
                       uint32_t checkpoint_every, float* pos_out);
int wcsph_grad_backward(wcsph_t* h, const float* grad_pos, float* grad_vel0, double* grad_alpha);
/* The backward pass uses the checkpoints, final particle order and physics
 * (alpha included) of the last successful wcsph_grad_forward, whatever was
 * stepped or set on the handle since; it can be run more than once.  Each
 * forward run gets a new generation number (0 = none yet): a caller that
 * keeps several rollouts checks it before a backward pass, since only the
 * last forward run's data is held. */
uint64_t wcsph_grad_generation(const wcsph_t* h);

/*
 * Project autotuner (replaces the KTT searcher and its per-frame tuning
 * budget: step_tuner.cuh:73-121, tuner.cpp:48-64, tuning_scheduler.h:12-26).
 * Times each candidate variant over steps_per_trial steps from the current
 * state (HIP events on the handle's stream, one untimed step first),
 * restores the state after every trial, applies the fastest variant and
 * returns it in *best.  candidates == NULL uses the default space: block
 * size {128, 256, 512} with the per-particle builder at the handle's
 * cell_size_mult; at cell_size_mult 1 also the cell builder with
 * {4096, 8192} workgroups, the cell-loop form and the brick-list form; and
 * the per-particle builder at block size 256 over the other cell sizes of
 * {1, 2/3, 1/2} x h (the reference's CELL_SIZE_MULT values,
 * step_tuner.cuh:88): 9 candidates from the default variant, at most 9 in
 * any case.  ms_per_step (may be NULL) receives one
 * time per candidate (INFINITY for a candidate that failed, e.g. by list
 * overflow).  Per-scene list builders (wcsph_set_scene_neighbor_modes)
 * stay in force during the trials: a candidate's neighbor_mode then only
 * chooses between the list forms and the cell-loop / brick-list forms.
 */
int wcsph_autotune(wcsph_t* h, float dt_hint, uint32_t steps_per_trial, const wcsph_variant* candidates,
                   uint32_t n_candidates, double* ms_per_step, wcsph_variant* best);

/*
 * wcsph_autotune without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory once set
 * up (serving loops, code under a no-hipMalloc rule) and still want to
 * retune a live handle, as the reference's tuning budget does between
 * frames.  Same arguments, candidates, timing and restored state as
 * wcsph_autotune, except:
 *   - the state saved across the trials (fluid positions, velocities and
 *     ids, and each scene's time state: 36 B per fluid particle plus 64 B
 *     per scene) is kept in a 64 MiB device array that the library holds
 *     on every GPU it runs on until the process ends; a larger handle
 *     (above about 1.86M fluid particles) gets WCSPH_ERR_INVALID_ARGUMENT;
 *   - every candidate must keep the handle's list capacities and cell-size
 *     multiplier (changing them reallocates the lists or the cell grid),
 *     else WCSPH_ERR_INVALID_ARGUMENT before any trial runs; the default
 *     space is the one above without the other cell sizes;
 *   - trials run as plain launches (no graph is captured), so the times
 *     include the launch gaps a graph replay hides, and profiling marks are
 *     off while they run;
 *   - steps still pending from wcsph_step_async complete first, and a list
 *     overflow among them is returned rather than recovered.
 * The array is shared by every handle on one GPU and keeps the last call's
 * contents: calls that overlap on one GPU (from several host threads) race
 * and must not be made.  A graph the handle had is dropped when the chosen
 * variant is applied; the next graph step captures a new one.
 */
int wcsph_autotune_noalloc(wcsph_t* h, float dt_hint, uint32_t steps_per_trial, const wcsph_variant* candidates,
                           uint32_t n_candidates, double* ms_per_step, wcsph_variant* best);

/*
 * .sphs snapshots (fluid positions + velocities, float4 each, after a
 * 76-byte header {"SPHS", u32 version 1, char app_name[64], u32 fluid_n}).
 * Host-only, no handle.  Replaces Snapshot::save / Snapshot::load
 * (src/cuda/SPH/snapshot.cuh:9-51, snapshot.cu:10-106) with the same layout,
 * checks (magic, version, app_name, fluid count) and messages;
 * wcsph_snapshot_error() returns the last message.  wcsph_snapshot_read
 * checks app_name unless it is NULL.  Pair with wcsph_get_state /
 * wcsph_set_state.
 */
int wcsph_snapshot_info(const char* path, char* app_name, size_t app_name_len, uint32_t* fluid_n);
int wcsph_snapshot_read(const char* path, const char* app_name, uint32_t fluid_n, float* pos, float* vel);
int wcsph_snapshot_write(const char* path, const char* app_name, uint32_t fluid_n, const float* pos,
                         const float* vel);
const char* wcsph_snapshot_error(void);

#ifdef __cplusplus
}
#endif
#endif /* WCSPH_H */
