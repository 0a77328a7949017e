// Device code of the WCSPH step for gfx950 (wave64).
//
// Data layout in HBM.  A handle holds S >= 1 independent scenes; fluid
// particles, boundary particles and grid cells of all scenes are
// concatenated scene-major, and a particle finds its scene by binary search
// over the scene offsets (pstart / bstart / cstart, S + 1 entries each).
// Per scene (N fluid, B boundary particles):
//   pos_a/vel_a/id_a  state between steps (float4, float4, u32) — the
//                     order of the last sort;
//   pos_b/vel_b/id_b  the same particles in this step's Morton order;
//   xp                (x, y, z, p/rho^2) per sorted particle, written by the
//                     density pass and gathered by the force pass;
//   vel_b.w           rho of the sorted particle (density pass);
//   bpos              boundary particles, Morton-sorted once, w = Psi;
//   cell grid         dense box of cells of size h over the scene; per cell
//                     a count and the ranges (start, end) of its fluid
//                     particles (rebuilt every step) and of its boundary
//                     particles (static), as uint2;
//   rank_to_lin       the grid's cells listed in reference Morton-key order;
//   nlist / blist     neighbour lists in slot quads: entry k of particle i
//                     at ((k / 4) * N + i) * 4 + k % 4, so one 16-byte load
//                     gives a particle four entries and a wavefront's 64
//                     quads of one slot group are one contiguous 1 KiB
//                     segment (list_at).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wcsph.h"

namespace wcsph_dev {

constexpr float kPi = 3.1415926535897932384626433f;

struct Grid {
    int c0x, c0y, c0z;  // cell coordinate of grid index (0, 0, 0)
    int nx, ny, nz;     // grid dims; lin = ((cx - c0x) * ny + (cy - c0y)) * nz + (cz - c0z)
    float cs;           // cell size (h times the cell-size multiplier; h for the boundary grid)
    int R;              // walk range: a neighbour lies within R cells per axis; interior cells are [R, n - 1 - R]
};

// Physics constants in the form the kernels use them.
struct Phys {
    float h, h2, inv_h, mass, w0;
    float cubic_factor;      // 8 / (pi h^3)
    float cubic_gfactor;     // 48 / (pi h^3)
    float spiky_gfactor;     // -45 / (pi h^9)
    float cohesion_factor;   // 32 / (pi h^3)
    float rho0, inv_rho0_sq, stiffness, exponent, max_ratio;
    float visc, visc_eps;    // alpha or nu; 0.01 h^2
    float art_visc_coef;     // 2 alpha h c_s (artificial viscosity)
    float surface_tension;
    float gx, gy, gz;
    float r, cbuf, elasticity;
    int grad_kind, visc_kind;
    int n_rules;                              // external force rules in use
    wcsph_force_rule rules[WCSPH_MAX_FORCE_RULES];
};

// External acceleration at (x, y, z): the first rule whose region holds the
// point (EXTERNAL_FORCE of compute_non_pressure_accel.cu:9-11, as rules).
__device__ inline void external_accel(const Phys& P, float x, float y, float z, float& ax, float& ay, float& az) {
#pragma clang fp contract(off)
    ax = ay = az = 0.f;
    for (int k = 0; k < P.n_rules; ++k) {
        const wcsph_force_rule& R = P.rules[k];
        const bool in = R.region == WCSPH_REGION_ALL ||
                        (y < R.y_below && (R.region == WCSPH_REGION_Y_BELOW || __builtin_fmaf(z, z, x * x) < R.radius2));
        if (!in) continue;
        if (R.kind == WCSPH_FORCE_CONSTANT) {
            ax = R.value[0];
            ay = R.value[1];
            az = R.value[2];
        } else {
            // magnitude * normalize(d) with d = (-x, 0, -z) or (z, 0, -x)
            const float dx = R.kind == WCSPH_FORCE_TOWARD_Y_AXIS ? -x : z;
            const float dz = -(R.kind == WCSPH_FORCE_TOWARD_Y_AXIS ? z : x);
            const float len = sqrtf(__builtin_fmaf(dz, dz, dx * dx));
            if (len > 0.f) {
                ax = R.magnitude * (dx / len);
                az = R.magnitude * (dz / len);
            }
        }
        return;
    }
}

// One scene: its fluid cell grid (cells c0 .. c0 + nx*ny*nz - 1 of the
// global fluid-cell arrays), its boundary cell grid (cell size h, like the
// reference's boundary search, common.cuh:81-87, 153-178; cells bc0 ... of
// the boundary-cell arrays; the same grid as the fluid one when the
// cell-size multiplier is 1) and its box.
struct SceneDev {
    Grid g;
    uint32_t c0;
    Grid bg;
    uint32_t bc0;
    float bmin[3], bmax[3];
    float model[16], model_inv[16];
    int box_identity;
    uint32_t form;  // the scene's list builder when the handle sets one per scene (WCSPH_NEIGHBORS_*)
};

// The list builders' scene filter: every scene, or only the scenes whose
// `form` is the given builder (a batch whose scenes use different builders
// runs each builder over its own scenes).
constexpr uint32_t kAllForms = 0xffffffffu;

// Per-scene time stepping state.  The saved_* fields hold the values
// before the most recent time step, so that a step lost to a neighbour-list
// overflow can be run again from exactly the state it started from.
struct SceneState {
    unsigned long long vmax2_bits;  // max |v|^2 of the last integration (FP64 bits)
    double sim_time;
    double last_dt;
    unsigned long long steps;
    float dt;      // dt of the current step
    float dt_np;   // non-pressure dt of the current step
    unsigned long long saved_vmax2_bits;
    double saved_sim_time;
    double saved_last_dt;
    unsigned long long saved_steps;
};

// Per-handle device scalars.
struct DevStatus {
    uint32_t flags;
    uint32_t max_fluid;
    uint32_t max_boundary;
    uint32_t n_occ;   // occupied cells of this step (written by the scan)
    uint32_t frozen;  // steps skipped whole after a neighbour-list overflow
};

// Flag bits of a neighbour-list overflow (WCSPH_FLAG_FLUID_OVERFLOW |
// WCSPH_FLAG_BOUNDARY_OVERFLOW).  Once one is set, the state stops
// advancing: force passes store nothing and time steps are skipped, until
// the host grows the lists and runs the lost steps again.
constexpr uint32_t kOverflowFlags = 3u;

struct DtRule {
    float hint, dt_min, dt_max, dt_np_max, cfl_factor, spacing;
    int adaptive, split;
};

// ------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------
__device__ __forceinline__ float dist2(float rx, float ry, float rz) {
    // dot(r, r) contracted as fma(z, z, fma(y, y, x*x)) — the form the
    // neighbour predicate is pinned to (common.cuh:39-42, math.cuh:54-56).
    return __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remapping the block index so that XCD x runs one contiguous range of
// blocks keeps Morton-neighbouring particles (and their neighbour data) in
// one L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t x = b & 7u, k = b >> 3;
    const uint32_t q = g >> 3, r = g & 7u;
    return (x < r) ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
}

__device__ __forceinline__ uint32_t particle_index() { return xcd_block() * blockDim.x + threadIdx.x; }

// The scene s with start[s] <= i < start[s + 1] (start has S + 1 entries).
__device__ __forceinline__ uint32_t find_scene(const uint32_t* __restrict__ start, uint32_t S, uint32_t i) {
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// 1-ulp hardware square root and reciprocal for the physics terms (the
// neighbour predicate and the cell division stay exact).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ int cell_coord(float p, float cs) {
    return (int)floorf(p / cs);  // IEEE-rounded division (hipcc default)
}

// a + b never fused with the product that forms b: the density sums of the
// list and cell-loop forms round identically whatever the surrounding code.
__device__ __forceinline__ float add_unfused(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}

// The physics helpers below fix every fused multiply-add in the source
// (contraction off, explicit fmaf), so that each kernel that inlines them
// rounds identically.
__device__ __forceinline__ float cubic_w(float q, const Phys& P) {
#pragma clang fp contract(off)
    // both pieces, then a select: no divergent branches in the neighbour loops
    const float q2 = q * q;
    const float inner = P.cubic_factor * __builtin_fmaf(6.f * q2, q, __builtin_fmaf(-6.f, q2, 1.f));
    const float a = 1.f - q;
    const float outer = P.cubic_factor * 2.f * (a * a * a);
    return q <= 0.5f ? inner : (q <= 1.f ? outer : 0.f);
}

// grad W = coef * r for the cubic spline (kernel.cuh:22-32).
__device__ __forceinline__ float cubic_grad(float q, const Phys& P) {
#pragma clang fp contract(off)
    const float rl = q * P.h;
    const float inv = frcp(P.h * rl);
    const float inner = P.cubic_gfactor * q * __builtin_fmaf(3.f, q, -2.f) * inv;
    const float a = 1.f - q;
    const float outer = -P.cubic_gfactor * a * a * inv;
    // selects, not branches (the pieces are finite or unused)
    return (q < 1e-9f || q > 1.f) ? 0.f : (q <= 0.5f ? inner : outer);
}

// Reference spiky gradient with its h^-9 factor (kernel.cuh:43-51).
__device__ __forceinline__ float spiky_ref_grad(float q, const Phys& P) {
#pragma clang fp contract(off)
    const float rl = q * P.h;
    const float a = 1.f - q;
    const float g = P.spiky_gfactor * a * a * frcp(P.h * rl);
    return (q < 1e-9f || q > 1.f) ? 0.f : g;
}

__device__ __forceinline__ float cohesion_w(float q, const Phys& P) {
#pragma clang fp contract(off)
    const float a = 1.f - q;
    const float iq3 = a * a * a;
    const float q3 = q * q * q;
    const float inner = P.cohesion_factor * __builtin_fmaf(2.f * q3, iq3, -1.f / 64.f);
    const float outer = P.cohesion_factor * q3 * iq3;
    return q <= 0.5f ? inner : (q <= 1.f ? outer : 0.f);
}

// One fluid neighbour's term of the surface normal (compute_pressure_accel_n_normal.cu:98-124).
__device__ __forceinline__ void add_normal(float& nx, float& ny, float& nz, const Phys& P, const float4& xi,
                                           const float4& xj, float dj) {
#pragma clang fp contract(off)
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float q = fsqrt(dist2(rx, ry, rz)) * P.inv_h;
    const float c = cubic_grad(q, P) * frcp(dj);
    nx = __builtin_fmaf(rx, c, nx);
    ny = __builtin_fmaf(ry, c, ny);
    nz = __builtin_fmaf(rz, c, nz);
}

// atomicMax that skips the atomic when the current value already reaches v.
// Per-wave maxima target one address; a maximum only grows while a kernel
// runs, so a stale read can cause an unneeded atomic but never skip a
// needed one, and the serialised same-address atomics drop to the few that
// raise the maximum.
__device__ __forceinline__ void atomic_max_if_above(uint32_t* p, uint32_t v) {
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) atomicMax(p, v);
}
__device__ __forceinline__ void atomic_max_if_above(unsigned long long* p, unsigned long long v) {
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) atomicMax(p, v);
}

// Sum of v over the workgroup, returned to thread 0 (every thread calls it).
// Reductions to one value per workgroup are written to a slot of their own
// and summed on the host in a fixed order: no same-address atomics, and the
// result does not depend on the order the workgroups ran in.
__device__ __forceinline__ double block_sum_f64(double v) {
    __shared__ double wsum[16];
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const uint32_t wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    if ((threadIdx.x & 63u) == 0) wsum[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < nw; ++w) s += wsum[w];
    __syncthreads();  // wsum may be reused by the next call
    return s;
}

// Maximum / minimum over the 64 lanes of a wavefront (every lane active):
// four DPP exchanges reduce each row of 16 lanes (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), then the four row
// results are read into scalars.  No LDS traffic, a wave-uniform result.
template <bool kMax>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v) {
    auto op = [](uint32_t a, uint32_t b) { return kMax ? max(a, b) : min(a, b); };
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xb1, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4e, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xf, 0xf, false));
    const uint32_t a = op((uint32_t)__builtin_amdgcn_readlane((int)v, 0), (uint32_t)__builtin_amdgcn_readlane((int)v, 16));
    const uint32_t b = op((uint32_t)__builtin_amdgcn_readlane((int)v, 32), (uint32_t)__builtin_amdgcn_readlane((int)v, 48));
    return op(a, b);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return wave_reduce_u32<true>(v); }
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) { return wave_reduce_u32<false>(v); }

__device__ __forceinline__ double wave_max_f64(double v) {
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over a 256-thread block (4 wavefronts); returns the
// thread's exclusive prefix and writes the block total.
__device__ __forceinline__ uint32_t block256_excl_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[4];
    const uint32_t inc = wave_incl_scan(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + inc - v;
}

// ------------------------------------------------------------------
// 0. time step (sph.cu:115-130, FP64) — one thread
// ------------------------------------------------------------------
// dt of one scene from the previous step's max |v|^2 (sph.cu:115-130).
// `frozen`: an earlier step overflowed its neighbour lists; the scene's time
// state stays as it is.
__device__ __forceinline__ void time_step_scene(SceneState* __restrict__ st, const DtRule& rule, bool frozen) {
    if (frozen) return;
    st->saved_vmax2_bits = st->vmax2_bits;
    st->saved_sim_time = st->sim_time;
    st->saved_last_dt = st->last_dt;
    st->saved_steps = st->steps;
    const double vmax2 = __longlong_as_double((long long)st->vmax2_bits);
    st->vmax2_bits = 0ull;
    double dt = rule.hint;
    if (rule.adaptive) {
        const double vmax = sqrt(vmax2);
        if (vmax < 1e-9) {
            dt = rule.dt_max;
        } else {
            const double cfl = (double)rule.cfl_factor * (double)rule.spacing / vmax;
            double c = rule.hint;
            c = c < rule.dt_min ? rule.dt_min : c;
            c = c > rule.dt_max ? rule.dt_max : c;
            dt = c < cfl ? c : cfl;
        }
    }
    const float dtf = (float)dt;
    float dtnp = dtf;
    if (rule.split) dtnp = rule.hint < rule.dt_np_max ? rule.hint : rule.dt_np_max;
    st->dt = dtf;
    st->dt_np = dtnp;
    st->sim_time += (double)dtf;
    st->last_dt = (double)dtf;
    st->steps += 1ull;
}

// The time step alone (a step without fluid particles runs no scan).
__global__ void k_time_step(SceneState* __restrict__ ss, uint32_t S, DtRule rule) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) time_step_scene(ss + s, rule, false);
}

__global__ void k_clear_flags(DevStatus* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->flags = 0;
        st->max_fluid = 0;
        st->max_boundary = 0;
        st->frozen = 0;
    }
}

// ------------------------------------------------------------------
// 1. cell of every particle + per-cell count (slot = arrival order)
// ------------------------------------------------------------------
// Outputs of the counting (the next sort's inputs); cnt == nullptr: the
// kernel that takes it counts nothing.
struct CellCount {
    const uint32_t* lin_to_rank;
    uint32_t* cnt;
    uint32_t* cell;
    uint32_t* slot;
};

// Cell of particle i (position p, scene sd) in its fluid grid, counted for
// the sort.  Every lane of the wavefront calls it (lanes without a particle
// with valid = false).
__device__ __forceinline__ void count_cell(const float4& p, bool valid, uint32_t i, const SceneDev& sd,
                                           const CellCount& C, DevStatus* st) {
    uint32_t lin = 0xffffffffu;
    if (valid) {
        const Grid g = sd.g;
        int cx = cell_coord(p.x, g.cs) - g.c0x;
        int cy = cell_coord(p.y, g.cs) - g.c0y;
        int cz = cell_coord(p.z, g.cs) - g.c0z;
        // Interior cells only: the neighbour walks then never leave the grid.
        const int R = g.R;
        const bool ok = cx >= R && cx <= g.nx - 1 - R && cy >= R && cy <= g.ny - 1 - R && cz >= R && cz <= g.nz - 1 - R;
        if (!ok) {
            atomicOr(&st->flags, 4u /* WCSPH_FLAG_DOMAIN */);
            cx = min(max(cx, R), g.nx - 1 - R);
            cy = min(max(cy, R), g.ny - 1 - R);
            cz = min(max(cz, R), g.nz - 1 - R);
        }
        lin = sd.c0 + ((uint32_t)cx * g.ny + cy) * g.nz + cz;
    }
    // The particles are in the last sort's cell order, so runs of
    // consecutive lanes share a cell: one atomic per run reserves the run's
    // slots (any distinct slots in [0, count) will do: the permutation orders
    // a cell's members by previous index).
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = (uint32_t)__shfl_up((int)lin, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || prev != lin || !valid);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t head = 63u - (uint32_t)__clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
    uint32_t base = 0;
    // counts are kept in Morton-rank order, so the scan streams them
    if (valid && head == lane) base = atomicAdd(&C.cnt[C.lin_to_rank[lin]], next - lane);
    base = (uint32_t)__shfl((int)base, (int)head, 64);
    if (valid) {
        C.cell[i] = lin;
        C.slot[i] = base + (lane - head);
    }
}

// The counting on its own: the first step after the positions were set
// from the host (or restored), and every step of the cell-loop and
// brick-list forms.  Otherwise the force pass of the previous step counted
// the positions it wrote.
__global__ void k_cell_count(const float4* __restrict__ pos, uint32_t n, const SceneDev* __restrict__ sc,
                             const uint32_t* __restrict__ pstart, uint32_t S, CellCount C, DevStatus* st) {
    const uint32_t i = particle_index();
    const bool valid = i < n;
    const uint32_t sid = find_scene(pstart, S, valid ? i : n - 1u);
    count_cell(valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f), valid, i, sc[sid], C, st);
}

// ------------------------------------------------------------------
// 2. exclusive scan of the counts in Morton-rank order -> start/end
// ------------------------------------------------------------------
constexpr int kScanItems = 4;
constexpr int kScanTile = 256 * kScanItems;

// partial[2b] = particles in tile b, partial[2b+1] = occupied cells in tile b
// One block of 256 threads: exclusive scans of both partial sequences in
// place; the number of occupied cells goes to st->n_occ.
__device__ __forceinline__ void scan_partials(uint32_t* partial, uint32_t nparts, uint32_t n, DevStatus* st) {
    const uint32_t per = (nparts + 255) / 256;
    const uint32_t b = threadIdx.x * per;
    uint32_t s = 0, o = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nparts) {
            s += partial[2 * (b + k)];
            o += partial[2 * (b + k) + 1];
        }
    uint32_t ts, to;
    uint32_t offs = block256_excl_scan(s, &ts);
    uint32_t offo = block256_excl_scan(o, &to);
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nparts) {
            const uint32_t vs = partial[2 * (b + k)], vo = partial[2 * (b + k) + 1];
            partial[2 * (b + k)] = offs;
            partial[2 * (b + k) + 1] = offo;
            offs += vs;
            offo += vo;
        }
    if (threadIdx.x == 0) {
        st->n_occ = to;
        if (ts != n) atomicOr(&st->flags, 32u /* WCSPH_FLAG_INTERNAL */);
    }
}

// First pass: per-tile totals of the counts and of the occupied cells.  It
// also empties the ranges of the previous step's occupied cells (every other
// cell's range is already empty), so that k_scan_apply writes only the
// occupied ones; st->n_occ still holds the previous count here.
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ cnt, uint32_t ncell,
                                                     uint32_t* __restrict__ partial, uint32_t nparts,
                                                     const uint32_t* __restrict__ occ, uint2* __restrict__ crange,
                                                     const DevStatus* __restrict__ st) {
    {
        const uint32_t nprev = st->n_occ;
        const uint32_t lo = (uint32_t)(((unsigned long long)nprev * blockIdx.x) / nparts);
        const uint32_t hi = (uint32_t)(((unsigned long long)nprev * (blockIdx.x + 1)) / nparts);
        for (uint32_t c = lo + threadIdx.x; c < hi; c += 256) crange[occ[c]] = make_uint2(0u, 0u);
    }
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0, o = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k * 256 + threadIdx.x;
        const uint32_t c = r < ncell ? cnt[r] : 0u;
        s += c;
        o += c ? 1u : 0u;
    }
    uint32_t ts, to;
    block256_excl_scan(s, &ts);
    block256_excl_scan(o, &to);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ts;
        partial[2 * blockIdx.x + 1] = to;
    }
}

// One block: scans the tile totals (and sets st->n_occ), then every scene's
// time step.  A kernel of its own: the tiles' totals need no device-scope
// fences or counters, which cost one serialised atomic per tile on grids of
// millions of cells.
__global__ __launch_bounds__(256) void k_scan_partials(uint32_t* partial, uint32_t nparts, uint32_t n, DevStatus* st,
                                                       SceneState* __restrict__ ss, uint32_t S, DtRule rule) {
    const bool frozen = (st->flags & kOverflowFlags) != 0u;  // read before scan_partials' barriers
    scan_partials(partial, nparts, n, st);
    for (uint32_t k = threadIdx.x; k < S; k += 256) time_step_scene(ss + k, rule, frozen);
    if (frozen && threadIdx.x == 0) st->frozen += 1u;
}

// Each thread owns kScanItems consecutive ranks of the tile.  Writes the
// cell ranges and the list of occupied cells in Morton order.
// kInline (grids of at most kInlineScanTiles tiles): the tile totals are
// still raw, and each workgroup sums those of the tiles before its own (a
// few KiB from L2) in place of a k_scan_partials launch; workgroup 0 then
// also takes the time step, and the last one the occupied count.
constexpr uint32_t kInlineScanTiles = 1024;

template <bool kInline>
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* __restrict__ cnt,
                                                    const uint32_t* __restrict__ rank_to_lin,
                                                    uint32_t ncell, uint32_t n, const uint32_t* __restrict__ partial,
                                                    uint2* __restrict__ crange, uint32_t* __restrict__ occ,
                                                    DevStatus* st, SceneState* __restrict__ ss, uint32_t S,
                                                    DtRule rule) {
    uint32_t pre_s, pre_o;  // particles and occupied cells in the tiles before this one
    if (kInline) {
        const bool frozen = (st->flags & kOverflowFlags) != 0u;  // read before the barriers below
        uint32_t s = 0, o = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256) {
            s += partial[2 * b];
            o += partial[2 * b + 1];
        }
        block256_excl_scan(s, &pre_s);
        block256_excl_scan(o, &pre_o);
        if (blockIdx.x == 0) {
            for (uint32_t k = threadIdx.x; k < S; k += 256) time_step_scene(ss + k, rule, frozen);
            if (frozen && threadIdx.x == 0) st->frozen += 1u;
        }
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
            const uint32_t tot_s = pre_s + partial[2 * blockIdx.x], tot_o = pre_o + partial[2 * blockIdx.x + 1];
            st->n_occ = tot_o;
            if (tot_s != n) atomicOr(&st->flags, 32u /* WCSPH_FLAG_INTERNAL */);
        }
    } else {
        pre_s = partial[2 * blockIdx.x];
        pre_o = partial[2 * blockIdx.x + 1];
    }
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t lin[kScanItems];
    uint32_t s = 0, o = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k;
        c[k] = r < ncell ? cnt[r] : 0u;
        lin[k] = c[k] ? rank_to_lin[r] : 0u;  // only occupied cells are written
        s += c[k];
        o += c[k] ? 1u : 0u;
    }
    uint32_t ts, to;
    uint32_t off = block256_excl_scan(s, &ts) + pre_s;
    uint32_t ooff = block256_excl_scan(o, &to) + pre_o;
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k;
        if (r < ncell && c[k]) {
            // a count defect (flagged by k_scan_partials) stays inside the arrays
            crange[lin[k]] = make_uint2(min(off, n), min(off + c[k], n));
            off += c[k];
            occ[ooff++] = lin[k];
            cnt[r] = 0u;  // ready for the next step's counting
        }
    }
}

// ------------------------------------------------------------------
// 3. scatter particle ids into their cell's range (arrival order)
// ------------------------------------------------------------------
__global__ void k_scatter(const uint32_t* __restrict__ cell, const uint32_t* __restrict__ slot,
                          const uint2* __restrict__ crange, uint32_t n, uint32_t* __restrict__ tmp) {
    const uint32_t i = particle_index();
    if (i >= n) return;
    const uint32_t f = crange[cell[i]].x + slot[i];
    if (f < n) tmp[f] = i;  // always, unless the counts are defective (flagged by the scan)
}

// ------------------------------------------------------------------
// 4. stable placement: rank of i among the cell's members by previous
//    index, then move x, v, id to the sorted buffers
// ------------------------------------------------------------------
// Positions as three arrays of floats (sorted fluid, boundary), beside the
// float4 copies: the neighbour search reads four consecutive candidates'
// x (y, z) with one 16-byte load, three loads for four candidates.
struct Coords {
    float *x, *y, *z;
};

__global__ void k_permute(const uint32_t* __restrict__ cell, const uint2* __restrict__ crange,
                          const uint32_t* __restrict__ tmp, uint32_t n,
                          const float4* __restrict__ pos_a, const float4* __restrict__ vel_a,
                          const uint32_t* __restrict__ id_a, float4* __restrict__ pos_b,
                          float4* __restrict__ vel_b, uint32_t* __restrict__ id_b,
                          uint32_t* __restrict__ perm, Coords xyz_b) {
    const uint32_t i = particle_index();
    if (i >= n) return;
    // the previous state is read once: streamed past L2 (the cached lines
    // are left to the sorted arrays the search gathers from)
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const f32x4 pv = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(pos_a) + i);
    const f32x4 vv = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(vel_a) + i);
    const float4 p = make_float4(pv.x, pv.y, pv.z, pv.w), v = make_float4(vv.x, vv.y, vv.z, vv.w);
    const uint32_t id = __builtin_nontemporal_load(id_a + i);
    const uint2 cr = crange[cell[i]];
    const uint32_t s = cr.x, e = cr.y;
    // the cell's members four at a time (tmp carries 3 padding entries)
    uint32_t r = 0;
    for (uint32_t k = s; k < e; k += 4) {
        uint4 t;
        __builtin_memcpy(&t, reinterpret_cast<const char*>(tmp) + (size_t)k * 4u, sizeof(t));
        const uint32_t lim = e - k;
        r += (t.x < i ? 1u : 0u) + (lim > 1u && t.y < i ? 1u : 0u) + (lim > 2u && t.z < i ? 1u : 0u) +
             (lim > 3u && t.w < i ? 1u : 0u);
    }
    const uint32_t f = s + r;
    pos_b[f] = p;
    xyz_b.x[f] = p.x;
    xyz_b.y[f] = p.y;
    xyz_b.z[f] = p.z;
    vel_b[f] = v;
    id_b[f] = id;
    perm[f] = i;
}

// ------------------------------------------------------------------
// 5. neighbour lists (fluid + boundary), reference enumeration order.
//    Lists are stored in slot quads (list_at): consumers read a particle's
//    four next entries with one 16-byte load, and 64 consecutive particles'
//    quads of one slot group are one 1 KiB segment.  Scattered 4-byte stores
//    would leave partially written lines for L2 to evict, so one wavefront
//    per workgroup stages its 64 lists in a circular LDS window of kWinFluid
//    (kWinBoundary) slot rows, [slot % window][lane].  After each row of
//    three cells the quads every lane has filled are written to HBM, one
//    16-byte store per lane, a whole segment per quad.
//    A lane whose next entry would overrun the window (far ahead of the
//    slowest lane) stores that entry and all its later ones to HBM directly
//    (its spill point `sf`/`sb` is sticky), and flushes copy only the slots
//    below a lane's spill point, so an LDS row is never read after it was
//    reused.
//    Candidates are loaded eight at a time (a cell of the benchmark lattice,
//    and of its three-layer walls, holds eight: one batch of six 16-byte
//    loads per cell).
// ------------------------------------------------------------------
constexpr uint32_t kWinFluid = 32;     // powers of two
constexpr int kDensityBatch = 8;        // neighbours whose loads a density pass issues together
constexpr uint32_t kWinBoundary = 8;
constexpr uint32_t kNeighborBlock = 64;
constexpr int kFluidGroup = 8;     // fluid candidates tested together (a cell holds ~8)
constexpr int kBoundaryGroup = 8;  // boundary candidates tested together (wall cells hold ~8)
// padding entries past the end of the coordinate arrays (a group's loads
// may reach G - 1 entries past its range)
constexpr uint32_t kFluidCoordPad = kFluidGroup;
constexpr uint32_t kBoundaryCoordPad = kBoundaryGroup;

// Element idx of an array through a 32-bit byte offset, so that the load
// uses the scalar-base + 32-bit vector-offset addressing form.  Creation
// limits particle arrays to 2^27 entries (float4: idx * 16 < 2^31) and the
// cell grids of a handle to kMaxCells cells (cell ranges are uint2:
// idx * 8 < 2^31), which keeps every offset formed here in range.
constexpr uint64_t kMaxCells = 1ull << 28;
template <class T>
__device__ __forceinline__ T ld(const T* __restrict__ base, uint32_t idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (uint32_t)sizeof(T));
}

// Element of entry k of particle i in a list of n particles (slot quads).
// Capacities are multiples of 4, so a list of capacity cap spans cap * n
// elements.
__device__ __forceinline__ size_t list_at(uint32_t k, uint32_t i, uint32_t n) {
    return (((size_t)(k >> 2) * n + i) << 2) | (k & 3u);
}

// Writes the quads [q0, q1) of this lane's list from a circular LDS window,
// for the entries held there (k < held).  A quad below `held`, or one that
// reaches past the lane's count (its tail is never read), is one 16-byte
// store; a quad that the spill point cuts is written entry by entry, since
// its entries past the spill point were stored to HBM directly.
template <uint32_t kWin>
__device__ __forceinline__ void flush_quads(const uint32_t (*win)[kNeighborBlock], uint32_t q0, uint32_t q1,
                                            uint32_t held, uint32_t count, uint32_t lane, uint32_t i, uint32_t n,
                                            bool valid, uint32_t* __restrict__ list) {
    for (uint32_t q = q0; q < q1; ++q) {
        const uint32_t k = 4u * q;
        if (!valid || k >= held) continue;
        if (k + 4u <= held || held >= count) {
            const uint4 v = make_uint4(win[k & (kWin - 1)][lane], win[(k + 1u) & (kWin - 1)][lane],
                                       win[(k + 2u) & (kWin - 1)][lane], win[(k + 3u) & (kWin - 1)][lane]);
            *reinterpret_cast<uint4*>(list + list_at(k, i, n)) = v;
        } else {
            for (uint32_t u = k; u < held; ++u) list[list_at(u, i, n)] = win[u & (kWin - 1)][lane];
        }
    }
}

// Appends entry j at slot cnt of a lane's list (LDS window or HBM).
template <uint32_t kWin>
__device__ __forceinline__ void push_entry(uint32_t (*win)[kNeighborBlock], uint32_t j, uint32_t& cnt,
                                           uint32_t& spill, uint32_t flushed, uint32_t cap, uint32_t lane,
                                           uint32_t i, uint32_t n, uint32_t* __restrict__ list) {
    if (cnt < cap) {
        if (cnt < spill && cnt >= flushed + kWin) spill = cnt;
        if (cnt < spill) win[cnt & (kWin - 1)][lane] = j;
        else list[list_at(cnt, i, n)] = j;
    }
    ++cnt;
}

// Four consecutive floats from a 4-byte aligned address, as one 16-byte
// load (global loads on gfx950 take any alignment).
__device__ __forceinline__ float4 ld4(const float* __restrict__ base, uint32_t idx) {
    float4 v;
    __builtin_memcpy(&v, reinterpret_cast<const char*>(base) + idx * 4u, sizeof(v));
    return v;
}

// Tests candidates [j0, j0 + G) of one cell range ending at e (j0 < e) and
// appends the hits (ascending) in one pass: candidate u is written to the
// slot after the hits before it, so a miss is overwritten by the next hit or
// left just past the lane's entries (the strict window bound keeps that
// slot's row free: it is the row of an entry already written out).  Each
// hit adds one to the running slot, straight from its compare: no hit mask
// is assembled on this path.  When the hits would pass the spill point, the
// window or the capacity -- rarely -- they go entry by entry.  G is 4 or 8
// (one or two 16-byte loads per coordinate); the coordinate arrays carry
// G - 1 padding entries past their end (kFluidCoordPad, kBoundaryCoordPad),
// so the loads need no clamping.  Only the particle's own cell can hold
// `self`.
template <int G, bool kSkipSelf, uint32_t kWin>
__device__ __forceinline__ void test_push(const Coords& src, uint32_t j0, uint32_t e, uint32_t self, float4 xi,
                                          float h2, uint32_t (*win)[kNeighborBlock], uint32_t& cnt,
                                          uint32_t& spill, uint32_t flushed, uint32_t cap, uint32_t lane,
                                          uint32_t i, uint32_t n, uint32_t* __restrict__ list) {
    static_assert(G == 4 || G == 8, "candidate groups of 4 or 8");
    float xs[G], ys[G], zs[G];
#pragma unroll
    for (int q = 0; q < G / 4; ++q) {
        const float4 cx = ld4(src.x, j0 + 4u * q), cy = ld4(src.y, j0 + 4u * q), cz = ld4(src.z, j0 + 4u * q);
        xs[4 * q] = cx.x, xs[4 * q + 1] = cx.y, xs[4 * q + 2] = cx.z, xs[4 * q + 3] = cx.w;
        ys[4 * q] = cy.x, ys[4 * q + 1] = cy.y, ys[4 * q + 2] = cy.z, ys[4 * q + 3] = cy.w;
        zs[4 * q] = cz.x, zs[4 * q + 1] = cz.y, zs[4 * q + 2] = cz.z, zs[4 * q + 3] = cz.w;
    }
    const uint32_t lim = e - j0;
    bool hit[G];
    uint32_t nh = 0;
#pragma unroll
    for (int u = 0; u < G; ++u) {
        hit[u] = dist2(xi.x - xs[u], xi.y - ys[u], xi.z - zs[u]) <= h2 && (uint32_t)u < lim &&
                 (!kSkipSelf || j0 + (uint32_t)u != self);
        nh += hit[u] ? 1u : 0u;
    }
    if (nh == 0) return;  // lanes without a hit write nothing
    const uint32_t end = cnt + nh;
    if (end <= spill && end < flushed + kWin && end <= cap) {
        uint32_t slot = cnt;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            win[slot & (kWin - 1)][lane] = j0 + (uint32_t)u;
            slot += hit[u] ? 1u : 0u;
        }
        cnt = end;
    } else {
#pragma unroll
        for (int u = 0; u < G; ++u)
            if (hit[u]) push_entry<kWin>(win, j0 + (uint32_t)u, cnt, spill, flushed, cap, lane, i, n, list);
    }
}

// Entry k of a lane's list once its search is done.  The lane stored its
// entries below `held` (its spill point or count) in the LDS window; a row
// is reused only by the entry kWin slots later, after it was written to
// HBM.  The branchless append may also leave a missed candidate in the slot
// just past the lane's last entry, which is the row of entry held - kWin,
// so the window reliably holds the last kWin - 1 entries.  Older entries
// and entries past the spill point are read back from the list rows this
// lane wrote (the same thread's stores: no fence needed).
template <uint32_t kWin>
__device__ __forceinline__ uint32_t list_entry(const uint32_t (*win)[kNeighborBlock],
                                               const uint32_t* __restrict__ list, uint32_t k, uint32_t held,
                                               uint32_t lane, uint32_t i, uint32_t n) {
    return (k < held && k + kWin > held) ? win[k & (kWin - 1)][lane] : list[list_at(k, i, n)];
}

// Tait EOS of particle i with density d (clamped ratio, compute_rho_p.cu:
// 291-317) and the density pass's outputs: rho, p, xp = (x, p / rho^2) and
// vel.w = rho.
__device__ __forceinline__ void store_density(const Phys& P, uint32_t i, const float4& xi, float d,
                                              float4* __restrict__ vel, float4* __restrict__ xp,
                                              float* __restrict__ rho_out, float* __restrict__ p_out) {
    float ratio = fmaxf(d / P.rho0, 1.f);
    ratio = fminf(ratio, P.max_ratio);
    const float p = P.stiffness * (powf(ratio, P.exponent) - 1.f);
    rho_out[i] = d;
    p_out[i] = p;
    xp[i] = make_float4(xi.x, xi.y, xi.z, p / (d * d));
    reinterpret_cast<float*>(vel)[4 * (size_t)i + 3] = d;
}

// Density + Tait EOS of particle i from its finished lists: the arithmetic
// and order of k_density (fluid terms in list order from W(0), times m, then
// the boundary terms in list order), so every form stays bit-identical.
template <int B>
__device__ __forceinline__ void density_from_lists(const float4* __restrict__ pos, const float4* __restrict__ bpos,
                                                   const uint32_t (*wf)[kNeighborBlock],
                                                   const uint32_t (*wb)[kNeighborBlock], const uint32_t* __restrict__ nlist,
                                                   const uint32_t* __restrict__ blist, uint32_t nf, uint32_t hf,
                                                   uint32_t nb, uint32_t hb, uint32_t lane, uint32_t i, uint32_t n,
                                                   float4 xi, const Phys& P, float4* __restrict__ vel,
                                                   float4* __restrict__ xp, float* __restrict__ rho_out,
                                                   float* __restrict__ p_out) {
    float d = P.w0;
    for (uint32_t k = 0; k < nf; k += B) {
        uint32_t j[B];
#pragma unroll
        for (int u = 0; u < B; ++u) j[u] = k + u < nf ? list_entry<kWinFluid>(wf, nlist, k + u, hf, lane, i, n) : i;
        float4 xj[B];
#pragma unroll
        for (int u = 0; u < B; ++u) xj[u] = ld(pos, j[u]);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const float q = fsqrt(dist2(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z)) * P.inv_h;
            if (k + u < nf) d = add_unfused(d, cubic_w(q, P));
        }
    }
    d = d * P.mass;
    for (uint32_t k = 0; k < nb; k += 4) {
        uint32_t j[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = k + u < nb ? list_entry<kWinBoundary>(wb, blist, k + u, hb, lane, i, n) : 0u;
        float4 xb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xb[u] = ld(bpos, j[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float q = fsqrt(dist2(xi.x - xb[u].x, xi.y - xb[u].y, xi.z - xb[u].z)) * P.inv_h;
            if (k + u < nb) d = add_unfused(d, cubic_w(q, P) * xb[u].w);
        }
    }
    store_density(P, i, xi, d, vel, xp, rho_out, p_out);
}

// A particle's cell in a grid, and the distances from the particle to the
// faces of that cell (negative values from rounding are clamped).
struct CellPos {
    uint32_t lin;
    int sx, sy;  // strides of x and y in cells
    float lox, hix, loy, hiy, loz, hiz, cs;
};

__device__ __forceinline__ CellPos cell_pos(const float4& xi, const Grid& g, uint32_t c0) {
    const int ax = cell_coord(xi.x, g.cs), ay = cell_coord(xi.y, g.cs), az = cell_coord(xi.z, g.cs);
    const int cx = min(max(ax - g.c0x, g.R), g.nx - 1 - g.R);
    const int cy = min(max(ay - g.c0y, g.R), g.ny - 1 - g.R);
    const int cz = min(max(az - g.c0z, g.R), g.nz - 1 - g.R);
    CellPos c;
    c.lin = c0 + ((uint32_t)cx * g.ny + cy) * g.nz + cz;
    c.sx = g.ny * g.nz;
    c.sy = g.nz;
    c.lox = fmaxf(xi.x - (float)ax * g.cs, 0.f);
    c.hix = fmaxf((float)(ax + 1) * g.cs - xi.x, 0.f);
    c.loy = fmaxf(xi.y - (float)ay * g.cs, 0.f);
    c.hiy = fmaxf((float)(ay + 1) * g.cs - xi.y, 0.f);
    c.loz = fmaxf(xi.z - (float)az * g.cs, 0.f);
    c.hiz = fmaxf((float)(az + 1) * g.cs - xi.z, 0.f);
    c.cs = g.cs;
    return c;
}

// Distance along one axis from the particle to the cell d cells away.
__device__ __forceinline__ float cell_gap(int d, float lo, float hi, float cs) {
    return d < 0 ? lo + (float)(-d - 1) * cs : (d > 0 ? hi + (float)(d - 1) * cs : 0.f);
}

// kDensity: the particle's density pass (k_density's outputs) follows its
// search in the same kernel, reading the list entries still in LDS and the
// neighbours' positions the search has just brought into the caches.
// R: walk range of the fluid grid (cells of size h * multiplier; R = 1 for
// multipliers >= 1, 2 for 0.5 and 2/3), common.cuh:126-151.  The boundary
// is searched in its own grid of cell size h, 27 cells (common.cuh:153-178).
template <bool kDensity, int R>
__global__ __launch_bounds__(kNeighborBlock) void k_neighbors(
    const float4* __restrict__ pos, uint32_t n, const SceneDev* __restrict__ sc, const uint32_t* __restrict__ pstart,
    uint32_t S, const uint2* __restrict__ crange,
    const float4* __restrict__ bpos, const uint2* __restrict__ brange, const uint8_t* __restrict__ bnear, float h2,
    uint32_t capf, uint32_t capb, uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount,
    uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount, DevStatus* st, Phys P, float4* __restrict__ vel,
    float4* __restrict__ xp, float* __restrict__ rho_out, float* __restrict__ p_out, Coords fxyz, Coords bxyz,
    uint32_t only) {
    constexpr int W = 2 * R + 1;  // cells per row
    __shared__ uint32_t wf[kWinFluid][kNeighborBlock];
    __shared__ uint32_t wb[kWinBoundary][kNeighborBlock];
    const uint32_t lane = threadIdx.x;
    const uint32_t i = particle_index();
    bool valid = i < n;
    const uint32_t ii = valid ? i : 0u;  // a readable particle for idle lanes
    const float4 xi = ld(pos, ii);
    // Cell of the particle in its scene's grids.  A neighbour cell whose box
    // is farther than h holds no neighbour and is skipped.  The 1e-3 relative
    // margin on h^2 covers the rounding of the cell bounds and of the
    // predicate, so a cell holding a pair the predicate accepts is never
    // skipped.  A wavefront inside one scene reads the grids through a
    // wave-uniform index.
    CellPos F, B;
    {
        const uint32_t sid = find_scene(pstart, S, ii);
        const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sid);
        const SceneDev& sd = sc[__ballot(sid != s0) == 0ull ? s0 : sid];
        F = cell_pos(xi, sd.g, sd.c0);
        B = cell_pos(xi, sd.bg, sd.bc0);
        if (only != kAllForms && sd.form != only) valid = false;  // the other builder's scene
    }
    if (only != kAllForms && __ballot(valid) == 0ull) return;  // a wavefront of the other builder's scenes
    const float h2m = h2 * 1.001f;
    uint32_t cf = 0, cb = 0;           // entries found (may exceed the capacity)
    uint32_t ff = 0, fb = 0;           // quads [0, ff / 4) / [0, fb / 4) already written to HBM
    uint32_t sf = capf, sb = capb;     // this lane's first slot stored to HBM directly
    for (int dx = -R; dx <= R; ++dx) {
        const float ex = cell_gap(dx, F.lox, F.hix, F.cs);
        for (int dy = -R; dy <= R; ++dy) {
            const float ey = cell_gap(dy, F.loy, F.hiy, F.cs);
            const float exy = ex * ex + ey * ey;
            // with R = 2 the corner rows are often out of reach of every lane
            if (R > 1 && __ballot(valid && exy <= h2m) == 0ull) continue;
            const uint32_t row = F.lin + dx * F.sx + dy * F.sy;
            uint2 fr[W];
#pragma unroll
            for (int dz = 0; dz < W; ++dz) fr[dz] = ld(crange, row + dz - R);
#pragma unroll
            for (int dz = 0; dz < W; ++dz) {
                const float ez = cell_gap(dz - R, F.loz, F.hiz, F.cs);
                const bool keep = valid && exy + ez * ez <= h2m;
                const uint32_t e0 = keep ? fr[dz].y : fr[dz].x;
                if (dx == 0 && dy == 0 && dz == R) {  // the own cell (uniform)
                    for (uint32_t j0 = fr[dz].x; j0 < e0; j0 += kFluidGroup)
                        test_push<kFluidGroup, true, kWinFluid>(fxyz, j0, e0, i, xi, h2, wf, cf, sf, ff, capf,
                                                                lane, i, n, nlist);
                } else {
                    for (uint32_t j0 = fr[dz].x; j0 < e0; j0 += kFluidGroup)
                        test_push<kFluidGroup, false, kWinFluid>(fxyz, j0, e0, i, xi, h2, wf, cf, sf, ff, capf,
                                                                 lane, i, n, nlist);
                }
            }
            // quads every lane has filled are complete: write them out (ff
            // stays a multiple of 4)
            const uint32_t lf = min(wave_min_u32(valid ? min(cf, capf) : 0xffffffffu), capf);
            if (lf >= ff + 4u) {
                const uint32_t q1 = min(lf, ff + kWinFluid) >> 2;
                flush_quads<kWinFluid>(wf, ff >> 2, q1, min(cf, sf), cf, lane, i, n, valid, nlist);
                ff = q1 << 2;
            }
        }
    }
    // a wavefront none of whose particles has a boundary particle in its 27
    // boundary cells (most of the fluid) skips the boundary walk
    if (__ballot(valid && bnear[B.lin] != 0) != 0ull) {
        for (int dx = -1; dx <= 1; ++dx) {
            const float ex = cell_gap(dx, B.lox, B.hix, B.cs);
            for (int dy = -1; dy <= 1; ++dy) {
                const float ey = cell_gap(dy, B.loy, B.hiy, B.cs);
                const float exy = ex * ex + ey * ey;
                const uint32_t row = B.lin + dx * B.sx + dy * B.sy;
                uint2 br[3];
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) br[dz] = ld(brange, row + dz - 1);
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    const float ez = cell_gap(dz - 1, B.loz, B.hiz, B.cs);
                    const bool keep = valid && exy + ez * ez <= h2m;
                    const uint32_t e1 = keep ? br[dz].y : br[dz].x;
                    for (uint32_t j0 = br[dz].x; j0 < e1; j0 += kBoundaryGroup)
                        test_push<kBoundaryGroup, false, kWinBoundary>(bxyz, j0, e1, 0u, xi, h2, wb, cb, sb, fb,
                                                                       capb, lane, i, n, blist);
                }
                const uint32_t lb = min(wave_min_u32(valid ? min(cb, capb) : 0xffffffffu), capb);
                if (lb >= fb + 4u) {
                    const uint32_t q1 = min(lb, fb + kWinBoundary) >> 2;
                    flush_quads<kWinBoundary>(wb, fb >> 2, q1, min(cb, sb), cb, lane, i, n, valid, blist);
                    fb = q1 << 2;
                }
            }
        }
    }
    const uint32_t mf = wave_max_u32(valid ? cf : 0u), mb = wave_max_u32(valid ? cb : 0u);
    // the remaining quads of the windows (partially filled across lanes)
    flush_quads<kWinFluid>(wf, ff >> 2, (min(min(mf, capf), ff + kWinFluid) + 3u) >> 2, min(cf, sf), cf, lane, i, n,
                           valid, nlist);
    flush_quads<kWinBoundary>(wb, fb >> 2, (min(min(mb, capb), fb + kWinBoundary) + 3u) >> 2, min(cb, sb), cb, lane,
                              i, n, valid, blist);
    if (valid) {
        ncount[i] = cf;
        bcount[i] = cb;
        if (kDensity)
            density_from_lists<kDensityBatch>(pos, bpos, wf, wb, nlist, blist, min(cf, capf), min(cf, sf),
                                              min(cb, capb), min(cb, sb), lane, i, n, xi, P, vel, xp, rho_out, p_out);
    }
    if (lane == 0) {
        if (mf) atomic_max_if_above(&st->max_fluid, mf);
        if (mb) atomic_max_if_above(&st->max_boundary, mb);
        uint32_t fl = 0;
        if (mf > capf) fl |= 1u;  // WCSPH_FLAG_FLUID_OVERFLOW
        if (mb > capb) fl |= 2u;  // WCSPH_FLAG_BOUNDARY_OVERFLOW
        if (fl) atomicOr(&st->flags, fl);
    }
}

// ------------------------------------------------------------------
// 5'. neighbour lists, cell-cooperative form.  One wavefront per workgroup
//    walks a contiguous range of occupied cells (Morton order).  For a cell,
//    lanes 0..26 load the fluid and boundary ranges of the 27 neighbour
//    cells in reference order; their prefix sums place every candidate at a
//    position t of the concatenation, and the candidate indices of a window
//    of kCoopWin positions are expanded into LDS.  Lane t then loads
//    candidate t, coalesced: a cell costs a few vector loads in all, where
//    the per-particle form issues one per candidate and home particle.
//    Each home particle of the cell, broadcast from a lane, is tested
//    against the window; the ballot of hits is in candidate order and is
//    appended at the home's running count in an LDS stage (kSlots entries
//    per home, later ones go to HBM directly).  The stage is written out
//    slot-major, the homes of one slot side by side.
// ------------------------------------------------------------------
constexpr uint32_t kCoopBlock = 64;
constexpr uint32_t kCoopHomes = 16;             // home particles per pass
constexpr uint32_t kCoopWin = 4 * kCoopBlock;   // candidate window (4 per lane)
constexpr uint32_t kCoopSlotsF = 64;            // staged entries per home
constexpr uint32_t kCoopSlotsB = 32;
constexpr uint32_t kCoopPadF = kCoopSlotsF + 8; // row strides: the flush's
constexpr uint32_t kCoopPadB = kCoopSlotsB + 8; // reads see <= 2-way bank conflicts

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float readlane_f32(float v, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Orders the LDS accesses of the lanes of a one-wavefront workgroup: the
// hardware performs one wavefront's LDS operations in issue order, so a
// compiler barrier suffices.  Unlike __syncthreads() it does not wait for
// the wavefront's outstanding global stores.
static_assert(kCoopBlock == 64, "the cell-cooperative builder is one wavefront per workgroup");
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tests the home particles [hb, hb + nh) (lane h holds xh of home h) against
// the candidates of one source whose 27 ranges lanes 0..26 hold in r, with
// exclusive prefix `excl` and total `tot`.  Hits are staged at
// stage[h * kPad + slot]; cnt (lane h) counts home h's entries, including
// those past the capacity, which are not stored.  Without kSpill the
// capacity fits the stage (cap <= kSlots) and nothing goes to HBM here.
template <bool kFluid, bool kSpill, uint32_t kSlots, uint32_t kPad>
__device__ __forceinline__ void coop_source(const float4* __restrict__ src, uint32_t nsrc, uint2 r, uint32_t excl,
                                            uint32_t tot,
                                            uint32_t lane, uint32_t hb, uint32_t nh, float4 xh, float h2, uint32_t n,
                                            uint32_t cap, uint32_t* __restrict__ cand, uint32_t* __restrict__ stage,
                                            uint32_t* __restrict__ list, uint32_t& cnt) {
    for (uint32_t t0 = 0; t0 < tot; t0 += kCoopWin) {
        wave_lds_sync();  // the previous window's indices are consumed
        if (lane < 27) {
            const uint32_t lo = max(t0, excl), hi = min(t0 + kCoopWin, excl + (r.y - r.x));
            for (uint32_t q = lo; q < hi; ++q) cand[q - t0] = r.x + (q - excl);
        }
        wave_lds_sync();
        const uint32_t nt = min(kCoopWin, tot - t0);
        const uint32_t nbat = (nt + kCoopBlock - 1u) / kCoopBlock;  // uniform
        uint32_t jj[4];
        float cx[4], cy[4], cz[4];
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t t = b * kCoopBlock + lane;
            // position t0 always holds a candidate; the clamp keeps every
            // load inside the source array whatever the LDS holds
            jj[b] = min(cand[t < nt ? t : 0u], nsrc - 1u);
            const float4 p = ld(src, jj[b]);
            cx[b] = p.x;
            cy[b] = p.y;
            cz[b] = p.z;
        }
        for (uint32_t h = 0; h < nh; ++h) {
            const float hx = readlane_f32(xh.x, h), hy = readlane_f32(xh.y, h), hz = readlane_f32(xh.z, h);
            const uint32_t ih = hb + h;
            uint32_t c = readlane_u32(cnt, h);
            uint32_t* row = stage + h * kPad;
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b) {
                if (b < nbat) {
                    // the hit predicate, and the same compares as a lane mask
                    // (ordered <=: a NaN distance is no hit)
                    const uint32_t t = b * kCoopBlock + lane;
                    const float d2 = dist2(hx - cx[b], hy - cy[b], hz - cz[b]);
                    const bool hit = (t < nt) & (!kFluid | (jj[b] != ih)) & (d2 <= h2);
                    unsigned long long m = __builtin_amdgcn_fcmpf(d2, h2, 5 /* OLE */) &
                                           __builtin_amdgcn_uicmp(t, nt, 36 /* ULT */);
                    if (kFluid) m &= __builtin_amdgcn_uicmp(jj[b], ih, 33 /* NE */);
                    if (m) {
                        const uint32_t slot = c + lanes_below(m);
                        if (!kSpill) {
                            if (hit & (slot < cap)) row[slot] = jj[b];
                        } else if (c + (uint32_t)__popcll(m) <= kSlots) {
                            if (hit) row[slot] = jj[b];
                        } else if (hit) {
                            if (slot < kSlots) row[slot] = jj[b];
                            else if (slot < cap) list[list_at(slot, ih, n)] = jj[b];
                        }
                        c += (uint32_t)__popcll(m);
                    }
                }
            }
            if (lane == h) cnt = c;
        }
    }
}

// Writes the staged entries of homes [hb, hb + nh) to the list: lane (h, q)
// of a pass stores slot quad q of home h as one 16-byte store, so the homes
// of one quad are adjacent in memory.  The stage holds slots below kSlots (a
// multiple of 4; later ones went to HBM directly), so a quad is cut only by
// the home's count, past which nothing is read.
template <uint32_t kSlots, uint32_t kPad>
__device__ __forceinline__ void coop_flush(const uint32_t* __restrict__ stage, uint32_t cnt, uint32_t lane,
                                           uint32_t hb, uint32_t nh, uint32_t n, uint32_t cap,
                                           uint32_t* __restrict__ list) {
    static_assert(kSlots % 4 == 0 && kPad % 4 == 0, "stage rows are whole 16-byte quads");
    wave_lds_sync();  // the stage writes of every lane are complete
    const uint32_t held = min(min(cnt, kSlots), cap);
    const uint32_t sh = nh <= 4u ? 2u : (nh <= 8u ? 3u : 4u);  // uniform
    const uint32_t h = lane & ((1u << sh) - 1u), qo = lane >> sh;
    const uint32_t hh = (uint32_t)__shfl((int)held, (int)h, 64);
    const uint32_t top = (wave_max_u32(lane < nh ? held : 0u) + 3u) >> 2;
    for (uint32_t q0 = 0; q0 < top; q0 += kCoopBlock >> sh) {
        const uint32_t q = q0 + qo;
        if (h < nh && 4u * q < hh)
            *reinterpret_cast<uint4*>(list + list_at(4u * q, hb + h, n)) =
                *reinterpret_cast<const uint4*>(stage + h * kPad + 4u * q);
    }
}

// Inclusive prefix sum over lanes 0..31 (lanes 32..63 hold no range):
// DPP row shifts scan each row of 16 lanes, row_bcast:15 carries row 0's
// total into row 1.
__device__ __forceinline__ uint32_t scan32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    return v;
}

// Fluid and boundary ranges of the 27 neighbour cells of cell `lin` (lane k
// < 27 takes neighbour k in reference order, dx outer, dz inner).
__device__ __forceinline__ void coop_ranges(const SceneDev* __restrict__ sc, const uint32_t* __restrict__ cstart,
                                            uint32_t S, const uint2* __restrict__ crange,
                                            const uint2* __restrict__ brange, uint32_t lin, uint32_t lane,
                                            uint2& fr, uint2& br) {
    const Grid g = sc[find_scene(cstart, S, lin)].g;
    const int kk = lane < 27 ? (int)lane : 13;
    const int noff = (kk / 9 - 1) * (g.ny * g.nz) + ((kk / 3) % 3 - 1) * g.nz + (kk % 3 - 1);
    fr = make_uint2(0u, 0u);
    br = make_uint2(0u, 0u);
    if (lane < 27) {
        fr = crange[lin + noff];
        br = brange[lin + noff];
    }
}

template <bool kSpill>
__global__ __launch_bounds__(kCoopBlock) void k_neighbors_coop(
    const float4* __restrict__ pos, uint32_t n, const SceneDev* __restrict__ sc, const uint32_t* __restrict__ cstart,
    uint32_t S, const uint2* __restrict__ crange,
    const float4* __restrict__ bpos, uint32_t nb, const uint2* __restrict__ brange, const uint32_t* __restrict__ occ,
    float h2, uint32_t capf, uint32_t capb, uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount,
    uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount, DevStatus* st, uint32_t only) {
    __shared__ uint32_t cand[kCoopWin];
    __shared__ __align__(16) uint32_t stf[kCoopHomes * kCoopPadF];
    __shared__ __align__(16) uint32_t stb[kCoopHomes * kCoopPadB];
    const uint32_t lane = threadIdx.x;
    const uint32_t nocc = st->n_occ;
    const uint32_t w = xcd_block(), G = gridDim.x;
    const uint32_t c0 = (uint32_t)(((unsigned long long)nocc * w) / G);
    const uint32_t c1 = (uint32_t)(((unsigned long long)nocc * (w + 1)) / G);
    uint32_t mf = 0, mb = 0;
    // the ranges of the next cell are loaded while this one is processed
    uint2 fr, br;
    bool have = false;  // fr, br hold cell c's ranges
    for (uint32_t c = c0; c < c1; ++c) {
        if (only != kAllForms) {
            // Occupied cells are grouped by scene, in scene order: the cells
            // of a scene another builder handles are skipped as one block
            // (the first cell of a later scene found by bisection).
            const uint32_t sid = find_scene(cstart, S, occ[c]);
            if (sc[sid].form != only) {
                const uint32_t bound = cstart[sid + 1];
                uint32_t lo = c + 1, hi = c1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (occ[mid] < bound) lo = mid + 1;
                    else hi = mid;
                }
                c = lo - 1u;
                have = false;
                continue;
            }
        }
        if (!have) coop_ranges(sc, cstart, S, crange, brange, occ[c], lane, fr, br);
        const uint2 cfr = fr, cbr = br;
        have = c + 1 < c1;
        if (have) coop_ranges(sc, cstart, S, crange, brange, occ[c + 1], lane, fr, br);
        const uint32_t szf = cfr.y - cfr.x, szb = cbr.y - cbr.x;
        const uint32_t incf = scan32(szf), incb = scan32(szb);
        const uint32_t totf = readlane_u32(incf, 26), totb = readlane_u32(incb, 26);
        // the home cell is neighbour 13 (offset 0, 0, 0)
        const uint32_t home0 = readlane_u32(cfr.x, 13), home1 = readlane_u32(cfr.y, 13);
        for (uint32_t hb = home0; hb < home1; hb += kCoopHomes) {
            const uint32_t nh = min(kCoopHomes, home1 - hb);
            const float4 xh = ld(pos, hb + min(lane, nh - 1u));
            uint32_t cf = 0, cb = 0;
            coop_source<true, kSpill, kCoopSlotsF, kCoopPadF>(pos, n, cfr, incf - szf, totf, lane, hb, nh, xh, h2, n,
                                                              capf, cand, stf, nlist, cf);
            coop_flush<kCoopSlotsF, kCoopPadF>(stf, cf, lane, hb, nh, n, capf, nlist);
            if (totb) {
                coop_source<false, kSpill, kCoopSlotsB, kCoopPadB>(bpos, nb, cbr, incb - szb, totb, lane, hb, nh, xh, h2,
                                                                   n, capb, cand, stb, blist, cb);
                coop_flush<kCoopSlotsB, kCoopPadB>(stb, cb, lane, hb, nh, n, capb, blist);
            }
            if (lane < nh) {
                ncount[hb + lane] = cf;
                bcount[hb + lane] = cb;
                mf = max(mf, cf);
                mb = max(mb, cb);
            }
        }
    }
    mf = wave_max_u32(mf);
    mb = wave_max_u32(mb);
    if (lane == 0) {
        if (mf) atomic_max_if_above(&st->max_fluid, mf);
        if (mb) atomic_max_if_above(&st->max_boundary, mb);
        uint32_t fl = 0;
        if (mf > capf) fl |= 1u;  // WCSPH_FLAG_FLUID_OVERFLOW
        if (mb > capb) fl |= 2u;  // WCSPH_FLAG_BOUNDARY_OVERFLOW
        if (fl) atomicOr(&st->flags, fl);
    }
}

// List entries k .. k+B-1 of particle i (k a multiple of 4), B / 4 quad
// loads; slots past the count are replaced by `self`, a valid index, and
// masked by the caller.  Quads past the last one holding an entry are not
// read (the quad index is clamped to it).
template <int B>
__device__ __forceinline__ void load_batch(const uint32_t* __restrict__ list, uint32_t n, uint32_t i, uint32_t k,
                                           uint32_t cnt, uint32_t self, uint32_t (&j)[B]) {
    static_assert(B % 4 == 0, "whole quads");
    const uint32_t qlast = cnt ? (cnt - 1u) >> 2 : 0u;
#pragma unroll
    for (int g = 0; g < B / 4; ++g) {
        const uint32_t q = min((k >> 2) + (uint32_t)g, qlast);
        // each list quad is read once per pass: streamed past L2, which is
        // left to the neighbour data the pass gathers
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 e = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(list + list_at(4u * q, i, n)));
        const uint32_t ev[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) j[4 * g + u] = k + 4u * g + u < cnt ? ev[u] : self;
    }
}

// ------------------------------------------------------------------
// 6. density + Tait EOS (compute_rho_p_with_boundary, compute_rho_p.cu:242-318)
//    Neighbours are taken in batches of 8: the next batch's list entries
//    are requested before this batch's gathers, so one round trip to memory
//    serves eight neighbours.
// ------------------------------------------------------------------
__global__ void k_density(const float4* __restrict__ pos, float4* __restrict__ vel, uint32_t n,
                          const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                          const float4* __restrict__ bpos, const uint32_t* __restrict__ blist,
                          const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb, Phys P,
                          float4* __restrict__ xp, float* __restrict__ rho_out, float* __restrict__ p_out,
                          const SceneDev* __restrict__ sc, const uint32_t* __restrict__ pstart, uint32_t S,
                          uint32_t only) {
    constexpr int B = kDensityBatch;
    const uint32_t i = particle_index();
    if (i >= n) return;
    if (only != kAllForms && sc[find_scene(pstart, S, i)].form != only) return;  // density fused in its search
    const float4 xi = ld(pos, i);
    // counts above the capacity are reported by k_neighbors; only the
    // stored slots are read
    const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
    float d = P.w0;
    uint32_t jn[B];
    load_batch<B>(nlist, n, i, 0, cf, i, jn);
    for (uint32_t k = 0; k < cf; k += B) {
        uint32_t j[B];
#pragma unroll
        for (int u = 0; u < B; ++u) j[u] = jn[u];
        if (k + B < cf) load_batch<B>(nlist, n, i, k + B, cf, i, jn);
        float4 xj[B];
#pragma unroll
        for (int u = 0; u < B; ++u) xj[u] = ld(pos, j[u]);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const float q = fsqrt(dist2(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z)) * P.inv_h;
            if (k + u < cf) d = add_unfused(d, cubic_w(q, P));
        }
    }
    d = d * P.mass;
    for (uint32_t k = 0; k < cb; k += 4) {
        uint32_t j[4];
        load_batch<4>(blist, n, i, k, cb, 0u, j);
        float4 xb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xb[u] = ld(bpos, j[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float q = fsqrt(dist2(xi.x - xb[u].x, xi.y - xb[u].y, xi.z - xb[u].z)) * P.inv_h;
            if (k + u < cb) d = add_unfused(d, cubic_w(q, P) * xb[u].w);
        }
    }
    store_density(P, i, xi, d, vel, xp, rho_out, p_out);
}

// ------------------------------------------------------------------
// 6b. normals for surface tension (compute_pressure_accel_n_normal.cu:98-124)
// ------------------------------------------------------------------
__global__ void k_normals(const float4* __restrict__ xp, const float4* __restrict__ vel, uint32_t n,
                          const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                          uint32_t capf, Phys P, float4* __restrict__ nrm) {
    const uint32_t i = particle_index();
    if (i >= n) return;
    const float4 xi = xp[i];
    const uint32_t cf = min(ncount[i], capf);
    float nx = 0.f, ny = 0.f, nz = 0.f;
    for (uint32_t k = 0; k < cf; k += 4) {
        uint32_t j[4];
        load_batch<4>(nlist, n, i, k, cf, i, j);
        float4 xj[4];
        float dj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            xj[u] = ld(xp, j[u]);
            dj[u] = ld(vel, j[u]).w;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < cf) add_normal(nx, ny, nz, P, xi, xj[u], dj[u]);
    }
    const float f = P.h * P.mass;
    nrm[i] = make_float4(f * nx, f * ny, f * nz, 0.f);
}

// ------------------------------------------------------------------
// 7. forces + symplectic Euler + box collisions + CFL max
//    (compute_pressure_accel_n_normal_with_boundary :337-365,
//     compute_non_pressure_accel :113-145, update_positions.cu:14-68)
// ------------------------------------------------------------------
__device__ __forceinline__ void mat4_apply(const float* m, float x, float y, float z, float w, float* o) {
#pragma clang fp contract(off)
    for (int r = 0; r < 3; ++r)
        o[r] = __builtin_fmaf(m[12 + r], w, __builtin_fmaf(m[8 + r], z, __builtin_fmaf(m[4 + r], y, m[r] * x)));
}

// Sums of one particle's fluid pair terms.
struct ForceSums {
    float pax, pay, paz;  // pressure (before the mass factor)
    float vlx, vly, vlz;  // viscosity
    float fx, fy, fz;     // surface tension
};

__device__ __forceinline__ ForceSums zero_sums() {
    ForceSums F;
    F.pax = F.pay = F.paz = 0.f;
    F.vlx = F.vly = F.vlz = 0.f;
    F.fx = F.fy = F.fz = 0.f;
    return F;
}

// One fluid neighbour j of i: pressure (compute_pressure_accel_n_normal.cu:
// 98-124), viscosity and surface tension (compute_non_pressure_accel.cu:
// 113-145).  xi/xj = (x, p / rho^2), vi/vj = (v, rho).  kGrad / kVisc fix
// the gradient and viscosity kinds at compile time (-1: read from P); the
// arithmetic is the same either way.  j == i adds exact zeros to every sum
// (q = 0 gives a zero gradient, vr = 0 no viscosity, r^2 = 0 no cohesion),
// so list batches are padded with i.
template <bool kNormals, int kGrad = -1, int kVisc = -1>
__device__ __forceinline__ void add_fluid_pair(ForceSums& F, const Phys& P, const float4& xi, const float4& vi,
                                               const float4& ni, const float4& xj, const float4& vj,
                                               const float4& nj) {
#pragma clang fp contract(off)
    const float di = vi.w, dpi = xi.w;
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = dist2(rx, ry, rz);
    const float q = fsqrt(r2) * P.inv_h;
    const float gc = cubic_grad(q, P);
    const bool spiky = kGrad < 0 ? P.grad_kind != 0 : kGrad != 0;
    const float gp = spiky ? spiky_ref_grad(q, P) : gc;
    const float cp = (dpi + xj.w) * gp;
    F.pax = __builtin_fmaf(-cp, rx, F.pax);
    F.pay = __builtin_fmaf(-cp, ry, F.pay);
    F.paz = __builtin_fmaf(-cp, rz, F.paz);
    const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
    const float vr = __builtin_fmaf(vz, rz, __builtin_fmaf(vy, ry, vx * rx));
    if (kVisc < 0 ? P.visc_kind != 0 : kVisc != 0) {
        const float c = vr * gc * frcp(vj.w * (r2 + P.visc_eps));
        F.vlx = __builtin_fmaf(c, rx, F.vlx);
        F.vly = __builtin_fmaf(c, ry, F.vly);
        F.vlz = __builtin_fmaf(c, rz, F.vlz);
    } else {
        // -Pi_ij * gc with Pi_ij = -nu vr / (r^2 + eps), nu = 2 alpha h c_s / (rho_i + rho_j),
        // for approaching pairs (vr < 0) only: a select, not a branch
        const float c = P.art_visc_coef * vr * frcp((di + vj.w) * (r2 + P.visc_eps)) * gc;
        const bool near = vr < 0.f;
        F.vlx = near ? __builtin_fmaf(c, rx, F.vlx) : F.vlx;
        F.vly = near ? __builtin_fmaf(c, ry, F.vly) : F.vly;
        F.vlz = near ? __builtin_fmaf(c, rz, F.vlz) : F.vlz;
    }
    if (kNormals) {  // pairs closer than 1e-3 m are skipped (a select)
        const bool use = r2 > 1e-6f;
        const float irl = frcp(fsqrt(r2));
        const float cq = P.mass * cohesion_w(q, P);
        const float inv = frcp(di + vj.w);
        F.fx = use ? __builtin_fmaf(__builtin_fmaf(cq, rx * irl, ni.x) - nj.x, inv, F.fx) : F.fx;
        F.fy = use ? __builtin_fmaf(__builtin_fmaf(cq, ry * irl, ni.y) - nj.y, inv, F.fy) : F.fy;
        F.fz = use ? __builtin_fmaf(__builtin_fmaf(cq, rz * irl, ni.z) - nj.z, inv, F.fz) : F.fz;
    }
}

// One boundary neighbour (Akinci pressure term, compute_pressure_accel_n_normal.cu:
// 337-365); xb.w = Psi.  `use` false adds a zero term.
template <int kGrad = -1>
__device__ __forceinline__ void add_boundary_pair(float& apx, float& apy, float& apz, const Phys& P, const float4& xi,
                                                  float bterm, const float4& xb, bool use) {
#pragma clang fp contract(off)
    const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
    const float q = fsqrt(dist2(rx, ry, rz)) * P.inv_h;
    const bool spiky = kGrad < 0 ? P.grad_kind != 0 : kGrad != 0;
    const float gp = spiky ? spiky_ref_grad(q, P) : cubic_grad(q, P);
    const float c = use ? xb.w * bterm * gp : 0.f;
    apx = __builtin_fmaf(-c, rx, apx);
    apy = __builtin_fmaf(-c, ry, apy);
    apz = __builtin_fmaf(-c, rz, apz);
}

// Akinci boundary factor p_i / rho_i^2 + p_i / rho0^2.
__device__ __forceinline__ float boundary_term(const Phys& P, float dpi, float p) {
    return __builtin_fmaf(p, P.inv_rho0_sq, dpi);
}


// Non-pressure acceleration: gravity + external force + viscosity -
// surface tension (compute_non_pressure_accel.cu:109-111, 142-145).
__device__ __forceinline__ void nonpressure_accel(const Phys& P, const float4& xi, const ForceSums& F, float& anx,
                                                  float& any, float& anz) {
#pragma clang fp contract(off)
    const float vfac = P.visc_kind ? P.visc * 10.f * P.mass : P.mass;
    const float sfac = P.surface_tension * 2.f * P.rho0;
    float ex, ey, ez;
    external_accel(P, xi.x, xi.y, xi.z, ex, ey, ez);
    anx = __builtin_fmaf(-sfac, F.fx, __builtin_fmaf(vfac, F.vlx, P.gx + ex));
    any = __builtin_fmaf(-sfac, F.fy, __builtin_fmaf(vfac, F.vly, P.gy + ey));
    anz = __builtin_fmaf(-sfac, F.fz, __builtin_fmaf(vfac, F.vlz, P.gz + ez));
}

// Symplectic Euler + box collisions of particle i (update_positions.cu:
// 14-68); stores x, v, id and returns |v|^2 (FP64) and the non-finite flag.
__device__ __forceinline__ void integrate_store(const Phys& P, const SceneDev& B, float dt, float dt_np,
                                                const float4& xi, const float4& vi, float apx, float apy, float apz,
                                                float anx, float any, float anz, uint32_t i, uint32_t id,
                                                float4* __restrict__ pos_out, float4* __restrict__ vel_out,
                                                uint32_t* __restrict__ id_out, uint32_t* __restrict__ coll,
                                                double& v2, bool& bad, float4& stored) {
#pragma clang fp contract(off)
    float v[3] = {vi.x + __builtin_fmaf(dt_np, anx, dt * apx), vi.y + __builtin_fmaf(dt_np, any, dt * apy),
                  vi.z + __builtin_fmaf(dt_np, anz, dt * apz)};
    float x[3] = {__builtin_fmaf(dt, v[0], xi.x), __builtin_fmaf(dt, v[1], xi.y), __builtin_fmaf(dt, v[2], xi.z)};
    // box collisions in the box frame (update_positions.cu:14-52)
    float bx[3], bv[3];
    if (B.box_identity) {
        for (int a = 0; a < 3; ++a) {
            bx[a] = x[a];
            bv[a] = v[a];
        }
    } else {
        mat4_apply(B.model_inv, x[0], x[1], x[2], 1.f, bx);
        mat4_apply(B.model_inv, v[0], v[1], v[2], 0.f, bv);
    }
    bool changed = false;
    uint32_t flags = 0;
    const float buf = P.r * P.cbuf;
    for (int a = 0; a < 3; ++a) {
        if (bx[a] - buf < B.bmin[a]) {
            bx[a] = B.bmin[a] + P.r;
            bv[a] *= -P.elasticity;
            changed = true;
            flags |= 1u << (2 * a);
        } else if (bx[a] + buf > B.bmax[a]) {
            bx[a] = B.bmax[a] - P.r;
            bv[a] *= -P.elasticity;
            changed = true;
            flags |= 2u << (2 * a);
        }
    }
    if (coll) coll[i] = flags;
    if (changed) {
        if (B.box_identity) {
            for (int a = 0; a < 3; ++a) {
                x[a] = bx[a];
                v[a] = bv[a];
            }
        } else {
            mat4_apply(B.model, bx[0], bx[1], bx[2], 1.f, x);
            mat4_apply(B.model, bv[0], bv[1], bv[2], 0.f, v);
        }
    }
    stored = make_float4(x[0], x[1], x[2], 1.f);
    pos_out[i] = stored;
    vel_out[i] = make_float4(v[0], v[1], v[2], 0.f);
    id_out[i] = id;
    const double dvx = v[0], dvy = v[1], dvz = v[2];
    v2 = __builtin_fma(dvz, dvz, __builtin_fma(dvy, dvy, dvx * dvx));
    bad = !(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]) && isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]));
    if (bad) v2 = 0.0;
}

// kCubicArt: the benchmark kinds (cubic gradient, artificial viscosity)
// fixed at compile time, so every batch is straight-line code.
template <bool kNormals, bool kCubicArt = false>
__global__ void k_force_integrate(const float4* __restrict__ xp, const float4* __restrict__ vel,
                                  const uint32_t* __restrict__ id_b, uint32_t n,
                                  const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                                  const float4* __restrict__ bpos, const uint32_t* __restrict__ blist,
                                  const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb,
                                  const float* __restrict__ p_in, const float4* __restrict__ nrm, Phys P,
                                  const SceneDev* __restrict__ sc, SceneState* __restrict__ ss,
                                  const uint32_t* __restrict__ pstart, uint32_t S, DevStatus* st,
                                  float4* __restrict__ pos_out, float4* __restrict__ vel_out,
                                  uint32_t* __restrict__ id_out, uint32_t* __restrict__ coll, CellCount C) {
    const uint32_t i = particle_index();
    // a neighbour list of this (or an earlier) step overflowed: leave the
    // state as it was before the step (uniform over the grid), and count it
    // for the next sort as it is
    if (st->flags & kOverflowFlags) {
        if (C.cnt) {
            const bool valid = i < n;
            const uint32_t sid = find_scene(pstart, S, valid ? i : n - 1u);
            count_cell(valid ? pos_out[i] : make_float4(0.f, 0.f, 0.f, 0.f), valid, i, sc[sid], C, st);
        }
        return;
    }
    // the scene of each lane; a wavefront inside one scene reads the scene's
    // dt and box through a wave-uniform index (scalar loads)
    const uint32_t sid = find_scene(pstart, S, i < n ? i : n - 1);
    const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sid);
    const bool uniform_scene = __ballot(sid != s0) == 0ull;
    double v2 = 0.0;
    bool bad = false;
    float4 stored = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
        const float4 xi = xp[i];
        const float4 vi = vel[i];
        float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
        if (kNormals) ni = nrm[i];
        const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
        ForceSums F = zero_sums();
        constexpr int B = kNormals ? 4 : 8;
        uint32_t jn[B];
        load_batch<B>(nlist, n, i, 0, cf, i, jn);
        for (uint32_t k = 0; k < cf; k += B) {
            uint32_t j[B];
#pragma unroll
            for (int u = 0; u < B; ++u) j[u] = jn[u];
            if (k + B < cf) load_batch<B>(nlist, n, i, k + B, cf, i, jn);
            float4 xj[B], vj[B], nj[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                xj[u] = ld(xp, j[u]);
                vj[u] = ld(vel, j[u]);
                nj[u] = kNormals ? ld(nrm, j[u]) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            // entries past the count are i itself: exact zeros
#pragma unroll
            for (int u = 0; u < B; ++u)
                add_fluid_pair<kNormals, kCubicArt ? 0 : -1, kCubicArt ? 0 : -1>(F, P, xi, vi, ni, xj[u], vj[u], nj[u]);
        }
        float apx = P.mass * F.pax, apy = P.mass * F.pay, apz = P.mass * F.paz;
        const float bterm = boundary_term(P, xi.w, p_in[i]);
        for (uint32_t k = 0; k < cb; k += 4) {
            uint32_t j[4];
            load_batch<4>(blist, n, i, k, cb, 0u, j);
            float4 xb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xb[u] = ld(bpos, j[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                add_boundary_pair<kCubicArt ? 0 : -1>(apx, apy, apz, P, xi, bterm, xb[u], k + u < cb);
        }
        float anx, any, anz;
        nonpressure_accel(P, xi, F, anx, any, anz);
        if (uniform_scene)
            integrate_store(P, sc[s0], ss[s0].dt, ss[s0].dt_np, xi, vi, apx, apy, apz, anx, any, anz, i, id_b[i],
                            pos_out, vel_out, id_out, coll, v2, bad, stored);
        else
            integrate_store(P, sc[sid], ss[sid].dt, ss[sid].dt_np, xi, vi, apx, apy, apz, anx, any, anz, i, id_b[i],
                            pos_out, vel_out, id_out, coll, v2, bad, stored);
    }
    // the next sort's counting, of the positions just stored
    if (C.cnt) count_cell(stored, i < n, i, uniform_scene ? sc[s0] : sc[sid], C, st);
    // max |v|^2 per scene in FP64: one atomic per wavefront when the wave
    // lies in one scene, one per lane at a scene boundary
    if (uniform_scene) {
        const double m = wave_max_f64(v2);
        if ((threadIdx.x & 63) == 0 && m > 0.0)
            atomic_max_if_above(&ss[s0].vmax2_bits, (unsigned long long)__double_as_longlong(m));
    } else if (v2 > 0.0) {
        atomic_max_if_above(&ss[sid].vmax2_bits, (unsigned long long)__double_as_longlong(v2));
    }
    if (bad) atomicOr(&st->flags, 8u /* WCSPH_FLAG_NONFINITE */);
}

// ------------------------------------------------------------------
// 8. cell-loop form (no neighbour lists).  Bricks of 4 x 4 x 4 cells,
//    aligned in absolute cell coordinates, are contiguous in Morton order
//    (the low 6 bits of a cell's key are the low 2 bits of its three
//    coordinates), so a brick's particles are one contiguous range of the
//    sorted arrays.  One workgroup per brick stages the candidates of the
//    brick's 6 x 6 x 6 halo cells in LDS -- halo cells in (x, y, z) order,
//    each cell's particles in ascending index order -- and every particle
//    of the brick then walks its 27 cells from LDS in the reference order
//    (dx, dy, dz from -1 to 1, ascending index), applying the same
//    predicate and the same pair terms, in the same order, as the list
//    form.  A cell whose candidates did not all fit in LDS is read from
//    HBM.  Density and forces are separate passes (the forces need every
//    neighbour's density), each walking the cells again.
// ------------------------------------------------------------------
constexpr int kHaloEdge = 6;
constexpr int kHaloCells = kHaloEdge * kHaloEdge * kHaloEdge;
constexpr uint32_t kBrickThreads = 512;
constexpr uint32_t kDensFluidCap = 3072;  // float4 per staged candidate
constexpr uint32_t kDensBoundCap = 1024;
constexpr uint32_t kForceFluidCap = 2048;  // 2 (3 with normals) float4 per staged candidate
constexpr uint32_t kForceBoundCap = 768;

struct BrickDev {
    uint32_t scene;
    uint32_t rank_lo, rank_hi;  // global Morton ranks of the brick's first and last grid cell
    int x0, y0, z0;             // scene-grid coordinates of the brick's corner cell (may be < 0)
};

struct Halo {
    uint32_t off[kHaloCells + 1];  // LDS slot of each halo cell's first candidate (prefix of the sizes)
    uint32_t gs[kHaloCells];       // index of that candidate in the sorted arrays
};

// Cell ranges of the halo (thread t < 216 takes halo cell t) and their
// prefix offsets.  Every thread of the block calls it (barriers inside).
__device__ __forceinline__ void halo_ranges(const BrickDev& b, const Grid& g, uint32_t c0, const uint2* __restrict__ range,
                                            Halo& H, uint32_t* wsum) {
    const uint32_t t = threadIdx.x;
    uint32_t size = 0;
    if (t < (uint32_t)kHaloCells) {
        const int cx = b.x0 - 1 + (int)(t / 36u), cy = b.y0 - 1 + (int)((t / 6u) % 6u), cz = b.z0 - 1 + (int)(t % 6u);
        uint32_t gs = 0;
        if (cx >= 0 && cx < g.nx && cy >= 0 && cy < g.ny && cz >= 0 && cz < g.nz) {
            const uint2 r = range[c0 + ((uint32_t)cx * g.ny + cy) * g.nz + cz];
            gs = r.x;
            size = r.y - r.x;
        }
        H.gs[t] = gs;
    }
    const uint32_t inc = wave_incl_scan(size);
    const uint32_t wave = t >> 6;
    if ((t & 63) == 63 && wave < 4) wsum[wave] = inc;
    __syncthreads();
    if (t < (uint32_t)kHaloCells) {
        uint32_t off = 0;
        for (uint32_t w = 0; w < wave; ++w) off += wsum[w];
        H.off[t] = off + inc - size;
        if (t == (uint32_t)kHaloCells - 1) H.off[kHaloCells] = off + inc;
    }
    __syncthreads();
}

// Copies the first min(total, kCap) halo candidates of src into lds (and
// of src2 into lds2 when given).
// One thread per halo cell copies the cell's range, four candidates per
// round so that the loads are in flight together.
template <uint32_t kCap>
__device__ __forceinline__ void halo_stage(const Halo& H, const float4* __restrict__ src, float4* lds,
                                           const float4* __restrict__ src2, float4* lds2,
                                           const float4* __restrict__ src3, float4* lds3) {
    const uint32_t c = threadIdx.x;
    if (c >= (uint32_t)kHaloCells) return;
    const uint32_t o = H.off[c], e = min(H.off[c + 1], kCap), g0 = H.gs[c];
    for (uint32_t k0 = o; k0 < e; k0 += 4) {
        float4 v[4], v2[4], v3[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t j = g0 + (min(k0 + u, e - 1) - o);
            v[u] = src[j];
            if (src2) v2[u] = src2[j];
            if (src3) v3[u] = src3[j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k0 + u >= e) break;
            lds[k0 + u] = v[u];
            if (src2) lds2[k0 + u] = v2[u];
            if (src3) lds3[k0 + u] = v3[u];
        }
    }
}

// A particle's own cell in the halo and the distances from it to its
// cell's faces (for the same cell culling as the list builder).
struct BrickCell {
    int hc;                  // halo index of the own cell
    float lox, hix, loy, hiy, loz, hiz;
};

__device__ __forceinline__ BrickCell brick_cell(const float4& xi, const BrickDev& b, const Grid& g) {
    const int ax = cell_coord(xi.x, g.cs), ay = cell_coord(xi.y, g.cs), az = cell_coord(xi.z, g.cs);
    const int cx = min(max(ax - g.c0x, 1), g.nx - 2);
    const int cy = min(max(ay - g.c0y, 1), g.ny - 2);
    const int cz = min(max(az - g.c0z, 1), g.nz - 2);
    // the sort placed the particle in this brick: its cell is one of the 4^3
    const int hx = min(max(cx - b.x0 + 1, 1), 4), hy = min(max(cy - b.y0 + 1, 1), 4),
              hz = min(max(cz - b.z0 + 1, 1), 4);
    BrickCell c;
    c.hc = (hx * kHaloEdge + hy) * kHaloEdge + hz;
    c.lox = xi.x - (float)ax * g.cs;
    c.hix = (float)(ax + 1) * g.cs - xi.x;
    c.loy = xi.y - (float)ay * g.cs;
    c.hiy = (float)(ay + 1) * g.cs - xi.y;
    c.loz = xi.z - (float)az * g.cs;
    c.hiz = (float)(az + 1) * g.cs - xi.z;
    return c;
}

// Calls fn(j, slot, staged, xj, d2) for every candidate j != skip within h
// of xi, in the reference order.  Candidates are taken four at a time.
template <bool kAll = false, class Fn>
__device__ __forceinline__ void brick_walk(const Halo& H, const float4* lds, uint32_t cap,
                                           const float4* __restrict__ glob, const BrickCell& bc, const float4& xi,
                                           float h2, uint32_t skip, Fn&& fn) {
    const float h2m = h2 * 1.001f;  // the list builder's culling margin
    for (int dx = -1; dx <= 1; ++dx) {
        const float ex = dx < 0 ? fmaxf(bc.lox, 0.f) : (dx > 0 ? fmaxf(bc.hix, 0.f) : 0.f);
        for (int dy = -1; dy <= 1; ++dy) {
            const float ey = dy < 0 ? fmaxf(bc.loy, 0.f) : (dy > 0 ? fmaxf(bc.hiy, 0.f) : 0.f);
            const float exy = ex * ex + ey * ey;
            for (int dz = -1; dz <= 1; ++dz) {
                const float ez = dz < 0 ? fmaxf(bc.loz, 0.f) : (dz > 0 ? fmaxf(bc.hiz, 0.f) : 0.f);
                if (exy + ez * ez > h2m) continue;
                const int c = bc.hc + (dx * kHaloEdge + dy) * kHaloEdge + dz;
                const uint32_t o = H.off[c], e = H.off[c + 1], g0 = H.gs[c];
                const bool staged = kAll || e <= cap;
                for (uint32_t k0 = o; k0 < e; k0 += 4) {
                    float4 xj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t k = k0 + u < e ? k0 + u : k0;
                        if (kAll) xj[u] = lds[k];
                        else xj[u] = staged ? lds[k] : ld(glob, g0 + (k - o));
                    }
                    float d2[4];
                    bool hit[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t k = k0 + u;
                        d2[u] = dist2(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
                        hit[u] = k < e && g0 + (k - o) != skip && d2[u] <= h2;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (hit[u]) fn(g0 + (k0 + u - o), k0 + u, staged, xj[u], d2[u]);
                }
            }
        }
    }
}

// Particle range of a brick (the union of its occupied cells' ranges, which
// are contiguous: the brick is one Morton range of at most 64 cells); false
// when it holds no fluid.  Every thread of the block calls it (one barrier).
__device__ __forceinline__ bool brick_particles(const BrickDev& b, const uint32_t* __restrict__ rank_to_lin,
                                                const uint2* __restrict__ crange, uint32_t& p0, uint32_t& p1) {
    __shared__ uint32_t range[2];
    if (threadIdx.x < 64) {  // the block's first wavefront, all lanes active
        const uint32_t r = b.rank_lo + threadIdx.x;
        uint32_t lo = 0xffffffffu, hi = 0u;
        if (r <= b.rank_hi) {
            const uint2 cr = crange[rank_to_lin[r]];
            if (cr.y > cr.x) {
                lo = cr.x;
                hi = cr.y;
            }
        }
        lo = wave_min_u32(lo);
        hi = wave_max_u32(hi);
        if (threadIdx.x == 0) {
            range[0] = lo;
            range[1] = hi;
        }
    }
    __syncthreads();
    p0 = range[0];
    p1 = range[1];
    return p0 < p1;
}

// Density + Tait EOS of a brick's particles (the outputs of k_density).
__global__ __launch_bounds__(kBrickThreads) void k_brick_density(
    const BrickDev* __restrict__ bricks, uint32_t nbricks, const uint32_t* __restrict__ rank_to_lin,
    const SceneDev* __restrict__ sc, const float4* __restrict__ pos, float4* __restrict__ vel,
    const uint2* __restrict__ crange, const float4* __restrict__ bpos, const uint2* __restrict__ brange, Phys P,
    float4* __restrict__ xp, float* __restrict__ rho_out, float* __restrict__ p_out, DevStatus* st) {
    __shared__ Halo HF, HB;
    __shared__ float4 lf[kDensFluidCap];
    __shared__ float4 lb[kDensBoundCap];
    __shared__ uint32_t wsum[4];
    const uint32_t bi = blockIdx.x;  // round-robin over the XCDs: bricks with fluid are spread evenly
    if (bi >= nbricks) return;
    const BrickDev b = bricks[bi];
    uint32_t p0, p1;
    if (!brick_particles(b, rank_to_lin, crange, p0, p1)) return;  // uniform over the block
    const Grid g = sc[b.scene].g;
    const uint32_t c0 = sc[b.scene].c0;
    halo_ranges(b, g, c0, crange, HF, wsum);
    halo_ranges(b, g, c0, brange, HB, wsum);
    halo_stage<kDensFluidCap>(HF, pos, lf, nullptr, nullptr, nullptr, nullptr);
    halo_stage<kDensBoundCap>(HB, bpos, lb, nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
    uint32_t mf = 0, mb = 0;
    const bool all = HF.off[kHaloCells] <= kDensFluidCap && HB.off[kHaloCells] <= kDensBoundCap;  // uniform
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads) {
        const float4 xi = pos[i];
        const BrickCell bc = brick_cell(xi, b, g);
        float d = P.w0;
        uint32_t cf = 0, cb = 0;
        auto fluid = [&](uint32_t, uint32_t, bool, const float4&, float d2) {
            d = add_unfused(d, cubic_w(fsqrt(d2) * P.inv_h, P));
            ++cf;
        };
        auto bound = [&](uint32_t, uint32_t, bool, const float4& xb, float d2) {
            d = add_unfused(d, cubic_w(fsqrt(d2) * P.inv_h, P) * xb.w);
            ++cb;
        };
        if (all) brick_walk<true>(HF, lf, kDensFluidCap, pos, bc, xi, P.h2, i, fluid);
        else brick_walk<false>(HF, lf, kDensFluidCap, pos, bc, xi, P.h2, i, fluid);
        d = d * P.mass;
        if (all) brick_walk<true>(HB, lb, kDensBoundCap, bpos, bc, xi, P.h2, 0xffffffffu, bound);
        else brick_walk<false>(HB, lb, kDensBoundCap, bpos, bc, xi, P.h2, 0xffffffffu, bound);
        store_density(P, i, xi, d, vel, xp, rho_out, p_out);
        mf = max(mf, cf);
        mb = max(mb, cb);
    }
    mf = wave_max_u32(mf);
    mb = wave_max_u32(mb);
    if ((threadIdx.x & 63) == 0) {
        if (mf) atomic_max_if_above(&st->max_fluid, mf);
        if (mb) atomic_max_if_above(&st->max_boundary, mb);
    }
}

// Normals for surface tension (the outputs of k_normals).
__global__ __launch_bounds__(kBrickThreads) void k_brick_normals(
    const BrickDev* __restrict__ bricks, uint32_t nbricks, const uint32_t* __restrict__ rank_to_lin,
    const SceneDev* __restrict__ sc, const float4* __restrict__ xp, const float4* __restrict__ vel,
    const uint2* __restrict__ crange, Phys P, float4* __restrict__ nrm) {
    __shared__ Halo HF;
    __shared__ float4 lx[kForceFluidCap];
    __shared__ float4 lv[kForceFluidCap];
    __shared__ uint32_t wsum[4];
    const uint32_t bi = blockIdx.x;  // round-robin over the XCDs: bricks with fluid are spread evenly
    if (bi >= nbricks) return;
    const BrickDev b = bricks[bi];
    uint32_t p0, p1;
    if (!brick_particles(b, rank_to_lin, crange, p0, p1)) return;
    const Grid g = sc[b.scene].g;
    halo_ranges(b, g, sc[b.scene].c0, crange, HF, wsum);
    halo_stage<kForceFluidCap>(HF, xp, lx, vel, lv, nullptr, nullptr);
    __syncthreads();
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads) {
        const float4 xi = xp[i];
        const BrickCell bc = brick_cell(xi, b, g);
        float nx = 0.f, ny = 0.f, nz = 0.f;
        brick_walk(HF, lx, kForceFluidCap, xp, bc, xi, P.h2, i,
                   [&](uint32_t j, uint32_t k, bool staged, const float4& xj, float) {
                       add_normal(nx, ny, nz, P, xi, xj, staged ? lv[k].w : ld(vel, j).w);
                   });
        const float f = P.h * P.mass;
        nrm[i] = make_float4(f * nx, f * ny, f * nz, 0.f);
    }
}

// Forces + symplectic Euler + box collisions + CFL max of a brick's
// particles (the outputs of k_force_integrate).
template <bool kNormals>
__global__ __launch_bounds__(kBrickThreads) void k_brick_force(
    const BrickDev* __restrict__ bricks, uint32_t nbricks, const uint32_t* __restrict__ rank_to_lin,
    const SceneDev* __restrict__ sc, SceneState* __restrict__ ss, const uint2* __restrict__ crange,
    const float4* __restrict__ xp, const float4* __restrict__ vel, const float4* __restrict__ nrm,
    const uint32_t* __restrict__ id_b, const float* __restrict__ p_in, const float4* __restrict__ bpos,
    const uint2* __restrict__ brange, Phys P, DevStatus* st, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ id_out) {
    __shared__ Halo HF, HB;
    __shared__ float4 lx[kForceFluidCap];
    if (st->flags & kOverflowFlags) return;  // uniform: see k_force_integrate
    __shared__ float4 lv[kForceFluidCap];
    __shared__ float4 ln[kNormals ? kForceFluidCap : 1];
    __shared__ float4 lb[kForceBoundCap];
    __shared__ uint32_t wsum[4];
    const uint32_t bi = blockIdx.x;  // round-robin over the XCDs: bricks with fluid are spread evenly
    if (bi >= nbricks) return;
    const BrickDev b = bricks[bi];
    uint32_t p0, p1;
    if (!brick_particles(b, rank_to_lin, crange, p0, p1)) return;
    const Grid g = sc[b.scene].g;
    const uint32_t c0 = sc[b.scene].c0;
    halo_ranges(b, g, c0, crange, HF, wsum);
    halo_ranges(b, g, c0, brange, HB, wsum);
    halo_stage<kForceFluidCap>(HF, xp, lx, vel, lv, kNormals ? nrm : nullptr, ln);
    halo_stage<kForceBoundCap>(HB, bpos, lb, nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
    const float dt = ss[b.scene].dt, dt_np = ss[b.scene].dt_np;
    double v2max = 0.0;
    bool anybad = false;
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads) {
        const float4 xi = xp[i];
        const float4 vi = vel[i];
        float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
        if (kNormals) ni = nrm[i];
        const BrickCell bc = brick_cell(xi, b, g);
        ForceSums F = zero_sums();
        brick_walk(HF, lx, kForceFluidCap, xp, bc, xi, P.h2, i,
                   [&](uint32_t j, uint32_t k, bool staged, const float4& xj, float) {
                       const float4 vj = staged ? lv[k] : ld(vel, j);
                       float4 nj = make_float4(0.f, 0.f, 0.f, 0.f);
                       if (kNormals) nj = staged ? ln[k] : ld(nrm, j);
                       add_fluid_pair<kNormals>(F, P, xi, vi, ni, xj, vj, nj);
                   });
        float apx = P.mass * F.pax, apy = P.mass * F.pay, apz = P.mass * F.paz;
        const float bterm = boundary_term(P, xi.w, p_in[i]);
        brick_walk(HB, lb, kForceBoundCap, bpos, bc, xi, P.h2, 0xffffffffu,
                   [&](uint32_t, uint32_t, bool, const float4& xb, float) {
                       add_boundary_pair(apx, apy, apz, P, xi, bterm, xb, true);
                   });
        float anx, any, anz;
        nonpressure_accel(P, xi, F, anx, any, anz);
        double v2 = 0.0;
        bool bad = false;
        float4 stored;
        integrate_store(P, sc[b.scene], dt, dt_np, xi, vi, apx, apy, apz, anx, any, anz, i, id_b[i], pos_out, vel_out,
                        id_out, nullptr, v2, bad, stored);
        v2max = fmax(v2max, v2);
        anybad = anybad || bad;
    }
    const double m = wave_max_f64(v2max);
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomic_max_if_above(&ss[b.scene].vmax2_bits, (unsigned long long)__double_as_longlong(m));
    if (anybad) atomicOr(&st->flags, 8u /* WCSPH_FLAG_NONFINITE */);
}

// ------------------------------------------------------------------
// 9. brick list form.  The neighbour search of the cell-loop form (one
//    workgroup per brick, halo staged in LDS, reference order), but each
//    lane records its hits in a particle-major list row [i * cap + k]:
//    entry = the neighbour's LDS slot in the force pass's staging when it
//    has one (slot < kForceFluidCap / kForceBoundCap), else kGlobalEntry |
//    sorted index.  The density is then summed over the lane's own row, and
//    the force pass, over the same bricks and the same halo order, gathers
//    neighbour data from LDS.  Lists of this form hold no global indices,
//    so the parity hook rebuilds reference lists when asked.
// ------------------------------------------------------------------
constexpr uint32_t kGlobalEntry = 0x80000000u;
// staging of the force pass from brick lists: 2 float4 per fluid candidate
// (3 with normals, then the smaller caps of the cell-loop force pass)
constexpr uint32_t kListForceFluidCap = 3072;
constexpr uint32_t kListForceBoundCap = 1024;

template <uint32_t kSlotCapF, uint32_t kSlotCapB, bool kAll>
__device__ __forceinline__ void brick_neighbors_particle(
    uint32_t i, const BrickDev& b, const Grid& g, const Halo& HF, const Halo& HB, const float4* lf, const float4* lb,
    const float4* __restrict__ pos, float4* __restrict__ vel, const float4* __restrict__ bpos, const Phys& P,
    uint32_t capf, uint32_t capb, uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount,
    uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount, float4* __restrict__ xp, float* __restrict__ rho_out,
    float* __restrict__ p_out, uint32_t& mf, uint32_t& mb) {
    const float4 xi = pos[i];
    const BrickCell bc = brick_cell(xi, b, g);
    uint32_t* __restrict__ rowf = nlist + (size_t)i * capf;
    uint32_t* __restrict__ rowb = blist + (size_t)i * capb;
    uint32_t cf = 0, cb = 0;
    brick_walk<kAll>(HF, lf, kDensFluidCap, pos, bc, xi, P.h2, i,
                     [&](uint32_t j, uint32_t k, bool staged, const float4&, float) {
                         if (cf < capf) rowf[cf] = (staged && k < kSlotCapF) ? k : (kGlobalEntry | j);
                         ++cf;
                     });
    brick_walk<kAll>(HB, lb, kDensBoundCap, bpos, bc, xi, P.h2, 0xffffffffu,
                     [&](uint32_t j, uint32_t k, bool staged, const float4&, float) {
                         if (cb < capb) rowb[cb] = (staged && k < kSlotCapB) ? k : (kGlobalEntry | j);
                         ++cb;
                     });
    // density over the row just written (the order of the list form)
    constexpr bool kNoFlags = kAll && kSlotCapF >= kDensFluidCap && kSlotCapB >= kDensBoundCap;
    float d = P.w0;
    const uint32_t nf = min(cf, capf), nbn = min(cb, capb);
    for (uint32_t k = 0; k < nf; ++k) {
        const uint32_t e = rowf[k];
        const float4 xj = (!kNoFlags && (e & kGlobalEntry)) ? ld(pos, e & ~kGlobalEntry) : lf[min(e, kDensFluidCap - 1)];
        d = add_unfused(d, cubic_w(fsqrt(dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z)) * P.inv_h, P));
    }
    d = d * P.mass;
    for (uint32_t k = 0; k < nbn; ++k) {
        const uint32_t e = rowb[k];
        const float4 xb = (!kNoFlags && (e & kGlobalEntry)) ? ld(bpos, e & ~kGlobalEntry) : lb[min(e, kDensBoundCap - 1)];
        d = add_unfused(d, cubic_w(fsqrt(dist2(xi.x - xb.x, xi.y - xb.y, xi.z - xb.z)) * P.inv_h, P) * xb.w);
    }
    store_density(P, i, xi, d, vel, xp, rho_out, p_out);
    ncount[i] = cf;
    bcount[i] = cb;
    mf = max(mf, cf);
    mb = max(mb, cb);
}

template <uint32_t kSlotCapF, uint32_t kSlotCapB>
__global__ __launch_bounds__(kBrickThreads) void k_brick_neighbors(
    const BrickDev* __restrict__ bricks, uint32_t nbricks, const uint32_t* __restrict__ rank_to_lin,
    const SceneDev* __restrict__ sc, const float4* __restrict__ pos, float4* __restrict__ vel,
    const uint2* __restrict__ crange, const float4* __restrict__ bpos, const uint2* __restrict__ brange, Phys P,
    uint32_t capf, uint32_t capb, uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount,
    uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount, float4* __restrict__ xp, float* __restrict__ rho_out,
    float* __restrict__ p_out, DevStatus* st) {
    __shared__ Halo HF, HB;
    __shared__ float4 lf[kDensFluidCap];
    __shared__ float4 lb[kDensBoundCap];
    __shared__ uint32_t wsum[4];
    const uint32_t bi = blockIdx.x;  // round-robin over the XCDs: bricks with fluid are spread evenly
    if (bi >= nbricks) return;
    const BrickDev b = bricks[bi];
    uint32_t p0, p1;
    if (!brick_particles(b, rank_to_lin, crange, p0, p1)) return;  // uniform over the block
    const Grid g = sc[b.scene].g;
    const uint32_t c0 = sc[b.scene].c0;
    halo_ranges(b, g, c0, crange, HF, wsum);
    halo_ranges(b, g, c0, brange, HB, wsum);
    halo_stage<kDensFluidCap>(HF, pos, lf, nullptr, nullptr, nullptr, nullptr);
    halo_stage<kDensBoundCap>(HB, bpos, lb, nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
    uint32_t mf = 0, mb = 0;
    if (HF.off[kHaloCells] <= kDensFluidCap && HB.off[kHaloCells] <= kDensBoundCap) {  // uniform
        for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads)
            brick_neighbors_particle<kSlotCapF, kSlotCapB, true>(i, b, g, HF, HB, lf, lb, pos, vel, bpos, P, capf, capb,
                                                                 nlist, ncount, blist, bcount, xp, rho_out, p_out, mf, mb);
    } else {
        for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads)
            brick_neighbors_particle<kSlotCapF, kSlotCapB, false>(i, b, g, HF, HB, lf, lb, pos, vel, bpos, P, capf, capb,
                                                                  nlist, ncount, blist, bcount, xp, rho_out, p_out, mf,
                                                                  mb);
    }
    mf = wave_max_u32(mf);
    mb = wave_max_u32(mb);
    if ((threadIdx.x & 63) == 0) {
        if (mf) atomic_max_if_above(&st->max_fluid, mf);
        if (mb) atomic_max_if_above(&st->max_boundary, mb);
        uint32_t fl = 0;
        if (mf > capf) fl |= 1u;  // WCSPH_FLAG_FLUID_OVERFLOW
        if (mb > capb) fl |= 2u;  // WCSPH_FLAG_BOUNDARY_OVERFLOW
        if (fl) atomicOr(&st->flags, fl);
    }
}

// Forces + integration from the brick lists, neighbour data from LDS.
template <bool kNormals>
struct ListForceCaps {
    static constexpr uint32_t kF = kNormals ? kForceFluidCap : kListForceFluidCap;
    static constexpr uint32_t kB = kNormals ? kForceBoundCap : kListForceBoundCap;
};

template <bool kNormals, bool kAll>
__device__ __forceinline__ void brick_force_particle(
    uint32_t i, const SceneDev& B, float dt, float dt_np, const float4* lx, const float4* lv, const float4* ln,
    const float4* lb, const float4* __restrict__ xp, const float4* __restrict__ vel, const float4* __restrict__ nrm,
    const uint32_t* __restrict__ id_b, const float* __restrict__ p_in, const float4* __restrict__ bpos,
    const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ blist,
    const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb, const Phys& P, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ id_out, double& v2max, bool& anybad) {
    constexpr uint32_t kF = ListForceCaps<kNormals>::kF, kB = ListForceCaps<kNormals>::kB;
    const float4 xi = xp[i];
    const float4 vi = vel[i];
    float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
    if (kNormals) ni = nrm[i];
    const uint32_t* __restrict__ rowf = nlist + (size_t)i * capf;
    const uint32_t* __restrict__ rowb = blist + (size_t)i * capb;
    const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
    ForceSums F = zero_sums();
    for (uint32_t k0 = 0; k0 < cf; k0 += 4) {
        const uint4 e4 = *reinterpret_cast<const uint4*>(rowf + k0);  // rows are 16-byte aligned (capf % 4 == 0)
        const uint32_t e[4] = {e4.x, e4.y, e4.z, e4.w};
        float4 xj[4], vj[4], nj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t ee = k0 + u < cf ? e[u] : 0u;
            nj[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (kAll) {  // every entry is a slot
                const uint32_t sl = min(ee, kF - 1);
                xj[u] = lx[sl];
                vj[u] = lv[sl];
                if (kNormals) nj[u] = ln[sl];
            } else if (ee & kGlobalEntry) {
                const uint32_t j = ee & ~kGlobalEntry;
                xj[u] = ld(xp, j);
                vj[u] = ld(vel, j);
                if (kNormals) nj[u] = ld(nrm, j);
            } else {
                const uint32_t sl = min(ee, kF - 1);
                xj[u] = lx[sl];
                vj[u] = lv[sl];
                if (kNormals) nj[u] = ln[sl];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k0 + u < cf) add_fluid_pair<kNormals>(F, P, xi, vi, ni, xj[u], vj[u], nj[u]);
    }
    float apx = P.mass * F.pax, apy = P.mass * F.pay, apz = P.mass * F.paz;
    const float bterm = boundary_term(P, xi.w, p_in[i]);
    for (uint32_t k = 0; k < cb; ++k) {
        const uint32_t e = rowb[k];
        const float4 xb = (!kAll && (e & kGlobalEntry)) ? ld(bpos, e & ~kGlobalEntry) : lb[min(e, kB - 1)];
        add_boundary_pair(apx, apy, apz, P, xi, bterm, xb, true);
    }
    float anx, any, anz;
    nonpressure_accel(P, xi, F, anx, any, anz);
    double v2 = 0.0;
    bool bad = false;
    float4 stored;
    integrate_store(P, B, dt, dt_np, xi, vi, apx, apy, apz, anx, any, anz, i, id_b[i], pos_out, vel_out, id_out,
                    nullptr, v2, bad, stored);
    v2max = fmax(v2max, v2);
    anybad = anybad || bad;
}

template <bool kNormals>
__global__ __launch_bounds__(kBrickThreads) void k_brick_force_lists(
    const BrickDev* __restrict__ bricks, uint32_t nbricks, const uint32_t* __restrict__ rank_to_lin,
    const SceneDev* __restrict__ sc, SceneState* __restrict__ ss, const uint2* __restrict__ crange,
    const float4* __restrict__ xp, const float4* __restrict__ vel, const float4* __restrict__ nrm,
    const uint32_t* __restrict__ id_b, const float* __restrict__ p_in, const float4* __restrict__ bpos,
    const uint2* __restrict__ brange, const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
    const uint32_t* __restrict__ blist, const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb, Phys P,
    DevStatus* st, float4* __restrict__ pos_out, float4* __restrict__ vel_out, uint32_t* __restrict__ id_out) {
    constexpr uint32_t kF = ListForceCaps<kNormals>::kF, kB = ListForceCaps<kNormals>::kB;
    if (st->flags & kOverflowFlags) return;  // uniform: see k_force_integrate
    __shared__ Halo HF, HB;
    __shared__ float4 lx[kF];
    __shared__ float4 lv[kF];
    __shared__ float4 ln[kNormals ? kF : 1];
    __shared__ float4 lb[kB];
    __shared__ uint32_t wsum[4];
    const uint32_t bi = blockIdx.x;  // round-robin over the XCDs: bricks with fluid are spread evenly
    if (bi >= nbricks) return;
    const BrickDev b = bricks[bi];
    uint32_t p0, p1;
    if (!brick_particles(b, rank_to_lin, crange, p0, p1)) return;
    const Grid g = sc[b.scene].g;
    const uint32_t c0 = sc[b.scene].c0;
    halo_ranges(b, g, c0, crange, HF, wsum);
    halo_ranges(b, g, c0, brange, HB, wsum);
    halo_stage<kF>(HF, xp, lx, vel, lv, kNormals ? nrm : nullptr, ln);
    halo_stage<kB>(HB, bpos, lb, nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
    const SceneDev& B = sc[b.scene];
    const float dt = ss[b.scene].dt, dt_np = ss[b.scene].dt_np;
    double v2max = 0.0;
    bool anybad = false;
    if (HF.off[kHaloCells] <= kF && HB.off[kHaloCells] <= kB) {  // uniform: the search stored slots only
        for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads)
            brick_force_particle<kNormals, true>(i, B, dt, dt_np, lx, lv, ln, lb, xp, vel, nrm, id_b, p_in, bpos, nlist,
                                                 ncount, blist, bcount, capf, capb, P, pos_out, vel_out, id_out, v2max,
                                                 anybad);
    } else {
        for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBrickThreads)
            brick_force_particle<kNormals, false>(i, B, dt, dt_np, lx, lv, ln, lb, xp, vel, nrm, id_b, p_in, bpos,
                                                  nlist, ncount, blist, bcount, capf, capb, P, pos_out, vel_out, id_out,
                                                  v2max, anybad);
    }
    const double m = wave_max_f64(v2max);
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomic_max_if_above(&ss[b.scene].vmax2_bits, (unsigned long long)__double_as_longlong(m));
    if (anybad) atomicOr(&st->flags, 8u /* WCSPH_FLAG_NONFINITE */);
}

// ------------------------------------------------------------------
// boundary mass Psi (compute_boundary_mass.cu:10-30) — once per scene
// ------------------------------------------------------------------
__global__ void k_boundary_mass(const float4* __restrict__ bsrc, uint32_t nb, const SceneDev* __restrict__ sc,
                                const uint32_t* __restrict__ bstart, uint32_t S, const uint2* __restrict__ brange,
                                Phys P, float4* __restrict__ bdst) {
    const uint32_t i = particle_index();
    if (i >= nb) return;
    const uint32_t sid = find_scene(bstart, S, i);
    const Grid g = sc[sid].bg;  // the boundary grid: cell size h, 27 cells
    const float4 xi = bsrc[i];
    int cx = min(max(cell_coord(xi.x, g.cs) - g.c0x, 1), g.nx - 2);
    int cy = min(max(cell_coord(xi.y, g.cs) - g.c0y, 1), g.ny - 2);
    int cz = min(max(cell_coord(xi.z, g.cs) - g.c0z, 1), g.nz - 2);
    const uint32_t lin = sc[sid].bc0 + ((uint32_t)cx * g.ny + cy) * g.nz + cz;
    float sum = P.w0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint32_t c = lin + dx * g.ny * g.nz + dy * g.nz + dz;
                const uint2 br = brange[c];
                for (uint32_t j = br.x; j < br.y; ++j) {
                    if (j == i) continue;
                    const float4 xj = bsrc[j];
                    const float d2 = dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
                    if (d2 <= P.h2) sum += cubic_w(sqrtf(d2) * P.inv_h, P);
                }
            }
    bdst[i] = make_float4(xi.x, xi.y, xi.z, P.rho0 / sum);
}

// ------------------------------------------------------------------
// Reverse mode of one step (benchmark physics: cubic-spline gradients,
// artificial viscosity, Akinci boundary pressure, one fixed dt, identity
// box).  The sort, the neighbour lists and the collision decisions of the
// step are replayed from a tape and held fixed.  All passes are gathers:
// particle i sums the derivatives of every pair term with respect to its
// own variables, so no atomics are needed (the lists are symmetric).
// Adjoint buffers are in the step's sorted order.
// ------------------------------------------------------------------

// d/dR of the cubic-spline gradient coefficient gc(R) (grad W = gc * r).
__device__ __forceinline__ float cubic_grad_prime(float R, const Phys& P) {
    const float q = R * P.inv_h;
    if (q < 1e-9f || q > 1.f) return 0.f;
    const float a = P.cubic_gfactor * P.inv_h * P.inv_h;  // 48 / (pi h^5)
    if (q <= 0.5f) return 3.f * a * P.inv_h;
    const float b = 1.f - q;
    const float ir = frcp(R);
    return a * P.h * (2.f * b * P.inv_h * ir + b * b * ir * ir);
}

// collisions and the symplectic update, reversed:
// y = x + dt v_new, v_new = v + dt a, (x_out, v_out) = collide(y, v_new)
__global__ void k_adj_integrate(const float4* __restrict__ lx_out, const float4* __restrict__ lv_out,
                                const uint32_t* __restrict__ coll, uint32_t n, float dt, float elasticity,
                                float4* __restrict__ ly, float4* __restrict__ lvn, float4* __restrict__ la) {
    const uint32_t i = particle_index();
    if (i >= n) return;
    const uint32_t c = coll[i];
    const float4 lx = lx_out[i], lv = lv_out[i];
    const float lxs[3] = {lx.x, lx.y, lx.z}, lvs[3] = {lv.x, lv.y, lv.z};
    float y[3], vn[3];
    for (int a = 0; a < 3; ++a) {
        const bool hit = (c >> (2 * a)) & 3u;
        y[a] = hit ? 0.f : lxs[a];
        vn[a] = (hit ? -elasticity * lvs[a] : lvs[a]) + dt * y[a];
    }
    ly[i] = make_float4(y[0], y[1], y[2], 0.f);
    lvn[i] = make_float4(vn[0], vn[1], vn[2], 0.f);
    la[i] = make_float4(dt * vn[0], dt * vn[1], dt * vn[2], 0.f);
}

// Pass A: pressure, viscosity and boundary pair terms.  Writes d L / d rho_i
// (viscosity part plus pressure through the Tait derivative), the position
// and velocity adjoints of the force terms, and accumulates d L / d alpha.
// x: sorted positions of the step; v: sorted velocities with w = rho;
// xp: w = p / rho^2; p: pressures.
__global__ void k_adj_pairs(const float4* __restrict__ x, const float4* __restrict__ v,
                            const float4* __restrict__ xp, const float* __restrict__ p, uint32_t n,
                            const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                            const float4* __restrict__ bpos, const uint32_t* __restrict__ blist,
                            const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb, Phys P,
                            float dcoef, const float4* __restrict__ la, float* __restrict__ lrho,
                            float4* __restrict__ gx, float4* __restrict__ gv, double* __restrict__ galpha) {
    const uint32_t i = particle_index();
    double ga = 0.0;
    if (i < n) {
        const float4 xi = x[i], vi = v[i], li = la[i];
        const float ri = vi.w, pi = p[i], Pi = xp[i].w;
        const float calpha = P.visc * dcoef;
        const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
        float gP = 0.f, gB = 0.f, grv = 0.f, gal = 0.f;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, w0 = 0.f, w1 = 0.f, w2 = 0.f;
        for (uint32_t k = 0; k < cf; k += 4) {
            uint32_t j[4];
            load_batch<4>(nlist, n, i, k, cf, i, j);
            float4 xj[4], vj[4], lj[4];
            float Pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xj[u] = ld(xp, j[u]);  // the sorted position with p/rho^2 in w: one gather for both
                Pj[u] = xj[u].w;
                vj[u] = ld(v, j[u]);
                lj[u] = ld(la, j[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k + u >= cf) continue;
                const float rx = xi.x - xj[u].x, ry = xi.y - xj[u].y, rz = xi.z - xj[u].z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                const float R = fsqrt(r2);
                const float gq = cubic_grad(R * P.inv_h, P), gpr = cubic_grad_prime(R, P);
                const float Lx = li.x - lj[u].x, Ly = li.y - lj[u].y, Lz = li.z - lj[u].z;
                const float t = Lx * rx + Ly * ry + Lz * rz;
                const float Sp = Pi + Pj[u];
                const float tg = R > 0.f ? t * gpr * frcp(R) : 0.f;
                gP += -P.mass * gq * t;
                g0 += -P.mass * Sp * (gq * Lx + tg * rx);
                g1 += -P.mass * Sp * (gq * Ly + tg * ry);
                g2 += -P.mass * Sp * (gq * Lz + tg * rz);
                const float vx = vi.x - vj[u].x, vy = vi.y - vj[u].y, vz = vi.z - vj[u].z;
                const float sv = vx * rx + vy * ry + vz * rz;
                if (sv < 0.f) {
                    const float iD = frcp(ri + vj[u].w), iE = frcp(r2 + P.visc_eps);
                    const float Kc = P.mass * calpha * iD;
                    const float f = Kc * sv * gq * t * iE;
                    gal += P.mass * dcoef * iD * sv * gq * t * iE;
                    grv += -f * iD;
                    const float cw = Kc * gq * t * iE;
                    w0 += cw * rx;
                    w1 += cw * ry;
                    w2 += cw * rz;
                    const float sg = R > 0.f ? sv * t * gpr * frcp(R) : 0.f;
                    const float c2 = 2.f * Kc * sv * gq * t * iE * iE;
                    g0 += Kc * iE * (gq * t * vx + sg * rx + sv * gq * Lx) - c2 * rx;
                    g1 += Kc * iE * (gq * t * vy + sg * ry + sv * gq * Ly) - c2 * ry;
                    g2 += Kc * iE * (gq * t * vz + sg * rz + sv * gq * Lz) - c2 * rz;
                }
            }
        }
        const float Bi = Pi + pi * P.inv_rho0_sq;
        for (uint32_t k = 0; k < cb; ++k) {
            const uint32_t b = blist[list_at(k, i, n)];
            const float4 xb = ld(bpos, b);
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            const float R = fsqrt(rx * rx + ry * ry + rz * rz);
            const float gq = cubic_grad(R * P.inv_h, P), gpr = cubic_grad_prime(R, P);
            const float t = li.x * rx + li.y * ry + li.z * rz;
            const float tg = R > 0.f ? t * gpr * frcp(R) : 0.f;
            gB += -xb.w * gq * t;
            g0 += -xb.w * Bi * (gq * li.x + tg * rx);
            g1 += -xb.w * Bi * (gq * li.y + tg * ry);
            g2 += -xb.w * Bi * (gq * li.z + tg * rz);
        }
        // Tait derivative, zero at both clamps
        const float ratio = ri / P.rho0;
        const float dp = (ratio > 1.f && ratio < P.max_ratio)
                             ? P.stiffness * P.exponent * powf(ratio, P.exponent - 1.f) / P.rho0
                             : 0.f;
        const float ir = 1.f / ri;
        const float dP = dp * ir * ir - 2.f * pi * ir * ir * ir;
        const float dB = dP + dp * P.inv_rho0_sq;
        lrho[i] = grv + gP * dP + gB * dB;
        gx[i] = make_float4(g0, g1, g2, 0.f);
        gv[i] = make_float4(w0, w1, w2, 0.f);
        ga = 0.5 * (double)gal;  // every unordered pair is seen from both ends
    }
    const double sum = block_sum_f64(ga);
    if (threadIdx.x == 0) galpha[blockIdx.x] += sum;  // this workgroup's slot, summed over the steps
}

// Pass B: the density sums, then the adjoints of the state before the step,
// scattered back to the order before the step's sort.
__global__ void k_adj_density(const float4* __restrict__ x, uint32_t n, const uint32_t* __restrict__ nlist,
                              const uint32_t* __restrict__ ncount, const float4* __restrict__ bpos,
                              const uint32_t* __restrict__ blist, const uint32_t* __restrict__ bcount,
                              uint32_t capf, uint32_t capb, Phys P, const float* __restrict__ lrho,
                              const float4* __restrict__ gx, const float4* __restrict__ gv,
                              const float4* __restrict__ ly, const float4* __restrict__ lvn,
                              const uint32_t* __restrict__ perm, float4* __restrict__ lx_in,
                              float4* __restrict__ lv_in) {
    const uint32_t i = particle_index();
    if (i >= n) return;
    const float4 xi = x[i];
    const float li = lrho[i];
    const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    for (uint32_t k = 0; k < cf; k += 4) {
        uint32_t j[4];
        load_batch<4>(nlist, n, i, k, cf, i, j);
        float4 xj[4];
        float lj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            xj[u] = ld(x, j[u]);
            lj[u] = ld(lrho, j[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k + u >= cf) continue;
            const float rx = xi.x - xj[u].x, ry = xi.y - xj[u].y, rz = xi.z - xj[u].z;
            const float R = fsqrt(rx * rx + ry * ry + rz * rz);
            const float c = P.mass * cubic_grad(R * P.inv_h, P) * (li + lj[u]);
            g0 += c * rx;
            g1 += c * ry;
            g2 += c * rz;
        }
    }
    for (uint32_t k = 0; k < cb; ++k) {
        const uint32_t b = blist[list_at(k, i, n)];
        const float4 xb = ld(bpos, b);
        const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
        const float R = fsqrt(rx * rx + ry * ry + rz * rz);
        const float c = xb.w * cubic_grad(R * P.inv_h, P) * li;
        g0 += c * rx;
        g1 += c * ry;
        g2 += c * rz;
    }
    const float4 a = gx[i], y = ly[i], vn = lvn[i], w = gv[i];
    const uint32_t dst = perm[i];
    lx_in[dst] = make_float4(y.x + a.x + g0, y.y + a.y + g1, y.z + a.z + g2, 0.f);
    lv_in[dst] = make_float4(vn.x + w.x, vn.y + w.y, vn.z + w.z, 0.f);
}

// out[id[k]] = in[k] (to creation order) and out[k] = in[id[k]] (from it)
// The per-particle arrays of one taped step, copied to its tape slot in one
// launch (sorted x, v with rho, x/p, pressure, permutation, list counts).
__global__ void k_tape_copy(const float4* __restrict__ pos, const float4* __restrict__ vel,
                            const float4* __restrict__ xp, const float* __restrict__ p,
                            const uint32_t* __restrict__ perm, const uint32_t* __restrict__ nc,
                            const uint32_t* __restrict__ bc, uint32_t n, float4* __restrict__ tpos,
                            float4* __restrict__ tvel, float4* __restrict__ txp, float* __restrict__ tp,
                            uint32_t* __restrict__ tperm, uint32_t* __restrict__ tnc, uint32_t* __restrict__ tbc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tpos[i] = pos[i];
    tvel[i] = vel[i];
    txp[i] = xp[i];
    tp[i] = p[i];
    tperm[i] = perm[i];
    tnc[i] = nc[i];
    tbc[i] = bc[i];
}

__global__ void k_to_id_order(const float4* __restrict__ in, const uint32_t* __restrict__ id, uint32_t n,
                              float4* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[id[k]] = in[k];
}
__global__ void k_from_id_order(const float4* __restrict__ in, const uint32_t* __restrict__ id, uint32_t n,
                                float4* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[id[k]];
}
__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = k;
}

// ------------------------------------------------------------------
// diagnostics: sum |v| and sum |v|^2 in FP64 (sph.cu:180-197)
// ------------------------------------------------------------------
__global__ void k_metrics(const float4* __restrict__ vel, uint32_t n, double* out) {
    const uint32_t i = particle_index();
    double a = 0.0, b = 0.0;
    if (i < n) {
        const float4 v = vel[i];
        const double vx = v.x, vy = v.y, vz = v.z;
        b = vx * vx + vy * vy + vz * vz;
        a = sqrt(b);
    }
    a = block_sum_f64(a);
    b = block_sum_f64(b);
    if (threadIdx.x == 0) {  // one slot pair per workgroup, summed on the host
        out[2 * blockIdx.x] = a;
        out[2 * blockIdx.x + 1] = b;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace wcsph_dev
