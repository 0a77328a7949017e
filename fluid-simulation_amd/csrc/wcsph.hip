// Host side of libwcsph: the C ABI declared in include/wcsph.h.
//
// One handle = S >= 1 independent scenes on the current HIP device, one
// stream, and every buffer of the step allocated once at creation (nothing
// is allocated per step).  Fluid particles, boundary particles and grid
// cells of all scenes are concatenated scene-major; one launch per kernel
// covers all scenes.  A step is a fixed sequence of kernels
// (wcsph_device.hpp) that can be replayed as a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/wcsph.h"
#include "wcsph_device.hpp"

using namespace wcsph_dev;

// Device memory of wcsph_autotune_noalloc: the state it saves across its
// trials lives here instead of in buffers allocated per call.  Every GPU
// the library runs on holds its own copy of the array until the process
// ends.  In host code `scratch` is only a symbol: its device address on the
// current GPU comes from hipGetSymbolAddress (scratch_address below), and
// only that address is ever passed to copies or kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

thread_local std::string g_create_error;

// Device address of `scratch` per HIP device, resolved once per device (with
// that device current) by the first wcsph_create on it.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

int scratch_address(int device, void** out, std::string& why) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return WCSPH_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return WCSPH_OK;
}

// Host view of one scene.
struct SceneHost {
    uint32_t p0 = 0, n = 0;     // fluid particles [p0, p0 + n)
    uint32_t b0 = 0, nb = 0;    // boundary particles [b0, b0 + nb)
    uint32_t c0 = 0, ncell = 0; // fluid-grid cells [c0, c0 + ncell)
    uint32_t bc0 = 0, nbcell = 0;  // boundary-grid cells [bc0, bc0 + nbcell)
    double lo[3] = {}, hi[3] = {};  // world extent the grids cover (box, domain, particles at creation)
    wcsph_box box{};
    SceneDev dev{};
};

}  // namespace

struct wcsph {
    wcsph_params prm{};
    Phys phys{};
    std::vector<SceneHost> scenes;
    uint32_t S = 0;
    uint32_t n = 0, nb = 0, ncell = 0, nparts = 0, nbcell = 0;
    wcsph_variant var{256, 128, 64, 1, WCSPH_NEIGHBORS_PARTICLE, 16384, 1.f};
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_hint = NAN;
    std::string err;

    // fluid state
    float4 *pos_a = nullptr, *vel_a = nullptr, *pos_b = nullptr, *vel_b = nullptr, *xp = nullptr,
           *nrm = nullptr;
    uint32_t *id_a = nullptr, *id_b = nullptr, *perm = nullptr;
    float *rho = nullptr, *prs = nullptr;
    // sort scratch
    uint32_t *cell = nullptr, *slot = nullptr, *tmp = nullptr;
    // grid
    uint32_t *cnt = nullptr, *rank_to_lin = nullptr, *lin_to_rank = nullptr, *partial = nullptr, *occ = nullptr;
    uint2 *crange = nullptr;  // (start, end) per fluid-grid cell (rebuilt every step)
    uint2 *brange = nullptr;  // (start, end) per boundary-grid cell (static)
    uint8_t* bnear = nullptr;  // per boundary-grid cell: a boundary particle in one of its 27 cells (static)
    // neighbour lists
    uint32_t *nlist = nullptr, *ncount = nullptr, *blist = nullptr, *bcount = nullptr;
    bool lists_valid = false;    // lists hold the last step's neighbours
    bool counts_valid = false;   // cnt / cell / slot count pos_a as it is (the last force pass counted it)
    bool per_scene = false;      // list builder per scene (SceneDev::form) instead of the variant's
    bool any_cell = false, any_particle = false;  // builders the scenes use, when per_scene
    // bricks of 4^3 cells (cell-loop form)
    std::vector<BrickDev> bricks_host;
    BrickDev* bricks = nullptr;
    uint32_t nbricks = 0;
    DevStatus* hook_status = nullptr;  // status of list builds made for wcsph_get_neighbors
    // boundary
    float4* bpos = nullptr;
    // coordinates as separate arrays (kFluidCoordPad / kBoundaryCoordPad
    // padding entries): the sorted fluid (written by every sort) and the
    // boundary (static)
    Coords fxyz{}, bxyz{};
    // scenes
    SceneDev* d_scenes = nullptr;
    SceneState* d_state = nullptr;
    uint32_t *pstart = nullptr, *bstart = nullptr, *cstart = nullptr;
    DevStatus* status = nullptr;
    double* metrics = nullptr;
    // steps enqueued by wcsph_step_async since the last sync, as runs of
    // (dt hint, count), and scene 0's step counter at that sync: what a sync
    // needs to run again the steps lost to a neighbour-list overflow
    std::vector<std::pair<float, uint32_t>> pending;
    uint64_t synced_steps = 0;
    // autotuner: timing events made at creation, and the device address of
    // the no-allocation scratch array on the handle's GPU
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    unsigned char* scratch = nullptr;

    std::vector<void*> allocations;

    // differentiable runs
    float4* pos0 = nullptr;         // creation positions (reset / grad runs)
    uint32_t* coll_out = nullptr;   // collision flags of the step being taped
    struct Grad {
        uint32_t steps = 0, every = 0, ncp = 0, tape_len = 0;
        float dt = 0.f;
        bool ready = false;
        float4 *ck_pos = nullptr, *ck_vel = nullptr;
        uint32_t* ck_id = nullptr;
        float4 *t_pos = nullptr, *t_vel = nullptr, *t_xp = nullptr;
        float* t_p = nullptr;
        uint32_t *t_perm = nullptr, *t_nc = nullptr, *t_bc = nullptr, *t_coll = nullptr, *t_nl = nullptr,
                 *t_bl = nullptr;
        float4 *lx = nullptr, *lv = nullptr, *lx2 = nullptr, *lv2 = nullptr, *ly = nullptr, *lvn = nullptr,
               *la = nullptr, *gx = nullptr, *gv = nullptr;
        float* lrho = nullptr;
        double* galpha = nullptr;
        uint32_t tape_capf = 0, tape_capb = 0;
        uint32_t* final_id = nullptr;  // particle order after the forward run's last step
        Phys phys{};                   // physics of the forward run (alpha included)
        uint64_t gen = 0;              // generation of the forward run the checkpoints hold
    } grad;
    uint64_t grad_runs = 0;  // forward runs started on this handle

    // profiling: events between consecutive kernels of eager steps
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_tag;  // kernel id that ends at event k (-1: start mark)
    size_t ev_used = 0;
    double kms[WCSPH_NUM_KERNELS] = {};
    uint64_t klaunch[WCSPH_NUM_KERNELS] = {};
};

namespace {

#define HIP_TRY(h, expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return e_ == hipErrorOutOfMemory ? WCSPH_ERR_OUT_OF_MEMORY : WCSPH_ERR_HIP;   \
        }                                                                                 \
    } while (0)

int fail(wcsph_t* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

template <class T>
int dalloc(wcsph_t* h, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        h->err = "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? WCSPH_ERR_OUT_OF_MEMORY : WCSPH_ERR_HIP;
    }
    h->allocations.push_back(q);
    *p = (T*)q;
    return WCSPH_OK;
}

void dfree(wcsph_t* h, void* p) {
    if (!p) return;
    auto it = std::find(h->allocations.begin(), h->allocations.end(), p);
    if (it != h->allocations.end()) h->allocations.erase(it);
    (void)hipFree(p);
}

// Reference Morton key of a cell coordinate (morton.cuh:7-34), spread by
// the standard magic-mask method.
uint64_t spread_bits21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

uint64_t morton_of(int cx, int cy, int cz) {
    const uint32_t off = 1u << 20;
    return spread_bits21((uint32_t)cx + off) | (spread_bits21((uint32_t)cy + off) << 1) |
           (spread_bits21((uint32_t)cz + off) << 2);
}

// Cell coordinate of a host-side position.  Coordinates beyond the 21-bit
// Morton cell range come back clamped just outside it (every caller refuses
// those), so the conversion to int is defined for any finite input.
inline int host_cell(float p, float cs) {
    const float c = std::floor(p / cs);
    constexpr float kOut = (float)(1 << 21);
    if (!(c > -kOut)) return -(1 << 21);
    if (!(c < kOut)) return 1 << 21;
    return (int)c;
}

bool finite4(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

void build_phys(const wcsph_params& p, Phys& P) {
    // derived constants formed in float as the reference's constexprs are
    // (common.cuh:18-23)
    const float r = p.particle_radius;
    const float s = 2.f * r;
    const float h = 2.f * s;
    const float vol = s * s * s * 0.8f;
    P.h = h;
    P.h2 = h * h;
    P.inv_h = 1.f / h;
    P.mass = p.rest_density * vol;
    P.cubic_factor = 8.f / (kPi * h * h * h);
    P.w0 = P.cubic_factor * 1.f;
    P.cubic_gfactor = 48.f / (kPi * h * h * h);
    P.spiky_gfactor = -45.f / (kPi * powf(h, 9.f));
    P.cohesion_factor = 32.f / (kPi * h * h * h);
    P.rho0 = p.rest_density;
    P.inv_rho0_sq = 1.f / (p.rest_density * p.rest_density);
    P.stiffness = p.stiffness;
    P.exponent = p.exponent;
    P.max_ratio = p.max_density_ratio;
    P.visc = p.viscosity;
    P.visc_eps = 0.01f * h * h;
    const double csound = std::sqrt((double)p.stiffness * (double)p.exponent / (double)p.rest_density);
    P.art_visc_coef = (float)(2.0 * (double)p.viscosity * (double)h * csound);
    P.surface_tension = p.surface_tension;
    P.gx = p.gravity[0];
    P.gy = p.gravity[1];
    P.gz = p.gravity[2];
    P.r = r;
    P.cbuf = p.collision_buffer;
    P.elasticity = p.elasticity;
    P.grad_kind = (int)p.pressure_gradient;
    P.visc_kind = (int)p.viscosity_model;
    P.n_rules = (int)std::min<uint32_t>(p.n_force_rules, WCSPH_MAX_FORCE_RULES);
    for (int k = 0; k < WCSPH_MAX_FORCE_RULES; ++k) P.rules[k] = p.force_rules[k];
}

// World-space extent the cell grid of a scene must cover: the box corners
// under its model, the boundary particles, the fluid particles and the
// optional domain region.  Non-finite particles fail when check is set
// (else they are skipped).
int scene_extent(const wcsph_scene& in, double lo[3], double hi[3], bool check, std::string& why) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    auto grow = [&](int a, double v) {
        lo[a] = std::min(lo[a], v);
        hi[a] = std::max(hi[a], v);
    };
    for (int c = 0; c < 8; ++c) {
        const float bx = (c & 1) ? in.box.max[0] : in.box.min[0];
        const float by = (c & 2) ? in.box.max[1] : in.box.min[1];
        const float bz = (c & 4) ? in.box.max[2] : in.box.min[2];
        const float* m = in.box.model;
        for (int a = 0; a < 3; ++a)
            grow(a, (double)m[a] * bx + (double)m[4 + a] * by + (double)m[8 + a] * bz + (double)m[12 + a]);
    }
    if (in.domain_max[0] > in.domain_min[0] && in.domain_max[1] > in.domain_min[1] &&
        in.domain_max[2] > in.domain_min[2]) {
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(in.domain_min[a]) || !std::isfinite(in.domain_max[a])) {
                why = "non-finite scene domain";
                return 1;
            }
            grow(a, in.domain_min[a]);
            grow(a, in.domain_max[a]);
        }
    }
    for (int set = 0; set < 2; ++set) {
        const uint32_t cnt = set ? in.n_fluid : in.n_boundary;
        const float* p = set ? in.fluid_pos : in.boundary_pos;
        for (uint32_t i = 0; i < cnt; ++i) {
            if (!finite4(p + 4 * (size_t)i)) {
                if (!check) continue;
                why = set ? "non-finite fluid position" : "non-finite boundary position";
                return 1;
            }
            for (int a = 0; a < 3; ++a) grow(a, p[4 * (size_t)i + a]);
        }
    }
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(lo[a]) || !std::isfinite(hi[a])) {
            why = "scene has no finite extent";
            return 1;
        }
    return 0;
}

void box_to_dev(const wcsph_box& box, SceneDev& d) {
    bool ident = true;
    for (int a = 0; a < 3; ++a) {
        d.bmin[a] = box.min[a];
        d.bmax[a] = box.max[a];
    }
    for (int i = 0; i < 16; ++i) {
        d.model[i] = box.model[i];
        d.model_inv[i] = box.model_inv[i];
        const float id = (i % 5 == 0) ? 1.f : 0.f;
        if (box.model[i] != id || box.model_inv[i] != id) ident = false;
    }
    d.box_identity = ident ? 1 : 0;
}

int check_params(const wcsph_params& p, std::string& why) {
    if (!(p.particle_radius > 0.f) || !std::isfinite(p.particle_radius)) { why = "particle_radius must be > 0"; return 1; }
    if (!(p.rest_density > 0.f)) { why = "rest_density must be > 0"; return 1; }
    if (!(p.max_density_ratio >= 1.f)) { why = "max_density_ratio must be >= 1"; return 1; }
    if (p.pressure_gradient > 1u) { why = "unknown pressure_gradient"; return 1; }
    if (p.viscosity_model > 1u) { why = "unknown viscosity_model"; return 1; }
    if (p.adaptive_dt && !(p.dt_min > 0.f && p.dt_max >= p.dt_min)) { why = "need 0 < dt_min <= dt_max"; return 1; }
    if (p.n_force_rules > WCSPH_MAX_FORCE_RULES) { why = "at most WCSPH_MAX_FORCE_RULES external force rules"; return 1; }
    for (uint32_t k = 0; k < p.n_force_rules; ++k) {
        const wcsph_force_rule& R = p.force_rules[k];
        if (R.region > WCSPH_REGION_XZ_DISC_Y_BELOW || R.kind > WCSPH_FORCE_SWIRL_Y_AXIS) {
            why = "unknown external force region or kind";
            return 1;
        }
        if (!std::isfinite(R.y_below) || !std::isfinite(R.radius2) || !std::isfinite(R.magnitude) ||
            !std::isfinite(R.value[0]) || !std::isfinite(R.value[1]) || !std::isfinite(R.value[2])) {
            why = "non-finite external force rule";
            return 1;
        }
    }
    return 0;
}

uint32_t grid_blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// Neighbour-list buffers for capacities capf / capb.  The new buffers are
// allocated before the old ones are freed: on failure the handle keeps its
// lists and capacities unchanged.
int alloc_lists(wcsph_t* h, uint32_t capf, uint32_t capb) {
    uint32_t *nl = nullptr, *bl = nullptr;
    int rc;
    if ((rc = dalloc(h, &nl, (size_t)capf * h->n))) return rc;
    if ((rc = dalloc(h, &bl, (size_t)capb * h->n))) {
        dfree(h, nl);
        return rc;
    }
    dfree(h, h->nlist);
    dfree(h, h->blist);
    h->nlist = nl;
    h->blist = bl;
    h->var.fluid_capacity = capf;
    h->var.boundary_capacity = capb;
    return WCSPH_OK;
}

void drop_graph(wcsph_t* h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_hint = NAN;
}

// Records an event after the kernel `tag` (or a start mark with tag -1).
void mark(wcsph_t* h, int tag) {
    if (!h->prof) return;
    if (h->ev_used == h->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            h->prof = false;
            return;
        }
        h->ev_pool.push_back(e);
        h->ev_tag.push_back(-1);
    }
    h->ev_tag[h->ev_used] = tag;
    (void)hipEventRecord(h->ev_pool[h->ev_used], h->stream);
    ++h->ev_used;
}

void fold_events(wcsph_t* h) {
    for (size_t k = 1; k < h->ev_used; ++k) {
        const int tag = h->ev_tag[k];
        if (tag < 0) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->ev_pool[k - 1], h->ev_pool[k]) == hipSuccess) {
            h->kms[tag] += ms;
            h->klaunch[tag] += 1;
        }
    }
    h->ev_used = 0;
}

// Enqueue the kernels of one step (all scenes) on h->stream.
// Handles of at most this many fluid particles start with the cell builder.
constexpr uint32_t kSmallHandle = 1u << 17;

CellCount cell_count_args(const wcsph_t* h) { return CellCount{h->lin_to_rank, h->cnt, h->cell, h->slot}; }

// Whether a step's force pass counts the cells of the positions it writes
// (k_force_integrate, the list forms) or the next step must count them.
bool force_counts(const wcsph_t* h, bool need_lists) {
    return need_lists ||
           (h->var.neighbor_mode != WCSPH_NEIGHBORS_CELL_LOOP && h->var.neighbor_mode != WCSPH_NEIGHBORS_BRICK);
}

void enqueue_step(wcsph_t* h, float hint, bool need_lists = false) {
    const uint32_t n = h->n, S = h->S;
    const uint32_t bs = h->var.block_size;
    const uint32_t gb = grid_blocks(n, bs);
    hipStream_t s = h->stream;
    DtRule rule;
    rule.hint = hint;
    rule.dt_min = h->prm.dt_min;
    rule.dt_max = h->prm.dt_max;
    rule.dt_np_max = h->prm.dt_np_max;
    rule.cfl_factor = h->prm.cfl_factor;
    rule.spacing = 2.f * h->prm.particle_radius;
    rule.adaptive = (int)h->prm.adaptive_dt;
    rule.split = (int)h->prm.split_np_dt;

    mark(h, -1);
    if (n == 0) {
        k_time_step<<<grid_blocks(S, 64), 64, 0, s>>>(h->d_state, S, rule);
        mark(h, WCSPH_K_TIME_STEP);
        return;
    }
    if (!h->counts_valid) {
        k_cell_count<<<gb, bs, 0, s>>>(h->pos_a, n, h->d_scenes, h->pstart, S, cell_count_args(h), h->status);
        mark(h, WCSPH_K_CELL_COUNT);
    }
    // the list forms' force pass counts the positions it writes for the next
    // step; the cell-loop and brick-list forms count at the start of each step
    h->counts_valid = force_counts(h, need_lists);
    k_scan_reduce<<<h->nparts, 256, 0, s>>>(h->cnt, h->ncell, h->partial, h->nparts, h->occ, h->crange, h->status);
    if (h->nparts <= kInlineScanTiles) {  // tile prefixes and time step inside k_scan_apply
        k_scan_apply<true><<<h->nparts, 256, 0, s>>>(h->cnt, h->rank_to_lin, h->ncell, n, h->partial, h->crange,
                                                     h->occ, h->status, h->d_state, S, rule);
    } else {
        k_scan_partials<<<1, 256, 0, s>>>(h->partial, h->nparts, n, h->status, h->d_state, S, rule);  // + time step
        k_scan_apply<false><<<h->nparts, 256, 0, s>>>(h->cnt, h->rank_to_lin, h->ncell, n, h->partial, h->crange,
                                                      h->occ, h->status, h->d_state, S, rule);
    }
    mark(h, WCSPH_K_SCAN);
    k_scatter<<<gb, bs, 0, s>>>(h->cell, h->slot, h->crange, n, h->tmp);
    mark(h, WCSPH_K_SCATTER);
    k_permute<<<gb, bs, 0, s>>>(h->cell, h->crange, h->tmp, n, h->pos_a, h->vel_a, h->id_a, h->pos_b,
                                 h->vel_b, h->id_b, h->perm, h->fxyz);
    mark(h, WCSPH_K_PERMUTE);
    const uint32_t capf = h->var.fluid_capacity, capb = h->var.boundary_capacity;
    if (h->var.neighbor_mode == WCSPH_NEIGHBORS_CELL_LOOP && !need_lists) {
        // cell-loop form: density and forces walk the cells from LDS, no lists
        h->lists_valid = false;
        const uint32_t nbk = h->nbricks;
        k_brick_density<<<nbk, kBrickThreads, 0, s>>>(h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->pos_b,
                                                        h->vel_b, h->crange, h->bpos, h->brange, h->phys, h->xp,
                                                        h->rho, h->prs, h->status);
        mark(h, WCSPH_K_DENSITY);
        if (h->prm.surface_tension != 0.f) {
            k_brick_normals<<<nbk, kBrickThreads, 0, s>>>(h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->xp,
                                                           h->vel_b, h->crange, h->phys, h->nrm);
            mark(h, WCSPH_K_NORMALS);
            k_brick_force<true><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->d_state, h->crange, h->xp, h->vel_b, h->nrm, h->id_b,
                h->prs, h->bpos, h->brange, h->phys, h->status, h->pos_a, h->vel_a, h->id_a);
        } else {
            k_brick_force<false><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->d_state, h->crange, h->xp, h->vel_b, h->nrm, h->id_b,
                h->prs, h->bpos, h->brange, h->phys, h->status, h->pos_a, h->vel_a, h->id_a);
        }
        mark(h, WCSPH_K_FORCE);
        return;
    }
    if (h->var.neighbor_mode == WCSPH_NEIGHBORS_BRICK && !need_lists) {
        // brick list form: LDS-staged search writing LDS-slot lists, density
        // from them, forces gathering from LDS
        h->lists_valid = false;
        const uint32_t nbk = h->nbricks;
        if (h->prm.surface_tension != 0.f)
            k_brick_neighbors<ListForceCaps<true>::kF, ListForceCaps<true>::kB><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->pos_b, h->vel_b, h->crange, h->bpos, h->brange,
                h->phys, capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->xp, h->rho, h->prs, h->status);
        else
            k_brick_neighbors<ListForceCaps<false>::kF, ListForceCaps<false>::kB><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->pos_b, h->vel_b, h->crange, h->bpos, h->brange,
                h->phys, capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->xp, h->rho, h->prs, h->status);
        mark(h, WCSPH_K_NEIGHBORS);
        if (h->prm.surface_tension != 0.f) {
            k_brick_normals<<<nbk, kBrickThreads, 0, s>>>(h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->xp,
                                                           h->vel_b, h->crange, h->phys, h->nrm);
            mark(h, WCSPH_K_NORMALS);
            k_brick_force_lists<true><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->d_state, h->crange, h->xp, h->vel_b, h->nrm, h->id_b,
                h->prs, h->bpos, h->brange, h->nlist, h->ncount, h->blist, h->bcount, capf, capb, h->phys, h->status,
                h->pos_a, h->vel_a, h->id_a);
        } else {
            k_brick_force_lists<false><<<nbk, kBrickThreads, 0, s>>>(
                h->bricks, nbk, h->rank_to_lin, h->d_scenes, h->d_state, h->crange, h->xp, h->vel_b, h->nrm, h->id_b,
                h->prs, h->bpos, h->brange, h->nlist, h->ncount, h->blist, h->bcount, capf, capb, h->phys, h->status,
                h->pos_a, h->vel_a, h->id_a);
        }
        mark(h, WCSPH_K_FORCE);
        return;
    }
    h->lists_valid = true;
    // The list builders this step runs: the variant's, or, when the handle
    // sets one per scene (wcsph_set_scene_neighbor_modes), each builder over
    // its own scenes.  Every builder writes the same lists.
    const bool per_scene = h->per_scene;
    const bool run_cell = per_scene ? h->any_cell : h->var.neighbor_mode == WCSPH_NEIGHBORS_CELL;
    const bool run_particle = per_scene ? h->any_particle : h->var.neighbor_mode != WCSPH_NEIGHBORS_CELL;
    const uint32_t only_cell = per_scene ? (uint32_t)WCSPH_NEIGHBORS_CELL : kAllForms;
    const uint32_t only_particle = per_scene ? (uint32_t)WCSPH_NEIGHBORS_PARTICLE : kAllForms;
    if (run_cell) {
        // capacities within the LDS stage: no entry is stored to HBM directly
        if (capf <= kCoopSlotsF && capb <= kCoopSlotsB)
            k_neighbors_coop<false><<<h->var.neighbor_groups, kCoopBlock, 0, s>>>(
                h->pos_b, n, h->d_scenes, h->cstart, S, h->crange, h->bpos, h->nb, h->brange, h->occ, h->phys.h2,
                capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->status, only_cell);
        else
            k_neighbors_coop<true><<<h->var.neighbor_groups, kCoopBlock, 0, s>>>(
                h->pos_b, n, h->d_scenes, h->cstart, S, h->crange, h->bpos, h->nb, h->brange, h->occ, h->phys.h2,
                capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->status, only_cell);
    }
    if (run_particle) {
        // the per-particle builder ends with each particle's density pass
        if (h->scenes[0].dev.g.R == 1)
            k_neighbors<true, 1><<<grid_blocks(n, kNeighborBlock), kNeighborBlock, 0, s>>>(
                h->pos_b, n, h->d_scenes, h->pstart, S, h->crange, h->bpos, h->brange, h->bnear, h->phys.h2, capf,
                capb, h->nlist, h->ncount, h->blist, h->bcount, h->status, h->phys, h->vel_b, h->xp, h->rho, h->prs,
                h->fxyz, h->bxyz, only_particle);
        else
            k_neighbors<true, 2><<<grid_blocks(n, kNeighborBlock), kNeighborBlock, 0, s>>>(
                h->pos_b, n, h->d_scenes, h->pstart, S, h->crange, h->bpos, h->brange, h->bnear, h->phys.h2, capf,
                capb, h->nlist, h->ncount, h->blist, h->bcount, h->status, h->phys, h->vel_b, h->xp, h->rho, h->prs,
                h->fxyz, h->bxyz, only_particle);
    }
    mark(h, WCSPH_K_NEIGHBORS);
    if (run_cell) {
        k_density<<<gb, bs, 0, s>>>(h->pos_b, h->vel_b, n, h->nlist, h->ncount, h->bpos, h->blist, h->bcount,
                                     capf, capb, h->phys, h->xp, h->rho, h->prs, h->d_scenes, h->pstart, S, only_cell);
        mark(h, WCSPH_K_DENSITY);
    }
    if (h->prm.surface_tension != 0.f) {
        k_normals<<<gb, bs, 0, s>>>(h->xp, h->vel_b, n, h->nlist, h->ncount, capf, h->phys, h->nrm);
        mark(h, WCSPH_K_NORMALS);
        k_force_integrate<true><<<gb, bs, 0, s>>>(h->xp, h->vel_b, h->id_b, n, h->nlist, h->ncount, h->bpos,
                                                  h->blist, h->bcount, capf, capb, h->prs, h->nrm, h->phys,
                                                  h->d_scenes, h->d_state, h->pstart, S, h->status, h->pos_a,
                                                  h->vel_a, h->id_a, h->coll_out, cell_count_args(h));
    } else if (h->phys.grad_kind == 0 && h->phys.visc_kind == 0) {  // the benchmark kinds
        k_force_integrate<false, true><<<gb, bs, 0, s>>>(h->xp, h->vel_b, h->id_b, n, h->nlist, h->ncount, h->bpos,
                                                         h->blist, h->bcount, capf, capb, h->prs, h->nrm, h->phys,
                                                         h->d_scenes, h->d_state, h->pstart, S, h->status, h->pos_a,
                                                         h->vel_a, h->id_a, h->coll_out, cell_count_args(h));
    } else {
        k_force_integrate<false><<<gb, bs, 0, s>>>(h->xp, h->vel_b, h->id_b, n, h->nlist, h->ncount, h->bpos,
                                                   h->blist, h->bcount, capf, capb, h->prs, h->nrm, h->phys,
                                                   h->d_scenes, h->d_state, h->pstart, S, h->status, h->pos_a,
                                                   h->vel_a, h->id_a, h->coll_out, cell_count_args(h));
    }
    mark(h, WCSPH_K_FORCE);
}

// Positions in pos_a were replaced (upload, restore of a saved state): the
// counts of the previous ones are cleared, and the next step counts again.
// The stream orders the clearing before any later step.
void uncount(wcsph_t* h) {
    if (h->counts_valid && h->ncell)
        k_fill_u32<<<grid_blocks(h->ncell, 256), 256, 0, h->stream>>>(h->cnt, 0u, h->ncell);
    h->counts_valid = false;
}

// Uploads positions/velocities of all scenes (concatenated), resets ids and
// each scene's max |v|^2 for the first CFL step.
int upload_fluid(wcsph_t* h, const float* pos, const float* vel) {
    const uint32_t n = h->n;
    std::vector<float> p4(4 * (size_t)n), v4(4 * (size_t)n, 0.f);
    std::vector<SceneState> st(h->S);
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipMemcpy(st.data(), h->d_state, sizeof(SceneState) * h->S, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < h->S; ++s) {
        const SceneHost& sc = h->scenes[s];
        const Grid& g = sc.dev.g;
        double vmax2 = 0.0;
        for (uint32_t i = sc.p0; i < sc.p0 + sc.n; ++i) {
            if (!finite4(pos + 4 * (size_t)i)) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "non-finite fluid position");
            for (int a = 0; a < 3; ++a) p4[4 * (size_t)i + a] = pos[4 * (size_t)i + a];
            p4[4 * (size_t)i + 3] = 1.f;
            // every particle must start inside its scene's cell grid
            for (int a = 0; a < 3; ++a) {
                const int c = host_cell(p4[4 * (size_t)i + a], g.cs) - (&g.c0x)[a];
                if (c < g.R || c > (&g.nx)[a] - 1 - g.R)
                    return fail(h, WCSPH_ERR_DOMAIN, "fluid particle outside its scene's cell grid");
            }
            if (vel) {
                if (!finite4(vel + 4 * (size_t)i))
                    return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "non-finite fluid velocity");
                double v2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    v4[4 * (size_t)i + a] = vel[4 * (size_t)i + a];
                    v2 += (double)vel[4 * (size_t)i + a] * (double)vel[4 * (size_t)i + a];
                }
                vmax2 = std::max(vmax2, v2);
            }
        }
        std::memcpy(&st[s].vmax2_bits, &vmax2, sizeof(double));
    }
    std::vector<uint32_t> ids(n);
    std::iota(ids.begin(), ids.end(), 0u);
    if (n) {
        HIP_TRY(h, hipMemcpyAsync(h->pos_a, p4.data(), p4.size() * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(h->vel_a, v4.data(), v4.size() * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(h, hipMemcpyAsync(h->id_a, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(h, hipMemcpyAsync(h->d_state, st.data(), sizeof(SceneState) * h->S, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    uncount(h);  // new positions
    return WCSPH_OK;
}

int status_to_code(wcsph_t* h, uint32_t flags) {
    if (flags & (WCSPH_FLAG_FLUID_OVERFLOW | WCSPH_FLAG_BOUNDARY_OVERFLOW))
        return fail(h, WCSPH_ERR_NEIGHBOR_OVERFLOW,
                    "neighbour list overflow: a particle had more neighbours than the list capacity "
                    "(raise fluid_capacity/boundary_capacity with wcsph_set_variant)");
    if (flags & WCSPH_FLAG_DOMAIN)
        return fail(h, WCSPH_ERR_DOMAIN, "a fluid particle left the scene's cell grid");
    if (flags & WCSPH_FLAG_NONFINITE) return fail(h, WCSPH_ERR_NONFINITE, "non-finite position or velocity");
    if (flags & WCSPH_FLAG_INTERNAL)
        return fail(h, WCSPH_ERR_HIP, "internal error: the sort's cell counts do not add up to the particle count");
    return WCSPH_OK;
}

// Walk range of a fluid grid of cell size cs: the smallest R with R * cs >= h,
// so every neighbour lies within R cells per axis.  For the reference's
// multipliers {0.5, 2/3, 1, 1.5, 2} this is its RANGE = (int)(h / CELL_SIZE
// + 0.5) (common.cuh:126) with the search-side cell size: 2, 2, 1, 1, 1.
int walk_range(float h, float cs) { return (int)std::ceil((double)h / (double)cs - 1e-9); }

// Cell grid over the extent [lo, hi] with cells of size cs and walk range R:
// R + 1 spare cells on each side, so every cell the extent touches is an
// interior cell ([R, n - 1 - R]) with one more to spare for the rounding of
// the box corners.
int make_grid(const double lo[3], const double hi[3], float cs, int R, Grid& g, uint64_t& ncell, std::string& why) {
    ncell = 1;
    for (int a = 0; a < 3; ++a) {
        const int c_lo = host_cell((float)lo[a], cs), c_hi = host_cell((float)hi[a], cs);
        const int c0 = c_lo - (R + 1);
        const int dim = c_hi - c_lo + 1 + 2 * (R + 1);
        if (c0 < -(1 << 20) || c0 + dim - 1 > (1 << 20) - 1) {
            why = "scene extends beyond the 21-bit Morton cell range";
            return 1;
        }
        (&g.c0x)[a] = c0;
        (&g.nx)[a] = dim;
        ncell *= (uint64_t)dim;
    }
    g.cs = cs;
    g.R = R;
    return 0;
}

// Cells of grid g (global cell offset c0) in reference Morton-key order
// (rank_to_lin[c0 + k] = global index of the cell of rank k) and, when
// `bricks` is given, the 4^3-cell bricks: runs of ranks sharing key >> 6 (the
// same 4^3 block of absolute cells), corner in grid coordinates.
void rank_cells(const Grid& g, uint32_t c0, uint32_t scene, uint32_t* rank_to_lin, std::vector<BrickDev>* bricks) {
    const uint64_t ncell = (uint64_t)g.nx * g.ny * g.nz;
    std::vector<std::pair<uint64_t, uint32_t>> kl(ncell);
    for (int x = 0; x < g.nx; ++x)
        for (int y = 0; y < g.ny; ++y)
            for (int z = 0; z < g.nz; ++z) {
                const uint32_t lin = ((uint32_t)x * g.ny + y) * g.nz + z;
                kl[lin] = {morton_of(g.c0x + x, g.c0y + y, g.c0z + z), lin};
            }
    std::sort(kl.begin(), kl.end());
    for (uint64_t k = 0; k < ncell; ++k) rank_to_lin[c0 + k] = c0 + kl[k].second;
    if (!bricks) return;
    for (uint64_t k = 0; k < ncell; ++k) {
        if (k == 0 || (kl[k].first >> 6) != (kl[k - 1].first >> 6)) {
            const uint32_t lin = kl[k].second;
            const int x = (int)(lin / ((uint32_t)g.ny * g.nz)), y = (int)((lin / g.nz) % g.ny), z = (int)(lin % g.nz);
            auto corner = [](int local, int cz0) { return (((local + cz0) + (1 << 20)) & ~3) - (1 << 20) - cz0; };
            BrickDev b{};
            b.scene = scene;
            b.rank_lo = b.rank_hi = c0 + (uint32_t)k;
            b.x0 = corner(x, g.c0x);
            b.y0 = corner(y, g.c0y);
            b.z0 = corner(z, g.c0z);
            bricks->push_back(b);
        } else {
            bricks->back().rank_hi = c0 + (uint32_t)k;
        }
    }
}

// Boundary particles of one scene, Morton-sorted once (stable, cell size h:
// particle_data.cuh:84-110 with the multiplier at its initial 1), into
// bsorted, and the static ranges of its boundary-grid cells.
void sort_boundary(const wcsph_scene& in, const SceneHost& sc, std::vector<uint32_t>& brange,
                   std::vector<float>& bsorted) {
    const Grid& g = sc.dev.bg;
    std::vector<uint64_t> keys(in.n_boundary);
    std::vector<uint32_t> lins(in.n_boundary);
    for (uint32_t i = 0; i < in.n_boundary; ++i) {
        const float* p = in.boundary_pos + 4 * (size_t)i;
        const int c[3] = {host_cell(p[0], g.cs), host_cell(p[1], g.cs), host_cell(p[2], g.cs)};
        keys[i] = morton_of(c[0], c[1], c[2]);
        lins[i] = sc.bc0 + ((uint32_t)(c[0] - g.c0x) * g.ny + (uint32_t)(c[1] - g.c0y)) * g.nz + (uint32_t)(c[2] - g.c0z);
    }
    std::vector<uint32_t> ord(in.n_boundary);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint32_t k = 0; k < in.n_boundary; ++k) {
        const uint32_t i = ord[k];
        const size_t dst = 4 * ((size_t)sc.b0 + k);
        for (int a = 0; a < 3; ++a) bsorted[dst + a] = in.boundary_pos[4 * (size_t)i + a];
        bsorted[dst + 3] = 1.f;
        const uint32_t lin = lins[i];
        if (k == 0 || lins[ord[k - 1]] != lin) brange[2 * (size_t)lin] = sc.b0 + k;
        brange[2 * (size_t)lin + 1] = sc.b0 + k + 1;
    }
}

// (Re)builds the fluid cell grids of every scene for cell-size multiplier
// `mult`: cell size fp32(h) * fp32(mult), the product the reference's sort
// forms (particle_data.cuh:116), used here by the sort and the search
// alike; the cells in Morton-key order; the bricks (multiplier 1 only: the
// brick and cell forms read the boundary ranges through fluid-cell indices,
// which the two grids share only then); and every array indexed by fluid
// cell, emptied.  New buffers are allocated first: on failure the handle
// keeps its grids.  The stream is idle on entry.
int build_fluid_grid(wcsph_t* h, float mult) {
    const float cs = h->phys.h * mult;
    const int R = walk_range(h->phys.h, cs);
    if (R < 1 || R > 2) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "cell_size_mult must be in [0.5, 2]");
    const uint32_t S = h->S;
    std::vector<Grid> grids(S);
    std::vector<uint32_t> c0(S + 1);
    uint64_t total = 0;
    for (uint32_t sidx = 0; sidx < S; ++sidx) {
        uint64_t nc = 0;
        std::string why;
        if (make_grid(h->scenes[sidx].lo, h->scenes[sidx].hi, cs, R, grids[sidx], nc, why))
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT, why);
        c0[sidx] = (uint32_t)total;
        total += nc;
        if (total > kMaxCells)
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                        "cell grids too large (" + std::to_string(total) + " cells, at most 2^28)");
    }
    c0[S] = (uint32_t)total;
    std::vector<uint32_t> r2l(total), l2r(total);
    std::vector<BrickDev> bricks;
    for (uint32_t sidx = 0; sidx < S; ++sidx)
        rank_cells(grids[sidx], c0[sidx], sidx, r2l.data(), mult == 1.f ? &bricks : nullptr);
    for (uint64_t r = 0; r < total; ++r) l2r[r2l[r]] = (uint32_t)r;
    const uint32_t nparts = (uint32_t)((total + kScanTile - 1) / kScanTile);
    uint32_t *cnt = nullptr, *rk = nullptr, *lr = nullptr, *partial = nullptr, *occ = nullptr;
    uint2* crange = nullptr;
    BrickDev* bk = nullptr;
    auto undo = [&](int code) {
        const std::string why = h->err;
        for (void* q : {(void*)cnt, (void*)rk, (void*)lr, (void*)partial, (void*)occ, (void*)crange, (void*)bk})
            dfree(h, q);
        h->err = why;
        return code;
    };
    int rc;
    if ((rc = dalloc(h, &cnt, total)) || (rc = dalloc(h, &crange, total)) || (rc = dalloc(h, &rk, total)) ||
        (rc = dalloc(h, &lr, total)) || (rc = dalloc(h, &occ, total)) || (rc = dalloc(h, &partial, 2 * (size_t)nparts)) ||
        (rc = dalloc(h, &bk, bricks.size())))
        return undo(rc);
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes) HIP_TRY(h, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return WCSPH_OK;
    };
    if ((rc = up(rk, r2l.data(), 4 * total)) || (rc = up(lr, l2r.data(), 4 * total)) ||
        (rc = up(bk, bricks.data(), sizeof(BrickDev) * bricks.size())))
        return undo(rc);
    if (total) {
        k_fill_u32<<<grid_blocks(total, 256), 256, 0, h->stream>>>(cnt, 0u, total);
        k_fill_u32<<<grid_blocks(2 * total, 256), 256, 0, h->stream>>>((uint32_t*)crange, 0u, 2 * total);
    }
    const uint32_t no_occ = 0;  // the new ranges are empty: nothing for the scan to clear
    if (hipMemcpyAsync(&h->status->n_occ, &no_occ, 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        h->err = "set-up of the fluid cell grid failed";
        return undo(WCSPH_ERR_HIP);
    }
    // commit: the new grids replace the old ones
    for (void* q : {(void*)h->cnt, (void*)h->rank_to_lin, (void*)h->lin_to_rank, (void*)h->partial, (void*)h->occ,
                    (void*)h->crange, (void*)h->bricks})
        dfree(h, q);
    h->cnt = cnt;
    h->crange = crange;
    h->rank_to_lin = rk;
    h->lin_to_rank = lr;
    h->occ = occ;
    h->partial = partial;
    h->bricks = bk;
    h->bricks_host = bricks;
    h->nbricks = (uint32_t)bricks.size();
    h->ncell = (uint32_t)total;
    h->nparts = nparts;
    h->var.cell_size_mult = mult;
    h->counts_valid = false;  // the new counts are zero
    std::vector<SceneDev> sdev(S);
    for (uint32_t sidx = 0; sidx < S; ++sidx) {
        SceneHost& sc = h->scenes[sidx];
        sc.dev.g = grids[sidx];
        sc.dev.c0 = sc.c0 = c0[sidx];
        sc.ncell = c0[sidx + 1] - c0[sidx];
        sdev[sidx] = sc.dev;
    }
    if ((rc = up(h->d_scenes, sdev.data(), sizeof(SceneDev) * S)) || (rc = up(h->cstart, c0.data(), 4 * c0.size())))
        return rc;
    drop_graph(h);
    h->lists_valid = false;
    return WCSPH_OK;
}

}  // namespace

extern "C" {

int wcsph_default_params(int preset, wcsph_params* out) {
    if (!out) return WCSPH_ERR_INVALID_ARGUMENT;
    wcsph_params p{};
    p.rest_density = 1000.f;
    p.exponent = 7.f;
    p.max_density_ratio = 1.5f;
    p.gravity[0] = 0.f;
    p.gravity[1] = -9.81f;
    p.gravity[2] = 0.f;
    p.cfl_factor = 0.4f;
    p.dt_min = 0.00001f;
    p.dt_max = 0.0005f;
    p.dt_np_max = 0.015f;
    p.collision_buffer = 0.75f;
    p.elasticity = 0.5f;
    if (preset == WCSPH_PRESET_BENCHMARK) {
        p.particle_radius = 0.01f;
        p.stiffness = 1000.f;
        p.viscosity = 0.01f;
        p.surface_tension = 0.f;
        p.pressure_gradient = WCSPH_GRAD_CUBIC;
        p.viscosity_model = WCSPH_VISC_ARTIFICIAL;
        p.adaptive_dt = 0;
        p.split_np_dt = 0;
    } else if (preset == WCSPH_PRESET_REFERENCE) {
        p.particle_radius = 0.02f;
        p.stiffness = 0.1f;
        p.viscosity = 0.001f;
        p.surface_tension = 0.13f;
        p.pressure_gradient = WCSPH_GRAD_SPIKY_REF;
        p.viscosity_model = WCSPH_VISC_LAPLACIAN;
        p.adaptive_dt = 1;
        p.split_np_dt = 1;
    } else {
        return WCSPH_ERR_INVALID_ARGUMENT;
    }
    *out = p;
    return WCSPH_OK;
}

int wcsph_box_aabb(const float min[3], const float max[3], wcsph_box* out) {
    if (!min || !max || !out) return WCSPH_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    for (int a = 0; a < 3; ++a) {
        if (!(max[a] > min[a])) return WCSPH_ERR_INVALID_ARGUMENT;
        out->min[a] = min[a];
        out->max[a] = max[a];
    }
    for (int i = 0; i < 4; ++i) {
        out->model[5 * i] = 1.f;
        out->model_inv[5 * i] = 1.f;
    }
    return WCSPH_OK;
}

// Counter-based jitter in [-1, 1) for lattice site idx, axis a.
static float dam_jitter_unit(uint64_t seed, uint64_t idx, int axis) {
    uint64_t z = seed + (idx * 3u + (uint64_t)axis + 1u) * 0x9E3779B97F4A7C15ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
    const float two_u = 2.0f * u;
    return two_u - 1.0f;
}

constexpr uint32_t kMaxWallLayers = 64;

int wcsph_make_dam_break(float r, const uint32_t dims[3], const float lattice_min[3], float jitter,
                         uint64_t seed, const float box_min[3], const float box_max[3], uint32_t layers,
                         float* fluid_pos, uint32_t* n_fluid, float* boundary_pos, uint32_t* n_boundary) {
    if (!dims || !lattice_min || !box_min || !box_max || !n_fluid || !n_boundary || !(r > 0.f) ||
        !std::isfinite(r) || !(jitter >= 0.f) || !std::isfinite(jitter) || layers > kMaxWallLayers)
        return WCSPH_ERR_INVALID_ARGUMENT;
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(lattice_min[a]) || !std::isfinite(box_min[a]) || !std::isfinite(box_max[a]) ||
            !(box_max[a] > box_min[a]))
            return WCSPH_ERR_INVALID_ARGUMENT;
    const uint64_t nf = (uint64_t)dims[0] * dims[1] * dims[2];
    if (nf > 0xffffffffull) return WCSPH_ERR_INVALID_ARGUMENT;
    const float s = 2.f * r;
    // wall sites per axis: `layers` below, the interior sites, `layers` above
    // (at most the 21-bit Morton range of interior sites); the count is
    // checked before anything is written
    uint32_t interior[3];
    uint64_t sites[3];
    for (int a = 0; a < 3; ++a) {
        const double ni = std::round((double)(box_max[a] - box_min[a]) / (double)s);
        if (!(ni <= (double)(1u << 21))) return WCSPH_ERR_INVALID_ARGUMENT;
        interior[a] = (uint32_t)ni;
        sites[a] = interior[a] + 2ull * layers;
    }
    const uint64_t nw = sites[0] * sites[1] * sites[2] - (uint64_t)interior[0] * interior[1] * interior[2];
    if (nw > 0xffffffffull) return WCSPH_ERR_INVALID_ARGUMENT;
    if (fluid_pos) {
        uint64_t i = 0;
        for (uint32_t ix = 0; ix < dims[0]; ++ix)
            for (uint32_t iy = 0; iy < dims[1]; ++iy)
                for (uint32_t iz = 0; iz < dims[2]; ++iz, ++i) {
                    const uint32_t idx[3] = {ix, iy, iz};
                    for (int a = 0; a < 3; ++a) {
                        const float off = (float)idx[a] * s;
                        const float site = lattice_min[a] + off;
                        const float jit = jitter * dam_jitter_unit(seed, i, a);
                        fluid_pos[4 * i + a] = site + jit;
                    }
                    fluid_pos[4 * i + 3] = 1.f;
                }
    }
    *n_fluid = (uint32_t)nf;
    // Wall shells: per axis, `layers` coordinates below the box at
    // min - r - s*l, the interior sites min + r + s*m, and `layers` above at
    // max + r + s*l.  A lattice point is a wall particle when any coordinate
    // is a shell coordinate.
    std::vector<float> axis[3];
    for (int a = 0; a < 3; ++a) {
        for (int l = (int)layers - 1; l >= 0; --l) {
            const float off = (float)l * s;
            const float lo = box_min[a] - r;
            axis[a].push_back(lo - off);
        }
        for (uint32_t m = 0; m < interior[a]; ++m) {
            const float off = (float)m * s;
            const float lo = box_min[a] + r;
            axis[a].push_back(lo + off);
        }
        for (uint32_t l = 0; l < layers; ++l) {
            const float off = (float)l * s;
            const float hi = box_max[a] + r;
            axis[a].push_back(hi + off);
        }
    }
    if (boundary_pos) {
        uint64_t w = 0;
        const uint32_t nz = (uint32_t)sites[2];
        for (uint32_t x = 0; x < axis[0].size(); ++x) {
            const bool inx = x >= layers && x < layers + interior[0];
            for (uint32_t y = 0; y < axis[1].size(); ++y) {
                const bool iny = y >= layers && y < layers + interior[1];
                for (uint32_t z = 0; z < nz; ++z) {
                    // inside the walls in x and y: skip the interior z sites
                    if (inx && iny && z == layers && interior[2]) z = layers + interior[2];
                    if (z >= nz) break;
                    boundary_pos[4 * w] = axis[0][x];
                    boundary_pos[4 * w + 1] = axis[1][y];
                    boundary_pos[4 * w + 2] = axis[2][z];
                    boundary_pos[4 * w + 3] = 1.f;
                    ++w;
                }
            }
        }
    }
    *n_boundary = (uint32_t)nw;
    return WCSPH_OK;
}

const char* wcsph_last_error(const wcsph_t* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int wcsph_create_batch(const wcsph_params* params, const wcsph_scene* scenes, uint32_t n_scenes, wcsph_t** out) {
    g_create_error.clear();
    if (!params || !scenes || !out || n_scenes == 0) {
        g_create_error = "wcsph_create: NULL argument or no scene";
        return WCSPH_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::string why;
    if (check_params(*params, why)) {
        g_create_error = "wcsph_create: " + why;
        return WCSPH_ERR_INVALID_ARGUMENT;
    }
    uint64_t nf = 0, nbt = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        if ((scenes[s].n_fluid && !scenes[s].fluid_pos) || (scenes[s].n_boundary && !scenes[s].boundary_pos)) {
            g_create_error = "wcsph_create: missing particle arrays";
            return WCSPH_ERR_INVALID_ARGUMENT;
        }
        nf += scenes[s].n_fluid;
        nbt += scenes[s].n_boundary;
    }
    if (nf > (1u << 27) || nbt > (1u << 27)) {
        g_create_error = "wcsph_create: at most 2^27 fluid and 2^27 boundary particles per handle";
        return WCSPH_ERR_INVALID_ARGUMENT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "wcsph_create: no HIP device";
        return WCSPH_ERR_NO_DEVICE;
    }
    wcsph_t* h = new wcsph_t();
    h->prm = *params;
    h->S = n_scenes;
    h->n = (uint32_t)nf;
    h->nb = (uint32_t)nbt;
    // Default search form by size: up to 131,072 fluid particles the
    // cell-cooperative builder, at about one occupied cell per workgroup,
    // has the shorter latency chain; above that the per-particle builder
    // is faster.
    if (h->n <= kSmallHandle) {
        h->var.neighbor_mode = WCSPH_NEIGHBORS_CELL;
        h->var.neighbor_groups = std::min(65536u, std::max(1024u, (h->n + 7u) / 8u));
    }
    build_phys(h->prm, h->phys);
    auto bail = [&](int code) {
        g_create_error = h->err;
        wcsph_destroy(h);
        return code;
    };

    // scene layout and the world extent each scene's grids cover
    h->scenes.resize(n_scenes);
    uint32_t p0 = 0, b0 = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        SceneHost& sc = h->scenes[s];
        sc.p0 = p0;
        sc.n = scenes[s].n_fluid;
        sc.b0 = b0;
        sc.nb = scenes[s].n_boundary;
        p0 += scenes[s].n_fluid;
        b0 += scenes[s].n_boundary;
        if (scene_extent(scenes[s], sc.lo, sc.hi, true, h->err)) return bail(WCSPH_ERR_INVALID_ARGUMENT);
        sc.box = scenes[s].box;
        box_to_dev(scenes[s].box, sc.dev);
    }
    // boundary grids: cell size h, walk range 1, fixed for the handle's life
    uint64_t btotal = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        SceneHost& sc = h->scenes[s];
        uint64_t nc = 0;
        if (make_grid(sc.lo, sc.hi, h->phys.h, 1, sc.dev.bg, nc, h->err)) return bail(WCSPH_ERR_INVALID_ARGUMENT);
        sc.bc0 = sc.dev.bc0 = (uint32_t)btotal;
        sc.nbcell = (uint32_t)nc;
        btotal += nc;
        if (btotal > kMaxCells) {
            h->err = "cell grids too large (" + std::to_string(btotal) + " cells, at most 2^28)";
            return bail(WCSPH_ERR_INVALID_ARGUMENT);
        }
    }
    h->nbcell = (uint32_t)btotal;
    std::vector<uint32_t> brange_h(2 * (size_t)btotal, 0);
    std::vector<float> bsorted(4 * (size_t)h->nb);
    for (uint32_t s = 0; s < n_scenes; ++s) sort_boundary(scenes[s], h->scenes[s], brange_h, bsorted);
    // boundary cells with a boundary particle in their 27-cell neighbourhood
    // (boundary particles never move): the list builder skips the boundary
    // walk of wavefronts that have none
    std::vector<uint8_t> bnear_h(btotal, 0);
    for (uint32_t s = 0; s < n_scenes; ++s) {
        const SceneHost& sc = h->scenes[s];
        const Grid& g = sc.dev.bg;
        for (int x = 0; x < g.nx; ++x)
            for (int y = 0; y < g.ny; ++y)
                for (int z = 0; z < g.nz; ++z) {
                    const size_t lin = sc.bc0 + ((size_t)x * g.ny + y) * g.nz + z;
                    if (brange_h[2 * lin + 1] <= brange_h[2 * lin]) continue;
                    for (int dx = -1; dx <= 1; ++dx)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dz = -1; dz <= 1; ++dz) {
                                const int a = x + dx, b = y + dy, c = z + dz;
                                if (a < 0 || a >= g.nx || b < 0 || b >= g.ny || c < 0 || c >= g.nz) continue;
                                bnear_h[sc.bc0 + ((size_t)a * g.ny + b) * g.nz + c] = 1;
                            }
                }
    }
    std::vector<uint32_t> ps(n_scenes + 1), bst(n_scenes + 1);
    for (uint32_t s = 0; s < n_scenes; ++s) {
        ps[s] = h->scenes[s].p0;
        bst[s] = h->scenes[s].b0;
    }
    ps[n_scenes] = h->n;
    bst[n_scenes] = h->nb;

    // --- device buffers --------------------------------------------------
    int rc;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->err = "hipStreamCreateWithFlags failed";
        return bail(WCSPH_ERR_HIP);
    }
    if (hipEventCreate(&h->tune_ev[0]) != hipSuccess || hipEventCreate(&h->tune_ev[1]) != hipSuccess) {
        h->err = "hipEventCreate failed";
        return bail(WCSPH_ERR_HIP);
    }
    {
        int dev = 0;
        void* p = nullptr;
        if (hipGetDevice(&dev) != hipSuccess) {
            h->err = "hipGetDevice failed";
            return bail(WCSPH_ERR_HIP);
        }
        if (int rc0 = scratch_address(dev, &p, h->err)) return bail(rc0);
        h->scratch = (unsigned char*)p;
    }
    const size_t n = h->n, nbcell = h->nbcell;
    if ((rc = dalloc(h, &h->pos_a, n)) || (rc = dalloc(h, &h->vel_a, n)) || (rc = dalloc(h, &h->pos_b, n + 4)) ||
        (rc = dalloc(h, &h->vel_b, n)) || (rc = dalloc(h, &h->xp, n)) || (rc = dalloc(h, &h->id_a, n)) ||
        (rc = dalloc(h, &h->id_b, n)) || (rc = dalloc(h, &h->perm, n)) || (rc = dalloc(h, &h->rho, n)) ||
        (rc = dalloc(h, &h->prs, n)) || (rc = dalloc(h, &h->cell, n)) || (rc = dalloc(h, &h->slot, n)) ||
        (rc = dalloc(h, &h->tmp, n + 4)) || (rc = dalloc(h, &h->ncount, n)) || (rc = dalloc(h, &h->bcount, n)) ||
        (rc = dalloc(h, &h->brange, nbcell)) || (rc = dalloc(h, &h->bnear, nbcell)) ||
        (rc = dalloc(h, &h->bpos, h->nb + 4)) || (rc = dalloc(h, &h->status, 1)) ||
        (rc = dalloc(h, &h->metrics, 2 * (size_t)grid_blocks(h->n, 256))) || (rc = dalloc(h, &h->d_scenes, n_scenes)) ||
        (rc = dalloc(h, &h->d_state, n_scenes)) || (rc = dalloc(h, &h->pstart, n_scenes + 1)) ||
        (rc = dalloc(h, &h->bstart, n_scenes + 1)) || (rc = dalloc(h, &h->cstart, n_scenes + 1)))
        return bail(rc);
    if (h->prm.surface_tension != 0.f && (rc = dalloc(h, &h->nrm, n))) return bail(rc);
    const size_t nfc = n + kFluidCoordPad, nbc = (size_t)h->nb + kBoundaryCoordPad;
    if ((rc = dalloc(h, &h->fxyz.x, nfc)) || (rc = dalloc(h, &h->fxyz.y, nfc)) ||
        (rc = dalloc(h, &h->fxyz.z, nfc)) || (rc = dalloc(h, &h->bxyz.x, nbc)) ||
        (rc = dalloc(h, &h->bxyz.y, nbc)) || (rc = dalloc(h, &h->bxyz.z, nbc)))
        return bail(rc);
    if ((rc = alloc_lists(h, h->var.fluid_capacity, h->var.boundary_capacity))) return bail(rc);
    if ((rc = dalloc(h, &h->hook_status, 1))) return bail(rc);

    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (!bytes) return WCSPH_OK;
        HIP_TRY(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
        return WCSPH_OK;
    };
    DevStatus st0{};
    std::vector<SceneState> ss0(n_scenes);
    std::memset(ss0.data(), 0, sizeof(SceneState) * n_scenes);
    if ((rc = up(h->brange, brange_h.data(), 8 * nbcell)) || (rc = up(h->bnear, bnear_h.data(), nbcell)) ||
        (rc = up(h->status, &st0, sizeof(st0))) ||
        (rc = up(h->d_state, ss0.data(), sizeof(SceneState) * n_scenes)) ||
        (rc = up(h->pstart, ps.data(), 4 * ps.size())) || (rc = up(h->bstart, bst.data(), 4 * bst.size())) ||
        (rc = up(h->hook_status, &st0, sizeof(st0))))
        return bail(rc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        h->err = "upload of the scene tables failed";
        return bail(WCSPH_ERR_HIP);
    }
    // the fluid cell grids (multiplier 1) and the scene table with both grids
    if ((rc = build_fluid_grid(h, 1.f))) return bail(rc);
    auto fill = [&](void* p, size_t count) {
        if (count) k_fill_u32<<<grid_blocks(count, 256), 256, 0, h->stream>>>((uint32_t*)p, 0u, count);
    };
    fill(h->ncount, n);
    fill(h->bcount, n);
    fill(h->perm, n);
    fill(h->pos_b, 4 * n);
    for (float* a : {h->fxyz.x, h->fxyz.y, h->fxyz.z}) fill(a, nfc);
    {
        // the boundary's coordinate arrays, in its sorted order (padding zeroed)
        std::vector<float> c(3 * nbc, 0.f);
        for (size_t k = 0; k < h->nb; ++k)
            for (int a = 0; a < 3; ++a) c[a * nbc + k] = bsorted[4 * k + a];
        float* dst[3] = {h->bxyz.x, h->bxyz.y, h->bxyz.z};
        for (int a = 0; a < 3; ++a)
            if ((rc = up(dst[a], c.data() + a * nbc, 4 * nbc))) return bail(rc);
        if (hipStreamSynchronize(h->stream) != hipSuccess) {  // c is freed at the end of this block
            h->err = "upload of the boundary coordinates failed";
            return bail(WCSPH_ERR_HIP);
        }
    }
    fill(h->vel_b, 4 * n);
    fill(h->rho, n);
    fill(h->prs, n);

    // boundary mass Psi on the device, from a staging copy of the sorted walls
    if (h->nb) {
        float4* bsrc = nullptr;
        if ((rc = dalloc(h, &bsrc, h->nb))) return bail(rc);
        if ((rc = up(bsrc, bsorted.data(), 16 * (size_t)h->nb))) return bail(rc);
        k_boundary_mass<<<grid_blocks(h->nb, 256), 256, 0, h->stream>>>(bsrc, h->nb, h->d_scenes, h->bstart,
                                                                        n_scenes, h->brange, h->phys, h->bpos);
        if (hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
            h->err = "boundary mass kernel failed";
            return bail(WCSPH_ERR_HIP);
        }
        dfree(h, bsrc);
    }
    // fluid state: concatenation of the scenes
    {
        std::vector<float> pos(4 * n), vel(4 * n, 0.f);
        bool any_vel = false;
        for (uint32_t s = 0; s < n_scenes; ++s) {
            const SceneHost& sc = h->scenes[s];
            if (sc.n) std::memcpy(pos.data() + 4 * (size_t)sc.p0, scenes[s].fluid_pos, 16 * (size_t)sc.n);
            if (sc.n && scenes[s].fluid_vel) {
                std::memcpy(vel.data() + 4 * (size_t)sc.p0, scenes[s].fluid_vel, 16 * (size_t)sc.n);
                any_vel = true;
            }
        }
        if ((rc = upload_fluid(h, pos.data(), any_vel ? vel.data() : nullptr))) return bail(rc);
        if ((rc = dalloc(h, &h->pos0, n))) return bail(rc);
        if (n && hipMemcpy(h->pos0, h->pos_a, 16 * n, hipMemcpyDeviceToDevice) != hipSuccess) {
            h->err = "copy of the creation positions failed";
            return bail(WCSPH_ERR_HIP);
        }
    }
    if (hipGetLastError() != hipSuccess) {
        h->err = "kernel launch failed during set-up";
        return bail(WCSPH_ERR_HIP);
    }
    *out = h;
    return WCSPH_OK;
}

int wcsph_create(const wcsph_params* params, const wcsph_scene* scene, wcsph_t** out) {
    if (!scene) {
        g_create_error = "wcsph_create: NULL argument";
        return WCSPH_ERR_INVALID_ARGUMENT;
    }
    return wcsph_create_batch(params, scene, 1, out);
}

int wcsph_destroy(wcsph_t* h) {
    if (!h) return WCSPH_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : h->tune_ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : h->allocations) (void)hipFree(p);
    h->allocations.clear();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return WCSPH_OK;
}

int wcsph_set_variant(wcsph_t* h, const wcsph_variant* v) {
    if (!h || !v) return WCSPH_ERR_INVALID_ARGUMENT;
    if (v->block_size < 64 || v->block_size > 1024 || v->block_size % 64)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 in [64, 1024]");
    if (v->neighbor_mode > WCSPH_NEIGHBORS_BRICK)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "unknown neighbor_mode");
    if (v->neighbor_groups < 1 || v->neighbor_groups > 65536)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "neighbor_groups must be in [1, 65536]");
    if (v->fluid_capacity < 4 || v->fluid_capacity > 4096 || v->boundary_capacity < 4 ||
        v->boundary_capacity > 4096 || v->fluid_capacity % 4 || v->boundary_capacity % 4)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "list capacities must be multiples of 4 in [4, 4096]");
    const float mult = v->cell_size_mult == 0.f ? 1.f : v->cell_size_mult;
    if (!(mult >= 0.5f && mult <= 2.f)) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "cell_size_mult must be in [0.5, 2]");
    if (mult != 1.f && v->neighbor_mode != WCSPH_NEIGHBORS_PARTICLE)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                    "a cell_size_mult other than 1 needs the per-particle builder (WCSPH_NEIGHBORS_PARTICLE)");
    if (mult != 1.f && h->per_scene && h->any_cell)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                    "a cell_size_mult other than 1 needs the per-particle builder for every scene");
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    // a captured graph holds the buffers replaced below
    drop_graph(h);
    if (v->fluid_capacity != h->var.fluid_capacity || v->boundary_capacity != h->var.boundary_capacity) {
        // on failure the handle keeps its lists and variant
        int rc = alloc_lists(h, v->fluid_capacity, v->boundary_capacity);
        if (rc) return rc;
    }
    if (mult != h->var.cell_size_mult) {
        // on failure the handle keeps its grid and the rest of its variant
        // (with the new capacities, whose lists are in place)
        int rc = build_fluid_grid(h, mult);
        if (rc) return rc;
    }
    h->var = *v;
    h->var.cell_size_mult = mult;
    return WCSPH_OK;
}

int wcsph_get_variant(const wcsph_t* h, wcsph_variant* v) {
    if (!h || !v) return WCSPH_ERR_INVALID_ARGUMENT;
    *v = h->var;
    return WCSPH_OK;
}

int wcsph_set_scene_neighbor_modes(wcsph_t* h, const uint32_t* modes) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    bool cell = false, particle = false;
    if (modes) {
        for (uint32_t k = 0; k < h->S; ++k) {
            if (modes[k] != WCSPH_NEIGHBORS_CELL && modes[k] != WCSPH_NEIGHBORS_PARTICLE)
                return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                            "per-scene list builders are WCSPH_NEIGHBORS_CELL or WCSPH_NEIGHBORS_PARTICLE");
            cell = cell || modes[k] == WCSPH_NEIGHBORS_CELL;
            particle = particle || modes[k] == WCSPH_NEIGHBORS_PARTICLE;
        }
        if (cell && h->var.cell_size_mult != 1.f)
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                        "a cell_size_mult other than 1 needs the per-particle builder for every scene");
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    std::vector<SceneDev> sdev(h->S);
    for (uint32_t k = 0; k < h->S; ++k) {
        h->scenes[k].dev.form = modes ? modes[k] : 0u;
        sdev[k] = h->scenes[k].dev;
    }
    HIP_TRY(h, hipMemcpy(h->d_scenes, sdev.data(), sizeof(SceneDev) * h->S, hipMemcpyHostToDevice));
    h->per_scene = modes != nullptr;
    h->any_cell = cell;
    h->any_particle = particle;
    drop_graph(h);
    return WCSPH_OK;
}

int wcsph_get_scene_neighbor_modes(const wcsph_t* h, uint32_t* modes) {
    if (!h || !modes) return WCSPH_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < h->S; ++k) modes[k] = h->per_scene ? h->scenes[k].dev.form : h->var.neighbor_mode;
    return WCSPH_OK;
}

}  // extern "C"

namespace {

// Enqueues n_steps steps (graph replays or eager launches) on h->stream.
int enqueue_steps(wcsph_t* h, float dt_hint, uint32_t n_steps) {
    if (n_steps == 0) return WCSPH_OK;
    if (h->var.use_graph && !h->prof) {
        // A graph of a list form holds no counting (its force pass counts
        // for the next step): the positions as they are now are counted
        // first if no force pass counted them.  A graph of the cell-loop or
        // brick-list form counts at its start: counts a list form's force
        // pass left are cleared first.
        const bool fused = force_counts(h, false);
        if (h->n && fused && !h->counts_valid) {
            k_cell_count<<<grid_blocks(h->n, h->var.block_size), h->var.block_size, 0, h->stream>>>(
                h->pos_a, h->n, h->d_scenes, h->pstart, h->S, cell_count_args(h), h->status);
            h->counts_valid = true;
        } else if (h->n && !fused && h->counts_valid) {
            k_fill_u32<<<grid_blocks(h->ncell, 256), 256, 0, h->stream>>>(h->cnt, 0u, h->ncell);
            h->counts_valid = false;
        }
        if (!h->graph_exec || h->graph_hint != dt_hint) {
            drop_graph(h);
            HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(h, dt_hint);
            HIP_TRY(h, hipStreamEndCapture(h->stream, &h->graph));
            HIP_TRY(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
            h->graph_hint = dt_hint;
        }
        for (uint32_t k = 0; k < n_steps; ++k) HIP_TRY(h, hipGraphLaunch(h->graph_exec, h->stream));
    } else {
        for (uint32_t k = 0; k < n_steps; ++k) enqueue_step(h, dt_hint);
    }
    HIP_TRY(h, hipGetLastError());
    return WCSPH_OK;
}

// Recovery from a neighbour-list overflow (the stream is idle).  The step
// that overflowed stored nothing and every later step was skipped whole, so
// the fluid state is the one that step started from; its time state is in
// the saved_* fields.  Grows the list capacities past the reported maximum,
// restores the time state and enqueues the lost steps again, with the dt
// hints they were enqueued with.
constexpr uint32_t kMaxCapacity = 4096;

int recover_overflow(wcsph_t* h, const DevStatus& st) {
    auto grow = [](uint32_t cap, uint32_t need) {
        const uint32_t c = (need + need / 8u + 4u + 3u) & ~3u;  // 1/8 headroom, a multiple of 4
        return std::max(cap, c);
    };
    std::vector<SceneState> ss(h->S);
    HIP_TRY(h, hipMemcpy(ss.data(), h->d_state, sizeof(SceneState) * h->S, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (const auto& run : h->pending) total += run.second;
    const uint64_t done = ss[0].saved_steps - h->synced_steps;  // steps completed before the lost one
    if (ss[0].saved_steps < h->synced_steps || done >= total)
        return fail(h, WCSPH_ERR_NEIGHBOR_OVERFLOW, "neighbour list overflow: the lost steps cannot be identified");
    // back to the time state the lost step started from, which the fluid
    // state already is (also when the lists cannot grow below)
    for (SceneState& x : ss) {
        x.vmax2_bits = x.saved_vmax2_bits;
        x.sim_time = x.saved_sim_time;
        x.last_dt = x.saved_last_dt;
        x.steps = x.saved_steps;
    }
    HIP_TRY(h, hipMemcpy(h->d_state, ss.data(), sizeof(SceneState) * h->S, hipMemcpyHostToDevice));
    h->synced_steps = ss[0].saved_steps;
    uint32_t capf = h->var.fluid_capacity, capb = h->var.boundary_capacity;
    if (st.flags & WCSPH_FLAG_FLUID_OVERFLOW) capf = grow(capf, st.max_fluid);
    if (st.flags & WCSPH_FLAG_BOUNDARY_OVERFLOW) capb = grow(capb, st.max_boundary);
    if (capf > kMaxCapacity || capb > kMaxCapacity)
        return fail(h, WCSPH_ERR_NEIGHBOR_OVERFLOW,
                    "neighbour list overflow beyond the largest capacity (4096 entries per particle)");
    int rc = alloc_lists(h, capf, capb);
    if (rc) return rc;
    drop_graph(h);
    k_clear_flags<<<1, 64, 0, h->stream>>>(h->status);
    std::vector<std::pair<float, uint32_t>> rest;
    uint64_t skip = done;
    for (const auto& run : h->pending) {
        const uint64_t k = std::min<uint64_t>(skip, run.second);
        skip -= k;
        if (run.second > k) rest.emplace_back(run.first, (uint32_t)(run.second - k));
    }
    h->pending = rest;
    for (const auto& run : rest)
        if ((rc = enqueue_steps(h, run.first, run.second))) return rc;
    return WCSPH_OK;
}

// Waits for the stream and reports the flags and maxima of every step since
// the previous sync.  With `recover`, steps lost to a neighbour-list
// overflow are run again with grown lists (a few attempts at most).
int sync_impl(wcsph_t* h, wcsph_stats* stats, bool recover) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    fold_events(h);
    DevStatus st{};
    HIP_TRY(h, hipMemcpy(&st, h->status, sizeof(st), hipMemcpyDeviceToHost));
    uint32_t extra = 0, maxf = st.max_fluid, maxb = st.max_boundary;
    for (int attempt = 0; recover && (st.flags & kOverflowFlags) && !h->pending.empty() && attempt < 8; ++attempt) {
        // flags other than the overflow stay reported; the overflow is replaced
        // by the grown-lists note
        extra |= (st.flags & ~kOverflowFlags) | WCSPH_FLAG_LISTS_GROWN;
        int rc = recover_overflow(h, st);
        if (rc) {
            std::string why = h->err;
            k_clear_flags<<<1, 64, 0, h->stream>>>(h->status);
            (void)hipStreamSynchronize(h->stream);
            h->pending.clear();
            if (stats) {
                SceneState s0{};
                if (hipMemcpy(&s0, h->d_state, sizeof(s0), hipMemcpyDeviceToHost) == hipSuccess) {
                    stats->sim_time = s0.sim_time;
                    stats->last_dt = s0.last_dt;
                    stats->steps = s0.steps;
                }
                stats->flags = st.flags | extra;
                stats->max_fluid_neighbors = maxf;
                stats->max_boundary_neighbors = maxb;
                stats->reserved = 0;
            }
            h->err = why;
            return rc;
        }
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        fold_events(h);
        HIP_TRY(h, hipMemcpy(&st, h->status, sizeof(st), hipMemcpyDeviceToHost));
        maxf = std::max(maxf, st.max_fluid);
        maxb = std::max(maxb, st.max_boundary);
    }
    SceneState s0{};
    HIP_TRY(h, hipMemcpy(&s0, h->d_state, sizeof(s0), hipMemcpyDeviceToHost));
    // flags and maxima cover the steps since the previous sync
    k_clear_flags<<<1, 64, 0, h->stream>>>(h->status);
    h->pending.clear();
    h->synced_steps = s0.steps;
    const uint32_t flags = st.flags | extra;
    if (stats) {
        stats->sim_time = s0.sim_time;
        stats->last_dt = s0.last_dt;
        stats->steps = s0.steps;
        stats->flags = flags;
        stats->max_fluid_neighbors = maxf;
        stats->max_boundary_neighbors = maxb;
        stats->reserved = 0;
    }
    return status_to_code(h, st.flags);
}

}  // namespace

extern "C" {

int wcsph_step_async(wcsph_t* h, float dt_hint, uint32_t n_steps) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    if (!(dt_hint > 0.f) || !std::isfinite(dt_hint)) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "dt_hint must be > 0");
    if (n_steps == 0) return WCSPH_OK;
    int rc = enqueue_steps(h, dt_hint, n_steps);
    if (rc) return rc;
    if (!h->pending.empty() && h->pending.back().first == dt_hint &&
        h->pending.back().second <= 0xffffffffu - n_steps)
        h->pending.back().second += n_steps;
    else
        h->pending.emplace_back(dt_hint, n_steps);
    return WCSPH_OK;
}

int wcsph_sync(wcsph_t* h, wcsph_stats* stats) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    return sync_impl(h, stats, true);
}

uint32_t wcsph_scene_count(const wcsph_t* h) { return h ? h->S : 0; }

int wcsph_scene_range(const wcsph_t* h, uint32_t scene, uint32_t* first, uint32_t* count, uint32_t* bfirst,
                      uint32_t* bcount) {
    if (!h || scene >= h->S) return WCSPH_ERR_INVALID_ARGUMENT;
    const SceneHost& sc = h->scenes[scene];
    if (first) *first = sc.p0;
    if (count) *count = sc.n;
    if (bfirst) *bfirst = sc.b0;
    if (bcount) *bcount = sc.nb;
    return WCSPH_OK;
}

int wcsph_scene_stats(wcsph_t* h, uint32_t scene, wcsph_stats* stats) {
    if (!h || !stats || scene >= h->S) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    SceneState ss{};
    HIP_TRY(h, hipMemcpy(&ss, h->d_state + scene, sizeof(ss), hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof(*stats));
    stats->sim_time = ss.sim_time;
    stats->last_dt = ss.last_dt;
    stats->steps = ss.steps;
    return WCSPH_OK;
}

int wcsph_step(wcsph_t* h, float dt_hint, uint32_t n_steps, wcsph_stats* stats) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    int rc = wcsph_step_async(h, dt_hint, n_steps);
    if (rc) return rc;
    return wcsph_sync(h, stats);
}

uint32_t wcsph_fluid_count(const wcsph_t* h) { return h ? h->n : 0; }
uint32_t wcsph_boundary_count(const wcsph_t* h) { return h ? h->nb : 0; }

int wcsph_get_state(wcsph_t* h, float* pos, float* vel, float* rho, float* p, uint32_t* ids) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const size_t n = h->n;
    if (!n) return WCSPH_OK;
    if (pos) HIP_TRY(h, hipMemcpy(pos, h->pos_a, 16 * n, hipMemcpyDeviceToHost));
    if (vel) HIP_TRY(h, hipMemcpy(vel, h->vel_a, 16 * n, hipMemcpyDeviceToHost));
    if (rho) HIP_TRY(h, hipMemcpy(rho, h->rho, 4 * n, hipMemcpyDeviceToHost));
    if (p) HIP_TRY(h, hipMemcpy(p, h->prs, 4 * n, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(h, hipMemcpy(ids, h->id_a, 4 * n, hipMemcpyDeviceToHost));
    return WCSPH_OK;
}

int wcsph_set_box(wcsph_t* h, uint32_t scene, const wcsph_box* box) {
    if (!h || !box || scene >= h->S) return WCSPH_ERR_INVALID_ARGUMENT;
    for (int a = 0; a < 3; ++a)
        if (!(box->max[a] > box->min[a]) || !std::isfinite(box->min[a]) || !std::isfinite(box->max[a]))
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "wcsph_set_box: need finite min < max on every axis");
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(box->model[i]) || !std::isfinite(box->model_inv[i]))
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "wcsph_set_box: non-finite model matrix");
    SceneHost& sc = h->scenes[scene];
    // the new box's world AABB must stay inside the grid's interior cells
    wcsph_scene probe{};
    probe.box = *box;
    double lo[3], hi[3];
    std::string why;
    if (scene_extent(probe, lo, hi, true, why)) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, why);
    for (const Grid* gp : {&sc.dev.g, &sc.dev.bg}) {
        const Grid& g = *gp;
        for (int a = 0; a < 3; ++a) {
            const int c0 = (&g.c0x)[a], nd = (&g.nx)[a];
            if (host_cell((float)lo[a], g.cs) < c0 + g.R || host_cell((float)hi[a], g.cs) > c0 + nd - 1 - g.R)
                return fail(h, WCSPH_ERR_DOMAIN,
                            "wcsph_set_box: the box leaves the scene's cell grid (give wcsph_scene a domain at "
                            "creation)");
        }
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    SceneDev d = sc.dev;
    box_to_dev(*box, d);
    HIP_TRY(h, hipMemcpy(h->d_scenes + scene, &d, sizeof(SceneDev), hipMemcpyHostToDevice));
    sc.dev = d;
    sc.box = *box;
    return WCSPH_OK;
}

int wcsph_get_box(const wcsph_t* h, uint32_t scene, wcsph_box* box) {
    if (!h || !box || scene >= h->S) return WCSPH_ERR_INVALID_ARGUMENT;
    *box = h->scenes[scene].box;
    return WCSPH_OK;
}

int wcsph_set_state(wcsph_t* h, const float* pos, const float* vel) {
    if (!h || (!pos && h->n)) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return upload_fluid(h, pos, vel);
}

int wcsph_get_permutation(wcsph_t* h, uint32_t* perm) {
    if (!h || !perm) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->n) HIP_TRY(h, hipMemcpy(perm, h->perm, 4 * (size_t)h->n, hipMemcpyDeviceToHost));
    return WCSPH_OK;
}

int wcsph_get_sorted_positions(wcsph_t* h, float* pos) {
    if (!h || !pos) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->n) HIP_TRY(h, hipMemcpy(pos, h->pos_b, 16 * (size_t)h->n, hipMemcpyDeviceToHost));
    return WCSPH_OK;
}

int wcsph_get_neighbors(wcsph_t* h, int which, uint32_t* counts, uint32_t* list, uint64_t list_len,
                        uint64_t* total) {
    if (!h || !counts || (which != 0 && which != 1)) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const size_t n = h->n;
    if (!h->lists_valid && n) {
        // the last step walked the cells without lists: build them now from
        // the positions that step searched (same builder as the list form)
        const uint32_t capf = h->var.fluid_capacity, capb = h->var.boundary_capacity;
        if (h->scenes[0].dev.g.R == 1)
            k_neighbors<false, 1><<<grid_blocks(n, kNeighborBlock), kNeighborBlock, 0, h->stream>>>(
                h->pos_b, (uint32_t)n, h->d_scenes, h->pstart, h->S, h->crange, h->bpos, h->brange, h->bnear,
                h->phys.h2, capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->hook_status, h->phys, nullptr,
                nullptr, nullptr, nullptr, h->fxyz, h->bxyz, kAllForms);
        else
            k_neighbors<false, 2><<<grid_blocks(n, kNeighborBlock), kNeighborBlock, 0, h->stream>>>(
                h->pos_b, (uint32_t)n, h->d_scenes, h->pstart, h->S, h->crange, h->bpos, h->brange, h->bnear,
                h->phys.h2, capf, capb, h->nlist, h->ncount, h->blist, h->bcount, h->hook_status, h->phys, nullptr,
                nullptr, nullptr, nullptr, h->fxyz, h->bxyz, kAllForms);
        HIP_TRY(h, hipGetLastError());
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        h->lists_valid = true;
    }
    const uint32_t cap = which ? h->var.boundary_capacity : h->var.fluid_capacity;
    if (n) HIP_TRY(h, hipMemcpy(counts, which ? h->bcount : h->ncount, 4 * n, hipMemcpyDeviceToHost));
    uint64_t sum = 0;
    uint32_t mx = 0;
    for (size_t i = 0; i < n; ++i) {
        sum += counts[i];
        mx = std::max(mx, counts[i]);
    }
    if (total) *total = sum;
    if (!list) return WCSPH_OK;
    if (mx > cap) return fail(h, WCSPH_ERR_NEIGHBOR_OVERFLOW, "neighbour list overflowed in the last step");
    if (list_len < sum) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "list buffer too small");
    // the first ceil(mx / 4) slot quads of every particle (list_at layout)
    const size_t quads = ((size_t)mx + 3) / 4;
    std::vector<uint32_t> rows(quads * 4 * n);
    if (mx) HIP_TRY(h, hipMemcpy(rows.data(), which ? h->blist : h->nlist, 4 * rows.size(), hipMemcpyDeviceToHost));
    uint64_t w = 0;
    for (size_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < counts[i]; ++k) list[w++] = rows[((k >> 2) * n + i) * 4 + (k & 3)];
    return WCSPH_OK;
}

int wcsph_get_boundary(wcsph_t* h, float* pos_psi) {
    if (!h || !pos_psi) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (h->nb) HIP_TRY(h, hipMemcpy(pos_psi, h->bpos, 16 * (size_t)h->nb, hipMemcpyDeviceToHost));
    return WCSPH_OK;
}

int wcsph_state_metrics(wcsph_t* h, double* mean_speed, double* kinetic_energy) {
    return wcsph_scene_metrics(h, 0, mean_speed, kinetic_energy);
}

int wcsph_scene_metrics(wcsph_t* h, uint32_t scene, double* mean_speed, double* kinetic_energy) {
    if (!h || scene >= h->S) return WCSPH_ERR_INVALID_ARGUMENT;
    const SceneHost& sc = h->scenes[scene];
    const uint32_t nblk = sc.n ? grid_blocks(sc.n, 256) : 0u;
    std::vector<double> part(2 * (size_t)nblk);
    double out[2] = {0.0, 0.0};
    if (nblk) {
        k_metrics<<<nblk, 256, 0, h->stream>>>(h->vel_a + sc.p0, sc.n, h->metrics);
        HIP_TRY(h, hipGetLastError());
        HIP_TRY(h, hipMemcpyAsync(part.data(), h->metrics, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                  h->stream));
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (uint32_t b = 0; b < nblk; ++b) {  // workgroup order: deterministic
        out[0] += part[2 * (size_t)b];
        out[1] += part[2 * (size_t)b + 1];
    }
    if (mean_speed) *mean_speed = sc.n ? out[0] / (double)sc.n : 0.0;
    if (kinetic_energy) *kinetic_energy = 0.5 * (double)h->phys.mass * out[1];
    return WCSPH_OK;
}

int wcsph_set_profiling(wcsph_t* h, int enable) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    fold_events(h);
    h->prof = enable != 0;
    return WCSPH_OK;
}

int wcsph_kernel_times(wcsph_t* h, double* ms_total, uint64_t* launches, int reset) {
    if (!h) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    fold_events(h);
    for (int k = 0; k < WCSPH_NUM_KERNELS; ++k) {
        if (ms_total) ms_total[k] = h->kms[k];
        if (launches) launches[k] = h->klaunch[k];
        if (reset) {
            h->kms[k] = 0.0;
            h->klaunch[k] = 0;
        }
    }
    return WCSPH_OK;
}

const char* wcsph_kernel_name(int id) {
    static const char* names[WCSPH_NUM_KERNELS] = {"time_step", "cell_count", "scan", "scatter", "permute",
                                                   "neighbors", "density", "normals", "force_integrate"};
    return (id >= 0 && id < WCSPH_NUM_KERNELS) ? names[id] : "";
}

int wcsph_set_viscosity(wcsph_t* h, float alpha) {
    if (!h || !(alpha >= 0.f) || !std::isfinite(alpha)) return WCSPH_ERR_INVALID_ARGUMENT;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->prm.viscosity = alpha;
    build_phys(h->prm, h->phys);
    drop_graph(h);
    return WCSPH_OK;
}

}  // extern "C"

namespace {

int grad_alloc(wcsph_t* h, uint32_t steps, uint32_t every) {
    auto& G = h->grad;
    const size_t n = h->n;
    const uint32_t ncp = (steps + every - 1) / every;
    const uint32_t len = std::min(every, steps);
    const bool same = G.ready && G.ncp >= ncp && G.tape_len >= len && G.tape_capf == h->var.fluid_capacity &&
                      G.tape_capb == h->var.boundary_capacity;
    if (same) return WCSPH_OK;
    void* olds[] = {G.ck_pos, G.ck_vel, G.ck_id, G.t_pos, G.t_vel, G.t_xp, G.t_p, G.t_perm, G.t_nc, G.t_bc,
                    G.t_coll, G.t_nl, G.t_bl, G.lx, G.lv, G.lx2, G.lv2, G.ly, G.lvn, G.la, G.gx, G.gv, G.lrho,
                    G.galpha, G.final_id};
    for (void* q : olds) dfree(h, q);
    G = decltype(h->grad){};
    int rc;
    const size_t capf = h->var.fluid_capacity, capb = h->var.boundary_capacity;
    if ((rc = dalloc(h, &G.ck_pos, ncp * n)) || (rc = dalloc(h, &G.ck_vel, ncp * n)) ||
        (rc = dalloc(h, &G.ck_id, ncp * n)) || (rc = dalloc(h, &G.t_pos, len * n)) ||
        (rc = dalloc(h, &G.t_vel, len * n)) || (rc = dalloc(h, &G.t_xp, len * n)) || (rc = dalloc(h, &G.t_p, len * n)) ||
        (rc = dalloc(h, &G.t_perm, len * n)) || (rc = dalloc(h, &G.t_nc, len * n)) ||
        (rc = dalloc(h, &G.t_bc, len * n)) || (rc = dalloc(h, &G.t_coll, len * n)) ||
        (rc = dalloc(h, &G.t_nl, len * capf * n)) || (rc = dalloc(h, &G.t_bl, len * capb * n)) ||
        (rc = dalloc(h, &G.lx, n)) || (rc = dalloc(h, &G.lv, n)) || (rc = dalloc(h, &G.lx2, n)) ||
        (rc = dalloc(h, &G.lv2, n)) || (rc = dalloc(h, &G.ly, n)) || (rc = dalloc(h, &G.lvn, n)) ||
        (rc = dalloc(h, &G.la, n)) || (rc = dalloc(h, &G.gx, n)) || (rc = dalloc(h, &G.gv, n)) ||
        (rc = dalloc(h, &G.lrho, n)) || (rc = dalloc(h, &G.galpha, grid_blocks(h->n, 64))) ||
        (rc = dalloc(h, &G.final_id, n)))
        return rc;
    G.ncp = ncp;
    G.tape_len = len;
    G.tape_capf = (uint32_t)capf;
    G.tape_capb = (uint32_t)capb;
    return WCSPH_OK;
}

// One taped step: the step itself, then its sorted state, lists and
// permutation copied into tape slot k (the collision flags are written by
// the force kernel straight into the slot).
int taped_step(wcsph_t* h, float dt, uint32_t k) {
    auto& G = h->grad;
    const size_t n = h->n;
    const size_t capf = G.tape_capf, capb = G.tape_capb;
    // the step builds its neighbour lists straight into tape slot k
    uint32_t* const nl = h->nlist;
    uint32_t* const bl = h->blist;
    h->nlist = G.t_nl + k * capf * n;
    h->blist = G.t_bl + k * capb * n;
    h->coll_out = G.t_coll + k * n;
    enqueue_step(h, dt, true);  // the adjoint replays the step's lists
    h->coll_out = nullptr;
    h->nlist = nl;
    h->blist = bl;
    h->lists_valid = false;  // the handle's own list buffers hold an older step
    if (n)
        k_tape_copy<<<grid_blocks(n, 256), 256, 0, h->stream>>>(
            h->pos_b, h->vel_b, h->xp, h->prs, h->perm, h->ncount, h->bcount, (uint32_t)n, G.t_pos + k * n,
            G.t_vel + k * n, G.t_xp + k * n, G.t_p + k * n, G.t_perm + k * n, G.t_nc + k * n, G.t_bc + k * n);
    HIP_TRY(h, hipGetLastError());
    return WCSPH_OK;
}

// Adjoint of tape slot k: (lx, lv) after the step -> (lx2, lv2) before it.
void adjoint_step(wcsph_t* h, uint32_t k) {
    auto& G = h->grad;
    const uint32_t n = h->n;
    const size_t N = n;
    const uint32_t bs = h->var.block_size, gb = grid_blocks(n, bs);
    const uint32_t capf = G.tape_capf, capb = G.tape_capb;
    hipStream_t s = h->stream;
    const double csound = std::sqrt((double)h->prm.stiffness * (double)h->prm.exponent / (double)h->prm.rest_density);
    const float dcoef = (float)(2.0 * (double)h->phys.h * csound);
    k_adj_integrate<<<gb, bs, 0, s>>>(G.lx, G.lv, G.t_coll + k * N, n, G.dt, h->phys.elasticity, G.ly, G.lvn, G.la);
    k_adj_pairs<<<gb, bs, 0, s>>>(G.t_pos + k * N, G.t_vel + k * N, G.t_xp + k * N, G.t_p + k * N, n,
                                   G.t_nl + k * capf * N, G.t_nc + k * N, h->bpos, G.t_bl + k * capb * N,
                                   G.t_bc + k * N, capf, capb, h->phys, dcoef, G.la, G.lrho, G.gx, G.gv, G.galpha);
    k_adj_density<<<gb, bs, 0, s>>>(G.t_pos + k * N, n, G.t_nl + k * capf * N, G.t_nc + k * N, h->bpos,
                                     G.t_bl + k * capb * N, G.t_bc + k * N, capf, capb, h->phys, G.lrho, G.gx,
                                     G.gv, G.ly, G.lvn, G.t_perm + k * N, G.lx2, G.lv2);
    std::swap(G.lx, G.lx2);
    std::swap(G.lv, G.lv2);
}

int grad_supported(wcsph_t* h) {
    if (h->S != 1) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "differentiable runs take a single scene");
    if (h->n == 0) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "differentiable runs need fluid particles");
    if (h->prm.adaptive_dt || h->prm.split_np_dt || h->prm.surface_tension != 0.f || h->prm.n_force_rules != 0 ||
        h->prm.pressure_gradient != WCSPH_GRAD_CUBIC || h->prm.viscosity_model != WCSPH_VISC_ARTIFICIAL)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                    "the adjoint covers the baseline physics with a fixed dt (adaptive_dt = 0) and no external force rules");
    if (!h->scenes[0].dev.box_identity)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "the adjoint needs an axis-aligned (identity) box");
    return WCSPH_OK;
}

}  // namespace

extern "C" {

int wcsph_grad_forward(wcsph_t* h, const float* vel0, float dt, uint32_t n_steps, uint32_t checkpoint_every,
                       float* pos_out) {
    if (!h || !pos_out || n_steps == 0 || checkpoint_every == 0 || !(dt > 0.f)) return WCSPH_ERR_INVALID_ARGUMENT;
    int rc;
    if ((rc = grad_supported(h))) return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if ((rc = grad_alloc(h, n_steps, checkpoint_every))) return rc;
    auto& G = h->grad;
    G.steps = n_steps;
    G.every = checkpoint_every;
    G.dt = dt;
    G.ready = false;
    G.phys = h->phys;
    const size_t n = h->n;
    hipStream_t s = h->stream;
    // start: creation positions, the given velocities, ids 0..n-1
    HIP_TRY(h, hipMemcpyAsync(h->pos_a, h->pos0, 16 * n, hipMemcpyDeviceToDevice, s));
    uncount(h);
    if (vel0) HIP_TRY(h, hipMemcpyAsync(h->vel_a, vel0, 16 * n, hipMemcpyDefault, s));
    else HIP_TRY(h, hipMemsetAsync(h->vel_a, 0, 16 * n, s));
    if (n) k_iota<<<grid_blocks(n, 256), 256, 0, s>>>(h->id_a, h->n);
    for (uint32_t k = 0; k < n_steps; ++k) {
        if (k % checkpoint_every == 0) {
            const size_t c = k / checkpoint_every;
            HIP_TRY(h, hipMemcpyAsync(G.ck_pos + c * n, h->pos_a, 16 * n, hipMemcpyDeviceToDevice, s));
            HIP_TRY(h, hipMemcpyAsync(G.ck_vel + c * n, h->vel_a, 16 * n, hipMemcpyDeviceToDevice, s));
            HIP_TRY(h, hipMemcpyAsync(G.ck_id + c * n, h->id_a, 4 * n, hipMemcpyDeviceToDevice, s));
        }
        enqueue_step(h, dt);
    }
    HIP_TRY(h, hipGetLastError());
    // final positions in creation order; the final order is kept for the
    // backward pass's seed (the handle's own order changes when it steps on)
    if (n) k_to_id_order<<<grid_blocks(n, 256), 256, 0, s>>>(h->pos_a, h->id_a, h->n, G.lx2);
    HIP_TRY(h, hipMemcpyAsync(pos_out, G.lx2, 16 * n, hipMemcpyDefault, s));
    HIP_TRY(h, hipMemcpyAsync(G.final_id, h->id_a, 4 * n, hipMemcpyDeviceToDevice, s));
    rc = sync_impl(h, nullptr, false);
    if (rc) return rc;
    G.ready = true;
    G.gen = ++h->grad_runs;
    return WCSPH_OK;
}

uint64_t wcsph_grad_generation(const wcsph_t* h) { return (h && h->grad.ready) ? h->grad.gen : 0; }

int wcsph_grad_backward(wcsph_t* h, const float* grad_pos, float* grad_vel0, double* grad_alpha) {
    if (!h || !grad_pos) return WCSPH_ERR_INVALID_ARGUMENT;
    auto& G = h->grad;
    if (!G.ready) return fail(h, WCSPH_ERR_INVALID_ARGUMENT, "wcsph_grad_backward needs a wcsph_grad_forward first");
    const size_t n = h->n;
    hipStream_t s = h->stream;
    // the recomputation and the adjoint use the forward run's physics (alpha
    // may have been changed since), restored on every return
    struct PhysSwap {
        wcsph_t* h;
        Phys keep;
        ~PhysSwap() { h->phys = keep; }
    } swap{h, h->phys};
    h->phys = G.phys;
    // seed: dL/dx(T) from creation order to the forward run's final order
    HIP_TRY(h, hipMemcpyAsync(G.lx2, grad_pos, 16 * n, hipMemcpyDefault, s));
    if (n) k_from_id_order<<<grid_blocks(n, 256), 256, 0, s>>>(G.lx2, G.final_id, h->n, G.lx);
    HIP_TRY(h, hipMemsetAsync(G.lv, 0, 16 * n, s));
    HIP_TRY(h, hipMemsetAsync(G.galpha, 0, sizeof(double) * grid_blocks(h->n, 64), s));
    for (uint32_t c = G.ncp; c-- > 0;) {
        const uint32_t k0 = c * G.every, k1 = std::min(k0 + G.every, G.steps);
        HIP_TRY(h, hipMemcpyAsync(h->pos_a, G.ck_pos + c * n, 16 * n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(h, hipMemcpyAsync(h->vel_a, G.ck_vel + c * n, 16 * n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(h, hipMemcpyAsync(h->id_a, G.ck_id + c * n, 4 * n, hipMemcpyDeviceToDevice, s));
        uncount(h);
        for (uint32_t k = k0; k < k1; ++k) {
            int rc = taped_step(h, G.dt, k - k0);
            if (rc) return rc;
        }
        for (uint32_t k = k1; k-- > k0;) adjoint_step(h, k - k0);
    }
    HIP_TRY(h, hipGetLastError());
    // the first step's order is the creation order (ids 0..n-1)
    if (grad_vel0) HIP_TRY(h, hipMemcpyAsync(grad_vel0, G.lv, 16 * n, hipMemcpyDefault, s));
    // per-workgroup partial sums of dL/dalpha, summed in workgroup order
    std::vector<double> part(grid_blocks(h->n, h->var.block_size));
    HIP_TRY(h, hipMemcpyAsync(part.data(), G.galpha, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    int rc = sync_impl(h, nullptr, false);
    double ga = 0.0;
    for (double v : part) ga += v;
    if (grad_alpha) *grad_alpha = ga;
    return rc;
}

}  // extern "C"

namespace {

// Where the autotuner keeps the state it restores after every trial.
struct TuneBackup {
    float4* pos;
    float4* vel;
    uint32_t* id;
    SceneState* state;
};

constexpr uint32_t kMaxDefaultCandidates = 9;

// The default space, all with the handle's other settings: block size
// {128, 256, 512} with the per-particle builder, the cell builder with
// {4096, 8192} workgroups, the cell-loop form and the brick-list form at the
// handle's cell size (the last four only at cell_size_mult 1, the one value
// they take); with `other_cell_sizes` (a grid rebuild, so not without
// allocation) also the per-particle builder at block size 256 over the other
// cell sizes of {1, 2/3, 1/2}.  Returns the number of candidates.
uint32_t default_candidates(const wcsph_variant& base, wcsph_variant* out, bool other_cell_sizes) {
    const float m = base.cell_size_mult == 0.f ? 1.f : base.cell_size_mult;
    uint32_t k = 0;
    for (uint32_t bs : {128u, 256u, 512u}) {
        out[k] = base;
        out[k].cell_size_mult = m;
        out[k].block_size = bs;
        out[k++].neighbor_mode = WCSPH_NEIGHBORS_PARTICLE;
    }
    if (m == 1.f) {
        for (uint32_t groups : {4096u, 8192u}) {
            out[k] = base;
            out[k].cell_size_mult = 1.f;
            out[k].neighbor_mode = WCSPH_NEIGHBORS_CELL;
            out[k++].neighbor_groups = groups;
        }
        for (uint32_t mode : {(uint32_t)WCSPH_NEIGHBORS_CELL_LOOP, (uint32_t)WCSPH_NEIGHBORS_BRICK}) {
            out[k] = base;
            out[k].cell_size_mult = 1.f;
            out[k++].neighbor_mode = mode;
        }
    }
    if (other_cell_sizes) {
        for (float cm : {1.f, 2.f / 3.f, 0.5f}) {
            if (cm == m) continue;
            out[k] = base;
            out[k].cell_size_mult = cm;
            out[k].block_size = 256;
            out[k++].neighbor_mode = WCSPH_NEIGHBORS_PARTICLE;
        }
    }
    return k;
}

// The trials of wcsph_autotune / wcsph_autotune_noalloc.  `eager` runs every
// trial as plain launches (no graph is captured or instantiated).  The
// stream is idle and no step is pending on entry.
int autotune_core(wcsph_t* h, float dt_hint, uint32_t steps_per_trial, const wcsph_variant* cand, uint32_t ncand,
                  double* ms_per_step, wcsph_variant* best, const TuneBackup& bk, bool eager) {
    const wcsph_variant orig = h->var;
    const size_t n = h->n;
    hipStream_t s = h->stream;
    auto copy_state = [&](bool to_backup) -> int {
        if (n) {
            HIP_TRY(h, hipMemcpyAsync(to_backup ? bk.pos : h->pos_a, to_backup ? h->pos_a : bk.pos, 16 * n,
                                      hipMemcpyDeviceToDevice, s));
            HIP_TRY(h, hipMemcpyAsync(to_backup ? bk.vel : h->vel_a, to_backup ? h->vel_a : bk.vel, 16 * n,
                                      hipMemcpyDeviceToDevice, s));
            HIP_TRY(h, hipMemcpyAsync(to_backup ? bk.id : h->id_a, to_backup ? h->id_a : bk.id, 4 * n,
                                      hipMemcpyDeviceToDevice, s));
        }
        HIP_TRY(h, hipMemcpyAsync(to_backup ? bk.state : h->d_state, to_backup ? h->d_state : bk.state,
                                  sizeof(SceneState) * h->S, hipMemcpyDeviceToDevice, s));
        HIP_TRY(h, hipStreamSynchronize(s));
        if (!to_backup) {  // the step counter is restored with the rest
            SceneState s0{};
            HIP_TRY(h, hipMemcpy(&s0, h->d_state, sizeof(s0), hipMemcpyDeviceToHost));
            h->synced_steps = s0.steps;
            uncount(h);
        }
        return WCSPH_OK;
    };
    int rc = copy_state(true);
    if (rc) return rc;
    const bool prof = h->prof;
    if (eager) h->prof = false;  // profiling marks would create events
    double best_ms = INFINITY;
    uint32_t best_k = 0;
    for (uint32_t k = 0; k < ncand; ++k) {
        double ms = INFINITY;
        if (wcsph_set_variant(h, &cand[k]) == WCSPH_OK) {
            if (eager) h->var.use_graph = 0;
            // one untimed step (graph capture, caches), then the timed steps;
            // a candidate whose lists overflow fails (no recovery in a trial)
            bool ok = copy_state(false) == WCSPH_OK && enqueue_steps(h, dt_hint, 1) == WCSPH_OK &&
                      sync_impl(h, nullptr, false) == WCSPH_OK && copy_state(false) == WCSPH_OK;
            if (ok) {
                (void)hipEventRecord(h->tune_ev[0], s);
                ok = enqueue_steps(h, dt_hint, steps_per_trial) == WCSPH_OK;
                (void)hipEventRecord(h->tune_ev[1], s);
                ok = sync_impl(h, nullptr, false) == WCSPH_OK && ok;
                float t = 0.f;
                if (ok && hipEventElapsedTime(&t, h->tune_ev[0], h->tune_ev[1]) == hipSuccess)
                    ms = (double)t / steps_per_trial;
            }
        }
        if (ms_per_step) ms_per_step[k] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best_k = k;
        }
    }
    h->prof = prof;
    const wcsph_variant chosen = std::isfinite(best_ms) ? cand[best_k] : orig;
    rc = wcsph_set_variant(h, &chosen);
    if (rc == WCSPH_OK) rc = copy_state(false);
    if (best) *best = chosen;
    if (rc == WCSPH_OK) h->err.clear();
    return rc;
}

}  // namespace

extern "C" {

int wcsph_autotune(wcsph_t* h, float dt_hint, uint32_t steps_per_trial, const wcsph_variant* candidates,
                   uint32_t n_candidates, double* ms_per_step, wcsph_variant* best) {
    if (!h || steps_per_trial == 0 || !(dt_hint > 0.f)) return WCSPH_ERR_INVALID_ARGUMENT;
    wcsph_variant defaults[kMaxDefaultCandidates];
    const wcsph_variant* cand = candidates;
    uint32_t ncand = n_candidates;
    if (!candidates || !n_candidates) {
        ncand = default_candidates(h->var, defaults, true);
        cand = defaults;
    }
    // steps still pending complete first, as in wcsph_sync (their error, if
    // any, is returned before any trial runs)
    int rc = sync_impl(h, nullptr, true);
    if (rc) return rc;
    const size_t n = h->n;
    TuneBackup bk{};
    if ((rc = dalloc(h, &bk.pos, n)) || (rc = dalloc(h, &bk.vel, n)) || (rc = dalloc(h, &bk.id, n)) ||
        (rc = dalloc(h, &bk.state, h->S))) {
        const std::string why = h->err;
        dfree(h, bk.pos);
        dfree(h, bk.vel);
        dfree(h, bk.id);
        dfree(h, bk.state);
        h->err = why;
        return rc;
    }
    rc = autotune_core(h, dt_hint, steps_per_trial, cand, ncand, ms_per_step, best, bk, false);
    dfree(h, bk.pos);
    dfree(h, bk.vel);
    dfree(h, bk.id);
    dfree(h, bk.state);
    return rc;
}

int wcsph_autotune_noalloc(wcsph_t* h, float dt_hint, uint32_t steps_per_trial, const wcsph_variant* candidates,
                           uint32_t n_candidates, double* ms_per_step, wcsph_variant* best) {
    if (!h || steps_per_trial == 0 || !(dt_hint > 0.f)) return WCSPH_ERR_INVALID_ARGUMENT;
    if (!h->scratch || !h->tune_ev[0] || !h->tune_ev[1])
        return fail(h, WCSPH_ERR_HIP, "wcsph_autotune_noalloc: the handle has no scratch address or events");
    wcsph_variant defaults[kMaxDefaultCandidates];
    const wcsph_variant* cand = candidates;
    uint32_t ncand = n_candidates;
    if (!candidates || !n_candidates) {
        ncand = default_candidates(h->var, defaults, false);
        cand = defaults;
    }
    // a capacity change would reallocate the neighbour lists
    for (uint32_t k = 0; k < ncand; ++k) {
        const float m = cand[k].cell_size_mult == 0.f ? 1.f : cand[k].cell_size_mult;
        if (cand[k].fluid_capacity != h->var.fluid_capacity || cand[k].boundary_capacity != h->var.boundary_capacity ||
            m != h->var.cell_size_mult)
            return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                        "wcsph_autotune_noalloc: every candidate must keep the handle's list capacities and cell size");
    }
    // layout in the scratch array: positions, velocities, ids, scene states,
    // each at a 256-byte boundary
    auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t n = h->n;
    const size_t o_vel = up256(16 * n), o_id = o_vel + up256(16 * n), o_state = o_id + up256(4 * n);
    const size_t need = o_state + sizeof(SceneState) * h->S;
    if (need > kNoAllocScratchBytes)
        return fail(h, WCSPH_ERR_INVALID_ARGUMENT,
                    "wcsph_autotune_noalloc: the handle's state (" + std::to_string(need) +
                        " B) exceeds the 64 MiB scratch array (about 1.86M fluid particles)");
    // pending steps complete first; an overflow among them is reported, not
    // recovered (recovery grows the lists)
    int rc = sync_impl(h, nullptr, false);
    if (rc) return rc;
    TuneBackup bk{(float4*)h->scratch, (float4*)(h->scratch + o_vel), (uint32_t*)(h->scratch + o_id),
                  (SceneState*)(h->scratch + o_state)};
    return autotune_core(h, dt_hint, steps_per_trial, cand, ncand, ms_per_step, best, bk, true);
}

}  // extern "C"
