// Driver of the oracle's address/undefined-behaviour sanitizer build.
// TEST INFRASTRUCTURE ONLY (built and run by tests/test_sanitizers.py).
//
// Calls every extern "C" entry point of wcsph_oracle.cpp on small scenes of
// both presets and both precisions; a sanitizer report aborts the process
// (-fno-sanitize-recover=all), so exit status 0 means a clean run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../include/wcsph.h"

extern "C" {
uint64_t orc_morton_key(int cx, int cy, int cz);
int orc_cell_coord_f32(float p, float cs);
void orc_sort_order_f32(const float* pos4, uint32_t n, float cs, uint32_t* order);
void orc_morton_keys_f32(const float* pos4, uint32_t n, float cs, uint64_t* keys);
uint64_t orc_neighbors_f32(const float* q4, uint32_t nq, const float* src4, uint32_t ns, float h, float cs,
                           int exclude_self, uint32_t* counts, uint32_t* list, uint64_t list_len);
double orc_cubic_w(double q, double h);
double orc_cubic_grad(double q, double h);
double orc_spiky_ref_grad(double q, double h);
double orc_cohesion(double q, double h);
double orc_tait(double rho, double rho0, double k, double gamma, double max_ratio);
int orc_make_dam_break(float r, const uint32_t dims[3], const float lattice_min[3], float jitter, uint64_t seed,
                       const float box_min[3], const float box_max[3], uint32_t layers, float* fluid,
                       uint32_t* n_fluid, float* bnd, uint32_t* n_bnd);
void* orc_create(int precision, const wcsph_params* p, const wcsph_scene* sc, int parallel);
void orc_destroy(void* h);
void orc_set_box(void* h, const wcsph_box* box);
void orc_set_cell_size_mult(void* h, float mult);
void orc_step(void* h, double hint, uint32_t n_steps);
void orc_set_state64(void* h, const double* pos, const double* vel);
void orc_get_state(void* h, double* pos, double* vel, double* rho, double* p, uint32_t* ids);
void orc_get_boundary(void* h, double* pos_psi);
void orc_get_perm(void* h, uint32_t* perm);
uint64_t orc_get_neighbors(void* h, int which, uint32_t* counts, uint32_t* list, uint64_t list_len);
void orc_eval(void* h, const double* rho_in, const double* p_in, double* rho, double* p, double* ap, double* anp,
              double* nrm);
void orc_get_time(void* h, double* sim_time, double* last_dt);
double orc_grad_run(void* h, double hint, uint32_t n_steps, const double* targets);
void orc_grad_backward(void* h, const double* targets, double* gv0, double* galpha);
void orc_set_viscosity(void* h, double alpha);
}

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "san_check: %s\n", what);
        ++failures;
    }
}

// The two physics presets, written out (the product library is not linked).
wcsph_params preset(bool reference) {
    wcsph_params p{};
    p.rest_density = 1000.f;
    p.exponent = 7.f;
    p.max_density_ratio = 1.5f;
    p.gravity[1] = -9.81f;
    p.cfl_factor = 0.4f;
    p.dt_min = 1e-5f;
    p.dt_max = 5e-4f;
    p.dt_np_max = 0.015f;
    p.collision_buffer = 0.75f;
    p.elasticity = 0.5f;
    p.particle_radius = reference ? 0.02f : 0.01f;
    p.stiffness = reference ? 0.1f : 1000.f;
    p.viscosity = reference ? 0.001f : 0.01f;
    p.surface_tension = reference ? 0.13f : 0.f;
    p.pressure_gradient = reference ? WCSPH_GRAD_SPIKY_REF : WCSPH_GRAD_CUBIC;
    p.viscosity_model = reference ? WCSPH_VISC_LAPLACIAN : WCSPH_VISC_ARTIFICIAL;
    p.adaptive_dt = reference ? 1u : 0u;
    p.split_np_dt = reference ? 1u : 0u;
    if (reference) {
        p.n_force_rules = 2;
        p.force_rules[0].region = WCSPH_REGION_XZ_DISC_Y_BELOW;
        p.force_rules[0].radius2 = 0.2f;
        p.force_rules[0].value[1] = 25.f;
        p.force_rules[1].region = WCSPH_REGION_Y_BELOW;
        p.force_rules[1].y_below = 0.3f;
        p.force_rules[1].kind = WCSPH_FORCE_SWIRL_Y_AXIS;
        p.force_rules[1].magnitude = 2.f;
    }
    return p;
}

wcsph_box aabb(const float lo[3], const float hi[3], float angle) {
    wcsph_box b{};
    for (int a = 0; a < 3; ++a) {
        b.min[a] = lo[a];
        b.max[a] = hi[a];
    }
    const float c = std::cos(angle), s = std::sin(angle);
    // rotation about z (column-major), and its transpose as the inverse
    const float m[16] = {c, s, 0, 0, -s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (int i = 0; i < 16; ++i) b.model[i] = m[i];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) b.model_inv[4 * k + r] = m[4 * r + k];
    return b;
}

void run_scene(bool reference, int precision, int parallel) {
    const wcsph_params prm = preset(reference);
    const float r = prm.particle_radius, s = 2.f * r;
    const uint32_t dims[3] = {6, 5, 4};
    const float lmin[3] = {0.2f, 0.2f, 0.2f};
    const float bmin[3] = {0.f, 0.f, 0.f};
    const float bmax[3] = {s * 12.f, s * 10.f, s * 8.f};
    uint32_t nf = 0, nb = 0;
    orc_make_dam_break(r, dims, lmin, 0.1f * r, 7, bmin, bmax, 2, nullptr, &nf, nullptr, &nb);
    std::vector<float> fl(4 * (size_t)nf), bd(4 * (size_t)nb), vel(4 * (size_t)nf, 0.f);
    orc_make_dam_break(r, dims, lmin, 0.1f * r, 7, bmin, bmax, 2, fl.data(), &nf, bd.data(), &nb);
    std::mt19937 rng(3);
    std::normal_distribution<float> nd(0.f, 0.2f);
    for (uint32_t i = 0; i < nf; ++i)
        for (int a = 0; a < 3; ++a) vel[4 * (size_t)i + a] = nd(rng);
    wcsph_scene sc{};
    sc.n_fluid = nf;
    sc.n_boundary = nb;
    sc.fluid_pos = fl.data();
    sc.fluid_vel = vel.data();
    sc.boundary_pos = bd.data();
    sc.box = aabb(bmin, bmax, 0.f);
    void* h = orc_create(precision, &prm, &sc, parallel);
    orc_step(h, 1e-4, 3);
    std::vector<double> pos(4 * (size_t)nf), v(4 * (size_t)nf), rho(nf), p(nf), ap(4 * (size_t)nf),
        anp(4 * (size_t)nf), nrm(4 * (size_t)nf), bpsi(4 * (size_t)nb);
    std::vector<uint32_t> ids(nf), perm(nf), counts(nf);
    orc_get_state(h, pos.data(), v.data(), rho.data(), p.data(), ids.data());
    orc_get_boundary(h, bpsi.data());
    orc_get_perm(h, perm.data());
    for (int which = 0; which < 2; ++which) {
        const uint64_t total = orc_get_neighbors(h, which, counts.data(), nullptr, 0);
        std::vector<uint32_t> lst(total + 1);
        expect(orc_get_neighbors(h, which, counts.data(), lst.data(), lst.size()) == total, "neighbour total");
    }
    double t = 0, last = 0;
    orc_get_time(h, &t, &last);
    expect(t > 0 && last > 0, "time advances");
    // a rotated box, an exact restart, then an evaluation with overrides
    const wcsph_box tilted = aabb(bmin, bmax, 0.05f);
    orc_set_box(h, &tilted);
    orc_set_state64(h, pos.data(), v.data());
    std::vector<double> rin(nf, 1100.0), pin(nf, 20.0);
    orc_eval(h, rin.data(), pin.data(), rho.data(), p.data(), ap.data(), anp.data(), nrm.data());
    orc_eval(h, nullptr, nullptr, rho.data(), p.data(), ap.data(), anp.data(), nrm.data());
    orc_step(h, 2e-4, 2);
    // the finer fluid grids (walk range 2) and back
    for (float m : {2.f / 3.f, 0.5f, 1.f}) {
        orc_set_cell_size_mult(h, m);
        orc_step(h, 2e-4, 1);
        const uint64_t total = orc_get_neighbors(h, 0, counts.data(), nullptr, 0);
        std::vector<uint32_t> lst(total + 1);
        expect(orc_get_neighbors(h, 0, counts.data(), lst.data(), lst.size()) == total, "neighbour total (mult)");
    }
    orc_get_state(h, pos.data(), v.data(), rho.data(), p.data(), ids.data());
    for (double x : pos) expect(std::isfinite(x), "finite positions");
    orc_destroy(h);
}

void run_adjoint() {
    const wcsph_params prm = preset(false);
    const uint32_t dims[3] = {5, 5, 5};
    const float lmin[3] = {0.1f, 0.1f, 0.1f};
    const float bmin[3] = {0.f, 0.f, 0.f};
    const float bmax[3] = {0.4f, 0.4f, 0.4f};
    uint32_t nf = 0, nb = 0;
    orc_make_dam_break(prm.particle_radius, dims, lmin, 0.001f, 5, bmin, bmax, 1, nullptr, &nf, nullptr, &nb);
    std::vector<float> fl(4 * (size_t)nf), bd(4 * (size_t)nb);
    orc_make_dam_break(prm.particle_radius, dims, lmin, 0.001f, 5, bmin, bmax, 1, fl.data(), &nf, bd.data(), &nb);
    wcsph_scene sc{};
    sc.n_fluid = nf;
    sc.n_boundary = nb;
    sc.fluid_pos = fl.data();
    sc.boundary_pos = bd.data();
    sc.box = aabb(bmin, bmax, 0.f);
    void* h = orc_create(64, &prm, &sc, 1);
    orc_set_viscosity(h, 0.02);
    std::vector<double> targets(4 * (size_t)nf);
    for (uint32_t i = 0; i < nf; ++i)
        for (int a = 0; a < 4; ++a) targets[4 * (size_t)i + a] = fl[4 * (size_t)i + a] + 0.001;
    const double L = orc_grad_run(h, 2e-4, 4, targets.data());
    std::vector<double> gv(4 * (size_t)nf);
    double ga = 0.0;
    orc_grad_backward(h, targets.data(), gv.data(), &ga);
    expect(std::isfinite(L) && std::isfinite(ga), "finite adjoint");
    orc_destroy(h);
}

void run_primitives() {
    std::mt19937 rng(11);
    std::uniform_real_distribution<float> u(-0.3f, 0.3f);
    const uint32_t n = 600;
    std::vector<float> pts(4 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) pts[4 * (size_t)i + a] = u(rng);
        pts[4 * (size_t)i + 3] = 1.f;
    }
    const float cs = 0.04f;
    std::vector<uint32_t> order(n);
    orc_sort_order_f32(pts.data(), n, cs, order.data());
    std::vector<float> sorted(4 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k) std::memcpy(&sorted[4 * (size_t)k], &pts[4 * (size_t)order[k]], 16);
    std::vector<uint64_t> keys(n);
    orc_morton_keys_f32(sorted.data(), n, cs, keys.data());
    for (uint32_t k = 1; k < n; ++k) expect(keys[k - 1] <= keys[k], "keys sorted");
    std::vector<uint32_t> counts(n);
    const uint64_t total = orc_neighbors_f32(sorted.data(), n, sorted.data(), n, cs, cs, 1, counts.data(), nullptr, 0);
    std::vector<uint32_t> lst(total + 1);
    orc_neighbors_f32(sorted.data(), n, sorted.data(), n, cs, cs, 1, counts.data(), lst.data(), lst.size());
    expect(orc_neighbors_f32(pts.data(), n, pts.data(), n, cs, cs, 1, counts.data(), nullptr, 0) == UINT64_MAX,
           "unsorted sources refused");
    for (double q : {0.0, 0.25, 0.5, 0.75, 1.0, 1.5}) {
        expect(std::isfinite(orc_cubic_w(q, 0.04) + orc_cubic_grad(q, 0.04) + orc_spiky_ref_grad(q, 0.04) +
                             orc_cohesion(q, 0.04)),
               "kernel values");
    }
    expect(orc_tait(1200.0, 1000.0, 1000.0, 7.0, 1.5) > 0.0, "tait");
    expect(orc_morton_key(-(1 << 20), -(1 << 20), -(1 << 20)) == 0, "morton");
    expect(orc_cell_coord_f32(-1e-9f, 0.04f) == -1, "cell coordinate");
}

}  // namespace

int main() {
    run_primitives();
    for (int reference = 0; reference < 2; ++reference)
        for (int precision : {64, 32})
            for (int parallel = 0; parallel < 2; ++parallel) run_scene(reference != 0, precision, parallel);
    run_adjoint();
    if (failures) return 1;
    std::printf("san_check ok\n");
    return 0;
}
