// CPU oracle of the WCSPH solver step.  TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker or the timed CPU baseline;
// the product path (libwcsph.so) never calls it.
//
// It restates the reference algorithm on the CPU, templated on the real
// type: T = double is the FP64 oracle of the step, T = float with OpenMP is
// the CPU baseline, and the *_f32 free functions are the bit-exact
// statement of the reference's keys, sort order and neighbour enumeration.
// Every function cites the reference file:line it follows (paths relative
// to the reference repository root).
//
// Parity pins (see tests/test_oracle.py): Morton key vs a closed-form bit
// loop and special cases; sort order vs a brute-force order by
// (key, previous index); neighbour sets vs O(N^2) brute force; cubic spline
// W vs its integral (= 1) and W(0); gradients vs finite differences of W;
// EOS clamps; boundary mass vs an independent FP64 lattice sum; free fall vs
// the closed-form symplectic-Euler sums; box collisions per face; CFL rule
// branches; momentum conservation of the pairwise fluid forces.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <parallel/algorithm>
#include <type_traits>
#include <vector>
#include <omp.h>

#include "../include/wcsph.h"

namespace {

// ---------------------------------------------------------------------
// Morton key (src/cuda/nsearch/morton.cuh:7-34).  The reference spreads the
// low 21 bits of c + 2^20 with magic masks; restated here as the plain
// bit-by-bit interleave it is equivalent to.
// ---------------------------------------------------------------------
uint64_t spread21(uint32_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 21; ++b)
        out |= (uint64_t)((v >> b) & 1u) << (3 * b);
    return out;
}

uint64_t morton_key(int cx, int cy, int cz) {
    const uint32_t off = 1u << 20;  // signed_to_unsigned, morton.cuh:18-21
    return spread21((uint32_t)cx + off) | (spread21((uint32_t)cy + off) << 1) |
           (spread21((uint32_t)cz + off) << 2);
}

// Cell coordinate: static_cast<int>(floorf(p / cs)) (particle_data.cuh:117-123,
// common.cuh:73-79), with IEEE division in the working precision.
inline int cell_of(float p, float cs) { return (int)std::floor(p / cs); }
inline int cell_of(double p, double cs) { return (int)std::floor(p / cs); }

// Neighbour predicate (common.cuh:39-47, math.cuh:54-56): dot(r, r) <= h*h,
// with the dot product contracted as fma(z, z, fma(y, y, x*x)).
template <class T>
inline T dist2(T rx, T ry, T rz) {
    return std::fma(rz, rz, std::fma(ry, ry, rx * rx));
}

template <class T> struct K;
template <> struct K<float> { static constexpr float pi = 3.1415926535897932384626433f; };
template <> struct K<double> { static constexpr double pi = 3.1415926535897932384626433; };

// Cubic spline W (src/cuda/SPH/kernel.cuh:12-19).
template <class T>
inline T cubic_w(T q, T h) {
    const T factor = T(8) / (K<T>::pi * h * h * h);
    if (q <= T(0.5)) return factor * (T(6) * q * q * q - T(6) * q * q + T(1));
    if (q <= T(1)) {
        const T a = T(1) - q;
        return factor * T(2) * (a * a * a);
    }
    return T(0);
}

// Cubic spline gradient coefficient (kernel.cuh:22-32): grad W = coef * r.
template <class T>
inline T cubic_grad(T q, T h) {
    if (q < T(1e-9) || q > T(1)) return T(0);
    const T gf = T(48) / (K<T>::pi * h * h * h);
    const T rl = q * h;
    if (q <= T(0.5)) return gf * q * (T(3) * q - T(2)) / (h * rl);
    return -gf * (T(1) - q) * (T(1) - q) / (h * rl);
}

// Spiky gradient with the reference's h^-9 normalisation (kernel.cuh:43-51).
template <class T>
inline T spiky_ref_grad(T q, T h) {
    if (q < T(1e-9) || q > T(1)) return T(0);
    const T h9 = std::pow(h, T(9));
    const T factor = T(-45) / (K<T>::pi * h9);
    const T rl = q * h;
    return factor * (T(1) - q) * (T(1) - q) / (h * rl);
}

// Cohesion kernel (kernel.cuh:53-63).
template <class T>
inline T cohesion(T q, T h) {
    const T factor = T(32) / (K<T>::pi * h * h * h);
    const T a = T(1) - q;
    const T iq3 = a * a * a;
    const T q3 = q * q * q;
    if (q <= T(0.5)) return factor * (T(2) * q3 * iq3 - T(1) / T(64));
    if (q <= T(1)) return factor * q3 * iq3;
    return T(0);
}

// Tait EOS with both clamps (compute_rho_p.cu:314-316):
// p = k * (min(max(rho/rho0, 1), max_ratio)^gamma - 1).
template <class T>
inline T tait(T rho, T rho0, T k, T gamma, T max_ratio) {
    T ratio = std::max(rho / rho0, T(1));
    ratio = std::min(ratio, max_ratio);
    return k * (std::pow(ratio, gamma) - T(1));
}

// Column-major 4x4 times (x, y, z, w) (math.cuh:66-87).
template <class T>
inline void mat4_mul(const T* m, const T* v, T* out) {
    for (int r = 0; r < 4; ++r)
        out[r] = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1] + m[2 * 4 + r] * v[2] +
                 m[3 * 4 + r] * v[3];
}

struct CellRun {
    uint64_t key;
    uint32_t start, end;
};

// Runs of equal keys in a key-sorted array: the cell_start/cell_end table
// the reference fills by boundary detection (rebuild_n_search.cu:10-36),
// kept as a key-sorted vector instead of a hash table.
std::vector<CellRun> build_runs(const std::vector<uint64_t>& sorted_keys) {
    std::vector<CellRun> runs;
    const uint32_t n = (uint32_t)sorted_keys.size();
    uint32_t i = 0;
    while (i < n) {
        uint32_t j = i + 1;
        while (j < n && sorted_keys[j] == sorted_keys[i]) ++j;
        runs.push_back({sorted_keys[i], i, j});
        i = j;
    }
    return runs;
}

// find_cell_in_table (common.cuh:89-102): the range of a cell, or empty.
inline bool find_run(const std::vector<CellRun>& runs, uint64_t key, uint32_t& s, uint32_t& e) {
    auto it = std::lower_bound(runs.begin(), runs.end(), key,
                               [](const CellRun& c, uint64_t k) { return c.key < k; });
    if (it == runs.end() || it->key != key) return false;
    s = it->start;
    e = it->end;
    return true;
}

// Stable order by key: thrust::sequence + sort_by_key (particle_data.cuh:146-161),
// assumed stable, so ties keep the previous order.
void stable_order(const std::vector<uint64_t>& keys, std::vector<uint32_t>& order, bool parallel) {
    const uint32_t n = (uint32_t)keys.size();
    order.resize(n);
    std::iota(order.begin(), order.end(), 0u);
    auto cmp = [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; };
    if (parallel && n > 65536)
        __gnu_parallel::stable_sort(order.begin(), order.end(), cmp);
    else
        std::stable_sort(order.begin(), order.end(), cmp);
}

// Counter-based jitter in [-1, 1): splitmix64 of (seed, index, axis).
float jitter_unit(uint64_t seed, uint64_t idx, int axis) {
    uint64_t z = seed + (idx * 3u + (uint64_t)axis + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
    const float two_u = 2.0f * u;
    return two_u - 1.0f;
}

// ---------------------------------------------------------------------
// The step, templated on the working precision.
// ---------------------------------------------------------------------
template <class T>
struct Solver {
    // parameters in T (sph.cuh:19-32, fluid_simulator.h:11-12, common.cuh:18-23)
    // cs: cell size of the boundary grid (h, cell_coord_boundary); fcs and R:
    // cell size and walk range of the fluid grid (CELL_SIZE = h * CELL_SIZE_MULT
    // and RANGE, common.cuh:126-130, step_tuner.cuh:88)
    T r, s, h, h2, cs, fcs, mass, rho0, stiff, gam, maxr, visc, st, g[3];
    float mult = 1.f;
    int R = 1;
    T lam, dtmin, dtmax, dtnpmax, cbuf, elast, csound;
    uint32_t grad_kind, visc_kind, adaptive, split;
    T bmin[3], bmax[3], model[16], model_inv[16];
    wcsph_force_rule rules[WCSPH_MAX_FORCE_RULES];
    uint32_t nrules = 0;
    bool parallel;

    uint32_t n = 0, nb = 0;
    std::vector<T> X, V;  // 4 per fluid particle
    std::vector<uint32_t> ids, perm;
    std::vector<T> B;     // 4 per boundary particle, w = Psi
    std::vector<T> rho, prs, nrm, ap, anp;
    std::vector<uint64_t> keys;
    std::vector<uint32_t> cell_of_particle;
    std::vector<CellRun> fruns, bruns;
    std::vector<uint32_t> fcnt, foff, flist, bcnt, boff, blist;
    std::vector<T> Xsorted;  // positions the last neighbour search saw
    double sim_time = 0.0, last_dt = 0.0;
    uint64_t steps = 0;

    void set_params(const wcsph_params& p, const wcsph_box& box) {
        // Derived quantities are formed in float exactly as the reference's
        // constexpr floats are, then widened (common.cuh:18-23).
        const float rf = p.particle_radius;
        const float sf = 2.f * rf;
        const float hf = 2.f * sf;
        const float volf = sf * sf * sf * 0.8f;
        const float mf = p.rest_density * volf;
        const float h2f = hf * hf;
        r = rf; s = sf; h = hf; cs = hf; mass = mf;
        set_cell_size_mult(mult);
        h2 = std::is_same<T, float>::value ? (T)h2f : (T)hf * (T)hf;
        rho0 = p.rest_density; stiff = p.stiffness; gam = p.exponent;
        maxr = p.max_density_ratio; visc = p.viscosity; st = p.surface_tension;
        for (int a = 0; a < 3; ++a) g[a] = p.gravity[a];
        lam = p.cfl_factor; dtmin = p.dt_min; dtmax = p.dt_max; dtnpmax = p.dt_np_max;
        cbuf = p.collision_buffer; elast = p.elasticity;
        csound = std::sqrt((T)p.stiffness * (T)p.exponent / (T)p.rest_density);
        grad_kind = p.pressure_gradient; visc_kind = p.viscosity_model;
        adaptive = p.adaptive_dt; split = p.split_np_dt;
        nrules = std::min<uint32_t>(p.n_force_rules, WCSPH_MAX_FORCE_RULES);
        for (uint32_t k = 0; k < nrules; ++k) rules[k] = p.force_rules[k];
        set_box(box);
    }

    // Fluid cell size fp32(h) * fp32(m) (the host-side product of the
    // reference's grid, fluid_simulator.cpp's cell size times
    // CELL_SIZE_MULT) and the smallest walk range R with R * cs >= h; the
    // reference's RANGE = (int)(h / cs + 0.5) (common.cuh:130) is the same
    // for its multipliers {0.5, 2/3, 1, 1.5, 2}.
    void set_cell_size_mult(float m) {
        mult = m;
        const float cf = (float)h * m;
        fcs = (T)cf;
        R = (int)std::ceil((double)(float)h / (double)cf - 1e-9);
    }

    void set_box(const wcsph_box& box) {
        for (int a = 0; a < 3; ++a) { bmin[a] = box.min[a]; bmax[a] = box.max[a]; }
        for (int i = 0; i < 16; ++i) { model[i] = box.model[i]; model_inv[i] = box.model_inv[i]; }
    }

    // EXTERNAL_FORCE(x) of the scene (compute_non_pressure_accel.cu:9-11,
    // 109), as the first matching rule: the fountain lambda
    // (app/fountain/application.cu:41-46) is "x*x + z*z < 0.2 and y < 0 ->
    // (0, 25, 0); y < -0.3 -> 2 normalize(-x, 0, -z)", the vortex lambda
    // (app/vortex/application.cu:45) "y < -0.5 -> 2 normalize(z, 0, -x)".
    void ext_force(const T* x, T out[3]) const {
        out[0] = out[1] = out[2] = T(0);
        for (uint32_t k = 0; k < nrules; ++k) {
            const wcsph_force_rule& R = rules[k];
            bool in = R.region == WCSPH_REGION_ALL;
            if (R.region == WCSPH_REGION_Y_BELOW) in = x[1] < (T)R.y_below;
            if (R.region == WCSPH_REGION_XZ_DISC_Y_BELOW)
                in = x[0] * x[0] + x[2] * x[2] < (T)R.radius2 && x[1] < (T)R.y_below;
            if (!in) continue;
            if (R.kind == WCSPH_FORCE_CONSTANT) {
                for (int a = 0; a < 3; ++a) out[a] = (T)R.value[a];
            } else {
                const T dx = R.kind == WCSPH_FORCE_TOWARD_Y_AXIS ? -x[0] : x[2];
                const T dz = R.kind == WCSPH_FORCE_TOWARD_Y_AXIS ? -x[2] : -x[0];
                const T len = std::sqrt(dx * dx + dz * dz);
                if (len > T(0)) {
                    out[0] = (T)R.magnitude * dx / len;
                    out[2] = (T)R.magnitude * dz / len;
                }
            }
            return;
        }
    }

    // sort_boundary + build_boundary_n_search + compute_boundary_mass
    // (particle_data.cuh:84-110, sph.cu:84-103, compute_boundary_mass.cu:10-30).
    void init_boundary(const float* bpos, uint32_t count) {
        nb = count;
        std::vector<uint64_t> bk(nb);
        for (uint32_t i = 0; i < nb; ++i)
            bk[i] = morton_key(cell_of((T)bpos[4 * i], cs), cell_of((T)bpos[4 * i + 1], cs),
                               cell_of((T)bpos[4 * i + 2], cs));
        std::vector<uint32_t> ord;
        stable_order(bk, ord, parallel);
        B.assign(4 * (size_t)nb, T(0));
        std::vector<uint64_t> sk(nb);
        for (uint32_t k = 0; k < nb; ++k) {
            for (int a = 0; a < 3; ++a) B[4 * k + a] = (T)bpos[4 * ord[k] + a];
            sk[k] = bk[ord[k]];
        }
        bruns = build_runs(sk);
        std::vector<T> psi(nb);
#pragma omp parallel for schedule(dynamic, 256) if (parallel)
        for (int64_t i = 0; i < (int64_t)nb; ++i) {
            const T* xi = &B[4 * i];
            const int c0 = cell_of(xi[0], cs), c1 = cell_of(xi[1], cs), c2 = cell_of(xi[2], cs);
            T sum = cubic_w(T(0), h);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        uint32_t st0, en0;
                        if (!find_run(bruns, morton_key(c0 + dx, c1 + dy, c2 + dz), st0, en0)) continue;
                        for (uint32_t j = st0; j < en0; ++j) {
                            if ((int64_t)j == i) continue;
                            const T rx = xi[0] - B[4 * j], ry = xi[1] - B[4 * j + 1], rz = xi[2] - B[4 * j + 2];
                            const T d2 = dist2(rx, ry, rz);
                            if (d2 <= h2) sum += cubic_w(std::sqrt(d2) / h, h);
                        }
                    }
            psi[i] = rho0 / sum;
        }
        for (uint32_t i = 0; i < nb; ++i) B[4 * i + 3] = psi[i];
    }

    void set_fluid(const float* pos, const float* vel, uint32_t count) {
        n = count;
        X.assign(4 * (size_t)n, T(0));
        V.assign(4 * (size_t)n, T(0));
        ids.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                X[4 * i + a] = (T)pos[4 * i + a];
                V[4 * i + a] = vel ? (T)vel[4 * i + a] : T(0);
            }
            X[4 * i + 3] = T(1);
            ids[i] = i;
        }
        rho.assign(n, T(0)); prs.assign(n, T(0));
        nrm.assign(4 * (size_t)n, T(0)); ap.assign(4 * (size_t)n, T(0)); anp.assign(4 * (size_t)n, T(0));
    }

    // adapt_time_step (sph.cu:115-130), max |v|^2 and dt in FP64.
    double adapt_dt(double hint) const {
        if (!adaptive) return hint;
        double vmax2 = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double vx = (double)V[4 * i], vy = (double)V[4 * i + 1], vz = (double)V[4 * i + 2];
            vmax2 = std::max(vmax2, vx * vx + vy * vy + vz * vz);
        }
        const double vmax = std::sqrt(vmax2);
        if (vmax < 1e-9) return (double)dtmax;
        const double cfl = (double)lam * (double)s / vmax;
        return std::min(std::clamp(hint, (double)dtmin, (double)dtmax), cfl);
    }

    // ParticleData::sort (particle_data.cuh:146-167): keys of the current
    // order, stable sort, gather x and v (and the particle ids).
    void sort() {
        keys.resize(n);
#pragma omp parallel for if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i)
            keys[i] = morton_key(cell_of(X[4 * i], fcs), cell_of(X[4 * i + 1], fcs), cell_of(X[4 * i + 2], fcs));
        stable_order(keys, perm, parallel);
        std::vector<T> X2(X.size()), V2(V.size());
        std::vector<uint32_t> id2(n);
        std::vector<uint64_t> k2(n);
#pragma omp parallel for if (parallel)
        for (int64_t k = 0; k < (int64_t)n; ++k) {
            const uint32_t i = perm[k];
            for (int a = 0; a < 4; ++a) { X2[4 * k + a] = X[4 * i + a]; V2[4 * k + a] = V[4 * i + a]; }
            id2[k] = ids[i];
            k2[k] = keys[i];
        }
        X.swap(X2); V.swap(V2); ids.swap(id2); keys.swap(k2);
        fruns = build_runs(keys);
        cell_of_particle.resize(n);
        for (uint32_t c = 0; c < (uint32_t)fruns.size(); ++c)
            for (uint32_t i = fruns[c].start; i < fruns[c].end; ++i) cell_of_particle[i] = c;
    }

    // count_neighbors / exclusive_scan / fill_neighbors
    // (compute_neighbor_list.cu:10-44, step_tuner.cuh:281-288) for the fluid,
    // and the boundary set for_boundary_neighbors visits (common.cuh:153-178).
    // Enumeration order: dx, dy, dz in [-R, R] on the fluid grid (in [-1, 1]
    // on the boundary grid), then ascending index.
    void build_neighbors() {
        const uint32_t nc = (uint32_t)fruns.size();
        const int W = 2 * R + 1, T3 = W * W * W;
        // (2R+1)^3 fluid ranges per occupied fluid cell
        std::vector<uint32_t> fr((size_t)nc * 2 * T3);
#pragma omp parallel for schedule(dynamic, 64) if (parallel)
        for (int64_t c = 0; c < (int64_t)nc; ++c) {
            const uint32_t i0 = fruns[c].start;
            const int c0 = cell_of(X[4 * i0], fcs), c1 = cell_of(X[4 * i0 + 1], fcs), c2 = cell_of(X[4 * i0 + 2], fcs);
            int t = 0;
            for (int dx = -R; dx <= R; ++dx)
                for (int dy = -R; dy <= R; ++dy)
                    for (int dz = -R; dz <= R; ++dz, ++t) {
                        uint32_t s0 = 0, e0 = 0;
                        find_run(fruns, morton_key(c0 + dx, c1 + dy, c2 + dz), s0, e0);
                        fr[(c * T3 + t) * 2] = s0; fr[(c * T3 + t) * 2 + 1] = e0;
                    }
        }
        // 27 boundary ranges per particle (its own cell on the boundary grid);
        // with the fluid grid equal to the boundary grid, per occupied cell
        const bool same = fcs == cs;
        const size_t nbr = same ? nc : n;
        std::vector<uint32_t> br(nbr * 54);
#pragma omp parallel for schedule(dynamic, 256) if (parallel)
        for (int64_t i = 0; i < (int64_t)nbr; ++i) {
            const T* xi = &X[4 * (same ? (int64_t)fruns[i].start : i)];
            const int c0 = cell_of(xi[0], cs), c1 = cell_of(xi[1], cs), c2 = cell_of(xi[2], cs);
            int t = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz, ++t) {
                        uint32_t s0 = 0, e0 = 0;
                        find_run(bruns, morton_key(c0 + dx, c1 + dy, c2 + dz), s0, e0);
                        br[i * 54 + 2 * t] = s0; br[i * 54 + 2 * t + 1] = e0;
                    }
        }
        fcnt.assign(n, 0); bcnt.assign(n, 0);
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) {
                foff.resize(n); boff.resize(n);
                uint64_t a = 0, b = 0;
                for (uint32_t i = 0; i < n; ++i) { foff[i] = (uint32_t)a; a += fcnt[i]; boff[i] = (uint32_t)b; b += bcnt[i]; }
                flist.resize(a); blist.resize(b);
            }
#pragma omp parallel for schedule(dynamic, 256) if (parallel)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                const T* xi = &X[4 * i];
                const size_t c = cell_of_particle[i];
                uint32_t cf = 0, cb = 0;
                for (int t = 0; t < T3; ++t) {
                    for (uint32_t j = fr[(c * T3 + t) * 2]; j < fr[(c * T3 + t) * 2 + 1]; ++j) {
                        if ((int64_t)j == i) continue;
                        if (dist2(xi[0] - X[4 * j], xi[1] - X[4 * j + 1], xi[2] - X[4 * j + 2]) <= h2) {
                            if (pass == 1) flist[foff[i] + cf] = j;
                            ++cf;
                        }
                    }
                }
                const size_t bi = same ? c : (size_t)i;
                for (int t = 0; t < 27; ++t) {
                    for (uint32_t j = br[bi * 54 + 2 * t]; j < br[bi * 54 + 2 * t + 1]; ++j) {
                        if (dist2(xi[0] - B[4 * j], xi[1] - B[4 * j + 1], xi[2] - B[4 * j + 2]) <= h2) {
                            if (pass == 1) blist[boff[i] + cb] = j;
                            ++cb;
                        }
                    }
                }
                if (pass == 0) { fcnt[i] = cf; bcnt[i] = cb; }
            }
        }
        Xsorted = X;
    }

    // compute_rho_p_with_boundary (compute_rho_p.cu:84-101, 291-317).
    void density() {
        const T w0 = cubic_w(T(0), h);
#pragma omp parallel for schedule(static) if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const T* xi = &X[4 * i];
            T d = w0;
            for (uint32_t k = foff[i]; k < foff[i] + fcnt[i]; ++k) {
                const uint32_t j = flist[k];
                const T q = std::sqrt(dist2(xi[0] - X[4 * j], xi[1] - X[4 * j + 1], xi[2] - X[4 * j + 2])) / h;
                d += cubic_w(q, h);
            }
            d = d * mass;
            for (uint32_t k = boff[i]; k < boff[i] + bcnt[i]; ++k) {
                const uint32_t j = blist[k];
                const T q = std::sqrt(dist2(xi[0] - B[4 * j], xi[1] - B[4 * j + 1], xi[2] - B[4 * j + 2])) / h;
                d += cubic_w(q, h) * B[4 * j + 3];
            }
            rho[i] = d;
            prs[i] = tait(d, rho0, stiff, gam, maxr);
        }
    }

    inline T pgrad(T q) const { return grad_kind == WCSPH_GRAD_SPIKY_REF ? spiky_ref_grad(q, h) : cubic_grad(q, h); }

    // Pressure acceleration (+ normals for surface tension):
    // compute_pressure_accel_n_normal[_with_boundary]
    // (compute_pressure_accel_n_normal.cu:98-124, 337-365).
    // Non-pressure acceleration: compute_non_pressure_accel
    // (compute_non_pressure_accel.cu:113-145) for the Laplacian model; the
    // artificial-viscosity model replaces its viscosity term with
    // -m * sum_j Pi_ij grad W_ij, Pi_ij = -nu_ij (v_ij.x_ij)/(r^2 + 0.01 h^2)
    // for v_ij.x_ij < 0, nu_ij = 2 alpha h c_s / (rho_i + rho_j),
    // c_s = sqrt(k gamma / rho0).
    void forces() {
        const bool need_normals = st != T(0);
        if (need_normals) {
#pragma omp parallel for schedule(static) if (parallel)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                const T* xi = &X[4 * i];
                T nx = 0, ny = 0, nz = 0;
                for (uint32_t k = foff[i]; k < foff[i] + fcnt[i]; ++k) {
                    const uint32_t j = flist[k];
                    const T rx = xi[0] - X[4 * j], ry = xi[1] - X[4 * j + 1], rz = xi[2] - X[4 * j + 2];
                    const T q = std::sqrt(dist2(rx, ry, rz)) / h;
                    const T c = cubic_grad(q, h) / rho[j];
                    nx += rx * c; ny += ry * c; nz += rz * c;
                }
                const T f = h * mass;
                nrm[4 * i] = f * nx; nrm[4 * i + 1] = f * ny; nrm[4 * i + 2] = f * nz;
            }
        }
#pragma omp parallel for schedule(static) if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const T* xi = &X[4 * i];
            const T* vi = &V[4 * i];
            const T di = rho[i];
            const T dpi = prs[i] / (di * di);
            T pax = 0, pay = 0, paz = 0;  // pressure sum
            T vlx = 0, vly = 0, vlz = 0;  // viscosity sum
            T fx = 0, fy = 0, fz = 0;     // surface tension sum
            for (uint32_t k = foff[i]; k < foff[i] + fcnt[i]; ++k) {
                const uint32_t j = flist[k];
                const T rx = xi[0] - X[4 * j], ry = xi[1] - X[4 * j + 1], rz = xi[2] - X[4 * j + 2];
                const T r2 = dist2(rx, ry, rz);
                const T q = std::sqrt(r2) / h;
                const T dj = rho[j];
                const T dpj = prs[j] / (dj * dj);
                const T gp = pgrad(q);
                pax -= (dpi + dpj) * gp * rx; pay -= (dpi + dpj) * gp * ry; paz -= (dpi + dpj) * gp * rz;
                const T vx = vi[0] - V[4 * j], vy = vi[1] - V[4 * j + 1], vz = vi[2] - V[4 * j + 2];
                const T vr = vx * rx + vy * ry + vz * rz;
                const T gc = cubic_grad(q, h);
                if (visc_kind == WCSPH_VISC_LAPLACIAN) {
                    const T c = vr * gc / (dj * (r2 + T(0.01) * h * h));
                    vlx += c * rx; vly += c * ry; vlz += c * rz;
                } else if (vr < T(0)) {
                    const T nu = T(2) * visc * h * csound / (di + dj);
                    const T pi_ij = -nu * vr / (r2 + T(0.01) * h * h);
                    vlx -= pi_ij * gc * rx; vly -= pi_ij * gc * ry; vlz -= pi_ij * gc * rz;
                }
                if (need_normals && r2 > T(1e-6)) {
                    const T rl = std::sqrt(r2);
                    const T cq = cohesion(q, h);
                    const T inv = T(1) / (di + dj);
                    fx += (mass * (rx / rl) * cq + nrm[4 * i] - nrm[4 * j]) * inv;
                    fy += (mass * (ry / rl) * cq + nrm[4 * i + 1] - nrm[4 * j + 1]) * inv;
                    fz += (mass * (rz / rl) * cq + nrm[4 * i + 2] - nrm[4 * j + 2]) * inv;
                }
            }
            T a_p[3] = {mass * pax, mass * pay, mass * paz};
            const T bterm = dpi + prs[i] / (rho0 * rho0);
            for (uint32_t k = boff[i]; k < boff[i] + bcnt[i]; ++k) {
                const uint32_t j = blist[k];
                const T rx = xi[0] - B[4 * j], ry = xi[1] - B[4 * j + 1], rz = xi[2] - B[4 * j + 2];
                const T q = std::sqrt(dist2(rx, ry, rz)) / h;
                const T gp = pgrad(q);
                const T c = B[4 * j + 3] * bterm * gp;
                a_p[0] -= c * rx; a_p[1] -= c * ry; a_p[2] -= c * rz;
            }
            T vfac = visc_kind == WCSPH_VISC_LAPLACIAN ? visc * T(10) * mass : mass;
            T sfac = st * T(2) * rho0;
            ap[4 * i] = a_p[0]; ap[4 * i + 1] = a_p[1]; ap[4 * i + 2] = a_p[2];
            T ext[3];
            ext_force(xi, ext);
            anp[4 * i] = (g[0] + ext[0]) + vfac * vlx - sfac * fx;
            anp[4 * i + 1] = (g[1] + ext[1]) + vfac * vly - sfac * fy;
            anp[4 * i + 2] = (g[2] + ext[2]) + vfac * vlz - sfac * fz;
        }
    }

    // update_positions + resolve_collisions (update_positions.cu:14-68).
    // coll (optional) receives per particle 2 bits per axis: 1 = low wall,
    // 2 = high wall.
    void integrate(T dt, T dt_np, std::vector<uint8_t>* coll = nullptr) {
        if (coll) coll->assign(n, 0);
#pragma omp parallel for schedule(static) if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            T* x = &X[4 * i];
            T* v = &V[4 * i];
            for (int a = 0; a < 3; ++a) v[a] += dt_np * anp[4 * i + a] + dt * ap[4 * i + a];
            for (int a = 0; a < 3; ++a) x[a] += dt * v[a];
            T xp[4] = {x[0], x[1], x[2], T(1)}, vp[4] = {v[0], v[1], v[2], T(0)};
            T bx[4], bv[4];
            mat4_mul(model_inv, xp, bx);
            mat4_mul(model_inv, vp, bv);
            bool changed = false;
            uint8_t flags = 0;
            for (int a = 0; a < 3; ++a) {
                if (bx[a] - r * cbuf < bmin[a]) { bx[a] = bmin[a] + r; bv[a] *= -elast; changed = true; flags |= (uint8_t)(1u << (2 * a)); }
                else if (bx[a] + r * cbuf > bmax[a]) { bx[a] = bmax[a] - r; bv[a] *= -elast; changed = true; flags |= (uint8_t)(2u << (2 * a)); }
            }
            if (coll) (*coll)[i] = flags;
            if (changed) {
                T wx[4], wv[4];
                mat4_mul(model, bx, wx);
                mat4_mul(model, bv, wv);
                for (int a = 0; a < 3; ++a) { x[a] = wx[a]; v[a] = wv[a]; }
            }
        }
    }

    // The passes of one step up to the accelerations, without integration
    // (sph.cu:31-60 minus update_positions): sort, lists, density + EOS,
    // then the force pass.  rho_in / p_in (per particle id, may be NULL)
    // replace the density pass's rho and p before the force pass reads them.
    void eval(const double* rho_in, const double* p_in) {
        sort();
        build_neighbors();
        density();
        for (uint32_t i = 0; i < n; ++i) {
            if (rho_in) rho[i] = (T)rho_in[ids[i]];
            if (p_in) prs[i] = (T)p_in[ids[i]];
        }
        forces();
    }

    // CUDASPHSimulator::update (sph.cu:31-60) + StepTuner::launch order
    // (step_tuner.cuh:259-295).
    void step(double hint_in) {
        // the reference's update(float delta) takes the hint as a float
        const double hint = (double)(float)hint_in;
        const double dt_np_d = split ? std::min(hint, (double)dtnpmax) : 0.0;
        const double dt_d = adapt_dt(hint);
        const T dt = (T)dt_d;
        const T dt_np = split ? (T)dt_np_d : dt;
        sort();
        build_neighbors();
        density();
        forces();
        if (taping) {
            tape.emplace_back();
            StepTape& tp = tape.back();
            tp.X = X; tp.V = V; tp.rho = rho; tp.prs = prs; tp.perm = perm;
            tp.fcnt = fcnt; tp.foff = foff; tp.flist = flist;
            tp.bcnt = bcnt; tp.boff = boff; tp.blist = blist;
            tp.dt = dt;
            integrate(dt, dt_np, &tp.coll);
        } else {
            integrate(dt, dt_np);
        }
        sim_time += (double)dt;
        last_dt = (double)dt;
        ++steps;
    }

    // ------------------------------------------------------------------
    // Reverse mode of the step for the benchmark physics (cubic-spline
    // gradients, artificial viscosity, Akinci boundary pressure, one dt,
    // identity box).  Neighbour sets, the sort and the collision decisions
    // are held fixed per step (the cubic spline is C^1 at q = 1, so pairs
    // entering or leaving the support contribute nothing at the switch).
    // Gather form: every particle sums the derivative of each pair term
    // with respect to its own variables, so no scatter is needed.
    // ------------------------------------------------------------------
    struct StepTape {
        std::vector<T> X, V, rho, prs;
        std::vector<uint32_t> perm, fcnt, foff, flist, bcnt, boff, blist;
        std::vector<uint8_t> coll;
        T dt;
    };
    std::vector<StepTape> tape;
    bool taping = false;

    // d/dR of the cubic-spline gradient coefficient gc(R) (grad W = gc * r)
    T gc_prime(T R) const {
        const T q = R / h;
        if (q < T(1e-9) || q > T(1)) return T(0);
        if (q <= T(0.5)) return T(3) * T(48) / (K<T>::pi * std::pow(h, T(6)));
        const T b0 = T(48) / (K<T>::pi * std::pow(h, T(4)));
        const T a = T(1) - q;
        return b0 * (T(2) * a / (h * R) + a * a / (R * R));
    }

    // lx, lv: adjoints of the state after the step (this step's sorted
    // order), 4 per particle; on return, adjoints of the state before the
    // step (the previous order).  galpha accumulates dL/d(alpha).
    void step_backward(const StepTape& tp, std::vector<T>& lx, std::vector<T>& lv, double& galpha) {
        const T dt = tp.dt;
        const T csd = std::sqrt(stiff * gam / rho0);
        const T dcoef = T(2) * h * csd;  // d(C_alpha)/d(alpha), C_alpha = 2 alpha h c_s
        const T calpha = visc * dcoef;
        const T eps = T(0.01) * h * h;
        std::vector<T> la(4 * (size_t)n, T(0)), lvn(4 * (size_t)n), ly(4 * (size_t)n);
        // collisions, y = x + dt v_new, v_new = v + dt a
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                const uint8_t c = (tp.coll[i] >> (2 * a)) & 3u;
                const T lyo = c ? T(0) : lx[4 * i + a];
                const T lvo = c ? -elast * lv[4 * i + a] : lv[4 * i + a];
                ly[4 * i + a] = lyo;
                lvn[4 * i + a] = lvo + dt * lyo;
                la[4 * i + a] = dt * lvn[4 * i + a];
            }
        std::vector<T> lrho(n, T(0)), gx(4 * (size_t)n, T(0)), gv(4 * (size_t)n, T(0));
        double galpha2 = 0.0;
        // pass A: pressure, viscosity and boundary pair terms
#pragma omp parallel for schedule(static) reduction(+ : galpha2) if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const T* xi = &tp.X[4 * i];
            const T* vi = &tp.V[4 * i];
            const T ri = tp.rho[i], pi_ = tp.prs[i];
            const T Pi = pi_ / (ri * ri);
            T gP = 0, gB = 0, grv = 0, ga = 0;
            T g[3] = {0, 0, 0}, w[3] = {0, 0, 0};
            for (uint32_t k = tp.foff[i]; k < tp.foff[i] + tp.fcnt[i]; ++k) {
                const uint32_t j = tp.flist[k];
                const T r[3] = {xi[0] - tp.X[4 * j], xi[1] - tp.X[4 * j + 1], xi[2] - tp.X[4 * j + 2]};
                const T r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                const T R = std::sqrt(r2);
                const T gq = cubic_grad(R / h, h), gpr = gc_prime(R);
                const T L[3] = {la[4 * i] - la[4 * j], la[4 * i + 1] - la[4 * j + 1], la[4 * i + 2] - la[4 * j + 2]};
                const T t = L[0] * r[0] + L[1] * r[1] + L[2] * r[2];
                const T Pj = tp.prs[j] / (tp.rho[j] * tp.rho[j]);
                const T S = Pi + Pj;
                gP += -mass * gq * t;
                for (int a = 0; a < 3; ++a) g[a] += -mass * S * (gq * L[a] + (R > 0 ? t * gpr * r[a] / R : T(0)));
                const T vij[3] = {vi[0] - tp.V[4 * j], vi[1] - tp.V[4 * j + 1], vi[2] - tp.V[4 * j + 2]};
                const T sv = vij[0] * r[0] + vij[1] * r[1] + vij[2] * r[2];
                if (sv < T(0)) {
                    const T D = ri + tp.rho[j], E = r2 + eps;
                    const T Kc = mass * calpha / D;
                    const T f = Kc * sv * gq * t / E;
                    ga += mass * dcoef / D * sv * gq * t / E;
                    grv += -f / D;
                    for (int a = 0; a < 3; ++a) {
                        w[a] += Kc * gq * t * r[a] / E;
                        g[a] += Kc / E * (gq * t * vij[a] + (R > 0 ? sv * t * gpr * r[a] / R : T(0)) + sv * gq * L[a]) -
                                T(2) * Kc * sv * gq * t * r[a] / (E * E);
                    }
                }
            }
            const T Bi = Pi + pi_ / (rho0 * rho0);
            for (uint32_t k = tp.boff[i]; k < tp.boff[i] + tp.bcnt[i]; ++k) {
                const uint32_t b = tp.blist[k];
                const T r[3] = {xi[0] - B[4 * b], xi[1] - B[4 * b + 1], xi[2] - B[4 * b + 2]};
                const T R = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const T gq = cubic_grad(R / h, h), gpr = gc_prime(R);
                const T psi = B[4 * b + 3];
                const T t = la[4 * i] * r[0] + la[4 * i + 1] * r[1] + la[4 * i + 2] * r[2];
                gB += -psi * gq * t;
                for (int a = 0; a < 3; ++a)
                    g[a] += -psi * Bi * (gq * la[4 * i + a] + (R > 0 ? t * gpr * r[a] / R : T(0)));
            }
            // Tait derivative (zero at both clamps)
            const T ratio = ri / rho0;
            const T dp = (ratio > T(1) && ratio < maxr) ? stiff * gam * std::pow(ratio, gam - T(1)) / rho0 : T(0);
            const T dP = dp / (ri * ri) - T(2) * pi_ / (ri * ri * ri);
            const T dB = dP + dp / (rho0 * rho0);
            lrho[i] = grv + gP * dP + gB * dB;
            for (int a = 0; a < 3; ++a) {
                gx[4 * i + a] = g[a];
                gv[4 * i + a] = w[a];
            }
            galpha2 += (double)ga;
        }
        galpha += 0.5 * galpha2;
        // pass B: density sums
#pragma omp parallel for schedule(static) if (parallel)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const T* xi = &tp.X[4 * i];
            T g[3] = {0, 0, 0};
            for (uint32_t k = tp.foff[i]; k < tp.foff[i] + tp.fcnt[i]; ++k) {
                const uint32_t j = tp.flist[k];
                const T r[3] = {xi[0] - tp.X[4 * j], xi[1] - tp.X[4 * j + 1], xi[2] - tp.X[4 * j + 2]};
                const T R = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const T c = mass * cubic_grad(R / h, h) * (lrho[i] + lrho[j]);
                for (int a = 0; a < 3; ++a) g[a] += c * r[a];
            }
            for (uint32_t k = tp.boff[i]; k < tp.boff[i] + tp.bcnt[i]; ++k) {
                const uint32_t b = tp.blist[k];
                const T r[3] = {xi[0] - B[4 * b], xi[1] - B[4 * b + 1], xi[2] - B[4 * b + 2]};
                const T R = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const T c = B[4 * b + 3] * cubic_grad(R / h, h) * lrho[i];
                for (int a = 0; a < 3; ++a) g[a] += c * r[a];
            }
            for (int a = 0; a < 3; ++a) gx[4 * i + a] += g[a];
        }
        // back to the order before this step's sort
        std::vector<T> lx0(4 * (size_t)n, T(0)), lv0(4 * (size_t)n, T(0));
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = tp.perm[k];
            for (int a = 0; a < 3; ++a) {
                lx0[4 * i + a] = ly[4 * k + a] + gx[4 * k + a];
                lv0[4 * i + a] = lvn[4 * k + a] + gv[4 * k + a];
            }
        }
        lx.swap(lx0);
        lv.swap(lv0);
    }
};

struct Handle {
    int precision;
    Solver<float> f;
    Solver<double> d;
};

template <class T>
void export_state(const Solver<T>& s, double* pos, double* vel, double* rho, double* p, uint32_t* ids) {
    for (uint32_t i = 0; i < s.n; ++i) {
        for (int a = 0; a < 4; ++a) {
            if (pos) pos[4 * i + a] = (double)s.X[4 * i + a];
            if (vel) vel[4 * i + a] = a < 3 ? (double)s.V[4 * i + a] : 0.0;
        }
        if (rho) rho[i] = (double)s.rho[i];
        if (p) p[i] = (double)s.prs[i];
        if (ids) ids[i] = s.ids[i];
    }
}

}  // namespace

extern "C" {

// --- reference-semantics primitives (FP32, bit-exact statement) -------
uint64_t orc_morton_key(int cx, int cy, int cz) { return morton_key(cx, cy, cz); }

int orc_cell_coord_f32(float p, float cs) { return cell_of(p, cs); }

// Order of a stable sort by the reference key of pos4[i] (cell size cs).
void orc_sort_order_f32(const float* pos4, uint32_t n, float cs, uint32_t* order) {
    std::vector<uint64_t> keys(n);
    for (uint32_t i = 0; i < n; ++i)
        keys[i] = morton_key(cell_of(pos4[4 * i], cs), cell_of(pos4[4 * i + 1], cs), cell_of(pos4[4 * i + 2], cs));
    std::vector<uint32_t> ord;
    stable_order(keys, ord, false);
    std::memcpy(order, ord.data(), sizeof(uint32_t) * n);
}

void orc_morton_keys_f32(const float* pos4, uint32_t n, float cs, uint64_t* keys) {
    for (uint32_t i = 0; i < n; ++i)
        keys[i] = morton_key(cell_of(pos4[4 * i], cs), cell_of(pos4[4 * i + 1], cs), cell_of(pos4[4 * i + 2], cs));
}

// Neighbour lists of key-sorted query points against key-sorted source
// points (both cell size cs, support h), reference enumeration order: cells
// dx, dy, dz in [-R, R] with R the smallest walk range R * cs >= h.
// exclude_self: skip j == i (fluid-fluid).  counts[nq]; list may be NULL.
// Returns the total number of pairs.
uint64_t orc_neighbors_f32(const float* q4, uint32_t nq, const float* src4, uint32_t ns,
                           float h, float cs, int exclude_self, uint32_t* counts,
                           uint32_t* list, uint64_t list_len) {
    std::vector<uint64_t> sk(ns);
    for (uint32_t j = 0; j < ns; ++j)
        sk[j] = morton_key(cell_of(src4[4 * j], cs), cell_of(src4[4 * j + 1], cs), cell_of(src4[4 * j + 2], cs));
    for (uint32_t j = 1; j < ns; ++j)
        if (sk[j] < sk[j - 1]) return UINT64_MAX;  // sources must be key-sorted
    const std::vector<CellRun> runs = build_runs(sk);
    const float h2 = h * h;
    const int R = (int)std::ceil((double)h / (double)cs - 1e-9);  // walk range: R * cs >= h
    uint64_t w = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        const float* xi = &q4[4 * i];
        const int c0 = cell_of(xi[0], cs), c1 = cell_of(xi[1], cs), c2 = cell_of(xi[2], cs);
        uint32_t c = 0;
        for (int dx = -R; dx <= R; ++dx)
            for (int dy = -R; dy <= R; ++dy)
                for (int dz = -R; dz <= R; ++dz) {
                    uint32_t s0, e0;
                    if (!find_run(runs, morton_key(c0 + dx, c1 + dy, c2 + dz), s0, e0)) continue;
                    for (uint32_t j = s0; j < e0; ++j) {
                        if (exclude_self && j == i) continue;
                        const float d2 = dist2(xi[0] - src4[4 * j], xi[1] - src4[4 * j + 1], xi[2] - src4[4 * j + 2]);
                        if (d2 <= h2) {
                            if (list && w < list_len) list[w] = j;
                            ++w;
                            ++c;
                        }
                    }
                }
        counts[i] = c;
    }
    return w;
}

// W, grad W and friends, for the closed-form tests.
double orc_cubic_w(double q, double h) { return cubic_w(q, h); }
double orc_cubic_grad(double q, double h) { return cubic_grad(q, h); }
double orc_spiky_ref_grad(double q, double h) { return spiky_ref_grad(q, h); }
double orc_cohesion(double q, double h) { return cohesion(q, h); }
float orc_cubic_w_f32(float q, float h) { return cubic_w(q, h); }
double orc_tait(double rho, double rho0, double k, double gamma, double max_ratio) {
    return tait(rho, rho0, k, gamma, max_ratio);
}

// Dam-break generator, restated: the lattice of fluid_simulator.cpp:24-54
// plus seeded jitter, and `layers` wall shells on the lattice continuation.
int orc_make_dam_break(float r, const uint32_t dims[3], const float lattice_min[3], float jitter,
                       uint64_t seed, const float box_min[3], const float box_max[3], uint32_t layers,
                       float* fluid, uint32_t* n_fluid, float* bnd, uint32_t* n_bnd) {
    const float s = 2.f * r;
    const uint64_t nf = (uint64_t)dims[0] * dims[1] * dims[2];
    if (fluid) {
        uint64_t i = 0;
        for (uint32_t x = 0; x < dims[0]; ++x)
            for (uint32_t y = 0; y < dims[1]; ++y)
                for (uint32_t z = 0; z < dims[2]; ++z, ++i) {
                    const uint32_t ix[3] = {x, y, z};
                    for (int a = 0; a < 3; ++a) {
                        const float t = (float)ix[a] * s;
                        const float base = lattice_min[a] + t;
                        const float j = jitter * jitter_unit(seed, i, a);
                        fluid[4 * i + a] = base + j;
                    }
                    fluid[4 * i + 3] = 1.f;
                }
    }
    *n_fluid = (uint32_t)nf;
    // per-axis site coordinates: L lower shells, interior, L upper shells
    std::vector<float> sites[3];
    uint32_t nin[3];
    for (int a = 0; a < 3; ++a) {
        const float len = box_max[a] - box_min[a];
        nin[a] = (uint32_t)std::lround((double)len / (double)s);
        for (int l = (int)layers - 1; l >= 0; --l) {
            const float t = (float)l * s;
            const float c = box_min[a] - r;
            sites[a].push_back(c - t);
        }
        for (uint32_t m = 0; m < nin[a]; ++m) {
            const float t = (float)m * s;
            const float c = box_min[a] + r;
            sites[a].push_back(c + t);
        }
        for (uint32_t l = 0; l < layers; ++l) {
            const float t = (float)l * s;
            const float c = box_max[a] + r;
            sites[a].push_back(c + t);
        }
    }
    uint64_t nbw = 0;
    for (uint32_t x = 0; x < sites[0].size(); ++x)
        for (uint32_t y = 0; y < sites[1].size(); ++y)
            for (uint32_t z = 0; z < sites[2].size(); ++z) {
                const bool inx = x >= layers && x < layers + nin[0];
                const bool iny = y >= layers && y < layers + nin[1];
                const bool inz = z >= layers && z < layers + nin[2];
                if (inx && iny && inz) continue;
                if (bnd) {
                    bnd[4 * nbw] = sites[0][x];
                    bnd[4 * nbw + 1] = sites[1][y];
                    bnd[4 * nbw + 2] = sites[2][z];
                    bnd[4 * nbw + 3] = 1.f;
                }
                ++nbw;
            }
    *n_bnd = (uint32_t)nbw;
    return 0;
}

// --- the solver ----------------------------------------------------------
void* orc_create(int precision, const wcsph_params* p, const wcsph_scene* sc, int parallel) {
    Handle* hd = new Handle();
    hd->precision = precision;
    if (precision == 32) {
        hd->f.parallel = parallel != 0;
        hd->f.set_params(*p, sc->box);
        hd->f.init_boundary(sc->boundary_pos, sc->n_boundary);
        hd->f.set_fluid(sc->fluid_pos, sc->fluid_vel, sc->n_fluid);
    } else {
        hd->d.parallel = parallel != 0;
        hd->d.set_params(*p, sc->box);
        hd->d.init_boundary(sc->boundary_pos, sc->n_boundary);
        hd->d.set_fluid(sc->fluid_pos, sc->fluid_vel, sc->n_fluid);
    }
    return hd;
}

void orc_destroy(void* h) { delete (Handle*)h; }

// Box::set_model for the following steps (src/render/box.cpp:25-29).
// Fluid cell size h * mult (the reference's CELL_SIZE_MULT); the boundary
// grid keeps cell size h.  Takes effect at the next sort.
void orc_set_cell_size_mult(void* h, float mult) {
    Handle* hd = (Handle*)h;
    if (hd->precision == 32) hd->f.set_cell_size_mult(mult);
    else hd->d.set_cell_size_mult(mult);
}

void orc_set_box(void* h, const wcsph_box* box) {
    Handle* hd = (Handle*)h;
    if (hd->precision == 32) hd->f.set_box(*box);
    else hd->d.set_box(*box);
}

void orc_step(void* h, double hint, uint32_t n_steps) {
    Handle* hd = (Handle*)h;
    for (uint32_t k = 0; k < n_steps; ++k) {
        if (hd->precision == 32) hd->f.step(hint);
        else hd->d.step(hint);
    }
}

// Sets positions/velocities of the fluid in FP64 (the FP64 oracle can be
// restarted from an exact double state).
void orc_set_state64(void* h, const double* pos, const double* vel) {
    Handle* hd = (Handle*)h;
    if (hd->precision == 32) {
        for (uint32_t i = 0; i < hd->f.n; ++i)
            for (int a = 0; a < 3; ++a) {
                hd->f.X[4 * i + a] = (float)pos[4 * i + a];
                hd->f.V[4 * i + a] = vel ? (float)vel[4 * i + a] : 0.f;
            }
    } else {
        for (uint32_t i = 0; i < hd->d.n; ++i)
            for (int a = 0; a < 3; ++a) {
                hd->d.X[4 * i + a] = pos[4 * i + a];
                hd->d.V[4 * i + a] = vel ? vel[4 * i + a] : 0.0;
            }
    }
}

void orc_get_state(void* h, double* pos, double* vel, double* rho, double* p, uint32_t* ids) {
    Handle* hd = (Handle*)h;
    if (hd->precision == 32) export_state(hd->f, pos, vel, rho, p, ids);
    else export_state(hd->d, pos, vel, rho, p, ids);
}

void orc_get_boundary(void* h, double* pos_psi) {
    Handle* hd = (Handle*)h;
    if (hd->precision == 32) {
        for (size_t i = 0; i < hd->f.B.size(); ++i) pos_psi[i] = hd->f.B[i];
    } else {
        for (size_t i = 0; i < hd->d.B.size(); ++i) pos_psi[i] = hd->d.B[i];
    }
}

void orc_get_perm(void* h, uint32_t* perm) {
    Handle* hd = (Handle*)h;
    const std::vector<uint32_t>& p = hd->precision == 32 ? hd->f.perm : hd->d.perm;
    std::memcpy(perm, p.data(), sizeof(uint32_t) * p.size());
}

// counts (n) and total of the last fluid (which=0) or boundary (which=1) lists
uint64_t orc_get_neighbors(void* h, int which, uint32_t* counts, uint32_t* list, uint64_t list_len) {
    Handle* hd = (Handle*)h;
    auto get = [&](auto& s) -> uint64_t {
        const auto& c = which ? s.bcnt : s.fcnt;
        const auto& l = which ? s.blist : s.flist;
        if (counts) std::memcpy(counts, c.data(), sizeof(uint32_t) * c.size());
        if (list) std::memcpy(list, l.data(), sizeof(uint32_t) * std::min<uint64_t>(list_len, l.size()));
        return l.size();
    };
    return hd->precision == 32 ? get(hd->f) : get(hd->d);
}

// One evaluation of the step's passes without integration (Solver::eval):
// rho_in / p_in (per particle id, NULL = computed) override the density
// pass; rho, p (n) and the pressure acceleration ap, the non-pressure
// acceleration anp and the normals nrm (4 per particle) are returned per
// particle id.  Positions, velocities and time are unchanged (the particle
// order becomes this evaluation's sort).
void orc_eval(void* h, const double* rho_in, const double* p_in, double* rho, double* p, double* ap, double* anp,
              double* nrm) {
    Handle* hd = (Handle*)h;
    auto run = [&](auto& s) {
        s.eval(rho_in, p_in);
        for (uint32_t i = 0; i < s.n; ++i) {
            const uint32_t id = s.ids[i];
            if (rho) rho[id] = (double)s.rho[i];
            if (p) p[id] = (double)s.prs[i];
            for (int a = 0; a < 4; ++a) {
                if (ap) ap[4 * (size_t)id + a] = a < 3 ? (double)s.ap[4 * (size_t)i + a] : 0.0;
                if (anp) anp[4 * (size_t)id + a] = a < 3 ? (double)s.anp[4 * (size_t)i + a] : 0.0;
                if (nrm) nrm[4 * (size_t)id + a] = a < 3 ? (double)s.nrm[4 * (size_t)i + a] : 0.0;
            }
        }
    };
    if (hd->precision == 32) run(hd->f);
    else run(hd->d);
}

void orc_get_time(void* h, double* sim_time, double* last_dt) {
    Handle* hd = (Handle*)h;
    *sim_time = hd->precision == 32 ? hd->f.sim_time : hd->d.sim_time;
    *last_dt = hd->precision == 32 ? hd->f.last_dt : hd->d.last_dt;
}

// Differentiable run: n_steps taped steps from the current state; returns
// L = (1/N) sum_i |x_i(T) - target_i|^2 (targets 4 per particle, in id order).
double orc_grad_run(void* h, double hint, uint32_t n_steps, const double* targets) {
    Handle* hd = (Handle*)h;
    auto run = [&](auto& s) -> double {
        s.tape.clear();
        s.taping = true;
        for (uint32_t k = 0; k < n_steps; ++k) s.step(hint);
        s.taping = false;
        double L = 0.0;
        for (uint32_t i = 0; i < s.n; ++i) {
            const uint32_t id = s.ids[i];
            for (int a = 0; a < 3; ++a) {
                const double d = (double)s.X[4 * i + a] - targets[4 * id + a];
                L += d * d;
            }
        }
        return L / (double)s.n;
    };
    return hd->precision == 32 ? run(hd->f) : run(hd->d);
}

// Gradients of the last orc_grad_run's loss with respect to the initial
// velocities (4 per particle, id order) and the viscosity alpha.
void orc_grad_backward(void* h, const double* targets, double* gv0, double* galpha) {
    Handle* hd = (Handle*)h;
    auto back = [&](auto& s) {
        using T = typename std::remove_reference<decltype(s.X)>::type::value_type;
        std::vector<T> lx(4 * (size_t)s.n, T(0)), lv(4 * (size_t)s.n, T(0));
        for (uint32_t i = 0; i < s.n; ++i) {
            const uint32_t id = s.ids[i];
            for (int a = 0; a < 3; ++a)
                lx[4 * i + a] = (T)(2.0 * ((double)s.X[4 * i + a] - targets[4 * id + a]) / (double)s.n);
        }
        double ga = 0.0;
        for (size_t k = s.tape.size(); k-- > 0;) s.step_backward(s.tape[k], lx, lv, ga);
        // the order before the first taped step is the state's id order at
        // that time; ids started as 0..n-1 when the run began from set_fluid
        for (uint32_t i = 0; i < s.n; ++i)
            for (int a = 0; a < 4; ++a) gv0[4 * i + a] = a < 3 ? (double)lv[4 * i + a] : 0.0;
        *galpha = ga;
    };
    if (hd->precision == 32) back(hd->f);
    else back(hd->d);
}

void orc_set_viscosity(void* h, double alpha) {
    Handle* hd = (Handle*)h;
    hd->f.visc = (float)alpha;
    hd->d.visc = alpha;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { omp_set_num_threads(t); }

}  // extern "C"
