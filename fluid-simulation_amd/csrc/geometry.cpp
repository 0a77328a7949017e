// Scene geometry on the host: the clustered fluid block, procedural
// triangle meshes and the voxel sampler that turns a triangle mesh into
// boundary particles.  Runs once per scene; no device is needed.
//
// The sampler follows the published behaviour of Open3D's
// VoxelGrid::CreateFromTriangleMesh, which the reference calls at voxel size
// 0.75 * 2r (src/simulation/SPH/boundary.cpp:8-35): a voxel grid anchored at
// the mesh minimum minus half a voxel, a voxel kept when its cube overlaps a
// triangle by the separating-axis test of Akenine-Moeller ("Fast 3D
// Triangle-Box Overlap Testing", 2001), and the voxel centre as the
// particle.  Arithmetic is double, as in Open3D (Eigen::Vector3d).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/wcsph.h"

namespace {

constexpr double kPiD = 3.14159265358979323846;

// Counter-based uniform in [0, 1) (SplitMix64 finaliser of seed + counter).
double unit_hash(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Jitter in [-1, 1) for site idx, axis a (float, like the dam-break lattice).
float jitter_unit(uint64_t seed, uint64_t idx, int axis) {
    const float u = (float)(unit_hash(seed ^ 0x5A5A5A5A5A5A5A5Aull, idx * 3u + (uint64_t)axis) * 16777216.0) *
                    (1.0f / 16777216.0f);
    return 2.0f * u - 1.0f;
}

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Separating-axis overlap of the cube (centre c, half edge hs) and the
// triangle (a, b, d); touching counts as overlap.
bool tri_box_overlap(V3 c, double hs, V3 a, V3 b, V3 d) {
    const V3 v[3] = {sub(a, c), sub(b, c), sub(d, c)};
    const V3 e[3] = {sub(v[1], v[0]), sub(v[2], v[1]), sub(v[0], v[2])};
    // box face normals: the triangle's AABB against the box
    for (int k = 0; k < 3; ++k) {
        const double p0 = (&v[0].x)[k], p1 = (&v[1].x)[k], p2 = (&v[2].x)[k];
        if (std::min({p0, p1, p2}) > hs || std::max({p0, p1, p2}) < -hs) return false;
    }
    // the nine edge x axis directions
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) {
            V3 unit{0.0, 0.0, 0.0};
            (&unit.x)[k] = 1.0;
            const V3 ax = cross(unit, e[i]);
            const double p0 = dot(ax, v[0]), p1 = dot(ax, v[1]), p2 = dot(ax, v[2]);
            const double rad = hs * (std::fabs(ax.x) + std::fabs(ax.y) + std::fabs(ax.z));
            if (std::min({p0, p1, p2}) > rad || std::max({p0, p1, p2}) < -rad) return false;
        }
    }
    // the triangle's plane
    const V3 nrm = cross(e[0], e[1]);
    const double dist = dot(nrm, v[0]);
    const double rad = hs * (std::fabs(nrm.x) + std::fabs(nrm.y) + std::fabs(nrm.z));
    return std::fabs(dist) <= rad;
}

void apply_model(const float* m, double x, double y, double z, float out[3]) {
    if (!m) {
        out[0] = (float)x;
        out[1] = (float)y;
        out[2] = (float)z;
        return;
    }
    for (int a = 0; a < 3; ++a)
        out[a] = (float)((double)m[a] * x + (double)m[4 + a] * y + (double)m[8 + a] * z + (double)m[12 + a]);
}

void push_tri(std::vector<float>& out, const float* m, const double p[3][3]) {
    for (int k = 0; k < 3; ++k) {
        float w[3];
        apply_model(m, p[k][0], p[k][1], p[k][2], w);
        out.insert(out.end(), w, w + 3);
    }
}

int build_mesh(int kind, const float* prm, uint32_t res, const float* model, std::vector<float>& out) {
    if (!prm) return WCSPH_ERR_INVALID_ARGUMENT;
    if (kind == WCSPH_MESH_CUBE) {
        const double hx = prm[0], hy = prm[1], hz = prm[2];
        if (!(hx > 0 && hy > 0 && hz > 0)) return WCSPH_ERR_INVALID_ARGUMENT;
        // each face as two triangles (corner signs per axis)
        for (int axis = 0; axis < 3; ++axis) {
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const int u = (axis + 1) % 3, w = (axis + 2) % 3;
                const double hh[3] = {hx, hy, hz};
                double q[4][3];
                const int su[4] = {-1, 1, 1, -1}, sw[4] = {-1, -1, 1, 1};
                for (int k = 0; k < 4; ++k) {
                    q[k][axis] = sgn * hh[axis];
                    q[k][u] = su[k] * hh[u];
                    q[k][w] = sw[k] * hh[w];
                }
                const double t0[3][3] = {{q[0][0], q[0][1], q[0][2]}, {q[1][0], q[1][1], q[1][2]}, {q[2][0], q[2][1], q[2][2]}};
                const double t1[3][3] = {{q[0][0], q[0][1], q[0][2]}, {q[2][0], q[2][1], q[2][2]}, {q[3][0], q[3][1], q[3][2]}};
                push_tri(out, model, t0);
                push_tri(out, model, t1);
            }
        }
        return WCSPH_OK;
    }
    if (res < 3) return WCSPH_ERR_INVALID_ARGUMENT;
    if (kind == WCSPH_MESH_SPHERE) {
        const double R = prm[0];
        if (!(R > 0)) return WCSPH_ERR_INVALID_ARGUMENT;
        const uint32_t nl = res, nphi = 2 * res;
        auto vert = [&](uint32_t i, uint32_t j, double* p) {
            const double th = kPiD * (double)i / nl, ph = 2.0 * kPiD * (double)(j % nphi) / nphi;
            p[0] = R * std::sin(th) * std::cos(ph);
            p[1] = R * std::cos(th);
            p[2] = R * std::sin(th) * std::sin(ph);
        };
        for (uint32_t i = 0; i < nl; ++i)
            for (uint32_t j = 0; j < nphi; ++j) {
                double a[3], b[3], c[3], d[3];
                vert(i, j, a);
                vert(i + 1, j, b);
                vert(i + 1, j + 1, c);
                vert(i, j + 1, d);
                if (i + 1 < nl) {  // the lower triangle degenerates at the south pole
                    const double t[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
                    push_tri(out, model, t);
                }
                if (i > 0) {  // the upper triangle degenerates at the north pole
                    const double t[3][3] = {{a[0], a[1], a[2]}, {c[0], c[1], c[2]}, {d[0], d[1], d[2]}};
                    push_tri(out, model, t);
                }
            }
        return WCSPH_OK;
    }
    if (kind == WCSPH_MESH_TORUS) {
        const double R = prm[0], rt = prm[1];
        if (!(R > 0 && rt > 0 && rt < R)) return WCSPH_ERR_INVALID_ARGUMENT;
        const uint32_t nu = 2 * res, nv = res;
        auto vert = [&](uint32_t i, uint32_t j, double* p) {
            const double u = 2.0 * kPiD * (double)(i % nu) / nu, v = 2.0 * kPiD * (double)(j % nv) / nv;
            const double rr = R + rt * std::cos(v);
            p[0] = rr * std::cos(u);
            p[1] = rt * std::sin(v);
            p[2] = rr * std::sin(u);
        };
        for (uint32_t i = 0; i < nu; ++i)
            for (uint32_t j = 0; j < nv; ++j) {
                double a[3], b[3], c[3], d[3];
                vert(i, j, a);
                vert(i + 1, j, b);
                vert(i + 1, j + 1, c);
                vert(i, j + 1, d);
                const double t0[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
                const double t1[3][3] = {{a[0], a[1], a[2]}, {c[0], c[1], c[2]}, {d[0], d[1], d[2]}};
                push_tri(out, model, t0);
                push_tri(out, model, t1);
            }
        return WCSPH_OK;
    }
    return WCSPH_ERR_INVALID_ARGUMENT;
}

}  // namespace

extern "C" {

int wcsph_mesh_procedural(int kind, const float params[3], uint32_t resolution, const float model[16],
                          float* vertices, uint32_t capacity, uint32_t* n_triangles) {
    if (!n_triangles) return WCSPH_ERR_INVALID_ARGUMENT;
    std::vector<float> tri;
    const int rc = build_mesh(kind, params, resolution, model, tri);
    if (rc) return rc;
    const size_t nt = tri.size() / 9;
    if (nt > 0xffffffffull) return WCSPH_ERR_INVALID_ARGUMENT;
    *n_triangles = (uint32_t)nt;
    if (!vertices) return WCSPH_OK;
    if (capacity < nt) return WCSPH_ERR_INVALID_ARGUMENT;
    std::memcpy(vertices, tri.data(), tri.size() * sizeof(float));
    return WCSPH_OK;
}

int wcsph_sample_mesh(const float* tv, uint32_t n_triangles, float voxel_size, float* out_pos, uint32_t capacity,
                      uint32_t* n_out) {
    if (!n_out || (n_triangles && !tv) || !(voxel_size > 0.f) || !std::isfinite(voxel_size))
        return WCSPH_ERR_INVALID_ARGUMENT;
    *n_out = 0;
    if (n_triangles == 0) return WCSPH_OK;
    const double vs = voxel_size, half = 0.5 * vs;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t k = 0; k < 3 * (size_t)n_triangles; ++k)
        for (int a = 0; a < 3; ++a) {
            const double v = tv[3 * k + a];
            if (!std::isfinite(v)) return WCSPH_ERR_INVALID_ARGUMENT;
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    // grid: origin = mesh min - half a voxel, round((extent + voxel) / voxel) voxels per axis
    double org[3];
    int64_t dim[3];
    for (int a = 0; a < 3; ++a) {
        org[a] = lo[a] - half;
        dim[a] = (int64_t)std::llround((hi[a] + half - org[a]) / vs);
        if (dim[a] < 1) dim[a] = 1;
        if (dim[a] > (1 << 21)) return WCSPH_ERR_INVALID_ARGUMENT;
    }
    // bounds the voxels a triangle's box can visit
    if ((uint64_t)dim[0] * (uint64_t)dim[1] * (uint64_t)dim[2] > (1ull << 31)) return WCSPH_ERR_INVALID_ARGUMENT;
    std::vector<uint64_t> hits;
    for (uint32_t t = 0; t < n_triangles; ++t) {
        const float* p = tv + 9 * (size_t)t;
        const V3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, d{p[6], p[7], p[8]};
        int64_t i0[3], i1[3];
        for (int ax = 0; ax < 3; ++ax) {
            const double mn = std::min({(&a.x)[ax], (&b.x)[ax], (&d.x)[ax]});
            const double mx = std::max({(&a.x)[ax], (&b.x)[ax], (&d.x)[ax]});
            // one extra voxel either side: a cube that only touches the AABB still counts
            i0[ax] = std::max<int64_t>(0, (int64_t)std::floor((mn - org[ax]) / vs) - 1);
            i1[ax] = std::min<int64_t>(dim[ax] - 1, (int64_t)std::floor((mx - org[ax]) / vs) + 1);
        }
        for (int64_t x = i0[0]; x <= i1[0]; ++x)
            for (int64_t y = i0[1]; y <= i1[1]; ++y)
                for (int64_t z = i0[2]; z <= i1[2]; ++z) {
                    const V3 c{org[0] + half + (double)x * vs, org[1] + half + (double)y * vs,
                               org[2] + half + (double)z * vs};
                    if (tri_box_overlap(c, half, a, b, d))
                        hits.push_back(((uint64_t)x * (uint64_t)dim[1] + (uint64_t)y) * (uint64_t)dim[2] + (uint64_t)z);
                }
    }
    std::sort(hits.begin(), hits.end());
    hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
    if (hits.size() > 0xffffffffull) return WCSPH_ERR_INVALID_ARGUMENT;
    *n_out = (uint32_t)hits.size();
    if (!out_pos) return WCSPH_OK;
    if (capacity < hits.size()) return WCSPH_ERR_INVALID_ARGUMENT;
    for (size_t k = 0; k < hits.size(); ++k) {
        const uint64_t lin = hits[k];
        const uint64_t z = lin % (uint64_t)dim[2], y = (lin / (uint64_t)dim[2]) % (uint64_t)dim[1],
                       x = lin / ((uint64_t)dim[2] * (uint64_t)dim[1]);
        out_pos[4 * k + 0] = (float)(org[0] + half + (double)x * vs);
        out_pos[4 * k + 1] = (float)(org[1] + half + (double)y * vs);
        out_pos[4 * k + 2] = (float)(org[2] + half + (double)z * vs);
        out_pos[4 * k + 3] = 1.f;
    }
    return WCSPH_OK;
}

int wcsph_make_clustered_block(float particle_radius, const float block_min[3], const float block_max[3],
                               float cluster_fraction, float cluster_radius, float cluster_spacing, float jitter,
                               uint64_t seed, float* fluid_pos, uint32_t* n_fluid) {
    if (!block_min || !block_max || !n_fluid || !(particle_radius > 0.f) || !(cluster_fraction >= 0.f) ||
        !(cluster_fraction < 1.f) || !(jitter >= 0.f))
        return WCSPH_ERR_INVALID_ARGUMENT;
    const float r = particle_radius, s = 2.f * r;
    const double sd = s;
    uint32_t dims[3];
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)block_max[a] - (double)block_min[a];
        if (!(ext >= sd)) return WCSPH_ERR_INVALID_ARGUMENT;
        // sites at min + r + s*i <= max - r
        const double n = std::floor((ext - sd) / sd + 1e-4) + 1.0;  // tolerance for float-rounded bounds
        if (n > 1e6) return WCSPH_ERR_INVALID_ARGUMENT;
        dims[a] = (uint32_t)n;
    }
    const uint64_t n_main = (uint64_t)dims[0] * dims[1] * dims[2];
    // cluster lattice offsets inside one ball (the same for every ball)
    std::vector<double> ball;
    const double Rc = cluster_radius;
    const double cut = Rc + 0.9 * sd;
    uint64_t n_cl = 0;
    if (cluster_fraction > 0.f) {
        if (!(cluster_radius > 0.f) || !(cluster_spacing > 0.f)) return WCSPH_ERR_INVALID_ARGUMENT;
        const double cs = cluster_spacing;
        // a ball spans at most 2 * 64 + 1 sites across
        if (!(Rc / cs <= 64.0)) return WCSPH_ERR_INVALID_ARGUMENT;
        const int m = (int)std::floor(Rc / cs);
        for (int i = -m; i <= m; ++i)
            for (int j = -m; j <= m; ++j)
                for (int k = -m; k <= m; ++k) {
                    const double x = i * cs, y = j * cs, z = k * cs;
                    if (x * x + y * y + z * z <= Rc * Rc) ball.insert(ball.end(), {x, y, z});
                }
        const double kc = (double)(ball.size() / 3);
        const double km = 4.0 / 3.0 * kPiD * cut * cut * cut / (sd * sd * sd);  // main sites a ball replaces
        const double f = cluster_fraction;
        const double need = f * (double)n_main / std::max(1.0, (1.0 - f) * kc + f * km);
        n_cl = (uint64_t)std::ceil(need);
    }
    // ball centres by seeded rejection sampling
    std::vector<double> ctr;
    uint64_t counter = 0;
    for (uint64_t b = 0; b < n_cl; ++b) {
        bool placed = false;
        for (int tries = 0; tries < 10000 && !placed; ++tries) {
            double c[3];
            bool fits = true;
            for (int a = 0; a < 3; ++a) {
                const double lo = (double)block_min[a] + Rc + sd, hi = (double)block_max[a] - Rc - sd;
                if (!(hi >= lo)) return WCSPH_ERR_INVALID_ARGUMENT;
                c[a] = lo + (hi - lo) * unit_hash(seed, counter++);
            }
            for (size_t q = 0; q < ctr.size() && fits; q += 3) {
                const double dx = c[0] - ctr[q], dy = c[1] - ctr[q + 1], dz = c[2] - ctr[q + 2];
                fits = dx * dx + dy * dy + dz * dz >= 4.0 * (Rc + sd) * (Rc + sd);
            }
            if (fits) {
                ctr.insert(ctr.end(), c, c + 3);
                placed = true;
            }
        }
        if (!placed) return WCSPH_ERR_INVALID_ARGUMENT;  // the block cannot hold that many balls
    }
    // main lattice minus the balls' surroundings, then the balls' lattices
    uint64_t w = 0;
    auto emit = [&](double x, double y, double z) {
        if (fluid_pos) {
            const double p[3] = {x, y, z};
            for (int a = 0; a < 3; ++a) fluid_pos[4 * w + a] = (float)p[a] + jitter * jitter_unit(seed, w, a);
            fluid_pos[4 * w + 3] = 1.f;
        }
        ++w;
    };
    std::vector<size_t> near;  // balls that reach the current x slab
    for (uint32_t ix = 0; ix < dims[0]; ++ix) {
        const float sx = block_min[0] + r + (float)ix * s;
        near.clear();
        for (size_t q = 0; q < ctr.size(); q += 3)
            if (std::fabs(sx - ctr[q]) < cut) near.push_back(q);
        for (uint32_t iy = 0; iy < dims[1]; ++iy)
            for (uint32_t iz = 0; iz < dims[2]; ++iz) {
                const float sy = block_min[1] + r + (float)iy * s;
                const float sz = block_min[2] + r + (float)iz * s;
                bool inside = false;
                for (size_t k = 0; k < near.size() && !inside; ++k) {
                    const size_t q = near[k];
                    const double dx = sx - ctr[q], dy = sy - ctr[q + 1], dz = sz - ctr[q + 2];
                    inside = dx * dx + dy * dy + dz * dz < cut * cut;
                }
                if (!inside) emit(sx, sy, sz);
            }
    }
    for (size_t q = 0; q < ctr.size(); q += 3)
        for (size_t k = 0; k < ball.size(); k += 3) emit(ctr[q] + ball[k], ctr[q + 1] + ball[k + 1], ctr[q + 2] + ball[k + 2]);
    if (w > 0xffffffffull) return WCSPH_ERR_INVALID_ARGUMENT;
    *n_fluid = (uint32_t)w;
    return WCSPH_OK;
}

}  // extern "C"
