// Device code of the WCSPH step for gfx950 (wave64).
//
// Data layout in HBM (one scene, N fluid, B boundary particles):
//   pos_a/vel_a/id_a  state between steps (float4, float4, u32) — the
//                     order of the last sort;
//   pos_b/vel_b/id_b  the same particles in this step's Morton order;
//   xp                (x, y, z, p/rho^2) per sorted particle, written by the
//                     density pass and gathered by the force pass;
//   vel_b.w           rho of the sorted particle (density pass);
//   bpos              boundary particles, Morton-sorted once, w = Psi;
//   cell grid         dense box of cells of size h over the scene; per cell
//                     count / start / end (fluid) and bstart / bend (boundary);
//   rank_to_lin       the grid's cells listed in reference Morton-key order;
//   nlist / blist     neighbour lists, slot-major: entry k of particle i at
//                     [k * N + i], so a wavefront's 64 reads of slot k are one
//                     contiguous 256-byte segment.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wcsph_dev {

constexpr float kPi = 3.1415926535897932384626433f;

struct Grid {
    int c0x, c0y, c0z;  // cell coordinate of grid index (0, 0, 0)
    int nx, ny, nz;     // grid dims; lin = ((cx - c0x) * ny + (cy - c0y)) * nz + (cz - c0z)
    float cs;           // cell size (= h)
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
    float bmin[3], bmax[3];
    float model[16], model_inv[16];
    int grad_kind, visc_kind, box_identity;
};

// Per-handle device scalars.
struct DevStatus {
    uint32_t flags;
    uint32_t max_fluid;
    uint32_t max_boundary;
    uint32_t pad;
    unsigned long long vmax2_bits;  // max |v|^2 of the last integration (FP64 bits)
    double sim_time;
    double last_dt;
    unsigned long long steps;
    float dt;      // dt of the current step
    float dt_np;   // non-pressure dt of the current step
};

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

__device__ __forceinline__ int cell_coord(float p, float cs) {
    return (int)floorf(p / cs);  // IEEE-rounded division (hipcc default)
}

__device__ __forceinline__ float cubic_w(float q, const Phys& P) {
    if (q <= 0.5f) return P.cubic_factor * (6.f * q * q * q - 6.f * q * q + 1.f);
    if (q <= 1.f) {
        const float a = 1.f - q;
        return P.cubic_factor * 2.f * (a * a * a);
    }
    return 0.f;
}

// grad W = coef * r for the cubic spline (kernel.cuh:22-32).
__device__ __forceinline__ float cubic_grad(float q, const Phys& P) {
    if (q < 1e-9f || q > 1.f) return 0.f;
    const float rl = q * P.h;
    if (q <= 0.5f) return P.cubic_gfactor * q * (3.f * q - 2.f) / (P.h * rl);
    const float a = 1.f - q;
    return -P.cubic_gfactor * a * a / (P.h * rl);
}

// Reference spiky gradient with its h^-9 factor (kernel.cuh:43-51).
__device__ __forceinline__ float spiky_ref_grad(float q, const Phys& P) {
    if (q < 1e-9f || q > 1.f) return 0.f;
    const float rl = q * P.h;
    const float a = 1.f - q;
    return P.spiky_gfactor * a * a / (P.h * rl);
}

__device__ __forceinline__ float cohesion_w(float q, const Phys& P) {
    const float a = 1.f - q;
    const float iq3 = a * a * a;
    const float q3 = q * q * q;
    if (q <= 0.5f) return P.cohesion_factor * (2.f * q3 * iq3 - 1.f / 64.f);
    if (q <= 1.f) return P.cohesion_factor * q3 * iq3;
    return 0.f;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return v;
}

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
__global__ void k_time_step(DevStatus* st, DtRule rule) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

__global__ void k_clear_flags(DevStatus* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->flags = 0;
        st->max_fluid = 0;
        st->max_boundary = 0;
    }
}

// ------------------------------------------------------------------
// 1. cell of every particle + per-cell count (slot = arrival order)
// ------------------------------------------------------------------
__global__ void k_cell_count(const float4* __restrict__ pos, uint32_t n, Grid g,
                             uint32_t* __restrict__ cnt, uint32_t* __restrict__ cell,
                             uint32_t* __restrict__ slot, DevStatus* st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    int cx = cell_coord(p.x, g.cs) - g.c0x;
    int cy = cell_coord(p.y, g.cs) - g.c0y;
    int cz = cell_coord(p.z, g.cs) - g.c0z;
    // Interior cells only: the 27-cell loops then never leave the grid.
    const bool ok = cx >= 1 && cx <= g.nx - 2 && cy >= 1 && cy <= g.ny - 2 && cz >= 1 && cz <= g.nz - 2;
    if (!ok) {
        atomicOr(&st->flags, 4u /* WCSPH_FLAG_DOMAIN */);
        cx = min(max(cx, 1), g.nx - 2);
        cy = min(max(cy, 1), g.ny - 2);
        cz = min(max(cz, 1), g.nz - 2);
    }
    const uint32_t lin = ((uint32_t)cx * g.ny + cy) * g.nz + cz;
    cell[i] = lin;
    slot[i] = atomicAdd(&cnt[lin], 1u);
}

// ------------------------------------------------------------------
// 2. exclusive scan of the counts in Morton-rank order -> start/end
// ------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ rank_to_lin,
                                                     uint32_t ncell, uint32_t* __restrict__ partial) {
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k * 256 + threadIdx.x;
        if (r < ncell) s += cnt[rank_to_lin[r]];
    }
    uint32_t total;
    block256_excl_scan(s, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block of 256 threads: exclusive scan of nparts partials in place.
__global__ __launch_bounds__(256) void k_scan_partials(uint32_t* partial, uint32_t nparts) {
    const uint32_t per = (nparts + 255) / 256;
    const uint32_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nparts) s += partial[b + k];
    uint32_t total;
    uint32_t off = block256_excl_scan(s, &total);
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nparts) {
            const uint32_t v = partial[b + k];
            partial[b + k] = off;
            off += v;
        }
}

// Each thread owns kScanItems consecutive ranks of the tile.
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* __restrict__ cnt,
                                                    const uint32_t* __restrict__ rank_to_lin,
                                                    uint32_t ncell, const uint32_t* __restrict__ partial,
                                                    uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t lin[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k;
        lin[k] = r < ncell ? rank_to_lin[r] : 0u;
        c[k] = r < ncell ? cnt[lin[k]] : 0u;
        s += c[k];
    }
    uint32_t total;
    uint32_t off = block256_excl_scan(s, &total) + partial[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t r = base + k;
        if (r < ncell) {
            start[lin[k]] = off;
            off += c[k];
            end[lin[k]] = off;
            cnt[lin[k]] = 0u;  // ready for the next step's counting
        }
    }
}

// ------------------------------------------------------------------
// 3. scatter particle ids into their cell's range (arrival order)
// ------------------------------------------------------------------
__global__ void k_scatter(const uint32_t* __restrict__ cell, const uint32_t* __restrict__ slot,
                          const uint32_t* __restrict__ start, uint32_t n, uint32_t* __restrict__ tmp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmp[start[cell[i]] + slot[i]] = i;
}

// ------------------------------------------------------------------
// 4. stable placement: rank of i among the cell's members by previous
//    index, then move x, v, id to the sorted buffers
// ------------------------------------------------------------------
__global__ void k_permute(const uint32_t* __restrict__ cell, const uint32_t* __restrict__ start,
                          const uint32_t* __restrict__ end, const uint32_t* __restrict__ tmp, uint32_t n,
                          const float4* __restrict__ pos_a, const float4* __restrict__ vel_a,
                          const uint32_t* __restrict__ id_a, float4* __restrict__ pos_b,
                          float4* __restrict__ vel_b, uint32_t* __restrict__ id_b,
                          uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t lin = cell[i];
    const uint32_t s = start[lin], e = end[lin];
    uint32_t r = 0;
    for (uint32_t k = s; k < e; ++k) r += tmp[k] < i ? 1u : 0u;
    const uint32_t f = s + r;
    pos_b[f] = pos_a[i];
    vel_b[f] = vel_a[i];
    id_b[f] = id_a[i];
    perm[f] = i;
}

// ------------------------------------------------------------------
// 5. neighbour lists (fluid + boundary), reference enumeration order
// ------------------------------------------------------------------
__global__ void k_neighbors(const float4* __restrict__ pos, uint32_t n, Grid g,
                            const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                            const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart,
                            const uint32_t* __restrict__ bend, float h2, uint32_t capf, uint32_t capb,
                            uint32_t* __restrict__ nlist, uint32_t* __restrict__ ncount,
                            uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount, DevStatus* st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cf = 0, cb = 0;
    if (i < n) {
        const float4 xi = pos[i];
        int cx = cell_coord(xi.x, g.cs) - g.c0x;
        int cy = cell_coord(xi.y, g.cs) - g.c0y;
        int cz = cell_coord(xi.z, g.cs) - g.c0z;
        cx = min(max(cx, 1), g.nx - 2);
        cy = min(max(cy, 1), g.ny - 2);
        cz = min(max(cz, 1), g.nz - 2);
        const uint32_t lin = ((uint32_t)cx * g.ny + cy) * g.nz + cz;
        const uint32_t sx = (uint32_t)g.ny * g.nz, sy = (uint32_t)g.nz;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const uint32_t c = lin + dx * (int)sx + dy * (int)sy + dz;
                    const uint32_t s0 = start[c], e0 = end[c];
                    for (uint32_t j = s0; j < e0; ++j) {
                        const float4 xj = pos[j];
                        if (j != i && dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z) <= h2) {
                            if (cf < capf) nlist[(size_t)cf * n + i] = j;
                            ++cf;
                        }
                    }
                    const uint32_t bs = bstart[c], be = bend[c];
                    for (uint32_t j = bs; j < be; ++j) {
                        const float4 xj = bpos[j];
                        if (dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z) <= h2) {
                            if (cb < capb) blist[(size_t)cb * n + i] = j;
                            ++cb;
                        }
                    }
                }
        ncount[i] = cf;
        bcount[i] = cb;
    }
    const uint32_t mf = wave_max_u32(cf), mb = wave_max_u32(cb);
    if ((threadIdx.x & 63) == 0) {
        if (mf) atomicMax(&st->max_fluid, mf);
        if (mb) atomicMax(&st->max_boundary, mb);
        uint32_t fl = 0;
        if (mf > capf) fl |= 1u;  // WCSPH_FLAG_FLUID_OVERFLOW
        if (mb > capb) fl |= 2u;  // WCSPH_FLAG_BOUNDARY_OVERFLOW
        if (fl) atomicOr(&st->flags, fl);
    }
}

// ------------------------------------------------------------------
// 6. density + Tait EOS (compute_rho_p_with_boundary, compute_rho_p.cu:242-318)
// ------------------------------------------------------------------
__global__ void k_density(const float4* __restrict__ pos, float4* __restrict__ vel, uint32_t n,
                          const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                          const float4* __restrict__ bpos, const uint32_t* __restrict__ blist,
                          const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb, Phys P,
                          float4* __restrict__ xp, float* __restrict__ rho_out, float* __restrict__ p_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = pos[i];
    // counts above the capacity are reported by k_neighbors; only the
    // stored slots are read
    const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
    float d = P.w0;
    for (uint32_t k = 0; k < cf; ++k) {
        const uint32_t j = nlist[(size_t)k * n + i];
        const float4 xj = pos[j];
        const float q = sqrtf(dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z)) * P.inv_h;
        d += cubic_w(q, P);
    }
    d = d * P.mass;
    for (uint32_t k = 0; k < cb; ++k) {
        const uint32_t j = blist[(size_t)k * n + i];
        const float4 xj = bpos[j];
        const float q = sqrtf(dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z)) * P.inv_h;
        d += cubic_w(q, P) * xj.w;
    }
    float ratio = fmaxf(d / P.rho0, 1.f);
    ratio = fminf(ratio, P.max_ratio);
    const float p = P.stiffness * (powf(ratio, P.exponent) - 1.f);
    rho_out[i] = d;
    p_out[i] = p;
    xp[i] = make_float4(xi.x, xi.y, xi.z, p / (d * d));
    reinterpret_cast<float*>(vel)[4 * (size_t)i + 3] = d;
}

// ------------------------------------------------------------------
// 6b. normals for surface tension (compute_pressure_accel_n_normal.cu:98-124)
// ------------------------------------------------------------------
__global__ void k_normals(const float4* __restrict__ xp, const float4* __restrict__ vel, uint32_t n,
                          const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                          uint32_t capf, Phys P, float4* __restrict__ nrm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = xp[i];
    const uint32_t cf = min(ncount[i], capf);
    float nx = 0.f, ny = 0.f, nz = 0.f;
    for (uint32_t k = 0; k < cf; ++k) {
        const uint32_t j = nlist[(size_t)k * n + i];
        const float4 xj = xp[j];
        const float dj = vel[j].w;
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float q = sqrtf(dist2(rx, ry, rz)) * P.inv_h;
        const float c = cubic_grad(q, P) / dj;
        nx += rx * c;
        ny += ry * c;
        nz += rz * c;
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
    o[0] = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
    o[1] = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
    o[2] = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
}

template <bool kNormals>
__global__ void k_force_integrate(const float4* __restrict__ xp, const float4* __restrict__ vel,
                                  const uint32_t* __restrict__ id_b, uint32_t n,
                                  const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                                  const float4* __restrict__ bpos, const uint32_t* __restrict__ blist,
                                  const uint32_t* __restrict__ bcount, uint32_t capf, uint32_t capb,
                                  const float* __restrict__ p_in, const float4* __restrict__ nrm, Phys P,
                                  DevStatus* st,
                                  float4* __restrict__ pos_out, float4* __restrict__ vel_out,
                                  uint32_t* __restrict__ id_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double v2 = 0.0;
    bool bad = false;
    if (i < n) {
        const float dt = st->dt, dt_np = st->dt_np;
        const float4 xi = xp[i];
        const float4 vi = vel[i];
        const float di = vi.w;
        const float dpi = xi.w;
        float4 ni = make_float4(0.f, 0.f, 0.f, 0.f);
        if (kNormals) ni = nrm[i];
        const uint32_t cf = min(ncount[i], capf), cb = min(bcount[i], capb);
        float pax = 0.f, pay = 0.f, paz = 0.f;
        float vlx = 0.f, vly = 0.f, vlz = 0.f;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        for (uint32_t k = 0; k < cf; ++k) {
            const uint32_t j = nlist[(size_t)k * n + i];
            const float4 xj = xp[j];
            const float4 vj = vel[j];
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float r2 = dist2(rx, ry, rz);
            const float q = sqrtf(r2) * P.inv_h;
            const float gc = cubic_grad(q, P);
            const float gp = P.grad_kind ? spiky_ref_grad(q, P) : gc;
            const float cp = (dpi + xj.w) * gp;
            pax -= cp * rx;
            pay -= cp * ry;
            paz -= cp * rz;
            const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
            const float vr = vx * rx + vy * ry + vz * rz;
            if (P.visc_kind) {
                const float c = vr * gc / (vj.w * (r2 + P.visc_eps));
                vlx += c * rx;
                vly += c * ry;
                vlz += c * rz;
            } else if (vr < 0.f) {
                const float nu = P.art_visc_coef / (di + vj.w);
                const float c = -(-nu * vr / (r2 + P.visc_eps)) * gc;
                vlx += c * rx;
                vly += c * ry;
                vlz += c * rz;
            }
            if (kNormals && r2 > 1e-6f) {
                const float4 nj = nrm[j];
                const float rl = sqrtf(r2);
                const float cq = P.mass * cohesion_w(q, P);
                const float inv = 1.f / (di + vj.w);
                fx += (cq * (rx / rl) + ni.x - nj.x) * inv;
                fy += (cq * (ry / rl) + ni.y - nj.y) * inv;
                fz += (cq * (rz / rl) + ni.z - nj.z) * inv;
            }
        }
        float apx = P.mass * pax, apy = P.mass * pay, apz = P.mass * paz;
        const float bterm = dpi + p_in[i] * P.inv_rho0_sq;
        for (uint32_t k = 0; k < cb; ++k) {
            const uint32_t j = blist[(size_t)k * n + i];
            const float4 xb = bpos[j];
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            const float q = sqrtf(dist2(rx, ry, rz)) * P.inv_h;
            const float gp = P.grad_kind ? spiky_ref_grad(q, P) : cubic_grad(q, P);
            const float c = xb.w * bterm * gp;
            apx -= c * rx;
            apy -= c * ry;
            apz -= c * rz;
        }
        const float vfac = P.visc_kind ? P.visc * 10.f * P.mass : P.mass;
        const float sfac = P.surface_tension * 2.f * P.rho0;
        const float anx = P.gx + vfac * vlx - sfac * fx;
        const float any = P.gy + vfac * vly - sfac * fy;
        const float anz = P.gz + vfac * vlz - sfac * fz;
        float v[3] = {vi.x + (dt_np * anx + dt * apx), vi.y + (dt_np * any + dt * apy),
                      vi.z + (dt_np * anz + dt * apz)};
        float x[3] = {xi.x + dt * v[0], xi.y + dt * v[1], xi.z + dt * v[2]};
        // box collisions in the box frame (update_positions.cu:14-52)
        float bx[3], bv[3];
        if (P.box_identity) {
            for (int a = 0; a < 3; ++a) { bx[a] = x[a]; bv[a] = v[a]; }
        } else {
            mat4_apply(P.model_inv, x[0], x[1], x[2], 1.f, bx);
            mat4_apply(P.model_inv, v[0], v[1], v[2], 0.f, bv);
        }
        bool changed = false;
        const float buf = P.r * P.cbuf;
        for (int a = 0; a < 3; ++a) {
            if (bx[a] - buf < P.bmin[a]) {
                bx[a] = P.bmin[a] + P.r;
                bv[a] *= -P.elasticity;
                changed = true;
            } else if (bx[a] + buf > P.bmax[a]) {
                bx[a] = P.bmax[a] - P.r;
                bv[a] *= -P.elasticity;
                changed = true;
            }
        }
        if (changed) {
            if (P.box_identity) {
                for (int a = 0; a < 3; ++a) { x[a] = bx[a]; v[a] = bv[a]; }
            } else {
                mat4_apply(P.model, bx[0], bx[1], bx[2], 1.f, x);
                mat4_apply(P.model, bv[0], bv[1], bv[2], 0.f, v);
            }
        }
        pos_out[i] = make_float4(x[0], x[1], x[2], 1.f);
        vel_out[i] = make_float4(v[0], v[1], v[2], 0.f);
        id_out[i] = id_b[i];
        const double dvx = v[0], dvy = v[1], dvz = v[2];
        v2 = dvx * dvx + dvy * dvy + dvz * dvz;
        bad = !(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]) && isfinite(v[0]) && isfinite(v[1]) &&
                isfinite(v[2]));
        if (bad) v2 = 0.0;
    }
    // block max of |v|^2 in FP64, one atomic per wavefront
    const double m = wave_max_f64(v2);
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomicMax(&st->vmax2_bits, (unsigned long long)__double_as_longlong(m));
    if (bad) atomicOr(&st->flags, 8u /* WCSPH_FLAG_NONFINITE */);
}

// ------------------------------------------------------------------
// boundary mass Psi (compute_boundary_mass.cu:10-30) — once per scene
// ------------------------------------------------------------------
__global__ void k_boundary_mass(const float4* __restrict__ bsrc, uint32_t nb, Grid g,
                                const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
                                Phys P, float4* __restrict__ bdst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const float4 xi = bsrc[i];
    int cx = min(max(cell_coord(xi.x, g.cs) - g.c0x, 1), g.nx - 2);
    int cy = min(max(cell_coord(xi.y, g.cs) - g.c0y, 1), g.ny - 2);
    int cz = min(max(cell_coord(xi.z, g.cs) - g.c0z, 1), g.nz - 2);
    const uint32_t lin = ((uint32_t)cx * g.ny + cy) * g.nz + cz;
    float sum = P.w0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint32_t c = lin + dx * g.ny * g.nz + dy * g.nz + dz;
                for (uint32_t j = bstart[c]; j < bend[c]; ++j) {
                    if (j == i) continue;
                    const float4 xj = bsrc[j];
                    const float d2 = dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
                    if (d2 <= P.h2) sum += cubic_w(sqrtf(d2) * P.inv_h, P);
                }
            }
    bdst[i] = make_float4(xi.x, xi.y, xi.z, P.rho0 / sum);
}

// ------------------------------------------------------------------
// diagnostics: sum |v| and sum |v|^2 in FP64 (sph.cu:180-197)
// ------------------------------------------------------------------
__global__ void k_metrics(const float4* __restrict__ vel, uint32_t n, double* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        const float4 v = vel[i];
        const double vx = v.x, vy = v.y, vz = v.z;
        b = vx * vx + vy * vy + vz * vz;
        a = sqrt(b);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], a);
        atomicAdd(&out[1], b);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace wcsph_dev
