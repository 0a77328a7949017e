"""Scene descriptions of the dam-break workloads, built by the library's own
generator (wcsph_make_dam_break).

Geometry notes (the stated benchmark geometry does not fit as written):
 * 1,048,576 particles: a 128 x 128 x 64 lattice (x, y, z) at spacing
   2r = 0.02 m spans 2.56 x 2.56 x 1.28 m, taller than a 2 m box, so the box
   is 4.0 x 2.6 x 2.0 m (y raised from 2.0 to 2.6 m).
 * Walls are 3 shells of particles at spacing 2r on the continuation of
   the fluid lattice, so the wall count follows from the box size
   (369,696 for the 4.0 x 2.6 x 2.0 m box).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi


def preset_params(preset=abi.PRESET_BENCHMARK, **overrides):
    lib = abi.load_library()
    p = abi.Params()
    rc = lib.wcsph_default_params(preset, ctypes.byref(p))
    if rc:
        raise ValueError(f"unknown preset {preset}")
    for k, v in overrides.items():
        if k == "gravity":
            for a in range(3):
                p.gravity[a] = v[a]
        else:
            setattr(p, k, v)
    return p


def make_box(bmin, bmax, model=None):
    box = abi.Box()
    for a in range(3):
        box.min[a] = bmin[a]
        box.max[a] = bmax[a]
    m = np.eye(4, dtype=np.float64) if model is None else np.asarray(model, np.float64)
    inv = np.linalg.inv(m)
    # column-major like glm
    for c in range(4):
        for r in range(4):
            box.model[c * 4 + r] = m[r, c]
            box.model_inv[c * 4 + r] = inv[r, c]
    return box


@dataclass
class DamBreak:
    dims: tuple = (32, 32, 32)
    box_max: tuple = (1.6, 1.0, 1.0)
    box_min: tuple = (0.0, 0.0, 0.0)
    r: float = 0.01
    jitter_frac: float = 0.1  # uniform jitter in [-0.1 r, +0.1 r)
    seed: int = 12345
    layers: int = 3

    def lattice_min(self):
        # first fluid site one radius inside the box, on the wall lattice
        return tuple(np.float32(self.box_min[a]) + np.float32(self.r) for a in range(3))


def dam_break(desc: DamBreak):
    """Returns (fluid_pos (n,4), boundary_pos (nb,4), box)."""
    lib = abi.load_library()
    dims = (ctypes.c_uint32 * 3)(*desc.dims)
    lmin = (ctypes.c_float * 3)(*desc.lattice_min())
    bmin = (ctypes.c_float * 3)(*desc.box_min)
    bmax = (ctypes.c_float * 3)(*desc.box_max)
    nf, nb = ctypes.c_uint32(), ctypes.c_uint32()
    r = np.float32(desc.r)
    jit = np.float32(desc.jitter_frac) * r
    null = ctypes.cast(None, ctypes.POINTER(ctypes.c_float))
    rc = lib.wcsph_make_dam_break(r, dims, lmin, jit, desc.seed, bmin, bmax, desc.layers, null,
                                  ctypes.byref(nf), null, ctypes.byref(nb))
    if rc:
        raise ValueError("bad dam-break description")
    fluid = np.empty((nf.value, 4), np.float32)
    bnd = np.empty((nb.value, 4), np.float32)
    rc = lib.wcsph_make_dam_break(r, dims, lmin, jit, desc.seed, bmin, bmax, desc.layers, abi.fptr(fluid),
                                  ctypes.byref(nf), abi.fptr(bnd), ctypes.byref(nb))
    if rc:
        raise ValueError("bad dam-break description")
    return fluid, bnd, make_box(desc.box_min, desc.box_max)


# The benchmark workloads.
DAM_BREAK_32K = DamBreak(dims=(32, 32, 32), box_max=(1.6, 1.0, 1.0))
DAM_BREAK_1M = DamBreak(dims=(128, 128, 64), box_max=(4.0, 2.6, 2.0))


def batch_dam_breaks(n_scenes=256, lo=2000, hi=200000, seed=2024, r=0.01):
    """The batched workload: n_scenes dam breaks whose fluid counts are drawn
    log-uniform in [lo, hi] (seeded); each lattice is about cubic and sits in
    a box scaled to it like the 32^3 scene (2.5 x, 1.5625 y and z of the
    lattice extent, rounded to whole wall spacings)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_scenes))
    s = 2.0 * r
    out = []
    for k, c in enumerate(counts):
        side = max(2, int(round(c ** (1.0 / 3.0))))
        nz = max(2, int(round(c / (side * side))))
        dims = (side, side, nz)
        ext = [d * s for d in dims]
        box = (np.ceil(2.5 * ext[0] / s) * s, np.ceil(1.5625 * ext[1] / s) * s, np.ceil(1.5625 * ext[2] / s) * s)
        out.append(DamBreak(dims=dims, box_max=tuple(float(b) for b in box), r=r, seed=seed + 1 + k))
    return out


# -- meshes, mesh sampling and the obstacle-collapse workload ---------------

def procedural_mesh(kind, params, resolution=64, model=None):
    """Triangle soup (n, 3, 3) float32 of a procedural cube / sphere / torus
    (wcsph_mesh_procedural); model is a 4x4 world matrix (row-major numpy)."""
    lib = abi.load_library()
    prm = (ctypes.c_float * 3)(*(list(params) + [0.0] * (3 - len(params))))
    mdl = None
    if model is not None:
        m = np.asarray(model, np.float32)
        mdl = (ctypes.c_float * 16)(*m.T.reshape(-1))  # column-major like glm
    mp = ctypes.cast(mdl, ctypes.POINTER(ctypes.c_float)) if mdl is not None else None
    nt = ctypes.c_uint32()
    rc = lib.wcsph_mesh_procedural(kind, prm, resolution, mp, None, 0, ctypes.byref(nt))
    if rc:
        raise ValueError("bad mesh description")
    out = np.empty((nt.value, 3, 3), np.float32)
    rc = lib.wcsph_mesh_procedural(kind, prm, resolution, mp, abi.fptr(out), nt.value, ctypes.byref(nt))
    if rc:
        raise ValueError("bad mesh description")
    return out


def sample_mesh(triangles, voxel_size):
    """Boundary particles (n, 4) on a triangle soup: centres of the voxels
    that intersect it (wcsph_sample_mesh; the reference's
    generate_boundary_particles, src/simulation/SPH/boundary.cpp:8-35,
    uses voxel_size = 0.75 * 2r)."""
    lib = abi.load_library()
    tri = np.ascontiguousarray(triangles, np.float32).reshape(-1, 9)
    n = ctypes.c_uint32()
    rc = lib.wcsph_sample_mesh(abi.fptr(tri), tri.shape[0], float(voxel_size), None, 0, ctypes.byref(n))
    if rc:
        raise ValueError("bad mesh or voxel size")
    out = np.empty((n.value, 4), np.float32)
    rc = lib.wcsph_sample_mesh(abi.fptr(tri), tri.shape[0], float(voxel_size), abi.fptr(out), n.value,
                               ctypes.byref(n))
    if rc:
        raise ValueError("bad mesh or voxel size")
    return out


def clustered_block(r, block_min, block_max, cluster_fraction=0.1, cluster_radius=0.1, cluster_spacing=None,
                    jitter_frac=0.1, seed=12345):
    """Fluid block with denser clusters (wcsph_make_clustered_block);
    cluster_spacing defaults to 1.8 r."""
    lib = abi.load_library()
    cs = np.float32(1.8) * np.float32(r) if cluster_spacing is None else cluster_spacing
    bmin = (ctypes.c_float * 3)(*block_min)
    bmax = (ctypes.c_float * 3)(*block_max)
    jit = float(np.float32(jitter_frac) * np.float32(r))
    n = ctypes.c_uint32()
    args = (float(r), bmin, bmax, float(cluster_fraction), float(cluster_radius), float(cs), jit, seed)
    rc = lib.wcsph_make_clustered_block(*args, None, ctypes.byref(n))
    if rc:
        raise ValueError("bad clustered-block description")
    out = np.empty((n.value, 4), np.float32)
    rc = lib.wcsph_make_clustered_block(*args, abi.fptr(out), ctypes.byref(n))
    if rc:
        raise ValueError("bad clustered-block description")
    return out


def translation(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


@dataclass
class ObstacleCollapse:
    """Column collapse onto a sphere + torus obstacle.  A 2 x 1 x 1 m fluid
    block (x from -2.1 to -0.1, y from 0 to 1, z from -0.5 to 0.5) with 10%
    of its particles in denser clusters (spacing 1.8 r) falls towards a
    sphere of radius 0.5 m centred at (1.1, 0.5, 0) ringed by a torus
    (R = 0.8, r = 0.2 m, axis y) centred at (1.1, 0.2, 0), inside the box
    [-2.2, 2.2] x [0, 2] x [-1.2, 1.2].  The obstacle is triangulated and
    voxel-sampled into boundary particles at spacing `obstacle_spacing_mult`
    * 2r; the box keeps the fluid by collision only."""
    r: float = 0.003125
    block_min: tuple = (-2.1, 0.0, -0.5)
    block_max: tuple = (-0.1, 1.0, 0.5)
    box_min: tuple = (-2.2, 0.0, -1.2)
    box_max: tuple = (2.2, 2.0, 1.2)
    sphere_centre: tuple = (1.1, 0.5, 0.0)
    sphere_radius: float = 0.5
    torus_centre: tuple = (1.1, 0.2, 0.0)
    torus_radii: tuple = (0.8, 0.2)
    cluster_fraction: float = 0.1
    cluster_radius: float = 0.1
    obstacle_spacing_mult: float = 1.0
    mesh_resolution: int = 96
    jitter_frac: float = 0.1
    seed: int = 4242

    def scaled(self, k):
        """The same scene with every length multiplied by k (particle count
        scales by k^3 at a fixed r)."""
        sc = lambda t: tuple(float(np.float32(k) * np.float32(v)) for v in t)  # noqa: E731
        return ObstacleCollapse(r=self.r, block_min=sc(self.block_min), block_max=sc(self.block_max),
                                box_min=sc(self.box_min), box_max=sc(self.box_max),
                                sphere_centre=sc(self.sphere_centre), sphere_radius=k * self.sphere_radius,
                                torus_centre=sc(self.torus_centre), torus_radii=sc(self.torus_radii),
                                cluster_fraction=self.cluster_fraction, cluster_radius=k * self.cluster_radius,
                                obstacle_spacing_mult=self.obstacle_spacing_mult,
                                mesh_resolution=self.mesh_resolution, jitter_frac=self.jitter_frac, seed=self.seed)


def obstacle_collapse(desc: ObstacleCollapse):
    """Returns (fluid_pos (n,4), boundary_pos (nb,4), box)."""
    fluid = clustered_block(desc.r, desc.block_min, desc.block_max, desc.cluster_fraction, desc.cluster_radius,
                            jitter_frac=desc.jitter_frac, seed=desc.seed)
    sphere = procedural_mesh(abi.MESH_SPHERE, (desc.sphere_radius,), desc.mesh_resolution,
                             translation(desc.sphere_centre))
    torus = procedural_mesh(abi.MESH_TORUS, desc.torus_radii, desc.mesh_resolution, translation(desc.torus_centre))
    voxel = float(np.float32(desc.obstacle_spacing_mult) * np.float32(2.0) * np.float32(desc.r))
    bnd = np.concatenate([sample_mesh(sphere, voxel), sample_mesh(torus, voxel)])
    return fluid, bnd, make_box(desc.box_min, desc.box_max)


OBSTACLE_COLLAPSE_8M = ObstacleCollapse()
