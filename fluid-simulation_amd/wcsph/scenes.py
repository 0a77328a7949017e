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


def preset_params(preset=abi.PRESET_BASELINE, **overrides):
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
