"""Python host of the gfx950 WCSPH step (over the C ABI of libwcsph.so).

`FluidSimulator` mirrors the reference's solver interface
(src/simulation/fluid_simulator.h:6-54 and CUDASPHSimulator,
src/cuda/SPH/sph.cuh:14-86): construct from options, `update(delta)` per
frame, `reset()`, `get_fluid_particles()`, `get_position_data()`,
`compute_state_metrics()`.  Errors raise `WcsphError` carrying the C status
code and the library's message.
"""
import ctypes

import numpy as np

from . import abi
from .abi import Box, Params, Scene, Stats, Variant, load_library
from .scenes import DamBreak, batch_dam_breaks, dam_break, make_box, preset_params

__all__ = ["FluidSimulator", "BatchSimulator", "WcsphError", "Params", "Box", "Variant", "Stats", "DamBreak",
           "dam_break", "batch_dam_breaks", "make_box", "preset_params", "load_library", "abi"]


class WcsphError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"wcsph error {code}: {message}")
        self.code = code


def _check(lib, handle, rc):
    if rc != abi.WCSPH_OK:
        msg = lib.wcsph_last_error(handle)
        raise WcsphError(rc, msg.decode() if msg else "")


class FluidSimulator:
    """One scene on the current HIP device.

    fluid_pos: (n, 4) or (n, 3) float32; boundary_pos: (nb, 4|3) float32;
    box: abi.Box.  The initial positions are kept for reset()
    (FluidSimulator::reset, fluid_simulator.cpp:64-66).
    """

    n_scenes = 1

    def __init__(self, params, fluid_pos, boundary_pos, box, fluid_vel=None, variant=None, domain=None):
        self.lib = load_library()
        self.params = params
        self.box = box
        self._init_pos = _as4(fluid_pos)
        self._init_vel = None if fluid_vel is None else _as4(fluid_vel)
        bpos = _as4(boundary_pos) if boundary_pos is not None and len(boundary_pos) else np.zeros((0, 4), np.float32)
        scene = Scene()
        scene.n_fluid = self._init_pos.shape[0]
        scene.n_boundary = bpos.shape[0]
        scene.fluid_pos = abi.fptr(self._init_pos)
        scene.fluid_vel = abi.fptr(self._init_vel)
        scene.boundary_pos = abi.fptr(bpos)
        scene.box = box
        if domain is not None:  # world region the cell grid must also cover (for set_box)
            for a in range(3):
                scene.domain_min[a] = domain[0][a]
                scene.domain_max[a] = domain[1][a]
        h = ctypes.c_void_p()
        rc = self.lib.wcsph_create(ctypes.byref(params), ctypes.byref(scene), ctypes.byref(h))
        _check(self.lib, None, rc)
        self.handle = h
        self.n = scene.n_fluid
        self.nb = scene.n_boundary
        if variant is not None:
            self.set_variant(variant)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            self.lib.wcsph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference interface -----------------------------------------
    def update(self, delta):
        """One step (FluidSimulator::update, fluid_simulator.h:33)."""
        return self.step(delta, 1)

    def step(self, delta, n_steps=1):
        st = Stats()
        rc = self.lib.wcsph_step(self.handle, float(delta), int(n_steps), ctypes.byref(st))
        _check(self.lib, self.handle, rc)
        return st

    def step_async(self, delta, n_steps=1):
        _check(self.lib, self.handle, self.lib.wcsph_step_async(self.handle, float(delta), int(n_steps)))

    def sync(self):
        st = Stats()
        _check(self.lib, self.handle, self.lib.wcsph_sync(self.handle, ctypes.byref(st)))
        return st

    def reset(self):
        """Back to the initial lattice with zero velocities (fluid_simulator.cpp:64-66, sph.cu:76-81)."""
        self.set_state(self._init_pos, None)

    def get_fluid_particles(self):
        return self.n

    def get_position_data(self):
        return self.get_state()["pos"]

    def compute_state_metrics(self):
        """(mean |v|, kinetic energy) in FP64 (compute_state_metrics, sph.cu:180-197)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(self.lib, self.handle, self.lib.wcsph_state_metrics(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # -- state / parity hooks --------------------------------------------
    def get_state(self):
        n = self.n
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        rho = np.empty(n, np.float32)
        p = np.empty(n, np.float32)
        ids = np.empty(n, np.uint32)
        rc = self.lib.wcsph_get_state(self.handle, abi.fptr(pos), abi.fptr(vel), abi.fptr(rho), abi.fptr(p),
                                      abi.uptr(ids))
        _check(self.lib, self.handle, rc)
        return {"pos": pos, "vel": vel, "rho": rho, "p": p, "ids": ids}

    def set_state(self, pos, vel=None):
        pos4 = _as4(pos)
        vel4 = None if vel is None else _as4(vel)
        _check(self.lib, self.handle, self.lib.wcsph_set_state(self.handle, abi.fptr(pos4), abi.fptr(vel4)))

    def permutation(self):
        perm = np.empty(self.n, np.uint32)
        _check(self.lib, self.handle, self.lib.wcsph_get_permutation(self.handle, abi.uptr(perm)))
        return perm

    def sorted_positions(self):
        pos = np.empty((self.n, 4), np.float32)
        _check(self.lib, self.handle, self.lib.wcsph_get_sorted_positions(self.handle, abi.fptr(pos)))
        return pos

    def neighbors(self, which=0):
        """(counts, CSR list) of the last step's fluid (0) or boundary (1) lists."""
        counts = np.empty(self.n, np.uint32)
        total = ctypes.c_uint64()
        rc = self.lib.wcsph_get_neighbors(self.handle, which, abi.uptr(counts), abi.uptr(None), 0,
                                          ctypes.byref(total))
        _check(self.lib, self.handle, rc)
        lst = np.empty(max(total.value, 1), np.uint32)
        rc = self.lib.wcsph_get_neighbors(self.handle, which, abi.uptr(counts), abi.uptr(lst), lst.size,
                                          ctypes.byref(total))
        _check(self.lib, self.handle, rc)
        return counts, lst[: total.value]

    def neighbor_total(self, which=0):
        """Number of fluid (0) or boundary (1) list entries of the last step
        (counts only: the lists are not copied)."""
        counts = np.empty(max(self.n, 1), np.uint32)
        total = ctypes.c_uint64()
        rc = self.lib.wcsph_get_neighbors(self.handle, which, abi.uptr(counts), abi.uptr(None), 0,
                                          ctypes.byref(total))
        _check(self.lib, self.handle, rc)
        return int(total.value)

    def boundary(self):
        out = np.empty((self.nb, 4), np.float32)
        _check(self.lib, self.handle, self.lib.wcsph_get_boundary(self.handle, abi.fptr(out)))
        return out

    # -- differentiable runs ---------------------------------------------
    def set_viscosity(self, alpha):
        _check(self.lib, self.handle, self.lib.wcsph_set_viscosity(self.handle, float(alpha)))

    def grad_forward(self, vel0, dt, n_steps, checkpoint_every, out=None):
        """Run from the creation positions with velocities vel0 ((n, 4) float32
        numpy array or torch tensor on the GPU); returns the final positions
        (n, 4) in creation order, in the same kind of array."""
        vin, vptr = _ptr4(vel0, self.n)
        if out is None:
            out = _empty_like4(vin, self.n)
        _, optr = _ptr4(out, self.n)
        rc = self.lib.wcsph_grad_forward(self.handle, vptr, float(dt), int(n_steps), int(checkpoint_every), optr)
        _check(self.lib, self.handle, rc)
        return out

    def grad_generation(self):
        """Generation of the last successful grad_forward (0: none yet)."""
        return int(self.lib.wcsph_grad_generation(self.handle))

    def grad_backward(self, grad_pos, out=None):
        """dL/d(final positions) (n, 4) -> (dL/d(vel0) (n, 4), dL/d(alpha))."""
        gin, gptr = _ptr4(grad_pos, self.n)
        if out is None:
            out = _empty_like4(gin, self.n)
        _, optr = _ptr4(out, self.n)
        ga = ctypes.c_double()
        rc = self.lib.wcsph_grad_backward(self.handle, gptr, optr, ctypes.byref(ga))
        _check(self.lib, self.handle, rc)
        return out, ga.value

    def set_box(self, box, scene=0):
        """Box::set_model for the following steps (src/render/box.cpp:25-29)."""
        _check(self.lib, self.handle, self.lib.wcsph_set_box(self.handle, scene, ctypes.byref(box)))
        if scene == 0:
            self.box = box

    def get_box(self, scene=0):
        box = abi.Box()
        _check(self.lib, self.handle, self.lib.wcsph_get_box(self.handle, scene, ctypes.byref(box)))
        return box

    def set_variant(self, variant):
        _check(self.lib, self.handle, self.lib.wcsph_set_variant(self.handle, ctypes.byref(variant)))

    def set_scene_neighbor_modes(self, modes):
        """Per-scene list builders (NEIGHBORS_CELL / NEIGHBORS_PARTICLE, one
        per scene), or None for the variant's builder everywhere."""
        if modes is None:
            rc = self.lib.wcsph_set_scene_neighbor_modes(self.handle, None)
        else:
            m = np.ascontiguousarray(modes, np.uint32)
            if m.shape != (self.n_scenes,):
                raise ValueError(f"one builder per scene: {self.n_scenes} expected")
            rc = self.lib.wcsph_set_scene_neighbor_modes(self.handle, abi.uptr(m))
        _check(self.lib, self.handle, rc)

    def get_scene_neighbor_modes(self):
        out = np.empty(self.n_scenes, np.uint32)
        _check(self.lib, self.handle, self.lib.wcsph_get_scene_neighbor_modes(self.handle, abi.uptr(out)))
        return out

    def get_variant(self):
        v = Variant()
        _check(self.lib, self.handle, self.lib.wcsph_get_variant(self.handle, ctypes.byref(v)))
        return v


class BatchSimulator(FluidSimulator):
    """Many independent scenes stepped together (one launch per kernel for
    all scenes, per-scene CFL dt).  `scenes` is a list of
    (fluid_pos, boundary_pos, box) or (fluid_pos, boundary_pos, box, fluid_vel).
    State arrays are the scenes concatenated in order; scene(s) gives the
    slice of scene s."""

    def __init__(self, params, scenes, variant=None):
        self.lib = load_library()
        self.params = params
        keep = []
        arr = (Scene * len(scenes))()
        init_pos, init_vel = [], []
        for k, sc in enumerate(scenes):
            fl, bd, box = sc[0], sc[1], sc[2]
            vel = sc[3] if len(sc) > 3 else None
            f4 = _as4(fl)
            b4 = _as4(bd) if bd is not None and len(bd) else np.zeros((0, 4), np.float32)
            v4 = None if vel is None else _as4(vel)
            keep += [f4, b4, v4]
            init_pos.append(f4)
            init_vel.append(v4 if v4 is not None else np.zeros_like(f4))
            arr[k].n_fluid = f4.shape[0]
            arr[k].n_boundary = b4.shape[0]
            arr[k].fluid_pos = abi.fptr(f4)
            arr[k].fluid_vel = abi.fptr(v4)
            arr[k].boundary_pos = abi.fptr(b4)
            arr[k].box = box
        h = ctypes.c_void_p()
        rc = self.lib.wcsph_create_batch(ctypes.byref(params), arr, len(scenes), ctypes.byref(h))
        _check(self.lib, None, rc)
        self.handle = h
        self._init_pos = np.ascontiguousarray(np.concatenate(init_pos))
        self._init_vel = None
        self.n = int(self.lib.wcsph_fluid_count(h))
        self.nb = int(self.lib.wcsph_boundary_count(h))
        self.n_scenes = len(scenes)
        self.ranges = []
        for k in range(self.n_scenes):
            a, b, c, d = (ctypes.c_uint32() for _ in range(4))
            _check(self.lib, h, self.lib.wcsph_scene_range(h, k, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                            ctypes.byref(d)))
            self.ranges.append((a.value, b.value, c.value, d.value))
        if variant is not None:
            self.set_variant(variant)

    def scene(self, s):
        """slice of scene s in the concatenated fluid arrays"""
        p0, n = self.ranges[s][0], self.ranges[s][1]
        return slice(p0, p0 + n)

    def scene_stats(self, s):
        st = Stats()
        _check(self.lib, self.handle, self.lib.wcsph_scene_stats(self.handle, s, ctypes.byref(st)))
        return st

    def scene_metrics(self, s):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(self.lib, self.handle, self.lib.wcsph_scene_metrics(self.handle, s, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def _ptr4(a, n):
    """(array, pointer) of an (n, 4) float32 C-contiguous numpy array or CUDA
    tensor (host and device pointers are both accepted by the C ABI)."""
    if a is None:
        return None, None
    if hasattr(a, "data_ptr"):  # torch tensor
        if a.dtype.__str__() != "torch.float32" or not a.is_contiguous() or tuple(a.shape) != (n, 4):
            raise ValueError("expected a contiguous float32 tensor of shape (n, 4)")
        return a, ctypes.c_void_p(a.data_ptr())
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != (n, 4):
        raise ValueError("expected an (n, 4) float32 array")
    return a, ctypes.c_void_p(a.ctypes.data)


def _empty_like4(a, n):
    if a is not None and hasattr(a, "data_ptr"):
        import torch

        return torch.empty((n, 4), dtype=torch.float32, device=a.device)
    return np.empty((n, 4), np.float32)


def _as4(a):
    a = np.asarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] not in (3, 4):
        raise ValueError("expected an (n, 3) or (n, 4) array")
    if a.shape[1] == 3:
        a = np.concatenate([a, np.ones((a.shape[0], 1), np.float32)], axis=1)
    return np.ascontiguousarray(a)
