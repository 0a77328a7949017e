"""Autotuning of the step's kernel variants per size bucket.

The library times candidate variants on the live scene (wcsph_autotune);
this module chooses the candidate space, keeps the winners in a JSON cache
keyed by a size bucket (scene count and log2 of the fluid count), and
reapplies a cached winner without re-timing.  It replaces the reference's
KTT searcher and frozen-config JSON (src/cuda/tuning/tuner.cpp:48-80,
config_loader.cpp:8-54).
"""
import ctypes
import json
import math
import os

from . import abi

FIELDS = ("block_size", "fluid_capacity", "boundary_capacity", "use_graph", "neighbor_mode", "neighbor_groups")


def bucket(n_fluid, n_scenes=1):
    return f"s{n_scenes}_n2^{int(math.floor(math.log2(max(n_fluid, 1))))}"


def variant_to_dict(v):
    return {f: int(getattr(v, f)) for f in FIELDS}


def dict_to_variant(d, base):
    v = abi.Variant()
    for f in FIELDS:
        setattr(v, f, int(d.get(f, getattr(base, f))))
    return v


def default_space(base):
    """Block size x neighbour form (x workgroups of the cell builder): the
    list forms, the cell-loop form and the brick-list form."""
    out = []
    for bs in (128, 256, 512):
        v = dict_to_variant(variant_to_dict(base), base)
        v.block_size, v.neighbor_mode = bs, abi.NEIGHBORS_PARTICLE
        out.append(v)
    for groups in (4096, 8192, 16384):
        v = dict_to_variant(variant_to_dict(base), base)
        v.neighbor_mode, v.neighbor_groups = abi.NEIGHBORS_CELL, groups
        out.append(v)
    for mode in (abi.NEIGHBORS_CELL_LOOP, abi.NEIGHBORS_BRICK):
        v = dict_to_variant(variant_to_dict(base), base)
        v.neighbor_mode = mode
        out.append(v)
    return out


def tune(sim, dt_hint, steps_per_trial=10, candidates=None, cache_path=None):
    """Times the candidates on `sim` (state unchanged), applies the fastest,
    records it in cache_path (if given) and returns (best, [(variant, ms)])."""
    base = sim.get_variant()
    cands = candidates if candidates is not None else default_space(base)
    arr = (abi.Variant * len(cands))(*cands)
    ms = (ctypes.c_double * len(cands))()
    best = abi.Variant()
    rc = sim.lib.wcsph_autotune(sim.handle, float(dt_hint), int(steps_per_trial), ctypes.cast(arr, ctypes.c_void_p),
                                len(cands), ms, ctypes.byref(best))
    if rc:
        from . import WcsphError

        raise WcsphError(rc, sim.lib.wcsph_last_error(sim.handle).decode())
    results = [(variant_to_dict(c), ms[k]) for k, c in enumerate(cands)]
    if cache_path:
        key = bucket(sim.n, getattr(sim, "n_scenes", 1))
        cache = load_cache(cache_path)
        cache[key] = {"variant": variant_to_dict(best), "ms_per_step": min(m for _, m in results),
                      "trials": [{"variant": d, "ms_per_step": m} for d, m in results]}
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, results


def load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply_cached(sim, cache_path):
    """Applies the cached winner of sim's bucket; returns it or None."""
    entry = load_cache(cache_path).get(bucket(sim.n, getattr(sim, "n_scenes", 1)))
    if not entry:
        return None
    v = dict_to_variant(entry["variant"], sim.get_variant())
    sim.set_variant(v)
    return v


def from_ktt_config(cfg, base):
    """Variant from a reference KTT frozen configuration (the JSON of
    measurements/configs/*.json): the block size of the force pass
    (compute_pressure_accel_n_normal_block) becomes the particle-pass block
    size (rounded up to a whole wavefront); lists are always built."""
    pairs = {}
    if isinstance(cfg, dict):
        for key in ("Pairs", "pairs", "Configuration", "configuration"):
            if key in cfg:
                cfg = cfg[key]
                break
    if isinstance(cfg, list):
        for p in cfg:
            if isinstance(p, dict) and "Name" in p and "Value" in p:
                pairs[p["Name"]] = p["Value"]
    elif isinstance(cfg, dict):
        pairs = dict(cfg)
    v = dict_to_variant(variant_to_dict(base), base)
    bs = pairs.get("compute_pressure_accel_n_normal_block") or pairs.get("compute_rho_p_block")
    if bs:
        v.block_size = max(64, min(1024, 64 * int(math.ceil(float(bs) / 64))))
    return v
