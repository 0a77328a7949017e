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

FIELDS = ("block_size", "fluid_capacity", "boundary_capacity", "use_graph", "neighbor_mode", "neighbor_groups",
          "cell_size_mult")


def bucket(n_fluid, n_scenes=1):
    return f"s{n_scenes}_n2^{int(math.floor(math.log2(max(n_fluid, 1))))}"


def _cast(f, x):
    return (float(x) if float(x) != 0.0 else 1.0) if f == "cell_size_mult" else int(x)


def variant_to_dict(v):
    return {f: _cast(f, getattr(v, f)) for f in FIELDS}


def dict_to_variant(d, base):
    v = abi.Variant()
    for f in FIELDS:
        setattr(v, f, _cast(f, d.get(f, getattr(base, f))))
    return v


def default_space(base, cell_size_mults=(2.0 / 3.0, 0.5)):
    """Block size x neighbour form (x workgroups of the cell builder): the
    list forms, the cell-loop form and the brick-list form; and the
    per-particle builder with the reference's finer cell sizes
    (CELL_SIZE_MULT 2/3 and 1/2, step_tuner.cuh:88)."""
    out = []
    for bs in (128, 256, 512):
        v = dict_to_variant(variant_to_dict(base), base)
        v.block_size, v.neighbor_mode, v.cell_size_mult = bs, abi.NEIGHBORS_PARTICLE, 1.0
        out.append(v)
    for groups in (4096, 8192, 16384):
        v = dict_to_variant(variant_to_dict(base), base)
        v.neighbor_mode, v.neighbor_groups, v.cell_size_mult = abi.NEIGHBORS_CELL, groups, 1.0
        out.append(v)
    for mode in (abi.NEIGHBORS_CELL_LOOP, abi.NEIGHBORS_BRICK):
        v = dict_to_variant(variant_to_dict(base), base)
        v.neighbor_mode = mode
        v.cell_size_mult = 1.0
        out.append(v)
    for mult in cell_size_mults:
        v = dict_to_variant(variant_to_dict(base), base)
        v.neighbor_mode, v.block_size, v.cell_size_mult = abi.NEIGHBORS_PARTICLE, 256, mult
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


# Scene-size buckets of a batch (fluid particles): [0, 2^14), [2^14, 2^16), [2^16, inf).
SCENE_BUCKET_EDGES = (1 << 14, 1 << 16)


def scene_buckets(counts, edges=SCENE_BUCKET_EDGES):
    """Bucket index of each scene (the number of edges at or below its fluid
    count) and the scenes of each nonempty bucket, {bucket: [scene, ...]}."""
    idx = [sum(1 for e in edges if c >= e) for c in counts]
    groups = {}
    for k, b in enumerate(idx):
        groups.setdefault(b, []).append(k)
    return idx, groups


def tune_scene_buckets(params, scenes, dt_hint, batch=None, edges=SCENE_BUCKET_EDGES, steps_per_trial=5):
    """Per-size-bucket list builders of a batch.  `scenes` is the batch's
    scene list (as given to BatchSimulator).  The scenes of each size bucket
    are stepped as a batch of their own, on which the library's autotuner
    times the per-particle builder and the cell builder (state restored);
    each bucket's scenes take the builder that was faster per particle-step.
    Applied to `batch` (wcsph_set_scene_neighbor_modes) when given: when
    the buckets chose different builders, the batch is timed with the mixed
    builders and with the builder of most particles alone, and the faster
    is kept; the batch's state is restored after every trial.  Returns
    (modes, report): modes[s] for every scene, and per bucket its scenes,
    particles and the trial times (plus the mixed-versus-one comparison
    when it ran)."""
    from . import BatchSimulator

    counts = [int(sc[0].shape[0]) for sc in scenes]
    _, groups = scene_buckets(counts, edges)
    modes = [abi.NEIGHBORS_PARTICLE] * len(scenes)
    report = []
    for b in sorted(groups):
        members = groups[b]
        sub = BatchSimulator(params, [scenes[k] for k in members])
        try:
            base = sub.get_variant()
            cands = []
            for mode, groups_ in ((abi.NEIGHBORS_PARTICLE, base.neighbor_groups), (abi.NEIGHBORS_CELL, 8192),
                                  (abi.NEIGHBORS_CELL, 16384)):
                v = dict_to_variant(variant_to_dict(base), base)
                v.neighbor_mode, v.neighbor_groups, v.cell_size_mult = mode, groups_, 1.0
                cands.append(v)
            best, trials = tune(sub, dt_hint, steps_per_trial=steps_per_trial, candidates=cands)
        finally:
            sub.close()
        for k in members:
            modes[k] = int(best.neighbor_mode)
        report.append({"bucket": b, "scenes": len(members), "fluid_particles": sum(counts[k] for k in members),
                       "builder": int(best.neighbor_mode), "trials": trials})
    if batch is not None:
        if len(set(modes)) > 1:
            # Mixed builders run both builders' launches over the whole batch:
            # kept only if the batch itself steps faster so than with the
            # builder that covers most particles (both timed on the batch,
            # its state restored).
            counts_by = {}
            for k, m in enumerate(modes):
                counts_by[m] = counts_by.get(m, 0) + counts[k]
            main = max(counts_by, key=counts_by.get)
            v = batch.get_variant()
            batch.set_scene_neighbor_modes(modes)
            _, t_mixed = tune(batch, dt_hint, steps_per_trial=steps_per_trial, candidates=[v])
            batch.set_scene_neighbor_modes(None)
            vu = dict_to_variant(variant_to_dict(v), v)
            vu.neighbor_mode = main
            _, t_one = tune(batch, dt_hint, steps_per_trial=steps_per_trial, candidates=[vu])
            report.append({"mixed_ms_per_step": t_mixed[0][1], "one_builder_ms_per_step": t_one[0][1],
                           "one_builder": int(main)})
            if t_mixed[0][1] >= t_one[0][1]:
                modes = [int(main)] * len(scenes)
        batch.set_scene_neighbor_modes(modes)
    return modes, report


# KTT parameters of the reference's step (step_tuner.cuh:80-121) that have no
# counterpart here, with the reason; from_ktt_config reports them.
KTT_IGNORED = (
    (r"_u_n$", "loop unrolling is fixed at compile time"),
    (r"_shared$|_cache_log2$", "the LDS forms are chosen by neighbor_mode"),
    (r"^(MIN_)?(BOUNDARY_)?TABLE_SIZE(_MULT)?$", "the cell grid is dense: no hash table"),
    (r"_block$", "one block size serves every particle pass (the force pass's is taken)"),
    (r"^EXTERNAL_FORCE$", "the external force is the scene's force rules"),
    (r"^KERNEL_DIR$", "no run-time compilation"),
)


def from_ktt_config(cfg, base, ignored=None):
    """Variant from a reference KTT frozen configuration (the JSON of
    measurements/configs/*.json): CELL_SIZE_MULT becomes cell_size_mult
    (with the per-particle builder when it is not 1); the block size of the
    force pass (compute_pressure_accel_n_normal_block) becomes the
    particle-pass block size (rounded up to a whole wavefront); lists are
    built unless every SPH pass had *_nlist = 0 (then the cell-loop form).
    `ignored`, a list, receives "name: reason" for every parameter without a
    counterpart."""
    import re

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
    mult = float(pairs.get("CELL_SIZE_MULT", 1.0))
    v.cell_size_mult = mult
    nlist = [pairs.get(k) for k in ("compute_rho_p_nlist", "compute_pressure_accel_n_normal_nlist",
                                    "compute_non_pressure_accel_nlist")]
    if mult != 1.0:
        v.neighbor_mode = abi.NEIGHBORS_PARTICLE
    elif all(x is not None and int(x) == 0 for x in nlist):
        v.neighbor_mode = abi.NEIGHBORS_CELL_LOOP
    used = {"CELL_SIZE_MULT", "compute_pressure_accel_n_normal_block", "compute_rho_p_nlist",
            "compute_pressure_accel_n_normal_nlist", "compute_non_pressure_accel_nlist"}
    if ignored is not None:
        for name in pairs:
            if name in used:
                continue
            why = next((r for pat, r in KTT_IGNORED if re.search(pat, name)), "no counterpart")
            ignored.append(f"{name}: {why}")
    return v
