"""Shared scene set-up of the tests (small dam breaks the oracle runs in seconds)."""
import numpy as np

from wcsph import scenes
from wcsph.abi import PRESET_BENCHMARK, PRESET_REFERENCE

SMALL = scenes.DamBreak(dims=(16, 16, 16), box_max=(0.64, 0.5, 0.5), seed=7)
SMALL_REF = scenes.DamBreak(dims=(12, 12, 12), box_max=(1.2, 1.0, 1.0), r=0.02, seed=11)


def small_scene(desc=SMALL):
    return scenes.dam_break(desc)


def baseline_params(**kw):
    return scenes.preset_params(PRESET_BENCHMARK, **kw)


def reference_params(**kw):
    return scenes.preset_params(PRESET_REFERENCE, **kw)


def cell_size(params):
    r = np.float32(params.particle_radius)
    return np.float32(np.float32(2.0) * np.float32(np.float32(2.0) * r))


def by_id(st):
    order = np.argsort(st["ids"], kind="stable")
    return {k: v[order] for k, v in st.items()}


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))
