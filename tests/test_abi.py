"""The C-ABI library without a GPU: it loads, exports every symbol the
header declares, and its host-only entry points behave (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import SMALL
from wcsph import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", hdr)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(wcsph_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds every one of them too
    assert names <= set(abi._PROTOS), sorted(names - set(abi._PROTOS))


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirror has the C compiler's sizes and field offsets."""
    import subprocess
    structs = {"wcsph_params": abi.Params, "wcsph_box": abi.Box, "wcsph_scene": abi.Scene,
               "wcsph_variant": abi.Variant, "wcsph_stats": abi.Stats, "wcsph_force_rule": abi.ForceRule}
    lines = ['#include <cstdio>', '#include <cstddef>', '#include "wcsph.h"', "int main() {"]
    for cname, py in structs.items():
        lines.append(f'std::printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'std::printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, val = line.split()
        py = structs[cname]
        got = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(val), (cname, field, got, val)


def test_library_carries_gfx950_code_only():
    """The device code embedded in libwcsph.so targets gfx950 and nothing
    else (a build line that loses --offload-arch falls back to another
    target, whose kernels cannot load on MI355X)."""
    blob = open(os.path.join(ROOT, "fluid-simulation_amd", "libwcsph.so"), "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}, targets


def test_presets_carry_the_reference_constants():
    ref = scenes.preset_params(abi.PRESET_REFERENCE)
    # common.cuh:18-23, compute_rho_p.cu:10-12, compute_non_pressure_accel.cu:14-16,
    # sph.cuh:28-32, update_positions.cu:10-12
    assert ref.particle_radius == np.float32(0.02)
    assert ref.stiffness == np.float32(0.1)
    assert ref.exponent == 7.0 and ref.max_density_ratio == 1.5
    assert ref.viscosity == np.float32(0.001) and ref.surface_tension == np.float32(0.13)
    assert ref.cfl_factor == np.float32(0.4)
    assert ref.dt_max == np.float32(0.0005) and ref.dt_min == np.float32(0.00001)
    assert ref.dt_np_max == np.float32(0.015)
    assert ref.collision_buffer == np.float32(0.75) and ref.elasticity == 0.5
    assert ref.pressure_gradient == abi.GRAD_SPIKY_REF and ref.viscosity_model == abi.VISC_LAPLACIAN
    assert ref.adaptive_dt == 1 and ref.split_np_dt == 1
    base = scenes.preset_params(abi.PRESET_BENCHMARK)
    assert base.particle_radius == np.float32(0.01) and base.stiffness == 1000.0
    assert base.viscosity == np.float32(0.01) and base.surface_tension == 0.0
    assert list(base.gravity) == [0.0, np.float32(-9.81), 0.0]
    with pytest.raises(ValueError):
        scenes.preset_params(7)


@pytest.mark.parametrize("desc", [SMALL, scenes.DAM_BREAK_32K,
                                  scenes.DamBreak(dims=(5, 7, 3), box_min=(-0.3, 0.1, -0.2),
                                                  box_max=(0.5, 0.6, 0.3), r=0.02, seed=99, layers=2),
                                  # one wall layer; no wall layers; a box thinner than a site in z
                                  scenes.DamBreak(dims=(3, 3, 3), box_max=(0.2, 0.2, 0.2), layers=1),
                                  scenes.DamBreak(dims=(3, 3, 3), box_max=(0.2, 0.2, 0.2), layers=0),
                                  scenes.DamBreak(dims=(2, 2, 1), box_max=(0.2, 0.2, 0.009), layers=2)])
def test_scene_generator_matches_oracle_restatement(desc):
    fl, bd, box = scenes.dam_break(desc)
    of, ob = O.make_dam_break(desc.r, desc.dims, desc.lattice_min(), np.float32(desc.jitter_frac) * np.float32(desc.r),
                              desc.seed, desc.box_min, desc.box_max, desc.layers)
    np.testing.assert_array_equal(fl, of)
    np.testing.assert_array_equal(bd, ob)
    assert box.model[0] == 1.0 and box.model_inv[15] == 1.0


def test_scene_generators_refuse_bad_sizes():
    """Refused before any site is generated: non-finite bounds, more than 64
    wall layers, more than 2^21 sites along a box edge, more than 2^32 - 1
    particles, and cluster balls more than 64 cluster sites in radius."""
    lib = abi.load_library()
    f3 = ctypes.c_float * 3

    def dam(box_max=(1.0, 1.0, 1.0), layers=3, dims=(2, 2, 2), r=0.01):
        nf, nb = ctypes.c_uint32(), ctypes.c_uint32()
        return lib.wcsph_make_dam_break(r, (ctypes.c_uint32 * 3)(*dims), f3(0.1, 0.1, 0.1), 0.0, 1, f3(0, 0, 0),
                                        f3(*box_max), layers, None, ctypes.byref(nf), None, ctypes.byref(nb))

    assert dam() == abi.WCSPH_OK
    assert dam(box_max=(float("inf"), 1.0, 1.0)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert dam(box_max=(float("nan"), 1.0, 1.0)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert dam(layers=65) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert dam(box_max=(1e6, 1.0, 1.0)) == abi.WCSPH_ERR_INVALID_ARGUMENT   # 5e7 sites along x
    assert dam(box_max=(400.0, 400.0, 400.0)) == abi.WCSPH_ERR_INVALID_ARGUMENT  # ~1e10 wall particles
    assert dam(dims=(2 ** 11, 2 ** 11, 2 ** 11)) == abi.WCSPH_ERR_INVALID_ARGUMENT  # 2^33 fluid sites
    assert dam(r=float("inf")) == abi.WCSPH_ERR_INVALID_ARGUMENT
    n = ctypes.c_uint32()
    ok = lib.wcsph_make_clustered_block(0.01, f3(0, 0, 0), f3(0.5, 0.5, 0.5), 0.1, 0.05, 0.018, 0.0, 3, None,
                                        ctypes.byref(n))
    assert ok == abi.WCSPH_OK and n.value > 0
    big = lib.wcsph_make_clustered_block(0.01, f3(0, 0, 0), f3(0.5, 0.5, 0.5), 0.1, 0.05, 1e-5, 0.0, 3, None,
                                         ctypes.byref(n))
    assert big == abi.WCSPH_ERR_INVALID_ARGUMENT
    # one 1 m triangle sampled at 10 um: 1e10 voxels over the mesh
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0.5], np.float32)
    assert lib.wcsph_sample_mesh(abi.fptr(tri), 1, 1e-5, None, 0, ctypes.byref(n)) == abi.WCSPH_ERR_INVALID_ARGUMENT


def test_benchmark_scene_sizes():
    lib = abi.load_library()
    d = scenes.DAM_BREAK_1M
    dims = (ctypes.c_uint32 * 3)(*d.dims)
    lm = (ctypes.c_float * 3)(*d.lattice_min())
    bmin = (ctypes.c_float * 3)(*d.box_min)
    bmax = (ctypes.c_float * 3)(*d.box_max)
    nf, nb = ctypes.c_uint32(), ctypes.c_uint32()
    null = ctypes.cast(None, ctypes.POINTER(ctypes.c_float))
    assert lib.wcsph_make_dam_break(np.float32(0.01), dims, lm, np.float32(0.001), 1, bmin, bmax, 3, null,
                                    ctypes.byref(nf), null, ctypes.byref(nb)) == 0
    assert nf.value == 1048576
    assert nb.value == 206 * 136 * 106 - 200 * 130 * 100 == 369696


def test_host_entry_points_reject_bad_arguments():
    lib = abi.load_library()
    b = abi.Box()
    lo = (ctypes.c_float * 3)(0, 0, 0)
    hi = (ctypes.c_float * 3)(1, -1, 1)
    assert lib.wcsph_box_aabb(lo, hi, ctypes.byref(b)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    hi = (ctypes.c_float * 3)(1, 2, 3)
    assert lib.wcsph_box_aabb(lo, hi, ctypes.byref(b)) == abi.WCSPH_OK
    assert list(b.model) == [1.0 if i % 5 == 0 else 0.0 for i in range(16)]
    p = abi.Params()
    assert lib.wcsph_default_params(5, ctypes.byref(p)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    assert lib.wcsph_create(None, None, ctypes.byref(h)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert b"NULL" in lib.wcsph_last_error(None)
    # invalid parameters are rejected before any device call
    lib.wcsph_default_params(0, ctypes.byref(p))
    p.particle_radius = -1.0
    sc = abi.Scene()
    assert lib.wcsph_create(ctypes.byref(p), ctypes.byref(sc), ctypes.byref(h)) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert b"particle_radius" in lib.wcsph_last_error(None)
    assert lib.wcsph_step(None, 1e-4, 1, None) == abi.WCSPH_ERR_INVALID_ARGUMENT
    assert lib.wcsph_destroy(None) == abi.WCSPH_OK
    assert lib.wcsph_kernel_name(8) == b"force_integrate"
    assert lib.wcsph_kernel_name(99) == b""


def _have_gpu():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_have_gpu(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    import wcsph

    fl, bd, box = scenes.dam_break(SMALL)
    with pytest.raises(wcsph.WcsphError) as e:
        wcsph.FluidSimulator(scenes.preset_params(), fl, bd, box)
    assert e.value.code == abi.WCSPH_ERR_NO_DEVICE
