"""The headless driver's command line and scene set-up (no device needed):
flags and messages of the reference CLI (src/cli.cu:7-101), KTT config
import, and the scenes' particle counts against the Python scene builders."""
import json
import os
import subprocess

import numpy as np
import pytest

from wcsph import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fluid-simulation_amd", "wcsph_headless")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args):
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout, p.stderr


def test_help_lists_every_scene():
    rc, out, err = run("--help")
    assert rc == 0
    for name in ["dam_break_32k", "dam_break_1m", "obstacle_collapse_8m", "dragon_collision", "surface_tension",
                 "tilting_box", "fountain", "vortex", "BIG"]:
        assert name in err
    for flag in ["--snapshot-load", "--snapshot-save", "--frozen-config", "--searcher", "--tuning-budget", "--stop",
                 "--fixed-dt", "--warmup-iters", "--ktt-output", "--log-csv", "--log-metrics"]:
        assert flag in err


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "Unknown argument: --bogus"),
    (["--headless"], "--stop is required in --headless mode"),
    (["--stop", "steps=3"], "Unknown --stop kind: steps"),
    (["--stop", "iters=3x"], "Invalid --stop value: 3x"),
    (["--stop", "iters"], "--stop must be iters=N, sim-time=T, or wall-time=T"),
    (["--searcher", "greedy"], "Unknown searcher: greedy"),
    (["--log-csv"], "--log-csv requires an argument"),
    (["--app", "nowhere", "--print-scene"], "Unknown app: nowhere"),
    (["--frozen-config", "/nonexistent.json", "--print-options"], "Cannot open config file: /nonexistent.json"),
])
def test_errors(args, msg):
    rc, out, err = run(*args)
    assert rc == 1 and msg in err


def test_frozen_config_with_a_bad_number(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text('{"Configuration": [{"Name": "CELL_SIZE_MULT", "Value": "2/3x", "ValueType": "double"}]}')
    rc, out, err = run("--frozen-config", str(path), "--print-options")
    assert rc == 1 and "Config value of CELL_SIZE_MULT is not a number: '2/3x'" in err


def test_options_parse():
    rc, out, err = run("--headless", "--app", "fountain", "--stop", "sim-time=0.25", "--fixed-dt", "0.002",
                       "--warmup-iters", "7", "--searcher", "full", "--tuning-budget", "0", "--snapshot-load", "a.sphs",
                       "--snapshot-save", "b.sphs", "--ktt-output", "kt", "--log-csv", "log.csv", "--log-metrics",
                       "--mesh", "d.obj", "--print-options")
    assert rc == 0, err
    o = json.loads(out.splitlines()[0])
    assert o == {"app": "fountain", "headless": True, "stop": "sim-time", "stop_value": 0.25,
                 "fixed_dt": float(np.float32(0.002)), "warmup_iters": 7, "searcher": "full", "tuning_budget": 0.0,
                 "snapshot_load": "a.sphs", "snapshot_save": "b.sphs", "frozen_config": None, "ktt_output": "kt",
                 "log_csv": "log.csv", "log_metrics": True, "mesh": "d.obj"}
    rc, out, _ = run("--print-options")
    d = json.loads(out)
    assert (d["stop"], d["fixed_dt"], d["tuning_budget"], d["searcher"]) == ("none", float(np.float32(0.01)),
                                                                             float(np.float32(0.1)), "mcmc")


@pytest.mark.parametrize("state", ["dragon_falling", "dragon_settled", "surface_settled", "tilting_chaotic",
                                   "fountain_active", "vortex_active"])
def test_frozen_config_import_matches_python(state):
    path = os.path.join(GOLD, f"ktt_{state}.json")
    rc, out, err = run("--frozen-config", path, "--print-options")
    assert rc == 0, err
    got = json.loads(out.splitlines()[1])
    from wcsph import autotune
    base = abi.Variant(256, 64, 32, 1, abi.NEIGHBORS_PARTICLE, 16384, 1.0)
    v = autotune.from_ktt_config(json.load(open(path)), base)
    assert got["frozen_block_size"] == v.block_size and got["frozen_neighbor_mode"] == v.neighbor_mode
    assert np.float32(got["frozen_cell_size_mult"]) == np.float32(v.cell_size_mult)  # printed with 9 digits


def test_frozen_config_reports_parameters_without_counterpart(tmp_path):
    """The driver names every KTT parameter it does not map (on stderr, before
    the device is touched, so the message comes also without a GPU)."""
    path = os.path.join(GOLD, "ktt_fountain_active.json")
    rc, out, err = run("--headless", "--app", "fountain", "--stop", "iters=1", "--frozen-config", path)
    import torch
    if torch.cuda.is_available():
        pytest.skip("the run steps on a device here")
    assert "KTT parameters have no counterpart here" in err, err
    assert "MIN_TABLE_SIZE: the cell grid is dense: no hash table" in err
    assert "CELL_SIZE_MULT" not in err and "compute_pressure_accel_n_normal_block" not in err


def scene_counts(app):
    rc, out, err = run("--app", app, "--print-scene")
    assert rc == 0, err
    return json.loads(out)


def test_benchmark_scenes_match_python_builders():
    for app, desc in [("dam_break_32k", scenes.DAM_BREAK_32K), ("dam_break_1m", scenes.DAM_BREAK_1M)]:
        fl, bd, _ = scenes.dam_break(desc)
        assert scene_counts(app) == {"app": app, "fluid": fl.shape[0], "boundary": bd.shape[0]}
    fl, bd, _ = scenes.obstacle_collapse(scenes.OBSTACLE_COLLAPSE_8M)
    assert scene_counts("obstacle_collapse_8m") == {"app": "obstacle_collapse_8m", "fluid": fl.shape[0],
                                                     "boundary": bd.shape[0]}


def test_reference_scenes():
    # lattice sizes of the apps (grid_dims 30 / 40 / 50 / 75 cubed)
    assert scene_counts("surface_tension") == {"app": "surface_tension", "fluid": 27000, "boundary": 0}
    assert scene_counts("tilting_box") == {"app": "tilting_box", "fluid": 27000, "boundary": 0}
    # box walls: the box mesh voxel-sampled at 0.75 * 2r = 0.03 m
    r = float(np.float32(0.02))
    vox = float(np.float32(2.0) * np.float32(r) * np.float32(0.75))
    cube = scenes.procedural_mesh(abi.MESH_CUBE, (2.5, 1.0, 0.82), model=scenes.translation((0.0, 0.3, 0.0)))
    assert scene_counts("fountain") == {"app": "fountain", "fluid": 64000,
                                        "boundary": scenes.sample_mesh(cube, vox).shape[0]}
    cube = scenes.procedural_mesh(abi.MESH_CUBE, (2.5, 0.7, 0.7))
    assert scene_counts("dragon_collision")["boundary"] == scenes.sample_mesh(cube, vox).shape[0]
    assert scene_counts("vortex")["fluid"] == 125000 and scene_counts("BIG")["fluid"] == 421875


def test_dragon_mesh_from_obj(tmp_path):
    # a tetrahedron as the dragon: its sampled surface joins the box walls
    obj = tmp_path / "tet.obj"
    obj.write_text("v 0 0 0\nv 0.3 0 0\nv 0 0.3 0\nv 0 0 0.3\nf 1 3 2\nf 1 2 4\nf 1 4 3\nf 2/1 3/1 4/1\n")
    without = scene_counts("dragon_collision")["boundary"]
    rc, out, err = run("--app", "dragon_collision", "--mesh", str(obj), "--print-scene")
    assert rc == 0, err
    assert json.loads(out)["boundary"] > without
    rc, out, err = run("--app", "dragon_collision", "--mesh", str(tmp_path / "missing.obj"), "--print-scene")
    assert rc == 1 and "Cannot open mesh file" in err


def test_snapshot_checked_before_device_work(tmp_path):
    snap = os.path.join(GOLD, "snap_surface_settled_2048.sphs")
    rc, _, err = run("--headless", "--stop", "iters=1", "--app", "surface_tension", "--snapshot-load", snap)
    assert rc == 1 and "fluid count mismatch (file=2048, expected=27000)" in err
    rc, _, err = run("--headless", "--stop", "iters=1", "--app", "tilting_box", "--snapshot-load", snap)
    assert rc == 1 and "app_name mismatch (file='surface_tension', expected='tilting_box')" in err
    rc, _, err = run("--headless", "--stop", "iters=1", "--snapshot-load", str(tmp_path / "none.sphs"))
    assert rc == 1 and "Snapshot load failed: snapshot: cannot open" in err


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    rc, _, err = run("--headless", "--stop", "iters=1", "--app", "surface_tension")
    assert rc == 1 and "no HIP device" in err
