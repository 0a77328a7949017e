"""The headless driver on the GPU: CSV log, snapshots (bit-identical to the
Python host stepping the same scene), a moving box, KTT frozen configs and
the one-shot autotuner."""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import wcsph
from wcsph import scenes, snapshot

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fluid-simulation_amd", "wcsph_headless")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args):
    print("wcsph_headless", *args, flush=True)
    p = subprocess.run([EXE, "--headless", *args], capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_csv_log_with_metrics(tmp_path):
    log = tmp_path / "log.csv"
    out = run("--app", "surface_tension", "--stop", "iters=20", "--log-csv", str(log), "--log-metrics",
              "--tuning-budget", "0")
    assert out["steps"] == 20 and out["fluid"] == 27000
    rows = list(csv.DictReader(open(log)))
    # the reference's columns (application.cu:114-116) plus the dt used
    assert list(rows[0].keys()) == ["step", "sim_time", "wall_dt_ms", "scheduled_tune", "mean_speed", "ke", "dt"]
    assert len(rows) == 20
    # sim_time is the sum of the dt the solver used, not 20 x --fixed-dt
    # (the reference's application.cu:124-125 adds --fixed-dt): the
    # reference preset's CFL rule keeps dt <= dt_max = 5e-4 < 0.01
    total = 0.0
    for k, r in enumerate(rows):
        assert int(r["step"]) == k
        dt = float(r["dt"])
        assert 0 < dt <= float(np.float32(5e-4)) * (1 + 1e-6)
        total += dt
        assert float(r["sim_time"]) == pytest.approx(total, rel=1e-6)
        assert float(r["wall_dt_ms"]) > 0 and r["scheduled_tune"] == "0"
        assert np.isfinite(float(r["ke"])) and float(r["mean_speed"]) >= 0
    assert out["sim_time"] == pytest.approx(total, rel=1e-6)


def test_stop_at_physical_sim_time():
    """--stop sim-time=T stops at the first step whose summed dt reaches T:
    at rest the CFL rule gives dt_max = 5e-4, so T = 0.005 takes 10 steps,
    not the 1 step that --fixed-dt 0.01 per step would count."""
    out = run("--app", "surface_tension", "--stop", "sim-time=0.005", "--tuning-budget", "0")
    assert out["sim_time"] >= 0.005 - 1e-9
    assert 9 <= out["steps"] <= 11
    assert out["sim_time"] - float(np.float32(5e-4)) < 0.005


def test_snapshot_matches_python_host(tmp_path):
    a = tmp_path / "a.sphs"
    run("--app", "dam_break_32k", "--fixed-dt", "1e-4", "--stop", "iters=10", "--tuning-budget", "0",
        "--snapshot-save", str(a))
    name, pos, vel = snapshot.load(a)
    assert name == "dam_break_32k"
    fl, bd, box = scenes.dam_break(scenes.DAM_BREAK_32K)
    with wcsph.FluidSimulator(scenes.preset_params(wcsph.abi.PRESET_BENCHMARK), fl, bd, box) as sim:
        sim.step(1e-4, 10)
        st = sim.get_state()
    np.testing.assert_array_equal(pos, st["pos"])
    np.testing.assert_array_equal(vel, st["vel"])
    # load + save without stepping reproduces the file
    b = tmp_path / "b.sphs"
    out = run("--app", "dam_break_32k", "--fixed-dt", "1e-4", "--stop", "iters=0", "--tuning-budget", "0",
              "--snapshot-load", str(a), "--snapshot-save", str(b))
    assert out["steps"] == 0
    assert open(a, "rb").read() == open(b, "rb").read()


def test_tilting_box_and_fountain_run():
    # sim-time counts the dt used: the reference preset's CFL dt is at most
    # dt_max = 5e-4, so 0.03 s takes at least 60 steps (the reference's
    # nominal count of --fixed-dt 0.01 would stop after 3)
    out = run("--app", "tilting_box", "--stop", "sim-time=0.03", "--tuning-budget", "0")
    assert out["sim_time"] >= 0.03 - 1e-9 and out["sim_time"] - float(np.float32(5e-4)) < 0.03
    assert out["steps"] >= 60
    out = run("--app", "fountain", "--stop", "iters=10", "--warmup-iters", "2", "--tuning-budget", "0")
    assert out["steps"] == 10 and out["fluid"] == 64000


def test_frozen_config_and_autotuner(tmp_path):
    out = run("--app", "dragon_collision", "--stop", "iters=5",
              "--frozen-config", os.path.join(GOLD, "ktt_dragon_falling.json"))
    assert out["steps"] == 5
    prefix = tmp_path / "tune"
    out = run("--app", "surface_tension", "--stop", "iters=3", "--searcher", "full", "--tuning-budget", "0.1",
              "--ktt-output", str(prefix))
    res = json.load(open(str(prefix) + ".json"))
    assert len(res) == 9 and all(r["ms_per_step"] > 0 for r in res)
    assert {r["neighbor_mode"] for r in res} == {0, 1, 2, 3}
    # the per-particle builder also at the reference's finer cell sizes
    assert sorted(r["cell_size_mult"] for r in res if r["cell_size_mult"] != 1) == pytest.approx([0.5, 2 / 3])
    assert all(r["neighbor_mode"] == 1 for r in res if r["cell_size_mult"] != 1)
