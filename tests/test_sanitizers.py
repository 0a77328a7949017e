"""Address and undefined-behaviour sanitizer runs of the CPU code (no GPU).

- The oracle, built with -fsanitize=address,undefined and a driver that
  calls every entry point on small scenes of both presets, both precisions,
  serial and OpenMP (oracle/san_check.cpp).
- The host code: the headless driver with the host-only library sources
  (geometry, snapshots) instrumented, run through every mode that needs no
  device: usage, argument errors, option parsing, KTT config import, scene
  set-up of every app (procedural and OBJ meshes, voxel sampling, clustered
  blocks), snapshot header checks, and the no-device error.
Every sanitizer report aborts the run (-fno-sanitize-recover=all), and
LeakSanitizer is on, so a non-zero exit or a report in stderr fails.
"""
import fcntl
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
ORACLE_SAN = os.path.join(ROOT, "oracle", "_build", "san_check")
HOST_SAN = os.path.join(ROOT, "fluid-simulation_amd", "_build", "san", "wcsph_headless")
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:")


def _build(directory):
    # one build at a time: pytest-xdist workers may each reach this fixture,
    # and two makes writing the same binary would race
    d = os.path.join(ROOT, directory, "_build")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, ".san.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, directory), "san"], check=True, capture_output=True,
                       timeout=600)


@pytest.fixture(scope="module")
def host_san():
    _build("fluid-simulation_amd")
    return HOST_SAN


def _clean(p):
    for r in REPORTS:
        assert r not in p.stderr, p.stderr[-4000:]


def test_oracle_under_sanitizers():
    _build("oracle")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    p = subprocess.run([ORACLE_SAN], capture_output=True, text=True, timeout=600, env=env)
    _clean(p)
    assert p.returncode == 0 and "san_check ok" in p.stdout, p.stderr[-4000:]


def _host(exe, *args, rc=0, tmp=None):
    p = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, cwd=tmp)
    _clean(p)
    assert p.returncode == rc, (args, p.stderr[-2000:])
    return p


def test_host_cli_under_sanitizers(host_san):
    _host(host_san, "--help")
    for args in (["--bogus"], ["--headless"], ["--stop", "steps=3"], ["--stop", "iters=3x"], ["--stop", "iters"],
                 ["--searcher", "greedy"], ["--log-csv"], ["--app", "nowhere", "--print-scene"],
                 ["--frozen-config", "/nonexistent.json", "--print-options"]):
        _host(host_san, *args, rc=1)
    _host(host_san, "--headless", "--app", "fountain", "--stop", "sim-time=0.25", "--fixed-dt", "0.002",
          "--warmup-iters", "7", "--searcher", "full", "--tuning-budget", "0", "--snapshot-load", "a.sphs",
          "--snapshot-save", "b.sphs", "--ktt-output", "kt", "--log-csv", "log.csv", "--log-metrics", "--mesh",
          "d.obj", "--print-options")
    for state in ("dragon_falling", "fountain_active", "vortex_active"):
        _host(host_san, "--frozen-config", os.path.join(GOLD, f"ktt_{state}.json"), "--print-options")


@pytest.mark.parametrize("app", ["dam_break_32k", "dam_break_1m", "obstacle_collapse_8m", "surface_tension",
                                 "tilting_box", "fountain", "vortex", "dragon_collision", "BIG"])
def test_host_scenes_under_sanitizers(host_san, app):
    _host(host_san, "--app", app, "--print-scene")


def test_host_mesh_snapshot_and_device_errors_under_sanitizers(host_san, tmp_path):
    obj = tmp_path / "tet.obj"
    obj.write_text("v 0 0 0\nv 0.3 0 0\nv 0 0.3 0\nv 0 0 0.3\nf 1 3 2\nf 1 2 4\nf 1 4 3\nf 2/1 3/1 4/1\n")
    _host(host_san, "--app", "dragon_collision", "--mesh", str(obj), "--print-scene")
    _host(host_san, "--app", "dragon_collision", "--mesh", str(tmp_path / "missing.obj"), "--print-scene", rc=1)
    snap = os.path.join(GOLD, "snap_surface_settled_2048.sphs")
    for app in ("surface_tension", "tilting_box"):
        _host(host_san, "--headless", "--stop", "iters=1", "--app", app, "--snapshot-load", snap, rc=1)
    _host(host_san, "--headless", "--stop", "iters=1", "--snapshot-load", str(tmp_path / "none.sphs"), rc=1)
    import torch
    if not torch.cuda.is_available():
        p = _host(host_san, "--headless", "--stop", "iters=1", "--app", "surface_tension", rc=1)
        assert "no HIP device" in p.stderr
