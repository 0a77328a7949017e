"""The .sphs reader/writer and the reference's own data files (snapshots and
KTT frozen configurations, as fixtures under tests/golden/)."""
import json
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle import brute_neighbors, brute_order
from wcsph import abi, autotune, snapshot

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATES = ["dragon_falling", "dragon_settled", "surface_settled", "tilting_chaotic", "fountain_active", "vortex_active"]
APPS = {"dragon_falling": "dragon_collision", "dragon_settled": "dragon_collision", "surface_settled": "surface_tension",
        "tilting_chaotic": "tilting_box", "fountain_active": "fountain", "vortex_active": "vortex"}
H_REF = np.float32(4.0) * np.float32(0.02)  # reference support radius = cell size (common.cuh:18-37)


def fixture(state):
    return os.path.join(GOLD, f"snap_{state}_2048.sphs")


@pytest.mark.parametrize("state", STATES)
def test_reference_snapshot_fixtures_load(state):
    name, n = snapshot.info(fixture(state))
    assert name == APPS[state] and n == 2048
    name, pos, vel = snapshot.load(fixture(state), app_name=APPS[state], fluid_n=2048)
    assert np.all(pos[:, 3] == 1.0) and np.all(vel[:, 3] == 0.0)
    man = json.load(open(os.path.join(GOLD, "manifest.json")))[state]
    assert np.all(pos[:, :3] >= np.array(man["pos_min"]) - 1e-6)
    assert np.all(pos[:, :3] <= np.array(man["pos_max"]) + 1e-6)
    assert np.any(pos[:, :3] < 0)  # negative cells exercise the Morton key offset


def test_snapshot_roundtrip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(37, 4)).astype(np.float32)
    vel = rng.normal(size=(37, 4)).astype(np.float32)
    p = tmp_path / "a.sphs"
    snapshot.save(p, "dam_break", pos, vel)
    assert os.path.getsize(p) == 76 + 2 * 37 * 16
    raw = open(p, "rb").read()
    assert struct.unpack("<4sI", raw[:8]) == (b"SPHS", 1)
    name, p2, v2 = snapshot.load(p, app_name="dam_break", fluid_n=37)
    np.testing.assert_array_equal(p2, pos)
    np.testing.assert_array_equal(v2, vel)
    with pytest.raises(snapshot.SnapshotError, match="app_name mismatch"):
        snapshot.load(p, app_name="other", fluid_n=37)
    with pytest.raises(snapshot.SnapshotError, match="fluid count mismatch"):
        snapshot.load(p, app_name="dam_break", fluid_n=36)
    with pytest.raises(snapshot.SnapshotError, match="app_name too long"):
        snapshot.save(tmp_path / "b.sphs", "x" * 64, pos, vel)
    with pytest.raises(snapshot.SnapshotError, match="cannot open"):
        snapshot.info(tmp_path / "missing.sphs")
    bad = bytearray(raw)
    bad[0:4] = b"XXXX"
    (tmp_path / "c.sphs").write_bytes(bytes(bad))
    with pytest.raises(snapshot.SnapshotError, match="bad magic"):
        snapshot.info(tmp_path / "c.sphs")
    bad = bytearray(raw)
    bad[4:8] = struct.pack("<I", 2)
    (tmp_path / "d.sphs").write_bytes(bytes(bad))
    with pytest.raises(snapshot.SnapshotError, match="version mismatch"):
        snapshot.info(tmp_path / "d.sphs")
    (tmp_path / "e.sphs").write_bytes(raw[: 76 + 100])
    with pytest.raises(snapshot.SnapshotError, match="body read failed"):
        snapshot.load(tmp_path / "e.sphs")
    # a corrupt count is refused by info(), before anyone sizes buffers by it
    bad = bytearray(raw)
    bad[72:76] = struct.pack("<I", 0xFFFFFFF0)
    (tmp_path / "f.sphs").write_bytes(bytes(bad))
    with pytest.raises(snapshot.SnapshotError, match="body read failed.*4294967280 particles need"):
        snapshot.info(tmp_path / "f.sphs")


@pytest.mark.parametrize("state", STATES)
def test_oracle_order_and_neighbors_on_reference_states(state):
    _, pos, _ = snapshot.load(fixture(state))
    np.testing.assert_array_equal(O.sort_order(pos, H_REF), brute_order(pos, H_REF))
    sp = pos[O.sort_order(pos, H_REF)]
    sub = sp[:600]  # brute force on a subset of queries against all sources
    counts, lst = O.neighbors(sp, sp, H_REF, H_REF, True)
    ref = brute_neighbors(sp, sp, H_REF, True)
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    for i in range(len(sub)):
        assert set(lst[off[i]:off[i + 1]].tolist()) == ref[i]
    assert counts.mean() > 10


@pytest.mark.parametrize("state", STATES)
def test_ktt_configs_map_to_variants(state):
    cfg = json.load(open(os.path.join(GOLD, f"ktt_{state}.json")))
    pairs = {p["Name"]: p["Value"] for p in cfg}
    assert pairs["compute_rho_p_nlist"] == 1 and pairs["compute_rho_p_shared"] == 0
    base = abi.Variant(256, 128, 64, 1, abi.NEIGHBORS_CELL, 16384, 1.0)
    ignored = []
    v = autotune.from_ktt_config(cfg, base, ignored)
    bs = pairs["compute_pressure_accel_n_normal_block"]
    assert v.block_size == max(64, ((bs + 63) // 64) * 64)
    assert v.fluid_capacity == 128
    # CELL_SIZE_MULT (2/3 in most of the shipped states) as the fp32 multiplier,
    # with the per-particle builder, the one form that takes it
    mult = float(pairs.get("CELL_SIZE_MULT", 1.0))
    assert v.cell_size_mult == np.float32(mult)
    assert v.neighbor_mode == (abi.NEIGHBORS_PARTICLE if mult != 1.0 else abi.NEIGHBORS_CELL)
    # every parameter is mapped or named with its reason
    names = {x.split(":")[0] for x in ignored}
    mapped = {"CELL_SIZE_MULT", "compute_pressure_accel_n_normal_block", "compute_rho_p_nlist",
              "compute_pressure_accel_n_normal_nlist", "compute_non_pressure_accel_nlist"}
    assert names | (mapped & set(pairs)) == set(pairs) and not names & mapped
    assert all("no counterpart" not in x for x in ignored), ignored


def test_ktt_configs_hold_both_cell_sizes():
    mults = set()
    for state in STATES:
        pairs = {p["Name"]: p["Value"] for p in json.load(open(os.path.join(GOLD, f"ktt_{state}.json")))}
        mults.add(pairs.get("CELL_SIZE_MULT", 1.0))
    assert {1.0, 0.6666666666666666} <= mults


def test_autotune_buckets(tmp_path):
    assert autotune.bucket(1048576) == "s1_n2^20"
    assert autotune.bucket(9907267, 256) == "s256_n2^23"
    assert autotune.load_cache(tmp_path / "none.json") == {}


def test_scene_buckets():
    idx, groups = autotune.scene_buckets([2000, 16383, 16384, 65535, 65536, 200000, 5000])
    assert idx == [0, 0, 1, 1, 2, 2, 0]
    assert groups == {0: [0, 1, 6], 1: [2, 3], 2: [4, 5]}
    assert autotune.scene_buckets([10, 20], edges=())[1] == {0: [0, 1]}
