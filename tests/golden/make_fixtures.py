"""Writes the test fixtures derived from the reference repository's data
files (run in a container that has the reference checkout; the fixtures are
committed, the reference is not needed afterwards).

* snap_<state>_2048.sphs: the first 2,048 fluid particles of
  measurements/snapshots/<state>.sphs, in the same .sphs layout (76-byte
  header, float4 positions, float4 velocities).  The reference saved its
  state in Morton-sorted order, so the first particles form one compact
  block of the flow.
* ref_states_positions.npz: the positions (x, y, z; float32) of every
  fluid particle of each saved state, in the saved order, for the
  full-size parity tests (w is 1 in every file and is not kept).
* ktt_<state>.json: measurements/configs/<state>.json unchanged (KTT frozen
  configurations of the reference's tuned kernels).
* manifest.json: header fields and coordinate ranges of the sources.

The .sphs files are parsed as plain binary data with numpy.
"""
import json
import os
import shutil
import struct
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
STATES = ["dragon_falling", "dragon_settled", "surface_settled", "tilting_chaotic", "fountain_active", "vortex_active"]
KEEP = 2048


def read_sphs(path):
    raw = open(path, "rb").read()
    magic, version, name, n = struct.unpack("<4sI64sI", raw[:76])
    assert magic == b"SPHS" and version == 1
    body = np.frombuffer(raw, dtype=np.float32, offset=76)
    pos = body[: 4 * n].reshape(n, 4)
    vel = body[4 * n: 8 * n].reshape(n, 4)
    return name.split(b"\0")[0].decode(), pos, vel


def write_sphs(path, name, pos, vel):
    hdr = struct.pack("<4sI64sI", b"SPHS", 1, name.encode().ljust(64, b"\0"), pos.shape[0])
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(pos, np.float32).tobytes())
        f.write(np.ascontiguousarray(vel, np.float32).tobytes())


def main():
    manifest = {}
    full = {}
    for st in STATES:
        name, pos, vel = read_sphs(os.path.join(REF, "measurements", "snapshots", st + ".sphs"))
        assert np.all(pos[:, 3] == 1.0)
        full[st] = np.ascontiguousarray(pos[:, :3])
        write_sphs(os.path.join(OUT, f"snap_{st}_{KEEP}.sphs"), name, pos[:KEEP], vel[:KEEP])
        manifest[st] = {"app_name": name, "fluid_n": int(pos.shape[0]), "kept": KEEP,
                        "pos_min": pos[:, :3].min(axis=0).tolist(), "pos_max": pos[:, :3].max(axis=0).tolist(),
                        "vmax": float(np.sqrt((vel[:, :3] ** 2).sum(axis=1)).max())}
        shutil.copy(os.path.join(REF, "measurements", "configs", st + ".json"), os.path.join(OUT, f"ktt_{st}.json"))
    np.savez_compressed(os.path.join(OUT, "ref_states_positions.npz"), **full)
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
