""".sphs snapshots through the library's reader/writer (host only).

Layout and checks follow the reference's Snapshot (src/cuda/SPH/snapshot.cuh:9-51):
76-byte header {"SPHS", version 1, app_name[64], fluid_n}, then float4
positions and float4 velocities of the fluid particles.
"""
import ctypes

import numpy as np

from . import abi


class SnapshotError(ValueError):
    pass


def _err(lib):
    return SnapshotError(lib.wcsph_snapshot_error().decode())


def info(path):
    lib = abi.load_library()
    name = ctypes.create_string_buffer(64)
    n = ctypes.c_uint32()
    if lib.wcsph_snapshot_info(str(path).encode(), name, 64, ctypes.byref(n)):
        raise _err(lib)
    return name.value.decode(), n.value


def load(path, app_name=None, fluid_n=None):
    """(app_name, positions (n, 4), velocities (n, 4)); checks app_name and
    fluid_n when given, like Snapshot::load."""
    lib = abi.load_library()
    name, n = info(path)
    if fluid_n is None:
        fluid_n = n
    pos = np.empty((fluid_n, 4), np.float32)
    vel = np.empty((fluid_n, 4), np.float32)
    rc = lib.wcsph_snapshot_read(str(path).encode(), None if app_name is None else app_name.encode(), fluid_n,
                                 abi.fptr(pos), abi.fptr(vel))
    if rc:
        raise _err(lib)
    return name, pos, vel


def save(path, app_name, pos, vel):
    lib = abi.load_library()
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    if pos.shape != vel.shape or pos.ndim != 2 or pos.shape[1] != 4:
        raise SnapshotError("snapshot: positions and velocities must both be (n, 4)")
    if lib.wcsph_snapshot_write(str(path).encode(), app_name.encode(), pos.shape[0], abi.fptr(pos), abi.fptr(vel)):
        raise _err(lib)
