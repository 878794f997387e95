"""Neutron files (.mcv) as mcni.neutron_storage writes and reads them.

Layout (checked against the reference's fixtures
tests/components/sources/singletestneutron*.mcv, written by
tests/components/sources/create_neutron_storage_for_test.py:12-17): a
1100-byte header holding the ASCII tag "Neutron" zero-padded to 64 bytes, the
neutron count as a little-endian uint64 at byte 64, and zeros; then 10
little-endian float64 per neutron: x, y, z, vx, vy, vz, s1, s2, t, p.
"""
import os

import numpy as np

HEADER_SIZE = 1100
TAG = b"Neutron"
NDBLS = 10


def _header(count):
    h = bytearray(HEADER_SIZE)
    h[:len(TAG)] = TAG
    h[64:72] = np.uint64(count).astype("<u8").tobytes()
    return bytes(h)


def read(path):
    """(N, 10) float64 array of the neutrons in an .mcv file."""
    with open(path, "rb") as f:
        head = f.read(HEADER_SIZE)
        if len(head) != HEADER_SIZE or not head.startswith(TAG):
            raise ValueError("%s is not a neutron file" % path)
        count = int(np.frombuffer(head[64:72], "<u8")[0])
        body = f.read()
    if len(body) != count * NDBLS * 8:
        raise ValueError("%s: header says %d neutrons, body holds %d bytes" % (path, count, len(body)))
    return np.frombuffer(body, "<f8").reshape(count, NDBLS).astype(np.float64)


def write(path, neutrons):
    """Write an (N, 10) array as an .mcv file."""
    arr = np.ascontiguousarray(neutrons, dtype="<f8").reshape(-1, NDBLS)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "wb") as f:
        f.write(_header(arr.shape[0]))
        f.write(arr.tobytes())
    os.replace(tmp, path)


def append(path, neutrons):
    """Append neutrons to an .mcv file (created if missing), updating the count."""
    arr = np.ascontiguousarray(neutrons, dtype="<f8").reshape(-1, NDBLS)
    if not os.path.exists(path):
        write(path, arr)
        return
    with open(path, "r+b") as f:
        head = f.read(HEADER_SIZE)
        if not head.startswith(TAG):
            raise ValueError("%s is not a neutron file" % path)
        count = int(np.frombuffer(head[64:72], "<u8")[0])
        f.seek(0, os.SEEK_END)
        if f.tell() != HEADER_SIZE + count * NDBLS * 8:
            raise ValueError("%s: size does not match its header" % path)
        f.write(arr.tobytes())
        f.seek(64)
        f.write(np.uint64(count + arr.shape[0]).astype("<u8").tobytes())


class storage:
    """Minimal counterpart of mcni.neutron_storage.storage(path, mode)."""

    def __init__(self, path, mode="r"):
        if mode not in ("r", "w", "a"):
            raise ValueError("mode must be 'r', 'w' or 'a'")
        self.path, self.mode = path, mode
        if mode == "w":
            write(path, np.zeros((0, NDBLS)))

    def read(self, asnpyarr=True, wrap=False):
        return read(self.path)

    def write(self, neutrons):
        if self.mode == "r":
            raise IOError("storage opened read-only")
        append(self.path, neutrons)

    def close(self):
        pass
