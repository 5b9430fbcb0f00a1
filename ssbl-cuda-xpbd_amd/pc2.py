"""Point Cache 2 (PC2) files: the on-disk output of a cloth session.

The layout follows the reference's cache writer (session_manager.py:4624-4641):
a 32-byte little-endian header ``<12siiffi`` = magic ``POINTCACHE2\\0``,
version 1, vertex count, start frame (float), sample rate 1.0 and sample
count, then one block of float32 object-local xyz triples per sample.  World
positions are brought to object space with the inverse world matrix in float32
(xpbd_core.py:652-655).

A cache being recorded is written to ``<path>.realtime.tmp`` with a sample
count of 0 in its header.  Committing rewrites the header with the real count
and moves the file into place (session_manager.py:4672-4728).
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = b"POINTCACHE2\0"
HEADER = struct.Struct("<12siiffi")


def header_bytes(vertex_count: int, start_frame: int, sample_count: int) -> bytes:
    return HEADER.pack(MAGIC, 1, int(vertex_count), float(start_frame), 1.0, int(sample_count))


def to_local(world: np.ndarray, matrix_world_inv: np.ndarray | None) -> np.ndarray:
    """World to object coordinates in float32 (xpbd_core.py:652-655); identity when no matrix."""
    world = np.asarray(world, dtype=np.float32).reshape(-1, 3)
    if matrix_world_inv is None:
        return world
    m = np.asarray(matrix_world_inv, dtype=np.float32)
    return world @ m[:3, :3].T + m[:3, 3]


class Pc2Writer:
    """Records samples of one cloth into ``path`` (via a temporary file)."""

    def __init__(self, path: str, vertex_count: int, start_frame: int, matrix_world_inv: np.ndarray | None = None):
        self.path = str(path)
        self.temp_path = f"{self.path}.realtime.tmp"
        self.vertex_count = int(vertex_count)
        self.start_frame = int(start_frame)
        self.matrix_world_inv = matrix_world_inv
        self.sample_count = 0
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        if os.path.exists(self.temp_path):
            os.remove(self.temp_path)
        self._handle = open(self.temp_path, "w+b")
        self._handle.write(header_bytes(self.vertex_count, self.start_frame, 0))

    def add_sample(self, world_positions: np.ndarray) -> None:
        local = to_local(world_positions, self.matrix_world_inv)
        if local.shape != (self.vertex_count, 3):
            raise ValueError(f"PC2 sample has {local.shape[0]} vertices, cache holds {self.vertex_count}")
        self._handle.write(np.ascontiguousarray(local, dtype="<f4").tobytes())
        self.sample_count += 1

    def commit(self) -> str:
        """Writes the final header and moves the cache into place; returns its path."""
        self._handle.seek(0)
        self._handle.write(header_bytes(self.vertex_count, self.start_frame, self.sample_count))
        self._handle.flush()
        self._handle.close()
        os.replace(self.temp_path, self.path)
        return self.path

    def discard(self) -> None:
        if not self._handle.closed:
            self._handle.close()
        if os.path.exists(self.temp_path):
            os.remove(self.temp_path)


def read_pc2(path: str) -> tuple[dict, np.ndarray]:
    """Header fields and the samples as float32 (samples, vertices, 3)."""
    with open(path, "rb") as f:
        raw = f.read()
    magic, version, count, start, rate, samples = HEADER.unpack_from(raw, 0)
    if magic != MAGIC:
        raise ValueError(f"{path} is not a PC2 file")
    data = np.frombuffer(raw, dtype="<f4", offset=HEADER.size)
    if data.size != samples * count * 3:
        raise ValueError(f"{path}: {data.size} floats for {samples} samples of {count} vertices")
    return ({"version": version, "vertex_count": count, "start_frame": start, "sample_rate": rate, "samples": samples},
            data.reshape(samples, count, 3).astype(np.float32))
