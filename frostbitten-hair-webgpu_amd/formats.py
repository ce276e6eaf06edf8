"""On-disk formats the chain consumes: TressFX .tfx hair and the SDF .bin.

.tfx  — tfxFileLoader.ts:1-69: f32 version, then u32 numHairStrands,
        numVerticesPerStrand, offsetVertexPosition, offsetStrandUV,
        offsetVertexUV, offsetStrandThickness, offsetVertexColor, 32 reserved
        u32 (160-byte header); float4 positions (.w = 1 movable, 0 root) at
        offsetVertexPosition.
SDF   — createSdfColliderFromBinary.ts:5-43: u32 dims at byte 0 (8-byte usize
        in the baker, main.rs:72-95), then 6 f32 bounds (min xyz, max xyz) at
        float index 2, then dims^3 f32 distances, x fastest.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

TFX_HEADER_BYTES = 160


@dataclass
class TfxFile:
    version: float
    num_strands: int
    points_per_strand: int
    positions: np.ndarray  # (S*P, 4) float32


def parse_tfx(data: bytes, scale: float = 1.0) -> TfxFile:
    """parseTfxFile — tfxFileLoader.ts:45-69"""
    version = struct.unpack_from("<f", data, 0)[0]
    u = struct.unpack_from("<7I", data, 4)
    n_strands, pps, off_pos = u[0], u[1], u[2]
    count = n_strands * pps * 4
    pos = np.frombuffer(data, dtype="<f4", count=count, offset=off_pos).astype(np.float32)
    pos = (pos * np.float32(scale)).reshape(-1, 4)
    return TfxFile(version, n_strands, pps, pos)


def write_tfx(positions: np.ndarray, num_strands: int, points_per_strand: int) -> bytes:
    pos = np.ascontiguousarray(positions, dtype="<f4").reshape(-1)
    assert pos.size == num_strands * points_per_strand * 4
    header = struct.pack("<f7I", 4.0, num_strands, points_per_strand, TFX_HEADER_BYTES, 0, 0, 0, 0)
    header += b"\0" * (TFX_HEADER_BYTES - len(header))
    return header + pos.tobytes()


@dataclass
class SdfFile:
    dims: int
    bounds_min: np.ndarray  # (3,) float32
    bounds_max: np.ndarray  # (3,) float32
    distances: np.ndarray   # (dims^3,) float32, x fastest


def parse_sdf(data: bytes) -> SdfFile:
    """createSdfColliderFromBinary — createSdfColliderFromBinary.ts:5-43"""
    dims = struct.unpack_from("<I", data, 0)[0]
    f = np.frombuffer(data, dtype="<f4", count=len(data) // 4)
    bmin = f[2:5].astype(np.float32)
    bmax = f[5:8].astype(np.float32)
    dist = f[8:].astype(np.float32)
    if dist.size != dims ** 3:
        raise ValueError(
            f"Invalid SDF binary file. With dims={dims} expected {dims ** 3} values. Got {dist.size}."
        )
    return SdfFile(dims, bmin, bmax, dist)


def write_sdf(dims: int, bounds_min, bounds_max, distances: np.ndarray) -> bytes:
    out = struct.pack("<Q", dims)
    out += np.asarray(list(bounds_min) + list(bounds_max), dtype="<f4").tobytes()
    d = np.ascontiguousarray(distances, dtype="<f4").reshape(-1)
    assert d.size == dims ** 3
    return out + d.tobytes()
