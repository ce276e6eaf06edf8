"""Helpers shared by the GPU parity tests: list walking and set comparison."""
from __future__ import annotations

import numpy as np


def walk_lists(tiles: np.ndarray, seg: np.ndarray, capacity: int) -> dict:
    """(tile, bin) -> sorted list of packed (segment << 24 | strand) values."""
    t = tiles.reshape(-1, 4)
    data = seg[1:1 + 2 * capacity].reshape(-1, 2) if capacity else np.zeros((0, 2), np.uint32)
    out = {}
    for tb in np.nonzero(t[:, 2])[0]:
        ptr = int(t[tb, 2]) - 1
        vals = []
        guard = 0
        while ptr != 0xFFFFFFFF and ptr < capacity and guard <= capacity:
            vals.append(int(data[ptr, 0]))
            ptr = int(data[ptr, 1])
            guard += 1
        out[int(tb)] = sorted(vals)
    return out


def to_u8(img: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(np.clip(img, 0.0, 1.0) * 255.0), 0, 255).astype(np.int32)
