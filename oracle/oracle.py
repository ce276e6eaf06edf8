"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


class RenderCfg(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "tileSize", "tileDepthBins", "slicesPerPixel", "avgFragmentsPerSlice", "processorCount",
        "invalidTilesPerSegmentThreshold", "sortBuckets", "sortBucketSize", "shadingPoints", "alphaQuadratic")]
    _fields_ = _fields_ + [("alphaMultiplier", C.c_float), ("reserved", C.c_uint32)]

    @classmethod
    def from_config(cls, rc):
        return cls(rc.tileSize, rc.tileDepthBins, rc.slicesPerPixel, rc.avgFragmentsPerSlice, rc.processorCount,
                   rc.invalidTilesPerSegmentThreshold, rc.sortBuckets, rc.sortBucketSize, rc.shadingPoints,
                   1 if rc.alphaQuadratic else 0, rc.alphaMultiplier, 0)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.or_sample_sdf.restype = C.c_float
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> None:
    """Host threads of the oracle's independent loops (strands, grid nodes,
    tiles); every result is the same for any count."""
    lib().or_set_threads(C.c_int(int(n)))


def get_threads() -> int:
    return int(lib().or_get_threads())


def sample_sdf(sdf, d, bmin, bmax, p) -> float:
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    return lib().or_sample_sdf(_p(f(sdf)), C.c_uint32(d), _p(f(bmin)), _p(f(bmax)), _p(f(p)))


def grid_pre_sim(d, sim_u, dv, sdf, sdf_dims):
    out = np.zeros(d ** 3 * 4, dtype=np.float32)
    lib().or_grid_pre_sim(C.c_uint32(d), _p(sim_u), _p(dv), _p(out), _p(sdf), C.c_uint32(sdf_dims))
    return out


def sim_integrate(d, sim_u, prev, now, seglen, sdf, sdf_dims, dv, dgw, initial, S, P):
    prev = np.ascontiguousarray(prev, dtype=np.float32).copy()
    lib().or_sim_integrate(C.c_uint32(d), _p(sim_u), _p(prev), _p(now), _p(seglen), _p(sdf), C.c_uint32(sdf_dims),
                           _p(dv), _p(dgw), _p(initial), C.c_uint32(S), C.c_uint32(P))
    return prev


def grid_post_sim(d, sim_u, prev, now, S, P):
    dv = np.zeros(d ** 3 * 4, dtype=np.int32)
    lib().or_grid_post_sim(C.c_uint32(d), _p(sim_u), _p(prev), _p(now), _p(dv), C.c_uint32(S), C.c_uint32(P))
    return dv


def debug_raster(render_u, positions, tangents, S, P, W, H):
    out = np.zeros(W * H, dtype=np.uint32)
    lib().or_debug_raster(_p(render_u), _p(positions), _p(tangents), C.c_uint32(S), C.c_uint32(P),
                          C.c_uint32(W), C.c_uint32(H), _p(out))
    return out


def tiles_draw(rc, W, H, render_u, hair_data, positions, tangents, depth, capacity, S, P):
    cfg = RenderCfg.from_config(rc)
    tiles_n = -(-W // rc.tileSize) * -(-H // rc.tileSize)
    tiles = np.zeros(tiles_n * rc.tileDepthBins * 4, dtype=np.uint32)
    seg = np.zeros(2 * (capacity + 1), dtype=np.uint32)
    counts = np.zeros(tiles_n, dtype=np.uint32)
    keys = np.zeros(max(1, capacity), dtype=np.uint32)
    lib().or_tiles_draw(C.byref(cfg), C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(hair_data), _p(positions),
                        _p(tangents), _p(tiles), _p(depth), _p(seg), C.c_uint32(capacity), _p(counts), _p(keys),
                        C.c_uint32(S), C.c_uint32(P))
    return tiles, seg, counts, keys


def tiles_sort(rc, W, H, counts):
    cfg = RenderCfg.from_config(rc)
    tiles_n = -(-W // rc.tileSize) * -(-H // rc.tileSize)
    tile_list = np.zeros(4 + tiles_n, dtype=np.uint32)
    buckets = np.zeros(rc.sortBuckets * 2, dtype=np.uint32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    lib().or_tiles_sort(C.byref(cfg), C.c_uint32(W), C.c_uint32(H), _p(counts), _p(tile_list), _p(buckets))
    return tile_list, buckets


def fine_raster(rc, W, H, render_u, positions, tiles, seg, capacity, depth, shading, tangents, tile_list, S, P,
                canonical=True, raster_in=None):
    cfg = RenderCfg.from_config(rc)
    raster = np.zeros(W * H * 4, dtype=np.float32) if raster_in is None else raster_in.copy()
    lib().or_fine_raster(C.byref(cfg), C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(positions), _p(tiles),
                         _p(seg), C.c_uint32(capacity), _p(raster), _p(depth), _p(shading), _p(tangents),
                         _p(tile_list), C.c_uint32(S), C.c_uint32(P), C.c_int(1 if canonical else 0))
    return raster


def shading(rc, render_u, positions, tangents, shadow, ao, depth, S, P):
    out = np.zeros(S * rc.shadingPoints * 4, dtype=np.float32)
    cfg = RenderCfg.from_config(rc)
    lib().or_shading(C.byref(cfg), _p(render_u), _p(positions), _p(tangents), _p(out), _p(shadow),
                     C.c_uint32(shadow.shape[1]), C.c_uint32(shadow.shape[0]), _p(ao), C.c_uint32(ao.shape[1]),
                     C.c_uint32(ao.shape[0]), _p(depth), C.c_uint32(depth.shape[1]), C.c_uint32(depth.shape[0]),
                     C.c_uint32(S), C.c_uint32(P))
    return out


def _setup_post():
    L = lib()
    if getattr(L, "_post_ready", False):
        return L
    L.or_f2h.restype = C.c_uint32
    L.or_f2h.argtypes = [C.c_float]
    L.or_h2f.restype = C.c_float
    L.or_h2f.argtypes = [C.c_uint32]
    L.or_bayer8.restype = C.c_uint32
    L.or_bayer8.argtypes = [C.c_uint32, C.c_uint32]
    L.or_aces.restype = C.c_float
    L.or_aces.argtypes = [C.c_float]
    L.or_srgb8.restype = C.c_uint32
    L.or_srgb8.argtypes = [C.c_float]
    L._post_ready = True
    return L


def f2h(x: float) -> int:
    return _setup_post().or_f2h(x)


def h2f(h: int) -> float:
    return _setup_post().or_h2f(h)


def bayer8(x: int, y: int) -> int:
    return _setup_post().or_bayer8(x, y)


def aces(x: float) -> float:
    return _setup_post().or_aces(x)


def srgb8(x: float) -> int:
    return _setup_post().or_srgb8(x)


def hair_combine(rc, W, H, render_u, tiles, seg, raster, counts, hdr):
    """Returns a blended copy of `hdr` (uint16 rgba16float bits, H*W*4)."""
    cfg = RenderCfg.from_config(rc)
    out = np.ascontiguousarray(hdr, dtype=np.uint16).copy()
    lib().or_hair_combine(C.byref(cfg), C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(tiles), _p(seg),
                          _p(raster), _p(counts), _p(out))
    return out


def present(W, H, render_u, hdr, depth, ao, normals=None, dither_per_pixel=False, pitch=0):
    """rgba8 sRGB screen bytes (H rows of `pitch` bytes, default 4*W)."""
    pitch = pitch or 4 * W
    out = np.zeros(H * pitch, dtype=np.uint8)
    nptr = _p(normals) if normals is not None else C.c_void_p(0)
    lib().or_present(C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(hdr), _p(depth), _p(ao),
                     C.c_uint32(ao.shape[1]), C.c_uint32(ao.shape[0]), nptr, C.c_int(1 if dither_per_pixel else 0),
                     _p(out), C.c_uint32(pitch))
    return out


def encode_oct(x: float, y: float, z: float) -> int:
    L = lib()
    L.or_encode_oct.restype = C.c_uint32
    L.or_encode_oct.argtypes = [C.c_float, C.c_float, C.c_float]
    return L.or_encode_oct(x, y, z)


def hw_hair(shadow, W, H, render_u, positions, tangents, depth, S, P, normals=None):
    """Returns (depth, normals) after drawing the hair into copies of the inputs
    (normals: H x W uint32 rg16float bits, or None for depth only)."""
    d = np.ascontiguousarray(depth, dtype=np.float32).copy()
    n = None if normals is None else np.ascontiguousarray(normals, dtype=np.uint32).copy()
    lib().or_hw_hair(C.c_int(1 if shadow else 0), C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(positions),
                     _p(tangents), _p(d), _p(n) if n is not None else C.c_void_p(0), C.c_uint32(S), C.c_uint32(P))
    return d, n


def calc_ao(render_u, depth, normals, ao_w, ao_h):
    out = np.zeros((ao_h, ao_w), dtype=np.float32)
    H, W = depth.shape
    lib().or_ao(_p(render_u), _p(np.ascontiguousarray(depth, dtype=np.float32)), C.c_uint32(W), C.c_uint32(H),
                _p(np.ascontiguousarray(normals, dtype=np.uint32)), _p(out), C.c_uint32(ao_w), C.c_uint32(ao_h))
    return out


def hw_coverage(shadow, W, H, render_u, positions, tangents, S, P):
    counts = np.zeros((H, W), dtype=np.uint32)
    lib().or_hw_coverage(C.c_int(1 if shadow else 0), C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(positions),
                         _p(tangents), _p(counts), C.c_uint32(S), C.c_uint32(P))
    return counts


class Mesh(C.Structure):
    """fh_mesh with host pointers."""
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
                ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
                ("is_collider_preview", C.c_uint32), ("reserved", C.c_uint32)]


def _mesh_table(meshes):
    """ctypes fh_mesh array for a list of (positions (V,3) f32, normals (V,3)
    f32, indices (T,3) u32, preview) and the arrays it points into."""
    keep, rows = [], []
    for m in meshes:
        pos = np.ascontiguousarray(m["positions"], dtype=np.float32)
        nrm = np.ascontiguousarray(m["normals"], dtype=np.float32)
        idx = np.ascontiguousarray(m["indices"], dtype=np.uint32)
        keep += [pos, nrm, idx]
        rows.append(Mesh(pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data, pos.shape[0], idx.shape[0],
                         1 if m.get("preview") else 0, 0))
    arr = (Mesh * max(1, len(rows)))(*rows)
    return arr, keep


def background(W, H, render_u):
    """rgba16float bits (H x W x 4 uint16) of the background gradient pass."""
    out = np.zeros((H, W, 4), dtype=np.uint16)
    lib().or_background(C.c_uint32(W), C.c_uint32(H), _p(render_u), _p(out))
    return out


def meshes(shadow, W, H, render_u, mesh_list, shadow_map=None, ao=None, hdr=None):
    """The mesh passes: shadow=True returns the meshes' shadow depth (H x W,
    W = H = the map size); shadow=False returns (depth, normals, hdr) of the
    camera pass drawn over a copy of hdr."""
    arr, keep = _mesh_table(mesh_list)
    depth = np.zeros((H, W), dtype=np.float32)
    if shadow:
        lib().or_meshes(C.c_int(1), C.c_uint32(W), C.c_uint32(H), _p(render_u), arr, C.c_uint32(len(mesh_list)),
                        C.c_void_p(0), C.c_uint32(0), C.c_void_p(0), C.c_uint32(0), C.c_uint32(0), _p(depth),
                        C.c_void_p(0), C.c_void_p(0))
        return depth
    sm = np.ascontiguousarray(shadow_map, dtype=np.float32)
    aot = np.ascontiguousarray(ao, dtype=np.float32)
    normals = np.zeros((H, W), dtype=np.uint32)
    out = np.ascontiguousarray(hdr, dtype=np.uint16).copy()
    lib().or_meshes(C.c_int(0), C.c_uint32(W), C.c_uint32(H), _p(render_u), arr, C.c_uint32(len(mesh_list)),
                    _p(sm), C.c_uint32(sm.shape[0]), _p(aot), C.c_uint32(aot.shape[1]), C.c_uint32(aot.shape[0]),
                    _p(depth), _p(normals), _p(out))
    return depth, normals, out


def bg_hash(x: float, y: float) -> float:
    L = lib()
    L.or_bg_hash.restype = C.c_float
    L.or_bg_hash.argtypes = [C.c_float, C.c_float]
    return L.or_bg_hash(x, y)


def bg_noise(x: float, y: float) -> float:
    L = lib()
    L.or_bg_noise.restype = C.c_float
    L.or_bg_noise.argtypes = [C.c_float, C.c_float]
    return L.or_bg_noise(x, y)


def bg_fractal(x: float, y: float, scale: float) -> float:
    L = lib()
    L.or_bg_fractal.restype = C.c_float
    L.or_bg_fractal.argtypes = [C.c_float, C.c_float, C.c_float]
    return L.or_bg_fractal(x, y, scale)


def disney_pbr(positionWS, normal, toEye, albedo, roughness, isMetallic, light_pos, light_color_energy):
    """disneyPBR of one light (pbr.wgsl.ts:140-160)."""

    class Light(C.Structure):
        _fields_ = [("position", C.c_float * 4), ("colorAndEnergy", C.c_float * 4)]

    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731
    light = Light((C.c_float * 4)(*[float(x) for x in light_pos]),
                  (C.c_float * 4)(*[float(x) for x in light_color_energy]))
    out = (C.c_float * 3)()
    lib().or_disney_pbr(f3(positionWS), f3(normal), f3(toEye), f3(albedo), C.c_float(roughness),
                        C.c_float(isMetallic), C.byref(light), out)
    return np.array(list(out), dtype=np.float32)


def directional_shadow(render_u, shadow_map, posWS, normalWS, pss) -> float:
    L = lib()
    L.or_directional_shadow.restype = C.c_float
    sm = np.ascontiguousarray(shadow_map, dtype=np.float32)
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731
    return L.or_directional_shadow(_p(render_u), _p(sm), C.c_uint32(sm.shape[0]), f3(posWS), f3(normalWS), f3(pss))
