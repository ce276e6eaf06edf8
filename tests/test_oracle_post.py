"""The oracle's combine and present passes against closed forms and the
reference's own table.

hairCombinePass.wgsl.ts:51-162 with the blend state of hairCombinePass.ts
:204-217; presentPass.wgsl.ts:58-103 with dither.wgsl.ts, aces.wgsl.ts and
the rgba8unorm-srgb store. Each case is chosen so that a plausible slip (a
missing y flip, a swapped blend factor, a transposed dither index, a wrong
transfer-function branch) fails it.
"""
import math

import numpy as np
import pytest

# DITHER_MATRIX of dither.wgsl.ts:12-21 (row-major 8x8), copied as data
REFERENCE_BAYER8 = [
    0, 32, 8, 40, 2, 34, 10, 42, 48, 16, 56, 24, 50, 18, 58, 26, 12, 44, 4, 36, 14, 46, 6, 38,
    60, 28, 52, 20, 62, 30, 54, 22, 3, 35, 11, 43, 1, 33, 9, 41, 51, 19, 59, 27, 49, 17, 57, 25,
    15, 47, 7, 39, 13, 45, 5, 37, 63, 31, 55, 23, 61, 29, 53, 21]


def _u(**kw):
    from frostbitten_hair_webgpu_amd import layouts as L
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    for k, v in kw.items():
        u[k] = v
    return u


def test_half_conversions_match_numpy(oracle):
    """binary16 <-> binary32 (round to nearest even) over every half value and
    over boundary and random floats."""
    allh = np.arange(65536, dtype=np.uint32)
    ref = allh.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([oracle.h2f(int(h)) for h in allh], dtype=np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    rng = np.random.default_rng(5)
    xs = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(100.0),
        rng.uniform(-1e-4, 1e-4, 20000).astype(np.float32),          # subnormal halves
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e5, -1e5, 2.0 ** -24, 2.0 ** -25,
                  3 * 2.0 ** -26, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32),
    ])
    got = np.array([oracle.f2h(float(x)) for x in xs], dtype=np.uint32)
    with np.errstate(over="ignore"):
        exp = xs.astype(np.float16).view(np.uint16).astype(np.uint32)
    assert np.array_equal(got, exp)


def test_bayer_construction_matches_reference_table(oracle):
    got = [oracle.bayer8(i % 8, i // 8) for i in range(64)]
    assert got == REFERENCE_BAYER8


def test_aces_closed_forms(oracle):
    assert oracle.aces(0.0) == 0.0
    assert math.isclose(oracle.aces(1.0), (2.51 + 0.03) / (2.43 + 0.59 + 0.14), rel_tol=1e-6)
    assert oracle.aces(1e4) == 1.0          # a / c > 1, saturated
    assert oracle.aces(0.5) < oracle.aces(0.6) < oracle.aces(0.7)


def test_srgb_store_round_trips_every_code(oracle):
    for k in range(256):
        s = k / 255.0
        lin = s / 12.92 if s <= 0.04045 else ((s + 0.055) / 1.055) ** 2.4
        assert oracle.srgb8(lin) == k, k
    assert oracle.srgb8(-1.0) == 0 and oracle.srgb8(2.0) == 255 and oracle.srgb8(float("nan")) == 0
    assert oracle.srgb8(0.5) == 188  # 1.055 * 0.5^(1/2.4) - 0.055 = 0.7354 -> 187.5 -> 188


class _Tiny:
    tileSize = 4
    tileDepthBins = 2
    slicesPerPixel = 4
    avgFragmentsPerSlice = 4
    processorCount = 4
    invalidTilesPerSegmentThreshold = 64
    sortBuckets = 4
    sortBucketSize = 4
    shadingPoints = 4
    alphaQuadratic = False
    alphaMultiplier = 1.1


def _combine_inputs(W=8, H=6):
    raster = np.zeros(W * H * 4, np.float32)     # 16 B header, then W*H - 1 pixels
    pix = raster[4:].reshape(-1, 4)

    def put(x, py, c):
        pix[py * W + x] = c

    put(1, 0, [0.25, 0.5, 0.75, 1.0])           # opaque, bottom row of the rasterizer
    put(2, 0, [1.0, 0.0, 0.5, 0.5])             # half transparent
    put(3, 5, [9.0, 9.0, 9.0, 0.0])             # alpha 0: leaves the scene as it is
    hdr = np.zeros((H, W, 4), np.float16)
    hdr[..., 0], hdr[..., 1], hdr[..., 2], hdr[..., 3] = 0.25, 0.5, 0.125, 1.0
    return raster, hdr.view(np.uint16).copy()


def test_combine_blend_and_y_flip(oracle):
    W, H = 8, 6
    raster, hdr = _combine_inputs(W, H)
    tiles = np.zeros(2 * 2 * 2 * 4, np.uint32)
    seg = np.zeros(2 * 5, np.uint32)
    counts = np.zeros(4, np.uint32)
    u = _u(viewport=[W, H, 0, 0], displayMode=0)
    out = oracle.hair_combine(_Tiny, W, H, u, tiles, seg, raster, counts, hdr).view(np.float16).astype(np.float32)
    base = hdr.view(np.float16).astype(np.float32)
    # rasterizer row 0 lands on texture row H - 1
    assert out[H - 1, 1].tolist() == [0.25, 0.5, 0.75, 2.0]
    assert out[H - 1, 2].tolist() == [1.0 * 0.5 + 0.25 * 0.5, 0.0 + 0.5 * 0.5, 0.25 + 0.0625, 1.5]
    # alpha 0: colour unchanged, alpha + 0
    assert np.array_equal(out[0, 3], base[0, 3])
    # the rest (including the last pixel, past the end of the array) unchanged
    mask = np.ones((H, W), bool)
    mask[H - 1, 1] = mask[H - 1, 2] = False
    assert np.array_equal(out[mask], base[mask])


def test_combine_debug_tiles(oracle):
    """dbgShowTiles: pixels without hair get the tile checker colour; TILES
    and TILES_PPLL draw per-tile segment counts and discard empty tiles."""
    from frostbitten_hair_webgpu_amd import layouts as L
    W, H = 8, 6
    raster, hdr = _combine_inputs(W, H)
    base = hdr.view(np.float16).astype(np.float32)
    tiles = np.zeros(2 * 2 * 2 * 4, np.uint32).reshape(-1, 4)
    counts = np.array([3, 0, 0, 7], np.uint32)
    u = _u(viewport=[W, H, 0, 0], displayMode=0 | (1 << 8))
    out = oracle.hair_combine(_Tiny, W, H, u, tiles.reshape(-1), np.zeros(10, np.uint32), raster, counts, hdr)
    out = out.view(np.float16).astype(np.float32)
    for ty in range(H):
        py = H - 1 - ty
        for x in range(W):
            if (x, py) in ((1, 0), (2, 0)):
                continue
            exp = [0.0, (x // 4 % 2) / 2, (py // 4 % 2) / 2]
            assert out[ty, x, :3].tolist() == exp, (x, ty)
            assert out[ty, x, 3] == 2.0
    # TILES: counts / 370 per tile, tiles with 0 segments discarded
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_TILES | (370 << 8))
    out = oracle.hair_combine(_Tiny, W, H, u, tiles.reshape(-1), np.zeros(10, np.uint32), raster, counts, hdr)
    out = out.view(np.float16).astype(np.float32)
    r3 = np.float32(3) / np.float32(370)
    assert out[H - 1, 0, 0] == np.float16(r3) and out[H - 1, 0, 1] == np.float16(np.float32(1) - r3)
    assert np.array_equal(out[H - 1, 5], base[H - 1, 5])            # tile (1, 0): 0 segments
    assert out[0, 5, 0] == np.float16(np.float32(7) / np.float32(370))  # tile (1, 1), rasterizer row 5
    # TILES_PPLL: tile 0 has 3 elements in bin 0 and 2 in bin 1; the walk stops at the maximum (4)
    seg = np.zeros(2 * (1 + 6), np.uint32)
    el = seg[1:1 + 2 * 6].reshape(-1, 2)
    el[:, 1] = 0xFFFFFFFF
    el[0, 1], el[1, 1] = 1, 2      # bin 0: 0 -> 1 -> 2
    el[3, 1] = 4                   # bin 1: 3 -> 4
    tiles[0, 2], tiles[1, 2] = 0 + 1, 3 + 1
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_TILES_PPLL | (4 << 8), maxDrawnHairSegments=7)
    out = oracle.hair_combine(_Tiny, W, H, u, tiles.reshape(-1), seg, raster, counts, hdr)
    out = out.view(np.float16).astype(np.float32)
    assert out[H - 1, 0, :2].tolist() == [1.0, 0.0]
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_TILES_PPLL | (370 << 8), maxDrawnHairSegments=7)
    out = oracle.hair_combine(_Tiny, W, H, u, tiles.reshape(-1), seg, raster, counts, hdr)
    out = out.view(np.float16).astype(np.float32)
    assert out[H - 1, 0, 0] == np.float16(np.float32(5) / np.float32(370))
    assert np.array_equal(out[0, 5], base[0, 5])                    # other tiles: empty lists


def _present_inputs(W=16, H=8):
    rng = np.random.default_rng(3)
    hdr = np.zeros((H, W, 4), np.float16)
    hdr[..., :3] = rng.uniform(0, 3, (H, W, 3))
    hdr[..., 3] = 1
    depth = rng.uniform(0.9, 1.0, (H, W)).astype(np.float32)
    ao = rng.uniform(0, 1, (H // 2, W // 2)).astype(np.float32)
    return hdr.view(np.uint16).copy(), depth, ao


def test_present_tone_map(oracle):
    W, H = 16, 8
    hdr, depth, ao = _present_inputs(W, H)
    c = hdr.view(np.float16).astype(np.float32)
    for gamma, exposure in ((1.0, 1.0), (2.2, 0.85)):
        u = _u(viewport=[W, H, 0, 0], colorMgmt=[gamma, exposure, 1.0, 0], displayMode=0)
        out = oracle.present(W, H, u, hdr, depth, ao).reshape(H, W, 4)
        assert (out[..., 3] == 255).all()
        for (y, x, k) in ((0, 0, 0), (3, 7, 1), (7, 15, 2), (5, 2, 0)):
            # Deno reads Bayer element 0 (= 0): no dither at all
            v = oracle.aces(float(np.float32(c[y, x, k]) * np.float32(exposure)))
            exp = oracle.srgb8(float(np.float32(v) ** np.float32(1.0 / gamma)))
            assert abs(int(out[y, x, k]) - exp) <= 1, (gamma, y, x, k)
    # per-pixel dither on a black scene: the code grows with the Bayer value
    black = np.zeros_like(hdr)
    u = _u(viewport=[W, H, 0, 0], colorMgmt=[1.0, 1.0, 256.0, 0], displayMode=0)
    out = oracle.present(W, H, u, black, depth, ao, dither_per_pixel=True).reshape(H, W, 4)
    for y in range(H):
        for x in range(W):
            m = REFERENCE_BAYER8[(y % 8) * 8 + x % 8]
            assert out[y, x, 0] == oracle.srgb8(oracle.aces(m / 63.0 * 256.0 / 256.0)), (x, y)
    assert out[0, 0, 0] == 0 and out[7, 0, 0] == out[..., 0].max()  # element 0 is 0; element 63 at (0, 7)
    out0 = oracle.present(W, H, u, black, depth, ao, dither_per_pixel=False).reshape(H, W, 4)
    assert (out0[..., :3] == 0).all()


def test_present_debug_views_and_pitch(oracle):
    from frostbitten_hair_webgpu_amd import layouts as L
    W, H = 16, 8
    hdr, depth, ao = _present_inputs(W, H)
    # DEPTH: linearizeDepth_0_1 mapped from [0.002, 0.01]; depth 0 -> near plane -> black, 1 -> far -> white
    n, f = 0.01, 100.0
    d = depth.copy()
    d[0, 0], d[0, 1] = 0.0, 1.0
    lin_to_d = lambda L01: (n * f / (L01 * (f - n)) - f) / (n - f)
    d[0, 2] = lin_to_d(0.006)   # the middle of the mapped range
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_DEPTH)
    out = oracle.present(W, H, u, hdr, d, ao).reshape(H, W, 4)
    assert out[0, 0, :3].tolist() == [0, 0, 0] and out[0, 1, :3].tolist() == [255, 255, 255]
    assert abs(int(out[0, 2, 0]) - oracle.srgb8(0.5)) <= 1
    # NORMALS: octahedral codes of +z, -z, +x, -y decode to the axes
    nrm = np.zeros((H, W, 2), np.float16)
    nrm[..., :] = 0.5
    nrm[0, 0] = [0.5, 0.5]
    nrm[0, 1] = [1.0, 1.0]
    nrm[0, 2] = [1.0, 0.5]
    nrm[0, 3] = [0.5, 0.0]
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_NORMALS)
    out = oracle.present(W, H, u, hdr, depth, ao, normals=nrm.view(np.uint32).reshape(H, W).copy()).reshape(H, W, 4)
    assert out[0, 0, :3].tolist() == [0, 0, 255]
    assert out[0, 1, :3].tolist() == [0, 0, 255]
    assert out[0, 2, :3].tolist() == [255, 0, 0]
    assert out[0, 3, :3].tolist() == [0, 255, 0]
    # AO: the half-resolution texel under the pixel centre
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_AO)
    out = oracle.present(W, H, u, hdr, depth, ao).reshape(H, W, 4)
    for y in range(H):
        for x in range(W):
            assert out[y, x, 0] == oracle.srgb8(float(ao[y // 2, x // 2]))
    # another mode (USED_SLICES) passes the HDR colour through; padded rows keep their padding
    u = _u(viewport=[W, H, 0, 0], displayMode=L.DISPLAY_MODE_USED_SLICES)
    out = oracle.present(W, H, u, hdr, depth, ao, pitch=4 * W + 12).reshape(H, 4 * W + 12)
    assert (out[:, 4 * W:] == 0).all()
    c = hdr.view(np.float16).astype(np.float32)
    px = out[:, :4 * W].reshape(H, W, 4)
    assert px[2, 9, 1] == oracle.srgb8(float(c[2, 9, 1]))
