"""Pins of the oracle's tile pass (K4) depth bins with more than one bin.

hairTilesPass.wgsl.ts:198-201 puts a (segment, tile) pair in bin
(B - 1) - getDepthBin(B, depthBoundsVS, z_view) of the tile, where z_view is
the view-space depth of the pixel (projMatrixInv applied to the pixel
position and its NDC depth: only the depth decides z) and depthBoundsVS is
the view-space z of the hair's bounding sphere +- its radius
(getHairDepthBoundsVS :242-246). getDepthBin (shaderSnippets.wgls.ts:148-156)
is clamp(u32((z - zmin) / |zmax - zmin| * B), 0, B - 1), so bin 0 holds the
nearest strands, and over the pixels a pair covers the least bin wins
(:198-201, "closest bin only").

A perspective camera at the origin looks down -z (view matrix = identity);
the model matrix moves the hair by MODEL_Z along z, so object-space positions
and the object-space bounding sphere reach view space only through it. Each
strand is a short segment whose projection lies inside one 16 px tile at a
chosen view depth, so its bin follows from the closed form above.
"""
import math

import numpy as np
import pytest

W, H, TILE, B = 128, 64, 16, 8
ZMIN, ZMAX = -7.0, -3.0  # bounding sphere centre z = -5 in view space, radius 2
MODEL_Z = -1.0
SPHERE_OBJ = [0.0, 0.0, -5.0 - MODEL_Z]


def _setup():
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import matrices as M
    proj = M.perspective(math.radians(60.0), W / H, 0.01, 100.0)
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    model = M.translation((0.0, 0.0, MODEL_Z))
    u["vpMatrix"] = M.to_f32_colmajor(proj)
    u["projMatrixInv"] = M.to_f32_colmajor(M.invert(proj))
    u["modelMatrix"] = M.to_f32_colmajor(model)
    u["modelViewMat"] = M.to_f32_colmajor(model)
    u["viewport"] = [W, H, 0, 0]
    u["cameraPosition"] = [0, 0, 0, 0]
    u["fiberRadius"] = 0.2
    u["maxDrawnHairSegments"] = 4096
    rc = L.RenderConfig(tileSize=TILE, tileDepthBins=B, invalidTilesPerSegmentThreshold=128)
    return proj, u, rc


def _world(proj, px, py, z):
    """The object-space point at view depth z that projects to pixel (px, py), y up."""
    nx, ny = px / W * 2 - 1, py / H * 2 - 1
    return [nx * -z / proj[0, 0], ny * -z / proj[1, 1], z - MODEL_Z]


def _expected_bin(z):
    t = (z - ZMIN) / abs(ZMAX - ZMIN)
    return (B - 1) - min(max(int(max(t * B, 0.0)), 0), B - 1)


def _lists(tiles, seg):
    """{(tile, bin): [strand, ...]} read back through the linked lists."""
    t = tiles.reshape(-1, 4)
    out = {}
    for i in range(t.shape[0]):
        ptr = int(t[i, 2])
        if ptr == 0:
            continue
        ptr -= 1
        items = []
        while ptr != 0xFFFFFFFF:
            items.append(int(seg[1 + 2 * ptr]) & 0xFFFFFF)
            ptr = int(seg[2 + 2 * ptr])
        out[(i // B, i % B)] = sorted(items)
    return out


# (tile x, tile y, view depth at the segment's start, at its end)
STRANDS = [
    (0, 0, -3.25, -3.25),  # t = 0.9375 -> bin 0
    (1, 0, -3.75, -3.75),  # t = 0.8125 -> bin 1
    (2, 0, -4.75, -4.75),  # t = 0.5625 -> bin 3
    (3, 0, -6.75, -6.75),  # t = 0.0625 -> bin 7
    (4, 0, -2.5, -2.5),    # nearer than the sphere: t > 1 clamps to bin 0
    (5, 0, -7.5, -7.5),    # farther: t < 0 saturates to 0 -> bin 7
    (0, 1, -3.25, -3.25),  # two strands over one tile keep their own bins: 0 ...
    (0, 1, -5.25, -5.25),  # ... and 4
    (1, 1, -3.3, -3.9),    # slanted across bins 0 and 1: the least (closest) bin, 0
    (2, 2, -4.3, -5.7),    # slanted across bins 2..4: bin 2
]


def test_tile_pass_depth_bins_closed_form(oracle):
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    proj, u, rc = _setup()
    S = len(STRANDS)
    pos = np.zeros((S, 2, 4), np.float32)
    for s, (tx, ty, z0, z1) in enumerate(STRANDS):
        cx, cy = tx * TILE + 8, ty * TILE + 8
        pos[s, 0, :3] = _world(proj, cx - 3, cy, z0)
        pos[s, 1, :3] = _world(proj, cx + 3, cy, z1)
    pos[:, :, 3] = 1.0
    pos = pos.reshape(-1, 4)
    tan = A.tangents(pos, S, 2)
    hd = L.pack_hair_data(SPHERE_OBJ, 2.0, S, 2)
    depth = np.ones((H, W), np.float32)
    tiles, seg, counts, _ = oracle.tiles_draw(rc, W, H, u, hd, pos, tan, depth, 4095, S, 2)
    lists = _lists(tiles, seg)
    tcx = W // TILE
    exp = {}
    for s, (tx, ty, z0, z1) in enumerate(STRANDS):
        exp.setdefault((ty * tcx + tx, _expected_bin(max(z0, z1))), []).append(s)
    assert lists == exp
    assert seg[0] == S and counts.sum() == S
    t = tiles.reshape(-1, 4).astype(np.float64)
    # a slanted strand's bin holds the nearest and farthest covered depths
    tx, ty, z0, z1 = STRANDS[8]
    row = t[(ty * tcx + tx) * B + 0]
    near, far = ((proj[2, 2] * z + proj[2, 3]) / -z for z in (z0, z1))
    assert near - 1e-3 < 1 - row[0] / 2 ** 32 < row[1] / 2 ** 32 < far + 1e-3


def test_tile_depth_bounds_of_a_flat_strand(oracle):
    """On the screen's middle row the billboard offset is along y only, so the
    quad has one NDC depth d = (P22 z + P23) / -z. Its bin's maximum is stored
    as u32(d 2^32) and its minimum inverted, u32((1 - d) 2^32)
    (hairTilesResultBuffer.ts:14-42); the quad straddles tile rows 1 and 2,
    and both tiles list it in the same bin."""
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    proj, u, rc = _setup()
    z = -4.75
    pos = np.zeros((2, 4), np.float32)
    pos[0, :3] = _world(proj, 40 - 3, H // 2, z)
    pos[1, :3] = _world(proj, 40 + 3, H // 2, z)
    pos[:, 3] = 1.0
    hd = L.pack_hair_data(SPHERE_OBJ, 2.0, 1, 2)
    tiles, seg, counts, _ = oracle.tiles_draw(rc, W, H, u, hd, pos, A.tangents(pos, 1, 2),
                                              np.ones((H, W), np.float32), 4095, 1, 2)
    tcx = W // TILE
    assert _lists(tiles, seg) == {(1 * tcx + 2, 3): [0], (2 * tcx + 2, 3): [0]}
    ndc = (proj[2, 2] * z + proj[2, 3]) / -z
    t = tiles.reshape(-1, 4).astype(np.float64)
    for ty in (1, 2):
        row = t[(ty * tcx + 2) * B + 3]
        assert row[1] / 2 ** 32 == pytest.approx(ndc, abs=2e-6)
        assert 1 - row[0] / 2 ** 32 == pytest.approx(ndc, abs=2e-6)


def test_tile_pass_depth_test_against_the_scene(oracle):
    """A pixel whose hair depth is above the depth texture's (read at row
    H - y, hairTilesPass.wgsl.ts:181-185) is skipped: a wall at NDC depth
    just in front of the strand hides it, one just behind does not."""
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    proj, u, rc = _setup()
    z = -4.0
    ndc = (proj[2, 2] * z + proj[2, 3]) / -z
    pos = np.zeros((2, 4), np.float32)
    pos[0, :3] = _world(proj, 40 - 3, 24, z)
    pos[1, :3] = _world(proj, 40 + 3, 24, z)
    pos[:, 3] = 1.0
    tan = A.tangents(pos, 1, 2)
    hd = L.pack_hair_data(SPHERE_OBJ, 2.0, 1, 2)
    for wall, listed in ((ndc - 1e-4, False), (ndc + 1e-4, True)):
        depth = np.full((H, W), wall, np.float32)
        tiles, seg, counts, _ = oracle.tiles_draw(rc, W, H, u, hd, pos, tan, depth, 4095, 1, 2)
        assert (seg[0] == 1) == listed and (counts.sum() == 1) == listed
    # a wall in front of the strand's rows only where the texture row is H - y:
    # rows 24 +- 4 in y-up pixels are texture rows 36..44
    depth = np.ones((H, W), np.float32)
    depth[H - 30:H - 18, :] = ndc - 1e-4
    tiles, seg, counts, _ = oracle.tiles_draw(rc, W, H, u, hd, pos, tan, depth, 4095, 1, 2)
    assert seg[0] == 0
    depth = np.ones((H, W), np.float32)
    depth[18:30, :] = ndc - 1e-4  # the same rows without the flip: the strand is not hidden
    tiles, seg, counts, _ = oracle.tiles_draw(rc, W, H, u, hd, pos, tan, depth, 4095, 1, 2)
    assert seg[0] == 1


def _overflow_scene(n_strands):
    """n horizontal strands over an 8x8 viewport with an identity camera, all
    over pixel rows 4 and 5 (16 fragments each: 8 columns x 2 rows), at depths
    0.1, 0.2, ... and in one slice; strand k is shaded with colour COLOURS[k]."""
    from frostbitten_hair_webgpu_amd import layouts as L
    pos = np.zeros((n_strands, 2, 4), np.float32)
    for k in range(n_strands):
        pos[k] = [[-2, 0.1, 0.1 * (k + 1), 1], [2, 0.1, 0.1 * (k + 1), 1]]
    pos = pos.reshape(-1, 4)
    tan = np.tile(np.array([[1, 0, 0, 0]], np.float32), (2 * n_strands, 1))
    eye = np.eye(4, dtype=np.float32).reshape(16)
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    for k in ("vpMatrix", "modelMatrix", "modelViewMat", "projMatrixInv"):
        u[k] = eye
    u["cameraPosition"] = [0.0, 0.1, 5.0, 0]
    u["viewport"] = [8, 8, 0, 0]
    u["fiberRadius"] = 0.2
    u["maxDrawnHairSegments"] = 64
    hd = L.pack_hair_data([0, 0, 0.5], 1.0, n_strands, 2)
    shading = np.repeat(np.asarray(COLOURS[:n_strands], np.float32), 2, axis=0)
    return pos, tan, u, hd, shading


COLOURS = [(1.0, 0.0, 0.0, 1.0), (0.0, 1.0, 0.0, 1.0), (0.0, 0.0, 1.0, 1.0), (1.0, 1.0, 0.0, 1.0),
           (0.0, 1.0, 1.0, 1.0)]


@pytest.mark.parametrize("avg,drawn", [(1, [1, 2, 3, 4]), (2, [0, 1, 2, 3, 4])])
def test_fine_pass_slice_pool_exhaustion(oracle, avg, drawn):
    """A bin may take avgFragmentsPerSlice * slices * tile^2 fragments
    (hairSlicesDataBuffer.ts:27-40): 64 here with avgFragmentsPerSlice = 1.
    The list is walked from its head, i.e. from the last appended segment
    (strand 4) down; four strands fill the 64 slots exactly, the loop then
    stops (hairFinePass.wgsl.ts:252-264) and strand 0 is never drawn. With
    128 slots all five are. Each pixel's slice list reads back from the last
    pushed fragment, so the blend order is ascending strand: colour =
    sum_k c_k (1 - a)^k over the drawn strands, alpha = 1 - (1 - a)^n, with
    coverage a = 0.5 on row 4 and 0.25 on row 5."""
    from frostbitten_hair_webgpu_amd import layouts as L
    rc = L.RenderConfig(tileSize=8, tileDepthBins=1, slicesPerPixel=1, shadingPoints=2, alphaMultiplier=1.0,
                        avgFragmentsPerSlice=avg)
    pos, tan, u, hd, shading = _overflow_scene(5)
    depth = np.ones((8, 8), np.float32)
    tiles, seg, counts, _ = oracle.tiles_draw(rc, 8, 8, u, hd, pos, tan, depth, 63, 5, 2)
    assert counts[0] == 5
    tl, _ = oracle.tiles_sort(rc, 8, 8, counts)
    img = oracle.fine_raster(rc, 8, 8, u, pos, tiles, seg, 63, depth, shading, tan, tl, 5, 2)
    px = np.append(img[4:], np.zeros(4, np.float32)).reshape(8, 8, 4)  # the last pixel is outside the buffer
    for y, a in ((4, 0.5), (5, 0.25)):
        rgb = sum(np.array(COLOURS[k][:3]) * (1 - a) ** i for i, k in enumerate(drawn))
        exp = [*rgb, 1 - (1 - a) ** len(drawn)]
        for x in range(8):
            assert px[y, x] == pytest.approx(exp, abs=2e-3), (y, x)
    assert np.all(px[:4] == 0) and np.all(px[6:] == 0)


def test_fine_pass_depth_test_and_canonical_order(oracle):
    """The fine pass tests each fragment against the depth texture at row
    H - y with GL_LESS (processHairSegment.wgsl.ts:31-37): texture rows 3 and 4
    in front of every strand hide pixel rows 5 and 4; the same values in rows
    4 and 5 hide pixel rows 4 and 3 instead. With canonical order the oracle
    sorts each list by (strand, segment), so a list whose entries were
    appended in another order gives the same image."""
    from frostbitten_hair_webgpu_amd import layouts as L
    rc = L.RenderConfig(tileSize=8, tileDepthBins=1, slicesPerPixel=1, shadingPoints=2, alphaMultiplier=1.0,
                        avgFragmentsPerSlice=2)
    pos, tan, u, hd, shading = _overflow_scene(5)
    ones = np.ones((8, 8), np.float32)
    tiles, seg, counts, _ = oracle.tiles_draw(rc, 8, 8, u, hd, pos, tan, ones, 63, 5, 2)
    tl, _ = oracle.tiles_sort(rc, 8, 8, counts)

    def image(depth, segbuf=seg, canonical=True):
        img = oracle.fine_raster(rc, 8, 8, u, pos, tiles, segbuf, 63, depth, shading, tan, tl, 5, 2,
                                 canonical=canonical)
        return np.append(img[4:], np.zeros(4, np.float32)).reshape(8, 8, 4)

    ref = image(ones)
    assert np.all(ref[4:6, :, 3] > 0)
    wall = ones.copy()
    wall[3:5] = 0.05
    got = image(wall)
    assert np.all(got[4:6] == 0)
    wall = ones.copy()
    wall[4:6] = 0.05
    got = image(wall)
    assert np.all(got[4] == 0) and np.array_equal(got[5], ref[5])
    # permute the strand ids over the list's slots, keeping the links
    shuffled = seg.copy()
    shuffled[1:11:2] = seg[1:11:2][[3, 0, 4, 1, 2]]
    assert np.array_equal(image(ones, shuffled), ref)
    assert not np.allclose(image(ones, shuffled, canonical=False), ref, atol=1e-3)
