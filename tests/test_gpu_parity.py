"""HIP path vs the CPU oracle on identical seeded inputs, through the C-ABI.

Bars: integer / index results bit-exact (P2G grid, tile heads, per-(tile, bin)
segment membership compared as sorted sets, bucket histogram); per-step
positions within 1e-5 relative; images within 2/255 per channel (the HIP fine
pass blends each slice in ascending (strand, segment) order, the oracle walks
the lists in that same order, so the images agree far tighter than that).
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import to_u8, walk_lists

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _renderer(kind="small", strands=512, width=320, height=240, tile=None, bins=None, slices=None, seed=None,
              points=None):
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload(kind, seed=seed, strands=strands, width=width, height=height, points=points)
    if tile:
        wl.render.tileSize = tile
    if bins:
        wl.render.tileDepthBins = bins
    if slices:
        wl.render.slicesPerPixel = slices
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=512)
    return wl, r


def _step_to(r, steps):
    import torch
    r.beforeFirstFrame()
    for _ in range(steps):
        r.frame()
    torch.cuda.synchronize()


def test_debug_raster_matches_reference_snapshot(gpu):
    """swRasterizer.test.ts:57-144 with its fixture, on the HIP raster helpers."""
    import torch
    from conftest import GOLDEN
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import native as N
    pos = np.array([[-0.7, -0.7, 0, 1], [-0.2, 0.2, 0, 1], [0.2, 0.2, 0, 1], [0.7, -0.7, 0, 1]], np.float32)
    tan = A.tangents(pos, 1, 4)
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    u["vpMatrix"][0] = eye
    u["modelMatrix"][0] = eye
    u["cameraPosition"][0] = [0.25, 1.6, 0.6, 0]  # CONFIG.camera.position (constants.ts:127-130)
    u["viewport"][0] = [128, 128, 0, 0]
    u["fiberRadius"] = 0.1
    ctx = N.Context(0)
    du = torch.from_numpy(u.view(np.uint8).copy()).to(gpu)
    dp, dt = torch.from_numpy(pos).to(gpu), torch.from_numpy(tan).to(gpu)
    out = torch.zeros(128 * 128, dtype=torch.int32, device=gpu)
    ctx.call("fh_debug_raster_barycentric", N.ptr(du), N.ptr(dp), N.ptr(dt), 1, 4, N.ptr(out))
    ctx.sync()
    snap = np.fromfile(GOLDEN / "swRasterizer.snapshot.bin", dtype=np.uint32)
    got = _np(out).view(np.uint32)
    assert (got != 0).sum() == 1116
    assert np.array_equal(got, snap)


@pytest.mark.parametrize("kind,strands", [("small", 512), ("c2", 600)])
def test_grid_post_sim_bit_exact(gpu, oracle, kind, strands):
    import torch
    wl, r = _renderer(kind, strands)
    _step_to(r, 2)
    prev, now = _np(r.hair.positionsPrev), _np(r.hair.positionsNow)
    r.gridPostSimPass.cmdUpdateGridsAfterSim(r.ctx(), r.hair)
    torch.cuda.synchronize()
    got = _np(r.grid.densityVelocity)
    exp = oracle.grid_post_sim(wl.sim.gridDims, r.sim_uniforms(r.frameIdx), prev, now, wl.strands, wl.points)
    assert got.any()
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("dims", [8, 200])
def test_grid_post_sim_other_dims(gpu, oracle, dims):
    """Grids of one block (8^3 nodes) and of more blocks than the LDS
    histogram holds (25^3 blocks of 8^3) take the other binning paths."""
    import torch
    from frostbitten_hair_webgpu_amd import native as N
    wl, r = _renderer("c2", 600)
    _step_to(r, 2)
    prev, now = _np(r.hair.positionsPrev), _np(r.hair.positionsNow)
    dv = torch.full((dims ** 3 * 4,), 7, dtype=torch.int32, device=gpu)
    r.ctx().native.call("fh_grid_post_sim", dims, N.ptr(r.simulationUniforms), N.ptr(r.hair.hairData),
                        N.ptr(r.hair.positionsPrev), N.ptr(r.hair.positionsNow), N.ptr(dv),
                        r.hair.strandsCount, r.hair.pointsPerStrand)
    torch.cuda.synchronize()
    exp = oracle.grid_post_sim(dims, r.sim_uniforms(r.frameIdx), prev, now, wl.strands, wl.points)
    got = _np(dv).reshape(exp.shape)
    assert got.any()
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("kind,strands", [("small", 512), ("c2", 600)])
def test_grid_pre_sim(gpu, oracle, kind, strands):
    import torch
    wl, r = _renderer(kind, strands)
    _step_to(r, 2)
    r.upload_sim_uniforms()
    dv = _np(r.grid.densityVelocity)
    r.gridPreSimPass.cmdUpdateGridsBeforeSim(r.ctx(), r.hair)
    torch.cuda.synchronize()
    got = _np(r.grid.densityGradAndWind)
    exp = oracle.grid_pre_sim(wl.sim.gridDims, r.sim_uniforms(r.frameIdx), dv, wl.sdf, wl.sdf_dims)
    assert np.array_equal(got[3::4], exp[3::4])  # wind strength is one of {0, 0.5, 1}
    assert np.allclose(got, exp, rtol=1e-6, atol=0.0)


@pytest.mark.parametrize("kind,strands,steps,points", [("small", 512, 0, None), ("small", 2000, 5, None),
                                                        ("c2", 700, 3, None), ("c2", 300, 2, 5), ("c2", 300, 2, 24),
                                                        ("c2", 200, 2, 40), ("c2", 130, 2, 64)])
def test_sim_integrate_step(gpu, oracle, kind, strands, steps, points):
    """One step from an identical state: positions within 1e-5 relative.
    Strands of 5 to 64 points, staged in LDS with 16 or 32 strands per
    workgroup (the small strand counts take 16, so that every CU gets two
    workgroups)."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    wl, r = _renderer(kind, strands, points=points)
    _step_to(r, steps)
    r.upload_sim_uniforms()
    ctx = r.ctx()
    r.gridPreSimPass.cmdUpdateGridsBeforeSim(ctx, r.hair)
    torch.cuda.synchronize()
    prev, now = _np(r.hair.positionsPrev), _np(r.hair.positionsNow)
    dv, dgw = _np(r.grid.densityVelocity), _np(r.grid.densityGradAndWind)
    r.simPass.cmdSimulateHairPositions(ctx, r.hair)
    torch.cuda.synchronize()
    got = _np(r.hair.positionsNow)
    exp = oracle.sim_integrate(wl.sim.gridDims, r.sim_uniforms(r.frameIdx), prev, now,
                               A.segment_lengths(wl.positions, wl.strands, wl.points), wl.sdf, wl.sdf_dims, dv,
                               dgw, wl.positions, wl.strands, wl.points)
    assert np.isfinite(exp).all()
    scale = np.maximum(np.abs(exp), 1e-1)
    rel = np.abs(got - exp) / scale
    assert rel.max() <= 1e-5, f"max rel err {rel.max():.3e}"
    assert not np.array_equal(got, now)  # the step moved the strands


def _tile_inputs(r):
    from frostbitten_hair_webgpu_amd import layouts as L
    return dict(render_u=r.render_uniforms, hair_data=L.pack_hair_data(r.hair.sphere[0], r.hair.sphere[1],
                                                                       r.hair.strandsCount, r.hair.pointsPerStrand),
                positions=_np(r.hair.positionsNow), tangents=_np(r.hair.tangents),
                depth=np.ascontiguousarray(r.inputs["meshDepth"].cpu().numpy()))


def _check_slot_tags(tags_ranks, counts, n):
    """Every slot names a (tile, bin) and a rank; each list's ranks are 0 .. count-1."""
    tags, ranks = tags_ranks[0:2 * n:2], tags_ranks[1:2 * n:2]
    assert np.array_equal(np.bincount(tags, minlength=counts.shape[0]), counts)
    order = np.lexsort((ranks, tags))
    st, sr = tags[order], ranks[order]
    first = np.r_[True, st[1:] != st[:-1]]
    run_start = np.maximum.accumulate(np.where(first, np.arange(n), 0))
    assert np.array_equal(sr, np.arange(n) - run_start)


@pytest.mark.parametrize("tile,bins,width,height,radius", [(16, 4, 320, 240, None), (8, 8, 320, 240, None),
                                                            (32, 4, 320, 240, None), (16, 4, 333, 171, None),
                                                            (8, 8, 250, 97, None), (8, 4, 320, 240, 0.06)])
def test_tiles_bit_exact(gpu, oracle, tile, bins, width, height, radius):
    """Tile depth ranges, per-(tile, bin) counts and list contents equal the
    oracle's; two cases have partial tiles on both edges, and the last one has
    fibers wide enough that most segments cover more than 8 tiles (the tile
    pass queues those in 8-tile pieces for its second kernel)."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    if radius is None:
        wl, r = _renderer("small", 800, width, height, tile=tile, bins=bins)
    else:
        wl = A.make_workload("small", strands=300, width=width, height=height)
        wl.render.tileSize, wl.render.tileDepthBins, wl.render.fiberRadius = tile, bins, radius
        wl.render.avgSegmentsPerTile = 1024
        r = PS.HairRenderer(wl, "cuda:0", shadow_size=256)
    _step_to(r, 2)
    ctx = r.ctx()
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    torch.cuda.synchronize()
    tp = r.tilesPass
    got_tiles = _np(tp.hairTilesBuffer).view(np.uint32)
    got_seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32)
    got_cnt = _np(tp.segmentCountPerTileBuffer).view(np.uint32)
    got_keys = _np(tp.slotTileBinBuffer).view(np.uint32)
    i = _tile_inputs(r)
    et, es, ec, ek = oracle.tiles_draw(wl.render, wl.width, wl.height, i["render_u"], i["hair_data"],
                                       i["positions"], i["tangents"], i["depth"], tp.segmentsCapacity,
                                       wl.strands, wl.points)
    assert got_seg[0] == es[0] and es[0] > 0
    g4, e4 = got_tiles.reshape(-1, 4), et.reshape(-1, 4)
    assert np.array_equal(g4[:, 0], e4[:, 0]) and np.array_equal(g4[:, 1], e4[:, 1])  # depth maxima
    assert np.array_equal(g4[:, 3], e4[:, 3])  # entries per (tile, bin)
    assert np.array_equal(got_cnt, ec)
    gl, el = walk_lists(got_tiles, got_seg, tp.segmentsCapacity), walk_lists(et, es, tp.segmentsCapacity)
    assert gl == el
    if radius is not None:
        tiles_of = {}
        for tb, vals in gl.items():
            for v in vals:
                tiles_of.setdefault(v, set()).add(tb // bins)
        wide = sum(len(t) > 8 for t in tiles_of.values())
        print(f"segments drawn into more than 8 tiles: {wide} of {len(tiles_of)}", flush=True)
        assert wide > 100
    n = min(int(got_seg[0]), tp.segmentsCapacity)
    # every slot is tagged with the (tile, bin) whose list holds it and with a
    # rank in that list; the ranks of one list are 0 .. count-1
    for tb, vals in gl.items():
        assert len(vals) == g4[tb, 3]
    _check_slot_tags(got_keys, g4[:, 3], n)


@pytest.mark.parametrize("form", ["x_in_z", "z_and_w", "affine"])
def test_tiles_depth_bins_projection_forms(gpu, oracle, form):
    """The tile pass takes a segment's depth bin per (segment, tile) pair from
    the pair's extreme depths when the view-space depth is provably monotone
    in the hair depth (a perspective inverse: clip z constant), and per pixel
    otherwise (hairTilesPass.wgsl.ts:197-206). Three other inverse
    projections, uploaded after two frames: one whose z row reads the pixel's
    x (per pixel), one whose clip z and w both vary with depth (per pixel),
    and an affine one with w constant (per pair). Lists and depth ranges
    equal the oracle's, which bins every pixel, and the bins are spread."""
    import torch
    wl, r = _renderer("small", 800, 320, 240, tile=16, bins=8)
    _step_to(r, 2)
    u = r.render_uniforms
    m = u["projMatrixInv"][0].reshape(-1)  # column-major 4x4
    if form == "x_in_z":
        m[2] = np.float32(2e-4)
    elif form == "z_and_w":
        m[10] = np.float32(0.25)
    else:
        # view z = alpha * z + beta, spreading the hair's depths over its
        # view-space bounds (hairTilesPass.wgsl.ts:242-246)
        mv = u["modelViewMat"][0].reshape(4, 4).T.astype(np.float64)
        c, rad = r.hair.sphere
        vz = (mv @ np.array([c[0], c[1], c[2], 1.0]))[2]
        proj = u["projMatrix"][0].reshape(4, 4).T.astype(np.float64)
        ndc = lambda v: (proj[2, 2] * v + proj[2, 3]) / (proj[3, 2] * v + proj[3, 3])
        za, zb = ndc(vz + rad), ndc(vz - rad)
        alpha = ((vz + rad) - (vz - rad)) / (za - zb)
        beta = (vz + rad) - alpha * za
        m[:] = 0.0
        m[0], m[5], m[10], m[14], m[15] = 1.0, 1.0, np.float32(alpha), np.float32(beta), 1.0
    r.globalUniforms.copy_(torch.from_numpy(u.view(np.uint8).reshape(-1).copy()).to(gpu))
    ctx = r.ctx()
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    torch.cuda.synchronize()
    tp = r.tilesPass
    got_tiles = _np(tp.hairTilesBuffer).view(np.uint32)
    got_seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32)
    i = _tile_inputs(r)
    et, es, _, _ = oracle.tiles_draw(wl.render, wl.width, wl.height, i["render_u"], i["hair_data"],
                                     i["positions"], i["tangents"], i["depth"], tp.segmentsCapacity,
                                     wl.strands, wl.points)
    assert got_seg[0] == es[0] and es[0] > 0
    g4, e4 = got_tiles.reshape(-1, 4), et.reshape(-1, 4)
    assert np.array_equal(g4[:, :2], e4[:, :2]) and np.array_equal(g4[:, 3], e4[:, 3])
    el = walk_lists(et, es, tp.segmentsCapacity)
    assert walk_lists(got_tiles, got_seg, tp.segmentsCapacity) == el
    used = np.bincount(np.array(sorted(el.keys())) % 8, minlength=8)
    print(f"{form}: pairs per bin {used.tolist()}", flush=True)
    assert (used > 0).sum() >= 3


def test_tile_sort(gpu, oracle):
    import torch
    wl, r = _renderer("small", 800, 320, 240)
    _step_to(r, 1)
    ctx = r.ctx()
    r.sortPass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdSortHairTiles(ctx)
    torch.cuda.synchronize()
    counts = _np(r.tilesPass.segmentCountPerTileBuffer).view(np.uint32)
    got_list = _np(r.sortPass.tileListBuffer).view(np.uint32)
    got_b = _np(r.sortPass.bucketsDataBuffer).view(np.uint32).reshape(-1, 2)
    el, eb = oracle.tiles_sort(wl.render, wl.width, wl.height, counts)
    eb = eb.reshape(-1, 2)
    assert np.array_equal(got_b, eb)
    n = got_list[0]
    assert n == el[0] == (counts > 0).sum()
    assert sorted(got_list[1:1 + n]) == sorted(el[1:1 + n])
    bucket = np.minimum(counts[got_list[1:1 + n]] // wl.render.sortBucketSize, wl.render.sortBuckets - 1)
    assert (np.diff(bucket.astype(np.int64)) <= 0).all()


@pytest.mark.parametrize("tile,bins,slices,width,height", [(16, 4, 4, 320, 240), (8, 8, 8, 320, 240),
                                                            (32, 4, 4, 320, 240), (16, 4, 4, 333, 171)])
def test_fine_pass_image(gpu, oracle, tile, bins, slices, width, height):
    import torch
    wl, r = _renderer("small", 1200, width, height, tile=tile, bins=bins, slices=slices)
    _step_to(r, 2)
    ctx = r.ctx()
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.finePass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    r.sortPass.cmdSortHairTiles(ctx)
    torch.cuda.synchronize()
    tp = r.tilesPass
    tiles = _np(tp.hairTilesBuffer).view(np.uint32).copy()
    seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32).copy()
    tl = _np(r.sortPass.tileListBuffer).view(np.uint32).copy()
    shading = _np(r.hair.shading)
    r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, tp)
    torch.cuda.synchronize()
    got = _np(r.finePass.hairRasterizerResultsBuffer)
    i = _tile_inputs(r)
    exp = oracle.fine_raster(wl.render, wl.width, wl.height, i["render_u"], i["positions"], tiles, seg,
                             tp.segmentsCapacity, i["depth"], shading, i["tangents"], tl, wl.strands,
                             wl.points, canonical=True)
    assert exp[4:].reshape(-1, 4)[:, 3].max() > 0.5
    assert got[0:1].view(np.uint32)[0] >= tl[0]  # the queue served every tile
    ge, ee = got[4:], exp[4:]
    diff = np.abs(ge - ee).reshape(-1, 4)
    bad = np.nonzero(diff.max(axis=1) > 1e-6)[0]
    if bad.size:
        k = bad[np.argmax(diff[bad].max(axis=1))]
        print(f"pixels differing by >1e-6: {bad.size} of {diff.shape[0]}; worst pixel {k} "
              f"(x={k % wl.width}, y={k // wl.width}) got {ge.reshape(-1, 4)[k]} exp {ee.reshape(-1, 4)[k]}")
    assert np.abs(to_u8(ge) - to_u8(ee)).max() <= 2
    # one f16 unit in the last place at 0.5 is 4.9e-4: a rounding flip of a packed fragment colour
    assert diff.max() <= 5e-4


def test_fine_pass_dense_bins(gpu, oracle):
    """6,000 strands on 160x120 with 32x32 tiles: (tile, bin) lists of up to
    ~4,200 segments, so the bin sort takes every path (one wave up to 64
    keys, registers up to 256, 512, 1024, 2048 and 4096 keys, global memory
    beyond), and the image still matches the oracle's. The slice pool is
    sized so that no bin overflows it."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=6000, width=160, height=120)
    rc = wl.render
    rc.tileSize, rc.tileDepthBins, rc.slicesPerPixel = 32, 4, 4
    rc.avgSegmentsPerTile, rc.avgFragmentsPerSlice, rc.processorCount = 4096, 64, 32
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=256)
    r.beforeFirstFrame()
    ctx = r.ctx()
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.finePass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    r.sortPass.cmdSortHairTiles(ctx)
    torch.cuda.synchronize()
    tp = r.tilesPass
    tiles = _np(tp.hairTilesBuffer).view(np.uint32).copy()
    seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32).copy()
    tl = _np(r.sortPass.tileListBuffer).view(np.uint32).copy()
    counts = tiles.reshape(-1, 4)[:, 3]
    print(f"dense bins: largest list {counts.max()}, lists over 256: {(counts > 256).sum()}, "
          f"over 1024: {(counts > 1024).sum()}, over 4096: {(counts > 4096).sum()}", flush=True)
    assert counts.max() > 4096 and ((counts > 64) & (counts <= 256)).any() and ((counts > 1024) & (counts <= 4096)).any()
    assert seg[0] <= tp.segmentsCapacity
    shading = _np(r.hair.shading)
    r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, tp)
    torch.cuda.synchronize()
    got = _np(r.finePass.hairRasterizerResultsBuffer)
    i = _tile_inputs(r)
    exp = oracle.fine_raster(rc, wl.width, wl.height, i["render_u"], i["positions"], tiles, seg,
                             tp.segmentsCapacity, i["depth"], shading, i["tangents"], tl, wl.strands,
                             wl.points, canonical=True)
    assert exp[4:].reshape(-1, 4)[:, 3].max() > 0.5
    ge, ee = got[4:], exp[4:]
    assert np.abs(to_u8(ge) - to_u8(ee)).max() <= 2
    assert np.abs(ge - ee).max() <= 5e-4


def test_rebuilt_slot_tags(gpu):
    """fh_tiles_rebuild_slot_keys (for lists another tile pass wrote) tags every
    slot with its list and a rank, and the fine pass draws the same image from
    those tags as from the tile pass's own."""
    import torch
    from frostbitten_hair_webgpu_amd import native as N
    wl, r = _renderer("small", 1200, 320, 240)
    _step_to(r, 2)
    ctx = r.ctx()
    tp = r.tilesPass
    tp.cmdClearBeforeRender(ctx)
    r.finePass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdClearBeforeRender(ctx)
    tp.cmdDrawHairToTiles(ctx, r.hair)
    r.sortPass.cmdSortHairTiles(ctx)
    r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, tp)
    torch.cuda.synchronize()
    img1 = _np(r.finePass.hairRasterizerResultsBuffer).copy()
    tags1 = _np(tp.slotTileBinBuffer).view(np.uint32).copy()
    counts = _np(tp.hairTilesBuffer).view(np.uint32).reshape(-1, 4)[:, 3].copy()
    n = min(int(_np(tp.hairTileSegmentsBuffer).view(np.uint32)[0]), tp.segmentsCapacity)
    assert n > 0
    rebuilt = torch.full_like(tp.slotTileBinBuffer, -1)
    ctx.native.call("fh_tiles_rebuild_slot_keys", ctx.cfg, wl.width, wl.height, N.ptr(tp.hairTilesBuffer),
                    N.ptr(tp.hairTileSegmentsBuffer), tp.segmentsCapacity, N.ptr(rebuilt))
    torch.cuda.synchronize()
    tags2 = _np(rebuilt).view(np.uint32).copy()
    assert np.array_equal(tags2[0:2 * n:2], tags1[0:2 * n:2])
    assert np.array_equal(_np(tp.hairTilesBuffer).view(np.uint32).reshape(-1, 4)[:, 3], counts)
    _check_slot_tags(tags2, counts, n)
    tp.slotTileBinBuffer = rebuilt
    r.finePass.cmdClearBeforeRender(ctx)
    r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, tp)
    torch.cuda.synchronize()
    assert np.array_equal(_np(r.finePass.hairRasterizerResultsBuffer), img1)


def test_shading(gpu, oracle):
    import torch
    wl, r = _renderer("small", 512, 320, 240)
    _step_to(r, 1)
    r.shadingPass.cmdComputeShadingPoints(r.ctx(), r.hair)
    torch.cuda.synchronize()
    got = _np(r.hair.shading)
    i = r.inputs
    exp = oracle.shading(wl.render, r.render_uniforms, _np(r.hair.positionsNow), _np(r.hair.tangents),
                         np.ascontiguousarray(i["shadowMap"].cpu().numpy()), np.ascontiguousarray(i["ao"].cpu().numpy()),
                         np.ascontiguousarray(i["hairDepth"].cpu().numpy()), wl.strands, wl.points).reshape(-1, 4)
    assert np.isfinite(got).all()
    assert np.allclose(got, exp, rtol=2e-4, atol=2e-6)


def test_full_frames(gpu):
    """Several sim+render frames of the small workload: finite state and a non-empty image."""
    import torch
    wl, r = _renderer("small", 2000, 640, 480)
    _step_to(r, 5)
    img = r.finePass.image(wl.width, wl.height).cpu().numpy()
    assert np.isfinite(img).all()
    assert (img[..., 3] > 0.5).sum() > 500
    assert np.isfinite(_np(r.hair.positionsNow)).all()


def test_graph_replay_matches_eager(gpu):
    """Frames replayed from HIP graphs (capture_frames / frame_graph) give the
    same positions, image and shading as the same frames launched eagerly,
    replay after replay, with eager work and synchronisations in between."""
    import torch
    _, ra = _renderer("small", 2000, 640, 480)
    _, rb = _renderer("small", 2000, 640, 480)
    for r in (ra, rb):
        r.beforeFirstFrame()
        r.frame()  # sizes the scratch before any capture
    torch.cuda.synchronize()
    rb.capture_frames()
    for k in range(6):
        ra.frame()
        rb.frame_graph()
        torch.cuda.synchronize()
        assert ra.frameIdx == rb.frameIdx
        assert torch.equal(ra.hair.positionsNow, rb.hair.positionsNow), f"positions differ at replay {k}"
        assert torch.equal(ra.hair.shading, rb.hair.shading), f"shading differs at replay {k}"
        ia = _np(ra.finePass.hairRasterizerResultsBuffer)[4:]
        ib = _np(rb.finePass.hairRasterizerResultsBuffer)[4:]
        assert np.array_equal(ia, ib), f"image differs at replay {k}"
    assert np.isfinite(_np(rb.hair.positionsNow)).all()
