"""Edge cases of the pass entry points: no strands, list-capacity overflow,
one-point-wide viewports."""
import numpy as np
import pytest

from gpu_util import walk_lists

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_zero_strands_is_a_no_op(gpu):
    """Every pass accepts zero strands: nothing is drawn, the frame's clears
    still happen (an empty framebuffer, empty tile lists)."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import native as N
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=64, width=160, height=120)
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=128)
    r.beforeFirstFrame()
    r.finePass.hairRasterizerResultsBuffer.fill_(7.0)
    ctx = r.ctx()
    g, s, h = r.grid, r.sdf, r.hair
    ctx.native.call("fh_grid_pre_sim", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(g.densityVelocity),
                    N.ptr(g.densityGradAndWind), N.ptr(s.texture), s.dims)
    ctx.native.call("fh_sim_integrate", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(h.hairData),
                    N.ptr(h.positionsPrev), N.ptr(h.positionsNow), N.ptr(h.segmentLengths), N.ptr(s.texture),
                    s.dims, N.ptr(g.densityVelocity), N.ptr(g.densityGradAndWind), N.ptr(h.initialPositions),
                    0, h.pointsPerStrand)
    ctx.native.call("fh_grid_post_sim", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(h.hairData),
                    N.ptr(h.positionsPrev), N.ptr(h.positionsNow), N.ptr(g.densityVelocity), 0, h.pointsPerStrand)
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.finePass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdClearBeforeRender(ctx)
    w, hh = ctx.viewport
    tp = r.tilesPass
    ctx.native.call("fh_tiles_draw", ctx.cfg, w, hh, N.ptr(ctx.globalUniforms), N.ptr(h.hairData),
                    N.ptr(h.positionsNow), N.ptr(h.tangents), N.ptr(tp.hairTilesBuffer), ctx.depthTexture,
                    N.ptr(tp.hairTileSegmentsBuffer), tp.segmentsCapacity, N.ptr(tp.segmentCountPerTileBuffer),
                    N.ptr(tp.slotTileBinBuffer), 0, h.pointsPerStrand)
    r.sortPass.cmdSortHairTiles(ctx)
    fp = r.finePass
    ctx.native.call("fh_fine_raster", ctx.cfg, w, hh, N.ptr(ctx.globalUniforms), N.ptr(h.hairData),
                    N.ptr(h.positionsNow), N.ptr(ctx.hairTilesBuffer), N.ptr(ctx.hairTileSegmentsBuffer),
                    tp.segmentsCapacity, N.ptr(fp.hairSlicesDataBuffer), N.ptr(fp.hairRasterizerResultsBuffer),
                    ctx.depthTexture, N.ptr(h.shading), N.ptr(h.tangents), N.ptr(ctx.hairTileListBuffer),
                    N.ptr(tp.slotTileBinBuffer), 0, h.pointsPerStrand)
    ctx.native.call("fh_shading", ctx.cfg, N.ptr(ctx.globalUniforms), N.ptr(h.hairData), N.ptr(h.positionsNow),
                    N.ptr(h.tangents), N.ptr(h.shading), ctx.shadowDepthTexture, ctx.aoTexture,
                    ctx.shadingDepthTexture, 0, h.pointsPerStrand)
    torch.cuda.synchronize()
    assert not _np(tp.hairTilesBuffer).any()
    assert _np(tp.hairTileSegmentsBuffer)[0] == 0
    assert not _np(tp.segmentCountPerTileBuffer).any()
    assert _np(r.sortPass.tileListBuffer)[0] == 0
    assert not _np(fp.hairRasterizerResultsBuffer)[4:].any()


def test_list_capacity_overflow(gpu, oracle):
    """With room for fewer list elements than hit (tile, segment) pairs
    (hairTileSegmentsBuffer.ts:81-104), the counter still counts every hit;
    exactly `capacity` pairs are stored, each in the list of its (tile, bin),
    and the per-tile counts add up to the stored pairs. Which pairs are kept
    follows the order of the atomic appends, as in the reference."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=800, width=320, height=240)
    wl.render.avgSegmentsPerTile = 0.5
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=256)
    r.beforeFirstFrame()
    r.frame()
    ctx = r.ctx()
    tp = r.tilesPass
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    torch.cuda.synchronize()
    cap = tp.segmentsCapacity
    tiles = _np(tp.hairTilesBuffer).view(np.uint32)
    seg = _np(tp.hairTileSegmentsBuffer).view(np.uint32)
    counts = _np(tp.segmentCountPerTileBuffer).view(np.uint32)
    hd = L.pack_hair_data(r.hair.sphere[0], r.hair.sphere[1], wl.strands, wl.points)
    _, es, _, _ = oracle.tiles_draw(wl.render, wl.width, wl.height, r.render_uniforms, hd,
                                    _np(r.hair.positionsNow), _np(r.hair.tangents),
                                    np.ascontiguousarray(r.inputs["meshDepth"].cpu().numpy()), cap, wl.strands,
                                    wl.points)
    assert seg[0] == es[0] and seg[0] > cap
    lists = walk_lists(tiles, seg, cap)
    assert sum(len(v) for v in lists.values()) == cap
    t4 = tiles.reshape(-1, 4)
    for tb, vals in lists.items():
        assert len(vals) == t4[tb, 3]
    assert counts.sum() == cap
    B = wl.render.tileDepthBins
    assert np.array_equal(counts, t4[:, 3].reshape(-1, B).sum(axis=1))


def test_clear_batch(gpu):
    """A clear batch issues its clears and copies as one launch, with 16-byte
    accesses where buffers and lengths allow and word accesses otherwise;
    any other kernel inside a batch fails loudly and leaves the batch open."""
    import torch
    from frostbitten_hair_webgpu_amd import native as N
    ctx = N.Context(0)
    ctx.set_stream(torch.cuda.current_stream())
    g = torch.Generator(device="cpu").manual_seed(5)
    src = torch.randint(0, 2**31 - 1, (8192,), generator=g, dtype=torch.int32).cuda()
    dst = torch.full((8192,), -1, dtype=torch.int32, device="cuda")
    odd = torch.full((1001,), 7, dtype=torch.int32, device="cuda")
    keep = torch.full((8,), 3, dtype=torch.int32, device="cuda")
    ctx.clear_batch_begin()
    ctx.call("fh_copy", N.ptr(dst), N.ptr(src), 4096 * 4)                # aligned: uint4
    ctx.call("fh_copy", N.ptr(dst[4097:]), N.ptr(src[4097:]), 3001 * 4)  # misaligned: words
    ctx.call("fh_clear", N.ptr(odd[1:]), 999 * 4)                        # misaligned, odd length
    ctx.clear_batch_end()
    torch.cuda.synchronize()
    d, s_ = dst.cpu().numpy(), src.cpu().numpy()
    assert np.array_equal(d[:4096], s_[:4096]) and d[4096] == -1
    assert np.array_equal(d[4097:7098], s_[4097:7098]) and (d[7098:] == -1).all()
    o = odd.cpu().numpy()
    assert o[0] == 7 and (o[1:1000] == 0).all() and o[1000] == 7
    from frostbitten_hair_webgpu_amd import assets as A
    cfg = N.RenderCfg.from_config(A.make_workload("small", strands=64, width=64, height=64).render)
    ctx.clear_batch_begin()  # the sort below is refused before it launches anything
    with pytest.raises(N.FhError, match="only clears and copies"):
        ctx.call("fh_tiles_sort", cfg, 64, 64, N.ptr(keep), N.ptr(keep), N.ptr(keep))
    with pytest.raises(N.FhError, match="clear batch is open"):
        ctx.sync()
    ctx.clear_batch_end()
    torch.cuda.synchronize()
    assert (keep.cpu().numpy() == 3).all()


def test_replay_refused_after_scratch_reallocation(gpu):
    """A captured graph and a merged program address the context's scratch as
    it was when they were recorded. Once a larger eager need (here
    fh_scratch_reserve) reallocates it, replaying either returns
    FH_ERR_CAPACITY before anything is enqueued, the device state is left
    as it was, and a fresh capture or recording replays again; a program
    whose recording context was destroyed is refused the same way."""
    import torch
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import batch as B
    from frostbitten_hair_webgpu_amd import native as N
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=256, width=160, height=120)
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=128)
    r.beforeFirstFrame()
    r.frame()  # every call once: the scratch has its size
    torch.cuda.synchronize()
    r.capture_frames()
    r.frame_graph()
    torch.cuda.synchronize()
    prog = B.record_frame(r)
    merged = r.native.program_merge([prog])
    merged.launch()
    torch.cuda.synchronize()
    before = r.hair.positionsNow.clone()
    r.native.scratch_reserve(r.native.scratch_bytes() * 2 + (64 << 20))
    with pytest.raises(N.FhError) as e:
        r.frame_graph()
    assert e.value.code == N.FH_ERR_CAPACITY
    with pytest.raises(N.FhError) as e:
        merged.launch()
    assert e.value.code == N.FH_ERR_CAPACITY
    torch.cuda.synchronize()
    assert torch.equal(r.hair.positionsNow, before)  # nothing ran
    r.capture_frames()
    r.frame_graph()
    merged2 = r.native.program_merge([B.record_frame(r)])
    merged2.launch()
    torch.cuda.synchronize()
    assert np.isfinite(_np(r.hair.positionsNow)).all()
    # a program whose recording context is gone
    r2 = PS.HairRenderer(wl, "cuda:0", shadow_size=128)
    r2.beforeFirstFrame()
    r2.frame()
    torch.cuda.synchronize()
    orphan = r.native.program_merge([B.record_frame(r2)])
    r2.native.close()
    with pytest.raises(N.FhError) as e:
        orphan.launch()
    assert e.value.code == N.FH_ERR_CAPACITY
