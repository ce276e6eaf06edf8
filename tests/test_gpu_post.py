"""HIP combine and present passes vs the CPU oracle, through the C-ABI.

Bars: the combine pass (f32 blend of f16 texels, no transcendental) is
bit-exact; the present pass is bit-exact in the debug views and at most one
8-bit code apart in the tone-mapped views (its gamma pow is exp2(log2(x) / g)
in hardware math, the oracle's is libm powf); a whole frame's presented image
(HIP chain vs the oracle's tile, fine, combine and present passes on the same
inputs) is within 2/255 per channel.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DISPLAY_MODE_OFFSET = 1132  # byte offset of Uniforms.displayMode (renderUniformsBuffer.ts:131)
COLOR_MGMT_OFFSET = 608


def _np(t):
    return t.detach().cpu().numpy()


def _renderer(strands=1200, width=320, height=240, tile=None, bins=None):
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=strands, width=width, height=height)
    if tile:
        wl.render.tileSize = tile
    if bins:
        wl.render.tileDepthBins = bins
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=512)
    return wl, r


def _set_u32(r, offset, value):
    import torch
    b = np.array([value], np.uint32).view(np.uint8)
    r.render_uniforms.view(np.uint8)[offset:offset + 4] = b
    r.globalUniforms[offset:offset + 4] = torch.from_numpy(b.copy()).to(r.globalUniforms.device)


def _set_color_mgmt(r, gamma, exposure, dither):
    import torch
    b = np.array([gamma, exposure, dither, 0], np.float32).view(np.uint8)
    r.render_uniforms.view(np.uint8)[COLOR_MGMT_OFFSET:COLOR_MGMT_OFFSET + 16] = b
    r.globalUniforms[COLOR_MGMT_OFFSET:COLOR_MGMT_OFFSET + 16] = torch.from_numpy(b.copy()).to(r.globalUniforms.device)


def _raster_frame(r):
    """Runs a frame up to the fine pass and returns the combine pass's inputs."""
    import torch
    r.beforeFirstFrame()
    r.frame()
    ctx = r.ctx()
    r.drawScene(ctx)
    r.tilesPass.cmdClearBeforeRender(ctx)
    r.finePass.cmdClearBeforeRender(ctx)
    r.sortPass.cmdClearBeforeRender(ctx)
    r.tilesPass.cmdDrawHairToTiles(ctx, r.hair)
    r.sortPass.cmdSortHairTiles(ctx)
    r.finePass.cmdRasterizeSlicesHair(ctx, r.hair, r.tilesPass)
    torch.cuda.synchronize()
    tp = r.tilesPass
    return ctx, dict(tiles=_np(tp.hairTilesBuffer).view(np.uint32).copy(),
                     seg=_np(tp.hairTileSegmentsBuffer).view(np.uint32).copy(),
                     raster=_np(r.finePass.hairRasterizerResultsBuffer).copy(),
                     counts=_np(tp.segmentCountPerTileBuffer).view(np.uint32).copy(),
                     hdr=_np(r.hdrRenderTexture).view(np.uint16).copy())


@pytest.mark.parametrize("tile,bins,width,height", [(16, 4, 320, 240), (8, 8, 333, 171)])
def test_combine_bit_exact(gpu, oracle, tile, bins, width, height):
    import torch
    from frostbitten_hair_webgpu_amd import layouts as L
    wl, r = _renderer(1200, width, height, tile, bins)
    ctx, i = _raster_frame(r)
    assert (i["raster"][4:].reshape(-1, 4)[:, 3] > 0.5).sum() > 100
    for mode in (L.DISPLAY_MODE_FINAL, L.DISPLAY_MODE_FINAL | (1 << 8), L.DISPLAY_MODE_TILES | (370 << 8),
                 L.DISPLAY_MODE_TILES_PPLL | (40 << 8), L.DISPLAY_MODE_HW_RENDER):
        print(f"combine mode {mode:#x} {width}x{height} tile {tile}", flush=True)
        _set_u32(r, DISPLAY_MODE_OFFSET, mode)
        r.hdrRenderTexture.copy_(torch.from_numpy(i["hdr"]).to(gpu))
        r.hairCombinePass.cmdCombineRasterResults(ctx)
        torch.cuda.synchronize()
        got = _np(r.hdrRenderTexture).view(np.uint16)
        exp = oracle.hair_combine(wl.render, width, height, r.render_uniforms, i["tiles"], i["seg"], i["raster"],
                                  i["counts"], i["hdr"])
        assert np.array_equal(got, exp), f"mode {mode:#x}: {(got != exp).sum()} half values differ"
        assert not np.array_equal(got, i["hdr"])


def test_present_within_one_code(gpu, oracle):
    import torch
    from frostbitten_hair_webgpu_amd import layouts as L
    wl, r = _renderer(800, 320, 240)
    ctx, i = _raster_frame(r)
    r.hairCombinePass.cmdCombineRasterResults(ctx)
    torch.cuda.synchronize()
    hdr = _np(r.hdrRenderTexture).view(np.uint16).copy()
    from frostbitten_hair_webgpu_amd import native as N
    W, H = wl.width, wl.height
    rng = np.random.default_rng(11)
    # depths whose linear depth spans the DEPTH view's range [0.002, 0.01] of the far plane and beyond
    n, f = 0.01, 100.0
    lin = rng.uniform(0.001, 0.011, (H, W))
    depth = ((n * f / (lin * (f - n)) - f) / (n - f)).astype(np.float32)
    depth_dev = torch.from_numpy(depth).to(gpu)
    ctx.shadingDepthTexture = N.tex(depth_dev, W, H)
    ao = rng.uniform(0, 1, (H // 2, W // 2)).astype(np.float32)
    ao_dev = torch.from_numpy(ao).to(gpu)
    ctx.aoTexture = N.tex(ao_dev, W // 2, H // 2)
    nrm = rng.uniform(0, 1, (H, W, 2)).astype(np.float16).view(np.uint32).reshape(H, W).copy()
    ctx.normalsTexture = torch.from_numpy(nrm.view(np.int32)).to(gpu)
    pitch = 4 * W + 64
    screen = torch.zeros(H * pitch, dtype=torch.uint8, device=gpu)
    cases = [(L.DISPLAY_MODE_FINAL, (1.0, 1.0, 1.0), False), (L.DISPLAY_MODE_FINAL, (2.2, 0.85, 1.0), True),
             (L.DISPLAY_MODE_HW_RENDER, (2.2, 1.3, 64.0), True), (L.DISPLAY_MODE_DEPTH, (1.0, 1.0, 1.0), False),
             (L.DISPLAY_MODE_NORMALS, (1.0, 1.0, 1.0), False), (L.DISPLAY_MODE_AO, (1.0, 1.0, 1.0), False),
             (L.DISPLAY_MODE_USED_SLICES, (1.0, 1.0, 1.0), False)]
    for mode, cm, per_pixel in cases:
        print(f"present mode {mode} colorMgmt {cm} per-pixel dither {per_pixel}", flush=True)
        _set_u32(r, DISPLAY_MODE_OFFSET, mode)
        _set_color_mgmt(r, *cm)
        r.presentPass.ditherPerPixel = per_pixel
        screen.fill_(0)
        r.presentPass.cmdDraw(ctx, screen, pitch)
        torch.cuda.synchronize()
        got = _np(screen).reshape(H, pitch)
        exp = oracle.present(W, H, r.render_uniforms, hdr, depth, ao, normals=nrm, dither_per_pixel=per_pixel,
                             pitch=pitch).reshape(H, pitch)
        assert (got[:, 4 * W:] == 0).all()  # row padding untouched
        d = np.abs(got.astype(np.int32) - exp.astype(np.int32))
        if mode in (L.DISPLAY_MODE_FINAL, L.DISPLAY_MODE_HW_RENDER):
            # tone mapping takes pow(x, 1 / gamma) as exp2(log2(x) / gamma) in
            # hardware math: at most one code apart where it meets a rounding edge
            assert d.max() <= 1, f"mode {mode}: max difference {d.max()} codes"
            assert (d > 0).mean() < 0.01
        else:
            # correctly rounded arithmetic and an exact sRGB store: identical codes
            assert d.max() == 0, f"mode {mode}: {(d > 0).sum()} channels differ"
        assert len(np.unique(got[:, :4 * W].reshape(H, W, 4)[..., :3])) > 8  # not a flat image


@pytest.mark.parametrize("tile,bins,width,height", [(16, 4, 320, 240), (8, 8, 333, 171)])
def test_combine_present_equals_two_passes(gpu, tile, bins, width, height):
    """fh_combine_present leaves the HDR texture and the screen texture bit
    for bit as fh_hair_combine and then fh_present do, in every display
    mode, blending onto the HDR texture in place and onto a separate scene
    texture (whose texels the debug views' discarded texels take). A scene
    that overlaps the HDR texture is refused."""
    import torch
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import native as N
    wl, r = _renderer(1200, width, height, tile, bins)
    ctx, i = _raster_frame(r)
    W, H = width, height
    rng = np.random.default_rng(3)
    ctx.aoTexture = N.tex(torch.from_numpy(rng.uniform(0, 1, (H // 2, W // 2)).astype(np.float32)).to(gpu),
                          W // 2, H // 2)
    nrm = rng.uniform(0, 1, (H, W, 2)).astype(np.float16).view(np.uint32).reshape(H, W).copy()
    ctx.normalsTexture = torch.from_numpy(nrm.view(np.int32)).to(gpu)
    pitch = 4 * W + 32
    scene = torch.from_numpy(i["hdr"]).to(gpu)
    for mode in (L.DISPLAY_MODE_FINAL, L.DISPLAY_MODE_FINAL | (1 << 8), L.DISPLAY_MODE_TILES | (370 << 8),
                 L.DISPLAY_MODE_TILES_PPLL | (40 << 8), L.DISPLAY_MODE_HW_RENDER, L.DISPLAY_MODE_DEPTH,
                 L.DISPLAY_MODE_NORMALS, L.DISPLAY_MODE_AO):
        print(f"combine+present mode {mode:#x} {W}x{H} tile {tile}", flush=True)
        _set_u32(r, DISPLAY_MODE_OFFSET, mode)
        r.presentPass.ditherPerPixel = mode == L.DISPLAY_MODE_HW_RENDER
        outs = []
        for form in ("two passes", "fused in place", "fused from the scene"):
            screen = torch.zeros(H * pitch, dtype=torch.uint8, device=gpu)
            if form == "fused from the scene":
                r.hdrRenderTexture.view(torch.int16).fill_(7)
            else:
                r.hdrRenderTexture.copy_(scene)
            if form == "two passes":
                r.hairCombinePass.cmdCombineRasterResults(ctx)
                r.presentPass.cmdDraw(ctx, screen, pitch)
            else:
                r.hairCombinePass.cmdCombineAndPresent(ctx, r.presentPass, screen,
                                                       scene if form == "fused from the scene" else None, pitch)
            torch.cuda.synchronize()
            outs.append((_np(r.hdrRenderTexture).view(np.uint16).copy(), _np(screen)))
        for (h, sc), form in zip(outs[1:], ("fused in place", "fused from the scene")):
            assert np.array_equal(h, outs[0][0]), f"mode {mode:#x}, {form}: HDR texels differ"
            assert np.array_equal(sc, outs[0][1]), f"mode {mode:#x}, {form}: screen texels differ"
        if mode in (L.DISPLAY_MODE_FINAL, L.DISPLAY_MODE_HW_RENDER):
            assert len(np.unique(outs[0][1])) > 8  # not a flat image
        if mode in (L.DISPLAY_MODE_TILES | (370 << 8), L.DISPLAY_MODE_TILES_PPLL | (40 << 8)):
            # some texels are discarded and keep the scene, others show counts
            same = (outs[0][0] == i["hdr"]).reshape(H * W, 4).all(axis=1)
            assert 0 < int(same.sum()) < H * W
    shifted = r.hdrRenderTexture.view(-1)[4:]  # 8 bytes into the HDR texture
    with pytest.raises(RuntimeError, match="overlaps"):
        r.hairCombinePass.cmdCombineAndPresent(ctx, r.presentPass, torch.zeros(H * W * 4, dtype=torch.uint8,
                                                                                device=gpu), shifted)


def test_presented_frame_vs_oracle_chain(gpu, oracle):
    """The screen texture of a whole frame against the oracle's tile, sort,
    fine, combine and present passes on the frame's own inputs."""
    import torch
    wl, r = _renderer(1500, 320, 240)
    r.beforeFirstFrame()
    r.frame()
    r.frame()
    torch.cuda.synchronize()
    # inputs of the last frame's render passes: positions after its sim step,
    # shading of the frame before (the fine pass runs before the shading pass)
    pos = _np(r.hair.positionsNow)
    shading_before = _np(r.hair.shading).copy()
    ctx = r.ctx()
    r.drawHair(ctx)
    torch.cuda.synchronize()
    got = _np(r.screenTexture).reshape(wl.height, wl.width, 4)
    from frostbitten_hair_webgpu_amd import layouts as L
    rc = wl.render
    hd = L.pack_hair_data(r.hair.sphere[0], r.hair.sphere[1], r.hair.strandsCount, r.hair.pointsPerStrand)
    tan = _np(r.hair.tangents)
    md = np.ascontiguousarray(r.inputs["meshDepth"].cpu().numpy())
    cap = r.tilesPass.segmentsCapacity
    tiles, seg, counts, _ = oracle.tiles_draw(rc, wl.width, wl.height, r.render_uniforms, hd, pos, tan, md, cap,
                                              wl.strands, wl.points)
    tl, _ = oracle.tiles_sort(rc, wl.width, wl.height, counts)
    raster = oracle.fine_raster(rc, wl.width, wl.height, r.render_uniforms, pos, tiles, seg, cap, md,
                                shading_before, tan, tl, wl.strands, wl.points, canonical=True)
    hdr = oracle.hair_combine(rc, wl.width, wl.height, r.render_uniforms, tiles, seg, raster, counts,
                              np.ascontiguousarray(r.inputs["hdrScene"].cpu().numpy()))
    exp = oracle.present(wl.width, wl.height, r.render_uniforms, hdr,
                         np.ascontiguousarray(r.inputs["hairDepth"].cpu().numpy()),
                         np.ascontiguousarray(r.inputs["ao"].cpu().numpy())).reshape(wl.height, wl.width, 4)
    d = np.abs(got.astype(np.int32) - exp.astype(np.int32))
    print(f"presented frame: max difference {d.max()} codes, {(d > 0).sum()} channels differ", flush=True)
    assert d.max() <= 2
    scene = oracle.present(wl.width, wl.height, r.render_uniforms,
                           np.ascontiguousarray(r.inputs["hdrScene"].cpu().numpy()),
                           np.ascontiguousarray(r.inputs["hairDepth"].cpu().numpy()),
                           np.ascontiguousarray(r.inputs["ao"].cpu().numpy())).reshape(wl.height, wl.width, 4)
    assert (np.abs(got.astype(np.int32) - scene).max(axis=2) > 8).sum() > 200  # hair is visible
