"""HIP background, mesh and mesh-shadow passes vs the CPU oracle, through the
C-ABI, and frames that draw them every frame.

Bars: the background and the rasterized depth, normals and shadow depth are
bit-exact (the same f32 operations in one fixed order on both sides, the
winner of each pixel independent of the order the GPU runs the triangles
in). The meshes' colour takes pow from the device's math library and the
oracle from libm: the rgba16float texels agree to within 2 half-precision
ulps, and almost all of them exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _renderer(kind="small", strands=1500, width=320, height=240, shadow_size=512, **kw):
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload(kind, strands=strands, width=width, height=height)
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=shadow_size, mesh_passes=True, **kw)
    return wl, r


def _h2f(bits):
    return np.asarray(bits).view(np.uint16).view(np.float16).astype(np.float32)


def _half_ulps(a_bits, b_bits):
    """Distance in half-precision ulps of two rgba16float arrays of finite,
    non-negative values (bit patterns order like the values)."""
    a = np.asarray(a_bits).view(np.uint16).astype(np.int64)
    b = np.asarray(b_bits).view(np.uint16).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("kind,width,height", [("small", 320, 240), ("c2", 1920, 1080)])
def test_background_bit_exact(gpu, oracle, kind, width, height):
    import torch
    wl, r = _renderer(kind, strands=500, width=width, height=height)
    print(f"background {kind} {width}x{height}", flush=True)
    r.backgroundPass.cmdDraw(r.ctx())
    torch.cuda.synchronize()
    got = _np(r.hdrRenderTexture).view(np.uint16).reshape(height, width, 4)
    exp = oracle.background(width, height, r.render_uniforms)
    assert np.unique(exp[..., 0]).size > 50
    assert np.array_equal(got, exp), f"{(got != exp).any(axis=-1).sum()} texels differ"


@pytest.mark.parametrize("kind,size", [("small", 512), ("c2", 2048)])
def test_shadow_map_meshes_bit_exact(gpu, oracle, kind, size):
    import torch
    wl, r = _renderer(kind, strands=500, shadow_size=size)
    print(f"shadow meshes {kind} {size}^2, {r.meshes.triangles()} triangles", flush=True)
    sm = r.ctx().shadowDepthTexture
    r.native.call("fh_shadow_map_meshes", r.globalUniforms.data_ptr(), r.meshes.handle, sm.data, size)
    torch.cuda.synchronize()
    got = _np(r.shadowTexture)
    exp = oracle.meshes(True, size, size, r.render_uniforms, r.meshList)
    assert (exp < 1).sum() > size * size // 400
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"{(got != exp).sum()} texels differ"


@pytest.mark.parametrize("kind,width,height", [("small", 320, 240), ("small", 333, 171), ("c2", 1920, 1080)])
def test_draw_meshes_vs_oracle(gpu, oracle, kind, width, height):
    """The camera pass over the background, with the GPU's shadow map of the
    meshes and the hair and a seeded AO texture: depth and normals bit-exact,
    colour within 2 half ulps."""
    import torch
    wl, r = _renderer(kind, strands=1500 if kind == "small" else 4000, width=width, height=height)
    print(f"draw meshes {kind} {width}x{height}, {r.meshes.triangles()} triangles", flush=True)
    rng = np.random.default_rng(5)
    r.aoTexture.copy_(torch.from_numpy(rng.uniform(0.2, 1.0, r.aoTexture.shape).astype(np.float16)
                                       .astype(np.float32)))
    ctx = r.ctx()
    r.backgroundPass.cmdDraw(ctx)
    r.shadowMapPass.cmdUpdateShadowMap(ctx, r.hair, meshes=r.meshes)
    torch.cuda.synchronize()
    bg = _np(r.hdrRenderTexture).view(np.uint16).reshape(height, width, 4).copy()
    r.drawMeshesPass.cmdDrawMeshes(ctx, r.meshes, r.meshDepthTexture, r.meshNormalsTexture)
    torch.cuda.synchronize()
    sm = _np(r.shadowTexture)
    d, n, c = oracle.meshes(False, width, height, r.render_uniforms, r.meshList, shadow_map=sm,
                            ao=_np(r.aoTexture), hdr=bg)
    got_d, got_n = _np(r.meshDepthTexture), _np(r.meshNormalsTexture).view(np.uint32)
    got_c = _np(r.hdrRenderTexture).view(np.uint16).reshape(height, width, 4)
    drawn = d < 1
    assert drawn.mean() > 0.01
    assert np.array_equal(got_d.view(np.uint32), d.view(np.uint32)), f"{(got_d != d).sum()} depths differ"
    assert np.array_equal(got_n, n), f"{(got_n != n).sum()} normals differ"
    ulps = _half_ulps(got_c, c)
    print(f"colour: max {ulps.max()} half ulps, {(ulps > 0).mean():.5f} of channels differ", flush=True)
    assert ulps.max() <= 2 and (ulps > 0).mean() < 0.01
    assert np.array_equal(got_c[~drawn], bg[~drawn])
    # the shadow map darkens part of the meshes: not every drawn pixel lit alike
    assert np.unique(_h2f(c[drawn][:, 0])).size > 100


def test_frames_read_this_frames_meshes(gpu, oracle):
    """Frames with the mesh passes (graph replays): the meshes' depth and
    normals that the tile pass and the hair draw read are this frame's,
    equal to the oracle's camera pass, and the hair draw adds the hair over
    them."""
    import torch
    wl, r = _renderer()
    r.beforeFirstFrame()
    r.frame()
    torch.cuda.synchronize()
    r.capture_frames()
    for _ in range(2):
        r.frame_graph()
    torch.cuda.synchronize()
    d, n, _ = oracle.meshes(False, wl.width, wl.height, r.render_uniforms, r.meshList,
                            shadow_map=_np(r.shadowTexture), ao=np.ones_like(_np(r.aoTexture)),
                            hdr=np.zeros((wl.height, wl.width, 4), np.uint16))
    assert (d < 1).mean() > 0.02
    assert np.array_equal(_np(r.meshDepthTexture).view(np.uint32), d.view(np.uint32))
    assert np.array_equal(_np(r.meshNormalsTexture).view(np.uint32), n)
    hair = _np(r.depthTexture) < d
    assert hair.sum() > 100 and np.array_equal(_np(r.depthTexture)[~hair], d[~hair])


def test_split_pass_equals_fused(gpu):
    """fh_draw_meshes_visibility + fh_shade_meshes leave exactly what
    fh_draw_meshes leaves; the visibility texture holds each drawn pixel's
    scene triangle and 0xffffffff elsewhere, and indices past the scene's
    triangles are skipped by the shading call."""
    import torch
    wl, r = _renderer(width=333, height=171)
    ctx = r.ctx()
    r.backgroundPass.cmdDraw(ctx)
    r.shadowMapPass.cmdUpdateShadowMap(ctx, r.hair, meshes=r.meshes)
    bg = r.hdrRenderTexture.clone()
    r.drawMeshesPass.cmdDrawMeshes(ctx, r.meshes, r.meshDepthTexture, r.meshNormalsTexture)
    torch.cuda.synchronize()
    fused = [t.clone() for t in (r.meshDepthTexture, r.meshNormalsTexture, r.hdrRenderTexture)]
    r.hdrRenderTexture.copy_(bg)
    r.meshDepthTexture.zero_()
    r.meshNormalsTexture.fill_(7)
    r.drawMeshesPass.cmdDrawMeshes(ctx, r.meshes, r.meshDepthTexture, r.meshNormalsTexture,
                                   visibility=r.meshVisibility)
    r.drawMeshesPass.cmdShadeMeshes(ctx, r.meshes, r.meshVisibility)
    torch.cuda.synchronize()
    for a, b, name in zip(fused, (r.meshDepthTexture, r.meshNormalsTexture, r.hdrRenderTexture),
                          ("depth", "normals", "hdr")):
        assert torch.equal(a, b), name
    vis = _np(r.meshVisibility).view(np.uint32)
    drawn = _np(r.meshDepthTexture) < 1
    assert (vis[~drawn] == 0xffffffff).all() and (vis[drawn] < r.meshes.triangles()).all()
    r.meshVisibility.fill_(-2)  # 0xfffffffe: past the scene's triangles
    before = r.hdrRenderTexture.clone()
    r.drawMeshesPass.cmdShadeMeshes(ctx, r.meshes, r.meshVisibility)
    torch.cuda.synchronize()
    assert torch.equal(before, r.hdrRenderTexture)


def test_band_queue_cap_same_result(gpu):
    """A triangle whose bands do not fit the large-triangle queue draws
    itself (fh_set_raster_band_cap forces that with a queue of 2 bands): the
    meshes' depth, normals, colour and shadow map and the hair's depth and
    normals at 4K equal those with the full queue, bit for bit."""
    import torch
    wl, r = _renderer("c3", strands=3000, width=3840, height=2160, shadow_size=4096)
    # the scene has triangles larger than one rasterizer work item (1,024 px boxes)
    mvp = np.frombuffer(r.render_uniforms["mvpMatrix"].tobytes(), np.float32).reshape(4, 4).T
    big = 0
    for m in r.meshList:
        if m.get("preview"):
            continue
        c = np.c_[m["positions"], np.ones(len(m["positions"]))] @ mvp.T
        ok = c[:, 3] > 0
        s = np.c_[(c[:, 0] / c[:, 3] * 0.5 + 0.5) * wl.width, (0.5 - c[:, 1] / c[:, 3] * 0.5) * wl.height]
        tri = m["indices"]
        sx, sy = s[tri, 0], s[tri, 1]
        area = (sx.max(1) - sx.min(1)) * (sy.max(1) - sy.min(1))
        big += int(((area > 1024) & ok[tri].all(1)).sum())
    print(f"band queue: {big} mesh triangles with boxes over 1,024 px", flush=True)
    assert big > 50

    def draw():
        ctx = r.ctx()
        r.backgroundPass.cmdDraw(ctx)
        r.shadowMapPass.cmdUpdateShadowMap(ctx, r.hair, meshes=r.meshes)
        r.drawMeshesPass.cmdDrawMeshes(ctx, r.meshes, r.meshDepthTexture, r.meshNormalsTexture)
        r.hwHairPass.cmdDrawHair(ctx, r.hair, r.meshDepthTexture, r.meshNormalsTexture)
        torch.cuda.synchronize()
        return [t.clone() for t in (r.shadowTexture, r.meshDepthTexture, r.meshNormalsTexture, r.hdrRenderTexture,
                                     r.depthTexture, r.normalsTexture)]

    full = draw()
    r.native.call("fh_set_raster_band_cap", 2)
    try:
        capped = draw()
    finally:
        r.native.call("fh_set_raster_band_cap", 0)
    for a, b, name in zip(full, capped, ("shadow", "mesh depth", "mesh normals", "hdr", "depth", "normals")):
        assert torch.equal(a, b), name
    assert (_np(r.meshDepthTexture) < 1).mean() > 0.01
