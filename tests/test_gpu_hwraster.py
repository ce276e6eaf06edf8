"""HIP hair depth / normals, shadow-map hair and GTAO vs the CPU oracle,
through the C-ABI.

Bars: the rasterizers are bit-exact (integer coverage, f32 depth and
normals in one fixed evaluation order on both sides, winner independent of
the order the GPU runs the triangles in); GTAO takes cos / sin / acos from
the device's math library and the oracle from libm, and a one-ulp difference
can move a horizon sample to the neighbouring pixel, so the AO texels agree
to 1/255 except for under 1% of them, with a mean difference below 1e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _renderer(strands, width, height, hw_passes=True, points=None, shadow_size=512):
    from frostbitten_hair_webgpu_amd import assets as A
    from frostbitten_hair_webgpu_amd import passes as PS
    wl = A.make_workload("small", strands=strands, width=width, height=height, points=points)
    r = PS.HairRenderer(wl, "cuda:0", shadow_size=shadow_size, hw_passes=hw_passes)
    return wl, r


def _step(r, frames=2):
    import torch
    r.beforeFirstFrame()
    for _ in range(frames):
        r.frame()
    torch.cuda.synchronize()


@pytest.mark.parametrize("strands,width,height", [(1200, 320, 240), (2000, 333, 171), (600, 640, 480)])
def test_hw_hair_bit_exact(gpu, oracle, strands, width, height):
    import torch
    wl, r = _renderer(strands, width, height)
    _step(r)
    print(f"hw hair {strands} strands {width}x{height}", flush=True)
    ctx = r.ctx()
    r.hwHairPass.cmdDrawHair(ctx, r.hair, r.inputs["meshDepth"], r.inputs["meshNormals"])
    torch.cuda.synchronize()
    got_d, got_n = _np(r.depthTexture), _np(r.normalsTexture).view(np.uint32)
    exp_d, exp_n = oracle.hw_hair(False, width, height, r.render_uniforms, _np(r.hair.positionsNow),
                                  _np(r.hair.tangents), _np(r.inputs["meshDepth"]), wl.strands, wl.points,
                                  normals=_np(r.inputs["meshNormals"]).view(np.uint32))
    hair = exp_d < _np(r.inputs["meshDepth"])
    assert hair.sum() > 200
    assert np.array_equal(got_d.view(np.uint32), exp_d.view(np.uint32)), f"{(got_d != exp_d).sum()} depths differ"
    assert np.array_equal(got_n, exp_n), f"{(got_n != exp_n).sum()} normals differ"


def test_shadow_map_hair_bit_exact(gpu, oracle):
    import torch
    wl, r = _renderer(1500, 320, 240, shadow_size=512)
    _step(r)
    ctx = r.ctx()
    r.shadowMapPass.cmdUpdateShadowMap(ctx, r.hair, r.inputs["shadowMap"])
    torch.cuda.synchronize()
    got = _np(r.shadowTexture)
    exp, _ = oracle.hw_hair(True, 512, 512, r.render_uniforms, _np(r.hair.positionsNow), _np(r.hair.tangents),
                            _np(r.inputs["shadowMap"]), wl.strands, wl.points)
    assert (exp < _np(r.inputs["shadowMap"])).sum() > 500
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"{(got != exp).sum()} texels differ"


def test_gtao_vs_oracle(gpu, oracle):
    import torch
    wl, r = _renderer(1500, 320, 240)
    _step(r)
    ctx = r.ctx()
    r.hwHairPass.cmdDrawHair(ctx, r.hair, r.inputs["meshDepth"], r.inputs["meshNormals"])
    r.aoPass.cmdCalcAo(ctx)
    torch.cuda.synchronize()
    got = _np(r.aoTexture)
    exp = oracle.calc_ao(r.render_uniforms, _np(r.depthTexture), _np(r.normalsTexture).view(np.uint32),
                         got.shape[1], got.shape[0])
    d = np.abs(got - exp)
    print(f"gtao: max {d.max():.4g}, mean {d.mean():.3g}, > 1/255: {(d > 1 / 255).mean():.4f}", flush=True)
    assert np.isfinite(got).all()
    assert (exp > 0.05).sum() > 100
    assert (d > 1 / 255).mean() < 0.01 and d.mean() < 1e-3


def test_frames_with_hw_passes(gpu, oracle):
    """Frames with the depth, normals, shadow map and AO computed every frame:
    graph replays equal eager frames, and the shading pass reads the live
    textures (its output equals the oracle's shading on them)."""
    import torch
    _, ra = _renderer(1500, 320, 240)
    wl, rb = _renderer(1500, 320, 240)
    for r in (ra, rb):
        r.beforeFirstFrame()
        r.frame()
    torch.cuda.synchronize()
    rb.capture_frames()
    for k in range(4):
        ra.frame()
        rb.frame_graph()
        torch.cuda.synchronize()
        for name in ("depthTexture", "normalsTexture", "aoTexture", "shadowTexture", "screenTexture"):
            assert torch.equal(getattr(ra, name), getattr(rb, name)), f"{name} differs at replay {k}"
        assert torch.equal(ra.hair.shading, rb.hair.shading)
    exp = oracle.shading(wl.render, rb.render_uniforms, _np(rb.hair.positionsNow), _np(rb.hair.tangents),
                         _np(rb.shadowTexture), _np(rb.aoTexture), _np(rb.depthTexture), wl.strands,
                         wl.points).reshape(-1, 4)
    assert np.allclose(_np(rb.hair.shading), exp, rtol=2e-4, atol=2e-6)
    assert not np.array_equal(_np(rb.aoTexture), np.ones_like(_np(rb.aoTexture)))
