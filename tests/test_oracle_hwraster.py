"""The oracle's hair depth / normals rasterizer, shadow-map hair draw and GTAO
against closed forms.

The reference draws these with the hardware rasterizer (hwHairPass,
shadowMapPass's hair draw) and a fragment shader (aoPass + gtao); its exact
rasterization rules are the WebGPU implementation's, so they are parity
unpinned against the reference itself. The restatement fixes one set of rules
(1/256 px snapping, top-left rule at pixel centres, 'less' in primitive order)
and these tests pin that it implements them: watertight shared edges, exact
pixel sets, planar depth, tie order, occlusion by the meshes, octahedral
normals, and GTAO's background and unoccluded-wall values.
"""
import numpy as np
import pytest

EYE = np.eye(4, dtype=np.float32).reshape(16)


def _u(W, H, radius, shadow_radius=None, **kw):
    from frostbitten_hair_webgpu_amd import layouts as L
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    for k in ("vpMatrix", "viewMatrix", "projMatrix", "projMatrixInv", "modelMatrix", "modelViewMat",
              "shadowSourceModelViewMat", "shadowSourceProjMatrix"):
        u[k] = EYE
    u["viewport"] = [W, H, 0, 0]
    u["fiberRadius"] = radius
    u["shadowSourcePosition"] = [0, 0, 0, radius if shadow_radius is None else shadow_radius]
    for k, v in kw.items():
        u[k] = v
    return u


def _strand(points, tangent=(0.0, 1.0, 0.0)):
    """One strand through `points` (NDC = object space under identity matrices)."""
    p = np.zeros((len(points), 4), np.float32)
    p[:, :3] = points
    p[:, 3] = 1
    t = np.zeros_like(p)
    t[:, :3] = tangent
    return p, t


def test_pixel_aligned_quad_exact_and_watertight(oracle):
    """A vertical segment of half-width 0.25 from y = -0.25 to 0.25 on a 16x16
    target is the pixel square [6, 10) x [6, 10); its two triangles share the
    diagonal through the centres of (6, 9), (7, 8), (8, 7), (9, 6): every
    pixel of the square is covered exactly once, none outside."""
    W = H = 16
    u = _u(W, H, 0.25)
    pos, tan = _strand([(0, -0.25, 0.5), (0, 0.25, 0.5)])
    cov = oracle.hw_coverage(False, W, H, u, pos, tan, 1, 2)
    exp = np.zeros((H, W), np.uint32)
    exp[6:10, 6:10] = 1
    assert np.array_equal(cov, exp)
    depth, normals = oracle.hw_hair(False, W, H, u, pos, tan, np.ones((H, W), np.float32), 1, 2,
                                    normals=np.zeros((H, W), np.uint32))
    assert np.allclose(depth[6:10, 6:10], 0.5, rtol=0, atol=1e-7) and (depth[exp == 0] == 1).all()
    # tangent (0, 1, 0) -> world tangent modelMatrix * (t, 1) = (0, 1, 0) -> octahedral (0.5, 1.0)
    assert (normals[6:10, 6:10] == oracle.encode_oct(0.0, 1.0, 0.0)).all()
    assert (normals[exp == 0] == 0).all()


def test_fan_of_strands_is_watertight(oracle):
    """Adjacent segments of a bent strand and neighbouring strands share
    vertices: summed coverage has no pixel counted twice within one segment's
    pair of triangles (the pair shares one edge)."""
    W, H = 64, 48
    rng = np.random.default_rng(1)
    for k in range(20):
        a = rng.uniform(-0.8, 0.8, 3).astype(np.float32)
        b = (a + rng.uniform(-0.3, 0.3, 3)).astype(np.float32)
        a[2], b[2] = 0.3, 0.6
        t = (b - a) / np.linalg.norm(b - a)
        pos, tan = _strand([a, b], tangent=tuple(t))
        u = _u(W, H, float(rng.uniform(0.01, 0.1)))
        cov = oracle.hw_coverage(False, W, H, u, pos, tan, 1, 2)
        assert cov.max() <= 1, k
        # the quad's area in pixels matches the number of covered centres to within its perimeter
        assert cov.sum() > 0 or np.linalg.norm((b - a)[:2]) < 0.05


def test_planar_depth_and_clip(oracle):
    """Depth is the screen-space interpolation of NDC z: on a vertical
    segment from z = 0.2 (y = -0.5) to 0.6 (y = 0.5), row py holds the z of
    its centre's y. Fragments with z outside [0, 1] are clipped."""
    W = H = 16
    u = _u(W, H, 0.25)
    pos, tan = _strand([(0, -0.5, 0.2), (0, 0.5, 0.6)])
    depth, _ = oracle.hw_hair(False, W, H, u, pos, tan, np.ones((H, W), np.float32), 1, 2)
    for py in range(4, 12):
        ndc_y = 1.0 - (py + 0.5) / 8.0
        z = 0.2 + (ndc_y + 0.5) * 0.4
        assert np.allclose(depth[py, 6:10], z, atol=2e-6), py
    pos, tan = _strand([(0, -0.5, -0.2), (0, 0.5, 0.2)])
    depth, _ = oracle.hw_hair(False, W, H, u, pos, tan, np.ones((H, W), np.float32), 1, 2)
    drawn = depth < 1
    assert drawn.sum() > 0 and (depth[drawn] >= 0).all()
    # z crosses 0 between rows 7 and 8 (NDC y = 0): the rows below are clipped
    assert drawn[4:8, 6:10].all() and not drawn[8:].any()


def test_less_in_primitive_order_and_mesh_occlusion(oracle):
    W = H = 16
    u = _u(W, H, 0.25)
    # two strands over the same pixels at the same depth, different tangents:
    # the first one's fragments stay (a later fragment must be strictly nearer)
    p1, t1 = _strand([(0, -0.25, 0.5), (0, 0.25, 0.5)], tangent=(0.0, 1.0, 0.0))
    p2, t2 = _strand([(0, -0.25, 0.5), (0, 0.25, 0.5)], tangent=(0.0, 1.0, 0.001))
    pos, tan = np.concatenate([p1, p2]), np.concatenate([t1, t2])
    _, n = oracle.hw_hair(False, W, H, u, pos, tan, np.ones((H, W), np.float32), 2, 2,
                          normals=np.zeros((H, W), np.uint32))
    assert (n[6:10, 6:10] == oracle.encode_oct(0.0, 1.0, 0.0)).all()
    # the second strand nearer: it wins everywhere
    p2[:, 2] = 0.4
    d, n = oracle.hw_hair(False, W, H, u, np.concatenate([p1, p2]), tan, np.ones((H, W), np.float32), 2, 2,
                          normals=np.zeros((H, W), np.uint32))
    # a constant depth interpolates to within an ulp or two (the weights sum to 1 in f32)
    assert np.allclose(d[6:10, 6:10], 0.4, rtol=0, atol=1e-7)
    assert (n[6:10, 6:10] != oracle.encode_oct(0.0, 1.0, 0.0)).all()
    # meshes at 0.5 hide hair at 0.5 (strict less) but not hair at 0.4
    mesh = np.ones((H, W), np.float32)
    mesh[6:8] = 0.5
    d, n = oracle.hw_hair(False, W, H, u, p1, t1, mesh, 1, 2, normals=np.zeros((H, W), np.uint32))
    assert (d[6:8, 6:10] == 0.5).all() and (n[6:8] == 0).all()
    assert np.allclose(d[8:10, 6:10], 0.5, rtol=0, atol=1e-7)
    assert (n[8:10, 6:10] != 0).all()


def test_shadow_pass_uses_shadow_source(oracle):
    """The shadow draw takes the shadow source's matrices and fiber radius
    (shadows.sourcePosition.w) and writes depth only."""
    W = H = 16
    pos, tan = _strand([(0, -0.25, 0.5), (0, 0.25, 0.5)])
    u = _u(W, H, 0.01, shadow_radius=0.25)
    d_shadow, _ = oracle.hw_hair(True, W, H, u, pos, tan, np.ones((H, W), np.float32), 1, 2)
    d_cam, _ = oracle.hw_hair(False, W, H, _u(W, H, 0.25), pos, tan, np.ones((H, W), np.float32), 1, 2)
    assert np.array_equal(d_shadow, d_cam)
    shift = EYE.reshape(4, 4).copy()
    shift[3, 0] = 0.5  # column-major: translation x of the shadow source's view
    u["shadowSourceModelViewMat"] = shift.reshape(16)
    d_shift, _ = oracle.hw_hair(True, W, H, u, pos, tan, np.ones((H, W), np.float32), 1, 2)
    # the camera pass with the same view gives the same depth; the hair moved right
    d_cam_shift, _ = oracle.hw_hair(False, W, H, _u(W, H, 0.25, modelViewMat=shift.reshape(16)), pos, tan,
                                    np.ones((H, W), np.float32), 1, 2)
    assert np.array_equal(d_shift, d_cam_shift)
    assert (d_shift[:, :8] == 1).all() and (d_shift[6:10] < 1).sum() >= 12


def test_octahedral_normals(oracle):
    h = lambda a, b: int(np.array([a, b], np.float16).view(np.uint16).astype(np.uint32) @ [1, 1 << 16])
    assert oracle.encode_oct(0, 0, 1) == h(0.5, 0.5)
    assert oracle.encode_oct(1, 0, 0) == h(1.0, 0.5)
    assert oracle.encode_oct(0, -1, 0) == h(0.5, 0.0)
    assert oracle.encode_oct(0, 0, -1) == h(1.0, 1.0)
    rng = np.random.default_rng(4)
    v = rng.standard_normal((200, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for n in v:
        e = oracle.encode_oct(*map(float, n))
        f = np.array([e & 0xFFFF, e >> 16], np.uint16).view(np.float16).astype(np.float64) * 2 - 1
        d = np.array([f[0], f[1], 1 - abs(f[0]) - abs(f[1])])
        t = max(0.0, -d[2])
        d[0] = d[0] - t if d[0] >= 0 else d[0] + t
        d[1] = d[1] - t if d[1] >= 0 else d[1] + t
        assert np.allclose(d / np.linalg.norm(d), n, atol=2e-3)


def _ao_uniforms(W, H):
    from frostbitten_hair_webgpu_amd import layouts as L
    from frostbitten_hair_webgpu_amd import matrices as M
    import math
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    proj = M.perspective(math.radians(30.0), W / H, 0.01, 100.0)
    u["projMatrixInv"] = np.linalg.inv(proj).T.astype(np.float32).reshape(16)
    u["viewMatrix"] = EYE
    u["viewport"] = [W, H, 0, 0]
    u["aoParams"] = [2.0, 0.0, 0.16, 4.0]  # radius, directionOffset, falloffStart2, falloffEnd2
    u["aoNumDirections"], u["aoNumSteps"], u["aoStrength"] = 12, 8, 0.4
    return u, proj


def test_gtao_background_and_flat_wall(oracle):
    """Background (depth 1, beyond 0.999 far) gives 0. A wall at constant
    depth whose normal faces the camera has its horizons on the wall
    (cos h = 0 near the view axis, gamma = 0): the integral is 1 there."""
    W, H = 64, 48
    u, proj = _ao_uniforms(W, H)
    ao_w, ao_h = 32, 24
    # octahedral (0.5, 0.5) = +z; view normal z is negated (gtao.wgsl.ts:35-37) -> facing the camera
    n = np.full((H, W), np.array([0.5, 0.5], np.float16).view(np.uint32)[0], np.uint32)
    bg = oracle.calc_ao(u, np.ones((H, W), np.float32), n, ao_w, ao_h)
    assert (bg == 0).all()
    z_view = 2.0
    clip = proj @ np.array([0, 0, -z_view, 1.0])
    wall = np.full((H, W), clip[2] / clip[3], np.float32)
    ao = oracle.calc_ao(u, wall, n, ao_w, ao_h)
    centre = ao[ao_h // 2 - 2: ao_h // 2 + 2, ao_w // 2 - 2: ao_w // 2 + 2]
    assert np.isfinite(ao).all()
    assert np.allclose(centre, 1.0, atol=0.05), centre
    # a pit (nearer rim around a deeper floor) darkens the floor next to the rim
    clip2 = proj @ np.array([0, 0, -(z_view - 0.3), 1.0])
    pit = np.full((H, W), clip2[2] / clip2[3], np.float32)
    pit[16:32, 24:40] = wall[0, 0]
    ao_pit = oracle.calc_ao(u, pit, n, ao_w, ao_h)
    assert ao_pit[8, 12] < centre.mean() - 0.1
