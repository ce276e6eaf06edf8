"""The oracle's background, mesh and mesh-shadow passes against closed forms.

The reference draws these with the hardware rasterizer and fragment shaders
(drawBackgroundGradientPass, drawMeshesPass, shadowMapPass's mesh draw). The
rasterization rules are the restatement's own (those of the hair passes,
pinned in test_oracle_hwraster.py); these tests pin what the passes add: the
background's hash, noise and gradient, back-face culling only in the camera
pass, primitive order across meshes, the collider preview's matrices and
colour, disneyPBR against a float64 evaluation of the formulas it states
(Karis13: GGX distribution, Smith-Schlick geometry with k = (r + 1)^2 / 8,
Schlick Fresnel; Lambert without 1/pi), the PCF count on a step shadow map,
and the shadow-strength and AO terms of doShading.
"""
import math

import numpy as np
import pytest

EYE = np.eye(4, dtype=np.float32).reshape(16)


def _u(W, H, **kw):
    from frostbitten_hair_webgpu_amd import layouts as L
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    for k in ("vpMatrix", "viewMatrix", "projMatrix", "projMatrixInv", "modelMatrix", "modelViewMat", "mvpMatrix",
              "collisionSphereModelMatrix", "shadowSourceModelViewMat", "shadowSourceProjMatrix", "shadowSourceMVP"):
        u[k] = EYE
    u["viewport"] = [W, H, 0, 0]
    u["cameraPosition"] = [0, 0, -5, 0]
    u["lightAmbient"] = [1, 1, 1, 0.05]
    u["shadowPCFRadius"] = 1
    u["shadowBias"] = 0.0005
    u["shadowStrength"] = 1.0
    for k, v in kw.items():
        u[k] = v
    return u


def _quad(x0, y0, x1, y1, z, normal=(0.0, 0.0, -1.0), ccw=True, preview=False):
    """A rectangle in NDC (identity matrices) as two triangles, counter-clockwise
    seen with y up when ccw."""
    p = np.array([[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]], np.float32)
    n = np.tile(np.asarray(normal, np.float32), (4, 1))
    i = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    if not ccw:
        i = i[:, ::-1].copy()
    return {"positions": p, "normals": n, "indices": i, "preview": preview}


def _h2f(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def _camera(oracle, W, H, u, meshes, sm=None, ao=None, hdr=None):
    sm = np.ones((8, 8), np.float32) if sm is None else sm
    ao = np.ones(((H + 1) // 2, (W + 1) // 2), np.float32) if ao is None else ao
    hdr = np.zeros((H, W, 4), np.uint16) if hdr is None else hdr
    return oracle.meshes(False, W, H, u, meshes, shadow_map=sm, ao=ao, hdr=hdr)


# --------------------------------------------------------------------------
# background
# --------------------------------------------------------------------------

def _hash64(x, y):
    """hash in float64, with a bound on what f32 evaluation moves it by: the
    rounding of 0.3183099 and of each operation of fract(p * c + k), times 50,
    carried through the cubic product, plus a few ulps of the product."""
    c = 0.3183099
    dc = abs(c - float(np.float32(c)))
    ulp = lambda t: float(np.spacing(np.float32(abs(t)) + np.float32(1e-30)))  # noqa: E731
    qx, qy = x * c + 0.71, y * c + 0.113
    px, py = 50.0 * (qx % 1.0), 50.0 * (qy % 1.0)
    ex = 50.0 * (abs(x) * dc + ulp(x * c) + ulp(qx) + ulp(qx % 1.0)) + ulp(px)
    ey = 50.0 * (abs(y) * dc + ulp(y * c) + ulp(qy) + ulp(qy % 1.0)) + ulp(py)
    v = px * py * (px + py)
    dvx, dvy = py * (px + py) + px * py, px * (px + py) + px * py
    err = dvx * ex + dvy * ey + 4.0 * ulp(v)
    return -1.0 + 2.0 * (v % 1.0), v, 2.0 * err


@pytest.mark.parametrize("x,y", [(0.0, 0.0), (1.0, 3.0), (1.1, 2.9), (1.2, 3.05), (-2.0, -0.3), (-2.1, 6.0)])
def test_bg_hash_against_float64(oracle, x, y):
    """hash (noise.wgsl.ts:14-18) against float64 within the f32 rounding
    bound (a wrong constant moves it by far more)."""
    h64, v, tol = _hash64(x, y)
    frac = v % 1.0
    if min(frac, 1.0 - frac) < tol or tol > 0.05:
        pytest.skip("f32 rounding bound too wide at this point")
    h = oracle.bg_hash(x, y)
    assert -1.0 <= h <= 1.0
    assert abs(h - h64) <= tol, (h, h64, tol)


def test_bg_noise_lattice_and_quintic(oracle):
    """noise (noise.wgsl.ts:21-33): the hash itself at lattice points (the
    interpolant is 0 there), and inside a cell the bilinear mix of the four
    corner hashes with the quintic weights 6f^5 - 15f^4 + 10f^3."""
    for i, j in [(0, 0), (2, -1), (-3, 4)]:
        assert oracle.bg_noise(float(i), float(j)) == oracle.bg_hash(float(i), float(j))
    q = lambda f: f ** 3 * (f * (f * 6 - 15) + 10)  # noqa: E731
    for i, j, fx, fy in [(0, 0, 0.25, 0.75), (1, 2, 0.5, 0.5), (-2, 1, 0.125, 0.375)]:
        h = {(a, b): oracle.bg_hash(float(i + a), float(j + b)) for a in (0, 1) for b in (0, 1)}
        ux, uy = q(fx), q(fy)
        a = h[0, 0] * (1 - ux) + h[1, 0] * ux
        b = h[0, 1] * (1 - ux) + h[1, 1] * ux
        exp = a * (1 - uy) + b * uy
        assert abs(oracle.bg_noise(i + fx, j + fy) - exp) < 2e-6, (i, j, fx, fy)


@pytest.mark.parametrize("x,y,scale", [(0.3, 0.7, 5.0), (0.91, 0.12, 13.0), (0.5, 0.5, 40.0)])
def test_bg_fractal_octaves(oracle, x, y, scale):
    """fractalNoise (noise.wgsl.ts:35-42): octaves of weight 1/2, 1/4, 1/8,
    1/16 at uv, M uv, M^2 uv, M^3 uv with uv = p * scale and M the rotation
    and scaling mat2x2f(1.6, 1.2, -1.2, 1.6) (columns), i.e. (x, y) ->
    (1.6 x - 1.2 y, 1.2 x + 1.6 y). The points are rotated in float64 here,
    so the noise is sampled a rounding away from the oracle's points."""
    uv = np.array([x * scale, y * scale])
    M = np.array([[1.6, -1.2], [1.2, 1.6]])
    exp = 0.0
    for w in (0.5, 0.25, 0.125, 0.0625):
        exp += w * oracle.bg_noise(float(uv[0]), float(uv[1]))
        uv = M @ uv
    assert abs(oracle.bg_fractal(x, y, scale) - exp) < 2e-4


def test_background_gradient_closed_form(oracle):
    """With noiseScale 0 every octave samples noise(0, 0) = hash(0, 0), so
    c = 0.9375 hash(0, 0) / 2 + 1/2 everywhere, and the colour is
    mix(mix(color0, color1, c), 1 - uv.y, gradientStrength) with uv.y the
    pixel centre's row from the top."""
    W, H = 8, 6
    c0, c1, g = (0.1, 0.4, 0.7), (0.9, 0.2, 0.3), 0.25
    u = _u(W, H, bgColor0=c0, bgColor1=c1, bgNoiseScale=0.0, bgGradientStrength=g)
    out = _h2f(oracle.background(W, H, u))
    c = 0.9375 * oracle.bg_hash(0.0, 0.0) * 0.5 + 0.5
    for py in range(H):
        col = (1 - c) * np.asarray(c0) + c * np.asarray(c1)
        col = col * (1 - g) + (1.0 - (py + 0.5) / H) * g
        assert np.allclose(out[py, :, :3], col, rtol=2e-3, atol=1e-4), py
    assert (out[..., 3] == 1.0).all()
    # with a scale the noise varies over the image
    u = _u(W, H, bgColor0=c0, bgColor1=c1, bgNoiseScale=40.0, bgGradientStrength=0.0)
    out = _h2f(oracle.background(W, H, u))
    assert np.unique(out[..., 0]).size > 10


# --------------------------------------------------------------------------
# meshes: rasterization, culling, order
# --------------------------------------------------------------------------

def test_quad_pixels_depth_normals(oracle):
    """x, y in [-0.5, 0.5] on a 16x16 target covers pixels [4, 12) in both
    axes (edges on pixel borders, centres off them); depth is the plane's z;
    normals are the octahedral encoding of the vertex normal; pixels without
    a fragment keep the clears (depth 1, normals 0) and the colour target."""
    W = H = 16
    u = _u(W, H, lights=np.zeros((3, 8), np.float32))
    hdr = np.full((H, W, 4), 0x3c00, np.uint16)  # 1.0 everywhere
    d, n, c = _camera(oracle, W, H, u, [_quad(-0.5, -0.5, 0.5, 0.5, 0.3)], hdr=hdr)
    exp = np.zeros((H, W), bool)
    exp[4:12, 4:12] = True
    assert np.array_equal(d < 1, exp)
    assert np.allclose(d[exp], 0.3, rtol=0, atol=1e-7)
    assert (n[exp] == oracle.encode_oct(0.0, 0.0, -1.0)).all() and (n[~exp] == 0).all()
    # lights off: the colour is the ambient term 0.05, alpha 1
    col = _h2f(c)
    assert np.allclose(col[exp][:, :3], np.float16(0.05), rtol=0, atol=0) and (col[exp][:, 3] == 1).all()
    assert (c[~exp] == 0x3c00).all()


def test_perspective_correct_attributes(oracle):
    """Vertex attributes are interpolated perspective-correctly (weights
    l_i / w_i, normalised). The planar quad (-0.5, -0.5, 1), (0.5, -0.5, 1),
    (1.5, 1.5, 3), (-1.5, 1.5, 3) under clip = (x, y, z / 2, z) covers the
    NDC square [-0.5, 0.5]^2; its normals are (1, 0, 0) at z = 1 and
    (0, 1, 0) at z = 3. At NDC row Y the quad's parameter along its depth is
    t = (Y + 1/2) / (2 (1 - Y)), so the octahedral halves of the normal
    normalize((1 - t, t, 0)) are (1 - t / 2, 1/2 + t / 2); screen-space
    interpolation would give t = Y + 1/2 instead."""
    W = H = 16
    R = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5, 0], [0, 0, 1, 0]], np.float32)
    u = _u(W, H, mvpMatrix=R.T.reshape(16))
    m = {"positions": np.array([[-0.5, -0.5, 1], [0.5, -0.5, 1], [1.5, 1.5, 3], [-1.5, 1.5, 3]], np.float32),
         "normals": np.array([[1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 0]], np.float32),
         "indices": np.array([[0, 1, 2], [0, 2, 3]], np.uint32)}
    d, n, _ = _camera(oracle, W, H, u, [m])
    assert (d[4:12, 4:12] < 1).all()
    for py in range(4, 12):
        Y = 1.0 - (py + 0.5) / 8.0
        t = (Y + 0.5) / (2.0 * (1.0 - Y))
        r = _h2f((n[py, 4:12] & 0xffff).astype(np.uint16))
        g = _h2f((n[py, 4:12] >> 16).astype(np.uint16))
        assert np.allclose(r, 1.0 - t / 2, atol=1e-3) and np.allclose(g, 0.5 + t / 2, atol=1e-3), py


def test_perspective_correct_position_and_shadow_space(oracle):
    """The same quad lit by one light and shadowed through a shadow map of
    constant depth 0.5 under the shadow MVP diag(1/4, 1/4, 1/4, 1): a pixel
    is in shadow when its interpolated shadow-space z = z_world / 4 exceeds
    0.5 (t > 1/2, NDC rows Y > 1/4; screen-space interpolation would shadow
    Y > 0 instead). Lit pixels have ambient + disneyPBR at the perspective-
    correct world position x = X z, y = -1/2 + 2t, z = 1 + 2t (float64
    here); shadowed ones (strength 1) the ambient term alone."""
    W = H = 16
    R = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5, 0], [0, 0, 1, 0]], np.float32)
    lights = np.zeros((3, 8), np.float32)
    light = np.array([0.3, -2.0, -1.0])
    lights[0] = [*light, 0, 1.0, 0.9, 0.8, 2.0]
    cam = np.array([0.0, 0.0, -2.0])
    u = _u(W, H, mvpMatrix=R.T.reshape(16), shadowSourceMVP=np.diag([0.25, 0.25, 0.25, 1]).astype(np.float32)
           .reshape(16), shadowPCFRadius=0, shadowStrength=1.0, lights=lights, cameraPosition=[*cam, 0],
           shadowSourcePosition=[0, 0, -10, 0])
    m = {"positions": np.array([[-0.5, -0.5, 1], [0.5, -0.5, 1], [1.5, 1.5, 3], [-1.5, 1.5, 3]], np.float32),
         "normals": np.array([[1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 0]], np.float32),
         "indices": np.array([[0, 1, 2], [0, 2, 3]], np.uint32)}
    sm = np.full((16, 16), 0.5, np.float32)
    _, _, c = _camera(oracle, W, H, u, [m], sm=sm)
    col = _h2f(c)[..., :3]
    for py in range(4, 12):
        Y = 1.0 - (py + 0.5) / 8.0
        t = (Y + 0.5) / (2.0 * (1.0 - Y))
        for px in range(4, 12):
            X = (px + 0.5) / 8.0 - 1.0
            z = 1.0 + 2.0 * t
            P = np.array([X * z, -0.5 + 2.0 * t, z])
            N = np.array([1.0 - t, t, 0.0]) / math.hypot(1.0 - t, t)
            V = (cam - P) / np.linalg.norm(cam - P)
            exp = np.full(3, 0.05)
            if t <= 0.5:
                exp = exp + _pbr64(P, N, V, (0.9, 0.9, 0.9), 0.8, 0.0, light, (1.0, 0.9, 0.8), 2.0)
            assert np.allclose(col[py, px], exp, rtol=3e-3, atol=1e-4), (py, px, col[py, px], exp)


def test_back_faces_culled_in_camera_pass_only(oracle):
    W = H = 16
    u = _u(W, H)
    cw = _quad(-0.5, -0.5, 0.5, 0.5, 0.3, ccw=False)
    d, n, _ = _camera(oracle, W, H, u, [cw])
    assert (d == 1).all() and (n == 0).all()
    sm = oracle.meshes(True, W, H, u, [cw])
    assert (sm < 1).sum() == 64


def test_primitive_order_across_meshes(oracle):
    """The same triangles twice (equal depths, bit for bit): the earlier mesh
    keeps every pixel ('less'); a later mesh nearer wins where it covers."""
    W = H = 16
    u = _u(W, H)
    a = _quad(-0.5, -0.5, 0.5, 0.5, 0.3, normal=(0.0, 0.0, -1.0))
    same = _quad(-0.5, -0.5, 0.5, 0.5, 0.3, normal=(1.0, 0.0, 0.0))
    na, nb = oracle.encode_oct(0.0, 0.0, -1.0), oracle.encode_oct(1.0, 0.0, 0.0)
    d, n, _ = _camera(oracle, W, H, u, [a, same])
    assert (n[4:12, 4:12] == na).all()
    d, n, _ = _camera(oracle, W, H, u, [same, a])
    assert (n[4:12, 4:12] == nb).all()
    b = _quad(-0.25, -0.25, 0.75, 0.75, 0.2, normal=(1.0, 0.0, 0.0))
    d, n, _ = _camera(oracle, W, H, u, [a, b])
    # b covers rows 2..9 (from the top) and columns 6..13
    assert (n[2:10, 6:14] == nb).all() and np.allclose(d[2:10, 6:14], 0.2, atol=1e-7)
    assert (n[10:12, 4:12] == na).all() and (n[4:10, 4:6] == na).all()


def test_collider_preview(oracle):
    """The preview is drawn with proj * (view * collisionSphereModelMatrix)
    (getMVP_Mat), shaded white, rough, 0.4 metallic, unshadowed, without AO,
    and mixed halfway to yellow; the shadow pass skips it. Model: x + 0.25;
    view: y + 0.25; proj: scale 1/2. The quad [-0.5, 0.5]^2 lands on
    [-0.125, 0.375]^2 in NDC, rows 5..8 and columns 7..10 of 16x16 (the
    other order of the product would put it on [0, 0.5]^2)."""
    W = H = 16
    model = np.eye(4, dtype=np.float32)
    model[3, 0] = 0.25  # column-major: element 12 is the x translation
    view = np.eye(4, dtype=np.float32)
    view[3, 1] = 0.25
    proj = np.diag([0.5, 0.5, 1.0, 1.0]).astype(np.float32)
    u = _u(W, H, collisionSphereModelMatrix=model.reshape(16), viewMatrix=view.reshape(16),
           projMatrix=proj.reshape(16), lights=np.zeros((3, 8), np.float32))
    m = _quad(-0.5, -0.5, 0.5, 0.5, 0.3, preview=True)
    d, n, c = _camera(oracle, W, H, u, [m])
    exp = np.zeros((H, W), bool)
    exp[5:9, 7:11] = True
    assert np.array_equal(d < 1, exp)
    col = _h2f(c)[exp][:, :3]
    assert np.allclose(col, np.float16([0.525, 0.525, 0.025]).astype(np.float32), atol=0)
    assert (oracle.meshes(True, W, H, u, [m]) == 1).all()


def test_shadow_pass_uses_source_mvp(oracle):
    W = H = 16
    scale = np.diag([0.5, 0.5, 1.0, 1.0]).astype(np.float32).reshape(16)
    u = _u(W, H, shadowSourceMVP=scale)
    sm = oracle.meshes(True, W, H, u, [_quad(-0.5, -0.5, 0.5, 0.5, 0.3)])
    exp = np.zeros((H, W), bool)
    exp[6:10, 6:10] = True
    assert np.array_equal(sm < 1, exp) and np.allclose(sm[exp], 0.3, atol=1e-7)


# --------------------------------------------------------------------------
# shading
# --------------------------------------------------------------------------

def _pbr64(P, N, V, albedo, rough, metal, light_pos, color, energy):
    P, N, V = (np.asarray(x, np.float64) for x in (P, N, V))
    L = np.asarray(light_pos, np.float64) - P
    L /= np.linalg.norm(L)
    H = (V + L) / np.linalg.norm(V + L)
    albedo = np.asarray(albedo, np.float64)
    F0 = 0.04 * (1 - metal) + albedo * metal
    F = F0 + (1 - F0) * (1 - max(0.0, H @ V)) ** 5
    nv, nl, nh = max(0.0, N @ V), max(0.0, N @ L), max(0.0, N @ H)
    k = (rough + 1) ** 2 / 8
    G = (nl / (nl * (1 - k) + k)) * (nv / (nv * (1 - k) + k))
    a2 = rough ** 4
    D = a2 / (math.pi * (nh * nh * (a2 - 1) + 1) ** 2)
    spec = D * G * F / max(4 * nv * nl, 0.001)
    kD = (1 - F) * (1 - metal)
    return (kD * albedo + spec) * np.asarray(color, np.float64) * energy * nl


@pytest.mark.parametrize("theta_l,theta_v,rough,metal", [
    (0.3, -0.4, 0.8, 0.0), (1.1, 0.2, 0.8, 0.0), (0.5, 0.5, 0.35, 0.0), (0.7, -1.2, 1.0, 0.4), (0.2, 0.9, 0.6, 1.0)])
def test_disney_pbr_against_float64(oracle, theta_l, theta_v, rough, metal):
    P = np.array([0.1, -0.2, 0.3])
    N = np.array([0.0, 0.0, 1.0])
    V = np.array([math.sin(theta_v), 0.1, math.cos(theta_v)])
    V /= np.linalg.norm(V)
    Lp = P + 2.0 * np.array([math.sin(theta_l), 0.3, math.cos(theta_l)])
    albedo, color, energy = (0.9, 0.7, 0.5), (1.0, 0.8, 0.6), 1.7
    got = oracle.disney_pbr(P, N, V, albedo, rough, metal, [*Lp, 0.0], [*color, energy])
    exp = _pbr64(P, N, V, albedo, rough, metal, Lp, color, energy)
    assert np.allclose(got, exp, rtol=3e-5, atol=1e-7), (got, exp)


def test_disney_pbr_back_lit_is_black(oracle):
    got = oracle.disney_pbr([0, 0, 0], [0, 0, 1], [0, 0.6, 0.8], (0.9, 0.9, 0.9), 0.8, 0.0, [0, 0, -2, 0], [1, 1, 1, 1])
    assert (got == 0).all()


@pytest.mark.parametrize("i0,radius", [(20, 1), (29, 1), (30, 2), (31, 2), (33, 3), (40, 1)])
def test_pcf_count_on_step_map(oracle, i0, radius):
    """A 64^2 map at depth 0.2 for texel columns < 32 and 1.0 from 32 on; a
    fragment at depth 0.5 whose map coordinate is texel i0 + 0.75 (so tap k
    gathers columns i0 + k and i0 + k + 1): a tap is in shadow when both
    columns are < 32, so the PCF result is #{k in [-r, r]: i0 + k + 1 < 32}
    / (2r + 1)."""
    N = 64
    sm = np.ones((N, N), np.float32)
    sm[:, :32] = 0.2
    u = _u(64, 64, shadowPCFRadius=radius, shadowBias=0.0005, shadowSourcePosition=[0, 0, -10, 0])
    pss = [(i0 + 0.75) / N, (20 + 0.75) / N, 0.5]
    got = oracle.directional_shadow(u, sm, [0, 0, 0], [0, 0, -1], pss)
    k = np.arange(-radius, radius + 1)
    assert got == pytest.approx(((i0 + k + 1) < 32).sum() / (2 * radius + 1), abs=1e-7)
    # outside the map or beyond the far plane: not in shadow
    assert oracle.directional_shadow(u, sm, [0, 0, 0], [0, 0, -1], [1.2, 0.3, 0.5]) == 0.0
    assert oracle.directional_shadow(u, sm, [0, 0, 0], [0, 0, -1], [0.1, 0.3, 1.5]) == 0.0


def test_bias_grows_with_angle(oracle):
    """actualBias = max(|bias| (1 - N.toSource), 0.0005): a fragment 0.01
    behind a flat map at depth 0.5 is lit once the bias exceeds 0.01."""
    N = 32
    sm = np.full((N, N), 0.5, np.float32)
    pss = [0.5 + 0.25 / N, 0.5 + 0.25 / N, 0.51]
    u = _u(N, N, shadowPCFRadius=0, shadowBias=0.02, shadowSourcePosition=[0, 0, -10, 0])
    assert oracle.directional_shadow(u, sm, [0, 0, 0], [0, 0, -1], pss) == 1.0  # facing: bias 0.0005
    assert oracle.directional_shadow(u, sm, [0, 0, 0], [1, 0, 0], pss) == 0.0   # grazing: bias 0.02


def test_shadow_strength_and_ao_terms(oracle):
    """doShading: radiance x saturate(ao) x clamp(shadow, 1 - strength, 1),
    plus ambient. One light, ambient 0: a fully shadowed quad keeps
    1 - strength of its radiance, and AO 0.5 at strength 0.5
    (mix(1, ao, strength) = 0.75) keeps three quarters."""
    W = H = 16
    lights = np.zeros((3, 8), np.float32)
    lights[0] = [0, 0, -3, 0, 1, 1, 1, 1]
    base = dict(lights=lights, lightAmbient=[1, 1, 1, 0], shadowSourcePosition=[0, 0, -10, 0],
                shadowSourceMVP=EYE)
    q = [_quad(-0.5, -0.5, 0.5, 0.5, 0.3)]
    lit = np.ones((16, 16), np.float32)
    dark = np.zeros((16, 16), np.float32)
    _, _, c_lit = _camera(oracle, W, H, _u(W, H, shadowStrength=0.4, aoStrength=0.0, **base), q, sm=lit)
    _, _, c_dark = _camera(oracle, W, H, _u(W, H, shadowStrength=0.4, aoStrength=0.0, **base), q, sm=dark)
    ao = np.full((8, 8), 0.5, np.float32)
    _, _, c_ao = _camera(oracle, W, H, _u(W, H, shadowStrength=0.4, aoStrength=0.5, **base), q, sm=lit, ao=ao)
    m = np.zeros((H, W), bool)
    m[4:12, 4:12] = True
    a, b, c = (_h2f(x)[m][:, 0] for x in (c_lit, c_dark, c_ao))
    assert (a > 0.05).all()
    assert np.allclose(b, 0.6 * a, rtol=2e-3)
    assert np.allclose(c, 0.75 * a, rtol=2e-3)


def test_scene_meshes_closed_and_outward():
    """assets.scene_meshes: every mesh is closed (each edge used once in each
    direction, seam vertices merged by position) and wound counter-clockwise
    seen from outside (positive signed volume)."""
    from collections import Counter

    from frostbitten_hair_webgpu_amd import assets as A
    ms = A.scene_meshes(A._head_collider("union"), (0.0, 0.0, 0.0, 0.1), rings=8, segments=12)
    assert [m.get("preview", False) for m in ms] == [False, False, False, True]
    for m in ms:
        p, i = m["positions"].astype(np.float64), m["indices"]
        key = {}
        rep = np.array([key.setdefault(tuple(np.round(v, 6)), k) for k, v in enumerate(p)])
        e = Counter()
        for t in rep[i]:
            for k in range(3):
                e[(int(t[k]), int(t[(k + 1) % 3]))] += 1
        assert all(e.get((b, a), 0) == c for (a, b), c in e.items())
        vol = sum(np.dot(p[a], np.cross(p[b], p[c])) for a, b, c in i) / 6.0
        assert vol > 0
        assert np.allclose(np.linalg.norm(m["normals"], axis=1), 1.0, atol=1e-6)
