"""Pins of the oracle's shading pass (K8) from the mathematics.

The reference holds no fixture for hairShadingPass.wgsl.ts, so each test puts
one shading point in a chosen geometry (identity model, camera and shadow
matrices, lights at chosen directions) where the result has a closed form:

- Kajiya-Kay diffuse times the fake multiple scattering
  (hairShadingPass.wgsl.ts:166-234) at chosen angles, and zero when the light
  is behind the tangent (the saturate of :186);
- each Marschner lobe alone (marschner.wgsl.ts:34-97, Karis16 slides 18-32),
  evaluated in float64 at chosen angles, including the special angles where
  sin(theta_i) = sin(theta_r) = 0 and the R lobe reduces to
  specular * exp(-(2 shift)^2 / (2 beta^2)) / (beta sqrt(2 pi)) * F0 cos(phi/2) / 4;
- the PCF shadow (sampleShadows.wgsl.ts:47-132) over a step shadow map whose
  hit count is known column by column, with PCF and with PCSS;
- the AO texel choice and strength (textureAo.wgsl.ts:7-12,
  hairShadingPass.wgsl.ts:131-133) and the attenuation from linearizeDepth's
  closed form (linearDepth.wgsl.ts:9-27, hairShadingPass.wgsl.ts:259-282);
- the per-strand colour terms (randomRGB noise.wgsl.ts:45-51, the root-to-tip
  mix, the luma jitter) and the thin-tip alpha.

Each is chosen so that a dropped term, a swapped operand or a wrong sign in
the oracle fails it.
"""
import math

import numpy as np
import pytest

VP = 64  # viewport (square) and the AO / depth texture sizes below
HAIR_F0 = ((1.0 - 1.55) / (1.0 + 1.55)) ** 2


def _cfg(ps=2):
    from frostbitten_hair_webgpu_amd import layouts as L
    return L.RenderConfig(shadingPoints=ps)


def _uniforms(cam, lights, material, shadow=None, ao_strength=0.0):
    from frostbitten_hair_webgpu_amd import layouts as L
    u = np.zeros(1, dtype=L.RENDER_UNIFORMS_DTYPE)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    for k in ("vpMatrix", "modelMatrix", "shadowSourceMVP"):
        u[k] = eye
    u["viewport"] = [VP, VP, 0, 0]
    u["cameraPosition"] = [*cam, 0]
    ls = np.zeros((3, 8), np.float32)
    for i, (pos, color, energy) in enumerate(lights):
        ls[i] = [*pos, 0, *color, energy]
    u["lights"] = ls
    m = dict(color0=(0.8, 0.5, 0.2), color1=(0.8, 0.5, 0.2), specular=0.0, weightTT=0.0, weightTRT=0.0,
             shift=0.0, roughness=0.3, colorRng=0.0, lumaRng=0.0, attenuation=0.0, shadows=0.0)
    m.update(material)
    for k, f in (("color0", "hairColor0"), ("color1", "hairColor1"), ("specular", "hairSpecular"),
                 ("shadows", "hairShadows"), ("weightTT", "hairWeightTT"), ("weightTRT", "hairWeightTRT"),
                 ("shift", "hairShift"), ("roughness", "hairRoughness"), ("colorRng", "hairColorRng"),
                 ("lumaRng", "hairLumaRng"), ("attenuation", "hairAttenuation")):
        u[f] = m[k]
    s = dict(usePCSS=0, radius=3, bias=0.0, position=(0.0, 0.0, 10.0))
    s.update(shadow or {})
    u["shadowUsePCSS"] = s["usePCSS"]
    u["shadowPCFRadius"] = s["radius"]
    u["shadowBias"] = s["bias"]
    u["shadowSourcePosition"] = [*s["position"], 0]
    u["aoStrength"] = ao_strength
    return u


def _strand(p0, tangent, strands=1, ps=2):
    """`strands` copies of a 2-point strand from p0 along `tangent`; with
    ps = 2, shading point 0 sits at p0 with tangent normalize(p1 - p0)."""
    p0 = np.asarray(p0, np.float32)
    t = np.asarray(tangent, np.float64)
    t = t / np.linalg.norm(t)
    pos = np.zeros((strands, 2, 4), np.float32)
    pos[:, 0, :3] = p0
    pos[:, 1, :3] = p0 + (0.1 * t).astype(np.float32)
    pos[:, :, 3] = 1.0
    from frostbitten_hair_webgpu_amd import assets as A
    pos = pos.reshape(-1, 4)
    return pos, A.tangents(pos, strands, 2)


def _shade(oracle, u, pos, tan, strands=1, ps=2, shadow_map=None, ao=None, depth=None):
    shadow_map = np.zeros((VP, VP), np.float32) if shadow_map is None else shadow_map
    ao = np.ones((VP // 2, VP // 2), np.float32) if ao is None else ao
    depth = np.ones((VP, VP), np.float32) if depth is None else depth
    out = oracle.shading(_cfg(ps), u, pos, tan, shadow_map, ao, depth, strands, 2)
    return out.reshape(strands, ps, 4).astype(np.float64)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _diffuse(base, T, L, V, shadow=1.0):
    """Kajiya-Kay diffuse times the fake multiple scattering, one white light
    of energy 1 (hairShadingPass.wgsl.ts:176-234)."""
    base = np.asarray(base, np.float64)
    TL = float(np.dot(T, L))
    kk = (0.25 + 0.75 * TL) * base
    n = _unit(V - T * np.dot(V, T))
    nol = min(max(float(np.dot(n, L)), 0.0), 1.0)
    luma = 0.2126729 * base[0] + 0.7151522 * base[1] + 0.0721750 * base[2]
    ms = np.sqrt(base) * (nol + 1.0) / (4.0 * math.pi) * (base / luma) ** (1.0 - shadow)
    return kk * ms * min(max(TL, 0.0), 1.0)


def _lobes(base, T, L, V, shift, rough):
    """The three Marschner lobes (Karis16 as stated in marschner.wgsl.ts:34-170), float64."""
    base = np.asarray(base, np.float64)
    si, sr = float(np.dot(T, L)), float(np.dot(T, V))
    ci, cr = math.sqrt(max(0.0, 1 - si * si)), math.sqrt(max(0.0, 1 - sr * sr))
    cd = math.sqrt(min(max(0.5 + 0.5 * (ci * cr + si * sr), 0.0), 1.0))
    lp, cp = L - si * T, V - sr * T
    cphi = float(np.dot(lp, cp)) / math.sqrt(np.linalg.norm(lp) * np.linalg.norm(cp) + 1e-4)
    chp = math.sqrt(min(max(0.5 + 0.5 * cphi, 0.0), 1.0))

    def M(alpha, beta):
        a = si + sr - alpha
        return math.exp(-a * a / (2 * beta * beta)) / (beta * math.sqrt(2 * math.pi))

    def F(c):
        return HAIR_F0 + (1 - HAIR_F0) * (1 - c) ** 5

    b = rough * rough
    R = M(-2 * shift, b) * 0.25 * chp * F(cd)
    eta = 1.19 / cd + 0.36 * cd
    a = 1 / eta
    h = (1 + a * (0.6 - 0.8 * cphi)) * chp
    TT = (base ** (math.sqrt(1 - h * h * a * a) / (2 * cd)) * math.exp(-3.65 * cphi - 3.98)
          * (1 - F(cd * math.sqrt(min(max(1 - h * h, 0.0), 1.0)))) ** 2 * M(shift, 0.5 * b))
    f = F(cd * 0.5)
    TRT = base ** (0.8 / cd) * math.exp(17 * cphi - 16.78) * (1 - f) ** 2 * f * M(4 * shift, 2 * b)
    return np.array([R] * 3), TT, TRT


P0 = np.array([0.0390625, 0.3, 0.6])  # shadow-map u = 0.51953125, AO/depth texels below
CAM = P0 + [0.0, 0.0, 2.0]  # toCamera = +z
WHITE = (1.0, 1.0, 1.0)
OFF = P0 + [0.0, 5.0, 0.0]  # the unused lights (energy 0) sit away from the point


def test_kajiya_kay_diffuse_and_multiple_scattering(oracle):
    """Diffuse only (no lobes, unshadowed, no AO): with T = +x, V = +z and the
    light at 45 degrees in the xz plane, T.L = N.L = 1/sqrt(2)."""
    T = np.array([1.0, 0.0, 0.0])
    light = P0 + [1.0, 0.0, 1.0]
    u = _uniforms(CAM, [(light, WHITE, 0.8), (light, WHITE, 0.0), (light, WHITE, 0.0)], {})
    pos, tan = _strand(P0, T)
    got = _shade(oracle, u, pos, tan)[0, 0]
    exp = 0.8 * _diffuse((0.8, 0.5, 0.2), T, _unit([1, 0, 1]), np.array([0, 0, 1.0]))
    assert got[:3] == pytest.approx(exp, rel=2e-6)
    assert got[3] == 1.0
    # light colour and energy scale the radiance per channel; lights add
    u2 = _uniforms(CAM, [(light, (1.0, 0.5, 0.25), 0.8), (light, WHITE, 0.4), (light, WHITE, 0.2)], {})
    got2 = _shade(oracle, u2, pos, tan)[0, 0]
    assert got2[:3] == pytest.approx(exp * (np.array([1.0, 0.5, 0.25]) + 0.75), rel=2e-6)
    # light behind the tangent: the diffuse term is cut by saturate(T.L)
    u3 = _uniforms(CAM, [(P0 + [-1.0, 0.0, 1.0], WHITE, 0.8), (light, WHITE, 0.0), (light, WHITE, 0.0)], {})
    assert np.all(_shade(oracle, u3, pos, tan)[0, 0, :3] == 0.0)


def test_marschner_R_lobe_at_normal_incidence(oracle):
    """sin(theta_i) = sin(theta_r) = 0 (L and V perpendicular to T, 60 degrees
    apart): cos(theta_d) = 1, the Fresnel term is F0 and the R lobe is
    specular * exp(-(2 shift)^2 / (2 beta^2)) / (beta sqrt(2 pi)) * F0 cos(phi/2) / 4,
    beta = roughness^2. T.L = 0, so the diffuse term vanishes."""
    T = np.array([1.0, 0.0, 0.0])
    L = np.array([0.0, math.sin(math.radians(60)), math.cos(math.radians(60))])
    for shift, rough, spec in ((0.0, 0.3, 1.0), (0.05, 0.4, 0.7)):
        u = _uniforms(CAM, [(P0 + 3 * L, WHITE, 1.0), (OFF, WHITE, 0.0), (OFF, WHITE, 0.0)],
                      dict(specular=spec, shift=shift, roughness=rough))
        pos, tan = _strand(P0, T)
        got = _shade(oracle, u, pos, tan)[0, 0, :3]
        beta = rough * rough
        cphi = 0.5 / math.sqrt(1.0 + 1e-4)
        exp = (spec * math.exp(-(2 * shift) ** 2 / (2 * beta * beta)) / (beta * math.sqrt(2 * math.pi))
               * 0.25 * math.sqrt(0.5 + 0.5 * cphi) * HAIR_F0)
        assert got == pytest.approx([exp] * 3, rel=2e-5)


@pytest.mark.parametrize("lobe", ["R", "TT", "TRT"])
def test_marschner_lobes_at_oblique_angles(oracle, lobe):
    """Each lobe alone at sin(theta_i) = -0.5 (light behind along the tangent,
    so saturate(T.L) = 0 removes the diffuse term) and a view off the normal
    plane, against the Karis16 formulas in float64. A non-zero cuticle shift
    makes the longitudinal terms depend on the sign of alpha."""
    T = _unit([1.0, 0.0, 0.0])
    L = _unit([-0.5, math.sqrt(0.75) * 0.6, math.sqrt(0.75) * 0.8])
    V = _unit([0.3, -0.2, math.sqrt(1 - 0.09 - 0.04)])
    base = (0.8, 0.5, 0.2)
    w = dict(R=(1.0, 0.0, 0.0), TT=(0.0, 1.0, 0.0), TRT=(0.0, 0.0, 1.0))[lobe]
    u = _uniforms(P0 + 2.5 * V, [(P0 + 3 * L, WHITE, 1.0), (OFF, WHITE, 0.0), (OFF, WHITE, 0.0)],
                  dict(color0=base, color1=base, specular=w[0], weightTT=w[1], weightTRT=w[2], shift=0.06,
                       roughness=0.35))
    pos, tan = _strand(P0, T)
    got = _shade(oracle, u, pos, tan)[0, 0, :3]
    R, TT, TRT = _lobes(base, T, L, V, 0.06, 0.35)
    exp = dict(R=R, TT=TT, TRT=TRT)[lobe]
    assert got == pytest.approx(exp, rel=5e-5)
    # the opposite shift changes the lobe (a sign slip on alpha would not)
    u2 = _uniforms(P0 + 2.5 * V, [(P0 + 3 * L, WHITE, 1.0), (OFF, WHITE, 0.0), (OFF, WHITE, 0.0)],
                   dict(color0=base, color1=base, specular=w[0], weightTT=w[1], weightTRT=w[2], shift=-0.06,
                        roughness=0.35))
    R2, TT2, TRT2 = _lobes(base, T, L, V, -0.06, 0.35)
    exp2 = dict(R=R2, TT=TT2, TRT=TRT2)[lobe]
    assert _shade(oracle, u2, pos, tan)[0, 0, :3] == pytest.approx(exp2, rel=5e-5)
    assert not np.allclose(exp, exp2, rtol=1e-3)


def test_diffuse_and_lobes_add(oracle):
    """All terms on: diffuse + R + TT + TRT, each weighted (hairShading's brdfFinal)."""
    T = _unit([1.0, 0.2, 0.0])
    L = _unit([0.4, 0.5, 0.77])
    V = _unit([-0.2, 0.1, 1.0])
    base = (0.6, 0.3, 0.9)
    u = _uniforms(P0 + 2.0 * V, [(P0 + 3 * L, WHITE, 0.9), (OFF, WHITE, 0.0), (OFF, WHITE, 0.0)],
                  dict(color0=base, color1=base, specular=0.9, weightTT=0.5, weightTRT=1.4, shift=0.02,
                       roughness=0.3))
    pos, tan = _strand(P0, T)
    got = _shade(oracle, u, pos, tan)[0, 0, :3]
    R, TT, TRT = _lobes(base, T, L, V, 0.02, 0.3)
    exp = 0.9 * (_diffuse(base, T, L, V) + 0.9 * R + 0.5 * TT + 1.4 * TRT)
    assert got == pytest.approx(exp, rel=5e-5)


def _step_shadow_map():
    """Columns 0..31 hold depth 0 (an occluder), columns 32..63 depth 1."""
    m = np.ones((VP, VP), np.float32)
    m[:, : VP // 2] = 0.0
    return m


@pytest.mark.parametrize("z,radius,pcss,bias,src,hits", [
    (0.6, 3, 0, 0.0, "z", 21),    # tap columns -3..-1 see averages 0, 0, 0.5 < z - bias: 3 columns x 7 rows
    (0.4, 3, 0, 0.0, "z", 14),    # the straddling column's average 0.5 is above z - bias: 2 x 7
    (0.52, 3, 0, 0.0, "z", 21),   # the least bias (0.0005) keeps 0.52 - bias above 0.5
    (0.6, 3, 0, 0.15, "z", 14),   # bias 0.15 (source on +z, tangent +x: 1 - T.toSource ~ 1) lifts the straddle
    (0.6, 3, 0, 0.15, "x", 21),   # source along the tangent: 1 - T.toSource = 0, the bias falls to 0.0005
    (0.6, 1, 0, 0.0, "z", 3),     # 3 x 3 taps, column -1 only
    (0.6, 3, 1, 0.0, "z", 3),     # PCSS: centre average 1 >= z, radius 1 + 0 -> 3 x 3 taps
])
def test_pcf_shadow_on_a_step_map(oracle, z, radius, pcss, bias, src, hits):
    """The shadow point projects to u = 0.51953125 (texel coordinate 32.75 on
    a 64-texel map): tap column x gathers texels 32+x and 33+x, so columns
    x >= 0 see the lit side, x = -1 straddles the step and x <= -2 see the
    occluder. The bias is max(|bias| (1 - T.toSource), 0.0005).
    shadow = 1 - hits / (2r+1)^2 only changes the tint
    (base / luma)^(1 - shadow) of the multiple-scattering term; with the
    material's shadow strength 1 it is not clamped."""
    p0 = np.array([P0[0], P0[1], z])
    T = np.array([1.0, 0.0, 0.0])
    light = p0 + [1.0, 0.0, 1.0]
    base = (0.8, 0.5, 0.2)
    srcpos = tuple(p0 + ([0.0, 0.0, 10.0] if src == "z" else [10.0, 0.0, 0.0]))
    sh = dict(radius=radius, usePCSS=pcss, bias=bias, position=srcpos)
    u = _uniforms(p0 + [0, 0, 2.0], [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                  dict(shadows=1.0), shadow=sh)
    pos, tan = _strand(p0, T)
    got = _shade(oracle, u, pos, tan, shadow_map=_step_shadow_map())[0, 0, :3]
    n = (2 * (1 if pcss else radius) + 1) ** 2
    exp = _diffuse(base, T, _unit([1, 0, 1]), np.array([0, 0, 1.0]), shadow=1.0 - hits / n)
    assert got == pytest.approx(exp, rel=2e-5)
    # the material's shadow strength clamps shadow to [1 - strength, 1]
    u2 = _uniforms(p0 + [0, 0, 2.0], [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                   dict(shadows=0.1), shadow=sh)
    got2 = _shade(oracle, u2, pos, tan, shadow_map=_step_shadow_map())[0, 0, :3]
    exp2 = _diffuse(base, T, _unit([1, 0, 1]), np.array([0, 0, 1.0]), shadow=max(1.0 - hits / n, 0.9))
    assert got2 == pytest.approx(exp2, rel=2e-5)


def test_shadow_outside_the_map_is_unshadowed(oracle):
    """_isValidShadowSample: beyond the far plane (z > 1) no shadow, even over an occluder."""
    p0 = np.array([P0[0], P0[1], 1.5])
    T = np.array([1.0, 0.0, 0.0])
    light = p0 + [1.0, 0.0, 1.0]
    u = _uniforms(p0 + [0, 0, 2.0], [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                  dict(shadows=1.0))
    pos, tan = _strand(p0, T)
    got = _shade(oracle, u, pos, tan, shadow_map=np.zeros((VP, VP), np.float32))[0, 0, :3]
    assert got == pytest.approx(_diffuse((0.8, 0.5, 0.2), T, _unit([1, 0, 1]), np.array([0, 0, 1.0])), rel=2e-5)


def test_ao_texel_and_strength(oracle):
    """sampleAo: the texel is i32(aoSize * px / viewport) of the point's
    y-down pixel position: x = 0.0390625 -> 32 * 0.51953125 = 16.625 -> 16,
    y = 0.3 -> 32 * 0.35 = 11.2 -> 11. ao = 1 - mix(1, texel, strength)
    scales the radiance by 1 - ao."""
    T = np.array([1.0, 0.0, 0.0])
    light = P0 + [1.0, 0.0, 1.0]
    aotex = (np.arange((VP // 2) ** 2, dtype=np.float32).reshape(VP // 2, VP // 2) + 1.0) / 2048.0
    pos, tan = _strand(P0, T)
    exp0 = _diffuse((0.8, 0.5, 0.2), T, _unit([1, 0, 1]), np.array([0, 0, 1.0]))
    for strength in (1.0, 0.5):
        u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)], {},
                      ao_strength=strength)
        got = _shade(oracle, u, pos, tan, ao=aotex)[0, 0, :3]
        a = float(aotex[11, 16])
        occl = 1.0 - ((1.0 - strength) + strength * a)
        assert got == pytest.approx(exp0 * (1.0 - occl), rel=2e-5)


def test_attenuation_from_linearized_depth(oracle):
    """getAttenuation: |lin(z) - lin(zbuf)| * factor with lin(d) =
    n f / (f + (n - f) d), n = 0.01, f = 100, read at the texel
    (u32(px.x), u32(px.y)) = (33, 22); it saturates together with the AO term."""
    T = np.array([1.0, 0.0, 0.0])
    light = P0 + [1.0, 0.0, 1.0]
    pos, tan = _strand(P0, T)
    depth = np.zeros((VP, VP), np.float32)
    depth[22, 33] = 0.9
    lin = lambda d: 0.01 * 100.0 / (100.0 + (0.01 - 100.0) * d)  # noqa: E731
    exp0 = _diffuse((0.8, 0.5, 0.2), T, _unit([1, 0, 1]), np.array([0, 0, 1.0]))
    for factor in (4.0, 40.0):
        u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                      dict(attenuation=factor))
        got = _shade(oracle, u, pos, tan, depth=depth)[0, 0, :3]
        att = min(abs(lin(0.6) - lin(0.9)) * factor, 1.0)
        assert got == pytest.approx(exp0 * (1.0 - att), rel=2e-5, abs=1e-12)


def test_strand_colour_terms_and_tip_alpha(oracle):
    """Base colour = mix(color0, color1, t) at t = id / Ps; randomRGB(strand)
    with colorRng = 1 replaces it by fract(s * (1.73, 1.17, 1.31)); lumaRng = 1
    multiplies it by fract(s / 255). The last shading point has alpha 0."""
    T = np.array([1.0, 0.0, 0.0])
    light = P0 + [1.0, 0.0, 1.0]
    L, V = _unit([1, 0, 1]), np.array([0, 0, 1.0])
    pos, tan = _strand(P0, T, strands=6)
    c0, c1 = np.array([0.9, 0.2, 0.1]), np.array([0.1, 0.6, 0.8])
    u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                  dict(color0=c0, color1=c1))
    out = _shade(oracle, u, pos, tan, strands=6, ps=2)
    assert out[5, 0, :3] == pytest.approx(_diffuse(c0, T, L, V), rel=2e-5)  # t = 0
    assert out[:, 0, 3].tolist() == [1.0] * 6 and out[:, 1, 3].tolist() == [0.0] * 6
    s = np.float32(5)
    rnd = np.array([(s * np.float32(k)) % np.float32(1.0) for k in (1.73, 1.17, 1.31)], np.float64)
    u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                  dict(color0=c0, color1=c1, colorRng=1.0))
    out = _shade(oracle, u, pos, tan, strands=6, ps=2)
    assert out[5, 0, :3] == pytest.approx(_diffuse(rnd, T, L, V), rel=2e-5)
    u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)],
                  dict(color0=c0, color1=c1, lumaRng=1.0))
    out = _shade(oracle, u, pos, tan, strands=6, ps=2)
    assert out[5, 0, :3] == pytest.approx(_diffuse(c0 * (5.0 / 255.0), T, L, V), rel=2e-5)
    assert np.all(out[0, 0, :3] == 0.0)  # strand 0: fract(0 / 255) = 0


def test_shading_point_position_along_the_strand(oracle):
    """remapToIndices: with P = 2 and Ps = 4, point id sits at t = id / 4 of the
    segment; base colour mixes over the same t. A point light placed on the
    strand's normal plane through the point at t = 0.25 gives T.L = 0 there
    (no diffuse) but not at the other points."""
    T = np.array([1.0, 0.0, 0.0])
    pos, tan = _strand(P0, T)
    x25 = P0[0] + 0.025
    light = np.array([x25, P0[1], P0[2] + 1.0])
    u = _uniforms(CAM, [(light, WHITE, 1.0), (light, WHITE, 0.0), (light, WHITE, 0.0)], {})
    out = oracle.shading(_cfg(4), u, pos, tan, np.zeros((VP, VP), np.float32),
                         np.ones((VP // 2, VP // 2), np.float32), np.ones((VP, VP), np.float32), 1, 2)
    out = out.reshape(4, 4)
    assert np.abs(out[1, :3]).max() < 1e-6
    assert out[0, 0] > 1e-4 and np.all(out[2:, :3] == 0.0)  # points 2, 3: light behind the tangent
