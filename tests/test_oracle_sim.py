"""Pins of the oracle's integration forces and SDF collisions (K2) from closed forms.

The reference's own tests hold nothing for the simulation, so each test puts
points at chosen places in a uniform grid, where the non-standard grid weights
(grids.wgsl.ts:61-92) give closed forms:

- at a grid node one corner is touched, with weights (1, 1, 1): the velocity
  gather returns v and the density gather |(1, 1, 1)| rho = sqrt(3) rho;
- inside a cell with fractional offsets (fx, fy, fz) the eight corners' per-axis
  weights sum to 4 on every axis (each axis weight pair sums to 1 and the other
  two axes choose 2 x 2 corners), so a uniform vector field gathers to 4 v,
  while the density gathers rho * sum over corners of |w|;

and then checks the friction/grid-velocity term (hairSimIntegrationPass.wgsl.ts
:146-153, integration.wgsl.ts:8-28), the density-gradient force (:141-145), the
wind with its per-strand jitter (:130-139), and the SDF push-out along the
normalised forward-difference gradient of an analytic plane
(collisions.wgsl.ts:26-61, applied to points j > 2 only, :246-255).
"""
import math

import numpy as np
import pytest

D = 9  # grid dims; grid bounds [0, 8]^3, so node i sits at coordinate i (t = p / 8 is exact)


def _uniforms(frame_idx=0, **over):
    from frostbitten_hair_webgpu_amd import layouts as L
    sc = L.SimConfig(gravity=0.0, friction=0.0, volumePreservation=0.0, constraintIterations=0, windStrength=0.0,
                     collisionSphere=(50.0, 50.0, 50.0, 0.01), gridDims=D, deltaTime=0.5,
                     stiffnessLengthConstr=0.0, stiffnessGlobalConstr=0.0, stiffnessLocalConstr=0.0)
    for k, v in over.items():
        setattr(sc, k, v)
    grid = (np.zeros(3, np.float32), np.full(3, D - 1, np.float32))
    sdf_b = (np.full(3, -1.0, np.float32), np.full(3, 1.0, np.float32))
    return sc, L.pack_sim_uniforms(sc, np.eye(4), sdf_b, grid, frame_idx)


def _step(oracle, u, now, prev=None, dv=None, dgw=None, sdf=None, sdf_dims=8, P=None):
    now = np.ascontiguousarray(now, np.float32)
    P = P or now.shape[0]
    S = now.shape[0] // P
    prev = now.copy() if prev is None else np.ascontiguousarray(prev, np.float32)
    dv = np.zeros(D ** 3 * 4, np.int32) if dv is None else dv
    dgw = np.zeros(D ** 3 * 4, np.float32) if dgw is None else dgw
    sdf = np.full(sdf_dims ** 3, 10.0, np.float32) if sdf is None else sdf
    seglen = np.full(now.shape[0], 0.1, np.float32)
    out = oracle.sim_integrate(D, u, prev, now, seglen, sdf, sdf_dims, dv.reshape(-1), dgw.reshape(-1), now, S, P)
    return out.astype(np.float64)


NODE = (3.0, 4.0, 5.0)
CELL = (3.25, 4.5, 4.75)


def _corner_weight_norm_sum(p):
    """sum over the 8 corners of |w|, w = per-axis (1 - |p - corner|)."""
    f = [c - math.floor(c) for c in p]
    tot = 0.0
    for wx in (1 - f[0], f[0]):
        for wy in (1 - f[1], f[1]):
            for wz in (1 - f[2], f[2]):
                tot += math.sqrt(wx * wx + wy * wy + wz * wz)
    return tot


def _points(*ps):
    """A root (w = 0) at the origin followed by movable points at `ps`."""
    a = np.zeros((1 + len(ps), 4), np.float32)
    a[1:, :3] = ps
    a[1:, 3] = 1.0
    return a


@pytest.mark.parametrize("friction", [1.0, 0.5])
def test_friction_moves_by_grid_velocity_over_density(oracle, friction):
    """Uniform grid velocity v (i32 x 1e6) and density rho: at rest (prev = now)
    a point moves by friction * v_gather / rho_gather: v / (sqrt(3) rho) at a
    node, 4 v / (rho * sum |w|) inside a cell. The root keeps prev."""
    v = np.array([0.3, -0.6, 0.9])
    rho = 2.0
    dv = np.zeros((D ** 3, 4), np.int32)
    dv[:, :3] = (v * 1e6).astype(np.int32)
    dv[:, 3] = int(rho * 1e6)
    _, u = _uniforms(friction=friction)
    now = _points(NODE, CELL)
    out = _step(oracle, u, now, dv=dv)
    assert np.array_equal(out[0], now[0])
    assert out[1, :3] - now[1, :3] == pytest.approx(friction * v / (math.sqrt(3) * rho), rel=1e-5)
    assert out[2, :3] - now[2, :3] == pytest.approx(friction * 4 * v / (rho * _corner_weight_norm_sum(CELL)),
                                                    rel=1e-5)
    # moving point: mix(now - prev, gridDisp, friction) (integration.wgsl.ts:18-21)
    prev = now.copy()
    prev[1:, :3] -= np.float32(0.01)
    out2 = _step(oracle, u, now, prev=prev, dv=dv)
    exp = (1 - friction) * 0.01 + friction * v / (math.sqrt(3) * rho)
    assert out2[1, :3] - now[1, :3] == pytest.approx(exp, rel=1e-5)
    # empty grid cells (density 0) give no grid displacement, whatever the velocity
    dv[:, 3] = 0
    out3 = _step(oracle, u, now, dv=dv)
    assert np.array_equal(out3[1:], now[1:].astype(np.float64))


def test_density_gradient_force(oracle):
    """force += densityGrad * volumePreservation: a uniform gradient g gathers
    to g at a node and 4 g inside a cell (per-axis weights); the position moves
    by force * dt^2."""
    g = np.array([0.1, -0.2, 0.3])
    dgw = np.zeros((D ** 3, 4), np.float32)
    dgw[:, :3] = g
    k = 2.5
    _, u = _uniforms(volumePreservation=k)
    out = _step(oracle, u, _points(NODE, CELL), dgw=dgw)
    dt2 = 0.25  # dt = 0.5: displacements well above the f32 spacing at the points' coordinates
    assert out[1, :3] - np.array(NODE) == pytest.approx(g * k * dt2, rel=1e-5)
    assert out[2, :3] - np.array(CELL) == pytest.approx(4 * g * k * dt2, rel=1e-5)


@pytest.mark.parametrize("frame_idx", [0, 3, 17])
def test_wind_force_with_strand_jitter(oracle, frame_idx):
    """force += wind.xyz * |strength * (1 + jitter) * saturate(mix(lull, 1, w))|
    with w = sum over corners of |w_c| windStrength_c (sqrt(3) ws at a node),
    jitter = mix(-J/2, J/2, fract((frameIdx * 0.73 + phase * strand) * freq)).
    Three one-point strands at the same node differ only by the strand index."""
    ws, lull, J, phase, freq, A = 0.2, 0.75, 0.7, 0.45, 1.8, 2.0
    dgw = np.zeros((D ** 3, 4), np.float32)
    dgw[:, 3] = ws
    sc, u = _uniforms(frame_idx=frame_idx, windStrength=A, windStrengthLull=lull, windStrengthJitter=J,
                      windPhaseOffset=phase, windStrengthFrequency=freq, windDirPhi=30.0, windDirTheta=60.0)
    from frostbitten_hair_webgpu_amd import matrices as M
    wdir = M.spherical_to_cartesian(30.0, 60.0)
    now = np.concatenate([_points(NODE)] * 3)
    out = _step(oracle, u, now, dgw=dgw, P=2).reshape(3, 2, 4)
    dt2 = 0.25  # dt = 0.5: displacements well above the f32 spacing at the points' coordinates
    cell_str = min(max(lull + (1 - lull) * math.sqrt(3) * ws, 0.0), 1.0)
    for s in range(3):
        t = (np.float32(frame_idx) * np.float32(0.73) + np.float32(phase) * np.float32(s)) * np.float32(freq)
        jit = float(t) - math.floor(float(t))
        jd = -0.5 * J + J * jit
        exp = wdir * abs(A * (1 + jd) * cell_str) * dt2
        assert out[s, 1, :3] - np.array(NODE) == pytest.approx(exp, rel=2e-5, abs=1e-10)
    assert not np.allclose(out[0, 1], out[1, 1])
    # the strength enters through abs(): a negative strength pushes the same way
    _, u2 = _uniforms(frame_idx=frame_idx, windStrength=-A, windStrengthLull=lull, windStrengthJitter=J,
                      windPhaseOffset=phase, windStrengthFrequency=freq, windDirPhi=30.0, windDirTheta=60.0)
    assert np.allclose(_step(oracle, u2, now, dgw=dgw, P=2), out.reshape(-1, 4), rtol=0, atol=0)


def test_forces_add_with_gravity(oracle):
    """Gravity, density gradient and grid velocity at once: the displacement is
    friction * gridDisp + (g_vec + grad * k) * dt^2."""
    v, rho, grad = np.array([0.2, 0.1, -0.1]), 1.5, np.array([0.0, 0.4, 0.0])
    dv = np.zeros((D ** 3, 4), np.int32)
    dv[:, :3] = (v * 1e6).astype(np.int32)
    dv[:, 3] = int(rho * 1e6)
    dgw = np.zeros((D ** 3, 4), np.float32)
    dgw[:, :3] = grad
    _, u = _uniforms(gravity=0.5, volumePreservation=3.0, friction=0.25)
    out = _step(oracle, u, _points(NODE), dv=dv, dgw=dgw)
    dt2 = 0.25  # dt = 0.5: displacements well above the f32 spacing at the points' coordinates
    exp = 0.25 * v / (math.sqrt(3) * rho) + (np.array([0, -0.5, 0]) + grad * 3.0) * dt2
    assert out[1, :3] - np.array(NODE) == pytest.approx(exp, rel=1e-5)


def _plane_sdf(n, c, d=16, lo=-1.0, hi=1.0):
    """sdf(p) = n.p - c stored at texel centres of a d^3 volume over [lo, hi]^3:
    trilinear filtering reproduces a linear field inside the centres' range."""
    i = (np.arange(d) + 0.5) / d * (hi - lo) + lo
    zz, yy, xx = np.meshgrid(i, i, i, indexing="ij")
    return (n[0] * xx + n[1] * yy + n[2] * zz - c).astype(np.float32).reshape(-1), d


def _strand_on_level(n, level, c, off, P=6):
    """A strand at constant sdf + offset = `level`: it runs along e, which is
    perpendicular to n, and bends along f = n x e (an evenly spaced straight
    strand makes the local-shape constraint divide 0 by 0, constraints.wgsl.ts:69-84)."""
    e = np.array([n[1], -n[0], 0.0])
    e /= np.linalg.norm(e)
    f = np.cross(n, e)
    p0 = n * (level + c - 0.01) - 0.25 * e
    a = np.zeros((P, 4), np.float32)
    for j in range(P):
        a[j, :3] = p0 + 0.1 * j * e + 0.004 * (j % 2) * f
    a[:, 3] = 1.0
    a[0, 3] = 0.0
    return a


@pytest.mark.parametrize("iters,stiff,keep", [(1, 1.0, 0.0), (1, 0.5, 0.5), (2, 1.0, 0.25)])
def test_sdf_pushes_points_along_the_plane_normal(oracle, iters, stiff, keep):
    """Analytic plane sdf = n.p - c with n = (0.6, 0.48, 0.64) (distinct
    components, so a swapped gradient axis moves the point off the normal).
    A point with sdf + offset < 0 moves by -(sdf + offset) n * stiffnessSDF /
    iterations per iteration, so after the step (sdf + offset) is `keep` times
    its start value (1 - k per iteration). Points 1 and 2 are never tested
    against the SDF. A strand at sdf + offset = +0.005 (sdf itself negative:
    the offset decides) and one far outside do not move."""
    n = np.array([0.6, 0.48, 0.64])
    c, off = 0.1, 0.01
    vol, d = _plane_sdf(n, c)
    from frostbitten_hair_webgpu_amd import layouts as L
    sc = L.SimConfig(gravity=0.0, friction=0.0, volumePreservation=0.0, constraintIterations=iters,
                     windStrength=0.0, collisionSphere=(50.0, 50.0, 50.0, 0.01), gridDims=D,
                     stiffnessLengthConstr=0.0, stiffnessGlobalConstr=0.0, stiffnessLocalConstr=0.0,
                     stiffnessSDF=stiff, sdfDistanceOffset=off)
    grid = (np.zeros(3, np.float32), np.full(3, D - 1, np.float32))
    u = L.pack_sim_uniforms(sc, np.eye(4), (np.full(3, -1.0, np.float32), np.full(3, 1.0, np.float32)), grid, 0)
    inside, edge, outside = (_strand_on_level(n, lv, c, off) for lv in (-0.3, 0.005, 0.2))
    now = np.concatenate([inside, edge, outside])
    out = _step(oracle, u, now, sdf=vol, sdf_dims=d, P=6).reshape(3, 6, 4)
    sd = lambda p: float(np.dot(n, p[:3]) - c + off)  # noqa: E731
    for j in range(6):
        if j <= 2:
            assert np.array_equal(out[0, j], inside[j].astype(np.float64))
        else:
            assert sd(out[0, j]) == pytest.approx(keep * sd(inside[j]), abs=2e-4)
            move = out[0, j, :3] - inside[j, :3]
            assert np.linalg.norm(np.cross(move, n)) < 1e-3 * np.linalg.norm(move)  # along n
            assert np.dot(move, n) > 0
    assert np.array_equal(out[1], edge.astype(np.float64))
    assert np.array_equal(out[2], outside.astype(np.float64))
