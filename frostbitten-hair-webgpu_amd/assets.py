"""Seeded synthetic hair assets and scene inputs for the benchmark workloads.

Writes nothing the reference does not read: hair as TressFX strands (the .tfx
layout, formats.write_tfx), the collider as an SDF volume baked analytically at
texel centres (the SDF .bin layout, formats.write_sdf), and the depth / shadow
/ AO textures that the reference's hardware-raster passes would produce
(drawMeshesPass, hwHairPass, shadowMapPass, aoPass), computed here on the host
by sphere tracing the same analytic colliders.

Workloads (seed 0xF00D0000 + k):
  k=1 small:  2,000 x 16, hemisphere roots on a 0.1 m sphere, 0.25 m strands
  k=2 c2:     10,000 x 32, scalp cap 0-70 deg, log-normal lengths (median 0.3)
  k=3 c3:     100,000 x 32, scalp cap, lengths U(0.2, 0.6)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import layouts as L
from . import matrices as M

HEAD_RADIUS = 0.1


@dataclass
class Collider:
    spheres: list = field(default_factory=list)   # (center(3), radius)
    capsules: list = field(default_factory=list)  # (a(3), b(3), radius)

    def sdf(self, p: np.ndarray) -> np.ndarray:
        """Signed distance of the union, p: (..., 3) float64."""
        d = np.full(p.shape[:-1], np.inf)
        for c, r in self.spheres:
            d = np.minimum(d, np.linalg.norm(p - np.asarray(c), axis=-1) - r)
        for a, b, r in self.capsules:
            a, b = np.asarray(a), np.asarray(b)
            ab = b - a
            t = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
            q = a + t[..., None] * ab
            d = np.minimum(d, np.linalg.norm(p - q, axis=-1) - r)
        return d

    def bounds(self, margin: float) -> tuple:
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        for c, r in self.spheres:
            lo = np.minimum(lo, np.asarray(c) - r)
            hi = np.maximum(hi, np.asarray(c) + r)
        for a, b, r in self.capsules:
            for e in (a, b):
                lo = np.minimum(lo, np.asarray(e) - r)
                hi = np.maximum(hi, np.asarray(e) + r)
        return lo - margin, hi + margin


def bake_sdf(col: Collider, dims: int, margin: float = 0.05) -> tuple:
    """Distances at texel centres (i + 0.5) / dims of the bounds, x fastest."""
    lo, hi = col.bounds(margin)
    axis = [(np.arange(dims) + 0.5) / dims * (hi[k] - lo[k]) + lo[k] for k in range(3)]
    zz, yy, xx = np.meshgrid(axis[2], axis[1], axis[0], indexing="ij")
    p = np.stack([xx, yy, zz], axis=-1)
    d = col.sdf(p).astype(np.float32).reshape(-1)
    return lo.astype(np.float32), hi.astype(np.float32), d


def _orthonormal(n: np.ndarray) -> tuple:
    a = np.where(np.abs(n[:, 1:2]) < 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    return t, b


def grow_strands(rng: np.random.Generator, roots: np.ndarray, normals: np.ndarray, lengths: np.ndarray,
                 points: int, curl_sigma: float = 0.005, droop: float = 1.2) -> np.ndarray:
    """Strands grown along the root normal, bending toward -y (gravity droop),
    with smooth seeded curl noise; equidistant points (tfx_exporter.py:82-140
    resamples the same way); .w = 0 at the root, 1 elsewhere."""
    S = roots.shape[0]
    t_axis, b_axis = _orthonormal(normals)
    seg = lengths / (points - 1)
    pos = np.zeros((S, points, 4), dtype=np.float64)
    pos[:, 0, :3] = roots
    cur = roots.copy()
    freq = rng.uniform(6.0, 14.0, size=(S, 1))
    ph1 = rng.uniform(0, 2 * math.pi, size=(S, 1))
    ph2 = rng.uniform(0, 2 * math.pi, size=(S, 1))
    amp = np.abs(rng.normal(1.0, 0.3, size=(S, 1))) * curl_sigma
    gravity = np.array([0.0, -1.0, 0.0])
    prev_off = np.zeros((S, 3))
    for k in range(1, points):
        s = k / (points - 1)
        d = normals + droop * s * gravity[None, :]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        cur = cur + d * seg[:, None]
        off = amp * (np.sin(freq * s * math.pi + ph1) * t_axis + np.cos(freq * s * math.pi + ph2) * b_axis) * s
        pos[:, k, :3] = cur + off
        prev_off = off
    pos[:, 1:, 3] = 1.0
    del prev_off
    return pos.reshape(-1, 4).astype(np.float32)


@dataclass
class Workload:
    name: str
    seed: int
    strands: int
    points: int
    width: int
    height: int
    steps: int
    render: L.RenderConfig
    sim: L.SimConfig
    sdf_dims: int
    collider: Collider
    camera: L.Camera
    positions: np.ndarray = None  # initial pose (S*P, 4)
    sdf_bounds: tuple = None
    sdf: np.ndarray = None


def _head_collider(kind: str) -> Collider:
    head = ((0.0, 0.0, 0.0), HEAD_RADIUS)
    if kind == "sphere":
        return Collider(spheres=[head])
    # face / neck proxy: a sphere plus two capsules
    face = ((0.0, -0.06, 0.05), (0.0, -0.12, 0.07), 0.055)
    neck = ((0.0, -0.08, -0.01), (0.0, -0.25, -0.02), 0.05)
    return Collider(spheres=[head], capsules=[face, neck])


def segment_list_budget(S: int, P: int, W: int, H: int, tile: int) -> float:
    """avgSegmentsPerTile (constants.ts:230; the segment list holds
    1 + tiles * avgSegmentsPerTile entries, hairTileSegmentsBuffer.ts:81-87)
    for a workload of these generators: the reference's 128 unless the
    expected (segment, tile) pairs need more. The scalp-cap generators draw
    about 1.5 pairs per segment with 16 px tiles at 1080p, proportionally
    more with the resolution (1080p: 0.41 M pairs for 10k x 32; 4K: 9.3 M for
    100k x 32) and at most twice as many with 8 px tiles; the budget keeps
    twice that, so that no pair is dropped. The oracle and the HIP path read
    the same value."""
    tiles = -(-W // tile) * -(-H // tile)
    expected = 1.5 * S * (P - 1) * max(W / 1920.0, H / 1080.0) * max(1.0, 16.0 / tile)
    return float(max(128, math.ceil(2.0 * expected / tiles)))


def make_workload(kind: str, seed: int | None = None, strands: int | None = None,
                  width: int | None = None, height: int | None = None, points: int | None = None,
                  collider: Collider | None = None, sdf: tuple | None = None) -> Workload:
    """kind in {small, c2, c3}; optional overrides for tests and for the
    batched and served instances (`collider` / `sdf` reuse a baked volume)."""
    specs = {
        "small": dict(k=1, S=2000, P=16, W=640, H=480, steps=30, sdf=32, grid=32, tile=16, bins=4, slices=4,
                      ps=4, col="sphere"),
        "c2": dict(k=2, S=10000, P=32, W=1920, H=1080, steps=120, sdf=64, grid=64, tile=16, bins=8, slices=4,
                   ps=8, col="union"),
        "c3": dict(k=3, S=100000, P=32, W=3840, H=2160, steps=60, sdf=128, grid=96, tile=16, bins=16,
                   slices=8, ps=8, col="union"),
    }
    sp = specs[kind]
    seed = (0xF00D0000 + sp["k"]) if seed is None else seed
    rng = np.random.default_rng(seed)
    S = strands if strands is not None else sp["S"]
    P = points if points is not None else sp["P"]
    W = width if width is not None else sp["W"]
    H = height if height is not None else sp["H"]
    # roots
    if kind == "small":
        u = rng.uniform(0.0, 1.0, S)
        cos_t = u  # upper hemisphere, uniform in area
        lengths = np.full(S, 0.25)
    else:
        cos_t = rng.uniform(math.cos(math.radians(70.0)), 1.0, S)
        if kind == "c2":
            lengths = np.exp(rng.normal(math.log(0.3), 0.25, S))
        else:
            lengths = rng.uniform(0.2, 0.6, S)
    phi = rng.uniform(0.0, 2 * math.pi, S)
    sin_t = np.sqrt(1.0 - cos_t ** 2)
    normals = np.stack([sin_t * np.cos(phi), cos_t, sin_t * np.sin(phi)], axis=1)
    roots = normals * HEAD_RADIUS
    positions = grow_strands(rng, roots, normals, lengths, P)

    render = L.RenderConfig(tileSize=sp["tile"], tileDepthBins=sp["bins"], slicesPerPixel=sp["slices"],
                            shadingPoints=sp["ps"])
    render.avgSegmentsPerTile = segment_list_budget(S, P, W, H, sp["tile"])
    sim = L.SimConfig(deltaTime=1.0 / 60.0, gridDims=sp["grid"], collisionSphere=(0.0, 0.0, 0.0, HEAD_RADIUS))
    if kind == "small":
        sim.constraintIterations = 2
        sim.windStrength = 1.0
        sim.windStrengthJitter = 0.0
    else:
        sim.windStrength = 2.0
        # period 2 s at 60 steps/s: fract(timer * f) with timer += 0.73 per step
        sim.windStrengthFrequency = 1.0 / (0.73 * 120.0)
        sim.windPhaseOffset = float(rng.uniform(0.0, 2 * math.pi))
    collider = collider if collider is not None else _head_collider(sp["col"])
    # camera orbit around the head, seeded
    az = rng.uniform(-0.6, 0.6) + (0.0 if kind != "c3" else 0.3)
    el = rng.uniform(0.05, 0.35)
    # distances picked so the hair covers ~30-50% of the screen tiles
    dist = {"small": 1.3, "c2": 2.2, "c3": 1.8}[kind]
    lo, hi = positions[:, :3].min(axis=0), positions[:, :3].max(axis=0)
    target = np.array([0.0, 0.5 * (lo[1] + hi[1]), 0.0])
    eye = target + dist * np.array([math.sin(az) * math.cos(el), math.sin(el), math.cos(az) * math.cos(el)])
    camera = L.Camera(eye=tuple(eye), target=tuple(target))
    wl = Workload(kind, seed, S, P, W, H, sp["steps"], render, sim, sp["sdf"], collider, camera, positions)
    lo, hi, d = sdf if sdf is not None else bake_sdf(collider, sp["sdf"])
    wl.sdf_bounds = (lo, hi)
    wl.sdf = d
    return wl


# ---------------------------------------------------------------------------
# Instances of the batched and served forms
# ---------------------------------------------------------------------------

BATCH_SEED = 0xF00D0004
SERVE_SEED = 0xF00D0005
BATCH_RESOLUTIONS = [(512, 512), (1280, 720), (1920, 1080)]
SERVE_RESOLUTIONS = [(1280, 720), (1920, 1080), (3840, 2160)]

_shared_volume = {}


def _union_volume(dims: int) -> tuple:
    """The sphere + capsules collider and its baked SDF, baked once per size
    (every instance gets its own device copy)."""
    if dims not in _shared_volume:
        col = _head_collider("union")
        _shared_volume[dims] = (col, bake_sdf(col, dims))
    return _shared_volume[dims]


def make_instance(instance: int, base_seed: int = BATCH_SEED) -> Workload:
    """Instance `instance` of the batched form: seed = base ^ instance;
    strands log-uniform in [1k, 50k], points from {8, 16, 24, 32},
    resolution from {512^2, 1280x720, 1920x1080}; scalp-cap hair with its
    own camera and wind, a 64^3 SDF of the sphere + capsules collider and a
    64^3 grid."""
    rng = np.random.default_rng(base_seed ^ instance)
    S = int(round(math.exp(rng.uniform(math.log(1000.0), math.log(50000.0)))))
    P = int(rng.choice([8, 16, 24, 32]))
    W, H = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
    col, vol = _union_volume(64)
    wl = make_workload("c2", seed=int(rng.integers(1 << 31)), strands=S, width=W, height=H, points=P,
                       collider=col, sdf=vol)
    wl.name = f"batch{instance}"
    return wl


def make_pool_asset(asset: int, base_seed: int = SERVE_SEED) -> Workload:
    """Asset `asset` of the served form's pool: strands log-uniform in
    [1k, 100k], points uniform in [16, 32], resolution from {720p, 1080p,
    4K}; otherwise as the batched instances."""
    rng = np.random.default_rng(base_seed ^ asset)
    S = int(round(math.exp(rng.uniform(math.log(1000.0), math.log(100000.0)))))
    P = int(rng.integers(16, 33))
    W, H = SERVE_RESOLUTIONS[int(rng.integers(len(SERVE_RESOLUTIONS)))]
    col, vol = _union_volume(64)
    wl = make_workload("c2", seed=int(rng.integers(1 << 31)), strands=S, width=W, height=H, points=P,
                       collider=col, sdf=vol)
    wl.name = f"asset{asset}"
    return wl


# ---------------------------------------------------------------------------
# Derived per-strand buffers (src/scene/hair/*)
# ---------------------------------------------------------------------------

def tangents(positions: np.ndarray, S: int, P: int) -> np.ndarray:
    """createHairTangentsBuffer — hairTangentsBuffer.ts:35-76 (tip uses the last segment)."""
    p = positions.reshape(S, P, 4)[:, :, :3].astype(np.float32)
    nxt = np.concatenate([p[:, 1:], p[:, -1:]], axis=1)
    cur = np.concatenate([p[:, :-1], p[:, -2:-1]], axis=1)
    t = nxt - cur
    t = t / np.linalg.norm(t, axis=2, keepdims=True)
    out = np.zeros((S, P, 4), dtype=np.float32)
    out[:, :, :3] = t
    return out.reshape(-1, 4)


def segment_lengths(positions: np.ndarray, S: int, P: int) -> np.ndarray:
    """createHairSegmentLengthsBuffer — hairSegmentLengthsBuffer.ts:29-55."""
    p = positions.reshape(S, P, 4)[:, :, :3].astype(np.float64)
    d = np.linalg.norm(p[:, 1:] - p[:, :-1], axis=2)
    out = np.concatenate([d, d[:, -1:]], axis=1)
    return out.astype(np.float32).reshape(-1)


def bounding_box(positions: np.ndarray) -> tuple:
    """calcBoundingBox over xyz (utils/bounds.ts)."""
    p = positions[:, :3]
    return p.min(axis=0).astype(np.float64), p.max(axis=0).astype(np.float64)


def derived(wl: Workload) -> dict:
    """Host buffers derived from a workload's initial pose (bounding box,
    tangents, rest lengths), computed once per workload object."""
    d = getattr(wl, "_derived", None)
    if d is None:
        d = dict(box=bounding_box(wl.positions), tangents=tangents(wl.positions, wl.strands, wl.points),
                 seglen=segment_lengths(wl.positions, wl.strands, wl.points))
        wl._derived = d
    return d


def bounding_sphere(box: tuple) -> tuple:
    """calcBoundingSphere — utils/bounds.ts:101-108."""
    c = (box[0] + box[1]) * 0.5
    return c, float(np.linalg.norm(box[1] - c))


def scale_box(box: tuple, s: float) -> tuple:
    """scaleBoundingBox — utils/bounds.ts:121-133."""
    c = (box[0] + box[1]) * 0.5
    v = (box[1] - c) * s
    return (c - v).astype(np.float32), (c + v).astype(np.float32)


def initial_shading(S: int, PS: int) -> np.ndarray:
    """createHairShadingBuffer — hairShadingBuffer.ts:56-76: red-green gradient."""
    p = np.arange(PS) / (PS - 1)
    g = np.stack([1.0 - p, p, np.zeros(PS), np.ones(PS)], axis=1).astype(np.float32)
    return np.tile(g, (S, 1))


# ---------------------------------------------------------------------------
# Inputs produced by the reference's hardware-raster passes
# ---------------------------------------------------------------------------

def _sphere_trace(col: Collider, origins: np.ndarray, dirs: np.ndarray, t_max: float, iters: int = 96) -> tuple:
    """Sphere tracing of the collider union. Rays start where they enter the
    bounding sphere of the collider and stop where they leave it, so rays
    that miss it take no steps."""
    lo, hi = col.bounds(0.0)
    c, R = 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo)) + 1e-3
    oc = origins - c
    b = np.einsum("ij,ij->i", oc, dirs)
    disc = b * b - (np.einsum("ij,ij->i", oc, oc) - R * R)
    root = np.sqrt(np.maximum(disc, 0.0))
    t_in, t_out = -b - root, -b + root
    t = np.maximum(t_in, 0.0)
    t_end = np.minimum(t_out, t_max)
    hit = np.zeros(origins.shape[0], dtype=bool)
    active = (disc > 0.0) & (t_out > 0.0) & (t < t_end)
    for _ in range(iters):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        p = origins[idx] + dirs[idx] * t[idx, None]
        d = col.sdf(p)
        h = d < 1e-5
        hit[idx[h]] = True
        active[idx[h]] = False
        t[idx] += np.where(h, 0.0, d)
        far = t[idx] > t_end[idx]
        active[idx[far]] = False
    return t, hit


def _oriented_triangles(pos: np.ndarray, outward: np.ndarray, quads: list) -> np.ndarray:
    """Triangles (T, 3) of the quads (a, b, c, d) of a surface grid, each
    wound counter-clockwise seen from outside (its geometric normal along the
    outward normal at its centroid); zero-area triangles at the poles are
    dropped."""
    tris = []
    for a, b, c, d in quads:
        for t in ((a, b, c), (a, c, d)):
            p0, p1, p2 = pos[t[0]], pos[t[1]], pos[t[2]]
            n = np.cross(p1 - p0, p2 - p0)
            if np.dot(n, n) < 1e-24:
                continue
            if np.dot(n, outward[list(t)].sum(axis=0)) < 0:
                t = (t[0], t[2], t[1])
            tris.append(t)
    return np.asarray(tris, dtype=np.uint32)


def _surface_mesh(center_of, axis: np.ndarray, radius: float, rings: int, segments: int) -> dict:
    """A closed surface of revolution about `axis`: vertex (i, j) sits at
    center_of(i) + radius * n(theta_i, phi_j) with n the unit sphere's normal,
    theta_i from 0 (along +axis) to pi in `rings` steps."""
    axis = axis / np.linalg.norm(axis)
    v = np.cross(axis, [1.0, 0.0, 0.0] if abs(axis[0]) < 0.9 else [0.0, 1.0, 0.0])
    v /= np.linalg.norm(v)
    w = np.cross(axis, v)
    pos, nrm = [], []
    thetas = center_of("thetas", rings)
    for i, th in enumerate(thetas):
        for j in range(segments + 1):
            ph = 2.0 * math.pi * j / segments
            n = math.cos(th) * axis + math.sin(th) * (math.cos(ph) * v + math.sin(ph) * w)
            pos.append(center_of(i, rings) + radius * n)
            nrm.append(n)
    pos, nrm = np.asarray(pos), np.asarray(nrm)
    quads = []
    for i in range(len(thetas) - 1):
        for j in range(segments):
            a = i * (segments + 1) + j
            quads.append((a, a + segments + 1, a + segments + 2, a + 1))
    idx = _oriented_triangles(pos, nrm, quads)
    return {"positions": pos.astype(np.float32), "normals": nrm.astype(np.float32), "indices": idx}


def sphere_mesh(center, radius: float, rings: int = 24, segments: int = 48) -> dict:
    """A UV sphere (positions, normals float32 (V, 3), indices uint32 (T, 3),
    counter-clockwise seen from outside)."""
    c = np.asarray(center, dtype=np.float64)

    def center_of(i, n):
        if i == "thetas":
            return [math.pi * k / n for k in range(n + 1)]
        return c

    return _surface_mesh(center_of, np.array([0.0, 1.0, 0.0]), radius, rings, segments)


def capsule_mesh(a, b, radius: float, rings: int = 24, segments: int = 48) -> dict:
    """A capsule (the cylinder a-b with hemispherical caps): the equator ring
    is repeated, once on each cap's centre, which makes the cylinder."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    half = rings // 2
    thetas = [math.pi * k / (2 * half) for k in range(half + 1)] + \
             [math.pi * k / (2 * half) for k in range(half, 2 * half + 1)]

    def center_of(i, n):
        if i == "thetas":
            return thetas
        return b if i <= half else a

    return _surface_mesh(center_of, b - a, radius, rings, segments)


def scene_meshes(col: Collider, collision_sphere=None, rings: int = 24, segments: int = 48) -> list:
    """The scene's objects (Scene.objects, loadScene.ts:25-60) for a
    collider: one mesh per sphere and capsule, in that order, then (when
    `collision_sphere` (x, y, z, r) is given) the collider preview, a unit
    sphere drawn with collisionSphereModelMatrix (loadScene.ts:36,
    CONFIG.drawColliders)."""
    out = [sphere_mesh(c, r, rings, segments) for c, r in col.spheres]
    out += [capsule_mesh(a, b, r, rings, segments) for a, b, r in col.capsules]
    if collision_sphere is not None:
        m = sphere_mesh((0.0, 0.0, 0.0), 1.0, rings, segments)
        m["preview"] = True
        out.append(m)
    return out


def mesh_depth(wl: Workload, uniforms: np.ndarray) -> np.ndarray:
    """Depth texture (H x W, texture rows top to bottom) of the collider meshes
    as drawMeshesPass would leave it; 1.0 where nothing is drawn."""
    return mesh_depth_normals(wl, uniforms)[0]


def octahedral_rg16(n: np.ndarray) -> np.ndarray:
    """encodeOctahedronNormal (shaderSnippets.wgls.ts:25-46) of unit normals
    (..., 3) as rg16float texels (uint32, r in the low half)."""
    n = n / np.abs(n).sum(axis=-1, keepdims=True)
    x, y = n[..., 0].copy(), n[..., 1].copy()
    neg = n[..., 2] < 0
    sx, sy = np.where(x >= 0, 1.0, -1.0), np.where(y >= 0, 1.0, -1.0)
    wx, wy = (1.0 - np.abs(y)) * sx, (1.0 - np.abs(x)) * sy
    x, y = np.where(neg, wx, x), np.where(neg, wy, y)
    r = (x * 0.5 + 0.5).astype(np.float16).view(np.uint16).astype(np.uint32)
    g = (y * 0.5 + 0.5).astype(np.float16).view(np.uint16).astype(np.uint32)
    return r | (g << 16)


def mesh_depth_normals(wl: Workload, uniforms: np.ndarray) -> tuple:
    """Depth (H x W f32, 1.0 where nothing is drawn) and rg16float octahedral
    world-space normals (H x W u32, 0 where nothing is drawn) of the collider
    meshes, as drawMeshesPass leaves them."""
    W, H = wl.width, wl.height
    vp = uniforms["vpMatrix"][0].reshape(4, 4).T.astype(np.float64)
    inv = np.linalg.inv(vp)
    xs = (np.arange(W) + 0.5) / W * 2 - 1
    ys = 1 - (np.arange(H) + 0.5) / H * 2
    X, Y = np.meshgrid(xs, ys)
    n = np.stack([X.ravel(), Y.ravel(), np.zeros(W * H), np.ones(W * H)], axis=1) @ inv.T
    f = np.stack([X.ravel(), Y.ravel(), np.ones(W * H), np.ones(W * H)], axis=1) @ inv.T
    n = n[:, :3] / n[:, 3:]
    f = f[:, :3] / f[:, 3:]
    dirs = f - n
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    eye = np.asarray(wl.camera.eye)
    t, hit = _sphere_trace(wl.collider, np.broadcast_to(eye, n.shape).copy(), dirs, 10.0)
    p = eye + dirs * t[:, None]
    clip = np.concatenate([p, np.ones((p.shape[0], 1))], axis=1) @ vp.T
    z = clip[:, 2] / clip[:, 3]
    depth = np.where(hit, z, 1.0).astype(np.float32)
    normals = np.zeros(W * H, dtype=np.uint32)
    if hit.any():
        q = p[hit]
        e = 1e-4
        g = np.stack([wl.collider.sdf(q + e * np.eye(3)[k]) - wl.collider.sdf(q - e * np.eye(3)[k])
                      for k in range(3)], axis=1)
        g /= np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-12)
        normals[hit] = octahedral_rg16(g)
    return depth.reshape(H, W), normals.reshape(H, W)


def hair_depth(wl: Workload, uniforms: np.ndarray, positions: np.ndarray, base: np.ndarray) -> np.ndarray:
    """Depth with the strands splatted in (what hwHairPass adds before the
    shading pass samples it): min of the mesh depth and each point's depth."""
    W, H = wl.width, wl.height
    vp = uniforms["vpMatrix"][0].reshape(4, 4).T.astype(np.float64)
    P = wl.points
    p = positions.reshape(-1, P, 4)[:, :, :3].astype(np.float64)
    mid = 0.5 * (p[:, 1:] + p[:, :-1])
    pts = np.concatenate([p.reshape(-1, 3), mid.reshape(-1, 3)], axis=0)
    clip = np.concatenate([pts, np.ones((pts.shape[0], 1))], axis=1) @ vp.T
    ndc = clip[:, :3] / clip[:, 3:]
    x = ((ndc[:, 0] * 0.5 + 0.5) * W).astype(np.int64)
    row = ((1.0 - (ndc[:, 1] * 0.5 + 0.5)) * H).astype(np.int64)
    ok = (x >= 0) & (x < W) & (row >= 0) & (row < H) & (ndc[:, 2] > 0) & (ndc[:, 2] < 1)
    out = base.copy().reshape(-1)
    np.minimum.at(out, row[ok] * W + x[ok], ndc[ok, 2].astype(np.float32))
    return out.reshape(H, W)


def shadow_map(wl: Workload, uniforms: np.ndarray, size: int) -> np.ndarray:
    """Shadow-source depth of the collider meshes (orthographic, as
    shadowMapPass renders it); 1.0 where nothing is drawn."""
    mvp = uniforms["shadowSourceMVP"][0].reshape(4, 4).T.astype(np.float64)
    inv = np.linalg.inv(mvp)
    xs = (np.arange(size) + 0.5) / size * 2 - 1
    ys = 1 - (np.arange(size) + 0.5) / size * 2
    X, Y = np.meshgrid(xs, ys)
    n = np.stack([X.ravel(), Y.ravel(), np.zeros(size * size), np.ones(size * size)], axis=1) @ inv.T
    f = np.stack([X.ravel(), Y.ravel(), np.ones(size * size), np.ones(size * size)], axis=1) @ inv.T
    n = n[:, :3] / n[:, 3:]
    f = f[:, :3] / f[:, 3:]
    dirs = f - n
    length = np.linalg.norm(dirs, axis=1)
    dirs /= length[:, None]
    t, hit = _sphere_trace(wl.collider, n, dirs, float(length.max()), iters=64)
    depth = np.where(hit, t / length, 1.0).astype(np.float32)
    return depth.reshape(size, size)


def hdr_scene(wl: Workload, mesh: np.ndarray) -> np.ndarray:
    """The rgba16float HDR texture (H x W x 4 half bits, rows top to bottom)
    that the background and mesh passes leave for the combine pass: the
    background gradient of drawBackgroundGradientPass.wgsl.ts:33-48 with its
    noise term held at its mean (c = 0.5), and the collider meshes as flat
    grey where the mesh depth is below 1."""
    W, H = wl.width, wl.height
    c0, c1 = np.array(L.col(22, 162, 188)), np.array(L.col(14, 103, 120))
    uv_y = (np.arange(H) + 0.5) / H
    base = c0 * 0.5 + c1 * 0.5
    col = base[None, :] * 0.5 + (1.0 - uv_y)[:, None] * 0.5  # mix(color, 1 - uv.y, gradientStrength = 0.5)
    img = np.empty((H, W, 4), dtype=np.float32)
    img[..., :3] = col[:, None, :]
    img[..., 3] = 1.0
    img[mesh < 1.0, :3] = 0.5
    return img.astype(np.float16).view(np.uint16)
