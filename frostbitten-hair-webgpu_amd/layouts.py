"""Byte layouts of the uniform blocks and the tunables the reference keeps in
CONFIG (src/constants.ts:86-349), plus the host-side packers that mirror
SimulationUniformsBuffer.update (simulationUniformsBuffer.ts:86-152) and
RenderUniformsBuffer.update (renderUniformsBuffer.ts:205-288).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import matrices as M

F4 = ("<f4", (4,))
M4 = ("<f4", (16,))

SIM_UNIFORMS_DTYPE = np.dtype({
    "names": ["modelMatrix", "modelMatrixInv", "collisionSphere", "wind", "sdfBoundsMin", "sdfBoundsMax",
              "gridBoundsMin", "gridBoundsMax", "gridDebugDisplayValue", "deltaTime", "gravity",
              "constraintIterations", "frameIdx", "windStrengthLull", "windColisionTraceOffset",
              "stiffnessLengthConstr", "stiffnessCollisions", "stiffnessSDF", "volumePreservation", "friction",
              "stiffnessGlobalConstr", "globalConstrExtent", "globalConstrFade", "stiffnessLocalConstr",
              "windPhaseOffset", "windStrengthFrequency", "windStrengthJitter", "padding0", "padding1"],
    "formats": [M4, M4, F4, F4, F4, F4, F4, ("<f4", (3,)), "<u4", "<f4", "<f4", "<u4", "<u4", "<f4", "<f4",
                "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4"],
    "offsets": [0, 64, 128, 144, 160, 176, 192, 208, 220, 224, 228, 232, 236, 240, 244, 248, 252, 256, 260,
                264, 268, 272, 276, 280, 284, 288, 292, 296, 300],
    "itemsize": 304,
})

RENDER_UNIFORMS_DTYPE = np.dtype({
    "names": ["vpMatrix", "vpMatrixInv", "viewMatrix", "projMatrix", "projMatrixInv", "modelMatrix",
              "modelViewMat", "mvpMatrix", "collisionSphereModelMatrix", "viewport", "cameraPosition",
              "colorMgmt", "lightAmbient", "lights", "shadowSourceModelViewMat", "shadowSourceProjMatrix",
              "shadowSourceMVP", "shadowSourcePosition", "shadowUsePCSS", "shadowPCFRadius", "shadowBias",
              "shadowStrength", "aoParams", "aoNumDirections", "aoNumSteps", "aoStrength", "aoPadding",
              "hairColor0", "hairSpecular", "hairColor1", "hairShadows", "hairWeightTT", "hairWeightTRT",
              "hairShift", "hairRoughness", "hairColorRng", "hairLumaRng", "hairAttenuation", "hairPadding",
              "sdfBoundsMin", "sdfBoundsMax", "gridBoundsMin", "gridBoundsMax", "gridDebugDisplayValue",
              "fiberRadius", "dbgShadowMapPreviewSize", "maxDrawnHairSegments", "displayMode",
              "bgColor0", "bgNoiseScale", "bgColor1", "bgGradientStrength", "collisionSpherePosition",
              "gizmoActiveState"],
    "formats": [M4, M4, M4, M4, M4, M4, M4, M4, M4, F4, F4, F4, F4, ("<f4", (3, 8)), M4, M4, M4, F4, "<u4",
                "<u4", "<f4", "<f4", F4, "<u4", "<u4", "<f4", "<f4", ("<f4", (3,)), "<f4", ("<f4", (3,)),
                "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", F4, F4, F4, ("<f4", (3,)),
                "<u4", "<f4", "<f4", "<u4", "<u4", ("<f4", (3,)), "<f4", ("<f4", (3,)), "<f4",
                ("<f4", (3,)), "<u4"],
    "offsets": [0, 64, 128, 192, 256, 320, 384, 448, 512, 576, 592, 608, 624, 640, 736, 800, 864, 928, 944,
                948, 952, 956, 960, 976, 980, 984, 988, 992, 1004, 1008, 1020, 1024, 1028, 1032, 1036, 1040,
                1044, 1048, 1052, 1056, 1072, 1088, 1104, 1116, 1120, 1124, 1128, 1132, 1136, 1148, 1152,
                1164, 1168, 1180],
    "itemsize": 1184,
})

HAIR_DATA_BYTES = 256  # WEBGPU_MINIMAL_BUFFER_SIZE (webgpu.ts:6), hairDataBuffer.ts:29-47

# DISPLAY_MODE (constants.ts:44-58)
DISPLAY_MODE_FINAL = 0
DISPLAY_MODE_TILES = 1
DISPLAY_MODE_TILES_PPLL = 2
DISPLAY_MODE_HW_RENDER = 3
DISPLAY_MODE_USED_SLICES = 4
DISPLAY_MODE_DEPTH = 5
DISPLAY_MODE_NORMALS = 6
DISPLAY_MODE_AO = 7


def col(*c):
    return [x / 255.0 for x in c]


@dataclass
class RenderConfig:
    """CONFIG.hairRender (constants.ts:193-265) and the render-side scene values."""
    fiberRadius: float = 0.0006
    tileSize: int = 8
    tileDepthBins: int = 32
    avgSegmentsPerTile: float = 128
    invalidTilesPerSegmentThreshold: int = 128
    sortBuckets: int = 64
    sortBucketSize: int = 16
    slicesPerPixel: int = 8
    avgFragmentsPerSlice: int = 16
    processorCount: int = 512
    alphaQuadratic: bool = False
    alphaMultiplier: float = 1.1
    shadingPoints: int = 64
    lodRenderPercent: float = 100.0
    displayMode: int = DISPLAY_MODE_FINAL
    dbgTileModeMaxSegments: int = 370
    dbgSlicesModeMaxSlices: int = 50
    dbgShowTiles: bool = False
    material: dict = field(default_factory=lambda: dict(
        color0=col(119, 43, 119), color1=col(76, 0, 255), specular=0.9, weightTT=0.0, weightTRT=1.4,
        shift=0.0, roughness=0.3, colorRng=0.1, lumaRng=0.1, attenuation=30.0, shadows=0.5))
    lights: list = field(default_factory=lambda: [
        dict(posPhi=60.0, posTheta=20.0, color=col(255, 244, 204), energy=0.8),
        dict(posPhi=100.0, posTheta=97.0, color=col(204, 249, 255), energy=0.8),
        dict(posPhi=-90.0, posTheta=30.0, color=col(255, 242, 204), energy=0.8),
    ])
    lightAmbient: tuple = (1.0, 1.0, 1.0, 0.05)
    shadows: dict = field(default_factory=lambda: dict(
        usePCSS=False, PCF_Radius=3, bias=0.0005, strength=0.4, textureSize=2048,
        posPhi=37.0, posTheta=45.0, distance=5.0, target=(0.0, 0.0, 0.0)))
    ao: dict = field(default_factory=lambda: dict(
        radius=2.0, directionOffset=0.0, falloffStart2=0.16, falloffEnd2=4.0, numDirections=12, numSteps=8,
        strength=0.4))
    colors: tuple = (2.2, 0.85, 1.0)  # gamma, exposure, ditherStrength
    camera_fov_deg: float = 30.0
    camera_near: float = 0.01
    camera_far: float = 100.0


@dataclass
class SimConfig:
    """CONFIG.hairSimulation (constants.ts:267-340)."""
    deltaTime: float = 1.0 / 30.0
    gravity: float = 0.03
    friction: float = 0.3
    volumePreservation: float = 0.00003
    collisionSphere: tuple = (-0.18, 1.56, 0.06, 0.06)
    constraintIterations: int = 7
    stiffnessLengthConstr: float = 1.0
    stiffnessGlobalConstr: float = 0.2
    globalExtent: float = 0.1
    globalFade: float = 0.75
    stiffnessLocalConstr: float = 0.3
    stiffnessCollisions: float = 1.0
    stiffnessSDF: float = 1.0
    windDirPhi: float = 18.0
    windDirTheta: float = 91.0
    windStrength: float = 0.0
    windStrengthLull: float = 0.75
    windStrengthFrequency: float = 1.8
    windStrengthJitter: float = 0.7
    windPhaseOffset: float = 0.45
    windColisionTraceOffset: float = 1.5
    gridDims: int = 64
    gridScale: float = 2.0
    sdfDistanceOffset: float = -0.0015


def pack_sim_uniforms(cfg: SimConfig, model: np.ndarray, sdf_bounds, grid_bounds, frame_idx: int) -> np.ndarray:
    """SimulationUniformsBuffer.update — simulationUniformsBuffer.ts:86-152."""
    u = np.zeros(1, dtype=SIM_UNIFORMS_DTYPE)
    inv = M.invert(model)
    u["modelMatrix"][0] = M.to_f32_colmajor(model)
    u["modelMatrixInv"][0] = M.to_f32_colmajor(inv)
    cs = cfg.collisionSphere
    cs_os = inv @ np.array([cs[0], cs[1], cs[2], 1.0])
    u["collisionSphere"][0] = [cs_os[0], cs_os[1], cs_os[2], cs[3]]
    wd = M.spherical_to_cartesian(cfg.windDirPhi, cfg.windDirTheta)
    wd_os = inv @ np.array([wd[0], wd[1], wd[2], 1.0])  # w = 1 as in :113 (vec4.transformMat4 of a vec3 slot)
    u["wind"][0] = [wd_os[0], wd_os[1], wd_os[2], cfg.windStrength]
    u["sdfBoundsMin"][0] = [*sdf_bounds[0], 0.5 + 2.0]
    u["sdfBoundsMax"][0] = [*sdf_bounds[1], cfg.sdfDistanceOffset]
    u["gridBoundsMin"][0] = [*grid_bounds[0], 0.5]
    u["gridBoundsMax"][0] = grid_bounds[1]
    u["gridDebugDisplayValue"] = 1 + 16
    u["deltaTime"] = cfg.deltaTime
    u["gravity"] = cfg.gravity
    u["constraintIterations"] = cfg.constraintIterations
    u["frameIdx"] = frame_idx
    u["windStrengthLull"] = cfg.windStrengthLull
    u["windColisionTraceOffset"] = cfg.windColisionTraceOffset
    u["stiffnessLengthConstr"] = cfg.stiffnessLengthConstr
    u["stiffnessCollisions"] = cfg.stiffnessCollisions
    u["stiffnessSDF"] = cfg.stiffnessSDF
    u["volumePreservation"] = cfg.volumePreservation
    u["friction"] = cfg.friction
    u["stiffnessGlobalConstr"] = cfg.stiffnessGlobalConstr
    u["globalConstrExtent"] = cfg.globalExtent
    u["globalConstrFade"] = cfg.globalFade
    u["stiffnessLocalConstr"] = cfg.stiffnessLocalConstr
    u["windPhaseOffset"] = cfg.windPhaseOffset
    u["windStrengthFrequency"] = cfg.windStrengthFrequency
    u["windStrengthJitter"] = cfg.windStrengthJitter
    return u


def segments_buffer_entries(cfg: RenderConfig, width: int, height: int) -> int:
    """getLengthOfHairTileSegmentsBuffer — hairTileSegmentsBuffer.ts:81-87."""
    tiles = math.ceil(width / cfg.tileSize) * math.ceil(height / cfg.tileSize)
    return 1 + math.ceil(tiles * cfg.avgSegmentsPerTile)


@dataclass
class Camera:
    eye: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)

    def view(self) -> np.ndarray:
        return M.look_at(self.eye, self.target, self.up)


def shadow_matrices(cfg: RenderConfig, model: np.ndarray, spheres: list) -> tuple:
    """getShadowSourceViewMatrix / ProjectionMatrix — getMVP_ShadowSourceMatrix.ts:14-79.
    `spheres` are the (center, radius) bounding spheres of the hair and meshes."""
    s = cfg.shadows
    pos = M.spherical_to_cartesian(s["posPhi"], s["posTheta"]) * s["distance"]
    target = M.project_point(model, s["target"])[:3]
    view = M.look_at(pos, target)
    mv = view @ model
    l, r, b, t = 9999999.0, -9999999.0, 9999999.0, -9999999.0
    for c, rad in spheres:
        p = mv @ np.array([c[0], c[1], c[2], 1.0])
        l, r = min(l, p[0] - rad), max(r, p[0] + rad)
        b, t = min(b, p[1] - rad), max(t, p[1] + rad)
    proj = M.ortho(l * 1.05, r * 1.05, b * 1.05, t * 1.05, 0.1, 20.0)
    return view, proj, pos


def pack_render_uniforms(cfg: RenderConfig, sim: SimConfig, camera: Camera, width: int, height: int,
                         model: np.ndarray, hair_sphere, mesh_spheres: list, sdf_bounds, grid_bounds) -> np.ndarray:
    """RenderUniformsBuffer.update — renderUniformsBuffer.ts:205-288."""
    u = np.zeros(1, dtype=RENDER_UNIFORMS_DTYPE)
    view = camera.view()
    proj = M.perspective(math.radians(cfg.camera_fov_deg), width / height, cfg.camera_near, cfg.camera_far)
    vp = proj @ view
    u["vpMatrix"][0] = M.to_f32_colmajor(vp)
    u["vpMatrixInv"][0] = M.to_f32_colmajor(M.invert(vp))
    u["viewMatrix"][0] = M.to_f32_colmajor(view)
    u["projMatrix"][0] = M.to_f32_colmajor(proj)
    u["projMatrixInv"][0] = M.to_f32_colmajor(M.invert(proj))
    u["modelMatrix"][0] = M.to_f32_colmajor(model)
    u["modelViewMat"][0] = M.to_f32_colmajor(view @ model)
    u["mvpMatrix"][0] = M.to_f32_colmajor(proj @ view @ model)
    cs = sim.collisionSphere
    r = cs[3] * 0.95
    u["collisionSphereModelMatrix"][0] = M.to_f32_colmajor(M.scale(M.translation(cs), (r, r, r)))
    u["viewport"][0] = [width, height, 0, 0]
    u["cameraPosition"][0] = [*camera.eye, 0]
    u["colorMgmt"][0] = [*cfg.colors, 0]
    u["lightAmbient"][0] = cfg.lightAmbient
    lights = np.zeros((3, 8), dtype=np.float32)
    for i, l in enumerate(cfg.lights):
        p = M.spherical_to_cartesian(l["posPhi"], l["posTheta"]) * 2.0
        lights[i] = [p[0], p[1], p[2], 0.0, *l["color"], l["energy"]]
    u["lights"][0] = lights
    sview, sproj, spos = shadow_matrices(cfg, model, [hair_sphere, *mesh_spheres])
    u["shadowSourceModelViewMat"][0] = M.to_f32_colmajor(sview @ model)
    u["shadowSourceProjMatrix"][0] = M.to_f32_colmajor(sproj)
    u["shadowSourceMVP"][0] = M.to_f32_colmajor(sproj @ sview @ model)
    u["shadowSourcePosition"][0] = [*spos, cfg.fiberRadius]
    s = cfg.shadows
    u["shadowUsePCSS"] = 1 if s["usePCSS"] else 0
    u["shadowPCFRadius"] = s["PCF_Radius"]
    u["shadowBias"] = s["bias"]
    u["shadowStrength"] = s["strength"]
    a = cfg.ao
    u["aoParams"][0] = [a["radius"], a["directionOffset"], a["falloffStart2"], a["falloffEnd2"]]
    u["aoNumDirections"] = a["numDirections"]
    u["aoNumSteps"] = a["numSteps"]
    u["aoStrength"] = a["strength"]
    m = cfg.material
    u["hairColor0"][0] = m["color0"]
    u["hairSpecular"] = m["specular"]
    u["hairColor1"][0] = m["color1"]
    u["hairShadows"] = m["shadows"]
    u["hairWeightTT"] = m["weightTT"]
    u["hairWeightTRT"] = m["weightTRT"]
    u["hairShift"] = m["shift"]
    u["hairRoughness"] = m["roughness"]
    u["hairColorRng"] = m["colorRng"]
    u["hairLumaRng"] = m["lumaRng"]
    u["hairAttenuation"] = m["attenuation"]
    u["sdfBoundsMin"][0] = [*sdf_bounds[0], 2.5]
    u["sdfBoundsMax"][0] = [*sdf_bounds[1], sim.sdfDistanceOffset]
    u["gridBoundsMin"][0] = [*grid_bounds[0], 0.5]
    u["gridBoundsMax"][0] = grid_bounds[1]
    u["gridDebugDisplayValue"] = 17
    u["fiberRadius"] = cfg.fiberRadius
    u["dbgShadowMapPreviewSize"] = math.floor(min(500, width / 3, height / 3))
    u["maxDrawnHairSegments"] = segments_buffer_entries(cfg, width, height)
    u["displayMode"] = encode_debug_mode(cfg)
    u["bgColor0"][0] = col(22, 162, 188)
    u["bgNoiseScale"] = 5.0
    u["bgColor1"][0] = col(14, 103, 120)
    u["bgGradientStrength"] = 0.5
    u["collisionSpherePosition"][0] = cs[:3]
    return u


def encode_debug_mode(cfg: RenderConfig) -> int:
    """encodeDebugMode — renderUniformsBuffer.ts:396-414: the display mode in
    the low 8 bits, the mode's extra value above them."""
    extra = 0
    if cfg.displayMode in (DISPLAY_MODE_TILES, DISPLAY_MODE_TILES_PPLL):
        extra = cfg.dbgTileModeMaxSegments
    elif cfg.displayMode == DISPLAY_MODE_USED_SLICES:
        extra = cfg.dbgSlicesModeMaxSlices
    elif cfg.displayMode == DISPLAY_MODE_FINAL:
        extra = 1 if cfg.dbgShowTiles else 0
    return cfg.displayMode | (extra << 8)


def pack_hair_data(center, radius: float, strands: int, points: int) -> np.ndarray:
    """createHairDataBuffer — hairDataBuffer.ts:29-47 (256-byte buffer)."""
    b = np.zeros(HAIR_DATA_BYTES // 4, dtype=np.uint32)
    f = b.view(np.float32)
    f[0:3] = center
    f[3] = radius
    b[4] = strands
    b[5] = points
    return b


def render_params(cfg: RenderConfig, sim: SimConfig, camera: Camera, width: int, height: int, model: np.ndarray,
                  hair_sphere, mesh_spheres: list, sdf_bounds, grid_bounds):
    """The fh_render_params of the same inputs as pack_render_uniforms (for
    the library's own packer). Keeps the mesh-sphere array alive on the
    returned struct."""
    from . import native as N
    p = N.RenderParams()
    p.width, p.height = width, height
    p.cameraEye[:] = list(camera.eye)
    p.cameraTarget[:] = list(camera.target)
    p.cameraUp[:] = list(camera.up)
    p.fovDeg, p.near, p.far = cfg.camera_fov_deg, cfg.camera_near, cfg.camera_far
    p.modelMatrix[:] = list(np.asarray(model, dtype=np.float64).T.reshape(16))
    p.hairSphere[:] = [*hair_sphere[0], hair_sphere[1]]
    spheres = (N.d * (4 * max(1, len(mesh_spheres))))(*[x for c, r in mesh_spheres for x in (*c, r)])
    p.meshSpheres = spheres
    p.meshSphereCount = len(mesh_spheres)
    p._keep = spheres
    p.tileSize, p.avgSegmentsPerTile = cfg.tileSize, cfg.avgSegmentsPerTile
    p.sdfBoundsMin[:], p.sdfBoundsMax[:] = list(sdf_bounds[0]), list(sdf_bounds[1])
    p.sdfDistanceOffset = sim.sdfDistanceOffset
    p.gridBoundsMin[:], p.gridBoundsMax[:] = list(grid_bounds[0]), list(grid_bounds[1])
    p.collisionSphere[:] = list(sim.collisionSphere)
    p.fiberRadius = cfg.fiberRadius
    p.displayMode, p.dbgTileModeMaxSegments = cfg.displayMode, cfg.dbgTileModeMaxSegments
    p.dbgSlicesModeMaxSlices, p.dbgShowTiles = cfg.dbgSlicesModeMaxSlices, 1 if cfg.dbgShowTiles else 0
    p.colors[:] = list(cfg.colors)
    p.lightAmbient[:] = list(cfg.lightAmbient)
    for i, l in enumerate(cfg.lights):
        p.lights[i].posPhi, p.lights[i].posTheta = l["posPhi"], l["posTheta"]
        p.lights[i].color[:] = list(l["color"])
        p.lights[i].energy = l["energy"]
    s = cfg.shadows
    p.shadowUsePCSS, p.shadowPCFRadius = 1 if s["usePCSS"] else 0, s["PCF_Radius"]
    p.shadowBias, p.shadowStrength = s["bias"], s["strength"]
    p.shadowPosPhi, p.shadowPosTheta, p.shadowDistance = s["posPhi"], s["posTheta"], s["distance"]
    p.shadowTarget[:] = list(s["target"])
    a = cfg.ao
    p.aoRadius, p.aoDirectionOffset = a["radius"], a["directionOffset"]
    p.aoFalloffStart2, p.aoFalloffEnd2 = a["falloffStart2"], a["falloffEnd2"]
    p.aoNumDirections, p.aoNumSteps, p.aoStrength = a["numDirections"], a["numSteps"], a["strength"]
    m = cfg.material
    p.hairColor0[:], p.hairColor1[:] = list(m["color0"]), list(m["color1"])
    for k in ("specular", "shadows", "weightTT", "weightTRT", "shift", "roughness", "colorRng", "lumaRng",
              "attenuation"):
        setattr(p, k, m[k])
    return p


def sim_params(cfg: SimConfig, model: np.ndarray, sdf_bounds, grid_bounds, frame_idx: int):
    """The fh_sim_params of the same inputs as pack_sim_uniforms."""
    from . import native as N
    p = N.SimParams()
    p.modelMatrix[:] = list(np.asarray(model, dtype=np.float64).T.reshape(16))
    p.collisionSphere[:] = list(cfg.collisionSphere)
    p.windDirPhi, p.windDirTheta, p.windStrength = cfg.windDirPhi, cfg.windDirTheta, cfg.windStrength
    p.sdfBoundsMin[:], p.sdfBoundsMax[:] = list(sdf_bounds[0]), list(sdf_bounds[1])
    p.sdfDistanceOffset = cfg.sdfDistanceOffset
    p.gridBoundsMin[:], p.gridBoundsMax[:] = list(grid_bounds[0]), list(grid_bounds[1])
    p.frameIdx, p.constraintIterations = frame_idx, cfg.constraintIterations
    for k in ("deltaTime", "gravity", "windStrengthLull", "windColisionTraceOffset", "stiffnessLengthConstr",
              "stiffnessCollisions", "stiffnessSDF", "volumePreservation", "friction", "stiffnessGlobalConstr",
              "stiffnessLocalConstr", "windPhaseOffset", "windStrengthFrequency", "windStrengthJitter"):
        setattr(p, k, getattr(cfg, k))
    p.globalConstrExtent, p.globalConstrFade = cfg.globalExtent, cfg.globalFade
    return p
