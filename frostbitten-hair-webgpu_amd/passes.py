"""Host mirror of the reference's pass API (src/passes/**, src/scene/hair,
src/renderer.ts) over the HIP C-ABI. Class and method names follow the
reference so a caller drives the chain the way Renderer does:

    GridPreSimPass.cmdUpdateGridsBeforeSim(ctx, hairObject)
    HairSimIntegrationPass.cmdSimulateHairPositions(ctx, hairObject)
    GridPostSimPass.cmdUpdateGridsAfterSim(ctx, hairObject)
    HairTilesPass.cmdClearBeforeRender(ctx) / cmdDrawHairToTiles(ctx, hairObject)
    HairTileSortPass.cmdClearBeforeRender(ctx) / cmdSortHairTiles(ctx)
    HairFinePass.cmdClearBeforeRender(ctx) / cmdRasterizeSlicesHair(ctx, hairObject)
    HairShadingPass.cmdComputeShadingPoints(ctx, hairObject)
    HairCombinePass.cmdCombineRasterResults(ctx)
    PresentPass.cmdDraw(ctx, screenTexture)

Device memory is torch-allocated; every pass enqueues on the context's stream
and returns, like recording into a GPUCommandEncoder.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import assets as A
from . import layouts as L
from . import native as N


def _dev_bytes(arr: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1).copy()).to(device)


def _dev(arr: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(arr).copy()).to(device)


class HairObject:
    """src/scene/hair/hairObject.ts:29-155 — strand buffers and the tick-tock
    position pair (swap after integration, :102-106)."""

    def __init__(self, name: str, positions: np.ndarray, strands: int, points: int, device,
                 derived: dict | None = None):
        self.name = name
        self.strandsCount = strands
        self.pointsPerStrand = points
        if derived is None:
            derived = dict(box=A.bounding_box(positions), tangents=A.tangents(positions, strands, points),
                           seglen=A.segment_lengths(positions, strands, points))
        self.box = derived["box"]
        self.sphere = A.bounding_sphere(self.box)
        self.initialPositions = _dev(positions.astype(np.float32), device)
        self.positions = [self.initialPositions.clone(), self.initialPositions.clone()]
        self.tangents = _dev(derived["tangents"], device)
        self.segmentLengths = _dev(derived["seglen"], device)
        self.hairData = _dev(L.pack_hair_data(self.sphere[0], self.sphere[1], strands, points), device)
        self.shading = None
        self.current = 0
        self.lodRenderPercent = 100.0

    def allocShading(self, shading_points: int, device) -> None:
        self.shading = _dev(A.initial_shading(self.strandsCount, shading_points), device)

    def getRenderedStrandCount(self) -> int:
        """hairObject.ts:52-57"""
        pct = min(max(self.lodRenderPercent, 0), 100)
        return min(max(math.ceil(self.strandsCount * pct / 100.0), 0), self.strandsCount)

    @property
    def positionsNow(self) -> torch.Tensor:
        return self.positions[self.current]

    @property
    def positionsPrev(self) -> torch.Tensor:
        return self.positions[1 - self.current]

    def swapPositionBuffersAfterSimIntegration(self) -> None:
        self.current = 1 - self.current

    def resetSimulation(self) -> None:
        """hairObject.ts:75-91"""
        self.positions[0].copy_(self.initialPositions)
        self.positions[1].copy_(self.initialPositions)


class _PositionsView:
    """A hair object whose `positionsNow` is another buffer (a copy of one
    frame's positions); every other attribute is the hair object's."""

    def __init__(self, hair: HairObject, now: torch.Tensor):
        self._hair = hair
        self.positionsNow = now

    def __getattr__(self, name):
        return getattr(self._hair, name)


class GridData:
    """src/passes/simulation/grids/gridData.ts:14-107"""

    def __init__(self, hair_box, dims: int, scale: float, device):
        self.dims = dims
        self.bounds = A.scale_box(hair_box, scale)
        n = dims ** 3
        self.densityVelocity = torch.zeros(n * 4, dtype=torch.int32, device=device)
        self.densityGradAndWind = torch.zeros(n * 4, dtype=torch.float32, device=device)


class SDFCollider:
    """src/scene/sdfCollider/sdfCollider.ts:9-92"""

    def __init__(self, dims: int, bounds, distances: np.ndarray, device):
        self.dims = dims
        self.bounds = bounds
        self.texture = _dev(distances.astype(np.float32), device)


@dataclass
class PassCtx:
    """src/passes/passCtx.ts:9-35 (fields used by the hair chain)."""
    native: N.Context
    cfg: N.RenderCfg
    viewport: tuple
    frameIdx: int
    globalUniforms: torch.Tensor
    simulationUniforms: torch.Tensor
    physicsForcesGrid: GridData
    sdfCollider: SDFCollider
    depthTexture: N.Tex2D          # mesh depth (drawMeshesPass)
    shadingDepthTexture: N.Tex2D   # mesh + hair depth (after hwHairPass)
    aoTexture: N.Tex2D
    shadowDepthTexture: N.Tex2D
    hairTilesBuffer: torch.Tensor = None
    hairTileSegmentsBuffer: torch.Tensor = None
    hairSegmentCountPerTileBuffer: torch.Tensor = None
    hairTileListBuffer: torch.Tensor = None
    hairRasterizerResultsBuffer: torch.Tensor = None
    hdrRenderTexture: torch.Tensor = None  # rgba16float, (H, W, 4) half bits
    normalsTexture: torch.Tensor = None    # rg16float, (H, W) u32 of two halves, or None


class GridPreSimPass:
    NAME = "GridPreSimPass"

    def cmdUpdateGridsBeforeSim(self, ctx: PassCtx, hair: HairObject) -> None:
        g, s = ctx.physicsForcesGrid, ctx.sdfCollider
        ctx.native.call("fh_grid_pre_sim", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(g.densityVelocity),
                        N.ptr(g.densityGradAndWind), N.ptr(s.texture), s.dims)


class HairSimIntegrationPass:
    NAME = "HairSimIntegrationPass"

    def cmdSimulateHairPositions(self, ctx: PassCtx, hair: HairObject) -> None:
        g, s = ctx.physicsForcesGrid, ctx.sdfCollider
        ctx.native.call("fh_sim_integrate", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(hair.hairData),
                        N.ptr(hair.positionsPrev), N.ptr(hair.positionsNow), N.ptr(hair.segmentLengths),
                        N.ptr(s.texture), s.dims, N.ptr(g.densityVelocity), N.ptr(g.densityGradAndWind),
                        N.ptr(hair.initialPositions), hair.strandsCount, hair.pointsPerStrand)
        hair.swapPositionBuffersAfterSimIntegration()  # hairSimIntegrationPass.ts:59


class GridPostSimPass:
    NAME = "GridPostSimPass"

    def cmdUpdateGridsAfterSim(self, ctx: PassCtx, hair: HairObject) -> None:
        g = ctx.physicsForcesGrid
        ctx.native.call("fh_grid_post_sim", g.dims, N.ptr(ctx.simulationUniforms), N.ptr(hair.hairData),
                        N.ptr(hair.positionsPrev), N.ptr(hair.positionsNow), N.ptr(g.densityVelocity),
                        hair.strandsCount, hair.pointsPerStrand)


class HairTilesPass:
    NAME = "HairTilesPass"

    def __init__(self):
        self.hairTilesBuffer = None
        self.hairTileSegmentsBuffer = None
        self.segmentCountPerTileBuffer = None
        self.slotTileBinBuffer = None
        self.segmentsCapacity = 0

    def onViewportResize(self, ctx_cfg: N.RenderCfg, rc: L.RenderConfig, width: int, height: int, device) -> None:
        tiles = math.ceil(width / rc.tileSize) * math.ceil(height / rc.tileSize)
        entries = L.segments_buffer_entries(rc, width, height)
        self.hairTilesBuffer = torch.zeros(tiles * rc.tileDepthBins * 4, dtype=torch.int32, device=device)
        # DrawnHairSegments: (1 + cnt) * 8 bytes; the element array at byte 4 holds cnt elements
        self.hairTileSegmentsBuffer = torch.zeros(entries * 2, dtype=torch.int32, device=device)
        self.segmentsCapacity = entries - 1
        self.segmentCountPerTileBuffer = torch.zeros(tiles, dtype=torch.int32, device=device)
        self.slotTileBinBuffer = torch.zeros(2 * max(1, self.segmentsCapacity), dtype=torch.int32, device=device)

    def cmdClearBeforeRender(self, ctx: PassCtx) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_tiles_clear", ctx.cfg, w, h, N.ptr(self.hairTilesBuffer),
                        N.ptr(self.hairTileSegmentsBuffer), N.ptr(self.segmentCountPerTileBuffer))

    def cmdDrawHairToTiles(self, ctx: PassCtx, hair: HairObject) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_tiles_draw", ctx.cfg, w, h, N.ptr(ctx.globalUniforms), N.ptr(hair.hairData),
                        N.ptr(hair.positionsNow), N.ptr(hair.tangents), N.ptr(self.hairTilesBuffer),
                        ctx.depthTexture, N.ptr(self.hairTileSegmentsBuffer), self.segmentsCapacity,
                        N.ptr(self.segmentCountPerTileBuffer), N.ptr(self.slotTileBinBuffer),
                        hair.getRenderedStrandCount(), hair.pointsPerStrand)


class HairTileSortPass:
    NAME = "HairTileSortPass"

    def __init__(self, rc: L.RenderConfig, device):
        self.bucketsDataBuffer = torch.zeros(rc.sortBuckets * 2, dtype=torch.int32, device=device)
        self.tileListBuffer = None

    def onViewportResize(self, rc: L.RenderConfig, width: int, height: int, device) -> None:
        tiles = math.ceil(width / rc.tileSize) * math.ceil(height / rc.tileSize)
        self.tileListBuffer = torch.zeros(4 + tiles, dtype=torch.int32, device=device)

    def cmdClearBeforeRender(self, ctx: PassCtx) -> None:
        ctx.native.call("fh_tiles_sort_clear", ctx.cfg, N.ptr(self.tileListBuffer), N.ptr(self.bucketsDataBuffer))

    def cmdSortHairTiles(self, ctx: PassCtx) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_tiles_sort", ctx.cfg, w, h, N.ptr(ctx.hairSegmentCountPerTileBuffer),
                        N.ptr(self.tileListBuffer), N.ptr(self.bucketsDataBuffer))


class HairFinePass:
    NAME = "HairFinePass"

    def __init__(self, rc: L.RenderConfig, device):
        per_proc = rc.avgFragmentsPerSlice * rc.slicesPerPixel * rc.tileSize * rc.tileSize
        self.hairSlicesDataBuffer = torch.zeros(rc.processorCount * per_proc * 4, dtype=torch.int32, device=device)
        self.hairRasterizerResultsBuffer = None

    def onViewportResize(self, width: int, height: int, device) -> None:
        self.hairRasterizerResultsBuffer = torch.zeros(width * height * 4, dtype=torch.float32, device=device)

    def cmdClearBeforeRender(self, ctx: PassCtx) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_fine_clear", w, h, N.ptr(self.hairRasterizerResultsBuffer))

    def cmdRasterizeSlicesHair(self, ctx: PassCtx, hair: HairObject, tiles: HairTilesPass) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_fine_raster", ctx.cfg, w, h, N.ptr(ctx.globalUniforms), N.ptr(hair.hairData),
                        N.ptr(hair.positionsNow), N.ptr(ctx.hairTilesBuffer), N.ptr(ctx.hairTileSegmentsBuffer),
                        tiles.segmentsCapacity, N.ptr(self.hairSlicesDataBuffer),
                        N.ptr(self.hairRasterizerResultsBuffer), ctx.depthTexture, N.ptr(hair.shading),
                        N.ptr(hair.tangents), N.ptr(ctx.hairTileListBuffer), N.ptr(tiles.slotTileBinBuffer),
                        hair.strandsCount, hair.pointsPerStrand)

    def image(self, width: int, height: int) -> torch.Tensor:
        """Pixels of the rasterizer result as (H, W, 4), y up; the last pixel
        lies past the end of the array and reads as 0 (hairRasterizerResultBuffer.ts:55-57)."""
        flat = self.hairRasterizerResultsBuffer.view(-1, 4)[1:]
        out = torch.zeros(width * height, 4, dtype=torch.float32, device=flat.device)
        out[: width * height - 1] = flat[: width * height - 1]
        return out.view(height, width, 4)


class HairShadingPass:
    NAME = "HairShadingPass"

    def cmdComputeShadingPoints(self, ctx: PassCtx, hair: HairObject) -> None:
        ctx.native.call("fh_shading", ctx.cfg, N.ptr(ctx.globalUniforms), N.ptr(hair.hairData),
                        N.ptr(hair.positionsNow), N.ptr(hair.tangents), N.ptr(hair.shading),
                        ctx.shadowDepthTexture, ctx.aoTexture, ctx.shadingDepthTexture,
                        hair.strandsCount, hair.pointsPerStrand)


class HairCombinePass:
    """src/passes/hairCombine/hairCombinePass.ts:178-270."""
    NAME = "HairCombinePass"

    def cmdCombineRasterResults(self, ctx: PassCtx) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_hair_combine", ctx.cfg, w, h, N.ptr(ctx.globalUniforms), N.ptr(ctx.hairTilesBuffer),
                        N.ptr(ctx.hairTileSegmentsBuffer), N.ptr(ctx.hairRasterizerResultsBuffer),
                        N.ptr(ctx.hairSegmentCountPerTileBuffer), N.ptr(ctx.hdrRenderTexture))

    def cmdCombineAndPresent(self, ctx: PassCtx, present: "PresentPass", screenTexture: torch.Tensor,
                             scene: torch.Tensor | None = None, rowBytes: int = 0) -> None:
        """cmdCombineRasterResults, then present.cmdDraw(ctx, screenTexture),
        as one pass over the texels (fh_combine_present), with the same
        results. `scene` (None: the HDR texture itself) is the HDR scene
        blended onto, read in place of a copy into the HDR texture."""
        w, h = ctx.viewport
        ctx.native.call("fh_combine_present", ctx.cfg, w, h, N.ptr(ctx.globalUniforms), N.ptr(ctx.hairTilesBuffer),
                        N.ptr(ctx.hairTileSegmentsBuffer), N.ptr(ctx.hairRasterizerResultsBuffer),
                        N.ptr(ctx.hairSegmentCountPerTileBuffer), N.ptr(scene) if scene is not None else None,
                        N.ptr(ctx.hdrRenderTexture), ctx.shadingDepthTexture, ctx.aoTexture,
                        N.ptr(ctx.normalsTexture), 1 if present.ditherPerPixel else 0, N.ptr(screenTexture), rowBytes)


class PresentPass:
    """src/passes/presentPass/presentPass.ts:125-210. `ditherPerPixel` False
    reads Bayer element 0 for every pixel, as the Deno build does
    (dither.wgsl.ts:34); True indexes the matrix by pixel (browser build)."""
    NAME = "PresentPass"

    def __init__(self, ditherPerPixel: bool = False):
        self.ditherPerPixel = ditherPerPixel

    def cmdDraw(self, ctx: PassCtx, screenTexture: torch.Tensor, rowBytes: int = 0) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_present", w, h, N.ptr(ctx.globalUniforms), N.ptr(ctx.hdrRenderTexture),
                        ctx.shadingDepthTexture, ctx.aoTexture, N.ptr(ctx.normalsTexture),
                        1 if self.ditherPerPixel else 0, N.ptr(screenTexture), rowBytes)


class ShadowMapPass:
    """src/passes/shadowMapPass/shadowMapPass.ts:83-133: the meshes (drawn
    first, :94-99), then the hair (:100-115) into the shadow map. With
    `meshes` None the meshes' depth is the precomputed `meshShadowDepth`."""
    NAME = "ShadowMapPass"

    def cmdUpdateShadowMap(self, ctx: PassCtx, hair: HairObject, meshShadowDepth: torch.Tensor | None = None,
                           meshes: "N.Meshes | None" = None) -> None:
        sm = ctx.shadowDepthTexture
        if meshes is not None:
            ctx.native.call("fh_shadow_map_meshes", N.ptr(ctx.globalUniforms), meshes.handle, N.Tex2D.ptr(sm),
                            sm.width)
        else:
            ctx.native.call("fh_copy", N.Tex2D.ptr(sm), N.ptr(meshShadowDepth), meshShadowDepth.numel() * 4)
        ctx.native.call("fh_shadow_map_hair", N.ptr(ctx.globalUniforms), N.ptr(hair.positionsNow),
                        N.ptr(hair.tangents), N.Tex2D.ptr(sm), sm.width, hair.getRenderedStrandCount(),
                        hair.pointsPerStrand)


class DrawBackgroundGradientPass:
    """src/passes/drawBackgroundGradient/drawBackgroundGradientPass.ts (the
    gradient with its noise into the HDR texture)."""
    NAME = "DrawBackgroundGradientPass"

    def cmdDraw(self, ctx: PassCtx) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_draw_background", w, h, N.ptr(ctx.globalUniforms), N.ptr(ctx.hdrRenderTexture))


class DrawMeshesPass:
    """src/passes/drawMeshes/drawMeshesPass.ts:58-115: the scene's meshes
    into the depth, normals and HDR textures (PBR colour with the shadow map
    and the AO texture)."""
    NAME = "DrawMeshesPass"

    def cmdDrawMeshes(self, ctx: PassCtx, meshes: "N.Meshes", depth: torch.Tensor, normals: torch.Tensor,
                      visibility: torch.Tensor | None = None) -> None:
        """The whole pass; with `visibility`, only its depth, normals and
        per-pixel triangle (cmdShadeMeshes adds the colour)."""
        w, h = ctx.viewport
        if visibility is not None:
            ctx.native.call("fh_draw_meshes_visibility", w, h, N.ptr(ctx.globalUniforms), meshes.handle,
                            N.ptr(depth), N.ptr(normals), N.ptr(visibility))
            return
        ctx.native.call("fh_draw_meshes", w, h, N.ptr(ctx.globalUniforms), meshes.handle, ctx.shadowDepthTexture,
                        ctx.aoTexture, N.ptr(depth), N.ptr(normals), N.ptr(ctx.hdrRenderTexture))

    def cmdShadeMeshes(self, ctx: PassCtx, meshes: "N.Meshes", visibility: torch.Tensor) -> None:
        w, h = ctx.viewport
        ctx.native.call("fh_shade_meshes", w, h, N.ptr(ctx.globalUniforms), meshes.handle, N.ptr(visibility),
                        ctx.shadowDepthTexture, ctx.aoTexture, N.ptr(ctx.hdrRenderTexture))


def upload_meshes(native: N.Context, mesh_list: list, device) -> "N.Meshes":
    """An fh_meshes object list from host meshes (assets.scene_meshes)."""
    dev = []
    for m in mesh_list:
        dev.append({"positions": _dev(m["positions"].astype(np.float32), device),
                    "normals": _dev(m["normals"].astype(np.float32), device),
                    "indices": _dev(m["indices"].astype(np.int32), device), "preview": m.get("preview", False)})
    return N.Meshes(native, dev)


class HwHairPass:
    """src/passes/hwHair/hwHairPass.ts:79-141: the strands' depth and normals
    drawn over the meshes' (both restored from the inputs first)."""
    NAME = "HwHairPass"

    def cmdDrawHair(self, ctx: PassCtx, hair: HairObject, meshDepth: torch.Tensor,
                    meshNormals: torch.Tensor) -> None:
        w, h = ctx.viewport
        d = ctx.shadingDepthTexture
        ctx.native.call("fh_copy", N.Tex2D.ptr(d), N.ptr(meshDepth), meshDepth.numel() * 4)
        ctx.native.call("fh_copy", N.ptr(ctx.normalsTexture), N.ptr(meshNormals), meshNormals.numel() * 4)
        ctx.native.call("fh_hw_hair", w, h, N.ptr(ctx.globalUniforms), N.ptr(hair.positionsNow),
                        N.ptr(hair.tangents), N.Tex2D.ptr(d), N.ptr(ctx.normalsTexture),
                        hair.getRenderedStrandCount(), hair.pointsPerStrand)


class AoPass:
    """src/passes/aoPass/aoPass.ts:45-86 (GTAO at half resolution)."""
    NAME = "AoPass"

    def cmdCalcAo(self, ctx: PassCtx) -> None:
        ao = ctx.aoTexture
        ctx.native.call("fh_calc_ao", N.ptr(ctx.globalUniforms), ctx.shadingDepthTexture, N.ptr(ctx.normalsTexture),
                        N.Tex2D.ptr(ao), ao.width, ao.height)


class HairRenderer:
    """The hair part of src/renderer.ts: beforeFirstFrame (:145-169),
    cmdRender -> simulateHair (:212-240) -> cmdDrawScene (:242-283) ->
    updateResourcesForNextFrame (:285-295), with the hardware-raster passes'
    outputs (depth, AO, shadow map) provided as precomputed inputs."""

    def __init__(self, wl: A.Workload, device="cuda", native: N.Context | None = None, shadow_size: int = 2048,
                 inputs: dict | None = None, hw_passes: bool = False, branches: bool = True,
                 mesh_passes: bool = False):
        self.wl = wl
        # branches: passes that neither read what the tile, sort and fine
        # passes write nor write what they read run beside them on the
        # context's second stream (fh_branch_*)
        self.branches = branches
        self.device = torch.device(device)
        self.native = native or N.Context(self.device.index or 0)
        self.native.set_stream(torch.cuda.current_stream(self.device))
        rc, sc = wl.render, wl.sim
        self.cfg = N.RenderCfg.from_config(rc)
        self.model = np.eye(4)
        self.hair = HairObject("hair", wl.positions, wl.strands, wl.points, self.device, derived=A.derived(wl))
        self.hair.allocShading(rc.shadingPoints, self.device)
        self.grid = GridData(self.hair.box, sc.gridDims, sc.gridScale, self.device)
        self.sdf = SDFCollider(wl.sdf_dims, wl.sdf_bounds, wl.sdf, self.device)
        self.frameIdx = 0
        W, H = wl.width, wl.height
        self.render_uniforms = workload_render_uniforms(wl, self.model)
        self.globalUniforms = _dev_bytes(self.render_uniforms, self.device)
        self.simulationUniforms = _dev_bytes(self.sim_uniforms(0), self.device)
        if inputs is None:
            inputs = make_inputs(wl, self.render_uniforms, shadow_size)
        self.inputs = {k: _dev(v, self.device) for k, v in inputs.items()}
        self.tilesPass = HairTilesPass()
        self.tilesPass.onViewportResize(self.cfg, rc, W, H, self.device)
        self.sortPass = HairTileSortPass(rc, self.device)
        self.sortPass.onViewportResize(rc, W, H, self.device)
        self.finePass = HairFinePass(rc, self.device)
        self.finePass.onViewportResize(W, H, self.device)
        self.gridPreSimPass = GridPreSimPass()
        self.simPass = HairSimIntegrationPass()
        self.gridPostSimPass = GridPostSimPass()
        self.shadingPass = HairShadingPass()
        self.hairCombinePass = HairCombinePass()
        self.presentPass = PresentPass()
        # the HDR scene texture, restored each frame from the scene the
        # out-of-scope background and mesh passes would draw; the screen
        # texture receives the presented rgba8 sRGB image
        self.hdrRenderTexture = torch.empty_like(self.inputs["hdrScene"])
        self.screenTexture = torch.zeros((H, W, 4), dtype=torch.uint8, device=self.device)
        # hw_passes: the hair's depth, normals, shadow-map draw and AO are
        # computed every frame (HwHairPass, ShadowMapPass, AoPass) from the
        # meshes' inputs instead of taken precomputed
        # mesh_passes (implies hw_passes): the background, the meshes' depth,
        # normals and colour and the meshes in the shadow map are drawn every
        # frame too (DrawBackgroundGradientPass, DrawMeshesPass, ShadowMapPass)
        # from the scene's meshes (assets.scene_meshes), so no input is
        # precomputed but the hair and the colliders themselves
        self.mesh_passes = mesh_passes
        hw_passes = hw_passes or mesh_passes
        self.hw_passes = hw_passes
        if mesh_passes:
            self.meshList = A.scene_meshes(wl.collider, wl.sim.collisionSphere)
            self.meshes = upload_meshes(self.native, self.meshList, self.device)
            self.backgroundPass = DrawBackgroundGradientPass()
            self.drawMeshesPass = DrawMeshesPass()
            self.meshDepthTexture = torch.ones_like(self.inputs["meshDepth"])
            self.meshNormalsTexture = torch.zeros_like(self.inputs["meshNormals"])
            # for the two-call form of the mesh pass (cmdDrawMeshes with visibility)
            self.meshVisibility = torch.zeros_like(self.inputs["meshNormals"])
        if hw_passes:
            self.shadowMapPass = ShadowMapPass()
            self.hwHairPass = HwHairPass()
            self.aoPass = AoPass()
            # zero-initialised like every WebGPU texture: the shading pass of
            # beforeFirstFrame reads them before any frame has drawn them
            self.depthTexture = torch.zeros_like(self.inputs["meshDepth"])
            self.normalsTexture = torch.zeros_like(self.inputs["meshNormals"])
            self.aoTexture = torch.zeros(((H + 1) // 2, (W + 1) // 2), dtype=torch.float32, device=self.device)
            self.shadowTexture = torch.zeros_like(self.inputs["shadowMap"])

    def sim_uniforms(self, frame_idx: int) -> np.ndarray:
        return L.pack_sim_uniforms(self.wl.sim, self.model, self.wl.sdf_bounds, self.grid.bounds, frame_idx)

    def ctx(self) -> PassCtx:
        W, H = self.wl.width, self.wl.height
        i = self.inputs
        if self.hw_passes:
            depth, ao, shadow = self.depthTexture, self.aoTexture, self.shadowTexture
        else:
            depth, ao, shadow = i["hairDepth"], i["ao"], i["shadowMap"]
        return PassCtx(
            native=self.native, cfg=self.cfg, viewport=(W, H), frameIdx=self.frameIdx,
            globalUniforms=self.globalUniforms, simulationUniforms=self.simulationUniforms,
            physicsForcesGrid=self.grid, sdfCollider=self.sdf,
            depthTexture=N.tex(self.meshDepthTexture if self.mesh_passes else i["meshDepth"], W, H),
            shadingDepthTexture=N.tex(depth, W, H),
            aoTexture=N.tex(ao, ao.shape[1], ao.shape[0]),
            shadowDepthTexture=N.tex(shadow, shadow.shape[1], shadow.shape[0]),
            normalsTexture=self.normalsTexture if self.hw_passes else None,
            hairTilesBuffer=self.tilesPass.hairTilesBuffer,
            hairTileSegmentsBuffer=self.tilesPass.hairTileSegmentsBuffer,
            hairSegmentCountPerTileBuffer=self.tilesPass.segmentCountPerTileBuffer,
            hairTileListBuffer=self.sortPass.tileListBuffer,
            hairRasterizerResultsBuffer=self.finePass.hairRasterizerResultsBuffer,
            hdrRenderTexture=self.hdrRenderTexture)

    def upload_sim_uniforms(self) -> None:
        u = torch.from_numpy(self.sim_uniforms(self.frameIdx).view(np.uint8).reshape(-1).copy())
        self.simulationUniforms.copy_(u, non_blocking=False)

    def beforeFirstFrame(self) -> None:
        ctx = self.ctx()
        self.gridPostSimPass.cmdUpdateGridsAfterSim(ctx, self.hair)
        self.shadingPass.cmdComputeShadingPoints(ctx, self.hair)

    def _branch_begin(self) -> None:
        if self.branches:
            self.native.branch_begin()

    def _branch_end(self) -> None:
        if self.branches:
            self.native.branch_end()

    def join_branch(self) -> None:
        if self.branches:
            self.native.branch_join()

    def use_uniform_ring(self, ring: torch.Tensor, start: int) -> None:
        """Per-frame simulation uniforms from a device ring (row k: the bytes
        of frame k's uniforms): every simulation step then starts by copying
        row (counter mod rows) into `simulationUniforms` on the device and
        advancing the counter (fh_copy_indexed), starting at row `start`, so
        that frames replayed from graphs need no host copy between them."""
        self.uniformRing = ring.contiguous()
        self.uniformCounter = torch.tensor([start], dtype=torch.int32, device=self.device)

    def _next_uniforms(self) -> None:
        ring = getattr(self, "uniformRing", None)
        if ring is not None:
            self.native.call("fh_copy_indexed", N.ptr(self.simulationUniforms), N.ptr(ring), ring.shape[1],
                             ring.shape[0], N.ptr(self.uniformCounter))

    def simulateHair(self, ctx: PassCtx) -> None:
        self._next_uniforms()
        self.gridPreSimPass.cmdUpdateGridsBeforeSim(ctx, self.hair)
        self.simPass.cmdSimulateHairPositions(ctx, self.hair)
        # the grid accumulation feeds only the next frame's grid passes
        self._branch_begin()
        self.gridPostSimPass.cmdUpdateGridsAfterSim(ctx, self.hair)
        self._branch_end()

    def simulateStep(self, ctx: PassCtx) -> None:
        """simulateHair's passes one after another on the current stream."""
        self._next_uniforms()
        self.gridPreSimPass.cmdUpdateGridsBeforeSim(ctx, self.hair)
        self.simPass.cmdSimulateHairPositions(ctx, self.hair)
        self.gridPostSimPass.cmdUpdateGridsAfterSim(ctx, self.hair)

    def drawScene(self, ctx: PassCtx) -> None:
        """The HDR scene of the background and mesh passes (renderer.ts:187-188,
        :244-246). Without mesh_passes their output is an input, copied into
        the HDR texture here for the combine pass alone (drawHair's fused
        combine and present pass reads the input in place); with them,
        drawMeshes() draws it after the clears."""
        if self.mesh_passes:
            return
        src = self.inputs["hdrScene"]
        self.native.call("fh_copy", N.ptr(self.hdrRenderTexture), N.ptr(src), src.numel() * src.element_size())

    def drawMeshes(self, ctx: PassCtx) -> None:
        """renderer.ts:189 and :244-246 in the reference's order: the
        background, the shadow map (meshes, then hair) and the meshes' depth,
        normals and colour, which reads the shadow map."""
        self.backgroundPass.cmdDraw(ctx)
        self.shadowMapPass.cmdUpdateShadowMap(ctx, self.hair, meshes=self.meshes)
        self.drawMeshesPass.cmdDrawMeshes(ctx, self.meshes, self.meshDepthTexture, self.meshNormalsTexture)

    def drawHair(self, ctx: PassCtx) -> None:
        """cmdDrawScene's hair part (renderer.ts:262-295) and the present pass
        (:205).

        The passes run in the reference's order except for where they overlap:
        the shadow map, the hair depth / normals and AO (updateResourcesForNextFrame,
        renderer.ts:285-295) read the positions and the meshes' inputs and write
        textures that only the shading pass and the present pass's debug views
        read, so they run on the branch beside the tile, sort and fine passes
        (which read the meshes' depth, not the hair's). The shading pass writes
        the shading points the fine pass reads, so it starts after the fine
        pass, beside the combine and present passes; a shading pass still
        running when the fine pass starts would keep the fine pass's
        persistent workgroups from the CUs."""
        # the three passes' clears touch disjoint buffers: one launch for all
        # of them
        self.native.clear_batch_begin()
        try:
            self.tilesPass.cmdClearBeforeRender(ctx)
            self.finePass.cmdClearBeforeRender(ctx)
            self.sortPass.cmdClearBeforeRender(ctx)
        finally:
            self.native.clear_batch_end()
        if self.mesh_passes:
            # on the stream: the tile pass reads the meshes' depth and the
            # combine pass their colour. Splitting off the colour (which waits
            # for the shadow map) onto the branch measured slower: the branch
            # already holds the simulation step, the hair draw and AO.
            self.drawMeshes(ctx)
        if self.hw_passes:
            self._branch_begin()
            if self.mesh_passes:
                self.hwHairPass.cmdDrawHair(ctx, self.hair, self.meshDepthTexture, self.meshNormalsTexture)
                self.aoPass.cmdCalcAo(ctx)
            elif self.hw_passes:
                self.shadowMapPass.cmdUpdateShadowMap(ctx, self.hair, self.inputs["shadowMap"])
                self.hwHairPass.cmdDrawHair(ctx, self.hair, self.inputs["meshDepth"], self.inputs["meshNormals"])
                self.aoPass.cmdCalcAo(ctx)
            self._branch_end()
        self.tilesPass.cmdDrawHairToTiles(ctx, self.hair)
        self.sortPass.cmdSortHairTiles(ctx)
        self.finePass.cmdRasterizeSlicesHair(ctx, self.hair, self.tilesPass)
        if self.hw_passes:  # the present pass's debug views read the hair depth, normals and AO
            self.join_branch()
        self._branch_begin()
        self.shadingPass.cmdComputeShadingPoints(ctx, self.hair)
        self._branch_end()
        # the combine and present passes in one (the scene blended onto read
        # where it lies: the input, or the meshes' colour in the HDR texture)
        self.hairCombinePass.cmdCombineAndPresent(ctx, self.presentPass, self.screenTexture,
                                                  None if self.mesh_passes else self.inputs["hdrScene"])
        self.join_branch()

    def frame(self, upload_uniforms: bool = True) -> None:
        """One sim step + render of the hair chain (Renderer.cmdRender)."""
        if upload_uniforms:
            self.upload_sim_uniforms()
        ctx = self.ctx()
        self.simulateHair(ctx)
        self.drawHair(ctx)
        self.frameIdx += 1

    # Pipelined frames. The simulation step of frame f writes only the older
    # buffer of the tick-tock position pair (the new "now"), and the render
    # passes of frame f-1 read only the newer one (frame f-1's "now"); the
    # grids and the shading points are read by one chain and written by the
    # other in an order that a frame already fixes. So one pipelined frame
    # runs the simulation step of frame f on the branch beside the render
    # passes of frame f-1, and each image is presented one frame after its
    # simulation step. Every pass computes exactly what it computes in
    # frame(): images and positions are the same, frame for frame.

    def pipeline_begin(self, upload_uniforms: bool = True) -> None:
        """The prologue: the simulation step of the first frame alone."""
        if upload_uniforms:
            self.upload_sim_uniforms()
        ctx = self.ctx()
        self.simulateHair(ctx)
        self.join_branch()
        self._pipelined = True
        self.frameIdx += 1

    def _pipelined_calls(self) -> None:
        ctx = self.ctx()
        drawn = self.hair.current  # frame f-1's positions
        self._branch_begin()
        self._next_uniforms()
        self.gridPreSimPass.cmdUpdateGridsBeforeSim(ctx, self.hair)
        self.simPass.cmdSimulateHairPositions(ctx, self.hair)  # swaps the pair
        self.gridPostSimPass.cmdUpdateGridsAfterSim(ctx, self.hair)
        self._branch_end()
        stepped = self.hair.current
        self.hair.current = drawn
        try:
            self.drawHair(ctx)
        finally:
            self.hair.current = stepped

    def frame_pipelined(self, upload_uniforms: bool = True) -> None:
        """The simulation step of frame frameIdx beside the render passes of
        frame frameIdx - 1 (after pipeline_begin)."""
        assert getattr(self, "_pipelined", False), "pipeline_begin() first"
        if upload_uniforms:
            self.upload_sim_uniforms()
        self._pipelined_calls()
        self.frameIdx += 1

    def pipeline_end(self) -> None:
        """The epilogue: the render passes of the last simulated frame."""
        assert getattr(self, "_pipelined", False), "pipeline_begin() first"
        self.drawHair(self.ctx())
        self._pipelined = False

    def capture_pipelined(self) -> list:
        """capture_frames() for pipelined frames (two graphs, one per parity
        of the position pair); replayed by frame_pipelined_graph()."""
        return self._capture(self._pipelined_calls, "_pgraphs", "_pgraphBase")

    def frame_pipelined_graph(self, upload_uniforms: bool = True) -> None:
        assert getattr(self, "_pipelined", False), "pipeline_begin() first"
        if upload_uniforms:
            self.upload_sim_uniforms()
        k = self.hair.current ^ self._pgraphBase
        self._pgraphs[k].launch()
        self.hair.swapPositionBuffersAfterSimIntegration()
        self.frameIdx += 1

    # Deep-pipelined frames: three frames in flight, one stage each. Frame f's
    # simulation step runs on branch lane 0, the tile and sort passes of frame
    # f-1 on lane 1, and the fine, combine-and-present and shading passes of
    # frame f-2 on the context's stream, all at once. The stages share no
    # buffer a frame writes: the tile passes of consecutive frames alternate
    # between two sets of tile buffers (the tile set of frame g is g mod 2),
    # and the fine and shading passes of frame g read a copy of frame g's
    # positions that its tile stage took (the simulation step of frame g+2
    # overwrites the older buffer of the tick-tock pair while they run). The
    # shading points are written by the fine stage after its fine pass, which
    # read the previous fine stage's. Each pass computes exactly what it
    # computes in frame(): after deep frame f the simulation state is frame
    # f's, the tile set (f-1) mod 2 holds frame f-1's tile lists, and the
    # image, screen and shading points are frame f-2's.

    def _deep_alloc(self) -> None:
        if getattr(self, "_deepSets", None) is not None:
            return
        assert not self.hw_passes, "deep-pipelined frames take the hair's depth, shadow map and AO as inputs"
        rc, W, H = self.wl.render, self.wl.width, self.wl.height
        tiles2 = HairTilesPass()
        tiles2.onViewportResize(self.cfg, rc, W, H, self.device)
        sort2 = HairTileSortPass(rc, self.device)
        sort2.onViewportResize(rc, W, H, self.device)
        self._deepSets = [(self.tilesPass, self.sortPass), (tiles2, sort2)]
        self._deepPos = [torch.empty_like(self.hair.positionsNow) for _ in range(2)]

    def deep_tile_set(self, frame: int) -> tuple:
        """(tiles pass, sort pass) holding frame `frame`'s tile lists."""
        return self._deepSets[frame % 2]

    def _deep_ctx(self, frame: int) -> PassCtx:
        ctx = self.ctx()
        tp, sp = self.deep_tile_set(frame)
        ctx.hairTilesBuffer = tp.hairTilesBuffer
        ctx.hairTileSegmentsBuffer = tp.hairTileSegmentsBuffer
        ctx.hairSegmentCountPerTileBuffer = tp.segmentCountPerTileBuffer
        ctx.hairTileListBuffer = sp.tileListBuffer
        return ctx

    def _deep_calls(self, f: int, sim: bool = True, tiles: bool = True, fine: bool = True) -> None:
        """Simulation step of frame f, tile stage of frame f-1, fine stage of frame f-2."""
        drawn = self.hair.current  # frame f-1's positions
        if sim:
            self.native.branch_begin(0)
            try:
                self.simulateStep(self.ctx())  # swaps the pair
            finally:
                self.native.branch_end()
        if tiles:
            g = f - 1
            tp, sp = self.deep_tile_set(g)
            ctx = self._deep_ctx(g)
            view = _PositionsView(self.hair, self.hair.positions[drawn])
            self.native.branch_begin(1)
            try:
                self.native.clear_batch_begin()
                try:
                    tp.cmdClearBeforeRender(ctx)
                    sp.cmdClearBeforeRender(ctx)
                finally:
                    self.native.clear_batch_end()
                tp.cmdDrawHairToTiles(ctx, view)
                sp.cmdSortHairTiles(ctx)
                pos = self._deepPos[g % 2]
                self.native.call("fh_copy", N.ptr(pos), N.ptr(view.positionsNow), pos.numel() * pos.element_size())
            finally:
                self.native.branch_end()
        if fine:
            g = f - 2
            tp, _ = self.deep_tile_set(g)
            ctx = self._deep_ctx(g)
            view = _PositionsView(self.hair, self._deepPos[g % 2])
            self.finePass.cmdClearBeforeRender(ctx)
            self.finePass.cmdRasterizeSlicesHair(ctx, view, tp)
            self.hairCombinePass.cmdCombineAndPresent(ctx, self.presentPass, self.screenTexture,
                                                      self.inputs["hdrScene"])
            self.shadingPass.cmdComputeShadingPoints(ctx, view)
        self.native.branch_join()

    def deep_pipeline_begin(self, upload_uniforms: bool = True) -> None:
        """The prologue: frame 0's simulation step, then frame 1's beside
        frame 0's tile stage."""
        self._deep_alloc()
        for stages in ((True, False, False), (True, True, False)):
            if upload_uniforms:
                self.upload_sim_uniforms()
            self._deep_calls(self.frameIdx, *stages)
            self.frameIdx += 1
        self._deepOn = True

    def frame_deep(self, upload_uniforms: bool = True) -> None:
        """Simulation step of frame frameIdx beside the tile stage of frame
        frameIdx - 1 and the fine stage of frame frameIdx - 2."""
        assert getattr(self, "_deepOn", False), "deep_pipeline_begin() first"
        if upload_uniforms:
            self.upload_sim_uniforms()
        self._deep_calls(self.frameIdx)
        self.frameIdx += 1

    def deep_pipeline_end(self) -> None:
        """The epilogue: the last two frames' remaining stages."""
        assert getattr(self, "_deepOn", False), "deep_pipeline_begin() first"
        f = self.frameIdx
        self._deep_calls(f, sim=False)
        self._deep_calls(f + 1, sim=False, tiles=False)
        self._deepOn = False

    def capture_deep(self) -> list:
        """Two HIP graphs of deep-pipelined frames (one per frame parity: the
        position pair and the tile sets alternate together); replayed by
        frame_deep_graph()."""
        state = {"f": self.frameIdx}

        def calls():
            self._deep_calls(state["f"])
            state["f"] += 1
        return self._capture(calls, "_dgraphs", "_dgraphBase")

    def frame_deep_graph(self, upload_uniforms: bool = True) -> None:
        assert getattr(self, "_deepOn", False), "deep_pipeline_begin() first"
        if upload_uniforms:
            self.upload_sim_uniforms()
        self._dgraphs[self.hair.current ^ self._dgraphBase].launch()
        self.hair.swapPositionBuffersAfterSimIntegration()
        self.frameIdx += 1

    def capture_frames(self) -> list:
        """Records the frame's calls as two HIP graphs, one per parity of the
        tick-tock position buffers (hairObject.ts:102-106), like recording
        the frame's GPUCommandEncoder once. Every call must have run once
        before (so no scratch grows while recording); the graphs read the
        simulation uniforms from `simulationUniforms`, whose contents the
        caller updates per frame, or which each simulation step fills from
        the device ring of use_uniform_ring()."""
        def calls():
            ctx = self.ctx()
            self.simulateHair(ctx)
            self.drawHair(ctx)
        return self._capture(calls, "_graphs", "_graphBase")

    def _capture(self, calls, attr: str, base_attr: str) -> list:
        graphs = []
        # recording needs a stream of its own (not the legacy default stream)
        main = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(main)
        self.native.set_stream(side)
        try:
            for _ in range(2):
                self.native.capture_begin()
                try:
                    calls()
                finally:
                    g = self.native.capture_end()
                graphs.append(g)
        finally:
            self.native.set_stream(main)
        setattr(self, attr, graphs)
        setattr(self, base_attr, self.hair.current)  # graphs[0] reads positions[current] as "now"
        return graphs

    def frame_graph(self, upload_uniforms: bool = True) -> None:
        """One frame replayed from capture_frames() (the graph of the current
        parity of the position buffers)."""
        if upload_uniforms:
            self.upload_sim_uniforms()
        self._graphs[self.hair.current ^ self._graphBase].launch()
        self.hair.swapPositionBuffersAfterSimIntegration()
        self.frameIdx += 1


def workload_render_uniforms(wl: A.Workload, model: np.ndarray | None = None) -> np.ndarray:
    """The render uniforms of a workload (renderUniformsBuffer.ts:205-288):
    its camera, the hair's bounding sphere, one light at the origin, the SDF
    and grid bounds."""
    box = A.derived(wl)["box"]
    return L.pack_render_uniforms(wl.render, wl.sim, wl.camera, wl.width, wl.height,
                                  np.eye(4) if model is None else model, A.bounding_sphere(box),
                                  [(np.zeros(3), 0.3)], wl.sdf_bounds, A.scale_box(box, wl.sim.gridScale))


def make_inputs(wl: A.Workload, render_uniforms: np.ndarray, shadow_size: int = 2048) -> dict:
    """The inputs the out-of-scope passes produce: the meshes' depth and
    normals, a precomputed depth with the hair splatted in, a flat AO, the
    meshes' shadow map and the HDR scene (background and meshes)."""
    md, mn = A.mesh_depth_normals(wl, render_uniforms)
    hd = A.hair_depth(wl, render_uniforms, wl.positions, md)
    ao = np.ones(((wl.height + 1) // 2, (wl.width + 1) // 2), dtype=np.float32)
    sm = A.shadow_map(wl, render_uniforms, shadow_size)
    hdr = A.hdr_scene(wl, md)
    return {"meshDepth": md, "meshNormals": mn.view(np.int32), "hairDepth": hd, "ao": ao, "shadowMap": sm,
            "hdrScene": hdr}
