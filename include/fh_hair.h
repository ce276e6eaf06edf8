/*
 * fh_hair.h — C-ABI of the MI355X (gfx950) per-frame hair compute chain.
 *
 * One entry point per method of the reference's pass classes under src/passes.
 * Every buffer argument is a plain device pointer whose byte layout is the
 * layout of the matching WGSL binding in the reference, so a pass can replace
 * its WGSL counterpart without touching the producers or consumers around it.
 * Citations are file:line in Scthe/frostbitten-hair-webgpu.
 *
 * Conventions
 *  - Every entry point returns FH_OK (0) or a non-zero FH_ERR_* code and never
 *    throws; fh_last_error() gives the message of the last failure on the
 *    calling thread. This replaces the reference's WebGPU error scopes
 *    (src/utils/errors.ts:3-41, index.deno.ts:157-163).
 *  - Work is enqueued on the context's HIP stream and the call returns at once,
 *    like recording into a GPUCommandEncoder; fh_sync() waits for it.
 *  - Values that the reference bakes into WGSL as compile-time constants
 *    (TILE_SIZE, TILE_DEPTH_BINS_COUNT, SLICES_PER_PIXEL, GRID_DIMS,
 *    SHADING_POINTS, ...) are run-time arguments here (fh_render_cfg,
 *    grid_dims), so instances with different settings can share one library.
 *  - Textures become dense f32 arrays in texture memory order (texel (x, y) at
 *    y * width + x, row 0 is the first row written by writeTexture). Loads
 *    outside the texture return 0 (robust image access).
 */
#ifndef FH_HAIR_H
#define FH_HAIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_ABI_VERSION 1

enum {
  FH_OK = 0,
  FH_ERR_INVALID_ARG = 1,
  FH_ERR_HIP = 2,
  FH_ERR_UNSUPPORTED = 3,
  FH_ERR_CAPACITY = 4,
};

typedef struct fh_ctx fh_ctx;

typedef struct fh_vec4 {
  float x, y, z, w;
} fh_vec4;

/* ------------------------------------------------------------------------ */
/* Uniform blocks (host-packed bytes, uploaded to a device buffer)           */
/* ------------------------------------------------------------------------ */

/* struct SimulationUniforms — simulationUniformsBuffer.ts:19-51, size :56-63,
 * packed by update() :86-152. Matrices are column-major mat4x4<f32>. */
typedef struct fh_sim_uniforms {
  float modelMatrix[16];          /*   0 */
  float modelMatrixInv[16];       /*  64 */
  float collisionSphere[4];       /* 128  object space xyz, radius */
  float wind[4];                  /* 144  direction.xyz (object space), strength */
  float sdfBoundsMin[4];          /* 160  sdfCollider.ts:174-177; .w debug slice */
  float sdfBoundsMax[4];          /* 176  .w = SDF distance offset */
  float gridBoundsMin[4];         /* 192  gridData.ts:264-268; .w debug slice */
  float gridBoundsMax[3];         /* 208 */
  uint32_t gridDebugDisplayValue; /* 220 */
  float deltaTime;                /* 224 */
  float gravity;
  uint32_t constraintIterations;
  uint32_t frameIdx;
  float windStrengthLull;         /* 240 */
  float windColisionTraceOffset;
  float stiffnessLengthConstr;
  float stiffnessCollisions;
  float stiffnessSDF;             /* 256 */
  float volumePreservation;
  float friction;
  float stiffnessGlobalConstr;
  float globalConstrExtent;       /* 272 */
  float globalConstrFade;
  float stiffnessLocalConstr;
  float windPhaseOffset;
  float windStrengthFrequency;    /* 288 */
  float windStrengthJitter;
  float padding0;
  float padding1;
} fh_sim_uniforms;                /* 304 */

/* struct Light — renderUniformsBuffer.ts:46-49 */
typedef struct fh_light {
  float position[4];
  float colorAndEnergy[4];
} fh_light;

/* struct Shadows — renderUniformsBuffer.ts:52-61 */
typedef struct fh_shadows {
  float sourceModelViewMat[16];
  float sourceProjMatrix[16];
  float sourceMVP_Matrix[16];
  float sourcePosition[4];        /* xyz, .w = shadow fiber radius */
  uint32_t usePCSS;
  uint32_t PCF_Radius;
  float bias;
  float strength;
} fh_shadows;

/* struct AmbientOcclusion — renderUniformsBuffer.ts:64-74 */
typedef struct fh_ao {
  float radius, directionOffset, falloffStart2, falloffEnd2;
  uint32_t numDirections, numSteps;
  float strength, padding0;
} fh_ao;

/* struct HairMaterialParams — renderUniformsBuffer.ts:76-93 */
typedef struct fh_hair_material {
  float color0[3];
  float specular;
  float color1[3];
  float shadows;
  float weightTT, weightTRT, shift, roughness;
  float colorRng, lumaRng, attenuation, padding1;
} fh_hair_material;

/* struct Uniforms — renderUniformsBuffer.ts:105-137, size :163-185,
 * packed by update() :205-288. */
typedef struct fh_render_uniforms {
  float vpMatrix[16];             /*    0 */
  float vpMatrixInv[16];          /*   64 */
  float viewMatrix[16];           /*  128 */
  float projMatrix[16];           /*  192 */
  float projMatrixInv[16];        /*  256 */
  float modelMatrix[16];          /*  320 */
  float modelViewMat[16];         /*  384 */
  float mvpMatrix[16];            /*  448 */
  float collisionSphereModelMatrix[16]; /* 512 */
  float viewport[4];              /*  576 */
  float cameraPosition[4];        /*  592 */
  float colorMgmt[4];             /*  608 */
  float lightAmbient[4];          /*  624 */
  fh_light light0;                /*  640 */
  fh_light light1;                /*  672 */
  fh_light light2;                /*  704 */
  fh_shadows shadows;             /*  736 */
  fh_ao ao;                       /*  960 */
  fh_hair_material hairMaterial;  /*  992 */
  float sdfBoundsMin[4];          /* 1056 */
  float sdfBoundsMax[4];          /* 1072 */
  float gridBoundsMin[4];         /* 1088 */
  float gridBoundsMax[3];         /* 1104 */
  uint32_t gridDebugDisplayValue; /* 1116 */
  float fiberRadius;              /* 1120 */
  float dbgShadowMapPreviewSize;
  uint32_t maxDrawnHairSegments;  /* 1128 */
  uint32_t displayMode;
  float backgroundColor0[3];      /* 1136 */
  float backgroundNoiseScale;
  float backgroundColor1[3];
  float backgroundGradientStrength;
  float collisionSpherePosition[3]; /* 1168 */
  uint32_t gizmoActiveState;
} fh_render_uniforms;             /* 1184 */

/* ------------------------------------------------------------------------ */
/* Storage-buffer element layouts                                            */
/* ------------------------------------------------------------------------ */

/* struct HairData — hairDataBuffer.ts:15-19 (buffer is 256 B) */
typedef struct fh_hair_data {
  float boundingSphere[4];
  uint32_t strandsCount;
  uint32_t pointsPerStrand;
} fh_hair_data;

/* struct DensityVelocityI32 — densityVelocityGrid.wgsl.ts:8-16; values are
 * i32 fixed point x1e6 (grids.wgsl.ts:3, 10-12). Index z*d*d + y*d + x. */
typedef struct fh_density_velocity {
  int32_t velocityX, velocityY, velocityZ, density;
} fh_density_velocity;

/* struct DensityGradAndWind — densityGradAndWindGrid.wgsl.ts:8-14 */
typedef struct fh_density_grad_wind {
  float densityGrad[3];
  float windStrength;
} fh_density_grad_wind;

/* struct HairTileResult — hairTilesResultBuffer.ts:77-82, T*B entries indexed
 * by getHairTileDepthBinIdx (tileUtils.wgsl.ts:483-490). padding0 carries the
 * number of list entries of this (tile, bin) when written by fh_tiles_draw. */
typedef struct fh_tile_result {
  uint32_t minDepth;      /* u32((1 - dmin) * 2^32), atomicMax */
  uint32_t maxDepth;      /* u32(dmax * 2^32), atomicMax */
  uint32_t tileSegmentPtr;/* head + 1, 0 = empty list */
  uint32_t padding0;
} fh_tile_result;

/* struct LinkedListElement — hairTileSegmentsBuffer.ts:60-63. The segments
 * buffer is struct DrawnHairSegments: u32 drawnSegmentsCount at byte 0 and
 * the element array at byte 4 (:65-68). */
typedef struct fh_list_element {
  uint32_t strandAndSegmentIdxs; /* (segmentIdx << 24) | strandIdx */
  uint32_t next;                 /* 0xffffffff = end of list */
} fh_list_element;

/* struct SortBucket — hairTileSortPass.countTiles.wgsl.ts:21-45 */
typedef struct fh_sort_bucket {
  uint32_t tileCount;
  uint32_t writeOffset;
} fh_sort_bucket;

/* Tile list (tileListBuffer.ts:15-22): u32 drawnTiles at byte 0, u32 tile ids
 * at byte 4; (4 + T) u32 allocated.
 * Rasterizer result (hairRasterizerResultBuffer.ts:11-34): u32
 * tileQueueAtomicIdx at byte 0, vec4f pixels at byte 16 indexed W*y + x with
 * y up; allocated W*H*16 B so the array holds W*H - 1 pixels.
 * Slice data (hairSlicesDataBuffer.ts:34-37): vec4u {rg f16x2, ba f16x2,
 * next, 0}, processorCount * avgFragmentsPerSlice * slices * tile^2 entries. */

/* Values the reference bakes into its WGSL from CONFIG.hairRender
 * (constants.ts:193-265). */
typedef struct fh_render_cfg {
  uint32_t tileSize;                        /* 8, 16 or 32 */
  uint32_t tileDepthBins;                   /* 1..32 */
  uint32_t slicesPerPixel;                  /* 1..16 */
  uint32_t avgFragmentsPerSlice;
  uint32_t processorCount;
  uint32_t invalidTilesPerSegmentThreshold;
  uint32_t sortBuckets;                     /* 1..64 */
  uint32_t sortBucketSize;
  uint32_t shadingPoints;                   /* 2..64 */
  uint32_t alphaQuadratic;                  /* 0 or 1 */
  float alphaMultiplier;
  uint32_t reserved;
} fh_render_cfg;

/* A 2D f32 texture (depth24plus read as f32, r16float AO widened to f32). */
typedef struct fh_tex2d {
  const float* data;
  uint32_t width;
  uint32_t height;
} fh_tex2d;

/* ------------------------------------------------------------------------ */
/* Context, memory and sizes                                                 */
/* ------------------------------------------------------------------------ */

/* Replaces createGpuDevice (src/utils/webgpu.ts:8-68): binds device `device`
 * and creates the stream all work of this context is enqueued on. */
int fh_create(int device, fh_ctx** out);
int fh_destroy(fh_ctx* ctx);
/* Use an existing hipStream_t (e.g. torch's current stream) instead. */
int fh_set_stream(fh_ctx* ctx, void* hip_stream);
void* fh_get_stream(fh_ctx* ctx);
/* Replaces queue.onSubmittedWorkDone (index.deno.ts:108-120). */
int fh_sync(fh_ctx* ctx);
const char* fh_last_error(void);
int fh_abi_version(void);

/* Device memory for hosts without their own allocator (a Deno host). */
int fh_alloc(fh_ctx* ctx, uint64_t bytes, void** out);
int fh_free(fh_ctx* ctx, void* ptr);
int fh_upload(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int fh_download(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* GPUCommandEncoder.clearBuffer (webgpu.ts:141-143); a fill kernel, so it is
 * safe inside a captured graph. bytes must be a multiple of 4. */
int fh_clear(fh_ctx* ctx, void* dst, uint64_t bytes);

/* Buffer sizes the reference computes on the host. */
/* getTileCount (shared/utils.ts:4-10) */
uint32_t fh_tile_count(const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h);
/* getLengthOfHairTileSegmentsBuffer (hairTileSegmentsBuffer.ts:81-87):
 * 1 + ceil(tiles * avgSegmentsPerTile); also maxDrawnHairSegments. */
uint32_t fh_segments_buffer_entries(const fh_render_cfg* cfg, uint32_t vp_w,
                                    uint32_t vp_h, float avg_segments_per_tile);
/* SLICE_DATA_PER_PROCESSOR_COUNT (hairSlicesDataBuffer.ts:16-20) */
uint32_t fh_slice_data_per_processor(const fh_render_cfg* cfg);

/* ------------------------------------------------------------------------ */
/* Simulation step (src/passes/simulation)                                   */
/* ------------------------------------------------------------------------ */

/* GridPreSimPass.cmdUpdateGridsBeforeSim (gridPreSimPass.ts:34-62; kernel
 * gridPreSimPass.wgsl.ts:38-159): wind strength and density gradient per grid
 * point. Every grid point is written, which covers the clear at :37. */
int fh_grid_pre_sim(fh_ctx* ctx, uint32_t grid_dims,
                    const fh_sim_uniforms* sim_uniforms,
                    const fh_density_velocity* density_velocity,
                    fh_density_grad_wind* density_grad_wind,
                    const float* sdf_texture, uint32_t sdf_dims);

/* HairSimIntegrationPass.cmdSimulateHairPositions (hairSimIntegrationPass.ts
 * :33-60; kernel hairSimIntegrationPass.wgsl.ts:79-273): Verlet with grid
 * friction and forces, constraintIterations x [length, global shape, local
 * shape, tip tangent, sphere + SDF collisions], result written to
 * positions_prev (the caller then swaps the tick-tock pair, :59). */
int fh_sim_integrate(fh_ctx* ctx, uint32_t grid_dims,
                     const fh_sim_uniforms* sim_uniforms,
                     const fh_hair_data* hair_data,
                     fh_vec4* positions_prev, const fh_vec4* positions_now,
                     const float* segment_lengths,
                     const float* sdf_texture, uint32_t sdf_dims,
                     const fh_density_velocity* density_velocity,
                     const fh_density_grad_wind* density_grad_wind,
                     const fh_vec4* positions_initial,
                     uint32_t strands_count, uint32_t points_per_strand);

/* GridPostSimPass.cmdUpdateGridsAfterSim (gridPostSimPass.ts:33-59; kernel
 * gridPostSimPass.wgsl.ts:45-72): clears density_velocity (:36), then scatters
 * velocity (now - prev) and density of every point into the grid with i32
 * fixed-point atomics (densityVelocityGrid.wgsl.ts:93-124). */
int fh_grid_post_sim(fh_ctx* ctx, uint32_t grid_dims,
                     const fh_sim_uniforms* sim_uniforms,
                     const fh_hair_data* hair_data,
                     const fh_vec4* positions_prev, const fh_vec4* positions_now,
                     fh_density_velocity* density_velocity,
                     uint32_t strands_count, uint32_t points_per_strand);

/* No counterpart in the reference: fh_grid_post_sim for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results; its
 * temporaries live in a fixed 16 MiB device array of the library, one per
 * GPU and held until the process ends, instead of the context's
 * grow-on-demand scratch, and it calls no allocating routine. Inputs needing
 * more (about 500,000 points; 4 * (4 * ceil(d / 8)^3 + 1 + 8 * points) + 256
 * bytes) are refused with FH_ERR_CAPACITY before anything is enqueued. Every
 * call on one GPU shares the array: calls that overlap on one GPU (on two
 * streams or two contexts) race. */
int fh_grid_post_sim_noalloc(fh_ctx* ctx, uint32_t grid_dims,
                             const fh_sim_uniforms* sim_uniforms,
                             const fh_hair_data* hair_data,
                             const fh_vec4* positions_prev, const fh_vec4* positions_now,
                             fh_density_velocity* density_velocity,
                             uint32_t strands_count, uint32_t points_per_strand);

/* ------------------------------------------------------------------------ */
/* Software rasterizer (src/passes/swHair)                                   */
/* ------------------------------------------------------------------------ */

/* HairTilesPass.cmdClearBeforeRender (hairTilesPass.ts:48-52). */
int fh_tiles_clear(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w,
                   uint32_t vp_h, fh_tile_result* tiles,
                   void* tile_segments, uint32_t* segment_count_per_tile);

/* HairTilesPass.cmdDrawHairToTiles (hairTilesPass.ts:78-114; kernel
 * hairTilesPass.wgsl.ts:70-246). tile_segments is DrawnHairSegments holding
 * segments_capacity elements. slot_tile_bin (may be NULL; 2 u32 per element)
 * receives, for each list slot k, slot_tile_bin[2k] = the (tile, bin) index
 * the element was linked into and slot_tile_bin[2k + 1] = its rank in that
 * list, in [0, HairTileResult.padding0). The fine pass uses both to gather
 * each list into a contiguous range without walking it or using atomics. */
int fh_tiles_draw(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w,
                  uint32_t vp_h, const fh_render_uniforms* uniforms,
                  const fh_hair_data* hair_data, const fh_vec4* positions,
                  const fh_vec4* tangents, fh_tile_result* tiles,
                  fh_tex2d depth, void* tile_segments,
                  uint32_t segments_capacity,
                  uint32_t* segment_count_per_tile, uint32_t* slot_tile_bin,
                  uint32_t rendered_strands, uint32_t points_per_strand);

/* For lists produced by another implementation of the tile pass: walks every
 * (tile, bin) list and fills slot_tile_bin (tag and rank, as fh_tiles_draw
 * writes them) and the per-bin counts (padding0). */
int fh_tiles_rebuild_slot_keys(fh_ctx* ctx, const fh_render_cfg* cfg,
                               uint32_t vp_w, uint32_t vp_h,
                               fh_tile_result* tiles, const void* tile_segments,
                               uint32_t segments_capacity,
                               uint32_t* slot_tile_bin);

/* HairTileSortPass.cmdClearBeforeRender (hairTileSortPass.ts:63-68). */
int fh_tiles_sort_clear(fh_ctx* ctx, const fh_render_cfg* cfg, void* tile_list,
                        fh_sort_bucket* buckets);

/* HairTileSortPass.cmdSortHairTiles (hairTileSortPass.ts:80-153; kernels
 * hairTileSortPass.countTiles.wgsl.ts:63-97 and .sort.wgsl.ts:37-75). */
int fh_tiles_sort(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w,
                  uint32_t vp_h, const uint32_t* segment_count_per_tile,
                  void* tile_list, fh_sort_bucket* buckets);

/* HairFinePass.cmdClearBeforeRender (hairFinePass.ts:50-55): clears the
 * rasterizer result, including the tile-queue counter. */
int fh_fine_clear(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h,
                  void* rasterizer_result);

/* HairFinePass.cmdRasterizeSlicesHair (hairFinePass.ts:70-90; kernel
 * hairFinePass.wgsl.ts:126-335): cfg->processorCount persistent workgroups
 * pull tiles from the tile list and blend each depth bin's slices front to
 * back into rasterizer_result. slice_data is the per-processor fragment pool
 * (used here only for fragments that do not fit in LDS). */
int fh_fine_raster(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w,
                   uint32_t vp_h, const fh_render_uniforms* uniforms,
                   const fh_hair_data* hair_data, const fh_vec4* positions,
                   const fh_tile_result* tiles, const void* tile_segments,
                   uint32_t segments_capacity, void* slice_data,
                   void* rasterizer_result, fh_tex2d depth,
                   const fh_vec4* shading, const fh_vec4* tangents,
                   const void* tile_list, const uint32_t* slot_tile_bin,
                   uint32_t strands_count, uint32_t points_per_strand);

/* ------------------------------------------------------------------------ */
/* Strand-space shading (src/passes/hairShadingPass)                         */
/* ------------------------------------------------------------------------ */

/* HairShadingPass.cmdComputeShadingPoints (hairShadingPass.ts:41-68; kernel
 * hairShadingPass.wgsl.ts:61-282): cfg->shadingPoints colours per strand.
 * shadow_map is the shadow source depth (square), ao the GTAO result. */
int fh_shading(fh_ctx* ctx, const fh_render_cfg* cfg,
               const fh_render_uniforms* uniforms,
               const fh_hair_data* hair_data, const fh_vec4* positions,
               const fh_vec4* tangents, fh_vec4* shading,
               fh_tex2d shadow_map, fh_tex2d ao, fh_tex2d depth,
               uint32_t strands_count, uint32_t points_per_strand);

/* ------------------------------------------------------------------------ */
/* Combine and present (src/passes/hairCombine, src/passes/presentPass)      */
/* ------------------------------------------------------------------------ */

/* HDR scene texture: rgba16float (HDR_RENDER_TEX_FORMAT, constants.ts:35),
 * 8 B per texel in texture memory order (texel (x, y) at y * width + x, row 0
 * at the top). Screen texture: rgba8unorm-srgb (index.deno.ts:42), 4 B per
 * texel, rows of screen_row_bytes bytes (0 = 4 * width), so that it can be
 * written straight into a row-padded readback layout (index.deno.ts:131-146). */

/* HairCombinePass.cmdCombineRasterResults (hairCombinePass.ts:228-249;
 * fragment shader hairCombinePass.wgsl.ts:51-162, blend state :204-217):
 * blends the rasterizer result over hdr_texture with src-alpha /
 * one-minus-src-alpha colour and one / one alpha, rasterizer row H - 1 - y
 * onto texture row y. Display modes TILES / TILES_PPLL draw per-tile segment
 * counts instead, as the reference's debug views do. */
int fh_hair_combine(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                    const fh_render_uniforms* uniforms, const fh_tile_result* tiles,
                    const void* tile_segments, const void* rasterizer_result,
                    const uint32_t* segment_count_per_tile, void* hdr_texture);

/* PresentPass.cmdDraw (presentPass.ts:182-210; fragment shader
 * presentPass.wgsl.ts:58-103): dither, exposure, ACES tone mapping and gamma
 * (colorMgmt) of hdr_texture into the sRGB screen texture. depth is the depth
 * texture (DEPTH view), ao the AO texture (AO view), normals_texture the
 * rg16float octahedral normals (NORMALS view; NULL reads as 0).
 * dither_per_pixel = 0 reads Bayer element 0 for every pixel, as the Deno
 * build does (dither.wgsl.ts:34); 1 indexes the matrix by pixel as in the
 * browser. */
int fh_present(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
               const void* hdr_texture, fh_tex2d depth, fh_tex2d ao, const void* normals_texture,
               uint32_t dither_per_pixel, void* screen_texture, uint32_t screen_row_bytes);

/* fh_hair_combine and then fh_present of the texture it wrote, in one pass
 * over the texels (each thread presents the texel it blended): leaves
 * hdr_texture and the screen texture exactly as those two calls do. The
 * reference records the two passes back to back in one command encoder
 * (hairCombinePass.ts:228-249, presentPass.ts:182-210); this entry point has
 * no single counterpart there. scene_texture (NULL: hdr_texture itself) holds
 * the HDR scene to blend onto, so that a caller whose background and meshes
 * are an input passes them here instead of copying them into hdr_texture
 * first; it must not overlap hdr_texture (FH_ERR_INVALID_ARG). */
int fh_combine_present(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                       const fh_render_uniforms* uniforms, const fh_tile_result* tiles, const void* tile_segments,
                       const void* rasterizer_result, const uint32_t* segment_count_per_tile,
                       const void* scene_texture, void* hdr_texture, fh_tex2d depth, fh_tex2d ao,
                       const void* normals_texture, uint32_t dither_per_pixel, void* screen_texture,
                       uint32_t screen_row_bytes);

/* ------------------------------------------------------------------------ */
/* Hair depth and normals, shadow map, AO (the reference's hardware-raster   */
/* passes that feed the shading pass)                                        */
/* ------------------------------------------------------------------------ */

/* The reference draws these with the hardware rasterizer; here they are a
 * software rasterizer with fixed rules: vertices snapped to 1/256 px,
 * top-left fill rule at pixel centres, screen-space interpolated NDC depth
 * clipped to [0, 1], the pipelines' 'less' test (shared.ts:22-26) in
 * primitive order, triangles with a vertex at w <= 0 not drawn. The result
 * does not depend on the order the GPU runs the triangles in. Uses the
 * context's scratch (8 B per target pixel). */

/* HwHairPass.cmdDrawHair (hwHairPass.ts:79-141; vertex shader
 * hwRasterizeHair.wgsl.ts:23-48, fragment shader hwHairPass.wgsl.ts:77-99):
 * the strands' two triangles per segment (hairIndexBuffer.ts:20-54) drawn
 * into depth_texture (in: the meshes' depth; out: with the hair) and, where
 * the hair wins, normals_texture (rg16float octahedral world-space tangent,
 * W x H u32; may be NULL). The colour target is not written (display mode
 * FINAL writes alpha 0 there). */
int fh_hw_hair(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
               const fh_vec4* positions, const fh_vec4* tangents, float* depth_texture,
               void* normals_texture, uint32_t rendered_strands, uint32_t points_per_strand);

/* The hair draw of ShadowMapPass.cmdUpdateShadowMap (shadowMapPass.ts
 * :100-115; shadowMapHairPass.wgsl.ts): the same triangles from the shadow
 * source (shadows.sourceModelViewMat / sourceProjMatrix, fiber radius in
 * shadows.sourcePosition.w) into shadow_map (shadow_size^2 f32; in: the
 * meshes' shadow depth). */
int fh_shadow_map_hair(fh_ctx* ctx, const fh_render_uniforms* uniforms, const fh_vec4* positions,
                       const fh_vec4* tangents, float* shadow_map, uint32_t shadow_size,
                       uint32_t rendered_strands, uint32_t points_per_strand);

/* AoPass.cmdCalcAo (aoPass.ts:45-86; aoPass.wgsl.ts:48-79, gtao.wgsl.ts
 * :26-136): GTAO from the depth texture (with the hair) and the normals
 * texture into the AO texture (ao_w x ao_h = ceil(W / 2) x ceil(H / 2),
 * CONFIG.ao.textureSizeMul, renderer.ts:406-416), r16float values widened
 * to f32. */
int fh_calc_ao(fh_ctx* ctx, const fh_render_uniforms* uniforms, fh_tex2d depth, const void* normals_texture,
               float* ao_texture, uint32_t ao_w, uint32_t ao_h);

/* ------------------------------------------------------------------------ */
/* Background and meshes (src/passes/drawBackgroundGradient,                 */
/* src/passes/drawMeshes, the meshes' draw of src/passes/shadowMapPass)      */
/* ------------------------------------------------------------------------ */

/* One object of the scene (GPUMesh, scene/gpuMesh.ts:5-16, as loadObjFile
 * leaves it, scene/objLoader.ts:21-66): device buffers of vertex positions
 * and normals (float32x3, 12 B per vertex, VERTEX_ATTRIBUTES gpuMesh.ts
 * :18-55) and u32 triangle indices (3 per triangle). is_collider_preview
 * marks the collider preview sphere (loadScene.ts:36), drawn with the
 * collision sphere's model matrix and not into the shadow map. */
typedef struct fh_mesh {
  const float* positions;
  const float* normals;
  const uint32_t* indices;
  uint32_t vertex_count;
  uint32_t triangle_count;
  uint32_t is_collider_preview;
  uint32_t reserved;
} fh_mesh;

/* The scene's object list (Scene.objects, scene/scene.ts), in draw order.
 * Created once (the reference creates its GPUMesh objects at load,
 * loadScene.ts:25-60): checks that every index addresses a vertex of its
 * mesh (FH_ERR_INVALID_ARG otherwise) and keeps a device copy of the
 * descriptors; the buffers themselves stay the caller's and must outlive the
 * object. Not while the context records or captures. At most 256 meshes and
 * 2^31 - 1 triangles. */
typedef struct fh_meshes fh_meshes;
int fh_meshes_create(fh_ctx* ctx, const fh_mesh* meshes, uint32_t count, fh_meshes** out);
int fh_meshes_destroy(fh_meshes* meshes);
int fh_meshes_triangles(const fh_meshes* meshes, uint32_t* n_out);

/* DrawBackgroundGradientPass.cmdDraw (renderer.ts:189; fragment shader
 * drawBackgroundGradientPass.wgsl.ts:33-48, fractalNoise noise.wgsl.ts
 * :14-42): the background gradient with its noise into hdr_texture
 * (rgba16float, the layout of fh_hair_combine). */
int fh_draw_background(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                       void* hdr_texture);

/* DrawMeshesPass.cmdDrawMeshes (drawMeshesPass.ts:58-115; shaders
 * drawMeshesPass.wgsl.ts:46-168, pbr.wgsl.ts:101-160, sampleShadows.wgsl.ts
 * :47-150, textureAo.wgsl.ts:7-12): clears depth_texture (to 1.0) and
 * normals_texture (rg16float octahedral world normals, to 0), draws every
 * object with back faces culled and the 'less' test, and writes each drawn
 * pixel's PBR colour (three lights, shadow from shadow_map, AO from ao) into
 * hdr_texture over what is there (loadOp 'load': the background). Rules and
 * scratch as fh_hw_hair. */
int fh_draw_meshes(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                   const fh_meshes* meshes, fh_tex2d shadow_map, fh_tex2d ao, float* depth_texture,
                   void* normals_texture, void* hdr_texture);

/* fh_draw_meshes in two calls, for callers that run the colour beside other
 * work (the colour needs the shadow map; the depth, which the tile pass
 * reads, does not). fh_draw_meshes_visibility clears and draws depth_texture
 * and normals_texture as fh_draw_meshes does and writes, per pixel, the scene
 * triangle index of the drawn fragment (0xffffffff where none; triangles
 * numbered over the meshes in draw order) into visibility (W x H u32).
 * fh_shade_meshes then writes those fragments' colour into hdr_texture
 * (indices past the scene's triangles are skipped). Together they leave
 * exactly what fh_draw_meshes leaves. */
int fh_draw_meshes_visibility(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                              const fh_meshes* meshes, float* depth_texture, void* normals_texture,
                              uint32_t* visibility);
int fh_shade_meshes(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                    const fh_meshes* meshes, const uint32_t* visibility, fh_tex2d shadow_map, fh_tex2d ao,
                    void* hdr_texture);

/* The meshes' draw of ShadowMapPass.cmdUpdateShadowMap (shadowMapPass.ts
 * :83-99, :117-133; shader shadowMapPass.wgsl.ts): clears shadow_map
 * (shadow_size^2 f32) to 1.0 and draws every object but the collider preview
 * with the shadow source's MVP (shadows.sourceMVP_Matrix), no culling.
 * fh_shadow_map_hair then draws the hair into the same map. */
int fh_shadow_map_meshes(fh_ctx* ctx, const fh_render_uniforms* uniforms, const fh_meshes* meshes,
                         float* shadow_map, uint32_t shadow_size);

/* Device-to-device copy as a kernel (bytes a multiple of 4), so that a
 * captured frame can restore its inputs (e.g. the HDR scene drawn by the
 * out-of-scope mesh passes) without memcpy nodes. */
int fh_copy(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* Copies record (*counter mod count) of src (count records of `bytes` bytes,
 * a multiple of 4) to dst and adds one to *counter (a device u32), all on the
 * device. The per-frame uploads of the reference (queue.writeBuffer of the
 * simulation uniforms, simulationUniformsBuffer.ts:86-152) become one such
 * call at the start of a frame over a device ring of the frames' uniforms,
 * so that replays of a captured frame need no host call between them. */
int fh_copy_indexed(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t count, uint32_t* counter);

/* ------------------------------------------------------------------------ */
/* Uniform packing (host side)                                               */
/* ------------------------------------------------------------------------ */

/* The inputs of RenderUniformsBuffer.update (renderUniformsBuffer.ts
 * :205-288): the camera (camera.ts, CONFIG.camera), the model matrix, the
 * bounding spheres the shadow source frames (getMVP_ShadowSourceMatrix.ts
 * :14-79) and the CONFIG values it copies. Reals are double: the matrices
 * are built in double and stored as f32, as wgpu-matrix stores into
 * Float32Array. */
typedef struct fh_light_params {
  double posPhi, posTheta;     /* degrees; position = 2 * sphericalToCartesian */
  double color[3];
  double energy;
} fh_light_params;

typedef struct fh_render_params {
  uint32_t width, height;
  double cameraEye[3], cameraTarget[3], cameraUp[3];
  double fovDeg, near, far;
  double modelMatrix[16];      /* column-major */
  double hairSphere[4];        /* centre xyz, radius */
  const double* meshSpheres;   /* meshSphereCount x (centre xyz, radius) */
  uint32_t meshSphereCount;
  uint32_t tileSize;
  double avgSegmentsPerTile;
  double sdfBoundsMin[3], sdfBoundsMax[3], sdfDistanceOffset;
  double gridBoundsMin[3], gridBoundsMax[3];
  double collisionSphere[4];   /* CONFIG.hairSimulation.collisionSphere: xyz, radius */
  double fiberRadius;
  uint32_t displayMode, dbgTileModeMaxSegments, dbgSlicesModeMaxSlices, dbgShowTiles;
  double colors[3];            /* gamma, exposure, ditherStrength */
  double lightAmbient[4];
  fh_light_params lights[3];
  uint32_t shadowUsePCSS, shadowPCFRadius;
  double shadowBias, shadowStrength;
  double shadowPosPhi, shadowPosTheta, shadowDistance, shadowTarget[3];
  double aoRadius, aoDirectionOffset, aoFalloffStart2, aoFalloffEnd2;
  uint32_t aoNumDirections, aoNumSteps;
  double aoStrength;
  double hairColor0[3], hairColor1[3];
  double specular, shadows, weightTT, weightTRT, shift, roughness, colorRng, lumaRng, attenuation;
} fh_render_params;

/* RenderUniformsBuffer.update (renderUniformsBuffer.ts:205-288) with
 * encodeDebugMode (:396-414): host-only, writes *out (host memory). */
int fh_pack_render_uniforms(const fh_render_params* params, fh_render_uniforms* out);

/* The inputs of SimulationUniformsBuffer.update
 * (simulationUniformsBuffer.ts:86-152). */
typedef struct fh_sim_params {
  double modelMatrix[16];      /* column-major */
  double collisionSphere[4];   /* world xyz, radius */
  double windDirPhi, windDirTheta, windStrength;
  double sdfBoundsMin[3], sdfBoundsMax[3], sdfDistanceOffset;
  double gridBoundsMin[3], gridBoundsMax[3];
  uint32_t frameIdx, constraintIterations;
  double deltaTime, gravity, windStrengthLull, windColisionTraceOffset;
  double stiffnessLengthConstr, stiffnessCollisions, stiffnessSDF, volumePreservation, friction;
  double stiffnessGlobalConstr, globalConstrExtent, globalConstrFade, stiffnessLocalConstr;
  double windPhaseOffset, windStrengthFrequency, windStrengthJitter;
} fh_sim_params;

/* SimulationUniformsBuffer.update (simulationUniformsBuffer.ts:86-152):
 * host-only, writes *out (host memory). */
int fh_pack_sim_uniforms(const fh_sim_params* params, fh_sim_uniforms* out);

/* ------------------------------------------------------------------------ */
/* Kernel timing                                                             */
/* ------------------------------------------------------------------------ */

/* Per-kernel GPU time, the role of the reference's per-pass timestamp
 * queries (src/gpuProfiler.ts:7-60, BEGIN/END timestamps per pass). */
typedef struct fh_kernel_time {
  char name[48];      /* kernel name, e.g. "fine_raster" */
  uint32_t launches;  /* launches timed since the last read */
  uint32_t reserved;
  double total_ms;    /* summed duration of those launches */
} fh_kernel_time;

/* max_launches > 0 starts timing: each kernel the library launches on this
 * context is bracketed by two HIP events on the context's stream, for the
 * next max_launches launches (later ones are not timed). 0 stops it. Waits
 * for the stream; the events are created here, not per launch. */
int fh_profiler_enable(fh_ctx* ctx, uint32_t max_launches);
/* Waits for the stream, sums the timed launches per kernel name into
 * out[0 .. *n_out) (at most max_out names) and forgets them. */
int fh_profiler_read(fh_ctx* ctx, fh_kernel_time* out, uint32_t max_out, uint32_t* n_out);

/* ------------------------------------------------------------------------ */
/* Frame capture                                                             */
/* ------------------------------------------------------------------------ */

/* A recorded sequence of entry-point calls, replayed as one HIP graph (the
 * role of the reference's one queue.submit of a recorded GPUCommandEncoder
 * per frame, renderer.ts:108-120). */
typedef struct fh_graph fh_graph;

/* Starts recording the context's stream: the entry points called until
 * fh_capture_end enqueue nothing and are recorded instead. Call each entry
 * point once without capture first, so that none needs to grow the
 * context's scratch while recording (that returns FH_ERR_UNSUPPORTED).
 * Kernel timing is paused while recording. */
int fh_capture_begin(fh_ctx* ctx);
/* Ends the recording and instantiates it. */
int fh_capture_end(fh_ctx* ctx, fh_graph** out);
/* Enqueues one replay on the context's stream. The recorded calls read and
 * write the same buffers every replay; per-frame inputs (uniforms) change
 * through those buffers' contents. */
int fh_graph_launch(fh_ctx* ctx, fh_graph* graph);
int fh_graph_destroy(fh_graph* graph);

/* Branches: the entry points called between fh_branch_begin and
 * fh_branch_end run on a second stream of the context, after everything
 * enqueued on the context's stream before fh_branch_begin, and may overlap
 * what the context's stream runs next; fh_branch_join orders the context's
 * later work after them. The reference records one command encoder per
 * frame (renderer.ts:108-120) and leaves overlap to the WebGPU
 * implementation; a caller uses a branch for passes that neither read what
 * the overlapped passes write nor write what they read (the grid
 * accumulation beside the tile, sort and fine passes). A branch has a
 * scratch of its own. Consecutive branches run in order on the second
 * stream. While the context records (fh_record_begin) or times kernels
 * (fh_profiler_enable) the calls stay on the context's stream. Inside a
 * capture the fork and join become edges of the graph; fh_capture_end
 * joins an unjoined branch and fails if one is still open. fh_sync waits
 * for both streams. */
int fh_branch_begin(fh_ctx* ctx);
int fh_branch_end(fh_ctx* ctx);
int fh_branch_join(fh_ctx* ctx);
/* A branch on lane 0 (fh_branch_begin's) or lane 1: each lane has its own
 * stream and scratch, so the work of both lanes and of the context's stream
 * may run at once (three stages of a pipelined frame). One branch is open at
 * a time; fh_branch_end leaves its lane pending, and fh_branch_join joins
 * both lanes. Consecutive branches of one lane run in order. */
int fh_branch_begin_lane(fh_ctx* ctx, uint32_t lane);

/* Clear batches: the clears and copies of the entry points called between
 * fh_clear_batch_begin and fh_clear_batch_end (fh_tiles_clear,
 * fh_tiles_sort_clear, fh_fine_clear, fh_clear, fh_copy) are issued together
 * as one launch by fh_clear_batch_end, in place of one launch each; the
 * per-pass clears of a frame (hairTilesPass.ts:48-52, hairFinePass.ts:50-55,
 * hairTileSortPass.ts:63-68) touch disjoint buffers, so their order does not
 * matter. Any other entry point called inside a batch fails
 * (FH_ERR_INVALID_ARG), and so do fh_sync, fh_record_end and the branch
 * calls; fh_capture_end ends an open batch. */
int fh_clear_batch_begin(fh_ctx* ctx);
int fh_clear_batch_end(fh_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Batching: many hair instances per launch                                 */
/* ------------------------------------------------------------------------ */

/* The batched form of the chain (no counterpart in the reference, which
 * renders one hair object per frame, renderer.ts:212-283): each instance
 * keeps its own context, buffers and configuration; its frame is recorded
 * once through the ordinary entry points, and the recorded frames of all
 * instances are merged so that every kernel of the chain runs once for all
 * instances that use it. A merged launch gives instance i its own block
 * range, arguments and grid size, so each instance computes exactly what its
 * own calls compute. */
typedef struct fh_program fh_program;

/* Device bytes currently held by the context's scratch (for memory
 * budgets: the scratch grows on first use at a new size). */
int fh_scratch_bytes(fh_ctx* ctx, uint64_t* bytes);
/* Grows the context's scratch (the one in use: the branch's inside an open
 * branch) to at least `bytes` now, as a first call at a larger size would,
 * so that a later capture or recording finds it large enough. Programs
 * recorded and graphs captured before a reallocation address the freed
 * scratch: fh_program_launch and fh_graph_launch then return
 * FH_ERR_CAPACITY before enqueuing anything, and so they do once a context
 * they address was destroyed. Fails while capturing or recording. */
int fh_scratch_reserve(fh_ctx* ctx, uint64_t bytes);

/* Caps the workgroups of the grid-stride kernels launched on `ctx` (the tile
 * pass's queued pieces, the fine pass's gather and sort, the grid
 * accumulation); 0 restores the default of 8 per CU. An instance batched
 * with many others needs only its share of the device: the results do not
 * depend on this value. */
int fh_set_persistent_blocks(fh_ctx* ctx, uint32_t blocks);

/* Caps the queue of large-triangle bands of the software rasterizer
 * (fh_hw_hair, fh_shadow_map_hair, fh_shadow_map_meshes, fh_draw_meshes*)
 * at `bands` (at most 2^20, the default; 0 restores it). A triangle whose
 * bands do not fit is drawn by its own wave instead: the results do not
 * depend on this value. No counterpart in the reference: it exists so that
 * the overflow path can be exercised without a scene of 2^20 bands. */
int fh_set_raster_band_cap(fh_ctx* ctx, uint32_t bands);

/* Starts recording: until fh_record_end, the entry points called on `ctx`
 * check their arguments and keep their kernel launches (grid, LDS bytes,
 * arguments) instead of enqueuing them. Nothing runs on the device, so the
 * instance's state does not advance. */
int fh_record_begin(fh_ctx* ctx);
/* Ends recording; *out holds the recorded launches of this context. If the
 * context's scratch had to grow during the recording, the launches recorded
 * before the growth point at the old scratch: the recording is discarded,
 * *out is NULL and FH_ERR_CAPACITY is returned; recording the same calls
 * again then succeeds. */
int fh_record_end(fh_ctx* ctx, fh_program** out);
/* Merges `count` recorded programs into one program of batched launches.
 * Each input's launches keep their order; a kernel that several inputs
 * launch at a matching position becomes one launch. The arguments and block
 * ranges are uploaded to device memory owned by the result. The inputs stay
 * valid and unchanged. */
int fh_program_merge(fh_ctx* ctx, const fh_program* const* programs, uint32_t count, fh_program** out);
/* Number of launches of a merged program. */
int fh_program_launches(const fh_program* program, uint32_t* n_out);
/* Enqueues a merged program on the context's stream (also under
 * fh_capture_begin, to replay the batch as one graph). The buffers of every
 * instance must stay allocated while the program is used. */
int fh_program_launch(fh_ctx* ctx, const fh_program* program);
int fh_program_destroy(fh_program* program);

/* Statistics (no counterpart in the reference): the number of (tile, bin)
 * pairs whose slice memory (avgFragmentsPerSlice * slices * tile^2
 * fragments, hairSlicesDataBuffer.ts:27-40) the context's last
 * fh_fine_raster used up, so that later fragments of the bin were dropped
 * (hairFinePass.wgsl.ts:252-264). Waits for the context's stream. */
int fh_fine_exhausted_bins(fh_ctx* ctx, uint32_t* out);

/* ------------------------------------------------------------------------ */
/* Test hook                                                                 */
/* ------------------------------------------------------------------------ */

/* The raster-helper test kernel of src/passes/swHair/swRasterizer.test.ts
 * :165-250: rasterises every segment and atomicMax-es debugBarycentric
 * colours (ABGR u32) into out[(H - y) * W + x]. The reference pins it with
 * __test__/swRasterizer.snapshot.bin. */
int fh_debug_raster_barycentric(fh_ctx* ctx, const fh_render_uniforms* uniforms,
                                const fh_vec4* positions, const fh_vec4* tangents,
                                uint32_t strands_count, uint32_t points_per_strand,
                                uint32_t* out);

/* Test hook (no counterpart in the reference): the pixel bounds that the
 * tile and fine passes enumerate, checked against the edge tests
 * (swRasterizeHair.wgsl.ts:127-129, processHairSegment.wgsl.ts:14-22). For
 * quad i (v00, v01, v10, v11 as 8 floats at quads[8 i]) every pixel that
 * passes the four edge tests within 3 px of the quad's pixel box must lie in
 * the box and in its row's x span; out[4 i .. 4 i + 3] receives the pixels
 * outside the box, the pixels outside the span, the candidates the bounds
 * leave and the inside pixels (0, 0, 0, 0xffffffff for a quad whose box is
 * unbounded or over 200 px, which is not checked). */
int fh_debug_pixel_bounds(fh_ctx* ctx, const float* quads, uint32_t count, uint32_t* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* FH_HAIR_H */
