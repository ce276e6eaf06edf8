// The scene's mesh passes on gfx950: the background gradient
// (DrawBackgroundGradientPass), the meshes' depth, normals and colour
// (DrawMeshesPass) and the meshes in the shadow map (ShadowMapPass's mesh
// draw). Reference: src/passes/drawBackgroundGradient/
// drawBackgroundGradientPass.wgsl.ts:33-48, _shaderSnippets/noise.wgsl.ts
// :14-42, drawMeshes/drawMeshesPass.ts:58-115, drawMeshes/
// drawMeshesPass.wgsl.ts:46-158, _shaderSnippets/pbr.wgsl.ts, shadowMapPass/
// shadowMapPass.ts:34-60, :83-133, shadowMapPass/shadowMapPass.wgsl.ts,
// shadowMapPass/shared/sampleShadows.wgsl.ts:21-150, aoPass/shared/
// textureAo.wgsl.ts:7-12, scene/gpuMesh.ts:5-55, scene/loadScene.ts:25-60.
//
// The meshes are triangle lists (float32x3 positions and normals, u32
// indices), rasterized by the rules of fh_tri_raster.hpp: one 64-bit
// atomicMin of (depth bits, triangle index) per fragment, the triangle index
// running over the meshes in scene order, so the 'less' test resolves in
// primitive order whatever order the GPU runs the triangles in. The camera
// pass culls back faces (PIPELINE_PRIMITIVE_TRIANGLE_LIST, cullMode 'back',
// frontFace 'ccw': counter-clockwise in NDC, y up), the shadow pass culls
// nothing (shadowMapPass.ts:53-56). A resolve kernel then shades each pixel's
// winning fragment once: the forward shading of an opaque 'less' pass leaves
// exactly that fragment's colour in the target.
#include <algorithm>
#include <vector>

#include "fh_device.hpp"
#include "fh_launch.hpp"
#include "fh_tri_raster.hpp"

namespace fh {
int fill_u32(fh_ctx* ctx, void* dst, uint64_t n_words, uint32_t value);
}

// One mesh of a scene object list as the kernels read it.
struct MeshDesc {
  const float* pos;       // 3 floats per vertex
  const float* nrm;       // 3 floats per vertex
  const uint32_t* idx;    // 3 per triangle
  uint32_t triStart;      // first triangle index of this mesh in the scene
  uint32_t triCount;
  uint32_t preview;       // isColliderPreview
  uint32_t vertexCount;
};

struct fh_meshes {
  int device = 0;
  MeshDesc* dtable = nullptr;  // device copy of `table`
  std::vector<MeshDesc> table;
  uint32_t nTri = 0;
};

namespace {

using namespace fh;
using namespace fh::tri;

constexpr float kPi = 3.14159265358979323846f;
constexpr uint32_t kMaxMeshes = 256;
constexpr uint32_t kNoTriangle = 0xffffffffu;

// --------------------------------------------------------------------------
// Background (drawBackgroundGradientPass.wgsl.ts:33-48; noise.wgsl.ts:14-42)
// --------------------------------------------------------------------------

__device__ __forceinline__ float bg_hash(f2 p0) {
  const f2 q = mk2(p0.x * 0.3183099f + 0.71f, p0.y * 0.3183099f + 0.113f);
  const f2 p = mk2(50.0f * fractf(q.x), 50.0f * fractf(q.y));
  return -1.0f + 2.0f * fractf(p.x * p.y * (p.x + p.y));
}

__device__ __forceinline__ float bg_noise(f2 p) {
  const f2 i = mk2(floorf(p.x), floorf(p.y));
  const f2 f = mk2(fractf(p.x), fractf(p.y));
  // quintic interpolant f*f*f*(f*(f*6-15)+10), per component
  const f2 u = mk2(f.x * f.x * f.x * (f.x * (f.x * 6.0f - 15.0f) + 10.0f),
                   f.y * f.y * f.y * (f.y * (f.y * 6.0f - 15.0f) + 10.0f));
  const float a = mixf(bg_hash(i + mk2(0.0f, 0.0f)), bg_hash(i + mk2(1.0f, 0.0f)), u.x);
  const float b = mixf(bg_hash(i + mk2(0.0f, 1.0f)), bg_hash(i + mk2(1.0f, 1.0f)), u.x);
  return mixf(a, b, u.y);
}

// mat2x2f(1.6, 1.2, -1.2, 1.6) * uv, columns (1.6, 1.2) and (-1.2, 1.6)
__device__ __forceinline__ f2 bg_rot(f2 uv) { return mk2(1.6f * uv.x + -1.2f * uv.y, 1.2f * uv.x + 1.6f * uv.y); }

__device__ __forceinline__ float fractal_noise(f2 p, float scale) {
  f2 uv = p * scale;
  float f = 0.5000f * bg_noise(uv); uv = bg_rot(uv);
  f = f + 0.2500f * bg_noise(uv); uv = bg_rot(uv);
  f = f + 0.1250f * bg_noise(uv); uv = bg_rot(uv);
  f = f + 0.0625f * bg_noise(uv);
  return f;
}

__device__ __forceinline__ void store_rgba16f(void* hdr, size_t p, f3 c, float a) {
  reinterpret_cast<uint2*>(hdr)[p] = make_uint2(pack_half2(c.x, c.y), pack_half2(c.z, a));
}

struct BgArgs {
  const fh_render_uniforms* u;
  void* hdr;
  uint32_t W, H;
};

struct Background {
  using Args = BgArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "background";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t p = block * kThreads + threadIdx.x;
    if (p >= a.W * a.H) return;
    const fh_render_uniforms* __restrict__ u = a.u;
    const uint32_t px = p % a.W, py = p / a.W;
    const f2 uv = mk2(((float)px + 0.5f) / u->viewport[0], ((float)py + 0.5f) / u->viewport[1]);
    const float c = fractal_noise(uv, u->backgroundNoiseScale) * 0.5f + 0.5f;
    f3 color = mix(mk3(u->backgroundColor0[0], u->backgroundColor0[1], u->backgroundColor0[2]),
                   mk3(u->backgroundColor1[0], u->backgroundColor1[1], u->backgroundColor1[2]), c);
    const float g = 1.0f - uv.y;
    color = mix(color, mk3(g, g, g), u->backgroundGradientStrength);
    store_rgba16f(a.hdr, p, color, 1.0f);
  }
};

// --------------------------------------------------------------------------
// Mesh rasterization
// --------------------------------------------------------------------------

struct MeshArgs {
  const fh_render_uniforms* u;
  const MeshDesc* meshes;
  uint32_t nMeshes, nTri;
  unsigned long long* key;  // per pixel: (depth bits << 32) | triangle, kNoFragment when empty
  uint32_t* bigCount;
  uint2* big;
  // resolve outputs (camera pass: depth, normals, visibility and hdr, each
  // but depth optional; shadow pass: depth)
  float* depth;
  uint32_t* normals;
  uint32_t* visibility;  // scene triangle of each pixel's fragment, kNoTriangle where none
  void* hdr;
  const float* shadow;
  const float* ao;
  uint32_t shadowW, aoW, aoH;
  uint32_t W, H, shadowPass;
  uint32_t bigCap;          // queued bands at most (<= kBigCap)
};

__device__ __forceinline__ uint32_t mesh_of(const MeshArgs& a, uint32_t t) {
  uint32_t m = 0;
  while (m + 1u < a.nMeshes && t >= a.meshes[m + 1u].triStart) m++;
  return m;
}

// mat4 product a * b (column-major), column by column
__device__ __forceinline__ void mat_mul(const float* a, const float* b, float* out) {
  for (int c = 0; c < 4; c++) {
    const f4 r = mul(a, mk4(b[4 * c], b[4 * c + 1], b[4 * c + 2], b[4 * c + 3]));
    out[4 * c] = r.x; out[4 * c + 1] = r.y; out[4 * c + 2] = r.z; out[4 * c + 3] = r.w;
  }
}

__device__ __forceinline__ f3 vtx3(const float* v, uint32_t i) {
  return mk3(v[3ull * i], v[3ull * i + 1], v[3ull * i + 2]);
}

// Clip positions of the three vertices of scene triangle t (main_vs of
// drawMeshesPass.wgsl.ts:56-79 or shadowMapPass.wgsl.ts:19-25). False when
// the pass does not draw the triangle (collider previews in the shadow pass).
__device__ __forceinline__ bool mesh_clip(const MeshArgs& a, uint32_t t, uint32_t& m, uint32_t* vi, f4* c) {
  m = mesh_of(a, t);
  const MeshDesc& d = a.meshes[m];
  if (a.shadowPass && d.preview) return false;
  const uint32_t lt = t - d.triStart;
  const fh_render_uniforms* __restrict__ u = a.u;
  float mvp[16];
  if (a.shadowPass) {
    for (int k = 0; k < 16; k++) mvp[k] = u->shadows.sourceMVP_Matrix[k];
  } else if (d.preview) {
    float vm[16];  // getMVP_Mat: projMat * (viewMat * modelMat)
    mat_mul(u->viewMatrix, u->collisionSphereModelMatrix, vm);
    mat_mul(u->projMatrix, vm, mvp);
  } else {
    for (int k = 0; k < 16; k++) mvp[k] = u->mvpMatrix[k];
  }
  for (int q = 0; q < 3; q++) {
    vi[q] = d.idx[3ull * lt + q];
    const f3 p = vtx3(d.pos, vi[q]);
    c[q] = mul(mvp, mk4(p.x, p.y, p.z, 1.0f));
  }
  return true;
}

__device__ __forceinline__ bool mesh_setup(const MeshArgs& a, uint32_t t, uint32_t W, uint32_t H, SetupTri& s,
                                           uint32_t& m, uint32_t* vi) {
  f4 c[3];
  if (!mesh_clip(a, t, m, vi, c)) return false;
  s = setup_from(W, H, c[0], c[1], c[2]);
  // back faces: clockwise in NDC (y up) is counter-clockwise on screen (y
  // down), which setup leaves unswapped
  if (!a.shadowPass && !s.swapped) return false;
  return s.ok;
}

struct MeshRaster {
  using Args = MeshArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "mesh_raster";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t gid = xcd_order(block, blocks) * kThreads + threadIdx.x;
    const uint32_t t = gid / kRowSplit, k = gid % kRowSplit;
    bool wave = false;
    if (t < a.nTri) {
      SetupTri s;
      uint32_t m, vi[3];
      if (mesh_setup(a, t, a.W, a.H, s, m, vi)) {
        const PixBox b = pix_box(a.W, a.H, s);
        const long long bw = b.x1 - b.x0 + 1, bh = b.y1 - b.y0 + 1;
        if (bw > 0 && bh > 0) {
          if (bw * bh <= kWaveRaster) {
            raster_rows(s, t, (long long)k, (long long)kRowSplit, a.W, a.H, nullptr, a.key);
          } else if (k == 0u) {
            const uint32_t bands = (uint32_t)((bh + kBandRows - 1) / kBandRows);
            const uint32_t first = atomicAdd(a.bigCount, bands);
            if (first + bands <= a.bigCap) {
              for (uint32_t q = 0; q < bands; q++) a.big[first + q] = make_uint2(t, q);
            } else {
              for (uint32_t q = first; q < a.bigCap; q++) a.big[q] = make_uint2(kBigSentinel, 0u);
              wave = true;
            }
          }
        }
      }
    }
    // queue full: the wave draws the triangle, one row per lane
    const uint32_t lane = __lane_id();
    uint64_t todo = __ballot(wave);
    while (todo) {
      const uint32_t src = (uint32_t)__builtin_ctzll(todo);
      todo &= todo - 1ull;
      const uint32_t tb = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)src);
      SetupTri sb;
      uint32_t m, vi[3];
      if (mesh_setup(a, tb, a.W, a.H, sb, m, vi))
        raster_rows(sb, tb, (long long)lane, 64, a.W, a.H, nullptr, a.key);
    }
  }
};

// The queued bands: a workgroup per band, its threads taking the band's
// (row, 64-pixel chunk) items in turn; every workgroup exits after the last
// band.
struct MeshRasterBig {
  using Args = MeshArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "mesh_raster_big";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t count = min(*a.bigCount, a.bigCap);
    for (uint32_t i = block; i < count; i += blocks) {
      const uint2 e = a.big[i];
      if (e.x >= a.nTri) continue;  // kBigSentinel
      SetupTri s;
      uint32_t m, vi[3];
      if (!mesh_setup(a, e.x, a.W, a.H, s, m, vi)) continue;
      const PixBox b = pix_box(a.W, a.H, s);
      const long long y0 = b.y0 + (long long)e.y * kBandRows, y1 = min(b.y1, y0 + kBandRows - 1);
      const long long cpr = (b.x1 - b.x0) / 64 + 1;
      const long long items = (y1 - y0 + 1) * cpr;
      for (long long it = threadIdx.x; it < items; it += kThreads) {
        const long long row = it / cpr, chunk = it - row * cpr;
        const long long xa = b.x0 + chunk * 64;
        raster_row(s, e.x, y0 + row, xa, min(b.x1, xa + 63), a.W, nullptr, a.key);
      }
    }
  }
};

// --------------------------------------------------------------------------
// Mesh shading (drawMeshesPass.wgsl.ts:88-158, pbr.wgsl.ts)
// --------------------------------------------------------------------------

struct Material {
  f3 positionWS, normal, toEye, albedo;
  float roughness, isMetallic, ao, shadow;
};

__device__ __forceinline__ float dot_max0(f3 a, f3 b) { return fmaxf(0.0f, dot(a, b)); }

// FresnelSchlick — pbr.wgsl.ts:18-20
__device__ __forceinline__ f3 fresnel_schlick(float cosTheta, f3 f0) {
  const float k = powf(1.0f - cosTheta, 5.0f);
  return mk3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}

// DistributionGGX — pbr.wgsl.ts:35-44
__device__ __forceinline__ float distribution_ggx(f3 N, f3 H, float roughness) {
  const float a = roughness * roughness;
  const float a2 = a * a;
  const float NdotH = dot_max0(N, H);
  const float NdotH2 = NdotH * NdotH;
  float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
  denom = kPi * denom * denom;
  return a2 / denom;
}

// GeometrySchlickGGX — pbr.wgsl.ts:54-59
__device__ __forceinline__ float geometry_schlick_ggx(float NdotV, float roughness) {
  const float r = roughness + 1.0f;
  const float k = (r * r) / 8.0f;
  const float denom = NdotV * (1.0f - k) + k;
  return NdotV / denom;
}

// disneyPBR — pbr.wgsl.ts:140-160 with pbr_CookTorrance :112-131 and
// pbr_mixDiffuseAndSpecular :133-138
__device__ __forceinline__ f3 disney_pbr(const Material& m, const fh_light& light) {
  const f3 N = m.normal, V = m.toEye;
  const f3 L = normalize(mk3(light.position[0], light.position[1], light.position[2]) - m.positionWS);
  const f3 lambert = m.albedo;
  const f3 H = normalize(V + L);
  const f3 F0 = mix(mk3(0.04f, 0.04f, 0.04f), m.albedo, m.isMetallic);
  const f3 F = fresnel_schlick(dot_max0(H, V), F0);
  const float NdotV = dot_max0(N, V), NdotL = dot_max0(N, L);
  const float G = geometry_schlick_ggx(NdotL, m.roughness) * geometry_schlick_ggx(NdotV, m.roughness);
  const float NDF = distribution_ggx(N, H, m.roughness);
  const f3 numerator = F * (NDF * G);
  const float denominator = 4.0f * dot_max0(N, V) * dot_max0(N, L);
  const f3 specular = numerator / fmaxf(denominator, 0.001f);
  const f3 kD = mix(mk3(1.0f, 1.0f, 1.0f) - F, mk3(0.0f, 0.0f, 0.0f), m.isMetallic);
  const f3 brdf = kD * lambert + specular;
  const f3 radiance = mk3(light.colorAndEnergy[0], light.colorAndEnergy[1], light.colorAndEnergy[2]) *
                      light.colorAndEnergy[3];
  return brdf * radiance * NdotL;
}

// calculateDirectionalShadow — sampleShadows.wgsl.ts:47-84 with
// _sampleShadowMap :95-131 and _isValidShadowSample :136-150
__device__ __forceinline__ float directional_shadow(const MeshArgs& a, f3 posWS, f3 normalWS, f3 pss) {
  const fh_shadows& sh = a.u->shadows;
  const f3 toSrc = normalize(mk3(sh.sourcePosition[0], sh.sourcePosition[1], sh.sourcePosition[2]) - posWS);
  const float bias = fmaxf(fabsf(sh.bias) * (1.0f - dot(normalWS, toSrc)), 0.0005f);
  int32_t radius = (int32_t)sh.PCF_Radius;
  if (sh.usePCSS > 0u) {
    const float sm = gather_avg(a.shadow, a.shadowW, a.shadowW, mk2(pss.x, pss.y));
    const float dd = fmaxf(pss.z - sm, 0.0f);
    radius = 1 + f2i(dd / sm * 10.0f);  // PCSS_PENUMBRA_BASE + depthDiff / depth * PCSS_PENUMBRA_WIDTH
  }
  const float texel = 1.0f / (float)a.shadowW;
  // the taps x-major as in the reference's loops, eight gathers issued
  // before their comparisons (a count of 0 / 1 results, exact in any order)
  const int32_t side = radius * 2 + 1, taps = side * side;
  float shadow = 0.0f;
  for (int32_t k0 = 0; k0 < taps; k0 += 8) {
    float g[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int32_t k = min(k0 + j, taps - 1);
      const int32_t x = -radius + k / side, y = -radius + k % side;
      const f2 uv = mk2(pss.x, pss.y) + mk2((float)x * texel, (float)y * texel);
      g[j] = gather_avg(a.shadow, a.shadowW, a.shadowW, uv);
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (k0 + j < taps) shadow += (pss.z - bias > g[j]) ? 1.0f : 0.0f;
  }
  const float pcf = (float)(radius * 2 + 1);
  const float res = shadow / (pcf * pcf);
  if (pss.z > 1.0f) return 0.0f;
  if (pss.x < 0.0f || pss.x > 1.0f || pss.y < 0.0f || pss.y > 1.0f) return 0.0f;
  return res;
}

// main_fs (drawMeshesPass.wgsl.ts:88-130, doShading :133-146) for the
// fragment of scene triangle t at pixel (px, py): the encoded normal and,
// when `colour`, the colour. False when the triangle is not drawn (not
// reached for a pixel's winner).
__device__ __forceinline__ bool mesh_fragment(const MeshArgs& a, uint32_t t, uint32_t px, uint32_t py, bool colour,
                                              uint32_t& normalOut, f3& colourOut) {
  SetupTri s;
  uint32_t m, vi[3];
  if (!mesh_setup(a, t, a.W, a.H, s, m, vi)) return false;
  const MeshDesc& d = a.meshes[m];
  const fh_render_uniforms* __restrict__ u = a.u;
  float l[3], wq[3];
  barycentrics(s, (long long)px * 256 + 128, (long long)py * 256 + 128, l);
  const float den = persp_weights(s, l, wq);
  const float* model = d.preview ? u->collisionSphereModelMatrix : u->modelMatrix;
  const float* smvp = u->shadows.sourceMVP_Matrix;
  f3 pos = mk3(0.0f, 0.0f, 0.0f), nrm = mk3(0.0f, 0.0f, 0.0f), pss = mk3(0.0f, 0.0f, 0.0f);
  for (int q = 0; q < 3; q++) {
    const uint32_t v = vi[s.vi[q]];
    const f3 pv = vtx3(d.pos, v), nv = vtx3(d.nrm, v);
    // transformNormalToWorldSpace (shaderSnippets.wgls.ts:15-19)
    const f3 nws = normalize(xyz(mul(model, mk4(nv.x, nv.y, nv.z, 0.0f))));
    // projectToShadowSpace (sampleShadows.wgsl.ts:21-34); the preview's
    // shadow matrix is mat4x4<f32>(), all zeros
    f4 lp = mk4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!d.preview) lp = mul(smvp, mk4(pv.x, pv.y, pv.z, 1.0f));
    const f3 sp = mk3(lp.x * 0.5f + 0.5f, 1.0f - (lp.y * 0.5f + 0.5f), lp.z);
    pos = pos + pv * wq[q];
    nrm = nrm + nws * wq[q];
    pss = pss + sp * wq[q];
  }
  pos = pos / den;
  nrm = nrm / den;
  pss = pss / den;
  // createDefaultMaterial (:148-168)
  Material mt;
  mt.normal = normalize(nrm);
  normalOut = encode_oct(mt.normal);
  if (!colour) return true;
  mt.positionWS = pos;
  mt.toEye = normalize(mk3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]) - pos);
  mt.albedo = mk3(0.9f, 0.9f, 0.9f);
  mt.roughness = 0.8f;
  mt.isMetallic = 0.0f;
  mt.shadow = 1.0f - directional_shadow(a, pos, mt.normal, pss);
  // sampleAo (textureAo.wgsl.ts:7-12) at the fragment's position
  const f2 at = mk2(((float)px + 0.5f) / u->viewport[0], ((float)py + 0.5f) / u->viewport[1]);
  const float aov = tex_load(a.ao, a.aoW, a.aoH, f2i((float)a.aoW * at.x), f2i((float)a.aoH * at.y));
  mt.ao = mixf(1.0f, aov, u->ao.strength);
  if (d.preview) {
    mt.albedo = mk3(1.0f, 1.0f, 1.0f);
    mt.shadow = 1.0f;
    mt.ao = 1.0f;
    mt.roughness = 1.0f;
    mt.isMetallic = 0.4f;
  }
  // doShading (:133-146): (0 + L0) + L1 + L2
  const f3 ambient = mk3(u->lightAmbient[0], u->lightAmbient[1], u->lightAmbient[2]) * u->lightAmbient[3];
  f3 rs = mk3(0.0f, 0.0f, 0.0f);
  rs = rs + disney_pbr(mt, u->light0);
  rs = rs + disney_pbr(mt, u->light1);
  rs = rs + disney_pbr(mt, u->light2);
  rs = rs * satf(mt.ao);
  rs = rs * clampf(mt.shadow, 1.0f - u->shadows.strength, 1.0f);
  f3 color = ambient + rs;
  if (d.preview) color = mix(color, mk3(1.0f, 1.0f, 0.0f), 0.5f);
  colourOut = color;
  return true;
}

// The winner of each pixel. Camera pass: its depth, its normal
// (encodeOctahedronNormal of the normalized interpolated world normal), its
// triangle (visibility, when asked for) and its colour (when hdr is set);
// pixels with no fragment get the clear values (depth 1, normals 0) and keep
// the colour already in the target ('load'). Shadow pass: the depth.
struct MeshResolve {
  using Args = MeshArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "mesh_resolve";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t p = xcd_order(block, blocks) * kThreads + threadIdx.x;
    if (p >= a.W * a.H) return;
    const unsigned long long key = a.key[p];
    if (key == kNoFragment) {
      a.depth[p] = 1.0f;
      if (a.shadowPass) return;
      a.normals[p] = 0u;
      if (a.visibility) a.visibility[p] = kNoTriangle;
      return;
    }
    a.depth[p] = __uint_as_float((uint32_t)(key >> 32));
    if (a.shadowPass) return;
    const uint32_t t = (uint32_t)key;
    if (a.visibility) a.visibility[p] = t;
    uint32_t n = 0u;
    f3 c;
    const bool drawn = mesh_fragment(a, t, p % a.W, p / a.W, a.hdr != nullptr, n, c);
    a.normals[p] = drawn ? n : 0u;
    if (drawn && a.hdr) store_rgba16f(a.hdr, p, c, 1.0f);
  }
};

// The colour of each pixel's fragment from a visibility texture
// (fh_shade_meshes): what MeshResolve writes with hdr set. Pixels without a
// triangle (or with an index past the scene's) keep the target's colour.
struct MeshShade {
  using Args = MeshArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "mesh_shade";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t p = xcd_order(block, blocks) * kThreads + threadIdx.x;
    if (p >= a.W * a.H) return;
    const uint32_t t = a.visibility[p];
    if (t >= a.nTri) return;
    uint32_t n;
    f3 c;
    if (mesh_fragment(a, t, p % a.W, p / a.W, true, n, c)) store_rgba16f(a.hdr, p, c, 1.0f);
  }
};

// Any index at or past its mesh's vertex count sets *bad.
struct MeshIndexCheck {
  struct Args {
    const uint32_t* idx;
    uint64_t n;
    uint32_t vertexCount;
    uint32_t* bad;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "mesh_index_check";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    for (uint64_t i = (uint64_t)block * kThreads + threadIdx.x; i < a.n; i += (uint64_t)blocks * kThreads)
      if (a.idx[i] >= a.vertexCount) atomicOr(a.bad, 1u);
  }
};

int raster_meshes(fh_ctx* ctx, MeshArgs& a) {
  const uint64_t px = (uint64_t)a.W * a.H;
  const size_t bigOff = align_up(px * 8, 256);
  int rc = ensure_scratch(ctx, bigOff + 256 + (size_t)kBigCap * 8);
  a.bigCap = ctx->band_cap ? ctx->band_cap : kBigCap;
  if (rc) return rc;
  a.key = static_cast<unsigned long long*>(ctx->scratch);
  a.bigCount = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->scratch) + bigOff);
  a.big = reinterpret_cast<uint2*>(static_cast<char*>(ctx->scratch) + bigOff + 256);
  if ((rc = fill_u32(ctx, a.key, px * 2, 0xffffffffu))) return rc;
  if ((rc = fill_u32(ctx, a.bigCount, 1, 0u))) return rc;
  if (a.nTri) {
    if ((rc = launch<MeshRaster>(ctx, div_ceil(a.nTri, MeshRaster::kThreads / kRowSplit), 0, a))) return rc;
    if ((rc = launch<MeshRasterBig>(ctx, persistent_blocks(ctx), 0, a))) return rc;
  }
  return launch<MeshResolve>(ctx, div_ceil((uint32_t)px, MeshResolve::kThreads), 0, a);
}

}  // namespace

extern "C" {

int fh_meshes_create(fh_ctx* ctx, const fh_mesh* meshes, uint32_t count, fh_meshes** out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(out != nullptr, "fh_meshes_create: out is null");
  *out = nullptr;
  FH_REQUIRE(count <= kMaxMeshes, "fh_meshes_create: at most 256 meshes");
  FH_REQUIRE(count == 0 || meshes != nullptr, "fh_meshes_create: null mesh list");
  FH_REQUIRE(!ctx->recording && !ctx->capturing, "fh_meshes_create: the context is recording or capturing");
  std::unique_ptr<fh_meshes> ms(new fh_meshes());
  ms->device = ctx->device;
  uint64_t total = 0;
  for (uint32_t i = 0; i < count; i++) {
    const fh_mesh& m = meshes[i];
    FH_REQUIRE(m.triangle_count == 0 || (m.positions && m.normals && m.indices), "fh_meshes_create: null buffer");
    MeshDesc d{m.positions, m.normals, m.indices, (uint32_t)total, m.triangle_count, m.is_collider_preview ? 1u : 0u,
               m.vertex_count};
    total += m.triangle_count;
    FH_REQUIRE(total < (1ull << 31), "fh_meshes_create: too many triangles");
    ms->table.push_back(d);
  }
  ms->nTri = (uint32_t)total;
  FH_HIP(hipSetDevice(ctx->device));
  // every index must address a vertex of its mesh (the kernels read
  // positions[index] unchecked)
  uint32_t* bad = nullptr;
  FH_HIP(hipMalloc(&bad, sizeof(uint32_t)));
  int rc = FH_OK;
  uint32_t hbad = 0;
  if (hipMemsetAsync(bad, 0, sizeof(uint32_t), ctx->stream) != hipSuccess) rc = fail(FH_ERR_HIP, "fh_meshes_create: memset");
  for (const MeshDesc& d : ms->table) {
    if (rc || d.triCount == 0) continue;
    MeshIndexCheck::Args ca{d.idx, 3ull * d.triCount, d.vertexCount, bad};
    rc = launch<MeshIndexCheck>(ctx, std::min<uint32_t>(div_ceil((uint32_t)std::min<uint64_t>(ca.n, 1u << 30), 256u), 4096u),
                                0, ca);
  }
  if (!rc && hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = fail(FH_ERR_HIP, "fh_meshes_create: read back");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_ERR_HIP, "fh_meshes_create: sync");
  hipFree(bad);
  if (rc) return rc;
  FH_REQUIRE(hbad == 0, "fh_meshes_create: an index is at or past its mesh's vertex_count");
  if (count) {
    FH_HIP(hipMalloc(&ms->dtable, count * sizeof(MeshDesc)));
    FH_HIP(hipMemcpy(ms->dtable, ms->table.data(), count * sizeof(MeshDesc), hipMemcpyHostToDevice));
  }
  *out = ms.release();
  return FH_OK;
}

int fh_meshes_destroy(fh_meshes* meshes) {
  if (!meshes) return FH_OK;
  if (meshes->dtable) {
    FH_HIP(hipSetDevice(meshes->device));
    FH_HIP(hipFree(meshes->dtable));
  }
  delete meshes;
  return FH_OK;
}

int fh_meshes_triangles(const fh_meshes* meshes, uint32_t* n_out) {
  FH_REQUIRE(meshes && n_out, "fh_meshes_triangles: null argument");
  *n_out = meshes->nTri;
  return FH_OK;
}

int fh_draw_background(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                       void* hdr_texture) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && hdr_texture, "fh_draw_background: null buffer");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), "fh_draw_background: viewport too large");
  if (vp_w == 0 || vp_h == 0) return FH_OK;
  BgArgs a{uniforms, hdr_texture, vp_w, vp_h};
  return launch<Background>(ctx, div_ceil(vp_w * vp_h, Background::kThreads), 0, a);
}

int fh_draw_meshes(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                   const fh_meshes* meshes, fh_tex2d shadow_map, fh_tex2d ao, float* depth_texture,
                   void* normals_texture, void* hdr_texture) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && meshes && depth_texture && normals_texture && hdr_texture, "fh_draw_meshes: null buffer");
  FH_REQUIRE(meshes->device == ctx->device, "fh_draw_meshes: meshes of another device");
  FH_REQUIRE(shadow_map.data && shadow_map.width > 0 && shadow_map.width == shadow_map.height,
             "fh_draw_meshes: the shadow map must be square");
  FH_REQUIRE(ao.data && ao.width > 0 && ao.height > 0, "fh_draw_meshes: null AO texture");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), "fh_draw_meshes: viewport too large");
  if (vp_w == 0 || vp_h == 0) return FH_OK;
  MeshArgs a{};
  a.u = uniforms;
  a.meshes = meshes->dtable;
  a.nMeshes = (uint32_t)meshes->table.size();
  a.nTri = meshes->nTri;
  a.depth = depth_texture;
  a.normals = static_cast<uint32_t*>(normals_texture);
  a.hdr = hdr_texture;
  a.shadow = shadow_map.data;
  a.shadowW = shadow_map.width;
  a.ao = ao.data;
  a.aoW = ao.width;
  a.aoH = ao.height;
  a.W = vp_w;
  a.H = vp_h;
  a.shadowPass = 0;
  return raster_meshes(ctx, a);
}

int fh_draw_meshes_visibility(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                              const fh_meshes* meshes, float* depth_texture, void* normals_texture,
                              uint32_t* visibility) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && meshes && depth_texture && normals_texture && visibility,
             "fh_draw_meshes_visibility: null buffer");
  FH_REQUIRE(meshes->device == ctx->device, "fh_draw_meshes_visibility: meshes of another device");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), "fh_draw_meshes_visibility: viewport too large");
  if (vp_w == 0 || vp_h == 0) return FH_OK;
  MeshArgs a{};
  a.u = uniforms;
  a.meshes = meshes->dtable;
  a.nMeshes = (uint32_t)meshes->table.size();
  a.nTri = meshes->nTri;
  a.depth = depth_texture;
  a.normals = static_cast<uint32_t*>(normals_texture);
  a.visibility = visibility;
  a.W = vp_w;
  a.H = vp_h;
  return raster_meshes(ctx, a);
}

int fh_shade_meshes(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                    const fh_meshes* meshes, const uint32_t* visibility, fh_tex2d shadow_map, fh_tex2d ao,
                    void* hdr_texture) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && meshes && visibility && hdr_texture, "fh_shade_meshes: null buffer");
  FH_REQUIRE(meshes->device == ctx->device, "fh_shade_meshes: meshes of another device");
  FH_REQUIRE(shadow_map.data && shadow_map.width > 0 && shadow_map.width == shadow_map.height,
             "fh_shade_meshes: the shadow map must be square");
  FH_REQUIRE(ao.data && ao.width > 0 && ao.height > 0, "fh_shade_meshes: null AO texture");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), "fh_shade_meshes: viewport too large");
  if (vp_w == 0 || vp_h == 0 || meshes->nTri == 0) return FH_OK;
  MeshArgs a{};
  a.u = uniforms;
  a.meshes = meshes->dtable;
  a.nMeshes = (uint32_t)meshes->table.size();
  a.nTri = meshes->nTri;
  a.visibility = const_cast<uint32_t*>(visibility);
  a.hdr = hdr_texture;
  a.shadow = shadow_map.data;
  a.shadowW = shadow_map.width;
  a.ao = ao.data;
  a.aoW = ao.width;
  a.aoH = ao.height;
  a.W = vp_w;
  a.H = vp_h;
  return launch<MeshShade>(ctx, div_ceil(vp_w * vp_h, MeshShade::kThreads), 0, a);
}

int fh_shadow_map_meshes(fh_ctx* ctx, const fh_render_uniforms* uniforms, const fh_meshes* meshes, float* shadow_map,
                         uint32_t shadow_size) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && meshes && shadow_map, "fh_shadow_map_meshes: null buffer");
  FH_REQUIRE(meshes->device == ctx->device, "fh_shadow_map_meshes: meshes of another device");
  FH_REQUIRE(shadow_size > 0 && shadow_size <= 16384, "fh_shadow_map_meshes: shadow_size in [1, 16384]");
  MeshArgs a{};
  a.u = uniforms;
  a.meshes = meshes->dtable;
  a.nMeshes = (uint32_t)meshes->table.size();
  a.nTri = meshes->nTri;
  a.depth = shadow_map;
  a.W = shadow_size;
  a.H = shadow_size;
  a.shadowPass = 1;
  return raster_meshes(ctx, a);
}

}  // extern "C"
