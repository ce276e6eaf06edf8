// Producers of the shading pass's inputs that the reference draws with the
// hardware rasterizer (gfx950): the hair's depth and normals (HwHairPass),
// the hair in the shadow map (ShadowMapPass's hair draw) and GTAO (AoPass).
// Reference: src/passes/hwHair/hwHairPass.wgsl.ts:46-99,
// hwHair/shaderImpl/hwRasterizeHair.wgsl.ts:15-60, hwHair/hwHairPass.ts
// :79-141, scene/hair/hairIndexBuffer.ts:20-54,
// shadowMapPass/shadowMapHairPass.wgsl.ts, shadowMapPass/shadowMapPass.ts
// :100-115, aoPass/aoPass.wgsl.ts:48-79, aoPass/gtao.wgsl.ts:26-136.
//
// Rasterization (the rules are in fh_tri_raster.hpp). Each segment is the
// reference's two triangles (indices 2i, 2i+1, 2i+2 and 2i+2, 2i+1, 2i+3 over
// the strand's left / right edge vertices). The depth test is the pipelines'
// 'less' (shared.ts:22-26) against the depth already in the texture (the
// meshes), and among the hair fragments of one pixel the least depth wins,
// the earliest triangle on a tie, as in API order; a resolve pass writes the
// winner's depth and its perspective-correct normal.
#include <algorithm>

#include "fh_device.hpp"
#include "fh_launch.hpp"
#include "fh_tri_raster.hpp"

namespace fh {
int fill_u32(fh_ctx* ctx, void* dst, uint64_t n_words, uint32_t value);
}

namespace {

using namespace fh;
using namespace fh::tri;

constexpr float kFloatEps = 1e-7f;  // FLOAT_EPSILON (shaderSnippets.wgls.ts:67)
constexpr float kPi = 3.14159265358979323846f;
constexpr float kHalfPi = 1.57079632679489661923f;
constexpr float kNear = 0.01f, kFar = 100.0f;  // CONFIG.camera.projection (constants.ts:131-135)

__device__ __forceinline__ f3 safe_normalize3(f3 v) {
  return dot(v, v) >= kFloatEps ? normalize(v) : mk3(0.0f, 0.0f, 0.0f);
}

struct RasterArgs {
  const fh_render_uniforms* u;
  const float4* pos;
  const float4* tan;
  float* depth;             // the depth texture: tested ('less'), then written by the resolve
  uint32_t* normals;        // rg16float, may be null (shadow map: depth only)
  unsigned long long* key;  // per pixel: (depth bits << 32) | triangle, kNoFragment when empty
  uint32_t* bigCount;       // queued bands of large triangles
  uint2* big;               // (triangle, band)
  uint32_t W, H, P, nTri, shadow;
  uint32_t bigCap;          // queued bands at most (<= kBigCap)
};

// Clip-space corner `side` (0: +thickness, 'right'; 1: -thickness, 'left') of
// point `idx` — hwRasterizeHair (hwRasterizeHair.wgsl.ts:23-48)
__device__ __forceinline__ f4 edge_vertex(const RasterArgs& a, uint32_t idx, uint32_t side) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const float* mv = a.shadow ? u->shadows.sourceModelViewMat : u->modelViewMat;
  const float* proj = a.shadow ? u->shadows.sourceProjMatrix : u->projMatrix;
  const float radius = a.shadow ? u->shadows.sourcePosition[3] : u->fiberRadius;  // getShadowFiberRadius
  const float4 p = a.pos[idx], t = a.tan[idx];
  const f4 pv = mul(mv, mk4(p.x, p.y, p.z, 1.0f));
  const f4 tv = mul(mv, mk4(t.x, t.y, t.z, 1.0f));
  const f3 right = safe_normalize3(cross(xyz(tv), mk3(0.0f, 0.0f, 1.0f)));
  const f3 th = right * radius;
  const f3 e = side ? xyz(pv) - th : xyz(pv) + th;
  return mul(proj, mk4(e.x, e.y, e.z, 1.0f));
}

// Point index and side of the three vertices of triangle t (hairIndexBuffer.ts:39-46)
__device__ __forceinline__ void tri_vertices(uint32_t t, uint32_t P, uint32_t* idx, uint32_t* side) {
  const uint32_t seg = t >> 1, strand = seg / (P - 1u), j = seg - strand * (P - 1u);
  const uint32_t id = strand * P + j;
  if ((t & 1u) == 0u) {  // 2id, 2id+1, 2id+2
    idx[0] = id; side[0] = 0u; idx[1] = id; side[1] = 1u; idx[2] = id + 1u; side[2] = 0u;
  } else {               // 2id+2, 2id+1, 2id+3
    idx[0] = id + 1u; side[0] = 0u; idx[1] = id; side[1] = 1u; idx[2] = id + 1u; side[2] = 1u;
  }
}

__device__ __forceinline__ SetupTri setup_tri(const RasterArgs& a, uint32_t t) {
  uint32_t idx[3], side[3];
  tri_vertices(t, a.P, idx, side);
  return setup_from(a.W, a.H, edge_vertex(a, idx[0], side[0]), edge_vertex(a, idx[1], side[1]),
                    edge_vertex(a, idx[2], side[2]));
}

// Both corners of point idx (0: +thickness, 1: -thickness), sharing the
// point's transform: the same values edge_vertex gives for each side.
__device__ __forceinline__ void point_corners(const RasterArgs& a, uint32_t idx, f4& plus, f4& minus) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const float* mv = a.shadow ? u->shadows.sourceModelViewMat : u->modelViewMat;
  const float* proj = a.shadow ? u->shadows.sourceProjMatrix : u->projMatrix;
  const float radius = a.shadow ? u->shadows.sourcePosition[3] : u->fiberRadius;
  const float4 p = a.pos[idx], t = a.tan[idx];
  const f4 pv = mul(mv, mk4(p.x, p.y, p.z, 1.0f));
  const f4 tv = mul(mv, mk4(t.x, t.y, t.z, 1.0f));
  const f3 right = safe_normalize3(cross(xyz(tv), mk3(0.0f, 0.0f, 1.0f)));
  const f3 th = right * radius;
  const f3 ep = xyz(pv) + th, em = xyz(pv) - th;
  plus = mul(proj, mk4(ep.x, ep.y, ep.z, 1.0f));
  minus = mul(proj, mk4(em.x, em.y, em.z, 1.0f));
}

// The hair triangles' rows: the texture's depth is the 'less' reference.
__device__ __forceinline__ void hair_rows(const RasterArgs& a, const SetupTri& s, uint32_t t, long long first,
                                          long long stride) {
  raster_rows(s, t, first, stride, a.W, a.H, a.depth, a.key);
}

struct HairRaster {
  using Args = RasterArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "hair_raster";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t gid = xcd_order(block, blocks) * kThreads + threadIdx.x;
    const uint32_t seg = gid / kRowSplit, k = gid % kRowSplit;
    // the segment's two triangles (hairIndexBuffer.ts:39-46): (A+, A-, B+), (B+, A-, B-)
    uint32_t bigMask = 0;
    if (2u * seg < a.nTri) {
      const uint32_t strand = seg / (a.P - 1u), j = seg - strand * (a.P - 1u);
      const uint32_t ia = strand * a.P + j;
      f4 ap, am, bp, bm;
      point_corners(a, ia, ap, am);
      point_corners(a, ia + 1u, bp, bm);
      for (uint32_t h = 0; h < 2u; h++) {
        const uint32_t t = 2u * seg + h;
        const SetupTri s = h == 0u ? setup_from(a.W, a.H, ap, am, bp) : setup_from(a.W, a.H, bp, am, bm);
        if (!s.ok) continue;
        const PixBox b = pix_box(a.W, a.H, s);
        const long long bw = b.x1 - b.x0 + 1, bh = b.y1 - b.y0 + 1;
        if (bw <= 0 || bh <= 0) continue;
        if (bw * bh <= kWaveRaster) {
          hair_rows(a, s, t, (long long)k, (long long)kRowSplit);
        } else if (k == 0u) {
          const uint32_t bands = (uint32_t)((bh + kBandRows - 1) / kBandRows);
          const uint32_t first = atomicAdd(a.bigCount, bands);
          if (first + bands <= a.bigCap) {
            for (uint32_t q = 0; q < bands; q++) a.big[first + q] = make_uint2(t, q);
          } else {
            // a reservation that crosses the cap leaves its slots below the
            // cap to a sentinel triangle id (skipped by HairRasterBig, which
            // reads every slot below min(count, cap)); this wave draws it
            for (uint32_t q = first; q < a.bigCap; q++) a.big[q] = make_uint2(kBigSentinel, 0u);
            bigMask |= 1u << h;
          }
        }
      }
    }
    // queue full: the wave draws the triangle, one row per lane
    const uint32_t lane = __lane_id();
    for (uint32_t h = 0; h < 2u; h++) {
      uint64_t todo = __ballot((bigMask >> h) & 1u);
      while (todo) {
        const uint32_t src = (uint32_t)__builtin_ctzll(todo);
        todo &= todo - 1ull;
        const uint32_t tb = 2u * (uint32_t)__builtin_amdgcn_readlane((int)seg, (int)src) + h;
        const SetupTri sb = setup_tri(a, tb);
        hair_rows(a, sb, tb, (long long)lane, 64);
      }
    }
  }
};

// The queued bands: a workgroup per band, its threads taking the band's
// (row, 64-pixel chunk) items in turn; every workgroup exits after the last
// band.
struct HairRasterBig {
  using Args = RasterArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "hair_raster_big";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t count = min(*a.bigCount, a.bigCap);
    for (uint32_t i = block; i < count; i += blocks) {
      const uint2 e = a.big[i];
      if (e.x >= a.nTri) continue;  // kBigSentinel
      const SetupTri s = setup_tri(a, e.x);
      const PixBox b = pix_box(a.W, a.H, s);
      const long long y0 = b.y0 + (long long)e.y * kBandRows, y1 = min(b.y1, y0 + kBandRows - 1);
      const long long cpr = (b.x1 - b.x0) / 64 + 1;
      const long long items = (y1 - y0 + 1) * cpr;
      for (long long it = threadIdx.x; it < items; it += kThreads) {
        const long long row = it / cpr, chunk = it - row * cpr;
        const long long xa = b.x0 + chunk * 64;
        raster_row(s, e.x, y0 + row, xa, min(b.x1, xa + 63), a.W, a.depth, a.key);
      }
    }
  }
};

// The winner of each pixel: its depth, and (camera pass) its normal — the
// world-space tangent modelMatrix * vec4(tangent, 1) interpolated with
// perspective correction, normalized and octahedron-encoded (hwHairPass.wgsl.ts:65, :95-96).
struct HairResolve {
  using Args = RasterArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "hair_resolve";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t p = block * kThreads + threadIdx.x;
    if (p >= a.W * a.H) return;
    const unsigned long long k = a.key[p];
    if (k == kNoFragment) return;
    const uint32_t t = (uint32_t)k;
    a.depth[p] = __uint_as_float((uint32_t)(k >> 32));
    if (!a.normals) return;
    const SetupTri s = setup_tri(a, t);
    const uint32_t px = p % a.W, py = p / a.W;
    float l[3];
    barycentrics(s, (long long)px * 256 + 128, (long long)py * 256 + 128, l);
    uint32_t idx[3], side[3];
    tri_vertices(t, a.P, idx, side);
    f3 num = mk3(0.0f, 0.0f, 0.0f);
    float den = 0.0f;
    for (int q = 0; q < 3; q++) {
      const float4 tg = a.tan[idx[s.vi[q]]];
      const f4 tw = mul(a.u->modelMatrix, mk4(tg.x, tg.y, tg.z, 1.0f));
      const float wq = l[q] / s.c[q].w;
      num = num + xyz(tw) * wq;
      den = den + wq;
    }
    a.normals[p] = encode_oct(normalize(num / den));
  }
};

struct AoArgs {
  const fh_render_uniforms* u;
  const float* depth;
  const uint32_t* normals;
  float* out;
  uint32_t W, H, aoW, aoH;
};

__device__ __forceinline__ float depth_at(const AoArgs& a, f2 px) {
  // textureLoad(vec2u(px)): u32() saturates at 0, texels past the edge read 0
  const uint32_t x = f2u(px.x), y = f2u(px.y);
  return (x < a.W && y < a.H) ? a.depth[(size_t)y * a.W + x] : 0.0f;
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// The loop-invariant parts of reprojectPositionToVS (gtao.wgsl.ts:8-25):
// projMatrixInv applied to (2 px / viewport - 1, depth, 1), x and y divided
// by w, z replaced by linearizeDepth(depth).
struct Reproj {
  float m[16];
  float sx, sy;  // 2 / viewport
};

__device__ __forceinline__ f3 reproject_vs(const AoArgs& a, const Reproj& R, f2 px) {
  const float d = depth_at(a, px);
  const float x = px.x * R.sx - 1.0f, y = px.y * R.sy - 1.0f;
  const float* m = R.m;
  const float X = m[0] * x + m[4] * y + m[8] * d + m[12];
  const float Y = m[1] * x + m[5] * y + m[9] * d + m[13];
  const float Wc = m[3] * x + m[7] * y + m[11] * d + m[15];
  const float iw = fast_rcp(Wc);
  return mk3(X * iw, Y * iw, (kNear * kFar) * fast_rcp(kFar + (kNear - kFar) * d));
}

__device__ __forceinline__ float round_f(float v) { return v < 0.5f ? floorf(v) : ceilf(v); }  // :130-133

// gtao (gtao.wgsl.ts:26-128). The sample loop uses the hardware's reciprocal,
// square root and sin / cos (about 1 ulp; the reference's WGSL built-ins are
// specified no tighter).
__device__ __forceinline__ float gtao(const AoArgs& a, const Reproj& R, f2 positionPx, f3 normalWS) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const fh_ao& prm = u->ao;
  const f3 positionVS = reproject_vs(a, R, positionPx);
  const f4 nv = mul(u->viewMatrix, mk4(normalWS.x, normalWS.y, normalWS.z, 1.0f));
  const f3 normalVS = mk3(nv.x, nv.y, -nv.z);
  const f3 toCamera = normalize(mk3(-positionVS.x, -positionVS.y, -positionVS.z));
  if (positionVS.z >= 0.999f * kFar) return 0.0f;
  const float radius = fmaxf((float)prm.numSteps, prm.radius * (positionVS.z / kFar));
  const float stepSize = radius / (float)prm.numSteps;
  const float fallScale = fast_rcp(prm.falloffEnd2 - prm.falloffStart2);
  float ao = 0.0f;
  const float rotJitter = kPi * fractf(52.9829189f * fractf(positionPx.x * 0.06711056f + positionPx.y * 0.00583715f));
  const float jitterMod = (positionPx.x + positionPx.y) * 0.25f;
  const float radiusJitter = fractf(jitterMod) * stepSize * 0.25f;
  for (uint32_t i = 0; i < prm.numDirections; i++) {
    float phi = (float)i * (kPi / (float)prm.numDirections) + prm.directionOffset * kPi;
    phi += rotJitter;
    float currStep = 1.0f + 0.25f * stepSize;
    currStep += radiusJitter;
    const f3 dir = mk3(__cosf(phi), __sinf(phi), 0.0f);
    float hx = -1.0f, hy = -1.0f;
    for (uint32_t j = 0; j < prm.numSteps; j++) {
      const f2 off = mk2(round_f(dir.x * currStep), round_f(dir.y * currStep));
      f3 ws = reproject_vs(a, R, positionPx + off) - positionVS;
      float q = dot(ws, ws);
      float dist = fast_sqrt(q);
      float cosh = dot(ws, toCamera) * fast_rsqrt(dist);
      hx = fmaxf(hx, cosh - 2.0f * satf((dist - prm.falloffStart2) * fallScale));
      ws = reproject_vs(a, R, positionPx - off) - positionVS;
      q = dot(ws, ws);
      dist = fast_sqrt(q);
      cosh = dot(ws, toCamera) * fast_rsqrt(dist);
      hy = fmaxf(hy, cosh - 2.0f * satf((dist - prm.falloffStart2) * fallScale));
      currStep += stepSize;
    }
    hx = acosf(hx);
    hy = acosf(hy);
    const f3 bitangent = normalize(cross(dir, toCamera));
    const f3 tangent = cross(toCamera, bitangent);
    const f3 nx = normalVS - bitangent * dot(normalVS, bitangent);
    const float nnx = length(nx);
    const float nnxInv = 1.0f / (nnx + 1e-6f);
    const float cosxi = dot(nx, tangent) * nnxInv;
    const float gamma = acosf(cosxi) - kHalfPi;
    const float cosgamma = dot(nx, toCamera) * nnxInv;
    const float singamma2 = -2.0f * cosxi;
    hx = gamma + fmaxf(-hx - gamma, -kHalfPi);
    hy = gamma + fminf(hy - gamma, kHalfPi);
    ao += nnx * 0.25f * ((hx * singamma2 + cosgamma - __cosf(2.0f * hx - gamma)) +
                         (hy * singamma2 + cosgamma - __cosf(2.0f * hy - gamma)));
  }
  return ao / (float)prm.numDirections;
}

// AoPass fragment shader (aoPass.wgsl.ts:48-79) per AO texel, stored as
// r16float (rounded to half, widened back to f32)
struct Gtao {
  using Args = AoArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "gtao";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    const uint32_t i = xcd_order(block, blocks) * kThreads + threadIdx.x;
    if (i >= a.aoW * a.aoH) return;
    const uint32_t x = i % a.aoW, y = i / a.aoW;
    Reproj R;
#pragma unroll
    for (int k = 0; k < 16; k++) R.m[k] = a.u->projMatrixInv[k];
    R.sx = 2.0f / a.u->viewport[0];
    R.sy = 2.0f / a.u->viewport[1];
    const f2 full = mk2(((float)x + 0.5f) / 0.5f, ((float)y + 0.5f) / 0.5f);  // AO_TEXTURE_SCALE = 0.5
    const uint32_t fx = f2u(full.x), fy = f2u(full.y);
    const uint32_t n = (fx < a.W && fy < a.H) ? a.normals[(size_t)fy * a.W + fx] : 0u;
    const float v = gtao(a, R, full, decode_oct(n));
    a.out[i] = (float)(_Float16)v;
  }
};

int raster_hair(fh_ctx* ctx, RasterArgs& a) {
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
  FH_REQUIRE((uint64_t)a.nTri * kRowSplit < (1ull << 32), "hair raster: too many triangles");
  if ((rc = launch<HairRaster>(ctx, div_ceil(a.nTri / 2u * kRowSplit, HairRaster::kThreads), 0, a))) return rc;
  if ((rc = launch<HairRasterBig>(ctx, persistent_blocks(ctx), 0, a))) return rc;
  return launch<HairResolve>(ctx, div_ceil((uint32_t)px, HairResolve::kThreads), 0, a);
}

}  // namespace

extern "C" {

int fh_hw_hair(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
               const fh_vec4* positions, const fh_vec4* tangents, float* depth_texture, void* normals_texture,
               uint32_t rendered_strands, uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && positions && tangents && depth_texture, "fh_hw_hair: null buffer");
  FH_REQUIRE(points_per_strand >= 2, "fh_hw_hair: points_per_strand >= 2");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), "fh_hw_hair: viewport too large");
  const uint64_t nTri = 2ull * rendered_strands * (points_per_strand - 1u);
  FH_REQUIRE(nTri < (1ull << 31), "fh_hw_hair: too many triangles");
  if (nTri == 0 || vp_w == 0 || vp_h == 0) return FH_OK;
  RasterArgs a;
  a.u = uniforms;
  a.pos = reinterpret_cast<const float4*>(positions);
  a.tan = reinterpret_cast<const float4*>(tangents);
  a.depth = depth_texture;
  a.normals = static_cast<uint32_t*>(normals_texture);
  a.W = vp_w; a.H = vp_h; a.P = points_per_strand; a.nTri = (uint32_t)nTri; a.shadow = 0;
  return raster_hair(ctx, a);
}

int fh_shadow_map_hair(fh_ctx* ctx, const fh_render_uniforms* uniforms, const fh_vec4* positions,
                       const fh_vec4* tangents, float* shadow_map, uint32_t shadow_size, uint32_t rendered_strands,
                       uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && positions && tangents && shadow_map, "fh_shadow_map_hair: null buffer");
  FH_REQUIRE(points_per_strand >= 2, "fh_shadow_map_hair: points_per_strand >= 2");
  FH_REQUIRE(shadow_size > 0 && shadow_size <= 16384, "fh_shadow_map_hair: shadow_size in [1, 16384]");
  const uint64_t nTri = 2ull * rendered_strands * (points_per_strand - 1u);
  FH_REQUIRE(nTri < (1ull << 31), "fh_shadow_map_hair: too many triangles");
  if (nTri == 0) return FH_OK;
  RasterArgs a;
  a.u = uniforms;
  a.pos = reinterpret_cast<const float4*>(positions);
  a.tan = reinterpret_cast<const float4*>(tangents);
  a.depth = shadow_map;
  a.normals = nullptr;
  a.W = shadow_size; a.H = shadow_size; a.P = points_per_strand; a.nTri = (uint32_t)nTri; a.shadow = 1;
  return raster_hair(ctx, a);
}

int fh_calc_ao(fh_ctx* ctx, const fh_render_uniforms* uniforms, fh_tex2d depth, const void* normals_texture,
          float* ao_texture, uint32_t ao_w, uint32_t ao_h) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && depth.data && normals_texture && ao_texture, "fh_calc_ao: null buffer");
  FH_REQUIRE((uint64_t)ao_w * ao_h < (1ull << 31), "fh_calc_ao: AO texture too large");
  if (ao_w == 0 || ao_h == 0) return FH_OK;
  AoArgs a{uniforms, depth.data, static_cast<const uint32_t*>(normals_texture), ao_texture, depth.width,
           depth.height, ao_w, ao_h};
  return launch<Gtao>(ctx, div_ceil(ao_w * ao_h, Gtao::kThreads), 0, a);
}

}  // extern "C"
