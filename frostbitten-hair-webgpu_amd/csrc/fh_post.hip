// Combine and present (gfx950): the hair rasterizer result blended over the
// HDR scene texture, then dither, exposure, ACES tone mapping and gamma into
// an 8-bit sRGB screen texture. The reference draws both as fullscreen
// triangles; here each is a compute kernel with one thread per texel.
// Reference: src/passes/hairCombine/hairCombinePass.wgsl.ts:51-162 (+ the
// pipeline's blend state, hairCombinePass.ts:25-40), src/passes/presentPass/
// presentPass.wgsl.ts:54-103, _shaderSnippets/dither.wgsl.ts:12-49,
// _shaderSnippets/aces.wgsl.ts:5-13.
//
// Both kernels are bound by HBM bandwidth: combine moves 16 B (rasterizer
// pixel) + 8 B + 8 B (HDR texel read and written) per texel, present 8 B in
// and 4 B out.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace {

using namespace fh;

// DISPLAY_MODE (constants.ts:44-58)
constexpr uint32_t kModeFinal = 0, kModeTiles = 1, kModeTilesPpll = 2, kModeHwRender = 3, kModeDepth = 5,
                   kModeNormals = 6, kModeAo = 7;
// linearizeDepth constants: CONFIG.camera.projection near / far (constants.ts:131-135)
constexpr float kNear = 0.01f, kFar = 100.0f;

// The 8x8 ordered-dither (Bayer) matrix of dither.wgsl.ts:12-21, row-major.
// Entry (x, y) is the bit reversal of the 6-bit interleave of (x ^ y, y):
// the textbook recursive Bayer construction.
__constant__ uint8_t kBayer8[64] = {
    0,  32, 8,  40, 2,  34, 10, 42, 48, 16, 56, 24, 50, 18, 58, 26, 12, 44, 4,  36, 14, 46,
    6,  38, 60, 28, 52, 20, 62, 30, 54, 22, 3,  35, 11, 43, 1,  33, 9,  41, 51, 19, 59, 27,
    49, 17, 57, 25, 15, 47, 7,  39, 13, 45, 5,  37, 63, 31, 55, 23, 61, 29, 53, 21};

__device__ __forceinline__ float h2f(uint32_t bits16) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}
__device__ __forceinline__ uint32_t f2h(float x) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

struct CombineArgs {
  const fh_render_uniforms* u;
  const fh_tile_result* tiles;
  const fh_list_element* segs;  // element array (byte 4 of DrawnHairSegments)
  const float4* pix;            // pixel array (byte 16 of the rasterizer result)
  const uint32_t* countPerTile;
  const uint2* scene;           // the texels blended onto (hdr itself, or a scene texture of its layout)
  uint2* hdr;                   // rgba16float texels, texture memory order
  uint32_t W, H, tile, bins, tilesX;
};

// One HDR texel (x, ty), ty counted from the top row as the texture is
// stored. The reference's fragment at pixel centre (x + .5, ty + .5) reads
// rasterizer row u32(H - (ty + .5)) = H - 1 - ty (hairCombinePass.wgsl.ts:60-63).
// Returns the texel as written to hdr.
__device__ __forceinline__ uint2 combine_texel(const CombineArgs& a, uint32_t x, uint32_t ty) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const uint32_t py = a.H - 1u - ty;
  const uint32_t tx = x / a.tile, tyl = py / a.tile;
  const uint32_t dm = u->displayMode;
  const uint32_t mode = dm & 0xffu, extra = (dm >> 8) & 0xffffu;
  const size_t at = (size_t)ty * a.W + x;
  const uint2 d = a.scene[at];
  f4 c;
  if (mode == kModeTiles || mode == kModeTilesPpll) {
    // renderTileSegmentCount :100-128
    const uint32_t maxSeg = extra;
    uint32_t segments = 0;
    if (mode == kModeTiles) {
      segments = a.countPerTile[tyl * a.tilesX + tx];  // getSegmentCountInTiles_Count :155-162
    } else {
      // getSegmentCountInTiles_PPLL :130-153: list lengths summed over the
      // bins, at most maxSeg in all
      const uint32_t maxDrawn = u->maxDrawnHairSegments;
      for (uint32_t b = 0; b < a.bins; b++) {
        const uint32_t head = a.tiles[(tyl * a.tilesX + tx) * a.bins + b].tileSegmentPtr;
        uint32_t p = head == 0u ? 0xffffffffu : head - 1u;
        while (segments < maxSeg) {
          if (p >= maxDrawn || p == 0xffffffffu) break;
          segments++;
          p = a.segs[p].next;
        }
      }
    }
    if (segments == 0u) {  // discard: the texel keeps the scene's value
      if (a.scene != a.hdr) a.hdr[at] = d;
      return d;
    }
    const float r = (float)segments / (float)maxSeg;
    c = mk4(r, 1.0f - r, 0.0f, 1.0f);
  } else {
    const uint32_t i = a.W * py + x;  // _getRasterizerResult; the array holds W*H - 1 pixels
    if (i < a.W * a.H - 1u) {
      const float4 v = a.pix[i];
      c = mk4(v.x, v.y, v.z, v.w);
    } else {
      c = mk4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (c.w <= 0.0f && mode == kModeFinal && (extra & 1u)) {
      // getDebugTileColor :92-98, then dbg + color * color.a
      const f4 d = mk4(0.0f, (float)(tx % 2u) / 2.0f, (float)(tyl % 2u) / 2.0f, 1.0f);
      c = mk4(d.x + c.x * c.w, d.y + c.y * c.w, d.z + c.z * c.w, d.w + c.w * c.w);
    }
  }
  // blend state: colour src * srcAlpha + dst * (1 - srcAlpha), alpha src + dst
  const float dr = h2f(d.x & 0xffffu), dg = h2f(d.x >> 16), db = h2f(d.y & 0xffffu), da = h2f(d.y >> 16);
  const float ia = 1.0f - c.w;
  const float r = c.x * c.w + dr * ia, g = c.y * c.w + dg * ia, b = c.z * c.w + db * ia, al = c.w + da;
  const uint2 o = make_uint2(f2h(r) | (f2h(g) << 16), f2h(b) | (f2h(al) << 16));
  a.hdr[at] = o;
  return o;
}

struct Combine {
  using Args = CombineArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "hair_combine";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t i = block * kThreads + threadIdx.x;
    if (i >= a.W * a.H) return;
    const uint32_t ty = i / a.W;
    (void)combine_texel(a, i - ty * a.W, ty);
  }
};

struct PresentArgs {
  const fh_render_uniforms* u;
  const uint2* hdr;            // rgba16float
  const float* depth;          // depth texture, W x H
  const float* ao;
  uint32_t aoW, aoH;
  const uint32_t* normals;     // rg16float, W x H (may be null)
  uint8_t* out;                // rgba8unorm-srgb rows of outPitch bytes
  uint32_t W, H, outPitch, ditherPerPixel;
  // srgbT[k] (k = 1..255): the least f32 linear value that an
  // rgba8unorm-srgb store rounds to code k or above (srgbT[0] = 0); device
  // memory of the context (an array in the arguments would be copied per
  // thread by a batched launch)
  const float* srgbT;
};

// doACES_Tonemapping (aces.wgsl.ts:5-13)
__device__ __forceinline__ float aces(float x) {
  return satf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f));
}

// pow(x, y) for x >= 0 as WGSL defines it: exp2(y * log2(x)), here with the
// hardware's base-2 logarithm and exponential (about 1 ulp each).
__device__ __forceinline__ float pow_pos(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// Linear value -> 8-bit sRGB code, as an rgba8unorm-srgb attachment stores it
// (IEC 61966-2-1 encoding, then round to nearest; NaN stores 0). The code of
// the hardware-math encoding is within one of the exact one; the thresholds
// then make it exact: the result is the number of thresholds <= c.
__device__ __forceinline__ uint32_t srgb8(float c, const float* __restrict__ T) {
  c = satf(c);
  const float s = c <= 0.0031308f ? 12.92f * c : 1.055f * pow_pos(c, 1.0f / 2.4f) - 0.055f;
  int32_t k = (int32_t)(s * 255.0f + 0.5f);
  k = k < 0 ? 0 : (k > 255 ? 255 : k);
  if (k > 0 && c < T[k]) k--;
  else if (k < 255 && c >= T[k + 1]) k++;
  return (uint32_t)k;
}

// One screen texel from its HDR texel h (texel (x, y) of hdr).
__device__ __forceinline__ void present_texel(const PresentArgs& a, uint32_t x, uint32_t y, uint2 h) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const float gamma = u->colorMgmt[0], exposure = u->colorMgmt[1], ditherStr = u->colorMgmt[2];
  f3 c = mk3(h2f(h.x & 0xffffu), h2f(h.x >> 16), h2f(h.y & 0xffffu));
  const uint32_t mode = u->displayMode & 0xffu;
  if (mode == kModeFinal || mode == kModeHwRender) {
    // ditherColor (dither.wgsl.ts:23-49); the Deno build reads element 0
    const uint32_t idx = a.ditherPerPixel ? (y % 8u) * 8u + (x % 8u) : 0u;
    const float dm = (float)kBayer8[idx] / 63.0f * ditherStr / 256.0f;
    c = mk3(c.x + dm, c.y + dm, c.z + dm);
    c = c * exposure;
    c = mk3(aces(c.x), aces(c.y), aces(c.z));  // saturate(ACES) is ACES: it saturates already
    const float e = 1.0f / gamma;
    c = mk3(pow_pos(c.x, e), pow_pos(c.y, e), pow_pos(c.z, e));
  } else if (mode == kModeDepth) {
    const float d = a.depth[(size_t)y * a.W + x];
    const float lin = kNear * kFar / (kFar + (kNear - kFar) * d) / (kFar - kNear);  // linearizeDepth_0_1
    const float t = satf((lin - 0.002f) / (0.01f - 0.002f));                         // mapRange :158-165
    const float v = mixf(0.0f, 1.0f, t);
    c = mk3(v, v, v);
  } else if (mode == kModeNormals) {
    float fx = 0.0f, fy = 0.0f;
    if (a.normals) {
      const uint32_t n = a.normals[(size_t)y * a.W + x];
      fx = h2f(n & 0xffffu);
      fy = h2f(n >> 16);
    }
    // decodeOctahedronNormal (shaderSnippets.wgls.ts:49-58)
    f3 n = mk3(fx * 2.0f - 1.0f, fy * 2.0f - 1.0f, 0.0f);
    n.z = 1.0f - fabsf(n.x) - fabsf(n.y);
    const float t = satf(-n.z);
    n.x = n.x >= 0.0f ? n.x - t : n.x + t;
    n.y = n.y >= 0.0f ? n.y - t : n.y + t;
    n = normalize(n);
    c = mk3(fabsf(n.x), fabsf(n.y), fabsf(n.z));
  } else if (mode == kModeAo) {
    // sampleAo (aoPass/shared/textureAo.wgsl.ts:7-12) at the pixel centre
    const float tx = ((float)x + 0.5f) / u->viewport[0], ty = ((float)y + 0.5f) / u->viewport[1];
    const float v = tex_load(a.ao, a.aoW, a.aoH, f2i((float)a.aoW * tx), f2i((float)a.aoH * ty));
    c = mk3(v, v, v);
  }
  const uint32_t rgba = srgb8(c.x, a.srgbT) | (srgb8(c.y, a.srgbT) << 8) | (srgb8(c.z, a.srgbT) << 16) | (255u << 24);
  *reinterpret_cast<uint32_t*>(a.out + (size_t)y * a.outPitch + 4u * x) = rgba;
}

struct Present {
  using Args = PresentArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "present";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t i = block * kThreads + threadIdx.x;
    if (i >= a.W * a.H) return;
    const uint32_t y = i / a.W;
    present_texel(a, i - y * a.W, y, a.hdr[i]);
  }
};

// The combine and present passes in one pass over the texels: each thread
// blends its texel, writes it to hdr and presents the value it wrote.
struct CombinePresent {
  struct Args {
    CombineArgs c;
    PresentArgs p;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "combine_present";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t i = block * kThreads + threadIdx.x;
    if (i >= a.c.W * a.c.H) return;
    const uint32_t y = i / a.c.W, x = i - y * a.c.W;
    present_texel(a.p, x, y, combine_texel(a.c, x, y));
  }
};

// Device-to-device copy as a kernel (a graph then holds no memcpy node), in
// 16-byte words when both ends allow it.
template <class W>
struct CopyWords {
  struct Args {
    W* dst;
    const W* src;
    uint64_t n;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "copy_u32";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    uint64_t i = (uint64_t)block * kThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)blocks * kThreads;
    for (; i < a.n; i += stride) a.dst[i] = a.src[i];
  }
};

// dst <- record (*counter mod count) of src, then ++*counter (one workgroup).
struct CopyIndexed {
  struct Args {
    uint32_t* dst;
    const uint32_t* src;
    uint32_t words, count;
    uint32_t* counter;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "copy_indexed";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)block;
    (void)blocks;
    const uint32_t c = *a.counter;
    const uint64_t base = (uint64_t)(c % a.count) * a.words;
    for (uint32_t i = threadIdx.x; i < a.words; i += kThreads) a.dst[i] = a.src[base + i];
    __syncthreads();  // every thread has read the counter
    if (threadIdx.x == 0) *a.counter = c + 1u;
  }
};

}  // namespace

namespace fh {

// Code of a linear value under the exact encoding (evaluated in double).
static uint32_t srgb_code_exact(float c) {
  double v = c > 0.0f ? (c < 1.0f ? (double)c : 1.0) : 0.0;
  double s = v <= 0.0031308 ? 12.92 * v : 1.055 * std::pow(v, 1.0 / 2.4) - 0.055;
  return (uint32_t)std::floor(s * 255.0 + 0.5);
}

// srgbT of PresentArgs: for each code k the least float with code >= k,
// found from the inverse transfer function and refined one ulp at a time.
const float* srgb_thresholds() {
  static const std::array<float, 256> table = [] {
    std::array<float, 256> t{};
    for (uint32_t k = 1; k < 256; k++) {
      const double sk = ((double)k - 0.5) / 255.0;
      double lin = sk <= 0.04045 ? sk / 12.92 : std::pow((sk + 0.055) / 1.055, 2.4);
      float c = (float)lin;
      while (c > 0.0f && srgb_code_exact(std::nextafter(c, 0.0f)) >= k) c = std::nextafter(c, 0.0f);
      while (srgb_code_exact(c) < k) c = std::nextafter(c, 2.0f);
      t[k] = c;
    }
    return t;
  }();
  return table.data();
}

}  // namespace fh

extern "C" {

int fh_copy_indexed(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t count, uint32_t* counter) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(dst && src && counter, "fh_copy_indexed: null buffer");
  FH_REQUIRE(bytes % 4 == 0 && bytes > 0 && bytes / 4 < (1ull << 32), "fh_copy_indexed: bytes must be a multiple of 4");
  FH_REQUIRE(count > 0, "fh_copy_indexed: count must be positive");
  return fh::launch<CopyIndexed>(ctx, 1, 0, CopyIndexed::Args{static_cast<uint32_t*>(dst),
                                                               static_cast<const uint32_t*>(src),
                                                               (uint32_t)(bytes / 4), count, counter});
}

int fh_copy(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(bytes % 4 == 0, "fh_copy: bytes must be a multiple of 4");
  if (bytes == 0) return FH_OK;
  FH_REQUIRE(dst && src, "fh_copy: null buffer");
  if (fh::batch_fill(ctx, dst, src, bytes / 4, 0u)) return FH_OK;
  if (bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0) {
    using K = CopyWords<uint4>;
    const uint64_t n = bytes / 16;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    return fh::launch<K>(ctx, blocks, 0, K::Args{(uint4*)dst, (const uint4*)src, n});
  }
  using K = CopyWords<uint32_t>;
  const uint64_t n = bytes / 4;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  return fh::launch<K>(ctx, blocks, 0, K::Args{(uint32_t*)dst, (const uint32_t*)src, n});
}

}  // extern "C"

namespace {

int combine_args(const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                 const fh_tile_result* tiles, const void* tile_segments, const void* rasterizer_result,
                 const uint32_t* segment_count_per_tile, const void* scene, void* hdr_texture, const char* what,
                 CombineArgs& a) {
  FH_REQUIRE(cfg && uniforms && tiles && tile_segments && rasterizer_result && segment_count_per_tile && scene &&
                 hdr_texture,
             std::string(what) + ": null buffer");
  FH_REQUIRE(cfg->tileSize > 0 && cfg->tileDepthBins > 0, std::string(what) + ": tileSize and tileDepthBins > 0");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), std::string(what) + ": viewport too large");
  a.u = uniforms;
  a.tiles = tiles;
  a.segs = reinterpret_cast<const fh_list_element*>(static_cast<const char*>(tile_segments) + 4);
  a.pix = reinterpret_cast<const float4*>(static_cast<const char*>(rasterizer_result) + 16);
  a.countPerTile = segment_count_per_tile;
  a.scene = static_cast<const uint2*>(scene);
  a.hdr = static_cast<uint2*>(hdr_texture);
  a.W = vp_w; a.H = vp_h; a.tile = cfg->tileSize; a.bins = cfg->tileDepthBins;
  a.tilesX = fh::div_ceil(vp_w, cfg->tileSize);
  return FH_OK;
}

int present_args(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
                 const void* hdr_texture, fh_tex2d depth, fh_tex2d ao, const void* normals_texture,
                 uint32_t dither_per_pixel, void* screen_texture, uint32_t screen_row_bytes, const char* what,
                 PresentArgs& a) {
  const std::string w(what);
  FH_REQUIRE(uniforms && hdr_texture && screen_texture && depth.data && ao.data, w + ": null buffer");
  FH_REQUIRE((uint64_t)vp_w * vp_h < (1ull << 31), w + ": viewport too large");
  if (vp_w == 0 || vp_h == 0) return FH_OK;  // nothing to draw (the caller launches nothing)
  const uint32_t pitch = screen_row_bytes ? screen_row_bytes : 4u * vp_w;
  FH_REQUIRE(pitch >= 4u * vp_w && pitch % 4u == 0, w + ": screen_row_bytes must be >= 4 * width, a multiple of 4");
  FH_REQUIRE(depth.width == vp_w && depth.height == vp_h, w + ": depth texture must match the viewport");
  FH_REQUIRE(ao.width > 0 && ao.height > 0, w + ": empty AO texture");
  a.u = uniforms;
  a.hdr = static_cast<const uint2*>(hdr_texture);
  a.depth = depth.data;
  a.ao = ao.data; a.aoW = ao.width; a.aoH = ao.height;
  a.normals = static_cast<const uint32_t*>(normals_texture);
  a.out = static_cast<uint8_t*>(screen_texture);
  a.W = vp_w; a.H = vp_h; a.outPitch = pitch; a.ditherPerPixel = dither_per_pixel ? 1u : 0u;
  a.srgbT = ctx->srgb_thresholds;
  return FH_OK;
}

}  // namespace

extern "C" {

int fh_hair_combine(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                    const fh_render_uniforms* uniforms, const fh_tile_result* tiles, const void* tile_segments,
                    const void* rasterizer_result, const uint32_t* segment_count_per_tile, void* hdr_texture) {
  FH_CHECK_CTX(ctx);
  CombineArgs a;
  int rc = combine_args(cfg, vp_w, vp_h, uniforms, tiles, tile_segments, rasterizer_result, segment_count_per_tile,
                        hdr_texture, hdr_texture, "fh_hair_combine", a);
  if (rc || vp_w == 0 || vp_h == 0) return rc;
  return fh::launch<Combine>(ctx, fh::div_ceil(vp_w * vp_h, Combine::kThreads), 0, a);
}

int fh_present(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, const fh_render_uniforms* uniforms,
               const void* hdr_texture, fh_tex2d depth, fh_tex2d ao, const void* normals_texture,
               uint32_t dither_per_pixel, void* screen_texture, uint32_t screen_row_bytes) {
  FH_CHECK_CTX(ctx);
  PresentArgs a;
  int rc = present_args(ctx, vp_w, vp_h, uniforms, hdr_texture, depth, ao, normals_texture, dither_per_pixel,
                        screen_texture, screen_row_bytes, "fh_present", a);
  if (rc || vp_w == 0 || vp_h == 0) return rc;
  return fh::launch<Present>(ctx, fh::div_ceil(vp_w * vp_h, Present::kThreads), 0, a);
}

int fh_combine_present(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                       const fh_render_uniforms* uniforms, const fh_tile_result* tiles, const void* tile_segments,
                       const void* rasterizer_result, const uint32_t* segment_count_per_tile,
                       const void* scene_texture, void* hdr_texture, fh_tex2d depth, fh_tex2d ao,
                       const void* normals_texture, uint32_t dither_per_pixel, void* screen_texture,
                       uint32_t screen_row_bytes) {
  FH_CHECK_CTX(ctx);
  CombinePresent::Args a;
  const void* scene = scene_texture ? scene_texture : hdr_texture;
  int rc = combine_args(cfg, vp_w, vp_h, uniforms, tiles, tile_segments, rasterizer_result, segment_count_per_tile,
                        scene, hdr_texture, "fh_combine_present", a.c);
  if (rc) return rc;
  rc = present_args(ctx, vp_w, vp_h, uniforms, hdr_texture, depth, ao, normals_texture, dither_per_pixel,
                    screen_texture, screen_row_bytes, "fh_combine_present", a.p);
  if (rc || vp_w == 0 || vp_h == 0) return rc;
  if (scene != hdr_texture) {
    // the scene is read where hdr is written: the two must not overlap
    const char* s0 = static_cast<const char*>(scene);
    const char* h0 = static_cast<const char*>(hdr_texture);
    const size_t bytes = (size_t)vp_w * vp_h * 8;
    FH_REQUIRE(s0 + bytes <= h0 || h0 + bytes <= s0, "fh_combine_present: scene_texture overlaps hdr_texture");
  }
  return fh::launch<CombinePresent>(ctx, fh::div_ceil(vp_w * vp_h, CombinePresent::kThreads), 0, a);
}

}  // extern "C"
