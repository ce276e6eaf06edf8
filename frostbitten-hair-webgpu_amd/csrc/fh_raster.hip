// Software hair rasterizer on gfx950: tile pass (K4), tile sort (K5/K6) and
// the fine OIT pass (K7), plus the raster-helper test kernel of the reference.
// Reference: src/passes/swHair/*.
#include "fh_launch.hpp"
#include "fh_raster_common.hpp"

namespace fh {
int fill_u32(fh_ctx* ctx, void* dst, uint64_t n_words, uint32_t value);
}

namespace {

using namespace fh;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

struct TileGeom {
  uint32_t vpW, vpH, tilesX, tilesY;
};

__device__ __forceinline__ TileGeom tile_geom(f2 vp, uint32_t T) {
  TileGeom g;
  g.vpW = f2u(vp.x);
  g.vpH = f2u(vp.y);
  g.tilesX = (g.vpW + T - 1u) / T;
  g.tilesY = (g.vpH + T - 1u) / T;
  return g;
}

// ---------------------------------------------------------------------------
// K4 — hairTilesPass.wgsl.ts:70-246, as one wavefront per 64 segments.
//
// The reference gives each thread one segment and loops over every pixel of
// every tile of its rectangle. With thread-private loops a wave runs as long
// as its longest segment, and the costly per-pixel branch (interpolation,
// depth test, view-space bin) executes whenever any lane takes it. Here the
// wave shares the work instead:
//   1. each lane projects its segment (swRasterizeHair.wgsl.ts:39-68) and
//      finds its tile rectangle (:94-117);
//   2. the wave flattens its (segment, tile) pairs and takes them 64 at a
//      time; every pair owns the pixels of its tile (getTileBoundsPx) that
//      can pass the edge tests (the segment's box, see seg_pixel_bbox);
//   3. row by row, a conservative x span (span_x) bounds the candidate
//      pixels; the wave enumerates them 64 at a
//      time, runs the four edge tests and queues the inside pixels in LDS;
//   4. every 64 queued pixels the whole wave interpolates depth, runs the
//      depth test and folds (min depth, max depth, min bin) into its pair
//      with LDS atomics (min/max are exact, so the order does not matter);
//   5. each pair with a hit is appended to the tile lists as in processTile
//      :205-238, with one drawnSegmentsCount atomic per 64 pairs.
// ---------------------------------------------------------------------------
struct TilesArgs {
  const fh_render_uniforms* u;
  const fh_hair_data* hd;
  const float4* pos;
  const float4* tan;
  fh_tile_result* tiles;
  const float* depth;
  uint32_t depthW, depthH;
  uint32_t* segCounter;          // DrawnHairSegments.drawnSegmentsCount
  fh_list_element* segData;      // DrawnHairSegments.data
  uint32_t limit;                // min(maxDrawnHairSegments, capacity)
  uint32_t* slotKey;
  uint32_t T, B, threshold;
  uint32_t S, P;
  uint2* pieces;           // overflow queue: (segment, first pair) of wide segments
  uint32_t* pieceCount;
  uint32_t pieceCap;
  const float2* depthBounds;  // per tile: (min, max) of the mesh depth its pixels read
  uint32_t nBounds;
};

// Per tile, the (min, max) of every depth texel that a pixel of the tile reads
// (tex_load at (x, vpH - y), 0 outside the texture), over the whole tile
// rectangle. A hair depth <= min passes the depth test and one > max fails it
// at every pixel of the tile, without a load; others load their texel. A NaN
// texel makes the bounds (-inf, +inf), so that every pixel loads.
struct TileDepthBounds {
  struct Args {
    const fh_render_uniforms* u;
    const float* depth;
    uint32_t depthW, depthH, T, nTiles;
    float2* bounds;
    uint32_t* zero;  // the tile pass's piece counter (set to 0 here)
  };
  // one wave per tile: every texel of the tile is loaded at once (T * T / 64
  // loads per lane, none waiting on another) and reduced across the wave
  static constexpr uint32_t kThreads = 256;
  static constexpr uint32_t kWaves = kThreads / 64;
  static constexpr const char* kName = "tile_depth_bounds";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    if (block == 0 && threadIdx.x == 0 && a.zero) *a.zero = 0u;
    const TileGeom tg = tile_geom(mk2(a.u->viewport[0], a.u->viewport[1]), a.T);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t tilesX = max(tg.tilesX, 1u), nt = a.T * a.T;
    for (uint32_t t = block * kWaves + w; t < a.nTiles; t += blocks * kWaves) {
      const uint32_t ty = t / tilesX, tx = t - ty * tilesX;
      float lo = INFINITY, hi = -INFINITY;
      bool nan = false;
      for (uint32_t i = lane; i < nt; i += 64u) {
        const uint32_t x = tx * a.T + i % a.T, y = ty * a.T + i / a.T;
        const float d = tex_load(a.depth, a.depthW, a.depthH, (int32_t)x, (int32_t)(tg.vpH - y));
        nan |= d != d;
        lo = fminf(lo, d);
        hi = fmaxf(hi, d);
      }
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
      }
      const bool anyNan = __ballot(nan) != 0ull;
      if (lane == 0) a.bounds[t] = anyNan ? make_float2(-INFINITY, INFINITY) : make_float2(lo, hi);
    }
  }
};

// A segment whose rectangle has more than kTilePiece tiles is not drawn by its
// own wave: it is queued in pieces of kTilePiece tiles that a second kernel
// takes 64 at a time, so one wide segment spreads over many waves.
constexpr uint32_t kTilePiece = 8;

// order-preserving map of f32 to u32 (for LDS min/max of non-NaN floats)
__device__ __forceinline__ uint32_t f32_ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_f32(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

constexpr uint32_t kTileQueue = 128;  // ring of queued inside pixels (power of two, >= 127)

// Conservative x ranges of a quad's inside pixels. The quad's inside pixels
// lie in its convex hull (seg_pixel_bbox), and the hull lies in the band
// |n . (p - c)| <= R around the axis through the two cap midpoints (n normal
// to the axis a, c its centre) and in the slab T0 <= a . (p - c) <= T1 along
// it, with R, T0, T1 taken over the four vertices. Both are widened by the
// quad's margin (raster_margin), which covers the f32 rounding of the edge
// tests and of these bounds. On a range of rows each constraint is an x
// interval with ends linear in y.
enum SpanField : uint32_t {
  kSpCx = 0, kSpCy, kSpNy, kSpR, kSpINx, kSpAy, kSpT0, kSpT1, kSpIAx, kSpOn, kSpFields
};

// The span frame of a quad; on = 0 when the bound is not usable (degenerate
// axis, collinear or far-off vertices, NaN): the caller keeps the whole box.
__device__ __forceinline__ void span_frame(const Seg& s, bool usable, float margin, float* f, uint32_t stride) {
  const f2 m0 = mk2((s.v00.x + s.v01.x) * 0.5f, (s.v00.y + s.v01.y) * 0.5f);
  const f2 m1 = mk2((s.v10.x + s.v11.x) * 0.5f, (s.v10.y + s.v11.y) * 0.5f);
  const f2 ax = m1 - m0;
  const f2 c = mk2((m0.x + m1.x) * 0.5f, (m0.y + m1.y) * 0.5f);
  const f2 n = mk2(-ax.y, ax.x);
  const float la = sqrtf(ax.x * ax.x + ax.y * ax.y);
  const f2 vs[4] = {s.v00, s.v01, s.v10, s.v11};
  float R = 0.0f, T0 = 3.0e38f, T1 = -3.0e38f;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const f2 q = vs[k] - c;
    R = fmaxf(R, fabsf(n.x * q.x + n.y * q.y));
    const float t = ax.x * q.x + ax.y * q.y;
    T0 = fminf(T0, t);
    T1 = fmaxf(T1, t);
  }
  const bool on = usable && la > 0.0f && isfinite(la) && isfinite(R) && isfinite(T0) && isfinite(T1);
  f[kSpCx * stride] = c.x;
  f[kSpCy * stride] = c.y;
  f[kSpNy * stride] = n.y;
  const float w = la * margin;  // n and a have length la
  f[kSpR * stride] = R + w;
  f[kSpINx * stride] = fabsf(n.x) > 1e-6f * la ? 1.0f / n.x : 0.0f;
  f[kSpAy * stride] = ax.y;
  f[kSpT0 * stride] = T0 - w;
  f[kSpT1 * stride] = T1 + w;
  f[kSpIAx * stride] = fabsf(ax.x) > 1e-6f * la ? 1.0f / ax.x : 0.0f;
  f[kSpOn * stride] = on ? 1.0f : 0.0f;
}

// x range (absolute) that may hold inside pixels on rows y in [y0, y1];
// empty when lo > hi. A constraint whose x coefficient vanishes bounds y only.
template <uint32_t STRIDE>
__device__ __forceinline__ void span_x(const float* f, float y0, float y1, float& lo, float& hi) {
  lo = -3.0e38f;
  hi = 3.0e38f;
  const float cx = f[kSpCx * STRIDE], cy = f[kSpCy * STRIDE];
  const float r0 = y0 - cy, r1 = y1 - cy;
  const float ny = f[kSpNy * STRIDE], R = f[kSpR * STRIDE], inx = f[kSpINx * STRIDE];
  if (inx != 0.0f) {
    const float a0 = (-R - ny * r0) * inx, b0 = (R - ny * r0) * inx;
    const float a1 = (-R - ny * r1) * inx, b1 = (R - ny * r1) * inx;
    lo = fmaxf(lo, fminf(fminf(a0, b0), fminf(a1, b1)));
    hi = fminf(hi, fmaxf(fmaxf(a0, b0), fmaxf(a1, b1)));
  } else if (fminf(ny * r0, ny * r1) > R || fmaxf(ny * r0, ny * r1) < -R) {
    lo = 1.0f; hi = 0.0f;
  }
  const float ay = f[kSpAy * STRIDE], T0 = f[kSpT0 * STRIDE], T1 = f[kSpT1 * STRIDE], iax = f[kSpIAx * STRIDE];
  if (iax != 0.0f) {
    const float a0 = (T0 - ay * r0) * iax, b0 = (T1 - ay * r0) * iax;
    const float a1 = (T0 - ay * r1) * iax, b1 = (T1 - ay * r1) * iax;
    lo = fmaxf(lo, fminf(fminf(a0, b0), fminf(a1, b1)));
    hi = fminf(hi, fmaxf(fmaxf(a0, b0), fmaxf(a1, b1)));
  } else if (fminf(ay * r0, ay * r1) > T1 || fmaxf(ay * r0, ay * r1) < T0) {
    lo = 1.0f; hi = 0.0f;
  }
  lo += cx;
  hi += cx;
}

// The view-space depth of a pixel's hair depth z, as the tile pass computes
// it (projectVertex of (x, y, z, 1) with projMatrixInv) but at x = y = 0:
// the clip z and w before the division.
__device__ __forceinline__ f2 clip_zw(const float* m, float z) {
  const f4 c = mul(m, mk4(0.0f, 0.0f, z, 1.0f));
  return mk2(c.z, c.w);
}

// Whether a segment's depth bin can be taken per (segment, tile) pair from
// the pair's smallest and largest hair depth instead of per pixel.
// getDepthBin (shaderSnippets.wgls.ts:148-156) is monotone non-decreasing in
// the view-space depth for any tile depth range (each of its operations is,
// including the infinite and NaN cases that saturate to 0 or n - 1), so if
// the view-space depth is monotone in the hair depth over every value a
// pixel of the segment can take, the closest bin over the pair's pixels
// (:201, :206) is the bin at one of the two extreme depths. It is when
//  - the x and y columns of projMatrixInv's z and w rows are exact zeros, so
//    the pixel position adds only signed zeros to clip z and w, which is
//    what clip_zw computes at x = y = 0;
//  - every hair depth lies in [lo, hi] (the segment's four depths widened by
//    the rounding of interp_f32's two mixes, which satf keeps convex), and
//    the depths are so small that no mix overflows;
//  - clip z and w are each a monotone function of the depth (one rounded
//    product and one rounded sum each); equal values at lo and hi make one
//    of them constant on [lo, hi], and w keeps one sign there, so the
//    correctly rounded quotient is monotone too.
// The per-pixel path handles every other case.
__device__ __forceinline__ bool pair_bin_ok(const float* m, f4 d) {
  if (!(m[2] == 0.0f && m[6] == 0.0f && m[3] == 0.0f && m[7] == 0.0f)) return false;
  const float dn = fminf(fminf(d.x, d.y), fminf(d.z, d.w));
  const float dx = fmaxf(fmaxf(d.x, d.y), fmaxf(d.z, d.w));
  // NaN fails both tests; |d| <= 998 also keeps every hair depth inside the
  // accumulators' +-999 sentinels, so the pair's min and max are pixel values
  if (!(dn >= -998.0f && dx <= 998.0f)) return false;
  const float r = 998.0f * 2e-6f + 1e-30f;  // two mixes' rounding, at most ~1e-6 of the magnitude
  const f2 c0 = clip_zw(m, dn - r), c1 = clip_zw(m, dx + r);
  if (!(isfinite(c0.x) && isfinite(c1.x))) return false;
  const bool oneSign = (c0.y > 0.0f && c1.y > 0.0f) || (c0.y < 0.0f && c1.y < 0.0f);
  return oneSign && (c0.x == c1.x || c0.y == c1.y);
}

struct TileWaveLds {
  // per segment (lane of the wave), field-major so that lanes reading
  // different segments hit different banks
  float v[8][64];          // v00, v01, v10, v11
  float d[4][64];          // projected depths
  float sp[kSpFields][64]; // span frame (span_frame)
  float ip[kIpFields][64]; // per-segment interpolation terms (interp_pre)
  int box[4][64];          // pixel box (seg_pixel_bbox)
  uint32_t pairBin[64];    // 1: the bin is decided per pair (pair_bin_ok), not per pixel
  uint32_t rect[3][64];    // first tile x, first tile y, tiles per row
  uint32_t gid[64];        // segment index
  uint32_t kBase[64];      // first pair (tile of the rectangle) of this lane
  uint32_t pairStart[64];
  int mark[64];            // lbs_owner scratch
  // per pair of the current window
  uint32_t owner[64];
  int px0[64], py0[64], px1[64];
  uint32_t rowStart[64];
  uint32_t accMin[64], accMax[64], accBin[64];
  // per pixel row of the current row window
  uint32_t itemSlot[64];
  int itemX[64], itemY[64];
  uint32_t candStart[64];
  uint32_t queue[kTileQueue];
};

// One step of a segmented inclusive (min, max, min) scan: combine with the
// lane the DPP control selects when it holds the same slot. Lanes without a
// source keep the identities.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_step(uint32_t slot, uint32_t& vmin, uint32_t& vmax, uint32_t& vbin) {
  const uint32_t os = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)slot, CTRL, ROWMASK, 0xf, false);
  const uint32_t omin = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)vmin, CTRL, ROWMASK, 0xf, false);
  const uint32_t omax = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)vmax, CTRL, ROWMASK, 0xf, false);
  const uint32_t obin = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)vbin, CTRL, ROWMASK, 0xf, false);
  if (os == slot) {
    vmin = min(vmin, omin);
    vmax = max(vmax, omax);
    vbin = min(vbin, obin);
  }
}

// Step 4: the queued inside pixels [head, head + n), n <= 64, one per lane.
// Queued pixels of one pair are consecutive, so the lanes of a run reduce
// their (min depth, max depth, min bin) among themselves and only the run's
// last lane touches the pair's LDS accumulators.
__device__ __forceinline__ void tiles_shade_queue(const TilesArgs& a, TileWaveLds& L, uint32_t head,
                                                  uint32_t n, const TileGeom& tg, f2 dbVS) {
  const uint32_t lane = __lane_id();
  const uint32_t tshift = (uint32_t)__builtin_ctz(a.T);
  uint32_t slot = 0xffffffffu, vmin = 0xffffffffu, vmax = 0u, vbin = 0xffffffffu;
  if (lane < n) {
    const uint32_t e = L.queue[(head + lane) & (kTileQueue - 1u)];
    slot = e >> 16;
    const uint32_t x = (uint32_t)L.px0[slot] + (e & 0xffu);
    const uint32_t y = (uint32_t)L.py0[slot] + ((e >> 8) & 0xffu);
    const uint32_t o = L.owner[slot];
    const f2 p = mk2((float)x, (float)y);
    const f2 w = interp_quad_pre<64>(&L.ip[0][o], mk2(L.v[0][o], L.v[1][o]), p);
    const float hdp = interp_f32(w, mk4(L.d[0][o], L.d[1][o], L.d[2][o], L.d[3][o]));
    // depth test with GL_LESS (:193-195), decided by the tile's depth bounds
    // when they settle it, by the texel otherwise
    const uint32_t tile = (y >> tshift) * tg.tilesX + (x >> tshift);
    const float2 bd = tile < a.nBounds ? a.depthBounds[tile] : make_float2(-INFINITY, INFINITY);
    bool pass;
    if (hdp <= bd.x) pass = true;
    else if (hdp > bd.y) pass = false;
    else pass = !(hdp > tex_load(a.depth, a.depthW, a.depthH, (int32_t)x, (int32_t)(tg.vpH - y)));
    if (pass) {
      if (!L.pairBin[o]) {
        const f3 vs = project(a.u->projMatrixInv, mk4(p.x, p.y, hdp, 1.0f));
        vbin = (a.B - 1u) - depth_bin(a.B, dbVS, vs.z);
      }
      // fminf/fmaxf in the reference skip a NaN depth; its bin still counts
      if (hdp == hdp) vmin = vmax = f32_ord(hdp);
    }
  }
  // segmented inclusive min/max over runs of equal slot, with the DPP steps
  // of wave_incl_scan (runs are contiguous, so a lane that matches the slot
  // of a lane before it belongs to the same run)
  seg_step<0x111, 0xf>(slot, vmin, vmax, vbin);
  seg_step<0x112, 0xf>(slot, vmin, vmax, vbin);
  seg_step<0x114, 0xf>(slot, vmin, vmax, vbin);
  seg_step<0x118, 0xf>(slot, vmin, vmax, vbin);
  seg_step<0x142, 0xa>(slot, vmin, vmax, vbin);
  seg_step<0x143, 0xc>(slot, vmin, vmax, vbin);
  const uint32_t nextSlot = (uint32_t)__shfl_down((int)slot, 1);
  if (lane < n && (lane == n - 1u || nextSlot != slot)) {
    if (vmin != 0xffffffffu) {
      atomicMin(&L.accMin[slot], vmin);
      atomicMax(&L.accMax[slot], vmax);
    }
    if (vbin != 0xffffffffu) atomicMin(&L.accBin[slot], vbin);
  }
  wave_sync();
}

// Steps 1-5 for one wave (the whole workgroup: LDS is ordered by wave_sync,
// which unlike a workgroup barrier does not wait for global atomics in
// flight). Each lane holds segment `gid` (when `have`) and
// draws the pairs [kBase, kBase + count) of its rectangle: in the main pass
// kBase = 0 and count = all tiles (or none, once queued as pieces); in the
// overflow pass one piece.
__device__ __forceinline__ void tiles_wave(const TilesArgs& a, TileWaveLds& L, uint32_t gid, bool have,
                                           uint32_t kBase, bool mainPass) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const uint32_t lane = threadIdx.x;
  const uint32_t segsPerStrand = a.P - 1u;
  const f2 vp = mk2(u->viewport[0], u->viewport[1]);
  const TileGeom tg = tile_geom(vp, a.T);
  const uint32_t T = a.T, B = a.B;
  const uint32_t limit = min(a.limit, u->maxDrawnHairSegments);

  // getHairDepthBoundsVS :242-246
  const f4 bsC = mul(u->modelViewMat, mk4(a.hd->boundingSphere[0], a.hd->boundingSphere[1],
                                          a.hd->boundingSphere[2], 1.0f));
  const f2 dbVS = mk2(bsC.z - a.hd->boundingSphere[3], bsC.z + a.hd->boundingSphere[3]);

  // 1. project
  uint32_t nTiles = 0;
  if (have) {
    const uint32_t strand = gid / segsPerStrand;
    const uint32_t seg = gid - strand * segsPerStrand;
    const Seg ps = project_segment(u, a.pos, a.tan, a.P, strand, seg, vp, u->fiberRadius);
    // getRasterizedHairBounds — swRasterizeHair.wgsl.ts:111-123
    f2 bmax = vmax(vmax(ps.v00, ps.v01), vmax(ps.v10, ps.v11));
    f2 bmin = vmin(vmin(ps.v00, ps.v01), vmin(ps.v10, ps.v11));
    bmax = vmin(mk2(ceilf(bmax.x), ceilf(bmax.y)), vp);
    bmin = vmax(mk2(floorf(bmin.x), floorf(bmin.y)), mk2(0.0f, 0.0f));
    const uint32_t tminx = f2u(bmin.x) / T, tminy = f2u(bmin.y) / T;
    const uint32_t tmaxx = f2u(bmax.x) / T, tmaxy = f2u(bmax.y) / T;
    const uint32_t tsx = (tmaxx - tminx) + 1u, tsy = (tmaxy - tminy) + 1u;
    // reject degenerate segments (:109-117); an inverted rectangle has no tile
    if (!(tsx * tsy > a.threshold) && tmaxx >= tminx && tmaxy >= tminy) nTiles = tsx * tsy;
    int cx0, cy0, cx1, cy1;
    const float margin = raster_margin(ps);
    seg_pixel_bbox(ps, margin, cx0, cy0, cx1, cy1);
    const float vv[8] = {ps.v00.x, ps.v00.y, ps.v01.x, ps.v01.y, ps.v10.x, ps.v10.y, ps.v11.x, ps.v11.y};
    float amax = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) { L.v[i][lane] = vv[i]; amax = fmaxf(amax, fabsf(vv[i])); }
    // collinear quads test whole tiles; far-off vertices skip the span
    span_frame(ps, cx0 > -(1 << 20) && amax < 65536.0f, margin, &L.sp[0][lane], 64);
    interp_pre(ps, &L.ip[0][lane], 64);
    L.d[0][lane] = ps.depths.x; L.d[1][lane] = ps.depths.y; L.d[2][lane] = ps.depths.z; L.d[3][lane] = ps.depths.w;
    L.pairBin[lane] = pair_bin_ok(u->projMatrixInv, ps.depths) ? 1u : 0u;
    L.box[0][lane] = cx0; L.box[1][lane] = cy0; L.box[2][lane] = cx1; L.box[3][lane] = cy1;
    L.rect[0][lane] = tminx; L.rect[1][lane] = tminy; L.rect[2][lane] = tsx;
    if (!mainPass) nTiles = kBase < nTiles ? min(kTilePiece, nTiles - kBase) : 0u;
  }
  if (mainPass && a.pieces) {
    // queue wide segments in pieces: one atomic per wave reserves the
    // wave's pieces; a lane whose pieces fall past the end of the queue
    // marks the part that fits as empty and draws its segment here
    const uint32_t np = nTiles > kTilePiece ? (nTiles + kTilePiece - 1u) / kTilePiece : 0u;
    const uint32_t incl = wave_incl_scan(np);
    const uint32_t total = wave_last(incl);
    if (total) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(a.pieceCount, total);
      base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
      const uint32_t first = base + incl - np;
      const bool fits = first + np <= a.pieceCap;
      for (uint32_t q = 0; q < np && first + q < a.pieceCap; q++)
        a.pieces[first + q] = fits ? make_uint2(gid, q * kTilePiece) : make_uint2(kInvalidPtr, 0u);
      if (np && fits) nTiles = 0;
    }
  }
  L.gid[lane] = gid;
  L.kBase[lane] = kBase;
  const uint32_t pairEnd = wave_incl_scan(nTiles);
  L.pairStart[lane] = pairEnd - nTiles;
  const uint32_t nPairs = wave_last(pairEnd);
  wave_sync();

  // 2. pairs, 64 at a time
  for (uint32_t base = 0; base < nPairs; base += 64u) {
    const uint32_t pidx = base + lane;
    uint32_t rows = 0, tx = 0, ty = 0;
    const uint32_t o = lbs_owner(L.mark, pairEnd, nTiles, base);
    if (pidx < nPairs) {
      const uint32_t k = L.kBase[o] + pidx - L.pairStart[o];
      const uint32_t tsx = L.rect[2][o];
      ty = L.rect[1][o] + k / tsx;
      tx = L.rect[0][o] + k % tsx;
      // getTileBoundsPx — tileUtils.wgsl.ts:44-54
      const uint32_t bx0 = clampu(tx * T, 0u, tg.vpW - 1u), by0 = clampu(ty * T, 0u, tg.vpH - 1u);
      const uint32_t bx1 = clampu((tx + 1u) * T, 0u, tg.vpW - 1u), by1 = clampu((ty + 1u) * T, 0u, tg.vpH - 1u);
      const int xs = max((int)bx0, L.box[0][o]), xe = min((int)bx1, L.box[2][o] + 1);
      const int ys = max((int)by0, L.box[1][o]), ye = min((int)by1, L.box[3][o] + 1);
      if (xe > xs && ye > ys) {
        rows = (uint32_t)(ye - ys);
        // a tile the span bound does not reach has no inside pixel
        if (L.sp[kSpOn][o] != 0.0f) {
          float xlo, xhi;
          span_x<64>(&L.sp[0][o], (float)ys, (float)(ye - 1), xlo, xhi);
          if (!(xlo <= xhi) || floorf(xhi) < (float)xs || ceilf(xlo) >= (float)xe) rows = 0;
        }
      }
      L.owner[lane] = o;
      L.px0[lane] = xs; L.py0[lane] = ys; L.px1[lane] = xe;
    }
    L.accMin[lane] = f32_ord(999.0f);
    L.accMax[lane] = f32_ord(-999.0f);
    L.accBin[lane] = B;
    const uint32_t rowEnd = wave_incl_scan(rows);
    L.rowStart[lane] = rowEnd - rows;
    const uint32_t nRows = wave_last(rowEnd);
    wave_sync();

    uint32_t head = 0, tail = 0;
    // 3a. pixel rows of the pairs, 64 at a time, each with its conservative
    // x span
    for (uint32_t r0 = 0; r0 < nRows; r0 += 64u) {
      const uint32_t ridx = r0 + lane;
      uint32_t cnt = 0;
      const uint32_t slot = lbs_owner(L.mark, rowEnd, rows, r0);
      if (ridx < nRows) {
        const int y = L.py0[slot] + (int)(ridx - L.rowStart[slot]);
        const uint32_t o = L.owner[slot];
        int xa = L.px0[slot], xb = L.px1[slot];
        if (L.sp[kSpOn][o] != 0.0f) {
          float xlo, xhi;
          span_x<64>(&L.sp[0][o], (float)y, (float)y, xlo, xhi);
          if (xlo <= xhi) {
            xa = max(xa, (int)fmaxf(ceilf(xlo), -1.0e9f));
            xb = min(xb, (int)fminf(floorf(xhi), 1.0e9f) + 1);
          } else {
            xb = xa;
          }
        }
        if (xb > xa) cnt = (uint32_t)(xb - xa);
        L.itemSlot[lane] = slot;
        L.itemX[lane] = xa;
        L.itemY[lane] = y;
      }
      const uint32_t candEnd = wave_incl_scan(cnt);
      L.candStart[lane] = candEnd - cnt;
      const uint32_t nCand = wave_last(candEnd);
      wave_sync();

      // 3b. candidate pixels, 64 at a time; 4. inside pixels, 64 at a time
      for (uint32_t c0 = 0; c0 < nCand; c0 += 64u) {
        const uint32_t c = c0 + lane;
        bool inside = false;
        uint32_t entry = 0;
        const uint32_t it = lbs_owner(L.mark, candEnd, cnt, c0);
        if (c < nCand) {
          const int x = L.itemX[it] + (int)(c - L.candStart[it]);
          const int y = L.itemY[it];
          const uint32_t slot = L.itemSlot[it];
          const uint32_t o = L.owner[slot];
          Seg ps;
          ps.v00 = mk2(L.v[0][o], L.v[1][o]); ps.v01 = mk2(L.v[2][o], L.v[3][o]);
          ps.v10 = mk2(L.v[4][o], L.v[5][o]); ps.v11 = mk2(L.v[6][o], L.v[7][o]);
          inside = inside_quad(ps, mk2((float)x, (float)y));
          entry = (slot << 16) | ((uint32_t)(y - L.py0[slot]) << 8) | (uint32_t)(x - L.px0[slot]);
        }
        const uint64_t m = __ballot(inside);
        if (inside) L.queue[(tail + mask_prefix(m)) & (kTileQueue - 1u)] = entry;
        tail += (uint32_t)__popcll(m);
        wave_sync();
        if (tail - head >= 64u) {
          tiles_shade_queue(a, L, head, 64u, tg, dbVS);
          head += 64u;
        }
      }
    }
    if (tail > head) tiles_shade_queue(a, L, head, tail - head, tg, dbVS);

    // 5. append the pairs with a hit (processTile :205-238)
    bool hit = false;
    float dmin = 999.0f, dmax = -999.0f;
    uint32_t dbin = B;
    if (pidx < nPairs) {
      dmin = ord_f32(L.accMin[lane]);
      dmax = ord_f32(L.accMax[lane]);
      dbin = L.accBin[lane];
      hit = !(dmin > 1.0f);
      if (hit && L.pairBin[L.owner[lane]]) {
        // the closest bin of the pair's passing pixels, from its extreme
        // depths (pair_bin_ok)
        const f2 c0 = clip_zw(u->projMatrixInv, dmin), c1 = clip_zw(u->projMatrixInv, dmax);
        const uint32_t b0 = depth_bin(B, dbVS, c0.x / c0.y), b1 = depth_bin(B, dbVS, c1.x / c1.y);
        dbin = (B - 1u) - max(b0, b1);
      }
    }
    const uint64_t mask = __ballot(hit);
    if (mask != 0) {
      uint32_t sbase = 0;
      const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
      if (lane == leader) sbase = atomicAdd(a.segCounter, (uint32_t)__popcll(mask));
      sbase = (uint32_t)__builtin_amdgcn_readlane((int)sbase, (int)leader);
      if (hit) {
        const uint32_t slot = sbase + mask_prefix(mask);
        if (slot < limit) {
          const uint32_t sgid = L.gid[L.owner[lane]];
          const uint32_t strand = sgid / segsPerStrand;
          const uint32_t seg = sgid - strand * segsPerStrand;
          const uint32_t tb = ty * tg.tilesX * B + tx * B + dbin;  // getHairTileDepthBinIdx
          fh_tile_result* t = &a.tiles[tb];
          atomicMax(&t->maxDepth, f2u(dmax * 4294967296.0f));
          atomicMax(&t->minDepth, f2u((1.0f - dmin) * 4294967296.0f));
          const uint32_t last = atomicExch(&t->tileSegmentPtr, slot + 1u);
          const uint32_t rank = atomicAdd(&t->padding0, 1u);
          const uint32_t prevPtr = last == 0u ? kInvalidPtr : last - 1u;
          *reinterpret_cast<uint2*>(&a.segData[slot]) = make_uint2((seg << 24) | strand, prevPtr);
          if (a.slotKey) *reinterpret_cast<uint2*>(&a.slotKey[2ull * slot]) = make_uint2(tb, rank);
        }
      }
    }
    wave_sync();
  }
}

__device__ __forceinline__ void tiles_draw_body(TilesArgs a, uint32_t block, uint32_t blocks) {
  __shared__ TileWaveLds L;
  const uint32_t gid = block * 64u + threadIdx.x;
  tiles_wave(a, L, gid, gid < a.S * (a.P - 1u), 0u, true);
}

struct TilesDraw {
  using Args = TilesArgs;
  static constexpr uint32_t kThreads = 64;
  static constexpr const char* kName = "tiles_draw";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { tiles_draw_body(a, xcd_order(block, blocks), blocks); }
};


// The queued pieces, `per` consecutive ones per wave: as few as spreads the
// queue over every workgroup (a short queue, one piece per wave), at most 64
// (one per lane). A wave's pixel work is spread over its lanes whatever the
// number of pieces it holds, so a short queue of wide segments runs on many
// waves at once instead of one after another on one. Every workgroup exits
// after the last piece.
__device__ __forceinline__ void tiles_pieces_body(TilesArgs a, uint32_t block, uint32_t blocks) {
  __shared__ TileWaveLds L;
  const uint32_t count = min(*a.pieceCount, a.pieceCap);
  const uint32_t per = min(64u, max(1u, (count + blocks - 1u) / blocks));
  for (uint32_t base = block * per; base < count; base += blocks * per) {
    const uint32_t i = base + threadIdx.x;
    const uint2 pc = threadIdx.x < per && i < count ? a.pieces[i] : make_uint2(kInvalidPtr, 0u);
    tiles_wave(a, L, pc.x, pc.x != kInvalidPtr, pc.y, false);
    wave_sync();
  }
}

struct TilesPieces {
  using Args = TilesArgs;
  static constexpr uint32_t kThreads = 64;
  static constexpr const char* kName = "tiles_pieces";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { tiles_pieces_body(a, block, blocks); }
};


// hairSegmentCountPerTile (processTile :236): the per-bin counts of each tile
// summed, which is what the per-append increments of the reference add up to.
__device__ __forceinline__ void tile_counts_body(const fh_tile_result* __restrict__ tiles,
                                                          uint32_t nTiles, uint32_t B,
                                                          uint32_t* __restrict__ countPerTile, uint32_t block, uint32_t blocks) {
  const uint32_t t = block * blockDim.x + threadIdx.x;
  if (t >= nTiles) return;
  uint32_t n = 0;
  for (uint32_t b = 0; b < B; b++) n += tiles[(size_t)t * B + b].padding0;
  countPerTile[t] = n;
}

struct TileCounts {
  struct Args {
    const fh_tile_result* tiles;
    uint32_t nTiles;
    uint32_t B;
    uint32_t* countPerTile;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "tile_counts";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { tile_counts_body(a.tiles, a.nTiles, a.B, a.countPerTile, block, blocks); }
};


// ---------------------------------------------------------------------------
// For lists written by another tile-pass implementation: one thread per
// (tile, bin) walks its list, tags each slot and records the list length.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rebuild_keys_body(fh_tile_result* __restrict__ tiles,
                                                           const fh_list_element* __restrict__ data,
                                                           uint32_t limit, uint32_t nTB,
                                                           uint32_t* __restrict__ slotKey, uint32_t block, uint32_t blocks) {
  const uint32_t tb = block * blockDim.x + threadIdx.x;
  if (tb >= nTB) return;
  uint32_t ptr = tiles[tb].tileSegmentPtr;
  ptr = ptr == 0u ? kInvalidPtr : ptr - 1u;
  uint32_t n = 0;
  while (ptr != kInvalidPtr && ptr < limit && n < limit) {
    slotKey[2ull * ptr] = tb;
    slotKey[2ull * ptr + 1] = n;
    n++;
    ptr = data[ptr].next;
  }
  tiles[tb].padding0 = n;
}

struct RebuildKeys {
  struct Args {
    fh_tile_result* tiles;
    const fh_list_element* data;
    uint32_t limit;
    uint32_t nTB;
    uint32_t* slotKey;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "rebuild_keys";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { rebuild_keys_body(a.tiles, a.data, a.limit, a.nTB, a.slotKey, block, blocks); }
};


// ---------------------------------------------------------------------------
// K5 + K6 — hairTileSortPass.countTiles.wgsl.ts:63-97 and
// hairTileSortPass.sort.wgsl.ts:37-75 in one workgroup of 1,024 threads: the
// bucket histogram in LDS, then the buckets' offsets (heavier buckets first,
// calculateOffsetsForEachBucket :68-75), then every drawn tile at its
// bucket's offset plus its rank from an LDS counter. Each thread loads
// kSortPer tiles' counts at once and issues their LDS atomics back to back.
// The buckets' tileCount / writeOffset and tileList[0] are added to as the
// reference's atomics add to them (a cleared buffer ends with each bucket's
// count in both), and no global atomic is left to wait behind other
// kernels' atomics in L2.
// ---------------------------------------------------------------------------
constexpr uint32_t kSortPer = 8;

struct SortTiles {
  struct Args {
    const uint32_t* counts;
    uint32_t nTiles;
    uint32_t nBuckets;
    uint32_t bucketSize;
    fh_sort_bucket* buckets;
    uint32_t* tileList;
  };
  static constexpr uint32_t kThreads = 1024;
  static constexpr const char* kName = "sort_tiles";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)block;
    (void)blocks;
    __shared__ uint32_t hist[64], cursor[64], offset[64];
    const uint32_t tid = threadIdx.x;
    if (tid < 64u) hist[tid] = 0u;
    __syncthreads();
    for (uint32_t t0 = tid; t0 < a.nTiles; t0 += kThreads * kSortPer) {
      uint32_t c[kSortPer];
#pragma unroll
      for (uint32_t k = 0; k < kSortPer; k++) {
        const uint32_t t = t0 + k * kThreads;
        c[k] = t < a.nTiles ? a.counts[t] : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < kSortPer; k++)
        if (c[k] > 0u) atomicAdd(&hist[min(c[k] / a.bucketSize, a.nBuckets - 1u)], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t off = 0, total = 0;
      for (uint32_t i = 0; i < a.nBuckets; i++) {
        const uint32_t b = a.nBuckets - 1u - i;
        const uint32_t count = a.buckets[b].tileCount + hist[b];
        a.buckets[b].tileCount = count;
        offset[b] = off;
        off += count;
        cursor[b] = a.buckets[b].writeOffset;
        a.buckets[b].writeOffset = cursor[b] + hist[b];
        total += hist[b];
      }
      a.tileList[0] += total;
    }
    __syncthreads();
    for (uint32_t t0 = tid; t0 < a.nTiles; t0 += kThreads * kSortPer) {
      uint32_t c[kSortPer];
#pragma unroll
      for (uint32_t k = 0; k < kSortPer; k++) {
        const uint32_t t = t0 + k * kThreads;
        c[k] = t < a.nTiles ? a.counts[t] : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < kSortPer; k++) {
        if (c[k] == 0u) continue;
        const uint32_t b = min(c[k] / a.bucketSize, a.nBuckets - 1u);
        const uint32_t o = offset[b] + atomicAdd(&cursor[b], 1u);
        // a caller that skipped the per-frame clear would push offsets past the list
        if (o < a.nTiles) a.tileList[1u + o] = t0 + k * kThreads;
      }
    }
  }
};


// ---------------------------------------------------------------------------
// Fine-pass gather: the per-(tile, bin) linked lists become one contiguous
// range per (tile, bin) of a scratch array, so the fine pass streams each
// list instead of chasing pointers. Step 1 is an exclusive scan of the list
// lengths (HairTileResult.padding0); step 2 writes every slot to the start
// of its range plus its rank in the list, both tagged by the tile pass.
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;

__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* warpSums) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t n = (uint32_t)__shfl_up((int)v, o);
    if (lane >= (uint32_t)o) v += n;
  }
  if (lane == 63u) warpSums[wid] = v;
  __syncthreads();
  if (wid == 0) {
    const uint32_t nw = blockDim.x >> 6;
    uint32_t s = lane < nw ? warpSums[lane] : 0u;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const uint32_t n = (uint32_t)__shfl_up((int)s, o);
      if (lane >= (uint32_t)o) s += n;
    }
    if (lane < nw) warpSums[lane] = s;
  }
  __syncthreads();
  if (wid > 0) v += warpSums[wid - 1];
  return v;
}

__device__ __forceinline__ void bin_scan_local_body(const fh_tile_result* __restrict__ tiles,
                                                                    uint32_t n, uint32_t* __restrict__ ends,
                                                                    uint32_t* __restrict__ blockSums, uint32_t block, uint32_t blocks) {
  __shared__ uint32_t warpSums[16];
  const uint32_t i = block * kScanBlock + threadIdx.x;
  const uint32_t v = i < n ? tiles[i].padding0 : 0u;
  const uint32_t s = block_inclusive_scan(v, warpSums);
  if (i < n) ends[i] = s - v;  // exclusive: first element of the range
  if (threadIdx.x == kScanBlock - 1) blockSums[block] = s;
}

struct BinScanLocal {
  struct Args {
    const fh_tile_result* tiles;
    uint32_t n;
    uint32_t* ends;
    uint32_t* blockSums;
    uint32_t* zero;   // a counter the later sort kernels count into (set to 0 here)
    uint32_t* zero2;  // the fine pass's count of exhausted bins (set to 0 here)
  };
  static constexpr uint32_t kThreads = kScanBlock;
  static constexpr const char* kName = "bin_scan_local";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    if (block == 0 && threadIdx.x == 0 && a.zero) *a.zero = 0u;
    if (block == 0 && threadIdx.x == 0 && a.zero2) *a.zero2 = 0u;
    bin_scan_local_body(a.tiles, a.n, a.ends, a.blockSums, block, blocks);
  }
};


__device__ __forceinline__ void bin_scan_add_body(uint32_t n, uint32_t* __restrict__ ends,
                                                                  const uint32_t* __restrict__ blockSums, uint32_t block, uint32_t blocks) {
  __shared__ uint32_t red[16];
  // prefix of this block = sum of the sums of all earlier blocks
  uint32_t acc = 0;
  for (uint32_t j = threadIdx.x; j < block; j += kScanBlock) acc += blockSums[j];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += (uint32_t)__shfl_xor((int)acc, o);
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  uint32_t prefix = 0;
  for (uint32_t w = 0; w < kScanBlock / 64; w++) prefix += red[w];
  const uint32_t i = block * kScanBlock + threadIdx.x;
  if (i < n && prefix) ends[i] += prefix;
}

struct BinScanAdd {
  struct Args {
    uint32_t n;
    uint32_t* ends;
    const uint32_t* blockSums;
  };
  static constexpr uint32_t kThreads = kScanBlock;
  static constexpr const char* kName = "bin_scan_add";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { bin_scan_add_body(a.n, a.ends, a.blockSums, block, blocks); }
};


// Ranges of more than this many keys are sorted by a workgroup each; shorter
// ones by a wave each.
constexpr uint32_t kWaveSortMax = 64;
// Scan blocks whose sums the scatter kernel prefixes itself (in LDS); beyond
// that a separate kernel adds the prefixes first.
constexpr uint32_t kMaxScanBlocks = 4096;

// Step 2, with the prefix of the scan blocks' sums taken in each workgroup:
// every range's global start goes to `starts` (and a range of more than
// kWaveSortMax keys onto the list `large` for the sort), and every tagged
// slot to its position. `ends` stays as the scan left it, so no workgroup
// reads a value another one rewrites.
struct BinScatter {
  struct Args {
    const fh_render_uniforms* u;
    const uint32_t* segCounter;
    const fh_list_element* data;
    uint32_t capacity;
    const uint2* slotTag;
    const uint32_t* ends;       // exclusive prefix within each scan block
    const uint32_t* blockSums;  // per scan block; null when `ends` already hold the global prefix
    uint32_t nSums;
    const fh_tile_result* tiles;
    uint32_t* starts;           // out: first element of each (tile, bin) range
    uint32_t nTB;
    uint32_t* binSeg;
    uint32_t* large;            // out: count in large[0] (zeroed by the scan), then the listed ranges
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "bin_scatter";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    __shared__ uint32_t pre[kMaxScanBlocks];
    __shared__ uint32_t warpSums[16];
    constexpr uint32_t per = kMaxScanBlocks / kThreads;
    const uint32_t tid = threadIdx.x;
    if (a.blockSums) {
      uint32_t v[per], sum = 0;
#pragma unroll
      for (uint32_t k = 0; k < per; k++) {
        const uint32_t j = tid * per + k;
        v[k] = j < a.nSums ? a.blockSums[j] : 0u;
        sum += v[k];
      }
      uint32_t run = block_inclusive_scan(sum, warpSums) - sum;
#pragma unroll
      for (uint32_t k = 0; k < per; k++) {
        pre[tid * per + k] = run;
        run += v[k];
      }
      __syncthreads();
    }
    const uint32_t lane = tid & 63u;
    for (uint32_t b0 = block * kThreads; b0 < a.nTB; b0 += blocks * kThreads) {
      const uint32_t tb = b0 + tid;
      bool big = false;
      if (tb < a.nTB) {
        const uint32_t st = a.ends[tb] + (a.blockSums ? pre[tb / kScanBlock] : 0u);
        a.starts[tb] = st;
        const uint32_t s0 = min(st, a.capacity);
        big = min(a.tiles[tb].padding0, a.capacity - s0) > kWaveSortMax;
      }
      const uint64_t m = __ballot(big);
      if (m) {
        uint32_t at = 0;
        const uint32_t leader = (uint32_t)__builtin_ctzll(m);
        if (lane == leader) at = atomicAdd(&a.large[0], (uint32_t)__popcll(m));
        at = (uint32_t)__builtin_amdgcn_readlane((int)at, (int)leader);
        if (big) a.large[1u + at + mask_prefix(m)] = tb;
      }
    }
    const uint32_t n = min(*a.segCounter, min(a.capacity, a.u->maxDrawnHairSegments));
    for (uint32_t sl = block * kThreads + tid; sl < n; sl += blocks * kThreads) {
      const uint2 tag = a.slotTag[sl];  // (tile-bin, rank in its list)
      if (tag.x >= a.nTB) continue;     // untagged slot: not linked into any list
      const uint32_t pos = a.ends[tag.x] + (a.blockSums ? pre[tag.x / kScanBlock] : 0u) + tag.y;
      if (pos < a.capacity) {
        const uint32_t packed = a.data[sl].strandAndSegmentIdxs;  // segment << 24 | strand
        a.binSeg[pos] = ((packed & 0x00ffffffu) << 8) | (packed >> 24);
      }
    }
  }
};


// ---------------------------------------------------------------------------
// Each gathered range sorted by blend key (strand << 8 | segment), largest
// first (bin_sort_kernel). The fine pass pushes fragments at the head of per-pixel lists in
// this order, so every list reads back in ascending key order: a fixed blend
// order for equal-depth-slice fragments, whatever order the tile pass's
// atomics appended the segments in. Keys are unique within a range (a
// segment enters each tile once, in one bin).
//   n <= 64: one wave, one key per lane, shuffle network in registers;
//   n <= kSortLds: bitonic network in LDS over the whole workgroup;
//   larger: the same network over a zero-padded copy in global scratch at
//   [2 * start, 2 * start + 2^ceil(log2 n)), disjoint between ranges.
// Padding with 0 is safe: it sorts last, and a real key 0 ties with it.
// ---------------------------------------------------------------------------
constexpr uint32_t kSortLds = 4096;

__device__ __forceinline__ void bitonic_desc(uint32_t* k, uint32_t n2, uint32_t tid, uint32_t nthr) {
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = tid; t < (n2 >> 1); t += nthr) {
        const uint32_t i = 2u * t - (t & (stride - 1u));
        const uint32_t j = i + stride;
        const bool desc = (i & size) == 0u;
        const uint32_t a = k[i], b = k[j];
        if (desc ? (a < b) : (a > b)) { k[i] = b; k[j] = a; }
      }
      __syncthreads();
    }
  }
}

// descending bitonic sort of one key per lane across the wave
__device__ __forceinline__ uint32_t wave_sort_desc(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)v, (int)stride);
      const bool desc = (lane & size) == 0u;
      const bool lower = (lane & stride) == 0u;
      v = (desc == lower) ? max(v, o) : min(v, o);
    }
  }
  return v;
}

// One workgroup per kSmallRanges consecutive ranges: ranges of up to
// kWaveSortMax keys are sorted one per wave in registers (each of the 4 waves
// takes every 4th short range, so a workgroup's ranges sort 4 at a time and
// the grid has enough waves to keep every SIMD busy); larger ones are on the
// list the scatter kernel made, so that the few heavy ranges, which cluster
// on screen, spread over many workgroups.
constexpr uint32_t kSmallRanges = 64;

__device__ __forceinline__ void bin_sort_small_body(const fh_tile_result* __restrict__ tiles,
                                                             const uint32_t* __restrict__ starts, uint32_t nTB,
                                                             uint32_t capacity, uint32_t* __restrict__ keys,
                                                             uint32_t block) {
  __shared__ uint32_t smallList[kSmallRanges];
  __shared__ uint32_t sStart[kSmallRanges], sCount[kSmallRanges];
  __shared__ uint32_t nSmall;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) nSmall = 0;
  __syncthreads();
  const uint32_t tb = block * kSmallRanges + tid;
  uint32_t start = 0, n = 0;
  if (tid < kSmallRanges && tb < nTB) {
    start = min(starts[tb], capacity);
    n = min(tiles[tb].padding0, capacity - start);
  }
  if (tid < kSmallRanges) {
    sStart[tid] = start;
    sCount[tid] = n;
  }
  if (n >= 2u && n <= kWaveSortMax) smallList[atomicAdd(&nSmall, 1u)] = tid;
  __syncthreads();
  const uint32_t lane = __lane_id(), w = tid >> 6;
  for (uint32_t i = w; i < nSmall; i += blockDim.x >> 6) {
    const uint32_t j = smallList[i];
    const uint32_t st = sStart[j], cnt = sCount[j];
    uint32_t v = lane < cnt ? keys[st + lane] : 0u;
    v = wave_sort_desc(v);
    if (lane < cnt) keys[st + lane] = v;
  }
}


// The same descending network over n2 = 256 * R keys (n2 <= kSortLds) held
// in registers by the 256 threads of a workgroup: thread t holds elements
// t * R .. t * R + R - 1. A compare-exchange at distance s runs
//   s < R:        between two registers of one thread;
//   s < 64 * R:   between lanes t and t ^ (s / R) of one wave (shuffle);
//   otherwise:    through LDS (only distances n2 / 4 and n2 / 2: 3 stages),
//                 rows padded to R + 1 words so that lanes hit distinct banks.
// Each stage compares the same pairs in the same direction as bitonic_desc.
template <uint32_t R>
__device__ __forceinline__ void bitonic_regs_desc(uint32_t (&k)[R], uint32_t* lds) {
  constexpr uint32_t n2 = 256u * R;
  constexpr uint32_t P = R == 1 ? 1u : R + 1u;  // LDS row pitch
  const uint32_t t = threadIdx.x;
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t s = size >> 1; s > 0; s >>= 1) {
      if (s < R) {
#pragma unroll
        for (uint32_t sb = 1; sb < R; sb <<= 1) {
          if (s != sb) continue;
#pragma unroll
          for (uint32_t r = 0; r < R; r++) {
            if (r & sb) continue;
            const bool desc = ((t * R + r) & size) == 0u;
            const uint32_t hi = max(k[r], k[r | sb]), lo = min(k[r], k[r | sb]);
            k[r] = desc ? hi : lo;
            k[r | sb] = desc ? lo : hi;
          }
        }
      } else {
        const uint32_t q = s / R;  // distance in threads
        const bool lower = (t & q) == 0u;
        if (q < 64u) {
#pragma unroll
          for (uint32_t r = 0; r < R; r++) {
            const uint32_t o = (uint32_t)__shfl_xor((int)k[r], (int)q);
            const bool desc = ((t * R + r) & size) == 0u;
            k[r] = (desc == lower) ? max(k[r], o) : min(k[r], o);
          }
        } else {
          __syncthreads();  // earlier reads of lds are done
#pragma unroll
          for (uint32_t r = 0; r < R; r++) lds[t * P + r] = k[r];
          __syncthreads();
          const uint32_t pt = t ^ q;
#pragma unroll
          for (uint32_t r = 0; r < R; r++) {
            const uint32_t o = lds[pt * P + r];
            const bool desc = ((t * R + r) & size) == 0u;
            k[r] = (desc == lower) ? max(k[r], o) : min(k[r], o);
          }
        }
      }
    }
  }
}

template <uint32_t R>
__device__ __forceinline__ void sort_range_regs(uint32_t* __restrict__ keys, uint32_t st, uint32_t cnt,
                                                uint32_t* lds) {
  uint32_t k[R];
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (uint32_t r = 0; r < R; r++) {
    const uint32_t e = t * R + r;
    k[r] = e < cnt ? keys[st + e] : 0u;
  }
  bitonic_regs_desc<R>(k, lds);
#pragma unroll
  for (uint32_t r = 0; r < R; r++) {
    const uint32_t e = t * R + r;
    if (e < cnt) keys[st + e] = k[r];
  }
}

// One workgroup of 256 threads per listed range (more than 64 keys).
__device__ __forceinline__ void bin_sort_large_body(const fh_tile_result* __restrict__ tiles,
                                                             const uint32_t* __restrict__ starts,
                                                             uint32_t capacity, uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ large,
                                                             uint32_t* __restrict__ big, uint32_t block, uint32_t blocks) {
  __shared__ uint32_t sk[256u * 17u];
  const uint32_t count = large[0];
  for (uint32_t i = block; i < count; i += blocks) {
    const uint32_t tb = large[1u + i];
    const uint32_t st = min(starts[tb], capacity);
    const uint32_t cnt = min(tiles[tb].padding0, capacity - st);
    uint32_t n2 = 256;
    while (n2 < cnt) n2 <<= 1;
    // every branch below is uniform over the workgroup (n2 is)
    __syncthreads();  // the previous range's reads of sk are done
    switch (n2) {
      case 256u: sort_range_regs<1>(keys, st, cnt, sk); break;
      case 512u: sort_range_regs<2>(keys, st, cnt, sk); break;
      case 1024u: sort_range_regs<4>(keys, st, cnt, sk); break;
      case 2048u: sort_range_regs<8>(keys, st, cnt, sk); break;
      case 4096u: sort_range_regs<16>(keys, st, cnt, sk); break;
      default: {
        uint32_t* k = big + 2ull * st;
        for (uint32_t e = threadIdx.x; e < n2; e += blockDim.x) k[e] = e < cnt ? keys[st + e] : 0u;
        __syncthreads();
        bitonic_desc(k, n2, threadIdx.x, blockDim.x);
        for (uint32_t e = threadIdx.x; e < cnt; e += blockDim.x) keys[st + e] = k[e];
      }
    }
  }
}

// Both sorts in one launch: the first nLargeBlocks workgroups take the listed
// long ranges (dispatched first, so that they start at once), the rest the
// short ranges, kSmallRanges (tile, bin) entries each.
struct BinSort {
  struct Args {
    const fh_tile_result* tiles;
    const uint32_t* starts;
    uint32_t nTB;
    uint32_t capacity;
    uint32_t* keys;
    const uint32_t* large;
    uint32_t* big;
    uint32_t nLargeBlocks;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "bin_sort";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    if (block < a.nLargeBlocks) bin_sort_large_body(a.tiles, a.starts, a.capacity, a.keys, a.large, a.big, block, a.nLargeBlocks);
    else bin_sort_small_body(a.tiles, a.starts, a.nTB, a.capacity, a.keys, block - a.nLargeBlocks);
  }
};


// ---------------------------------------------------------------------------
// K7 — hairFinePass.wgsl.ts:126-335 with processHairSegment.wgsl.ts:5-70 and
// reduceHairSlices.wgsl.ts:9-78. Persistent workgroups ("processors") pop
// tiles from the sorted tile list. A workgroup has NW waves; wave w owns the
// pixel rows [w * TILE / NW, (w + 1) * TILE / NW) of the tile. Per depth bin,
// the bin's segments arrive sorted by blend key (largest first) in batches:
//  * one thread per segment projects it into LDS (field-major records: quad,
//    depths, span frame, pixel box relative to the tile, nearest depth); a
//    tile with at most one batch of segments in all its bins is projected
//    once, before its first bin;
//  * each wave enumerates the (segment, pixel) candidates of its rows,
//    record by record (box rows x row spans), 64 per step, runs the edge
//    tests and a conservative depth pre-test, and queues the covered pixels
//    in its LDS ring in that order;
//  * every 64 queued pixels the wave computes the fragments together:
//    segment-space coordinates, depth test, then a slot of the slice memory
//    (perProc per bin, _hasMoreSliceDataSlots), coverage alpha, shading and
//    slice index; each fragment is linked at the head of its (pixel, slice)
//    list. Two lanes linking to one pixel in one step do it in lane order
//    (queue order), so each list reads back in ascending blend key;
//  * each pixel then blends its slices front to back.
// Slots live in LDS first, then in the processor's region of the reference's
// slice buffer.
// ---------------------------------------------------------------------------
struct FineArgs {
  const fh_render_uniforms* u;
  const float4* pos;
  const float4* tan;
  const fh_tile_result* tiles;
  const uint32_t* binStart;  // first element of each (tile, bin) range
  const uint32_t* binKey;    // blend keys (strand << 8 | segment), sorted largest first per range
  uint4* slicePool;          // processorCount * perProc entries
  uint32_t* queue;           // HairRasterResult.tileQueueAtomicIdx
  float4* fb;                // HairRasterResult.data
  const float* depth;
  uint32_t depthW, depthH;
  const float4* shading;
  const uint32_t* tileList;
  uint32_t S, P;
  uint32_t B, NS, PS;
  uint32_t perProc, ldsPool, batch;
  uint32_t capacity;  // elements of the gathered segment array
  uint32_t alphaQuadratic;
  float alphaMultiplier;
  uint32_t* exhausted;  // bins whose slice memory ran out (statistics)
};

// Segment records of one batch, field-major in LDS (field f of record j at
// f * batch + j): lanes reading different records hit different banks.
enum RecField : uint32_t {
  kRecV = 0,                     // 8: v00, v01, v10, v11
  kRecD = 8,                     // 4: projected depths
  kRecSpan = 12,                 // kSpFields: span frame (span_frame)
  kRecKey = kRecSpan + kSpFields,  // strand << 8 | segment
  kRecBox0,                      // pixel box relative to the tile, inclusive, clamped to
  kRecBox1,                      //   [-1, TILE]: (x0 | y0 << 16), (x1 | y1 << 16) as int16
  kRecNear,                      // lower bound of the interpolated depth (-inf when unknown)
  kRecIp,                        // kIpFields: per-segment interpolation terms (interp_pre)
  kRecFields = kRecIp + kIpFields
};

template <uint32_t RB>
struct RecStore {
  float* f;
  __device__ __forceinline__ float& at(uint32_t field, uint32_t j) const { return f[field * RB + j]; }
  __device__ __forceinline__ uint32_t u(uint32_t field, uint32_t j) const { return __float_as_uint(f[field * RB + j]); }
  __device__ __forceinline__ int lo16(uint32_t field, uint32_t j) const { return (int)(int16_t)(u(field, j) & 0xffffu); }
  __device__ __forceinline__ int hi16(uint32_t field, uint32_t j) const { return (int)(int16_t)(u(field, j) >> 16); }
};

template <int TILE>
constexpr uint32_t fine_batch() { return TILE * TILE < 256 ? TILE * TILE : 256; }
// waves per workgroup: 2 pixel rows per wave for 16x16 and 32x32 tiles
template <int TILE>
constexpr uint32_t fine_waves() { return TILE == 8 ? 1u : (uint32_t)TILE / 2u; }

constexpr uint32_t kFineRing = 128;  // per-wave ring of queued pixels

template <int TILE>
struct FineLds {
  static constexpr uint32_t NPX = TILE * TILE;
  static constexpr uint32_t NW = fine_waves<TILE>();
  static constexpr uint32_t PXW = NPX / NW;  // pixels per wave
  RecStore<fine_batch<TILE>()> recs;
  uint32_t* heads;   // NS * NPX: newest slot of each (slice, pixel) list
  uint32_t* smask;   // NPX: the slices whose list is not empty (bit s: slice s)
  float* db;         // NPX: mesh depth
  uint32_t* done;    // NPX: pixel saturated
  uint32_t* ring;    // NW * kFineRing
  uint32_t* wsc;     // NW * 6 * 64: per-wave enumeration scratch
  uint2* poolColor;  // ldsPool: f16 rg, ba
  uint32_t* poolNext;
};

// The queued pixels [head, head + n) of wave `w`, n <= 64, one per lane.
template <int TILE>
__device__ __forceinline__ void fine_shade_queue(const FineArgs& a, const FineLds<TILE>& L, uint4* gpool,
                                                 uint32_t w, uint32_t head, uint32_t n, uint32_t bx0,
                                                 uint32_t by0, f2 td, float invSegs, uint32_t* sliceOff) {
  constexpr uint32_t PXW = FineLds<TILE>::PXW;
  const uint32_t lane = __lane_id();
  bool accept = false;
  uint32_t j = 0, pix = 0;
  f2 wq = mk2(0.0f, 0.0f);
  float hdp = 0.0f, fr = 0.0f;
  float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0;
  const auto& R = L.recs;
  if (lane < n) {
    const uint32_t e = L.ring[w * kFineRing + ((head + lane) & (kFineRing - 1u))];
    j = e >> 8;
    pix = w * PXW + (e & 0xffu);
    const f2 p = mk2((float)(bx0 + pix % TILE), (float)(by0 + pix / TILE));
    // interpolateHairQuad from the record's per-segment terms (bit-identical)
    wq = interp_quad_pre<fine_batch<TILE>()>(&R.at(kRecIp, j), mk2(R.at(kRecV + 0, j), R.at(kRecV + 1, j)), p);
    hdp = interp_f32(wq, mk4(R.at(kRecD + 0, j), R.at(kRecD + 1, j), R.at(kRecD + 2, j), R.at(kRecD + 3, j)));
    // _sampleShading (hairShadingBuffer.ts:29-38): the two shading points are
    // loaded now, before the depth test and the slot, so that their latency
    // overlaps that work (a rejected fragment's loads are simply unused)
    const uint32_t key = R.u(kRecKey, j);
    const uint32_t strand = key >> 8, seg = key & 0xffu;
    const float tFull = ((float)seg + wq.y) / invSegs;
    const float sa = tFull * ((float)a.PS - 1.0f);
    const uint32_t au = f2u(sa);
    const uint32_t i0 = clampu(au, 0u, a.PS - 1u), i1 = clampu(au + 1u, 0u, a.PS - 1u);
    fr = fractf(sa);
    c0 = a.shading[(size_t)strand * a.PS + i0];
    c1 = a.shading[(size_t)strand * a.PS + i1];
    accept = !(hdp >= L.db[pix]);  // depth test with GL_LESS (:31-34)
  }
  // a slot of the slice memory: _hasMoreSliceDataSlots (hairFinePass.wgsl.ts:252-264)
  const uint64_t m = __ballot(accept);
  if (m == 0) return;
  uint32_t base = 0;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  if (lane == leader) base = atomicAdd(sliceOff, (uint32_t)__popcll(m));
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader);
  const uint32_t slot = base + mask_prefix(m);
  if (accept && slot >= a.perProc) accept = false;
  uint32_t tag = 0;
  if (accept) {
    // getAlphaCoverage (processHairSegment.wgsl.ts:61-70)
    float alpha = 1.0f - fabsf(wq.x * 2.0f - 1.0f);
    if (a.alphaQuadratic) alpha = sqrtf(alpha);
    alpha = satf(alpha * a.alphaMultiplier);
    const float cr = mixf(c0.x, c1.x, fr), cg = mixf(c0.y, c1.y, fr);
    const float cb = mixf(c0.z, c1.z, fr), ca = mixf(c0.w, c1.w, fr) * alpha;
    const uint2 col = make_uint2(pack_half2(cr, cg), pack_half2(cb, ca));
    tag = depth_bin(a.NS, td, hdp);  // getSliceIdx (:72-75)
    if (slot < a.ldsPool) L.poolColor[slot] = col;
    else *reinterpret_cast<uint2*>(&gpool[slot - a.ldsPool]) = col;
  }
  // Link at the head of the (slice, pixel) lists in lane order (queue order):
  // the lanes of one list find each other with one ballot per bit of the
  // list's index within the wave's pixels; each links to the next lower such
  // lane (the lowest to the old head) and the highest becomes the head.
  uint64_t grp = __ballot(accept);
  const uint32_t lk = tag * PXW + (pix - w * PXW);
  const uint32_t nbits = 32u - (uint32_t)__builtin_clz(a.NS * PXW - 1u | 1u);
  for (uint32_t b = 0; b < nbits; b++) {
    const uint64_t bal = __ballot((lk >> b) & 1u);
    grp &= ((lk >> b) & 1u) ? bal : ~bal;
  }
  const uint64_t below = grp & ((1ull << lane) - 1ull);
  const uint64_t above = lane == 63u ? 0ull : (grp >> (lane + 1u));
  const uint32_t prev = below ? 63u - (uint32_t)__builtin_clzll(below) : lane;
  const uint32_t prevSlot = (uint32_t)__shfl((int)slot, (int)prev);
  if (accept) {
    uint32_t* hp = &L.heads[tag * FineLds<TILE>::NPX + pix];
    const uint32_t next = below ? prevSlot : *hp;
    if (slot < a.ldsPool) L.poolNext[slot] = next;
    else gpool[slot - a.ldsPool].z = next;
    if (above == 0ull) {
      *hp = slot;
      atomicOr(&L.smask[pix], 1u << tag);  // lanes of one pixel may set different slices
    }
  }
  wave_sync();
}

// Cheap tests of the records [rec0, rec0 + nb) for wave `w`. Candidates are
// enumerated record by record (records are in blend order), row by row within
// the record's box and its row span, 64 per step; covered pixels are queued in
// that order.
template <int TILE>
__device__ __forceinline__ void fine_cheap_batch(const FineArgs& a, const FineLds<TILE>& L, uint4* gpool,
                                                 uint32_t rec0, uint32_t nb, uint32_t bx0, uint32_t by0,
                                                 f2 td, float invSegs, uint32_t* sliceOff) {
  constexpr uint32_t PXW = FineLds<TILE>::PXW;
  constexpr int RW = (int)(PXW / TILE);
  const uint32_t lane = __lane_id();
  const uint32_t w = threadIdx.x >> 6;
  const int wy0 = (int)w * RW, wy1 = wy0 + RW - 1;
  uint32_t* rowStart = L.wsc + (w * 6u + 0u) * 64u;
  uint32_t* itemRec = L.wsc + (w * 6u + 1u) * 64u;
  uint32_t* itemY = L.wsc + (w * 6u + 2u) * 64u;
  uint32_t* itemX = L.wsc + (w * 6u + 3u) * 64u;
  uint32_t* candStart = L.wsc + (w * 6u + 4u) * 64u;
  int* mark = reinterpret_cast<int*>(L.wsc + (w * 6u + 5u) * 64u);
  const auto& R = L.recs;
  uint32_t qh = 0, qt = 0;
  for (uint32_t rb = 0; rb < nb; rb += 64u) {
    const uint32_t jr = rec0 + rb + lane;
    uint32_t nrows = 0;
    if (rb + lane < nb) {
      const int ya = max(R.hi16(kRecBox0, jr), wy0), yb = min(R.hi16(kRecBox1, jr), wy1);
      const int xa = max(R.lo16(kRecBox0, jr), 0), xb = min(R.lo16(kRecBox1, jr), TILE - 1);
      if (yb >= ya && xb >= xa) nrows = (uint32_t)(yb - ya + 1);
    }
    const uint32_t re = wave_incl_scan(nrows);
    rowStart[lane] = re - nrows;
    const uint32_t nRows = wave_last(re);
    wave_sync();
    for (uint32_t r0 = 0; r0 < nRows; r0 += 64u) {
      const uint32_t ridx = r0 + lane;
      uint32_t cnt = 0;
      const uint32_t k = lbs_owner(mark, re, nrows, r0);
      if (ridx < nRows) {
        const uint32_t j = rec0 + rb + k;
        const int y = max(R.hi16(kRecBox0, j), wy0) + (int)(ridx - rowStart[k]);
        int xa = max(R.lo16(kRecBox0, j), 0), xb = min(R.lo16(kRecBox1, j), TILE - 1) + 1;
        if (R.at(kRecSpan + kSpOn, j) != 0.0f) {
          float xlo, xhi;
          const float yf = (float)(by0 + (uint32_t)y);
          span_x<fine_batch<TILE>()>(&R.at(kRecSpan, j), yf, yf, xlo, xhi);
          if (xlo <= xhi) {
            xa = max(xa, (int)fmaxf(ceilf(xlo), -1.0e9f) - (int)bx0);
            xb = min(xb, (int)fminf(floorf(xhi), 1.0e9f) + 1 - (int)bx0);
          } else {
            xb = xa;
          }
        }
        if (xb > xa) cnt = (uint32_t)(xb - xa);
        itemRec[lane] = j;
        itemY[lane] = (uint32_t)y;
        itemX[lane] = (uint32_t)xa;
      }
      const uint32_t ce = wave_incl_scan(cnt);
      candStart[lane] = ce - cnt;
      const uint32_t nCand = wave_last(ce);
      wave_sync();
      for (uint32_t c0 = 0; c0 < nCand; c0 += 64u) {
        const uint32_t c = c0 + lane;
        bool push = false;
        uint32_t entry = 0;
        const uint32_t it = lbs_owner(mark, ce, cnt, c0);
        if (c < nCand) {
          const uint32_t x = itemX[it] + (c - candStart[it]);
          const uint32_t y = itemY[it];
          const uint32_t j = itemRec[it];
          const uint32_t pix = y * TILE + x;
          if (!L.done[pix] && !(R.at(kRecNear, j) >= L.db[pix])) {
            Seg sg;
            sg.v00 = mk2(R.at(kRecV + 0, j), R.at(kRecV + 1, j)); sg.v01 = mk2(R.at(kRecV + 2, j), R.at(kRecV + 3, j));
            sg.v10 = mk2(R.at(kRecV + 4, j), R.at(kRecV + 5, j)); sg.v11 = mk2(R.at(kRecV + 6, j), R.at(kRecV + 7, j));
            push = inside_quad(sg, mk2((float)(bx0 + x), (float)(by0 + y)));
          }
          entry = (j << 8) | (pix - w * PXW);
        }
        const uint64_t m = __ballot(push);
        if (m == 0) continue;
        if (push) L.ring[w * kFineRing + ((qt + mask_prefix(m)) & (kFineRing - 1u))] = entry;
        qt += (uint32_t)__popcll(m);
        wave_sync();
        if (qt - qh >= 64u) {
          fine_shade_queue<TILE>(a, L, gpool, w, qh, 64u, bx0, by0, td, invSegs, sliceOff);
          qh += 64u;
        }
      }
    }
  }
  if (qt > qh) fine_shade_queue<TILE>(a, L, gpool, w, qh, qt - qh, bx0, by0, td, invSegs, sliceOff);
}

// Projects segment `key` into record `j`: quad, depths, span frame, pixel box
// relative to the tile, nearest depth and the per-segment interpolation terms.
template <int TILE>
__device__ __forceinline__ void fine_record(const FineArgs& a, const FineLds<TILE>& L, uint32_t j, uint32_t key,
                                            const Seg& sg, uint32_t bx0, uint32_t by0) {
  constexpr uint32_t RB = fine_batch<TILE>();
  int x0, y0, x1, y1;
  const float margin = raster_margin(sg);
  seg_pixel_bbox(sg, margin, x0, y0, x1, y1);
  const auto& R = L.recs;
  const float vv[8] = {sg.v00.x, sg.v00.y, sg.v01.x, sg.v01.y, sg.v10.x, sg.v10.y, sg.v11.x, sg.v11.y};
  float amax = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; i++) { R.at(kRecV + i, j) = vv[i]; amax = fmaxf(amax, fabsf(vv[i])); }
  // collinear quads test the whole box; far-off vertices skip the span
  span_frame(sg, x0 > -(1 << 20) && amax < 65536.0f, margin, &R.at(kRecSpan, j), RB);
  R.at(kRecD + 0, j) = sg.depths.x; R.at(kRecD + 1, j) = sg.depths.y;
  R.at(kRecD + 2, j) = sg.depths.z; R.at(kRecD + 3, j) = sg.depths.w;
  R.at(kRecKey, j) = __uint_as_float(key);
  // relative to the tile origin, clamped to one pixel beyond the tile
  const int rx0 = max(min(x0 - (int)bx0, TILE), -1), ry0 = max(min(y0 - (int)by0, TILE), -1);
  const int rx1 = max(min(x1 - (int)bx0, TILE), -1), ry1 = max(min(y1 - (int)by0, TILE), -1);
  R.at(kRecBox0, j) = __uint_as_float(((uint32_t)rx0 & 0xffffu) | ((uint32_t)ry0 << 16));
  R.at(kRecBox1, j) = __uint_as_float(((uint32_t)rx1 & 0xffffu) | ((uint32_t)ry1 << 16));
  // The interpolated depth is a two-level lerp of the four depths with
  // weights in [0, 1]: at least their minimum, less the rounding of two lerps
  // (a few ulp of the largest magnitude). Non-finite depths can interpolate
  // to NaN, which passes the depth test.
  const bool finite = isfinite(sg.depths.x) && isfinite(sg.depths.y) && isfinite(sg.depths.z) &&
                      isfinite(sg.depths.w);
  const float dn = fminf(fminf(sg.depths.x, sg.depths.y), fminf(sg.depths.z, sg.depths.w));
  const float dm = fmaxf(fmaxf(fabsf(sg.depths.x), fabsf(sg.depths.y)), fmaxf(fabsf(sg.depths.z), fabsf(sg.depths.w)));
  R.at(kRecNear, j) = finite ? dn - (dm * 1e-6f + 1e-30f) : -INFINITY;
  interp_pre(sg, &R.at(kRecIp, j), RB);
}

template <int TILE>
__device__ __forceinline__ void fine_project(const FineArgs& a, const FineLds<TILE>& L, uint32_t j, uint32_t key,
                                             f2 vpf, uint32_t bx0, uint32_t by0) {
  const fh_render_uniforms* __restrict__ u = a.u;
  const Seg sg = project_segment(u, a.pos, a.tan, a.P, key >> 8, key & 0xffu, vpf, u->fiberRadius);
  fine_record<TILE>(a, L, j, key, sg, bx0, by0);
}

// The two points and tangents of a segment, loaded ahead of its projection.
struct SegIn {
  uint32_t key;
  float4 p0, t0, p1, t1;
};
__device__ __forceinline__ void seg_load(const FineArgs& a, uint32_t key, SegIn& s) {
  const size_t k = (size_t)(key >> 8) * a.P + (key & 0xffu);
  s.key = key;
  s.p0 = a.pos[k]; s.t0 = a.tan[k];
  s.p1 = a.pos[k + 1]; s.t1 = a.tan[k + 1];
}

template <int TILE>
__device__ __forceinline__ void fine_raster_body(FineArgs a, uint32_t block, uint32_t blocks) {
  constexpr uint32_t NPX = TILE * TILE;
  constexpr uint32_t NW = fine_waves<TILE>();
  constexpr uint32_t RB = fine_batch<TILE>();
  extern __shared__ __align__(16) unsigned char smem[];
  FineLds<TILE> L;
  L.recs.f = reinterpret_cast<float*>(smem);
  L.heads = reinterpret_cast<uint32_t*>(smem + (size_t)RB * kRecFields * 4);
  L.smask = L.heads + (size_t)a.NS * NPX;
  L.db = reinterpret_cast<float*>(L.smask + NPX);
  L.done = reinterpret_cast<uint32_t*>(L.db + NPX);
  L.ring = L.done + NPX;
  L.wsc = L.ring + NW * kFineRing;
  L.poolColor = reinterpret_cast<uint2*>(L.wsc + NW * 6u * 64u);
  L.poolNext = reinterpret_cast<uint32_t*>(L.poolColor + a.ldsPool);
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_sliceOff;  // slots of the slice memory taken in the bin
  // the tile's bins, loaded together: minDepth, maxDepth, segments, first
  // element; and the first record of each bin when the tile is projected whole
  __shared__ uint4 s_bin[64];
  __shared__ uint32_t s_binRec[65];

  const fh_render_uniforms* __restrict__ u = a.u;
  const uint32_t tid = threadIdx.x;
  const bool pixThread = tid < NPX;
  const uint32_t lx = tid % TILE, ly = tid / TILE;
  const f2 vpf = mk2(u->viewport[0], u->viewport[1]);
  const TileGeom tg = tile_geom(vpf, TILE);
  const uint32_t tileCount = tg.tilesX * tg.tilesY;
  const uint32_t drawnTiles = a.tileList[0];
  const uint32_t fbLen = tg.vpW * tg.vpH - 1u;  // W*H*16 B minus the 16 B header
  const uint32_t dispMode = u->displayMode & 0xffu;
  const uint32_t dbgMax = dispMode == 4u ? ((u->displayMode >> 8) & 0xffffu) : 0u;  // USED_SLICES
  const bool isDbg = dbgMax != 0u;
  const float invSegs = (float)(a.P - 1u);
  uint4* gpool = a.slicePool + (size_t)block * a.perProc;

  for (uint32_t i = tid; i < a.NS * NPX; i += blockDim.x) L.heads[i] = kInvalidPtr;
  for (uint32_t i = tid; i < NPX; i += blockDim.x) L.smask[i] = 0u;

  while (true) {
    __syncthreads();
    if (tid == 0) {
      const uint32_t q = atomicAdd(a.queue, 1u);  // getNextTileIdx :294-303
      s_tile = q < drawnTiles ? a.tileList[1 + q] : kInvalidPtr;
    }
    __syncthreads();
    const uint32_t tileIdx = s_tile;
    if (tileIdx == kInvalidPtr) break;
    if (tid < 64u) {
      // the tile's bins: segment ranges, and where each bin's records start
      // when the whole tile is projected at once
      uint4 tr = make_uint4(0u, 0u, 0u, 0u);
      uint32_t n = 0;
      if (tid < a.B && tileIdx < tileCount) {
        const uint32_t tb = tileIdx * a.B + tid;
        const fh_tile_result r = a.tiles[tb];
        tr = make_uint4(r.minDepth, r.maxDepth, r.padding0, min(a.binStart[tb], a.capacity));
        if (tr.y != 0u) n = min(tr.z, a.capacity - tr.w);
      }
      const uint32_t e = wave_incl_scan(n);
      s_bin[tid] = make_uint4(tr.x, tr.y, n, tr.w);
      s_binRec[tid] = e - n;
      if (tid == 63u) s_binRec[64] = e;
    }
    const uint32_t row = tileIdx / tg.tilesX;
    const uint32_t tx = tileIdx - tg.tilesX * row, ty = row;
    const uint32_t bx0 = clampu(tx * TILE, 0u, tg.vpW - 1u), by0 = clampu(ty * TILE, 0u, tg.vpH - 1u);
    const uint32_t px = bx0 + lx, py = by0 + ly;
    const f2 p = mk2((float)px, (float)py);
    const bool readable = pixThread && px < tg.vpW && py < tg.vpH && (tg.vpW * py + px) < fbLen;
    const uint32_t fbIdx = tg.vpW * py + px;
    if (pixThread) {
      L.db[tid] = tex_load(a.depth, a.depthW, a.depthH, (int)px, f2i(vpf.y - p.y));
      L.done[tid] = 0u;
    }
    bool done = !pixThread;
    bool blended = false;  // a bin's slices were blended into fc
    f4 fc = mk4(0.0f, 0.0f, 0.0f, 0.0f);
    if (readable) { const float4 v = a.fb[fbIdx]; fc = mk4(v.x, v.y, v.z, v.w); }

    __syncthreads();  // s_bin, s_binRec
    // A tile whose segments fit in one batch is projected once, all bins
    // together, so a bin costs no memory round trips before its fragments.
    const uint32_t tileSegs = s_binRec[64];
    const bool whole = tileSegs <= a.batch;
    if (whole) {
      if (tid < tileSegs) {
        uint32_t lo = 0, hi = a.B;  // the last bin whose first record is <= tid
        while (hi - lo > 1u) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_binRec[mid] <= tid) lo = mid;
          else hi = mid;
        }
        fine_project<TILE>(a, L, tid, a.binKey[s_bin[lo].w + (tid - s_binRec[lo])], vpf, bx0, by0);
      }
      __syncthreads();
    }
    for (uint32_t bin = 0; bin < a.B && tileIdx < tileCount; bin++) {
      const uint4 tr = s_bin[bin];
      if (tr.y == 0u) continue;  // _getTileDepth == 0 means an empty bin (:215)
      const f2 td = mk2((float)(0xffffffffu - tr.x) / 4294967296.0f, (float)tr.y / 4294967296.0f);
      const uint32_t start = tr.w, n = tr.z;
      if (tid == 0) s_sliceOff = 0;
      if (whole) {
        __syncthreads();
        fine_cheap_batch<TILE>(a, L, gpool, s_binRec[bin], n, bx0, by0, td, invSegs, &s_sliceOff);
      }
      // Batches of a bin of several: each batch's points and tangents are
      // loaded while the previous batch is rasterized, and its keys one batch
      // before that, so a projection waits on no memory.
      SegIn cur;
      uint32_t nextKey = 0;
      if (!whole) {
        if (tid < min(a.batch, n)) seg_load(a, a.binKey[start + tid], cur);
        if (tid < a.batch && a.batch + tid < n) nextKey = a.binKey[start + a.batch + tid];
      }
      for (uint32_t b0 = 0; !whole && b0 < n; b0 += a.batch) {
        const uint32_t nb = min(a.batch, n - b0);
        __syncthreads();
        if (tid < nb)
          fine_record<TILE>(a, L, tid, cur.key,
                            project_segment_v(u, cur.p0, cur.t0, cur.p1, cur.t1, vpf, u->fiberRadius), bx0, by0);
        __syncthreads();
        if (tid < a.batch && b0 + a.batch + tid < n) seg_load(a, nextKey, cur);
        if (tid < a.batch && b0 + 2u * a.batch + tid < n) nextKey = a.binKey[start + b0 + 2u * a.batch + tid];
        fine_cheap_batch<TILE>(a, L, gpool, 0u, nb, bx0, by0, td, invSegs, &s_sliceOff);
        __syncthreads();
        if (s_sliceOff >= a.perProc) break;  // slice memory exhausted (:252-264)
      }
      __syncthreads();
      if (tid == 0 && a.exhausted && s_sliceOff >= a.perProc) atomicAdd(a.exhausted, 1u);
      // reduceHairSlices — reduceHairSlices.wgsl.ts:9-73
      if (!done) {
        if (!readable) fc = mk4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t cnt = isDbg ? f2u(fc.x * (float)dbgMax) : 0u;
        // only the slices with fragments: the others' heads are empty
        uint32_t sm = L.smask[tid];
        L.smask[tid] = 0u;
        while (sm != 0u) {
          const uint32_t s = (uint32_t)__builtin_ctz(sm);
          sm &= sm - 1u;
          uint32_t cur = L.heads[s * NPX + tid];
          L.heads[s * NPX + tid] = kInvalidPtr;
          if (!isDbg && fc.w >= 0.999f) continue;
          while (cur != kInvalidPtr) {
            if (!isDbg && fc.w >= 0.999f) break;
            uint2 col;
            if (cur < a.ldsPool) {
              col = L.poolColor[cur];
              cur = L.poolNext[cur];
            } else {
              const uint4 e = gpool[cur - a.ldsPool];
              col = make_uint2(e.x, e.y);
              cur = e.z;
            }
            cnt++;
            const float oneMinusA = 1.0f - fc.w;
            fc.x = fc.x + half_lo(col.x) * oneMinusA;
            fc.y = fc.y + half_hi(col.x) * oneMinusA;
            fc.z = fc.z + half_lo(col.y) * oneMinusA;
            fc.w = fc.w + half_hi(col.y) * oneMinusA;
          }
        }
        if (isDbg) {
          const float c = satf((float)cnt / (float)dbgMax);
          fc = mk4(c, 0.0f, 0.0f, 1.0f);
        }
        blended = true;
        done = !isDbg && fc.w >= 0.999f;
        L.done[tid] = done ? 1u : 0u;
      }
      // checkAllPixelsInWkgrpDone :305-327
      if (__syncthreads_and(done)) break;
    }
    // The pixel's value after its last blended bin. Stored once per tile: a
    // barrier waits for the workgroup's outstanding stores, so a store per
    // bin would put a memory round trip between bins.
    if (readable && blended) a.fb[fbIdx] = make_float4(fc.x, fc.y, fc.z, fc.w);
  }
}

template <uint32_t TILE>
struct FineRaster {
  using Args = FineArgs;
  static constexpr uint32_t kThreads = fine_waves<TILE>() * 64;
  static constexpr const char* kName = "fine_raster";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { fine_raster_body<TILE>(a, block, blocks); }
};

// swRasterizer.test.ts:184-248 — the reference's raster-helper test kernel.
__device__ __forceinline__ void debug_raster_body(const fh_render_uniforms* __restrict__ u,
                                                          const float4* __restrict__ pos,
                                                          const float4* __restrict__ tan, uint32_t S,
                                                          uint32_t P, uint32_t* __restrict__ out, uint32_t block, uint32_t blocks) {
  const uint32_t gid = block * blockDim.x + threadIdx.x;
  const uint32_t nSeg = S * (P - 1u);
  if (gid >= nSeg) return;
  const uint32_t strand = gid / (P - 1u), seg = gid - strand * (P - 1u);
  const f2 vp = mk2(u->viewport[0], u->viewport[1]);
  const uint32_t W = f2u(vp.x), H = f2u(vp.y);
  const Seg ps = project_segment(u, pos, tan, P, strand, seg, vp, u->fiberRadius);
  f2 bmax = vmax(vmax(ps.v00, ps.v01), vmax(ps.v10, ps.v11));
  f2 bmin = vmin(vmin(ps.v00, ps.v01), vmin(ps.v10, ps.v11));
  bmax = vmin(mk2(ceilf(bmax.x), ceilf(bmax.y)), vp);
  bmin = vmax(mk2(floorf(bmin.x), floorf(bmin.y)), mk2(0.0f, 0.0f));
  for (float y = bmin.y; y < bmax.y; y += 1.0f)
    for (float x = bmin.x; x < bmax.x; x += 1.0f) {
      const f2 p = mk2(x, y);
      if (!inside_quad(ps, p)) continue;
      const f2 w = interp_quad(ps, p);
      const uint32_t c0 = f2u(satf(w.x) * 255.0f), c1 = f2u(satf(w.y) * 255.0f);
      const uint32_t c2 = f2u(satf(0.1f) * 255.0f);
      const uint32_t value = 0xff000000u | c0 | (c1 << 8) | (c2 << 16);
      const uint32_t ix = f2u(x), iy = f2u(y);
      if (ix >= W || iy >= H) continue;
      const size_t idx = (size_t)(H - iy) * W + ix;
      if (idx >= (size_t)W * H) continue;
      atomicMax(&out[idx], value);
    }
}

struct DebugRaster {
  struct Args {
    const fh_render_uniforms* u;
    const float4* pos;
    const float4* tan;
    uint32_t S;
    uint32_t P;
    uint32_t* out;
  };
  static constexpr uint32_t kThreads = 64;
  static constexpr const char* kName = "debug_raster";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { debug_raster_body(a.u, a.pos, a.tan, a.S, a.P, a.out, block, blocks); }
};


// Test hook: for each quad, every pixel that passes the four edge tests in a
// window 3 px beyond its pixel box must lie in the box and in the x span of
// its row (the bounds the tile and fine passes enumerate). Writes per quad
// (pixels outside the box, pixels outside the span, candidates, inside).
__device__ __forceinline__ void pixel_bounds_body(const float* __restrict__ quads, uint32_t n,
                                                  uint4* __restrict__ out, uint32_t block, uint32_t blocks) {
  const uint32_t i = block * 64u + threadIdx.x;
  if (i >= n) return;
  const float* q = quads + (size_t)i * 8u;
  Seg s;
  s.v00 = mk2(q[0], q[1]); s.v01 = mk2(q[2], q[3]); s.v10 = mk2(q[4], q[5]); s.v11 = mk2(q[6], q[7]);
  s.depths = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  const float margin = raster_margin(s);
  int x0, y0, x1, y1;
  seg_pixel_bbox(s, margin, x0, y0, x1, y1);
  float amax = 0.0f;
  for (int k = 0; k < 8; k++) amax = fmaxf(amax, fabsf(q[k]));
  float sp[kSpFields];
  span_frame(s, x0 > -(1 << 20) && amax < 65536.0f, margin, sp, 1);
  uint4 r = make_uint4(0u, 0u, 0u, 0u);
  // quads whose box is the unbounded one (collinear) are not checked
  if (x0 <= -(1 << 20) || x1 - x0 > 200 || y1 - y0 > 200) { out[i] = make_uint4(0u, 0u, 0u, 0xffffffffu); return; }
  for (int y = y0 - 3; y <= y1 + 3; y++) {
    int xa = x0, xb = x1 + 1;
    if (sp[kSpOn] != 0.0f) {
      float xlo, xhi;
      span_x<1>(sp, (float)y, (float)y, xlo, xhi);
      if (xlo <= xhi) {
        xa = max(xa, (int)fmaxf(ceilf(xlo), -1.0e9f));
        xb = min(xb, (int)fminf(floorf(xhi), 1.0e9f) + 1);
      } else {
        xb = xa;
      }
    }
    const bool rowIn = y >= y0 && y <= y1;
    if (rowIn && xb > xa) r.z += (uint32_t)(xb - xa);
    for (int x = x0 - 3; x <= x1 + 3; x++) {
      if (!inside_quad(s, mk2((float)x, (float)y))) continue;
      r.w++;
      if (!rowIn || x < x0 || x > x1) r.x++;
      else if (x < xa || x >= xb) r.y++;
    }
  }
  out[i] = r;
}

struct PixelBounds {
  struct Args {
    const float* quads;
    uint32_t n;
    uint4* out;
  };
  static constexpr uint32_t kThreads = 64;
  static constexpr const char* kName = "pixel_bounds";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { pixel_bounds_body(a.quads, a.n, a.out, block, blocks); }
};

int check_cfg(const fh_render_cfg* cfg, const char* fn) {
  if (!cfg) return fail(FH_ERR_INVALID_ARG, std::string(fn) + ": null cfg");
  if (!(cfg->tileSize == 8 || cfg->tileSize == 16 || cfg->tileSize == 32))
    return fail(FH_ERR_UNSUPPORTED, std::string(fn) + ": tileSize must be 8, 16 or 32");
  if (cfg->tileDepthBins < 1 || cfg->tileDepthBins > 64)
    return fail(FH_ERR_INVALID_ARG, std::string(fn) + ": tileDepthBins must be in [1, 64]");
  if (cfg->slicesPerPixel < 1 || cfg->slicesPerPixel > 32)
    return fail(FH_ERR_INVALID_ARG, std::string(fn) + ": slicesPerPixel must be in [1, 32]");
  if (cfg->sortBuckets < 1 || cfg->sortBuckets > 64 || cfg->sortBucketSize < 1)
    return fail(FH_ERR_INVALID_ARG, std::string(fn) + ": sortBuckets must be in [1, 64], sortBucketSize >= 1");
  if (cfg->shadingPoints < 2 || cfg->shadingPoints > 1024)
    return fail(FH_ERR_INVALID_ARG, std::string(fn) + ": shadingPoints must be in [2, 1024]");
  return FH_OK;
}

}  // namespace

extern "C" {

int fh_tiles_clear(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                   fh_tile_result* tiles, void* tile_segments, uint32_t* segment_count_per_tile) {
  FH_CHECK_CTX(ctx);
  int rc = check_cfg(cfg, "fh_tiles_clear");
  if (rc) return rc;
  FH_REQUIRE(tiles && tile_segments && segment_count_per_tile, "fh_tiles_clear: null buffer");
  const uint64_t nt = fh_tile_count(cfg, vp_w, vp_h);
  // drawnSegmentsCount only of the segment buffer (hairTilesPass.ts:51)
  void* dst[3] = {tiles, segment_count_per_tile, tile_segments};
  const uint64_t n[3] = {nt * cfg->tileDepthBins * 4, nt, 1};
  return zero_many(ctx, 3, dst, n);
}

int fh_tiles_draw(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                  const fh_render_uniforms* uniforms, const fh_hair_data* hair_data,
                  const fh_vec4* positions, const fh_vec4* tangents, fh_tile_result* tiles,
                  fh_tex2d depth, void* tile_segments, uint32_t segments_capacity,
                  uint32_t* segment_count_per_tile, uint32_t* slot_tile_bin,
                  uint32_t rendered_strands, uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  int rc = check_cfg(cfg, "fh_tiles_draw");
  if (rc) return rc;
  FH_REQUIRE(uniforms && hair_data && positions && tangents && tiles && tile_segments &&
                 segment_count_per_tile && depth.data,
             "fh_tiles_draw: null buffer");
  FH_REQUIRE(points_per_strand >= 2 && points_per_strand <= 256, "fh_tiles_draw: points_per_strand in [2, 256]");
  FH_REQUIRE(rendered_strands < (1u << 24), "fh_tiles_draw: strand index must fit 24 bits (hairTileSegmentsBuffer.ts:26)");
  FH_REQUIRE(vp_w > 0 && vp_h > 0, "fh_tiles_draw: empty viewport");
  const uint64_t nSeg = (uint64_t)rendered_strands * (points_per_strand - 1u);
  if (nSeg == 0) return FH_OK;
  FH_REQUIRE(nSeg < (1ull << 31), "fh_tiles_draw: too many segments");
  TilesArgs a;
  a.u = uniforms; a.hd = hair_data;
  a.pos = reinterpret_cast<const float4*>(positions);
  a.tan = reinterpret_cast<const float4*>(tangents);
  a.tiles = tiles; a.depth = depth.data; a.depthW = depth.width; a.depthH = depth.height;
  a.segCounter = reinterpret_cast<uint32_t*>(tile_segments);
  a.segData = reinterpret_cast<fh_list_element*>(reinterpret_cast<char*>(tile_segments) + 4);
  a.limit = segments_capacity;  // the kernel also caps by uniforms->maxDrawnHairSegments
  a.slotKey = slot_tile_bin;
  a.T = cfg->tileSize; a.B = cfg->tileDepthBins; a.threshold = cfg->invalidTilesPerSegmentThreshold;
  a.S = rendered_strands; a.P = points_per_strand;
  // overflow queue in the context's scratch: room for every segment's pieces
  // up to 4 Mi pieces; a segment whose pieces do not fit is drawn by its wave
  const uint64_t maxTiles = std::min<uint64_t>(cfg->invalidTilesPerSegmentThreshold, fh_tile_count(cfg, vp_w, vp_h));
  a.pieceCap = (uint32_t)std::min<uint64_t>(nSeg * ((maxTiles + kTilePiece - 1) / kTilePiece), 4u << 20);
  const uint32_t nTiles = fh_tile_count(cfg, vp_w, vp_h);
  const size_t boundsOff = fh::align_up(256 + (size_t)a.pieceCap * 8, 256);
  if ((rc = ensure_scratch(ctx, boundsOff + (size_t)nTiles * sizeof(float2)))) return rc;
  a.pieceCount = reinterpret_cast<uint32_t*>(ctx->scratch);
  a.pieces = reinterpret_cast<uint2*>(reinterpret_cast<char*>(ctx->scratch) + 256);
  a.depthBounds = reinterpret_cast<const float2*>(reinterpret_cast<char*>(ctx->scratch) + boundsOff);
  a.nBounds = nTiles;
  // the depth-bounds pass also zeroes the piece counter (nTiles >= 1: it always runs)
  if ((rc = launch<TileDepthBounds>(ctx, fh::div_ceil(nTiles, TileDepthBounds::kWaves), 0,
                                    {uniforms, depth.data, depth.width, depth.height, cfg->tileSize, nTiles,
                                     const_cast<float2*>(a.depthBounds), a.pieceCount})))
    return rc;
  if ((rc = launch<TilesDraw>(ctx, fh::div_ceil((uint32_t)nSeg, 64), 0, a))) return rc;
  if ((rc = launch<TilesPieces>(ctx, std::min(fh::div_ceil((uint32_t)nSeg, 64), fh::persistent_blocks(ctx)), 0, a)))
    return rc;
  return launch<TileCounts>(ctx, fh::div_ceil(nTiles, 256), 0, {tiles, nTiles, cfg->tileDepthBins, segment_count_per_tile});
}

int fh_tiles_rebuild_slot_keys(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                               fh_tile_result* tiles, const void* tile_segments,
                               uint32_t segments_capacity, uint32_t* slot_tile_bin) {
  FH_CHECK_CTX(ctx);
  int rc = check_cfg(cfg, "fh_tiles_rebuild_slot_keys");
  if (rc) return rc;
  FH_REQUIRE(tiles && tile_segments && slot_tile_bin, "fh_tiles_rebuild_slot_keys: null buffer");
  const uint32_t nTB = fh_tile_count(cfg, vp_w, vp_h) * cfg->tileDepthBins;
  return launch<RebuildKeys>(
      ctx, fh::div_ceil(nTB, 256), 0,
      {tiles, reinterpret_cast<const fh_list_element*>(reinterpret_cast<const char*>(tile_segments) + 4),
       segments_capacity, nTB, slot_tile_bin});
}

int fh_tiles_sort_clear(fh_ctx* ctx, const fh_render_cfg* cfg, void* tile_list, fh_sort_bucket* buckets) {
  FH_CHECK_CTX(ctx);
  int rc = check_cfg(cfg, "fh_tiles_sort_clear");
  if (rc) return rc;
  FH_REQUIRE(tile_list && buckets, "fh_tiles_sort_clear: null buffer");
  void* dst[2] = {tile_list, buckets};
  const uint64_t n[2] = {1, (uint64_t)cfg->sortBuckets * 2};
  return zero_many(ctx, 2, dst, n);
}

int fh_tiles_sort(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                  const uint32_t* segment_count_per_tile, void* tile_list, fh_sort_bucket* buckets) {
  FH_CHECK_CTX(ctx);
  int rc = check_cfg(cfg, "fh_tiles_sort");
  if (rc) return rc;
  FH_REQUIRE(segment_count_per_tile && tile_list && buckets, "fh_tiles_sort: null buffer");
  const uint32_t nt = fh_tile_count(cfg, vp_w, vp_h);
  if (nt == 0) return FH_OK;
  return launch<SortTiles>(ctx, 1, 0,
                           {segment_count_per_tile, nt, cfg->sortBuckets, cfg->sortBucketSize, buckets,
                            reinterpret_cast<uint32_t*>(tile_list)});
}

int fh_fine_clear(fh_ctx* ctx, uint32_t vp_w, uint32_t vp_h, void* rasterizer_result) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(rasterizer_result, "fh_fine_clear: null buffer");
  return fill_u32(ctx, rasterizer_result, (uint64_t)vp_w * vp_h * 4, 0u);
}

int fh_fine_raster(fh_ctx* ctx, const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                   const fh_render_uniforms* uniforms, const fh_hair_data* hair_data,
                   const fh_vec4* positions, const fh_tile_result* tiles, const void* tile_segments,
                   uint32_t segments_capacity, void* slice_data, void* rasterizer_result,
                   fh_tex2d depth, const fh_vec4* shading, const fh_vec4* tangents,
                   const void* tile_list, const uint32_t* slot_tile_bin, uint32_t strands_count,
                   uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  (void)hair_data;
  int rc = check_cfg(cfg, "fh_fine_raster");
  if (rc) return rc;
  FH_REQUIRE(uniforms && positions && tiles && tile_segments && slice_data && rasterizer_result &&
                 depth.data && shading && tangents && tile_list && slot_tile_bin,
             "fh_fine_raster: null buffer");
  FH_REQUIRE(points_per_strand >= 2 && points_per_strand <= 256, "fh_fine_raster: points_per_strand in [2, 256]");
  FH_REQUIRE(cfg->processorCount >= 1, "fh_fine_raster: processorCount must be >= 1");
  FH_REQUIRE(vp_w > 0 && vp_h > 0, "fh_fine_raster: empty viewport");
  const uint32_t T = cfg->tileSize;
  const uint32_t nTB = fh_tile_count(cfg, vp_w, vp_h) * cfg->tileDepthBins;
  const uint32_t limit = segments_capacity;
  const uint32_t scanBlocks = fh::div_ceil(nTB, kScanBlock);
  // scratch: per-bin range ends and starts, per-block sums, gathered segment ids
  const size_t offEnds = 0;
  const size_t offStarts = fh::align_up(offEnds + (size_t)nTB * 4, 256);
  const size_t offSums = fh::align_up(offStarts + (size_t)nTB * 4, 256);
  const size_t offSeg = fh::align_up(offSums + (size_t)scanBlocks * 4, 256);
  const size_t offBig = fh::align_up(offSeg + (size_t)limit * 4 + 4, 256);
  const size_t offLarge = fh::align_up(offBig + (size_t)limit * 8 + 8, 256);
  const size_t total = fh::align_up(offLarge + ((size_t)nTB + 1) * 4, 256);
  if ((rc = ensure_scratch(ctx, total))) return rc;
  char* base = reinterpret_cast<char*>(ctx->scratch);
  uint32_t* ends = reinterpret_cast<uint32_t*>(base + offEnds);
  uint32_t* starts = reinterpret_cast<uint32_t*>(base + offStarts);
  uint32_t* sums = reinterpret_cast<uint32_t*>(base + offSums);
  uint32_t* binSeg = reinterpret_cast<uint32_t*>(base + offSeg);
  uint32_t* big = reinterpret_cast<uint32_t*>(base + offBig);
  uint32_t* large = reinterpret_cast<uint32_t*>(base + offLarge);
  const fh_list_element* data =
      reinterpret_cast<const fh_list_element*>(reinterpret_cast<const char*>(tile_segments) + 4);

  // the scan also zeroes the long-range counter of the bin sort (large[0])
  if ((rc = launch<BinScanLocal>(ctx, scanBlocks, 0, {tiles, nTB, ends, sums, large, ctx->stats}))) return rc;
  // up to kMaxScanBlocks scan blocks the scatter kernel adds their prefixes itself
  const bool addFirst = scanBlocks > kMaxScanBlocks;
  if (addFirst && (rc = launch<BinScanAdd>(ctx, scanBlocks, 0, {nTB, ends, sums}))) return rc;
  const uint32_t scatterBlocks =
      std::max(1u, std::min(fh::div_ceil(std::max(limit, nTB), 256), fh::persistent_blocks(ctx)));
  if ((rc = launch<BinScatter>(ctx, scatterBlocks, 0,
                               {uniforms, reinterpret_cast<const uint32_t*>(tile_segments), data, limit,
                                reinterpret_cast<const uint2*>(slot_tile_bin), ends, addFirst ? nullptr : sums,
                                scanBlocks, tiles, starts, nTB, binSeg, large})))
    return rc;
  const uint32_t largeBlocks = std::min(nTB, fh::persistent_blocks(ctx));
  if ((rc = launch<BinSort>(ctx, largeBlocks + fh::div_ceil(nTB, kSmallRanges), 0,
                            {tiles, starts, nTB, limit, binSeg, large, big, largeBlocks})))
    return rc;

  FineArgs a;
  a.u = uniforms;
  a.pos = reinterpret_cast<const float4*>(positions);
  a.tan = reinterpret_cast<const float4*>(tangents);
  a.tiles = tiles; a.binStart = starts; a.binKey = binSeg; a.capacity = limit;
  a.slicePool = reinterpret_cast<uint4*>(slice_data);
  a.queue = reinterpret_cast<uint32_t*>(rasterizer_result);
  a.fb = reinterpret_cast<float4*>(reinterpret_cast<char*>(rasterizer_result) + 16);
  a.depth = depth.data; a.depthW = depth.width; a.depthH = depth.height;
  a.shading = reinterpret_cast<const float4*>(shading);
  a.tileList = reinterpret_cast<const uint32_t*>(tile_list);
  a.S = strands_count; a.P = points_per_strand;
  a.B = cfg->tileDepthBins; a.NS = cfg->slicesPerPixel; a.PS = cfg->shadingPoints;
  a.perProc = fh_slice_data_per_processor(cfg);
  a.alphaQuadratic = cfg->alphaQuadratic; a.alphaMultiplier = cfg->alphaMultiplier;
  a.exhausted = ctx->stats;
  const uint32_t npx = T * T;
  a.batch = std::min(npx, 256u);
  // LDS: segment batch, slice heads, per-pixel state, fragment rings and
  // enumeration scratch, then the LDS part of the slice memory (12 B per
  // fragment). Under 80 KiB per workgroup keeps two 16x16 processors per CU
  // (160 KiB LDS on gfx950).
  const uint32_t nw = T == 8 ? 1u : T / 2u;
  const size_t ldsTarget = std::min<size_t>(T == 8 ? 40 * 1024 : (T == 16 ? 78 * 1024 : 150 * 1024),
                                            ctx->max_lds - 256);
  const size_t fixed = (size_t)a.batch * kRecFields * 4 + (size_t)npx * 4 * (a.NS + 3) +
                       (size_t)nw * (kFineRing + 6 * 64) * 4;
  FH_REQUIRE(fixed + 12 * 64 <= ldsTarget, "fh_fine_raster: slicesPerPixel too large for the tile size");
  a.ldsPool = (uint32_t)std::min<size_t>((ldsTarget - fixed) / 12, a.perProc);
  const size_t lds = fixed + (size_t)a.ldsPool * 12;
  const uint32_t grid = cfg->processorCount;
  if (T == 8) return launch<FineRaster<8>>(ctx, grid, (uint32_t)lds, a);
  if (T == 16) return launch<FineRaster<16>>(ctx, grid, (uint32_t)lds, a);
  return launch<FineRaster<32>>(ctx, grid, (uint32_t)lds, a);
}

// Test hook: the raster-helper kernel of swRasterizer.test.ts:165-250, which
// the reference pins with __test__/swRasterizer.snapshot.bin.
int fh_debug_raster_barycentric(fh_ctx* ctx, const fh_render_uniforms* uniforms,
                                const fh_vec4* positions, const fh_vec4* tangents,
                                uint32_t strands_count, uint32_t points_per_strand, uint32_t* out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(uniforms && positions && tangents && out, "fh_debug_raster_barycentric: null buffer");
  FH_REQUIRE(points_per_strand >= 2, "fh_debug_raster_barycentric: points_per_strand >= 2");
  const uint32_t nSeg = strands_count * (points_per_strand - 1u);
  if (nSeg == 0) return FH_OK;
  return launch<DebugRaster>(ctx, fh::div_ceil(nSeg, 64), 0,
                             {uniforms, reinterpret_cast<const float4*>(positions),
                              reinterpret_cast<const float4*>(tangents), strands_count, points_per_strand, out});
}

int fh_debug_pixel_bounds(fh_ctx* ctx, const float* quads, uint32_t count, uint32_t* out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(quads && out, "fh_debug_pixel_bounds: null buffer");
  if (count == 0) return FH_OK;
  return launch<PixelBounds>(ctx, fh::div_ceil(count, 64), 0, {quads, count, reinterpret_cast<uint4*>(out)});
}

int fh_fine_exhausted_bins(fh_ctx* ctx, uint32_t* out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(out != nullptr, "fh_fine_exhausted_bins: out is null");
  FH_HIP(hipMemcpyAsync(out, ctx->stats, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  FH_HIP(hipStreamSynchronize(ctx->stream));
  return FH_OK;
}

}  // extern "C"
