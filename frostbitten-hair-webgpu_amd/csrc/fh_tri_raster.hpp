// The software triangle rasterizer shared by the passes that the reference
// draws with the hardware rasterizer: the hair's depth and normals
// (HwHairPass), the hair and the meshes in the shadow map (ShadowMapPass) and
// the meshes' depth, normals and colour (DrawMeshesPass).
//
// Rules. Vertices are snapped to 1/256 px; a pixel centre is covered when the
// three integer edge functions are positive, or zero on a top or left edge, so
// triangles sharing an edge cover each pixel once. Depth is the screen-space
// barycentric interpolation of NDC z; fragments outside [0, 1] are clipped,
// triangles with a vertex at w <= 0 or beyond 2^20 px are not drawn. The depth
// test is 'less' (shared.ts:22-26) in primitive order: each fragment does one
// 64-bit atomicMin of (depth bits, triangle index) into a per-pixel key, so
// among a pixel's fragments the least depth wins, the earliest triangle on a
// tie, whatever order the GPU runs the triangles in; a resolve pass then
// computes the winner's attributes.
#pragma once

#include "fh_device.hpp"

namespace fh {
namespace tri {

constexpr unsigned long long kNoFragment = ~0ull;

struct SetupTri {
  f4 c[3];           // clip positions, counter-clockwise order after setup
  long long X[3], Y[3];
  long long area;
  uint32_t vi[3];    // original vertex slot of each setup vertex
  bool ok;
  bool swapped;      // the snapped vertices were clockwise on screen (y down)
};

// Snap to 1/256 px (framebuffer coordinates, y down) and orient so that the
// edge functions are positive inside.
__device__ __forceinline__ SetupTri setup_from(uint32_t W, uint32_t H, f4 c0, f4 c1, f4 c2) {
  SetupTri s;
  s.c[0] = c0; s.c[1] = c1; s.c[2] = c2;
  s.ok = true;
  s.swapped = false;
  for (int k = 0; k < 3; k++) {
    s.vi[k] = (uint32_t)k;
    const f4 c = s.c[k];
    s.X[k] = s.Y[k] = 0;
    if (!(c.w > 0.0f)) { s.ok = false; continue; }
    const float nx = c.x / c.w, ny = c.y / c.w;
    const float xs = (float)W * (nx * 0.5f + 0.5f);
    const float ys = (float)H * (1.0f - (ny * 0.5f + 0.5f));
    if (!(fabsf(xs) < 1048576.0f && fabsf(ys) < 1048576.0f)) { s.ok = false; continue; }
    s.X[k] = (long long)rintf(xs * 256.0f);
    s.Y[k] = (long long)rintf(ys * 256.0f);
  }
  s.area = 0;
  if (!s.ok) return s;
  long long area = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
  if (area < 0) {
    const f4 c = s.c[1]; s.c[1] = s.c[2]; s.c[2] = c;
    long long q = s.X[1]; s.X[1] = s.X[2]; s.X[2] = q;
    q = s.Y[1]; s.Y[1] = s.Y[2]; s.Y[2] = q;
    const uint32_t v = s.vi[1]; s.vi[1] = s.vi[2]; s.vi[2] = v;
    area = -area;
    s.swapped = true;
  }
  s.area = area;
  s.ok = area > 0;
  return s;
}

__device__ __forceinline__ long long edge_fn_i(long long ax, long long ay, long long bx, long long by, long long px,
                                               long long py) {
  return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}
// top edge (horizontal, interior below) or left edge (interior to the right)
__device__ __forceinline__ bool top_left(long long ax, long long ay, long long bx, long long by) {
  const long long dx = bx - ax, dy = by - ay;
  return (dy == 0 && dx > 0) || dy < 0;
}
__device__ __forceinline__ bool covers(long long e, bool tl) { return e > 0 || (e == 0 && tl); }

// screen-space barycentrics of a covered pixel centre
__device__ __forceinline__ void barycentrics(const SetupTri& s, long long px, long long py, float* l) {
  const long long e0 = edge_fn_i(s.X[1], s.Y[1], s.X[2], s.Y[2], px, py);
  const long long e1 = edge_fn_i(s.X[2], s.Y[2], s.X[0], s.Y[0], px, py);
  const long long e2 = edge_fn_i(s.X[0], s.Y[0], s.X[1], s.Y[1], px, py);
  const float ar = (float)s.area;
  l[0] = (float)e0 / ar;
  l[1] = (float)e1 / ar;
  l[2] = (float)e2 / ar;
}

__device__ __forceinline__ float frag_depth(const SetupTri& s, const float* l) {
  const float z0 = s.c[0].z / s.c[0].w, z1 = s.c[1].z / s.c[1].w, z2 = s.c[2].z / s.c[2].w;
  return l[0] * z0 + l[1] * z1 + l[2] * z2;
}

// Perspective-correct weights of the setup vertices: l_q / w_q, and their sum.
__device__ __forceinline__ float persp_weights(const SetupTri& s, const float* l, float* wq) {
  float den = 0.0f;
  for (int q = 0; q < 3; q++) {
    wq[q] = l[q] / s.c[q].w;
    den = den + wq[q];
  }
  return den;
}

__device__ __forceinline__ long long floor_div256(long long v) { return v >= 0 ? v / 256 : -((-v + 255) / 256); }

// Conservative pixel range [x0, x1] of row centre cy inside edge a -> b
// (E(cx) = dx (cy - ay) - dy (cx - ax) >= 0), two pixels wider on each side
// than the bound evaluated in f32 (whose error is far below a pixel); the
// caller still tests each pixel exactly.
__device__ __forceinline__ void edge_span(long long ax, long long ay, long long bx, long long by, long long cy,
                                          long long& x0, long long& x1) {
  const long long dx = bx - ax, dy = by - ay;
  const long long c = dx * (cy - ay) + dy * ax;  // E(cx) = c - dy * cx
  if (dy == 0) {
    if (c < 0) { x0 = 1; x1 = 0; }
    return;
  }
  const float cxb = (float)c * __builtin_amdgcn_rcpf((float)dy);  // the cx where E = 0
  const float pxb = fminf(fmaxf((cxb - 128.0f) * (1.0f / 256.0f), -1.0e9f), 1.0e9f);
  if (dy < 0) x0 = max(x0, (long long)floorf(pxb) - 2);  // E grows with cx
  else x1 = min(x1, (long long)floorf(pxb) + 2);          // E falls with cx
}

// Pixel box of a set-up triangle: centres 256 p + 128 inside its snapped
// bounds, clipped to the W x H target (empty when x0 > x1 or y0 > y1).
struct PixBox {
  long long x0, x1, y0, y1;
};
__device__ __forceinline__ PixBox pix_box(uint32_t W, uint32_t H, const SetupTri& s) {
  const long long minX = min(s.X[0], min(s.X[1], s.X[2])), maxX = max(s.X[0], max(s.X[1], s.X[2]));
  const long long minY = min(s.Y[0], min(s.Y[1], s.Y[2])), maxY = max(s.Y[0], max(s.Y[1], s.Y[2]));
  PixBox b;
  b.x0 = max(0ll, -floor_div256(128 - minX));
  b.x1 = min((long long)W - 1, floor_div256(maxX - 128));
  b.y0 = max(0ll, -floor_div256(128 - minY));
  b.y1 = min((long long)H - 1, floor_div256(maxY - 128));
  return b;
}

// The fragments of triangle t on row py between pixels xa and xb: a
// conservative span from the three edge functions, then the exact edge tests
// stepped along the row, the depth clip, the 'less' test against `depthIn`
// (the depth already in the texture; null: the pass cleared it to 1.0) and
// one atomicMin per fragment into `key` (W keys per row).
__device__ __forceinline__ void raster_row(const SetupTri& s, uint32_t t, long long py, long long xa, long long xb,
                                           uint32_t W, const float* depthIn, unsigned long long* key) {
  const long long cy = py * 256 + 128;
  edge_span(s.X[1], s.Y[1], s.X[2], s.Y[2], cy, xa, xb);
  edge_span(s.X[2], s.Y[2], s.X[0], s.Y[0], cy, xa, xb);
  edge_span(s.X[0], s.Y[0], s.X[1], s.Y[1], cy, xa, xb);
  if (xa > xb) return;
  const bool tl0 = top_left(s.X[1], s.Y[1], s.X[2], s.Y[2]);
  const bool tl1 = top_left(s.X[2], s.Y[2], s.X[0], s.Y[0]);
  const bool tl2 = top_left(s.X[0], s.Y[0], s.X[1], s.Y[1]);
  // edge values at the first centre, then one step of -256 dy per pixel
  const long long cx0 = xa * 256 + 128;
  long long e0 = edge_fn_i(s.X[1], s.Y[1], s.X[2], s.Y[2], cx0, cy);
  long long e1 = edge_fn_i(s.X[2], s.Y[2], s.X[0], s.Y[0], cx0, cy);
  long long e2 = edge_fn_i(s.X[0], s.Y[0], s.X[1], s.Y[1], cx0, cy);
  const long long d0 = -256 * (s.Y[2] - s.Y[1]), d1 = -256 * (s.Y[0] - s.Y[2]), d2 = -256 * (s.Y[1] - s.Y[0]);
  for (long long px = xa; px <= xb; px++, e0 += d0, e1 += d1, e2 += d2) {
    if (!(covers(e0, tl0) && covers(e1, tl1) && covers(e2, tl2))) continue;
    float l[3];
    barycentrics(s, px * 256 + 128, cy, l);
    float z = frag_depth(s, l);
    if (!(z >= 0.0f && z <= 1.0f)) continue;  // depth clip
    z = z + 0.0f;                              // -0 -> +0
    const size_t p = (size_t)py * W + (size_t)px;
    const float dIn = depthIn ? depthIn[p] : 1.0f;
    if (!(z < dIn)) continue;                  // 'less'
    atomicMin(&key[p], ((unsigned long long)__float_as_uint(z) << 32) | t);
  }
}

__device__ __forceinline__ void raster_rows(const SetupTri& s, uint32_t t, long long first, long long stride,
                                            uint32_t W, uint32_t H, const float* depthIn, unsigned long long* key) {
  const PixBox b = pix_box(W, H, s);
  for (long long py = b.y0 + first; py <= b.y1; py += stride) raster_row(s, t, py, b.x0, b.x1, W, depthIn, key);
}

// kRowSplit threads per triangle, thread k taking rows k, k + kRowSplit, ...
// (a tall triangle does not hold one lane while the rest of its wave idles).
// A triangle whose pixel box is larger than kWaveRaster pixels is queued in
// bands of kBandRows rows, each drawn by a whole workgroup of a second
// kernel; when the queue is full, by the triangle's own wave, one row per
// lane.
constexpr uint32_t kRowSplit = 8;
constexpr long long kWaveRaster = 1024;
constexpr uint32_t kBandRows = 64;
constexpr uint32_t kBigCap = 1u << 20;          // queued bands
constexpr uint32_t kBigSentinel = 0xffffffffu;  // an unused band slot

// encodeOctahedronNormal (shaderSnippets.wgls.ts:25-46), packed as rg16float
__device__ __forceinline__ uint32_t encode_oct(f3 n0) {
  const float s = fabsf(n0.x) + fabsf(n0.y) + fabsf(n0.z);
  f3 n = mk3(n0.x / s, n0.y / s, n0.z / s);
  if (n.z < 0.0f) {
    const float sx = n.x >= 0.0f ? 1.0f : -1.0f, sy = n.y >= 0.0f ? 1.0f : -1.0f;
    const float ax = (1.0f - fabsf(n.y)) * sx, ay = (1.0f - fabsf(n.x)) * sy;
    n.x = ax;
    n.y = ay;
  }
  return pack_half2(n.x * 0.5f + 0.5f, n.y * 0.5f + 0.5f);
}

__device__ __forceinline__ f3 decode_oct(uint32_t packed) {
  const float fx = half_lo(packed) * 2.0f - 1.0f, fy = half_hi(packed) * 2.0f - 1.0f;
  f3 n = mk3(fx, fy, 1.0f - fabsf(fx) - fabsf(fy));
  const float t = satf(-n.z);
  n.x = n.x >= 0.0f ? n.x - t : n.x + t;
  n.y = n.y >= 0.0f ? n.y - t : n.y + t;
  return normalize(n);
}

}  // namespace tri
}  // namespace fh
