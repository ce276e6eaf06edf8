// Raster helpers of the software hair rasterizer, shared by the tile pass,
// the fine pass and the raster test kernel.
// Reference: src/passes/swHair/shaderImpl/swRasterizeHair.wgsl.ts,
// src/passes/_shaderSnippets/shaderSnippets.wgls.ts, tileUtils.wgsl.ts.
#pragma once

#include "fh_device.hpp"

namespace fh {

constexpr uint32_t kInvalidPtr = 0xffffffffu;

struct Seg {
  f2 v00, v01, v10, v11;
  f4 depths;
};

// projectHairPoint — swRasterizeHair.wgsl.ts:76-105. The tangent is
// transformed with w = 1, as in the reference.
__device__ __forceinline__ void project_point_v(const fh_render_uniforms* __restrict__ u, float4 pk, float4 tk,
                                                f2 vp, float fiberRadius, f2& v0, f2& v1, f2& dd) {
  const f4 p0 = mul(u->modelMatrix, mk4(pk.x, pk.y, pk.z, 1.0f));
  const f4 t0 = mul(u->modelMatrix, mk4(tk.x, tk.y, tk.z, 1.0f));
  const f3 cam = mk3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]);
  const f3 toCam = normalize(cam - xyz(p0));
  const f3 right = normalize(cross(xyz(t0), toCam)) * fiberRadius;
  const f3 a = xyz(p0) - right, b = xyz(p0) + right;
  const f3 an = project(u->vpMatrix, mk4(a.x, a.y, a.z, 1.0f));
  const f3 bn = project(u->vpMatrix, mk4(b.x, b.y, b.z, 1.0f));
  // ndc2viewportPx — shaderSnippets.wgls.ts:104-107
  v0 = mk2((an.x * 0.5f + 0.5f) * vp.x, (an.y * 0.5f + 0.5f) * vp.y);
  v1 = mk2((bn.x * 0.5f + 0.5f) * vp.x, (bn.y * 0.5f + 0.5f) * vp.y);
  dd = mk2(an.z, bn.z);
}

__device__ __forceinline__ void project_point(const fh_render_uniforms* __restrict__ u,
                                              const float4* __restrict__ pos,
                                              const float4* __restrict__ tan, size_t k, f2 vp,
                                              float fiberRadius, f2& v0, f2& v1, f2& dd) {
  project_point_v(u, pos[k], tan[k], vp, fiberRadius, v0, v1, dd);
}

// projectHairSegment — swRasterizeHair.wgsl.ts:39-68
__device__ __forceinline__ Seg project_segment(const fh_render_uniforms* __restrict__ u,
                                               const float4* __restrict__ pos,
                                               const float4* __restrict__ tan, uint32_t P,
                                               uint32_t strand, uint32_t seg, f2 vp,
                                               float fiberRadius) {
  Seg r;
  f2 a, b, dd;
  const size_t k = (size_t)strand * P + seg;
  project_point(u, pos, tan, k, vp, fiberRadius, a, b, dd);
  r.v00 = a; r.v01 = b; r.depths.x = dd.x; r.depths.y = dd.y;
  project_point(u, pos, tan, k + 1, vp, fiberRadius, a, b, dd);
  r.v10 = a; r.v11 = b; r.depths.z = dd.x; r.depths.w = dd.y;
  return r;
}

// projectHairSegment from the segment's two points and tangents, loaded by
// the caller
__device__ __forceinline__ Seg project_segment_v(const fh_render_uniforms* __restrict__ u, float4 p0, float4 t0,
                                                 float4 p1, float4 t1, f2 vp, float fiberRadius) {
  Seg r;
  f2 a, b, dd;
  project_point_v(u, p0, t0, vp, fiberRadius, a, b, dd);
  r.v00 = a; r.v01 = b; r.depths.x = dd.x; r.depths.y = dd.y;
  project_point_v(u, p1, t1, vp, fiberRadius, a, b, dd);
  r.v10 = a; r.v11 = b; r.depths.z = dd.x; r.depths.w = dd.y;
  return r;
}

// edgeFunction — swRasterizeHair.wgsl.ts:127-129
__device__ __forceinline__ float edge_fn(f2 v0, f2 v1, f2 p) {
  return (p.x - v0.x) * (v1.y - v0.y) - (p.y - v0.y) * (v1.x - v0.x);
}

__device__ __forceinline__ bool inside_quad(const Seg& s, f2 p) {
  const float c0 = edge_fn(s.v01, s.v00, p);
  const float c1 = edge_fn(s.v11, s.v01, p);
  const float c2 = edge_fn(s.v10, s.v11, p);
  const float c3 = edge_fn(s.v00, s.v10, p);
  return c0 >= 0.0f && c1 >= 0.0f && c2 >= 0.0f && c3 >= 0.0f;
}

// projectPointToLine — shaderSnippets.wgls.ts:95-102
__device__ __forceinline__ f2 project_to_line(f2 l1, f2 l2, f2 p) {
  const f2 ab = l2 - l1, ac = p - l1;
  return l1 + (ab * dot(ab, ac)) / dot(ab, ab);
}

// interpolateHairQuad — swRasterizeHair.wgsl.ts:170-197
__device__ __forceinline__ f2 interp_quad(const Seg& s, f2 c) {
  const f2 sm = (s.v00 + s.v01) / 2.0f;
  const f2 em = (s.v10 + s.v11) / 2.0f;
  const f2 cp = project_to_line(sm, em, c);
  const float d1 = length(cp - sm) / length(sm - em);
  const float ws = length(s.v00 - s.v01);
  const float we = ws * dot(normalize(s.v00 - s.v01), normalize(s.v10 - s.v11));
  const float ew = mixf(ws, we, d1);
  const f2 e1 = project_to_line(s.v00, s.v10, c);
  const float d0 = length(c - e1) / ew;
  return mk2(satf(d0), satf(d1));
}

// interp_quad split into its per-segment terms (computed once) and its
// per-pixel terms: the same operations in the same order, so the results are
// bit-identical.
enum InterpField : uint32_t {
  kIpSmx = 0, kIpSmy, kIpAb1x, kIpAb1y, kIpQ1, kIpL1, kIpWs, kIpWe, kIpAb2x, kIpAb2y, kIpQ2, kIpFields
};

__device__ __forceinline__ void interp_pre(const Seg& s, float* f, uint32_t stride) {
  const f2 sm = (s.v00 + s.v01) / 2.0f;
  const f2 em = (s.v10 + s.v11) / 2.0f;
  const f2 ab1 = em - sm;
  const float ws = length(s.v00 - s.v01);
  const float we = ws * dot(normalize(s.v00 - s.v01), normalize(s.v10 - s.v11));
  const f2 ab2 = s.v10 - s.v00;
  f[kIpSmx * stride] = sm.x;
  f[kIpSmy * stride] = sm.y;
  f[kIpAb1x * stride] = ab1.x;
  f[kIpAb1y * stride] = ab1.y;
  f[kIpQ1 * stride] = dot(ab1, ab1);
  f[kIpL1 * stride] = length(sm - em);
  f[kIpWs * stride] = ws;
  f[kIpWe * stride] = we;
  f[kIpAb2x * stride] = ab2.x;
  f[kIpAb2y * stride] = ab2.y;
  f[kIpQ2 * stride] = dot(ab2, ab2);
}

template <uint32_t STRIDE>
__device__ __forceinline__ f2 interp_quad_pre(const float* f, f2 v00, f2 c) {
  const f2 sm = mk2(f[kIpSmx * STRIDE], f[kIpSmy * STRIDE]);
  const f2 ab1 = mk2(f[kIpAb1x * STRIDE], f[kIpAb1y * STRIDE]);
  const f2 cp = sm + (ab1 * dot(ab1, c - sm)) / f[kIpQ1 * STRIDE];
  const float d1 = length(cp - sm) / f[kIpL1 * STRIDE];
  const float ew = mixf(f[kIpWs * STRIDE], f[kIpWe * STRIDE], d1);
  const f2 ab2 = mk2(f[kIpAb2x * STRIDE], f[kIpAb2y * STRIDE]);
  const f2 e1 = v00 + (ab2 * dot(ab2, c - v00)) / f[kIpQ2 * STRIDE];
  const float d0 = length(c - e1) / ew;
  return mk2(satf(d0), satf(d1));
}

// interpolateHairF32 — swRasterizeHair.wgsl.ts:199-203
__device__ __forceinline__ float interp_f32(f2 w, f4 v) {
  const float a = mixf(v.x, v.y, w.x);
  const float b = mixf(v.z, v.w, w.x);
  return mixf(a, b, w.y);
}

// getDepthBin — shaderSnippets.wgls.ts:148-156
__device__ __forceinline__ uint32_t depth_bin(uint32_t n, f2 td, float d) {
  const float span = fabsf(td.y - td.x);
  const float t = (d - td.x) / span;
  return clampu(f2u(t * (float)n), 0u, n - 1u);
}

// How far beyond the convex hull of a quad's vertices a pixel may still pass
// the four f32 edge tests. Each test rounds a difference of two products of
// differences: an error below 4 units of 2^-24 of |p - a| |b - a| per product,
// i.e. a distance error below 6e-7 |p - a| from the edge's line; for pixels
// within the quad's box |p - a| is at most twice the coordinates' magnitude
// M, so the error is below 2^-19.7 M. Near a convex corner of interior angle
// t the lines' errors add up to at most 1 / sin(t / 2) times that beyond the
// vertex. For a convex quad whose corners all have |sin t| >= 1/4 (t between
// 14.5 and 165.5 degrees) the margin, 1/64 px plus 2^-16 M, is over twice the
// worst case; any other quad (concave, self-intersecting, a sharp corner, a
// NaN or infinite vertex, or M so large that the margin would exceed it)
// keeps a whole pixel.
__device__ __forceinline__ float raster_margin(const Seg& s) {
  const float mag = fmaxf(fmaxf(fabsf(s.v00.x) + fabsf(s.v00.y), fabsf(s.v01.x) + fabsf(s.v01.y)),
                          fmaxf(fabsf(s.v10.x) + fabsf(s.v10.y), fabsf(s.v11.x) + fabsf(s.v11.y)));
  // the loop v00 -> v10 -> v11 -> v01 of the edge tests
  const f2 e0 = s.v10 - s.v00, e1 = s.v11 - s.v10, e2 = s.v01 - s.v11, e3 = s.v00 - s.v01;
  const float c0 = e3.x * e0.y - e3.y * e0.x, c1 = e0.x * e1.y - e0.y * e1.x;
  const float c2 = e1.x * e2.y - e1.y * e2.x, c3 = e2.x * e3.y - e2.y * e3.x;
  const float l0 = dot(e0, e0), l1 = dot(e1, e1), l2 = dot(e2, e2), l3 = dot(e3, e3);
  const bool convex = (c0 > 0.0f && c1 > 0.0f && c2 > 0.0f && c3 > 0.0f) ||
                      (c0 < 0.0f && c1 < 0.0f && c2 < 0.0f && c3 < 0.0f);
  const bool open = 16.0f * c0 * c0 >= l3 * l0 && 16.0f * c1 * c1 >= l0 * l1 &&
                    16.0f * c2 * c2 >= l1 * l2 && 16.0f * c3 * c3 >= l2 * l3;
  const float m = 1.0f / 64.0f + mag * (1.0f / 65536.0f);
  return (convex && open && m < 1.0f) ? m : 1.0f;
}

// Pixel bounding box of a projected quad (getRasterizedHairBounds before the
// scissor), as integers clamped to a safe range. A pixel can only pass all
// four edge tests if it lies within `margin` (raster_margin) of the box of
// the vertices.
//
// Why: the four edge functions E_i(p) are the orientations of the triangles
// (p, a_i, b_i) over the closed loop v00 -> v10 -> v11 -> v01. For p outside
// the convex hull of the vertices, all of them are seen from p within an open
// half-turn, so the signed angles swept along the loop sum to zero; all E_i >= 0
// then forces every E_i = 0, i.e. p and the four vertices collinear. Any quad
// (convex, concave or self-intersecting) thus has its inside pixels in the
// hull; the margin absorbs f32 rounding near the hull. An exactly collinear
// quad gets the unbounded box (the caller tests every pixel).
__device__ __forceinline__ void seg_pixel_bbox(const Seg& s, float margin, int& x0, int& y0, int& x1, int& y1) {
  const f2 mx = vmax(vmax(s.v00, s.v01), vmax(s.v10, s.v11));
  const f2 mn = vmin(vmin(s.v00, s.v01), vmin(s.v10, s.v11));
  const float lo = -1048576.0f, hi = 1048576.0f;
  x0 = (int)clampf(ceilf(mn.x - margin), lo, hi);
  y0 = (int)clampf(ceilf(mn.y - margin), lo, hi);
  x1 = (int)clampf(floorf(mx.x + margin), lo, hi);
  y1 = (int)clampf(floorf(mx.y + margin), lo, hi);
  const f2 e1 = s.v10 - s.v00, e2 = s.v11 - s.v00, e3 = s.v01 - s.v00;
  const float k1 = e1.x * e2.y - e1.y * e2.x;
  const float k2 = e1.x * e3.y - e1.y * e3.x;
  const float k3 = e2.x * e3.y - e2.y * e3.x;
  if (k1 == 0.0f && k2 == 0.0f && k3 == 0.0f) {
    x0 = y0 = -(1 << 21);
    x1 = y1 = 1 << 21;
  }
  // all vertices NaN: no pixel passes an edge test (a NaN vertex fails the
  // two edges it belongs to, so partly-NaN quads are empty as well)
  if (!(mx.x == mx.x && mx.y == mx.y && mn.x == mn.x && mn.y == mn.y)) { x0 = 1; x1 = 0; }
}

}  // namespace fh
