// Device-side f32 helpers shared by the hair kernels (gfx950).
//
// The reference's results depend on discrete decisions taken in f32: inside /
// outside of a rasterised quad, depth tests, tile and depth-bin indices, grid
// cells and the i32 fixed-point grid values. For those to be reproducible, the
// whole library is built with -ffp-contract=off and every helper below fixes
// one evaluation order:
//   mix(a, b, t)        = a * (1 - t) + b * t           (WGSL spec formula)
//   length(v)           = sqrt(x*x + y*y [+ z*z]), summed left to right
//   normalize(v)        = v / length(v), per component (IEEE division)
//   mat4 * vec4         = col0*x + col1*y + col2*z + col3*w, left to right
//   float -> int        = truncate toward zero, saturate, NaN -> 0
//   min / max           = IEEE minNum / maxNum (the non-NaN operand wins)
//   trilinear sample    = clamp-to-edge, texel centres at (i + 0.5) / n,
//                         lerp(a, b, f) = a + f * (b - a), x, then y, then z
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fh_hair.h"

namespace fh {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

__device__ __forceinline__ f2 mk2(float x, float y) { return {x, y}; }
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f4 mk4(float x, float y, float z, float w) { return {x, y, z, w}; }
__device__ __forceinline__ f3 xyz(f4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f4 ld4(const float4* p) { float4 v = *p; return {v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ f4 ld4(const fh_vec4* p) {
  float4 v = *reinterpret_cast<const float4*>(p);
  return {v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ float fmn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float satf(float x) { return clampf(x, 0.0f, 1.0f); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ uint32_t clampu(uint32_t x, uint32_t lo, uint32_t hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// WGSL u32(f32) / i32(f32): truncation with saturation, NaN -> 0.
__device__ __forceinline__ uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}
__device__ __forceinline__ int32_t f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

// vec2
__device__ __forceinline__ f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ f2 operator-(f2 a, f2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ f2 operator*(f2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ f2 operator/(f2 a, float s) { return {a.x / s, a.y / s}; }
__device__ __forceinline__ float dot(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float length(f2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ f2 normalize(f2 a) { float l = length(a); return {a.x / l, a.y / l}; }
__device__ __forceinline__ f2 vmin(f2 a, f2 b) { return {fmn(a.x, b.x), fmn(a.y, b.y)}; }
__device__ __forceinline__ f2 vmax(f2 a, f2 b) { return {fmx(a.x, b.x), fmx(a.y, b.y)}; }

// vec3
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ f3 normalize(f3 a) { float l = length(a); return {a.x / l, a.y / l, a.z / l}; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 mix(f3 a, f3 b, float t) {
  return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)};
}
__device__ __forceinline__ f3 sat(f3 a) { return {satf(a.x), satf(a.y), satf(a.z)}; }

// column-major mat4 (WGSL mat4x4<f32>) times vec4
__device__ __forceinline__ f4 mul(const float* m, f4 v) {
  f4 r;
  r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return r;
}
// projectVertex — shaderSnippets.wgls.ts:89-93
__device__ __forceinline__ f3 project(const float* m, f4 p) {
  f4 c = mul(m, p);
  return {c.x / c.w, c.y / c.w, c.z / c.w};
}

// Trilinear sample of a d^3 r32float volume (index (z*d + y)*d + x) at
// normalised coordinates t, linear filter, clamp-to-edge addressing.
__device__ __forceinline__ float sample_volume(const float* __restrict__ vol, uint32_t d, f3 t) {
  const float df = (float)d;
  const int dm1 = (int)d - 1;
  float cx = t.x * df - 0.5f, cy = t.y * df - 0.5f, cz = t.z * df - 0.5f;
  float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
  float fx = cx - flx, fy = cy - fly, fz = cz - flz;
  int ax = (int)flx, ay = (int)fly, az = (int)flz;
  int x0 = min(max(ax, 0), dm1), x1 = min(max(ax + 1, 0), dm1);
  int y0 = min(max(ay, 0), dm1), y1 = min(max(ay + 1, 0), dm1);
  int z0 = min(max(az, 0), dm1), z1 = min(max(az + 1, 0), dm1);
  const size_t sz0 = (size_t)z0 * d, sz1 = (size_t)z1 * d;
  const float* r00 = vol + (sz0 + y0) * d;
  const float* r01 = vol + (sz0 + y1) * d;
  const float* r10 = vol + (sz1 + y0) * d;
  const float* r11 = vol + (sz1 + y1) * d;
  float c000 = r00[x0], c100 = r00[x1], c010 = r01[x0], c110 = r01[x1];
  float c001 = r10[x0], c101 = r10[x1], c011 = r11[x0], c111 = r11[x1];
  float a0 = c000 + fx * (c100 - c000);
  float a1 = c010 + fx * (c110 - c010);
  float a2 = c001 + fx * (c101 - c001);
  float a3 = c011 + fx * (c111 - c011);
  float b0 = a0 + fy * (a1 - a0);
  float b1 = a2 + fy * (a3 - a2);
  return b0 + fz * (b1 - b0);
}

// 2D f32 texture load with zero for out-of-range texels.
__device__ __forceinline__ float tex_load(const float* __restrict__ t, uint32_t w, uint32_t h,
                                          int32_t x, int32_t y) {
  if (x < 0 || y < 0 || (uint32_t)x >= w || (uint32_t)y >= h) return 0.0f;
  return t[(size_t)y * w + (uint32_t)x];
}

// textureGather on a clamp-to-edge sampler, averaged as (x + y + z + w) / 4
// with components (i0,j1), (i1,j1), (i1,j0), (i0,j0).
__device__ __forceinline__ float gather_avg(const float* __restrict__ t, uint32_t w, uint32_t h, f2 uv) {
  const float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
  int i0 = (int)floorf(x), j0 = (int)floorf(y);
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = min(max(i0, 0), (int)w - 1); i1 = min(max(i1, 0), (int)w - 1);
  j0 = min(max(j0, 0), (int)h - 1); j1 = min(max(j1, 0), (int)h - 1);
  const float gx = t[(size_t)j1 * w + i0], gy = t[(size_t)j1 * w + i1];
  const float gz = t[(size_t)j0 * w + i1], gw = t[(size_t)j0 * w + i0];
  return (gx + gy + gz + gw) / 4.0f;
}

// IEEE binary16 pack / unpack, round to nearest even (pack2x16float).
__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
  _Float16 ha = (_Float16)a, hb = (_Float16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}
__device__ __forceinline__ float half_lo(uint32_t v) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t v) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
}

// number of set bits of `mask` below the calling lane
__device__ __forceinline__ uint32_t mask_prefix(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive prefix sum over the 64 lanes with DPP: row_shr:1,2,4,8
// within each row of 16 lanes, then row_bcast:15 (into rows 1 and 3) and
// row_bcast:31 (into rows 2 and 3). GFX9 DPP controls: row_shr:n = 0x110 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143. A lane whose source is outside
// its row keeps `old`, the identity.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  return (uint32_t)x;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Inclusive prefix max over the 64 lanes with DPP, the same steps as
// wave_incl_scan.
__device__ __forceinline__ int wave_incl_max(int x) {
  constexpr int lo = -2147483647 - 1;
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x143, 0xc, 0xf, false));
  return x;
}

__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Load-balancing search: lane s owns the items [end_s - cnt_s, end_s) of a
// sequence (ends from wave_incl_scan of cnt). Returns the lane owning item
// c0 + lane, valid while c0 + lane < the last end. Every lane whose range
// meets [c0, c0 + 64) marks its first item there (distinct positions, since
// the ranges are disjoint), then a prefix max spreads the marks. `mark` is 64
// ints of wave-private LDS.
__device__ __forceinline__ uint32_t lbs_owner(int* mark, uint32_t end, uint32_t cnt, uint32_t c0) {
  const uint32_t lane = __lane_id();
  mark[lane] = -1;
  wave_sync();
  const uint32_t start = end - cnt;
  if (cnt && end > c0 && start < c0 + 64u) mark[(start > c0 ? start : c0) - c0] = (int)lane;
  wave_sync();
  return (uint32_t)wave_incl_max(mark[lane]);
}

}  // namespace fh
