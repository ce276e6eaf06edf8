/*
 * Test infrastructure, not product code: scalar f32 helpers of the CPU oracle.
 *
 * They restate the WGSL built-ins the reference's shaders use, with one fixed
 * evaluation order per built-in (the WGSL spec leaves the order, and FMA
 * contraction, to the implementation). The oracle is compiled with
 * -ffp-contract=off so every product and sum below rounds to f32 on its own.
 *
 *  mix(a, b, t)         a * (1 - t) + b * t             (WGSL spec formula)
 *  length(v)            sqrt(x*x + y*y [+ z*z])         summed left to right
 *  normalize(v)         v / length(v), per component
 *  mat4 * vec4          col0*x + col1*y + col2*z + col3*w, left to right
 *  u32(f), i32(f)       truncate toward zero, saturate to the type's range,
 *                       NaN -> 0 (WGSL float->int conversion rule)
 *  min/max              the non-NaN operand wins (IEEE minNum/maxNum)
 *  textureSampleLevel   linear filter, clamp-to-edge, texel centres at
 *                       (i + 0.5) / size, lerp(a, b, f) = a + f * (b - a),
 *                       x first, then y, then z
 */
#ifndef FH_ORACLE_MATH_H
#define FH_ORACLE_MATH_H

#include <math.h>
#include <stdint.h>

typedef struct { float x, y; } ov2;
typedef struct { float x, y, z; } ov3;
typedef struct { float x, y, z, w; } ov4;

static inline ov2 o2(float x, float y) { ov2 r = {x, y}; return r; }
static inline ov3 o3(float x, float y, float z) { ov3 r = {x, y, z}; return r; }
static inline ov4 o4(float x, float y, float z, float w) { ov4 r = {x, y, z, w}; return r; }

static inline float omin(float a, float b) { return fminf(a, b); }
static inline float omax(float a, float b) { return fmaxf(a, b); }
static inline float oclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float osat(float x) { return oclamp(x, 0.0f, 1.0f); }
static inline float omix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float ofract(float x) { return x - floorf(x); }

static inline uint32_t ou32(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}
static inline int32_t oi32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
static inline uint32_t oclampu(uint32_t x, uint32_t lo, uint32_t hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

/* vec2 */
static inline ov2 add2(ov2 a, ov2 b) { return o2(a.x + b.x, a.y + b.y); }
static inline ov2 sub2(ov2 a, ov2 b) { return o2(a.x - b.x, a.y - b.y); }
static inline ov2 muls2(ov2 a, float s) { return o2(a.x * s, a.y * s); }
static inline ov2 divs2(ov2 a, float s) { return o2(a.x / s, a.y / s); }
static inline float dot2(ov2 a, ov2 b) { return a.x * b.x + a.y * b.y; }
static inline float len2(ov2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static inline ov2 norm2(ov2 a) { float l = len2(a); return o2(a.x / l, a.y / l); }
static inline ov2 min2(ov2 a, ov2 b) { return o2(omin(a.x, b.x), omin(a.y, b.y)); }
static inline ov2 max2(ov2 a, ov2 b) { return o2(omax(a.x, b.x), omax(a.y, b.y)); }

/* vec3 */
static inline ov3 add3(ov3 a, ov3 b) { return o3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline ov3 sub3(ov3 a, ov3 b) { return o3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline ov3 mul3(ov3 a, ov3 b) { return o3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline ov3 div3(ov3 a, ov3 b) { return o3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline ov3 muls3(ov3 a, float s) { return o3(a.x * s, a.y * s, a.z * s); }
static inline ov3 divs3(ov3 a, float s) { return o3(a.x / s, a.y / s, a.z / s); }
static inline float dot3(ov3 a, ov3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len3(ov3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline ov3 norm3(ov3 a) { float l = len3(a); return o3(a.x / l, a.y / l, a.z / l); }
static inline ov3 cross3(ov3 a, ov3 b) {
  return o3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline ov3 mix3(ov3 a, ov3 b, float t) {
  return o3(omix(a.x, b.x, t), omix(a.y, b.y, t), omix(a.z, b.z, t));
}
static inline ov3 sat3(ov3 a) { return o3(osat(a.x), osat(a.y), osat(a.z)); }
static inline ov3 xyz(ov4 a) { return o3(a.x, a.y, a.z); }

/* column-major mat4 * vec4 */
static inline ov4 mulmv(const float* m, ov4 v) {
  ov4 r;
  r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return r;
}

/* projectVertex — shaderSnippets.wgls.ts:89-93 */
static inline ov3 oproject(const float* m, ov4 p) {
  ov4 c = mulmv(m, p);
  return o3(c.x / c.w, c.y / c.w, c.z / c.w);
}

/* IEEE binary16 <-> binary32, round to nearest even (pack2x16float). */
static inline uint32_t of2h(float f) {
  union { float f; uint32_t u; } v; v.f = f;
  uint32_t x = v.u, sign = (x >> 16) & 0x8000u;
  uint32_t e = (x >> 23) & 0xffu, m = x & 0x7fffffu;
  if (e == 0xffu) return sign | 0x7c00u | (m ? 0x200u : 0u);
  int32_t ee = (int32_t)e - 127 + 15;
  if (ee >= 31) return sign | 0x7c00u;
  if (ee <= 0) {
    if (ee < -10) return sign;
    m |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - ee);
    uint32_t h = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return sign | h;
  }
  uint32_t h = ((uint32_t)ee << 10) | (m >> 13), rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return sign | h;
}
static inline float oh2f(uint32_t h) {
  uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  union { float f; uint32_t u; } v;
  if (e == 0) {
    if (m == 0) { v.u = sign; return v.f; }
    while (!(m & 0x400u)) { m <<= 1; e--; }
    e++; m &= 0x3ffu;
    v.u = sign | ((e + 112u) << 23) | (m << 13);
    return v.f;
  }
  if (e == 31) { v.u = sign | 0x7f800000u | (m << 13); return v.f; }
  v.u = sign | ((e + 112u) << 23) | (m << 13);
  return v.f;
}
static inline uint32_t opack2h(float a, float b) { return of2h(a) | (of2h(b) << 16); }

#endif
