// Strand-space shading (K8) on gfx950: Kajiya-Kay diffuse with fake multiple
// scattering, Marschner R/TT/TRT (Karis16), PCF/PCSS shadow from the shadow
// map, AO and the fake Beer-Lambert attenuation, per (strand, shading point).
// Reference: src/passes/hairShadingPass/hairShadingPass.wgsl.ts:61-282,
// marschner.wgsl.ts:19-170, shadowMapPass/shared/sampleShadows.wgsl.ts:21-150,
// aoPass/shared/textureAo.wgsl.ts:7-12, _shaderSnippets/*.
#include "fh_device.hpp"
#include "fh_launch.hpp"

namespace {

using namespace fh;

constexpr float kPi = 3.14159265358979323846f;
constexpr float kSqrtTwoPi = 2.5066282746310002f;
// HAIR_F0 = (1 - eta)^2 / (1 + eta)^2 with eta = 1.55 (marschner.wgsl.ts:89-91)
constexpr float kHairF0 = (float)((1.0 - 1.55) * (1.0 - 1.55) / ((1.0 + 1.55) * (1.0 + 1.55)));
// linearizeDepth constants: CONFIG.camera.projection near / far (constants.ts:131-135)
constexpr float kNear = 0.01f, kFar = 100.0f;

// The lobes' powers and exponentials as WGSL defines pow (exp2(y * log2(x)))
// on the hardware's base-2 logarithm and exponential (about 1 ulp each; the
// relative error of the result grows with |y| by about |y| * 2^-24). The
// shading points are checked against the oracle's libm powf / expf within a
// relative 2e-4.
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
// x >= 0 (NaN passes through); x^0 = 1^e = 1 as powf has them
__device__ __forceinline__ float pow_hw(float x, float e) {
  if (e == 0.0f || x == 1.0f) return 1.0f;
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float fresnel1(float cosT, float f0) {
  const float x = 1.0f - cosT, x2 = x * x;  // pow(x, 5)
  return f0 + (1.0f - f0) * (x2 * x2 * x);
}
__device__ __forceinline__ float cos_half(float c) { return sqrtf(satf(0.5f + 0.5f * c)); }
__device__ __forceinline__ float long_gauss(float si, float sr, float alpha, float beta) {
  const float a = si + sr - alpha;
  const float b = -(a * a) / (2.0f * beta * beta);
  return exp_hw(b) / (beta * kSqrtTwoPi);
}
__device__ __forceinline__ f3 powv(f3 b, float e) { return mk3(pow_hw(b.x, e), pow_hw(b.y, e), pow_hw(b.z, e)); }

struct Marschner {
  f3 base;
  float specular, weightTT, weightTRT, shift, roughness;
};

// hairSpecularMarschner — marschner.wgsl.ts:34-97
__device__ f3 marschner(const Marschner& p, f3 L, f3 V, f3 T) {
  const float si = dot(T, L), sr = dot(T, V);
  const float ci = sqrtf(fmx(0.0f, 1.0f - si * si));
  const float cr = sqrtf(fmx(0.0f, 1.0f - sr * sr));
  const float cd = cos_half(ci * cr + si * sr);
  const f3 lp = L - T * si, cp = V - T * sr;
  const float cphi = dot(lp, cp) * (1.0f / sqrtf(length(lp) * length(cp) + 1e-4f));
  const float chp = cos_half(cphi);
  // R
  const float beta2 = p.roughness * p.roughness;
  const float Mr = long_gauss(si, sr, -2.0f * p.shift, beta2);
  const float Nr = 0.25f * chp * fresnel1(cd, kHairF0);
  // TT — azimuthalScattering_TT :412-434
  const float Mtt = long_gauss(si, sr, p.shift, 0.5f * beta2);
  const float eta = (1.19f / cd) + (0.36f * cd);
  const float a = 1.0f / eta;
  const float hTT = (1.0f + a * (0.6f - 0.8f * cphi)) * chp;
  const f3 Ttt = powv(p.base, sqrtf(1.0f - hTT * hTT * a * a) / (2.0f * cd));
  const float Dtt = exp_hw(-3.65f * cphi - 3.98f);
  const float f = fresnel1(cd * sqrtf(satf(1.0f - hTT * hTT)), kHairF0);
  const f3 Ntt = (Ttt * ((1.0f - f) * (1.0f - f))) * Dtt;
  // TRT — azimuthalScattering_TRT :437-454
  const float Mtrt = long_gauss(si, sr, 4.0f * p.shift, 2.0f * beta2);
  const f3 Ttrt = powv(p.base, 0.8f / cd);
  const float Dtrt = exp_hw(17.0f * cphi - 16.78f);
  const float f2v = fresnel1(cd * 0.5f, kHairF0);
  const f3 Ntrt = (Ttrt * ((1.0f - f2v) * (1.0f - f2v) * f2v)) * Dtrt;
  const float r = p.specular * Mr * Nr;
  return (mk3(r, r, r) + Ntt * (p.weightTT * Mtt)) + Ntrt * (p.weightTRT * Mtrt);
}

// hairShading — hairShadingPass.wgsl.ts:166-205
__device__ f3 hair_light(const Marschner& p, const fh_light& light, f3 toCam, f3 T, f3 posWS,
                         float shadow, float ao) {
  const f3 L = normalize(mk3(light.position[0], light.position[1], light.position[2]) - posWS);
  const f3 diffuse = p.base * mixf(0.25f, 1.0f, dot(T, L));  // KajiyaKayDiffuse :209-213
  // fakeMultipleScattering :216-234
  const f3 fn = normalize(toCam - T * dot(toCam, T));
  const float ds = (satf(dot(fn, L)) + 1.0f) / (4.0f * kPi);
  const float luma = 0.2126729f * p.base.x + 0.7151522f * p.base.y + 0.0721750f * p.base.z;
  const f3 tint = powv(p.base / luma, 1.0f - shadow);
  const f3 ms = (mk3(sqrtf(p.base.x), sqrtf(p.base.y), sqrtf(p.base.z)) * ds) * tint;
  const f3 dt = (diffuse * ms) * satf(dot(T, L));
  const f3 brdf = dt + marschner(p, L, toCam, T);
  const f3 radiance = mk3(light.colorAndEnergy[0], light.colorAndEnergy[1], light.colorAndEnergy[2]) *
                      light.colorAndEnergy[3];
  return (brdf * radiance) * (1.0f - ao);
}

__device__ __forceinline__ float linearize_depth(float d) { return kNear * kFar / (kFar + (kNear - kFar) * d); }

struct ShadeArgs {
  const fh_render_uniforms* u;
  const float4* pos;
  const float4* tan;
  float4* out;
  const float* shadow;
  uint32_t shadowW, shadowH;
  const float* ao;
  uint32_t aoW, aoH;
  const float* depth;
  uint32_t depthW, depthH;
  uint32_t S, P, PS;
};

// kShadeLanes lanes per shading point split the PCF taps (a count of 0/1
// comparisons, exact in any order). The 32 points of a workgroup then hand
// their three lights to 96 threads, one (point, light) each, and the lights
// are added back in the reference's order.
constexpr uint32_t kShadeLanes = 8;
constexpr uint32_t kShadePoints = 256 / kShadeLanes;

struct LightIn {
  float base[3], toCam[3], T[3], pos[3], shadow, ao;
};

__device__ __forceinline__ void shading_body(ShadeArgs a, uint32_t block, uint32_t blocks) {
  __shared__ LightIn lin[kShadePoints];
  __shared__ float lout[kShadePoints * 3][3];
  const uint32_t pi = threadIdx.x / kShadeLanes;
  const uint32_t gid = block * kShadePoints + pi;
  const uint32_t sub = threadIdx.x % kShadeLanes;
  const bool valid = gid < a.S * a.PS;
  const uint32_t s = valid ? gid / a.PS : 0u, id = valid ? gid - s * a.PS : 0u;
  const fh_render_uniforms* __restrict__ u = a.u;
  const fh_hair_material& m = u->hairMaterial;
  const f2 vp = mk2(u->viewport[0], u->viewport[1]);
  const float t = (float)id / (float)a.PS;
  // remapToIndices — shaderSnippets.wgls.ts:120-126
  const float ra = t * ((float)a.P - 1.0f);
  const uint32_t au = f2u(ra);
  const uint32_t ix = clampu(au, 0u, a.P - 1u), iy = clampu(au + 1u, 0u, a.P - 1u);
  const float tIn = fractf(ra);
  const float4 q0 = a.pos[(size_t)s * a.P + ix], q1 = a.pos[(size_t)s * a.P + iy];
  const float4 tq = a.tan[(size_t)s * a.P + iy];
  const f3 s0 = mk3(q0.x, q0.y, q0.z), s1 = mk3(q1.x, q1.y, q1.z);
  const f3 pObj = mix(s0, s1, tIn);
  const f3 tObj = mix(normalize(s1 - s0), mk3(tq.x, tq.y, tq.z), tIn);
  const f4 posWS = mul(u->modelMatrix, mk4(pObj.x, pObj.y, pObj.z, 1.0f));
  const f4 tanWS = mul(u->modelMatrix, mk4(tObj.x, tObj.y, tObj.z, 1.0f));  // w = 1 as in :95
  const f3 pproj = project(u->vpMatrix, posWS);
  const f2 p01 = mk2(pproj.x * 0.5f + 0.5f, pproj.y * 0.5f + 0.5f);
  const f2 texPx = mk2(vp.x * p01.x, vp.y * (1.0f - p01.y));
  // base colour :105-114
  f3 base = mix(mk3(m.color0[0], m.color0[1], m.color0[2]), mk3(m.color1[0], m.color1[1], m.color1[2]), t);
  const float fs = (float)s;
  const f3 rnd = sat(mk3(fractf(fs * 1.73f) * 1.0f, fractf(fs * 1.17f) * 1.0f, fractf(fs * 1.31f) * 1.0f));
  base = mix(base, rnd, m.colorRng);
  base = mix(base, base * fractf(fs / 255.0f), m.lumaRng);
  const f3 toCam = normalize(mk3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]) - xyz(posWS));
  // shadow — getShadow :236-254 -> calculateDirectionalShadow (sampleShadows.wgsl.ts:47-84)
  const f4 lpos = mul(u->shadows.sourceMVP_Matrix, posWS);
  const f3 pss = mk3(lpos.x * 0.5f + 0.5f, 1.0f - (lpos.y * 0.5f + 0.5f), lpos.z);
  const f3 ssp = mk3(u->shadows.sourcePosition[0], u->shadows.sourcePosition[1], u->shadows.sourcePosition[2]);
  const f3 toSrc = normalize(ssp - xyz(posWS));
  const float bias = fmx(fabsf(u->shadows.bias) * (1.0f - dot(xyz(tanWS), toSrc)), 0.0005f);
  int32_t radius = (int32_t)u->shadows.PCF_Radius;
  if (u->shadows.usePCSS > 0u) {
    const float sm = gather_avg(a.shadow, a.shadowW, a.shadowH, mk2(pss.x, pss.y));
    const float dd = fmx(pss.z - sm, 0.0f);
    radius = 1 + f2i(dd / sm * 10.0f);  // PCSS_PENUMBRA_BASE + depthDiff/depth * WIDTH
  }
  const float texel = 1.0f / (float)a.shadowW;
  // taps x-major as in the reference loops, dealt round-robin to the lanes
  const int32_t side = radius * 2 + 1;
  uint32_t hits = 0;
  for (int32_t k = (int32_t)sub; k < side * side; k += (int32_t)kShadeLanes) {
    const int32_t x = -radius + k / side, y = -radius + k % side;
    const f2 uv = mk2(pss.x, pss.y) + mk2((float)x * texel, (float)y * texel);
    hits += (pss.z - bias > gather_avg(a.shadow, a.shadowW, a.shadowH, uv)) ? 1u : 0u;
  }
#pragma unroll
  for (uint32_t o = 1; o < kShadeLanes; o <<= 1) hits += (uint32_t)__shfl_xor((int)hits, (int)o);
  const float sh = (float)hits;
  const float pcf = (float)(radius * 2 + 1);
  float shadowRes = sh / (pcf * pcf);
  const bool validSample = !(pss.z > 1.0f) && !(pss.x < 0.0f || pss.x > 1.0f || pss.y < 0.0f || pss.y > 1.0f);
  if (!validSample) shadowRes = 0.0f;
  const float shadow = clampf(1.0f - shadowRes, 1.0f - m.shadows, 1.0f);
  // AO — textureAo.wgsl.ts:7-12
  const f2 at = mk2(texPx.x / vp.x, texPx.y / vp.y);
  const float aov = tex_load(a.ao, a.aoW, a.aoH, f2i((float)a.aoW * at.x), f2i((float)a.aoH * at.y));
  float ao = 1.0f - mixf(1.0f, aov, u->ao.strength);
  // getAttenuation :259-282
  const float zb = tex_load(a.depth, a.depthW, a.depthH, (int32_t)f2u(texPx.x), (int32_t)f2u(texPx.y));
  const float att = fabsf(linearize_depth(pproj.z) - linearize_depth(zb)) * m.attenuation;
  ao = satf(ao + att);
  const f3 T = xyz(tanWS);
  if (sub == 0u) {
    LightIn& li = lin[pi];
    li.base[0] = base.x; li.base[1] = base.y; li.base[2] = base.z;
    li.toCam[0] = toCam.x; li.toCam[1] = toCam.y; li.toCam[2] = toCam.z;
    li.T[0] = T.x; li.T[1] = T.y; li.T[2] = T.z;
    li.pos[0] = posWS.x; li.pos[1] = posWS.y; li.pos[2] = posWS.z;
    li.shadow = shadow; li.ao = ao;
  }
  __syncthreads();
  if (threadIdx.x < kShadePoints * 3u) {
    const uint32_t q = threadIdx.x / 3u, l = threadIdx.x - q * 3u;
    const LightIn& li = lin[q];
    const Marschner lp{mk3(li.base[0], li.base[1], li.base[2]), m.specular, m.weightTT, m.weightTRT, m.shift,
                       m.roughness};
    const fh_light& light = l == 0u ? u->light0 : (l == 1u ? u->light1 : u->light2);
    const f3 r = hair_light(lp, light, mk3(li.toCam[0], li.toCam[1], li.toCam[2]), mk3(li.T[0], li.T[1], li.T[2]),
                            mk3(li.pos[0], li.pos[1], li.pos[2]), li.shadow, li.ao);
    lout[threadIdx.x][0] = r.x; lout[threadIdx.x][1] = r.y; lout[threadIdx.x][2] = r.z;
  }
  __syncthreads();
  if (valid && sub == 0u) {
    // (0 + L0) + L1 + L2, as the reference accumulates
    f3 rs = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t l = 0; l < 3u; l++)
      rs = rs + mk3(lout[pi * 3u + l][0], lout[pi * 3u + l][1], lout[pi * 3u + l][2]);
    const float alpha = id >= a.PS - 1u ? 0.0f : 1.0f;  // thin-tip alpha :160
    a.out[(size_t)s * a.PS + id] = make_float4(rs.x, rs.y, rs.z, alpha);
  }
}

struct Shading {
  using Args = ShadeArgs;
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "shading";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { shading_body(a, xcd_order(block, blocks), blocks); }
};


}  // namespace

extern "C" int fh_shading(fh_ctx* ctx, const fh_render_cfg* cfg, const fh_render_uniforms* uniforms,
                          const fh_hair_data* hair_data, const fh_vec4* positions,
                          const fh_vec4* tangents, fh_vec4* shading, fh_tex2d shadow_map, fh_tex2d ao,
                          fh_tex2d depth, uint32_t strands_count, uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  (void)hair_data;
  FH_REQUIRE(cfg && uniforms && positions && tangents && shading && shadow_map.data && ao.data && depth.data,
             "fh_shading: null buffer");
  FH_REQUIRE(cfg->shadingPoints >= 2 && cfg->shadingPoints <= 1024, "fh_shading: shadingPoints in [2, 1024]");
  FH_REQUIRE(points_per_strand >= 2, "fh_shading: points_per_strand >= 2");
  FH_REQUIRE(shadow_map.width > 0 && shadow_map.height > 0 && ao.width > 0 && ao.height > 0,
             "fh_shading: empty texture");
  const uint64_t n = (uint64_t)strands_count * cfg->shadingPoints;
  if (n == 0) return FH_OK;
  FH_REQUIRE(n < (1ull << 31), "fh_shading: too many shading points");
  ShadeArgs a;
  a.u = uniforms;
  a.pos = reinterpret_cast<const float4*>(positions);
  a.tan = reinterpret_cast<const float4*>(tangents);
  a.out = reinterpret_cast<float4*>(shading);
  a.shadow = shadow_map.data; a.shadowW = shadow_map.width; a.shadowH = shadow_map.height;
  a.ao = ao.data; a.aoW = ao.width; a.aoH = ao.height;
  a.depth = depth.data; a.depthW = depth.width; a.depthH = depth.height;
  a.S = strands_count; a.P = points_per_strand; a.PS = cfg->shadingPoints;
  return fh::launch<Shading>(ctx, fh::div_ceil((uint32_t)n, kShadePoints), 0, a);
}
