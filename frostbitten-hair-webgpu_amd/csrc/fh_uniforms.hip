// Host-side packing of the uniform blocks: camera and shadow-source
// matrices, lights, material and settings into fh_render_uniforms and
// fh_sim_uniforms, for hosts without the reference's TypeScript packers.
// Reference: src/passes/renderUniformsBuffer.ts:205-288 (update) and :396-414
// (encodeDebugMode), src/passes/simulation/simulationUniformsBuffer.ts:86-152,
// src/passes/shadowMapPass/shared/getMVP_ShadowSourceMatrix.ts:14-79,
// src/utils/index.ts:60-74 (sphericalToCartesian), and the wgpu-matrix 2.9.0
// functions they call (mat4.perspective, ortho, lookAt, inverse; WebGPU clip
// space, z in [0, 1]). Matrices are built in double and stored as f32, as the
// library stores into Float32Array.
#include <cmath>
#include <cstring>

#include "fh_internal.hpp"

namespace {

struct M4 {
  double m[4][4];  // [row][col]
};

M4 identity() {
  M4 r{};
  for (int i = 0; i < 4; i++) r.m[i][i] = 1.0;
  return r;
}

M4 mul(const M4& a, const M4& b) {
  M4 r{};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

void mul_vec(const M4& a, const double* v, double* out) {
  for (int i = 0; i < 4; i++) out[i] = a.m[i][0] * v[0] + a.m[i][1] * v[1] + a.m[i][2] * v[2] + a.m[i][3] * v[3];
}

// Gauss-Jordan elimination with partial pivoting; false when singular.
bool invert(const M4& a, M4& out) {
  double w[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      w[i][j] = a.m[i][j];
      w[i][j + 4] = i == j ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(w[r][c]) > std::fabs(w[p][c])) p = r;
    if (w[p][c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < 8; j++) std::swap(w[p][j], w[c][j]);
    const double inv = 1.0 / w[c][c];
    for (int j = 0; j < 8; j++) w[c][j] *= inv;
    for (int r = 0; r < 4; r++) {
      if (r == c || w[r][c] == 0.0) continue;
      const double f = w[r][c];
      for (int j = 0; j < 8; j++) w[r][j] -= f * w[c][j];
    }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out.m[i][j] = w[i][j + 4];
  return true;
}

void store(const M4& a, float* out16) {  // column-major, as WGSL mat4x4<f32>
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) out16[c * 4 + r] = (float)a.m[r][c];
}

M4 load_colmajor(const double* m16) {
  M4 r;
  for (int c = 0; c < 4; c++)
    for (int rr = 0; rr < 4; rr++) r.m[rr][c] = m16[c * 4 + rr];
  return r;
}

constexpr double kPi = 3.14159265358979323846;
double radians(double deg) { return deg * (kPi / 180.0); }

// mat4.perspective: f = tan(pi / 2 - fovy / 2)
M4 perspective(double fovy, double aspect, double n, double f) {
  M4 r{};
  const double t = std::tan(kPi * 0.5 - 0.5 * fovy);
  const double rangeInv = 1.0 / (n - f);
  r.m[0][0] = t / aspect;
  r.m[1][1] = t;
  r.m[2][2] = f * rangeInv;
  r.m[3][2] = -1.0;
  r.m[2][3] = f * n * rangeInv;
  return r;
}

M4 ortho(double l, double rr, double b, double t, double n, double f) {
  M4 r{};
  r.m[0][0] = 2.0 / (rr - l);
  r.m[1][1] = 2.0 / (t - b);
  r.m[2][2] = 1.0 / (n - f);
  r.m[0][3] = (rr + l) / (l - rr);
  r.m[1][3] = (t + b) / (b - t);
  r.m[2][3] = n / (n - f);
  r.m[3][3] = 1.0;
  return r;
}

void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
void normalize3(double* v) {
  const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= l; v[1] /= l; v[2] /= l;
}

// mat4.lookAt as a view matrix (world -> camera)
M4 look_at(const double* eye, const double* target, const double* up) {
  double z[3] = {eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]};
  normalize3(z);
  double x[3];
  cross(up, z, x);
  normalize3(x);
  double y[3];
  cross(z, x, y);
  normalize3(y);
  M4 r = identity();
  for (int k = 0; k < 3; k++) {
    r.m[0][k] = x[k];
    r.m[1][k] = y[k];
    r.m[2][k] = z[k];
  }
  r.m[0][3] = -(x[0] * eye[0] + x[1] * eye[1] + x[2] * eye[2]);
  r.m[1][3] = -(y[0] * eye[0] + y[1] * eye[1] + y[2] * eye[2]);
  r.m[2][3] = -(z[0] * eye[0] + z[1] * eye[1] + z[2] * eye[2]);
  return r;
}

void spherical_to_cartesian(double phiDeg, double thetaDeg, double* o) {
  const double phi = radians(phiDeg), theta = radians(thetaDeg);
  o[0] = std::cos(phi) * std::sin(theta);
  o[1] = std::cos(theta);
  o[2] = std::sin(phi) * std::sin(theta);
}

uint32_t segments_buffer_entries(uint32_t tile, double avg, uint32_t w, uint32_t h) {
  const double tiles = (double)((w + tile - 1) / tile) * (double)((h + tile - 1) / tile);
  return 1u + (uint32_t)std::ceil(tiles * avg);
}

}  // namespace

extern "C" {

int fh_pack_render_uniforms(const fh_render_params* p, fh_render_uniforms* out) {
  FH_REQUIRE(p && out, "fh_pack_render_uniforms: null argument");
  FH_REQUIRE(p->width > 0 && p->height > 0 && p->tileSize > 0, "fh_pack_render_uniforms: empty viewport or tile");
  FH_REQUIRE(p->meshSphereCount == 0 || p->meshSpheres, "fh_pack_render_uniforms: null mesh spheres");
  fh_render_uniforms u;
  std::memset(&u, 0, sizeof(u));
  const M4 model = load_colmajor(p->modelMatrix);
  const M4 view = look_at(p->cameraEye, p->cameraTarget, p->cameraUp);
  const M4 proj = perspective(radians(p->fovDeg), (double)p->width / (double)p->height, p->near, p->far);
  const M4 vp = mul(proj, view);
  M4 vpInv, projInv;
  FH_REQUIRE(invert(vp, vpInv) && invert(proj, projInv), "fh_pack_render_uniforms: singular camera matrix");
  store(vp, u.vpMatrix);
  store(vpInv, u.vpMatrixInv);
  store(view, u.viewMatrix);
  store(proj, u.projMatrix);
  store(projInv, u.projMatrixInv);
  store(model, u.modelMatrix);
  const M4 mv = mul(view, model);
  store(mv, u.modelViewMat);
  store(mul(mul(proj, view), model), u.mvpMatrix);
  // collision sphere gizmo: translation(center) * scale(0.95 r)
  M4 cs = identity();
  for (int k = 0; k < 3; k++) cs.m[k][3] = p->collisionSphere[k];
  const double rs = p->collisionSphere[3] * 0.95;
  M4 sc = identity();
  sc.m[0][0] = sc.m[1][1] = sc.m[2][2] = rs;
  store(mul(cs, sc), u.collisionSphereModelMatrix);
  u.viewport[0] = (float)p->width;
  u.viewport[1] = (float)p->height;
  for (int k = 0; k < 3; k++) u.cameraPosition[k] = (float)p->cameraEye[k];
  for (int k = 0; k < 3; k++) u.colorMgmt[k] = (float)p->colors[k];
  for (int k = 0; k < 4; k++) u.lightAmbient[k] = (float)p->lightAmbient[k];
  fh_light* lights[3] = {&u.light0, &u.light1, &u.light2};
  for (int i = 0; i < 3; i++) {
    double lp[3];
    spherical_to_cartesian(p->lights[i].posPhi, p->lights[i].posTheta, lp);
    for (int k = 0; k < 3; k++) lights[i]->position[k] = (float)(lp[k] * 2.0);
    for (int k = 0; k < 3; k++) lights[i]->colorAndEnergy[k] = (float)p->lights[i].color[k];
    lights[i]->colorAndEnergy[3] = (float)p->lights[i].energy;
  }
  // shadow source (getMVP_ShadowSourceMatrix.ts:14-79): orthographic box
  // around the hair's and meshes' bounding spheres, 5% margin
  double spos[3];
  spherical_to_cartesian(p->shadowPosPhi, p->shadowPosTheta, spos);
  for (int k = 0; k < 3; k++) spos[k] *= p->shadowDistance;
  double tgt4[4] = {p->shadowTarget[0], p->shadowTarget[1], p->shadowTarget[2], 1.0}, tgt[4];
  mul_vec(model, tgt4, tgt);
  const double up[3] = {0.0, 1.0, 0.0};
  const M4 sview = look_at(spos, tgt, up);
  const M4 smv = mul(sview, model);
  double l = 9999999.0, r = -9999999.0, b = 9999999.0, t = -9999999.0;
  for (uint32_t i = 0; i <= p->meshSphereCount; i++) {
    const double* s = i == 0 ? p->hairSphere : p->meshSpheres + 4 * (i - 1);
    const double c4[4] = {s[0], s[1], s[2], 1.0};
    double q[4];
    mul_vec(smv, c4, q);
    l = std::fmin(l, q[0] - s[3]);
    r = std::fmax(r, q[0] + s[3]);
    b = std::fmin(b, q[1] - s[3]);
    t = std::fmax(t, q[1] + s[3]);
  }
  const M4 sproj = ortho(l * 1.05, r * 1.05, b * 1.05, t * 1.05, 0.1, 20.0);
  store(smv, u.shadows.sourceModelViewMat);
  store(sproj, u.shadows.sourceProjMatrix);
  store(mul(mul(sproj, sview), model), u.shadows.sourceMVP_Matrix);
  for (int k = 0; k < 3; k++) u.shadows.sourcePosition[k] = (float)spos[k];
  u.shadows.sourcePosition[3] = (float)p->fiberRadius;
  u.shadows.usePCSS = p->shadowUsePCSS ? 1u : 0u;
  u.shadows.PCF_Radius = p->shadowPCFRadius;
  u.shadows.bias = (float)p->shadowBias;
  u.shadows.strength = (float)p->shadowStrength;
  u.ao.radius = (float)p->aoRadius;
  u.ao.directionOffset = (float)p->aoDirectionOffset;
  u.ao.falloffStart2 = (float)p->aoFalloffStart2;
  u.ao.falloffEnd2 = (float)p->aoFalloffEnd2;
  u.ao.numDirections = p->aoNumDirections;
  u.ao.numSteps = p->aoNumSteps;
  u.ao.strength = (float)p->aoStrength;
  fh_hair_material& m = u.hairMaterial;
  for (int k = 0; k < 3; k++) {
    m.color0[k] = (float)p->hairColor0[k];
    m.color1[k] = (float)p->hairColor1[k];
  }
  m.specular = (float)p->specular;
  m.shadows = (float)p->shadows;
  m.weightTT = (float)p->weightTT;
  m.weightTRT = (float)p->weightTRT;
  m.shift = (float)p->shift;
  m.roughness = (float)p->roughness;
  m.colorRng = (float)p->colorRng;
  m.lumaRng = (float)p->lumaRng;
  m.attenuation = (float)p->attenuation;
  for (int k = 0; k < 3; k++) {
    u.sdfBoundsMin[k] = (float)p->sdfBoundsMin[k];
    u.sdfBoundsMax[k] = (float)p->sdfBoundsMax[k];
    u.gridBoundsMin[k] = (float)p->gridBoundsMin[k];
    u.gridBoundsMax[k] = (float)p->gridBoundsMax[k];
  }
  u.sdfBoundsMin[3] = 2.5f;  // debug slice (0.5 + 2.0)
  u.sdfBoundsMax[3] = (float)p->sdfDistanceOffset;
  u.gridBoundsMin[3] = 0.5f;
  u.gridDebugDisplayValue = 17;
  u.fiberRadius = (float)p->fiberRadius;
  u.dbgShadowMapPreviewSize = (float)std::floor(std::fmin(500.0, std::fmin(p->width / 3.0, p->height / 3.0)));
  u.maxDrawnHairSegments = segments_buffer_entries(p->tileSize, p->avgSegmentsPerTile, p->width, p->height);
  // encodeDebugMode: the mode in the low 8 bits, its extra value above them
  uint32_t extra = 0;
  if (p->displayMode == 1u || p->displayMode == 2u) extra = p->dbgTileModeMaxSegments;
  else if (p->displayMode == 4u) extra = p->dbgSlicesModeMaxSlices;
  else if (p->displayMode == 0u) extra = p->dbgShowTiles ? 1u : 0u;
  u.displayMode = p->displayMode | (extra << 8);
  // CONFIG.background (constants.ts:104-109)
  const double c0[3] = {22.0, 162.0, 188.0}, c1[3] = {14.0, 103.0, 120.0};
  for (int k = 0; k < 3; k++) {
    u.backgroundColor0[k] = (float)(c0[k] / 255.0);
    u.backgroundColor1[k] = (float)(c1[k] / 255.0);
    u.collisionSpherePosition[k] = (float)p->collisionSphere[k];
  }
  u.backgroundNoiseScale = 5.0f;
  u.backgroundGradientStrength = 0.5f;
  *out = u;
  return FH_OK;
}

int fh_pack_sim_uniforms(const fh_sim_params* p, fh_sim_uniforms* out) {
  FH_REQUIRE(p && out, "fh_pack_sim_uniforms: null argument");
  fh_sim_uniforms u;
  std::memset(&u, 0, sizeof(u));
  const M4 model = load_colmajor(p->modelMatrix);
  M4 inv;
  FH_REQUIRE(invert(model, inv), "fh_pack_sim_uniforms: singular model matrix");
  store(model, u.modelMatrix);
  store(inv, u.modelMatrixInv);
  const double cs4[4] = {p->collisionSphere[0], p->collisionSphere[1], p->collisionSphere[2], 1.0};
  double cs[4];
  mul_vec(inv, cs4, cs);
  for (int k = 0; k < 3; k++) u.collisionSphere[k] = (float)cs[k];
  u.collisionSphere[3] = (float)p->collisionSphere[3];
  double wd[3];
  spherical_to_cartesian(p->windDirPhi, p->windDirTheta, wd);
  const double wd4[4] = {wd[0], wd[1], wd[2], 1.0};  // w = 1 (simulationUniformsBuffer.ts:113)
  double w[4];
  mul_vec(inv, wd4, w);
  for (int k = 0; k < 3; k++) u.wind[k] = (float)w[k];
  u.wind[3] = (float)p->windStrength;
  for (int k = 0; k < 3; k++) {
    u.sdfBoundsMin[k] = (float)p->sdfBoundsMin[k];
    u.sdfBoundsMax[k] = (float)p->sdfBoundsMax[k];
    u.gridBoundsMin[k] = (float)p->gridBoundsMin[k];
    u.gridBoundsMax[k] = (float)p->gridBoundsMax[k];
  }
  u.sdfBoundsMin[3] = 2.5f;
  u.sdfBoundsMax[3] = (float)p->sdfDistanceOffset;
  u.gridBoundsMin[3] = 0.5f;
  u.gridDebugDisplayValue = 1 + 16;
  u.deltaTime = (float)p->deltaTime;
  u.gravity = (float)p->gravity;
  u.constraintIterations = p->constraintIterations;
  u.frameIdx = p->frameIdx;
  u.windStrengthLull = (float)p->windStrengthLull;
  u.windColisionTraceOffset = (float)p->windColisionTraceOffset;
  u.stiffnessLengthConstr = (float)p->stiffnessLengthConstr;
  u.stiffnessCollisions = (float)p->stiffnessCollisions;
  u.stiffnessSDF = (float)p->stiffnessSDF;
  u.volumePreservation = (float)p->volumePreservation;
  u.friction = (float)p->friction;
  u.stiffnessGlobalConstr = (float)p->stiffnessGlobalConstr;
  u.globalConstrExtent = (float)p->globalConstrExtent;
  u.globalConstrFade = (float)p->globalConstrFade;
  u.stiffnessLocalConstr = (float)p->stiffnessLocalConstr;
  u.windPhaseOffset = (float)p->windPhaseOffset;
  u.windStrengthFrequency = (float)p->windStrengthFrequency;
  u.windStrengthJitter = (float)p->windStrengthJitter;
  *out = u;
  return FH_OK;
}

}  // extern "C"
