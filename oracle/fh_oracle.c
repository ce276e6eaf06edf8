/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A sequential C restatement of the reference's per-frame hair compute chain
 * (the WGSL compute kernels of Scthe/frostbitten-hair-webgpu). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker. Each function cites the reference file:line it follows.
 *
 * Serialisation: where the reference runs threads concurrently and lets atomic
 * order decide a result, this oracle runs them in ascending index order
 * (strand-major, then segment; tiles in tile-list order). That is one legal
 * schedule of the reference. In the fine pass it fixes the order in which
 * fragments of one slice are blended: ascending (strand, segment).
 *
 * Policies the reference leaves to the WebGPU implementation (robust access):
 * out-of-bounds texture loads return 0, out-of-bounds storage reads return 0
 * and out-of-bounds storage writes are dropped.
 *
 * Parity status per function is listed in the header comment of each block:
 * "pinned" names the check in tests/ that would fail on a plausible mistake.
 */
#include "fh_oracle.h"

#include <stdlib.h>
#include <string.h>

#include "fh_oracle_math.h"

#define FLOAT_EPSILON 1e-7f

/* Host threads for the loops whose iterations are independent (strands,
 * grid nodes, tiles). Results do not depend on it: every order-dependent
 * step (list appends, i32 grid sums) is done in the sequential order or
 * with associative integer additions. */
static int g_threads = 1;

void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }
#define MAX_U32 0xffffffffu
#define INVALID_PTR 0xffffffffu

/* ======================================================================== */
/* Shared helpers                                                            */
/* ======================================================================== */

/* sampleSDFCollider — sdfCollider.ts:36-44 with the linear clamp-to-edge
 * sampler of sdfUtils.ts:33-42. Pinned: test_oracle_golden.py
 * test_sdf_roundtrip_and_texel_centres (texel centres, a midpoint, clamp to
 * edge) and test_oracle_sim.py test_sdf_pushes_points_along_the_plane_normal
 * (a linear field reproduced through forward differences). */
static float sample_sdf(const float* sdf, uint32_t d, ov3 bmin, ov3 bmax, ov3 p) {
  ov3 t = sat3(div3(sub3(p, bmin), sub3(bmax, bmin)));
  float c[3] = {t.x * (float)d - 0.5f, t.y * (float)d - 0.5f, t.z * (float)d - 0.5f};
  int i0[3], i1[3];
  float f[3];
  for (int k = 0; k < 3; k++) {
    float fl = floorf(c[k]);
    f[k] = c[k] - fl;
    int a = (int)fl;
    int b = a + 1;
    i0[k] = a < 0 ? 0 : (a > (int)d - 1 ? (int)d - 1 : a);
    i1[k] = b < 0 ? 0 : (b > (int)d - 1 ? (int)d - 1 : b);
  }
#define S(X, Y, Z) sdf[((size_t)(Z) * d + (size_t)(Y)) * d + (size_t)(X)]
  float c00 = S(i0[0], i0[1], i0[2]), c10 = S(i1[0], i0[1], i0[2]);
  float c01 = S(i0[0], i1[1], i0[2]), c11 = S(i1[0], i1[1], i0[2]);
  float c02 = S(i0[0], i0[1], i1[2]), c12 = S(i1[0], i0[1], i1[2]);
  float c03 = S(i0[0], i1[1], i1[2]), c13 = S(i1[0], i1[1], i1[2]);
#undef S
  float x0 = c00 + f[0] * (c10 - c00);
  float x1 = c01 + f[0] * (c11 - c01);
  float x2 = c02 + f[0] * (c12 - c02);
  float x3 = c03 + f[0] * (c13 - c03);
  float y0 = x0 + f[1] * (x1 - x0);
  float y1 = x2 + f[1] * (x3 - x2);
  return y0 + f[2] * (y1 - y0);
}

float or_sample_sdf(const float* sdf, uint32_t d, const float* bmin,
                    const float* bmax, const float* p) {
  return sample_sdf(sdf, d, o3(bmin[0], bmin[1], bmin[2]),
                    o3(bmax[0], bmax[1], bmax[2]), o3(p[0], p[1], p[2]));
}

/* _getGridIdx — grids.wgsl.ts:15-21 */
static uint32_t grid_idx(uint32_t d, uint32_t x, uint32_t y, uint32_t z) {
  return oclampu(z, 0, d - 1) * d * d + oclampu(y, 0, d - 1) * d + oclampu(x, 0, d - 1);
}

/* _getGridCell — grids.wgsl.ts:61-74 */
typedef struct { uint32_t mn[3], mx[3]; ov3 pin; } grid_cell;
static grid_cell get_grid_cell(uint32_t d, ov3 bmin, ov3 bmax, ov3 p) {
  grid_cell r;
  ov3 t = sat3(div3(sub3(p, bmin), sub3(bmax, bmin)));
  float dm1 = (float)(d - 1u);
  r.pin = muls3(t, dm1);
  r.mn[0] = ou32(floorf(t.x * dm1)); r.mn[1] = ou32(floorf(t.y * dm1)); r.mn[2] = ou32(floorf(t.z * dm1));
  r.mx[0] = ou32(ceilf(t.x * dm1)); r.mx[1] = ou32(ceilf(t.y * dm1)); r.mx[2] = ou32(ceilf(t.z * dm1));
  return r;
}

/* _getGridCellWeights — grids.wgsl.ts:76-84 */
static ov3 cell_weights(uint32_t x, uint32_t y, uint32_t z, ov3 pin) {
  return o3(oclamp(1.0f - fabsf(pin.x - (float)x), 0.0f, 1.0f),
            oclamp(1.0f - fabsf(pin.y - (float)y), 0.0f, 1.0f),
            oclamp(1.0f - fabsf(pin.z - (float)z), 0.0f, 1.0f));
}

static float grid_decode(int32_t v) { return (float)v / 1000000.0f; }

/* ======================================================================== */
/* K1 — gridPreSimPass.wgsl.ts:38-159                                        */
/* Pinned: test_oracle_golden.py test_grid_pre_sim_known_answers (uniform   */
/* density -> zero gradient, symmetric spike -> zero, ramp -> -step, lone    */
/* corner -> (1,1,1), wind 0 / 0.5 / 1 regions around a sphere SDF).        */
/* ======================================================================== */

void or_grid_pre_sim(uint32_t d, const fh_sim_uniforms* u,
                     const fh_density_velocity* dv, fh_density_grad_wind* out,
                     const float* sdf, uint32_t sdf_dims) {
  ov3 gmin = o3(u->gridBoundsMin[0], u->gridBoundsMin[1], u->gridBoundsMin[2]);
  ov3 gmax = o3(u->gridBoundsMax[0], u->gridBoundsMax[1], u->gridBoundsMax[2]);
  ov3 smin = o3(u->sdfBoundsMin[0], u->sdfBoundsMin[1], u->sdfBoundsMin[2]);
  ov3 smax = o3(u->sdfBoundsMax[0], u->sdfBoundsMax[1], u->sdfBoundsMax[2]);
  ov3 windDir = o3(u->wind[0], u->wind[1], u->wind[2]);
  uint32_t total = d * d * d;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (uint32_t id = 0; id < total; id++) {
    /* deconstructId :77-83 */
    uint32_t z = id / (d * d);
    uint32_t idXY = id - z * d * d;
    uint32_t y = idXY / d, x = idXY % d;
    /* getGridPointPositionWS — grids.wgsl.ts:23-36 */
    ov3 cell = divs3(sub3(gmax, gmin), (float)(d - 1u));
    ov3 pos = add3(gmin, mul3(cell, o3((float)x, (float)y, (float)z)));
    /* getWindStrength :89-108 */
    float wind;
    float sd = sample_sdf(sdf, sdf_dims, smin, smax, pos);
    if (sd < 0.0f) {
      wind = 0.0f;
    } else {
      ov3 toward = muls3(norm3(windDir), -1.0f);
      ov3 rp = add3(pos, muls3(muls3(toward, sd), u->windColisionTraceOffset));
      float sd2 = sample_sdf(sdf, sdf_dims, smin, smax, rp);
      wind = sd2 < 0.0f ? 0.5f : 1.0f;
    }
    /* getDensityGradient :111-159 */
    uint32_t idx0 = grid_idx(d, x, y, z);
    float dens0 = grid_decode(dv[idx0].density);
    ov3 res = o3(0.0f, 0.0f, 0.0f);
    static const int offs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int k = 0; k < 6; k++) {
      int ox = (int)x + offs[k][0], oy = (int)y + offs[k][1], oz = (int)z + offs[k][2];
      ov3 g = o3(0.0f, 0.0f, 0.0f);
      if (!(ox < 0 || oy < 0 || oz < 0)) {
        uint32_t idxO = grid_idx(d, (uint32_t)ox, (uint32_t)oy, (uint32_t)oz);
        if (idxO != idx0) {
          float densO = grid_decode(dv[idxO].density);
          float diff = dens0 - densO;
          float m = omax(diff, 0.0f);
          g = o3(m * (float)offs[k][0], m * (float)offs[k][1], m * (float)offs[k][2]);
        }
      }
      res = add3(res, g);
    }
    out[idx0].densityGrad[0] = res.x;
    out[idx0].densityGrad[1] = res.y;
    out[idx0].densityGrad[2] = res.z;
    out[idx0].windStrength = wind;
  }
}

/* ======================================================================== */
/* K2 — hairSimIntegrationPass.wgsl.ts:79-273 with shaderImpl/               */
/* integration.wgsl.ts:8-28, constraints.wgsl.ts:5-120,                      */
/* collisions.wgsl.ts:4-61.                                                  */
/* Pinned: test_oracle_golden.py (test_integration_free_flight,            */
/* test_length_constraint_restores_rest_length,                              */
/* test_global_shape_stiffness_one_returns_initial,                          */
/* test_sphere_collision_pushes_to_shell, and                                */
/* test_integration_matches_double_precision_restatement: a float64 step of  */
/* a 6-point strand) and test_oracle_sim.py (grid velocity / friction,       */
/* density gradient, wind with strand jitter, forces together, SDF push-out  */
/* along an analytic plane's normal, j > 2 only).                            */
/* ======================================================================== */

static ov4 ld4(const fh_vec4* a, size_t i) { return o4(a[i].x, a[i].y, a[i].z, a[i].w); }

void or_sim_integrate(uint32_t d, const fh_sim_uniforms* u,
                      fh_vec4* prev, const fh_vec4* now, const float* seglen,
                      const float* sdf, uint32_t sdf_dims,
                      const fh_density_velocity* dv,
                      const fh_density_grad_wind* dgw, const fh_vec4* initial,
                      uint32_t S, uint32_t P) {
  ov3 gmin = o3(u->gridBoundsMin[0], u->gridBoundsMin[1], u->gridBoundsMin[2]);
  ov3 gmax = o3(u->gridBoundsMax[0], u->gridBoundsMax[1], u->gridBoundsMax[2]);
  ov3 smin = o3(u->sdfBoundsMin[0], u->sdfBoundsMin[1], u->sdfBoundsMin[2]);
  ov3 smax = o3(u->sdfBoundsMax[0], u->sdfBoundsMax[1], u->sdfBoundsMax[2]);
  float sdfOffset = u->sdfBoundsMax[3];
  ov3 gravityForce = o3(0.0f, -u->gravity, 0.0f);
  ov4 sphere = o4(u->collisionSphere[0], u->collisionSphere[1], u->collisionSphere[2], u->collisionSphere[3]);
  float iters = (float)u->constraintIterations;
  const float SDF_GRAD_OFFSET = (float)(0.0033918858971446753 * 0.1);

#pragma omp parallel num_threads(g_threads)
  {
  ov4* pos = (ov4*)malloc(sizeof(ov4) * P);
#pragma omp for schedule(dynamic, 64)
  for (uint32_t s = 0; s < S; s++) {
    size_t base = (size_t)s * P;
    pos[0] = ld4(now, base);
    /* verlet + forces :125-162 */
    for (uint32_t i = 1; i < P; i++) {
      ov4 pp = ld4(prev, base + i), pn = ld4(now, base + i);
      ov3 p = xyz(pn);
      ov3 force = gravityForce;
      grid_cell co = get_grid_cell(d, gmin, gmax, p);
      /* _getGridDensityGradAndWind — densityGradAndWindGrid.wgsl.ts:20-44 */
      ov3 dgrad = o3(0, 0, 0);
      float wstr = 0.0f;
      for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
        for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
          for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
            ov3 w = cell_weights(x, y, z, co.pin);
            const fh_density_grad_wind* v = &dgw[grid_idx(d, x, y, z)];
            dgrad = add3(dgrad, mul3(w, o3(v->densityGrad[0], v->densityGrad[1], v->densityGrad[2])));
            wstr = wstr + len3(w) * v->windStrength;
          }
      /* _getGridDensityVelocity — densityVelocityGrid.wgsl.ts:30-68 */
      ov3 vel = o3(0, 0, 0);
      float dens = 0.0f;
      for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
        for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
          for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
            ov3 w = cell_weights(x, y, z, co.pin);
            const fh_density_velocity* v = &dv[grid_idx(d, x, y, z)];
            ov3 vv = o3(grid_decode(v->velocityX), grid_decode(v->velocityY), grid_decode(v->velocityZ));
            vel = add3(vel, mul3(w, vv));
            dens = dens + len3(w) * grid_decode(v->density);
          }
      /* wind :133-139 */
      float timer = (float)u->frameIdx * 0.73f + u->windPhaseOffset * (float)s;
      float windJitter = ofract(timer * u->windStrengthFrequency);
      float jitterDelta = omix(-0.5f * u->windStrengthJitter, 0.5f * u->windStrengthJitter, windJitter);
      float jitterStr = 1.0f + jitterDelta;
      float windCellStr = osat(omix(u->windStrengthLull, 1.0f, wstr));
      force = add3(force, muls3(o3(u->wind[0], u->wind[1], u->wind[2]),
                                fabsf(u->wind[3] * jitterStr * windCellStr)));
      /* density gradient :145 */
      force = add3(force, muls3(dgrad, u->volumePreservation));
      /* grid velocity :149-153 */
      ov3 gridDisp = dens > 0.0f ? divs3(vel, dens) : o3(0, 0, 0);
      /* verletIntegration — integration.wgsl.ts:8-28 */
      ov3 pointDisp = sub3(xyz(pn), xyz(pp));
      ov3 fdisp = mix3(pointDisp, gridDisp, u->friction);
      ov3 next = add3(add3(xyz(pn), fdisp), muls3(muls3(force, u->deltaTime), u->deltaTime));
      ov3 r = mix3(xyz(pp), next, pn.w);
      pos[i] = o4(r.x, r.y, r.z, pn.w);
    }

    /* constraints :168-262 */
    for (uint32_t it = 0; it < u->constraintIterations; it++) {
      /* length — constraints.wgsl.ts:5-28 */
      float kLen = u->stiffnessLengthConstr;
      ov4 p0 = pos[0];
      for (uint32_t j = 0; j + 1 < P; j++) {
        ov4 p1 = pos[j + 1];
        float expected = seglen[base + j];
        float w0 = p0.w, w1 = p1.w;
        ov3 tg = sub3(xyz(p1), xyz(p0));
        float actual = len3(tg);
        float corr = 1.0f - expected / actual;
        float df = corr * kLen / (w0 + w1 + FLOAT_EPSILON);
        ov3 delta = muls3(tg, df);
        ov3 a = add3(xyz(p0), muls3(delta, w0));
        ov3 b = sub3(xyz(p1), muls3(delta, w1));
        p0 = o4(a.x, a.y, a.z, p0.w);
        p1 = o4(b.x, b.y, b.z, p1.w);
        pos[j] = p0;
        pos[j + 1] = p1;
        p0 = p1;
      }
      /* global shape — constraints.wgsl.ts:31-49 */
      float kGlobal = u->stiffnessGlobalConstr / iters;
      for (uint32_t j = 1; j < P; j++) {
        float x = (float)j / (float)(P - 1u);
        float att = 1.0f - osat((x - u->globalConstrExtent) / u->globalConstrFade);
        ov4 ini = ld4(initial, base + j);
        float k = kGlobal * att;
        ov4 p = pos[j];
        ov3 dvv = sub3(xyz(ini), xyz(p));
        ov3 r = add3(xyz(p), muls3(muls3(dvv, p.w), k));
        pos[j] = o4(r.x, r.y, r.z, p.w);
      }
      /* local shape — constraints.wgsl.ts:54-87 */
      float kLocal = u->stiffnessLocalConstr / iters;
      uint32_t jj = 0;
      for (; jj < P - 3u; jj++) {
        ov4 i0 = ld4(initial, base + jj), i1 = ld4(initial, base + jj + 1), i2 = ld4(initial, base + jj + 2);
        ov3 ci = divs3(add3(add3(xyz(i0), xyz(i1)), xyz(i2)), 3.0f);
        float h0 = len3(sub3(ci, xyz(i1)));
        ov4 a = pos[jj], b = pos[jj + 1], c = pos[jj + 2];
        ov3 cc = divs3(add3(add3(xyz(a), xyz(b)), xyz(c)), 3.0f);
        ov3 hv = sub3(xyz(b), cc);
        float h = len3(hv);
        float wT = i0.w + 2.0f * i1.w + i2.w;
        float w0 = i0.w / wT * 2.0f;
        float w1 = i1.w / wT * -4.0f;
        float w2 = i2.w / wT * 2.0f;
        ov3 delta = muls3(hv, 1.0f - h0 / h);
        ov3 ra = add3(xyz(a), muls3(delta, kLocal * w0));
        ov3 rb = add3(xyz(b), muls3(delta, kLocal * w1));
        ov3 rc = add3(xyz(c), muls3(delta, kLocal * w2));
        pos[jj] = o4(ra.x, ra.y, ra.z, a.w);
        pos[jj + 1] = o4(rb.x, rb.y, rb.z, b.w);
        pos[jj + 2] = o4(rc.x, rc.y, rc.z, c.w);
      }
      /* tip tangent — constraints.wgsl.ts:90-120 */
      for (; jj < P; jj++) {
        ov4 i0 = ld4(initial, base + jj - 1), i1 = ld4(initial, base + jj);
        ov3 tgi = sub3(xyz(i1), xyz(i0));
        ov4 a = pos[jj - 1], b = pos[jj];
        ov3 proj = add3(xyz(a), tgi);
        ov3 delta = sub3(proj, xyz(b));
        ov3 r = add3(xyz(b), muls3(delta, kLocal * b.w));
        pos[jj] = o4(r.x, r.y, r.z, b.w);
      }
      /* collisions :237-261, collisions.wgsl.ts:4-61 */
      float kColl = u->stiffnessCollisions / iters;
      float kSdf = u->stiffnessSDF / iters;
      for (uint32_t j = 1; j < P; j++) {
        ov4 p = pos[j];
        ov3 v = sub3(xyz(p), xyz(sphere));
        float dist = len3(v) - sphere.w;
        float toShell = omax(0.0f, -dist);
        ov3 r = add3(xyz(p), muls3(muls3(norm3(v), toShell), kColl));
        p = o4(r.x, r.y, r.z, p.w);
        if (j > 2u) {
          ov3 q = xyz(p);
          float sd = sample_sdf(sdf, sdf_dims, smin, smax, q);
          float sdo = sd + sdfOffset;
          if (!(sdo >= 0.0f)) {
            ov3 cd;
            cd.x = sample_sdf(sdf, sdf_dims, smin, smax, add3(q, o3(SDF_GRAD_OFFSET, 0.0f, 0.0f)));
            cd.y = sample_sdf(sdf, sdf_dims, smin, smax, add3(q, o3(0.0f, SDF_GRAD_OFFSET, 0.0f)));
            cd.z = sample_sdf(sdf, sdf_dims, smin, smax, add3(q, o3(0.0f, 0.0f, SDF_GRAD_OFFSET)));
            cd = sub3(cd, o3(sd, sd, sd));
            ov3 corr = muls3(norm3(cd), -sdo);
            ov3 rr = add3(q, muls3(corr, kSdf));
            p = o4(rr.x, rr.y, rr.z, p.w);
          }
        }
        pos[j] = p;
      }
    }
    /* write back :268-272 */
    for (uint32_t i = 0; i < P; i++) {
      prev[base + i].x = pos[i].x;
      prev[base + i].y = pos[i].y;
      prev[base + i].z = pos[i].z;
      prev[base + i].w = pos[i].w;
    }
  }
  free(pos);
  }
}

/* ======================================================================== */
/* K3 — gridPostSimPass.wgsl.ts:45-72, _addGridDensityVelocity               */
/* densityVelocityGrid.wgsl.ts:93-124; host clear gridPostSimPass.ts:36.     */
/* Pinned: test_oracle_golden.py test_p2g_closed_forms (point on a node ->  */
/* one corner of weight |(1,1,1)|, cell centre -> 8 corners of              */
/* |(.5,.5,.5)|, density sum, velocity truncated toward zero).               */
/* ======================================================================== */

/* The contributions of strands [s0, s1) added into `g` (u32 wrap-around
 * adds: the i32 sum of the reference's atomics, in any order). */
static void p2g_range(uint32_t d, ov3 gmin, ov3 gmax, const fh_vec4* prev, const fh_vec4* now,
                      fh_density_velocity* g, uint32_t s0, uint32_t s1, uint32_t P) {
  for (uint32_t s = s0; s < s1; s++)
    for (uint32_t i = 0; i < P; i++) {
      size_t k = (size_t)s * P + i;
      ov3 pp = o3(prev[k].x, prev[k].y, prev[k].z);
      ov3 pn = o3(now[k].x, now[k].y, now[k].z);
      ov3 vel = sub3(pn, pp);
      grid_cell co = get_grid_cell(d, gmin, gmax, pn);
      for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
        for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
          for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
            uint32_t idx = grid_idx(d, x, y, z);
            ov3 w = cell_weights(x, y, z, co.pin);
            ov3 cv = mul3(vel, w);
            fh_density_velocity* c = &g[idx];
            c->velocityX = (int32_t)((uint32_t)c->velocityX + (uint32_t)oi32(cv.x * 1000000.0f));
            c->velocityY = (int32_t)((uint32_t)c->velocityY + (uint32_t)oi32(cv.y * 1000000.0f));
            c->velocityZ = (int32_t)((uint32_t)c->velocityZ + (uint32_t)oi32(cv.z * 1000000.0f));
            c->density = (int32_t)((uint32_t)c->density + (uint32_t)oi32(len3(w) * 1000000.0f));
          }
    }
}

void or_grid_post_sim(uint32_t d, const fh_sim_uniforms* u, const fh_vec4* prev,
                      const fh_vec4* now, fh_density_velocity* dv, uint32_t S,
                      uint32_t P) {
  ov3 gmin = o3(u->gridBoundsMin[0], u->gridBoundsMin[1], u->gridBoundsMin[2]);
  ov3 gmax = o3(u->gridBoundsMax[0], u->gridBoundsMax[1], u->gridBoundsMax[2]);
  const size_t nodes = (size_t)d * d * d;
  memset(dv, 0, sizeof(fh_density_velocity) * nodes);
  int T = g_threads;
  if (T <= 1 || S < 2u * (uint32_t)T) {
    p2g_range(d, gmin, gmax, prev, now, dv, 0, S, P);
    return;
  }
  /* one partial grid per thread over a contiguous strand range, then the
   * node-wise sum of the partials */
  fh_density_velocity* part = (fh_density_velocity*)calloc((size_t)T * nodes, sizeof(fh_density_velocity));
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; t++) {
    uint32_t s0 = (uint32_t)((uint64_t)S * t / T), s1 = (uint32_t)((uint64_t)S * (t + 1) / T);
    p2g_range(d, gmin, gmax, prev, now, part + (size_t)t * nodes, s0, s1, P);
  }
#pragma omp parallel for num_threads(T) schedule(static)
  for (size_t n = 0; n < nodes; n++) {
    uint32_t a = 0, b = 0, c = 0, e = 0;
    for (int t = 0; t < T; t++) {
      const fh_density_velocity* q = &part[(size_t)t * nodes + n];
      a += (uint32_t)q->velocityX; b += (uint32_t)q->velocityY;
      c += (uint32_t)q->velocityZ; e += (uint32_t)q->density;
    }
    dv[n].velocityX = (int32_t)a; dv[n].velocityY = (int32_t)b;
    dv[n].velocityZ = (int32_t)c; dv[n].density = (int32_t)e;
  }
  free(part);
}

/* ======================================================================== */
/* Raster helpers — swRasterizeHair.wgsl.ts:39-203, shaderSnippets.wgls.ts   */
/* :89-156, tileUtils.wgsl.ts:5-54.                                          */
/* Pinned: test_oracle_golden.py test_raster_helpers_match_reference_      */
/* snapshot, against the reference's own fixture                             */
/* src/passes/swHair/__test__/swRasterizer.snapshot.bin (bit for bit).       */
/* ======================================================================== */

typedef struct {
  ov2 v00, v01, v10, v11;
  ov4 depths;
} proj_seg;

/* projectHairPoint — swRasterizeHair.wgsl.ts:76-105 */
static void project_point(const fh_render_uniforms* u, const fh_vec4* positions,
                          const fh_vec4* tangents, uint32_t P, uint32_t strand,
                          uint32_t point, ov2 vp, float fiberRadius, ov2* v0,
                          ov2* v1, ov2* dd) {
  size_t k = (size_t)strand * P + point;
  ov4 p0 = mulmv(u->modelMatrix, o4(positions[k].x, positions[k].y, positions[k].z, 1.0f));
  ov4 t0 = mulmv(u->modelMatrix, o4(tangents[k].x, tangents[k].y, tangents[k].z, 1.0f));
  ov3 cam = o3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]);
  ov3 toCam = norm3(sub3(cam, xyz(p0)));
  ov3 right = muls3(norm3(cross3(xyz(t0), toCam)), fiberRadius);
  ov3 a = sub3(xyz(p0), right), b = add3(xyz(p0), right);
  ov3 an = oproject(u->vpMatrix, o4(a.x, a.y, a.z, 1.0f));
  ov3 bn = oproject(u->vpMatrix, o4(b.x, b.y, b.z, 1.0f));
  /* ndc2viewportPx — shaderSnippets.wgls.ts:104-107 */
  *v0 = o2((an.x * 0.5f + 0.5f) * vp.x, (an.y * 0.5f + 0.5f) * vp.y);
  *v1 = o2((bn.x * 0.5f + 0.5f) * vp.x, (bn.y * 0.5f + 0.5f) * vp.y);
  *dd = o2(an.z, bn.z);
}

/* projectHairSegment — swRasterizeHair.wgsl.ts:39-68 */
static proj_seg project_segment(const fh_render_uniforms* u, const fh_vec4* positions,
                                const fh_vec4* tangents, uint32_t P, uint32_t strand,
                                uint32_t seg, ov2 vp, float fiberRadius) {
  proj_seg r;
  ov2 a, b, dd;
  project_point(u, positions, tangents, P, strand, seg, vp, fiberRadius, &a, &b, &dd);
  r.v00 = a; r.v01 = b; r.depths.x = dd.x; r.depths.y = dd.y;
  project_point(u, positions, tangents, P, strand, seg + 1, vp, fiberRadius, &a, &b, &dd);
  r.v10 = a; r.v11 = b; r.depths.z = dd.x; r.depths.w = dd.y;
  return r;
}

/* edgeFunction — swRasterizeHair.wgsl.ts:127-129 */
static float edge_fn(ov2 v0, ov2 v1, ov2 p) {
  return (p.x - v0.x) * (v1.y - v0.y) - (p.y - v0.y) * (v1.x - v0.x);
}

static int inside_quad(const proj_seg* s, ov2 p) {
  float c0 = edge_fn(s->v01, s->v00, p);
  float c1 = edge_fn(s->v11, s->v01, p);
  float c2 = edge_fn(s->v10, s->v11, p);
  float c3 = edge_fn(s->v00, s->v10, p);
  return c0 >= 0.0f && c1 >= 0.0f && c2 >= 0.0f && c3 >= 0.0f;
}

/* projectPointToLine — shaderSnippets.wgls.ts:95-102 */
static ov2 project_to_line(ov2 l1, ov2 l2, ov2 p) {
  ov2 ab = sub2(l2, l1), ac = sub2(p, l1);
  ov2 ad = divs2(muls2(ab, dot2(ab, ac)), dot2(ab, ab));
  return add2(l1, ad);
}

/* interpolateHairQuad — swRasterizeHair.wgsl.ts:170-197 */
static ov2 interp_quad(const proj_seg* s, ov2 c) {
  ov2 sm = divs2(add2(s->v00, s->v01), 2.0f);
  ov2 em = divs2(add2(s->v10, s->v11), 2.0f);
  ov2 cp = project_to_line(sm, em, c);
  float d1 = len2(sub2(cp, sm)) / len2(sub2(sm, em));
  float ws = len2(sub2(s->v00, s->v01));
  float we = ws * dot2(norm2(sub2(s->v00, s->v01)), norm2(sub2(s->v10, s->v11)));
  float ew = omix(ws, we, d1);
  ov2 e1 = project_to_line(s->v00, s->v10, c);
  float d0 = len2(sub2(c, e1)) / ew;
  return o2(osat(d0), osat(d1));
}

/* interpolateHairF32 — swRasterizeHair.wgsl.ts:199-203 */
static float interp_f32(ov2 w, ov4 v) {
  float a = omix(v.x, v.y, w.x);
  float b = omix(v.z, v.w, w.x);
  return omix(a, b, w.y);
}

/* getDepthBin — shaderSnippets.wgls.ts:148-156 */
static uint32_t depth_bin(uint32_t n, ov2 td, float d) {
  float span = fabsf(td.y - td.x);
  float t = (d - td.x) / span;
  return oclampu(ou32(t * (float)n), 0u, n - 1u);
}

static float tex_load(const float* tex, uint32_t w, uint32_t h, int32_t x, int32_t y) {
  if (x < 0 || y < 0 || (uint32_t)x >= w || (uint32_t)y >= h) return 0.0f;
  return tex[(size_t)y * w + (uint32_t)x];
}

/* swRasterizer.test.ts:184-248 — the reference's raster-helper test kernel. */
void or_debug_raster(const fh_render_uniforms* u, const fh_vec4* positions,
                     const fh_vec4* tangents, uint32_t S, uint32_t P,
                     uint32_t W, uint32_t H, uint32_t* out) {
  ov2 vp = o2(u->viewport[0], u->viewport[1]);
  for (uint32_t s = 0; s < S; s++)
    for (uint32_t g = 0; g + 1 < P; g++) {
      proj_seg ps = project_segment(u, positions, tangents, P, s, g, vp, u->fiberRadius);
      /* getRasterizedHairBounds — swRasterizeHair.wgsl.ts:111-123 */
      ov2 bmax = max2(max2(ps.v00, ps.v01), max2(ps.v10, ps.v11));
      ov2 bmin = min2(min2(ps.v00, ps.v01), min2(ps.v10, ps.v11));
      bmax = min2(o2(ceilf(bmax.x), ceilf(bmax.y)), vp);
      bmin = max2(o2(floorf(bmin.x), floorf(bmin.y)), o2(0.0f, 0.0f));
      for (float y = bmin.y; y < bmax.y; y += 1.0f)
        for (float x = bmin.x; x < bmax.x; x += 1.0f) {
          ov2 p = o2(x, y);
          if (!inside_quad(&ps, p)) continue;
          ov2 w = interp_quad(&ps, p);
          uint32_t c0 = ou32(osat(w.x) * 255.0f);
          uint32_t c1 = ou32(osat(w.y) * 255.0f);
          uint32_t c2 = ou32(osat(0.1f) * 255.0f);
          uint32_t value = 0xff000000u | c0 | (c1 << 8) | (c2 << 16);
          uint32_t px = ou32(x), py = ou32(y);
          if (px >= W || py >= H) continue;
          uint32_t yy = H - py;
          size_t idx = (size_t)yy * W + px;
          if (idx >= (size_t)W * H) continue; /* dropped out-of-bounds write */
          if (out[idx] < value) out[idx] = value;
        }
    }
}

/* ======================================================================== */
/* K4 — hairTilesPass.wgsl.ts:70-246 with hairTilesResultBuffer.ts:14-42,   */
/* hairTileSegmentsBuffer.ts:22-29, segmentCountPerTileBuffer.ts:15-20.      */
/* Pinned: test_oracle_raster.py (depth bins with B = 8 at chosen view      */
/* depths, clamping on both sides, the least bin of a slanted strand, two    */
/* strands of one tile in their own bins, depth words of a flat strand, the  */
/* depth test at texture row H - y) and test_oracle_golden.py                */
/* test_fine_pass_closed_forms (list membership and depth words, one bin).   */
/* ======================================================================== */

static uint32_t div_ceil(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }

/* One (segment, tile) pair that passed the raster: what the append needs. */
typedef struct {
  uint32_t s, g, tx, ty, dbin;
  float dmin, dmax;
} tile_hit;

typedef struct {
  tile_hit* v;
  size_t n, cap;
} hit_vec;

static void hit_push(hit_vec* h, tile_hit x) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 256;
    h->v = (tile_hit*)realloc(h->v, h->cap * sizeof(tile_hit));
  }
  h->v[h->n++] = x;
}

void or_tiles_draw(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                   const fh_render_uniforms* u, const fh_hair_data* hd,
                   const fh_vec4* positions, const fh_vec4* tangents,
                   fh_tile_result* tiles, const float* depth, void* segbuf,
                   uint32_t capacity, uint32_t* countPerTile,
                   uint32_t* slotKey, uint32_t S, uint32_t P) {
  const uint32_t T = cfg->tileSize, B = cfg->tileDepthBins;
  uint32_t* counter = (uint32_t*)segbuf;
  fh_list_element* data = (fh_list_element*)((char*)segbuf + 4);
  ov2 vp = o2(u->viewport[0], u->viewport[1]);
  uint32_t vpW = ou32(vp.x), vpH = ou32(vp.y);
  (void)W; (void)H;
  uint32_t tcx = div_ceil(vpW, T);
  uint32_t maxDrawn = u->maxDrawnHairSegments;
  uint32_t limit = maxDrawn < capacity ? maxDrawn : capacity;
  /* getHairDepthBoundsVS :242-246 */
  ov4 bs = o4(hd->boundingSphere[0], hd->boundingSphere[1], hd->boundingSphere[2], hd->boundingSphere[3]);
  ov4 bsc = mulmv(u->modelViewMat, o4(bs.x, bs.y, bs.z, 1.0f));
  ov2 dbVS = o2(bsc.z - bs.w, bsc.z + bs.w);

  /* The raster work of each (segment, tile) pair is independent: it runs
   * over chunks of strands (in parallel when threads > 1), collecting the
   * pairs that pass in (strand, segment, tile row, tile column) order. The
   * appends then run in that order, one after another, which is the
   * schedule a sequential run of the reference's threads produces. */
  const uint32_t CHUNK = 64;
  const uint32_t nChunks = div_ceil(S, CHUNK);
  hit_vec* hits = (hit_vec*)calloc(nChunks ? nChunks : 1, sizeof(hit_vec));
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
  for (uint32_t c = 0; c < nChunks; c++) {
    uint32_t s1 = (c + 1) * CHUNK < S ? (c + 1) * CHUNK : S;
    for (uint32_t s = c * CHUNK; s < s1; s++)
      for (uint32_t g = 0; g + 1 < P; g++) {
        proj_seg ps = project_segment(u, positions, tangents, P, s, g, vp, u->fiberRadius);
        ov2 bmax = max2(max2(ps.v00, ps.v01), max2(ps.v10, ps.v11));
        ov2 bmin = min2(min2(ps.v00, ps.v01), min2(ps.v10, ps.v11));
        bmax = min2(o2(ceilf(bmax.x), ceilf(bmax.y)), vp);
        bmin = max2(o2(floorf(bmin.x), floorf(bmin.y)), o2(0.0f, 0.0f));
        uint32_t tminx = ou32(bmin.x) / T, tminy = ou32(bmin.y) / T;
        uint32_t tmaxx = ou32(bmax.x) / T, tmaxy = ou32(bmax.y) / T;
        uint32_t tsx = (tmaxx - tminx) + 1u, tsy = (tmaxy - tminy) + 1u;
        if (tsx * tsy > cfg->invalidTilesPerSegmentThreshold) continue;
        for (uint32_t ty = tminy; ty <= tmaxy; ty++)
          for (uint32_t tx = tminx; tx <= tmaxx; tx++) {
            /* processTile :139-238; getTileBoundsPx tileUtils.wgsl.ts:44-54 */
            uint32_t x0 = oclampu(tx * T, 0, vpW - 1), y0 = oclampu(ty * T, 0, vpH - 1);
            uint32_t x1 = oclampu((tx + 1) * T, 0, vpW - 1), y1 = oclampu((ty + 1) * T, 0, vpH - 1);
            float dmin = 999.0f, dmax = -999.0f;
            uint32_t dbin = B;
            for (uint32_t y = y0; y < y1; y++)
              for (uint32_t x = x0; x < x1; x++) {
                ov2 p = o2((float)x, (float)y);
                if (!inside_quad(&ps, p)) continue;
                ov2 w = interp_quad(&ps, p);
                float hdp = interp_f32(w, ps.depths);
                float db = tex_load(depth, vpW, vpH, (int32_t)x, (int32_t)(vpH - y));
                if (hdp > db) continue;
                ov3 vs = oproject(u->projMatrixInv, o4(p.x, p.y, hdp, 1.0f));
                uint32_t bin = (B - 1u) - depth_bin(B, dbVS, vs.z);
                dmin = omin(dmin, hdp);
                dmax = omax(dmax, hdp);
                dbin = bin < dbin ? bin : dbin;
              }
            if (dmin > 1.0f) continue;
            tile_hit h = {s, g, tx, ty, dbin, dmin, dmax};
            hit_push(&hits[c], h);
          }
      }
  }
  for (uint32_t c = 0; c < nChunks; c++) {
    for (size_t k = 0; k < hits[c].n; k++) {
      const tile_hit* h = &hits[c].v[k];
      uint32_t nextPtr = (*counter)++;
      if (nextPtr < limit) {
        uint32_t tileIdx = h->ty * tcx * B + h->tx * B + h->dbin;
        fh_tile_result* t = &tiles[tileIdx];
        uint32_t dMaxU = ou32(h->dmax * 4294967296.0f);
        uint32_t dMinU = ou32((1.0f - h->dmin) * 4294967296.0f);
        if (t->maxDepth < dMaxU) t->maxDepth = dMaxU;
        if (t->minDepth < dMinU) t->minDepth = dMinU;
        uint32_t last = t->tileSegmentPtr;
        t->tileSegmentPtr = nextPtr + 1u;
        t->padding0 += 1u;
        uint32_t prevPtr = last == 0u ? INVALID_PTR : last - 1u;
        data[nextPtr].strandAndSegmentIdxs = (h->g << 24) | h->s;
        data[nextPtr].next = prevPtr;
        if (slotKey) slotKey[nextPtr] = tileIdx;
        countPerTile[h->ty * tcx + h->tx] += 1u;
      }
    }
    free(hits[c].v);
  }
  free(hits);
}

/* ======================================================================== */
/* K5 + K6 — hairTileSortPass.countTiles.wgsl.ts:63-97,                      */
/* hairTileSortPass.sort.wgsl.ts:37-75.                                      */
/* Pinned: test_oracle_golden.py test_tile_sort_reference_vectors, the     */
/* inline vectors of hairTileSortPass.test.ts:47-53 (histogram closed form   */
/* :160-183 and non-increasing bucket order :122-127).                       */
/* ======================================================================== */

void or_tiles_sort(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                   const uint32_t* countPerTile, uint32_t* tileList,
                   fh_sort_bucket* buckets) {
  uint32_t n = div_ceil(W, cfg->tileSize) * div_ceil(H, cfg->tileSize);
  uint32_t NB = cfg->sortBuckets;
  for (uint32_t t = 0; t < n; t++) {
    uint32_t c = countPerTile[t];
    if (c > 0u) buckets[oclampu(c / cfg->sortBucketSize, 0, NB - 1)].tileCount += 1u;
  }
  uint32_t* offs = (uint32_t*)calloc(NB, sizeof(uint32_t));
  uint32_t off = 0;
  for (uint32_t i = 0; i < NB; i++) {
    uint32_t idx = NB - 1u - i;
    offs[idx] = off;
    off += buckets[idx].tileCount;
  }
  for (uint32_t t = 0; t < n; t++) {
    uint32_t c = countPerTile[t];
    if (c > 0u) {
      uint32_t b = oclampu(c / cfg->sortBucketSize, 0, NB - 1);
      uint32_t o = offs[b] + buckets[b].writeOffset++;
      tileList[1 + o] = t;
      tileList[0] += 1u;
    }
  }
  free(offs);
}

/* ======================================================================== */
/* K7 — hairFinePass.wgsl.ts:126-335, processHairSegment.wgsl.ts:5-70,       */
/* reduceHairSlices.wgsl.ts:9-78, hairSlicesDataBuffer.ts:42-78,             */
/* hairSliceHeadsBuffer.ts:136-180, hairRasterizerResultBuffer.ts:14-34.     */
/* Pinned: test_oracle_golden.py test_fine_pass_closed_forms (two strands  */
/* in different slices -> front-to-back "over" at coverage 0.5 and 0.25,     */
/* the a >= 0.999 cutoff) and test_oracle_raster.py (slice-pool exhaustion:  */
/* the strands that fit the pool are blended and the rest of the bin is      */
/* dropped; the depth test at texture row H - y; canonical list order). The  */
/* f16 pack is pinned with the combine's converters in test_oracle_post.py.  */
/* canonical != 0 walks each (tile, bin) list in descending (strand, seg)   */
/* order, i.e. the order this oracle's own tile pass produces.              */
/* ======================================================================== */

static int cmp_desc_key(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  uint32_t kx = ((x & 0xffffffu) << 8) | (x >> 24), ky = ((y & 0xffffffu) << 8) | (y >> 24);
  return kx < ky ? 1 : (kx > ky ? -1 : 0);
}

void or_fine_raster(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                    const fh_render_uniforms* u, const fh_vec4* positions,
                    const fh_tile_result* tiles, const void* segbuf,
                    uint32_t capacity, void* rasterResult, const float* depth,
                    const fh_vec4* shading, const fh_vec4* tangents,
                    const uint32_t* tileList, uint32_t S, uint32_t P,
                    int canonical) {
  const uint32_t T = cfg->tileSize, B = cfg->tileDepthBins, NS = cfg->slicesPerPixel;
  const uint32_t PS = cfg->shadingPoints;
  const uint32_t PER_PROC = cfg->avgFragmentsPerSlice * NS * T * T;
  const fh_list_element* data = (const fh_list_element*)((const char*)segbuf + 4);
  uint32_t* queue = (uint32_t*)rasterResult;
  float* fb = (float*)((char*)rasterResult + 16);
  const size_t fbLen = (size_t)W * H - 1u; /* W*H*16 B buffer minus header */
  ov2 vpf = o2(u->viewport[0], u->viewport[1]);
  uint32_t vpW = ou32(vpf.x), vpH = ou32(vpf.y);
  (void)W; (void)H;
  uint32_t tcx = div_ceil(vpW, T), tcy = div_ceil(vpH, T);
  uint32_t tileCount = tcx * tcy;
  uint32_t maxDrawn = u->maxDrawnHairSegments;
  uint32_t maxProcessed = S * P;
  uint32_t dispMode = u->displayMode & 0xffu;
  uint32_t dbgMax = dispMode == 4u ? ((u->displayMode >> 8) & 0xffffu) : 0u;
  int isDbg = dbgMax != 0u;
  const uint32_t NPX = T * T;
  uint32_t drawnTiles = tileList[0];
  /* Tiles are independent (each writes only its own pixels, and its
   * processor state is reset per bin and per tile): they run in parallel
   * when threads > 1, each thread with its own processor state. */
#pragma omp parallel num_threads(g_threads)
  {
  uint32_t* heads = (uint32_t*)malloc(sizeof(uint32_t) * NPX * NS);
  uint32_t (*pool)[4] = (uint32_t(*)[4])malloc(sizeof(uint32_t) * 4 * PER_PROC);
  int* done = (int*)malloc(sizeof(int) * NPX);
  uint32_t listCap = 1024;
  uint32_t* listBuf = (uint32_t*)malloc(sizeof(uint32_t) * listCap);
  for (uint32_t i = 0; i < NPX * NS; i++) heads[i] = INVALID_PTR;

#pragma omp for schedule(dynamic, 1)
  for (uint32_t qi = 0; qi < drawnTiles; qi++) {
    uint32_t tileIdx = tileList[1 + qi];
    uint32_t row = tileIdx / tcx;
    uint32_t tx = tileIdx - tcx * row, ty = row;
    uint32_t bx0 = oclampu(tx * T, 0, vpW - 1), by0 = oclampu(ty * T, 0, vpH - 1);
    for (uint32_t i = 0; i < NPX; i++) done[i] = 0;
    for (uint32_t bin = 0; bin < B && tileIdx < tileCount; bin++) {
      uint32_t sliceOff = 0;
      const fh_tile_result* tr = &tiles[ty * tcx * B + tx * B + bin];
      /* _getTileDepth — hairTilesResultBuffer.ts:47-54 */
      ov2 tdepth = o2((float)(MAX_U32 - tr->minDepth) / 4294967296.0f,
                      (float)tr->maxDepth / 4294967296.0f);
      int binVisited = 0;
      if (tdepth.y != 0.0f) {
        binVisited = 1;
        /* gather the list */
        uint32_t n = 0;
        uint32_t ptr = tr->tileSegmentPtr == 0u ? INVALID_PTR : tr->tileSegmentPtr - 1u;
        while (n < maxProcessed && ptr < maxDrawn && ptr != INVALID_PTR && ptr < capacity) {
          if (n == listCap) {
            listCap *= 2;
            listBuf = (uint32_t*)realloc(listBuf, sizeof(uint32_t) * listCap);
          }
          listBuf[n] = data[ptr].strandAndSegmentIdxs;
          n++;
          ptr = data[ptr].next;
        }
        if (canonical) qsort(listBuf, n, sizeof(uint32_t), cmp_desc_key);
        for (uint32_t li = 0; li < n; li++) {
          uint32_t strand = listBuf[li] & 0x00ffffffu, seg = listBuf[li] >> 24;
          proj_seg ps = project_segment(u, positions, tangents, P, strand, seg, vpf, u->fiberRadius);
          for (uint32_t pix = 0; pix < NPX; pix++) {
            if (done[pix]) continue;
            uint32_t lx = pix % T, ly = pix / T;
            /* processHairSegment — processHairSegment.wgsl.ts:5-54 */
            ov2 p = o2((float)(bx0 + lx), (float)(by0 + ly));
            if (!inside_quad(&ps, p)) continue;
            ov2 w = interp_quad(&ps, p);
            float hdp = interp_f32(w, ps.depths);
            float db = tex_load(depth, vpW, vpH, oi32(p.x), oi32(vpf.y - p.y));
            if (hdp >= db) continue;
            uint32_t slot = sliceOff++;
            if (!(slot < PER_PROC)) continue;
            /* getAlphaCoverage :61-70 */
            float alpha = 1.0f - fabsf(w.x * 2.0f - 1.0f);
            if (cfg->alphaQuadratic) alpha = sqrtf(alpha);
            alpha = osat(alpha * cfg->alphaMultiplier);
            float tFull = ((float)seg + w.y) / (float)(P - 1u);
            /* _sampleShading — hairShadingBuffer.ts:29-38 */
            float a = tFull * ((float)PS - 1.0f);
            uint32_t au = ou32(a);
            uint32_t i0 = oclampu(au, 0, PS - 1), i1 = oclampu(au + 1u, 0, PS - 1);
            float fr = ofract(a);
            const fh_vec4* c0 = &shading[(size_t)strand * PS + i0];
            const fh_vec4* c1 = &shading[(size_t)strand * PS + i1];
            float cr = omix(c0->x, c1->x, fr), cg = omix(c0->y, c1->y, fr);
            float cb = omix(c0->z, c1->z, fr), ca = omix(c0->w, c1->w, fr);
            ca = ca * alpha;
            uint32_t sl = depth_bin(NS, tdepth, hdp);
            uint32_t hi = pix * NS + sl;
            uint32_t prevPtr = heads[hi];
            heads[hi] = slot;
            pool[slot][0] = opack2h(cr, cg);
            pool[slot][1] = opack2h(cb, ca);
            pool[slot][2] = prevPtr;
            pool[slot][3] = 0u;
          }
          if (!(sliceOff < PER_PROC)) break;
        }
        /* after the segment loop the reference sets the offset to the pool
         * size when the list ended, so the reduce below always runs for a
         * visited bin (hairFinePass.wgsl.ts:252-275) */
        for (uint32_t pix = 0; pix < NPX; pix++) {
          if (done[pix]) continue;
          /* reduceHairSlices — reduceHairSlices.wgsl.ts:9-73 */
          uint32_t lx = pix % T, ly = pix / T;
          uint32_t px = bx0 + lx, py = by0 + ly;
          size_t idx = (size_t)vpW * py + px;
          float fc[4] = {0, 0, 0, 0};
          if (idx < fbLen) memcpy(fc, &fb[idx * 4], 16);
          uint32_t cnt = isDbg ? ou32(fc[0] * (float)dbgMax) : 0u;
          uint32_t s = 0;
          for (; s < NS; s++) {
            if (!isDbg && fc[3] >= 0.999f) break;
            int clr = 0;
            uint32_t sp = heads[pix * NS + s];
            while (sp != INVALID_PTR && sp < PER_PROC) {
              clr = 1;
              if (!isDbg && fc[3] >= 0.999f) break;
              uint32_t* e = pool[sp];
              sp = e[2];
              cnt++;
              float sc[4] = {oh2f(e[0] & 0xffffu), oh2f(e[0] >> 16), oh2f(e[1] & 0xffffu), oh2f(e[1] >> 16)};
              float oneMinusA = 1.0f - fc[3];
              for (int k = 0; k < 4; k++) fc[k] = fc[k] + sc[k] * oneMinusA;
            }
            if (clr) heads[pix * NS + s] = INVALID_PTR;
          }
          for (; s < NS; s++) heads[pix * NS + s] = INVALID_PTR;
          if (isDbg) {
            float c = osat((float)cnt / (float)dbgMax);
            fc[0] = c; fc[1] = 0.0f; fc[2] = 0.0f; fc[3] = 1.0f;
          }
          if (px < vpW && py < vpH && idx < fbLen) memcpy(&fb[idx * 4], fc, 16);
          done[pix] = !isDbg && fc[3] >= 0.999f;
        }
      }
      (void)binVisited;
      /* checkAllPixelsInWkgrpDone :305-327 */
      int all = 1;
      for (uint32_t i = 0; i < NPX; i++) all &= done[i];
      if (all) break;
    }
  }
  free(heads); free(pool); free(done); free(listBuf);
  }
  /* getNextTileIdx: every processor took one queue index per tile */
  *queue += drawnTiles;
}

/* ======================================================================== */
/* K8 — hairShadingPass.wgsl.ts:61-282, marschner.wgsl.ts:19-170,            */
/* pbr.wgsl.ts:22-24, noise.wgsl.ts:45-51, sampleShadows.wgsl.ts:21-150,     */
/* textureAo.wgsl.ts:7-12, linearDepth.wgsl.ts:9-27.                         */
/* Pinned: test_oracle_shading.py (Kajiya-Kay diffuse and the fake         */
/* multiple scattering at chosen angles; each Marschner lobe against the     */
/* Karis16 formulas in float64, with the R lobe's closed form at normal      */
/* incidence; PCF / PCSS hit counts over a step shadow map and the bias;     */
/* the AO texel and strength; attenuation from linearizeDepth; randomRGB,    */
/* the root-to-tip colour mix, the luma jitter and the tip alpha).           */
/* ======================================================================== */

#define O_PI 3.14159265358979323846f

static float fresnel1(float cosT, float f0) {
  return f0 + (1.0f - f0) * powf(1.0f - cosT, 5.0f);
}
static float cos_half(float c) { return sqrtf(osat(0.5f + 0.5f * c)); }
static float long_gauss(float si, float sr, float alpha, float beta) {
  const float SQRT_TWO_PI = 2.5066282746310002f;
  float a = si + sr - alpha;
  float b = -(a * a) / (2.0f * beta * beta);
  return expf(b) / (beta * SQRT_TWO_PI);
}

typedef struct {
  ov3 baseColor;
  float specular, weightTT, weightTRT, shift, roughness;
} marschner_params;

static ov3 marschner(const marschner_params* p, ov3 L, ov3 V, ov3 T) {
  const float HAIR_F0 = (float)((1.0 - 1.55) * (1.0 - 1.55) / ((1.0 + 1.55) * (1.0 + 1.55)));
  float si = dot3(T, L), sr = dot3(T, V);
  float ci = sqrtf(omax(0.0f, 1.0f - si * si));
  float cr = sqrtf(omax(0.0f, 1.0f - sr * sr));
  float cdf = ci * cr + si * sr;
  float cd = cos_half(cdf);
  ov3 lp = sub3(L, muls3(T, si)), cp = sub3(V, muls3(T, sr));
  float cphi = dot3(lp, cp) * (1.0f / sqrtf(len3(lp) * len3(cp) + 1e-4f));
  float chp = cos_half(cphi);
  float alpha = -2.0f * p->shift, beta = p->roughness * p->roughness;
  float Mr = long_gauss(si, sr, alpha, beta);
  float Nr = 0.25f * chp * fresnel1(cd, HAIR_F0);
  alpha = p->shift; beta = 0.5f * p->roughness * p->roughness;
  float Mtt = long_gauss(si, sr, alpha, beta);
  /* azimuthalScattering_TT :412-434 */
  float eta = (1.19f / cd) + (0.36f * cd);
  float a = 1.0f / eta;
  float hTT = (1.0f + a * (0.6f - 0.8f * cphi)) * chp;
  float Tnum = sqrtf(1.0f - hTT * hTT * a * a);
  float Tden = 2.0f * cd;
  ov3 TTT = o3(powf(p->baseColor.x, Tnum / Tden), powf(p->baseColor.y, Tnum / Tden), powf(p->baseColor.z, Tnum / Tden));
  float DTT = expf(-3.65f * cphi - 3.98f);
  float fa = cd * sqrtf(osat(1.0f - hTT * hTT));
  float f = fresnel1(fa, HAIR_F0);
  float fTT = (1.0f - f) * (1.0f - f);
  ov3 Ntt = muls3(muls3(TTT, fTT), DTT);
  alpha = 4.0f * p->shift; beta = 2.0f * p->roughness * p->roughness;
  float Mtrt = long_gauss(si, sr, alpha, beta);
  /* azimuthalScattering_TRT :437-454 */
  float e = 0.8f / cd;
  ov3 TTRT = o3(powf(p->baseColor.x, e), powf(p->baseColor.y, e), powf(p->baseColor.z, e));
  float DTRT = expf(17.0f * cphi - 16.78f);
  float f2 = fresnel1(cd * 0.5f, HAIR_F0);
  float fTRT = (1.0f - f2) * (1.0f - f2) * f2;
  ov3 Ntrt = muls3(muls3(TTRT, fTRT), DTRT);
  float r = p->specular * Mr * Nr;
  ov3 res = add3(add3(o3(r, r, r), muls3(Ntt, p->weightTT * Mtt)), muls3(Ntrt, p->weightTRT * Mtrt));
  return res;
}

/* textureGather on a clamp-to-edge sampler: texels (i0,j1),(i1,j1),(i1,j0),(i0,j0) */
static float gather_avg(const float* tex, uint32_t w, uint32_t h, ov2 uv) {
  float x = uv.x * (float)w - 0.5f, y = uv.y * (float)h - 0.5f;
  int i0 = (int)floorf(x), j0 = (int)floorf(y);
  int i1 = i0 + 1, j1 = j0 + 1;
#define CL(v, n) ((v) < 0 ? 0 : ((v) > (int)(n) - 1 ? (int)(n) - 1 : (v)))
  i0 = CL(i0, w); i1 = CL(i1, w); j0 = CL(j0, h); j1 = CL(j1, h);
#undef CL
  float gx = tex[(size_t)j1 * w + i0], gy = tex[(size_t)j1 * w + i1];
  float gz = tex[(size_t)j0 * w + i1], gw = tex[(size_t)j0 * w + i0];
  return (gx + gy + gz + gw) / 4.0f;
}

static float linearize_depth(float d) {
  const float n = 0.01f, f = 100.0f;
  return n * f / (f + (n - f) * d);
}

static ov3 hair_shading(const marschner_params* p, const fh_light* light, ov3 toCam,
                        ov4 tangentWS, ov4 posWS, float shadow, float ao) {
  ov3 L = norm3(sub3(o3(light->position[0], light->position[1], light->position[2]), xyz(posWS)));
  ov3 T = xyz(tangentWS);
  /* KajiyaKayDiffuse :209-213 */
  ov3 diffuse = muls3(p->baseColor, omix(0.25f, 1.0f, dot3(T, L)));
  /* fakeMultipleScattering :216-234 */
  ov3 fn = norm3(sub3(toCam, muls3(T, dot3(toCam, T))));
  float NoL = osat(dot3(fn, L));
  float ds = (NoL + 1.0f) / (4.0f * O_PI);
  float luma = 0.2126729f * p->baseColor.x + 0.7151522f * p->baseColor.y + 0.0721750f * p->baseColor.z;
  ov3 bl = divs3(p->baseColor, luma);
  ov3 tint = o3(powf(bl.x, 1.0f - shadow), powf(bl.y, 1.0f - shadow), powf(bl.z, 1.0f - shadow));
  ov3 ms = mul3(muls3(o3(sqrtf(p->baseColor.x), sqrtf(p->baseColor.y), sqrtf(p->baseColor.z)), ds), tint);
  ov3 dt = muls3(mul3(diffuse, ms), osat(dot3(T, L)));
  ov3 spec = marschner(p, L, toCam, T);
  ov3 brdf = add3(dt, spec);
  ov3 radiance = muls3(o3(light->colorAndEnergy[0], light->colorAndEnergy[1], light->colorAndEnergy[2]),
                       light->colorAndEnergy[3]);
  return muls3(mul3(brdf, radiance), 1.0f - ao);
}

void or_shading(const fh_render_cfg* cfg, const fh_render_uniforms* u,
                const fh_vec4* positions, const fh_vec4* tangents,
                fh_vec4* out, const float* shadowTex, uint32_t shadowW,
                uint32_t shadowH, const float* aoTex, uint32_t aoW,
                uint32_t aoH, const float* depth, uint32_t dW, uint32_t dH,
                uint32_t S, uint32_t P) {
  const uint32_t PS = cfg->shadingPoints;
  ov2 vp = o2(u->viewport[0], u->viewport[1]);
  ov3 cam = o3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]);
  const fh_hair_material* m = &u->hairMaterial;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 64)
  for (uint32_t s = 0; s < S; s++)
    for (uint32_t id = 0; id < PS; id++) {
      float t = (float)id / (float)PS;
      /* remapToIndices — shaderSnippets.wgls.ts:120-126 */
      float a = t * ((float)P - 1.0f);
      uint32_t au = ou32(a);
      uint32_t ix = oclampu(au, 0, P - 1), iy = oclampu(au + 1u, 0, P - 1);
      float tIn = ofract(a);
      const fh_vec4* q0 = &positions[(size_t)s * P + ix];
      const fh_vec4* q1 = &positions[(size_t)s * P + iy];
      ov3 s0 = o3(q0->x, q0->y, q0->z), s1 = o3(q1->x, q1->y, q1->z);
      ov3 t0 = norm3(sub3(s1, s0));
      const fh_vec4* tt = &tangents[(size_t)s * P + iy];
      ov3 t1 = o3(tt->x, tt->y, tt->z);
      ov3 pObj = mix3(s0, s1, tIn), tObj = mix3(t0, t1, tIn);
      ov4 posWS = mulmv(u->modelMatrix, o4(pObj.x, pObj.y, pObj.z, 1.0f));
      ov4 tanWS = mulmv(u->modelMatrix, o4(tObj.x, tObj.y, tObj.z, 1.0f));
      ov3 pproj = oproject(u->vpMatrix, posWS);
      ov2 p01 = o2(pproj.x * 0.5f + 0.5f, pproj.y * 0.5f + 0.5f);
      ov2 texPx = o2(vp.x * p01.x, vp.y * (1.0f - p01.y));
      /* base colour :105-114 */
      ov3 c0 = o3(m->color0[0], m->color0[1], m->color0[2]);
      ov3 c1 = o3(m->color1[0], m->color1[1], m->color1[2]);
      ov3 base = mix3(c0, c1, t);
      float fs = (float)s;
      ov3 rnd = sat3(o3(ofract(fs * 1.73f) * 1.0f, ofract(fs * 1.17f) * 1.0f, ofract(fs * 1.31f) * 1.0f));
      base = mix3(base, rnd, m->colorRng);
      float rb = ofract(fs / 255.0f);
      base = mix3(base, muls3(base, rb), m->lumaRng);
      ov3 toCam = norm3(sub3(cam, xyz(posWS)));
      marschner_params mp = {base, m->specular, m->weightTT, m->weightTRT, m->shift, m->roughness};
      /* shadow — getShadow :236-254, calculateDirectionalShadow
       * sampleShadows.wgsl.ts:47-84 */
      ov4 lp = mulmv(u->shadows.sourceMVP_Matrix, posWS);
      ov3 pss = o3(lp.x * 0.5f + 0.5f, 1.0f - (lp.y * 0.5f + 0.5f), lp.z);
      ov3 ssp = o3(u->shadows.sourcePosition[0], u->shadows.sourcePosition[1], u->shadows.sourcePosition[2]);
      ov3 toSrc = norm3(sub3(ssp, xyz(posWS)));
      float bias = omax(fabsf(u->shadows.bias) * (1.0f - dot3(xyz(tanWS), toSrc)), 0.0005f);
      int32_t radius = (int32_t)u->shadows.PCF_Radius;
      if (u->shadows.usePCSS > 0u) {
        float sm = gather_avg(shadowTex, shadowW, shadowH, o2(pss.x, pss.y));
        float dd = omax(pss.z - sm, 0.0f);
        radius = 1 + oi32(dd / sm * 10.0f);
      }
      float texel = 1.0f / (float)shadowW;
      float sh = 0.0f;
      for (int32_t x = -radius; x <= radius; x++)
        for (int32_t y = -radius; y <= radius; y++) {
          ov2 off = o2((float)x * texel, (float)y * texel);
          float sm = gather_avg(shadowTex, shadowW, shadowH, add2(o2(pss.x, pss.y), off));
          sh += (pss.z - bias > sm) ? 1.0f : 0.0f;
        }
      float pcf = (float)(radius * 2 + 1);
      float shadowRes = sh / (pcf * pcf);
      int valid = !(pss.z > 1.0f) && !(pss.x < 0.0f || pss.x > 1.0f || pss.y < 0.0f || pss.y > 1.0f);
      if (!valid) shadowRes = 0.0f;
      float shadow = 1.0f - shadowRes;
      shadow = oclamp(shadow, 1.0f - m->shadows, 1.0f);
      /* ao — textureAo.wgsl.ts:7-12 */
      ov2 at = o2(texPx.x / vp.x, texPx.y / vp.y);
      int32_t ax = oi32((float)aoW * at.x), ay = oi32((float)aoH * at.y);
      float aov = tex_load(aoTex, aoW, aoH, ax, ay);
      float ao = 1.0f - omix(1.0f, aov, u->ao.strength);
      /* attenuation :259-282 */
      float myD = linearize_depth(pproj.z);
      float zb = tex_load(depth, dW, dH, (int32_t)ou32(texPx.x), (int32_t)ou32(texPx.y));
      float att = fabsf(myD - linearize_depth(zb)) * m->attenuation;
      ao = osat(ao + att);
      ov3 rs = o3(0, 0, 0);
      rs = add3(rs, hair_shading(&mp, &u->light0, toCam, tanWS, posWS, shadow, ao));
      rs = add3(rs, hair_shading(&mp, &u->light1, toCam, tanWS, posWS, shadow, ao));
      rs = add3(rs, hair_shading(&mp, &u->light2, toCam, tanWS, posWS, shadow, ao));
      float alpha = id >= PS - 1u ? 0.0f : 1.0f;
      fh_vec4* o = &out[(size_t)s * PS + id];
      o->x = rs.x; o->y = rs.y; o->z = rs.z; o->w = alpha;
    }
}

/* ======================================================================== */
/* Combine and present                                                       */
/* ======================================================================== */

/* IEEE binary16 conversions as exported test hooks. Pinned:
 * test_oracle_post.py compares them with numpy's float16 over every half
 * and over random and boundary floats. */
uint32_t or_f2h(float f) { return of2h(f); }
float or_h2f(uint32_t h) { return oh2f(h); }

/* HairCombinePass fragment shader (hairCombinePass.wgsl.ts:51-162) with the
 * pipeline's blend state (hairCombinePass.ts:204-217) for every texel of the
 * rgba16float HDR texture (texture row ty <- rasterizer row H - 1 - ty).
 * Pinned: test_oracle_post.py (y flip, alpha 0 / 1 / 0.5 closed forms, the
 * debug tile pattern, list-walk counts against segment_count_per_tile). */
void or_hair_combine(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                     const fh_render_uniforms* u, const fh_tile_result* tiles,
                     const void* segbuf, const void* rasterResult,
                     const uint32_t* countPerTile, uint16_t* hdr) {
  const fh_list_element* segs =
      (const fh_list_element*)((const char*)segbuf + 4);
  const float* pix = (const float*)((const char*)rasterResult + 16);
  const uint32_t tile = cfg->tileSize, bins = cfg->tileDepthBins;
  const uint32_t tilesX = (W + tile - 1) / tile;
  const uint32_t mode = u->displayMode & 0xffu;
  const uint32_t extra = (u->displayMode >> 8) & 0xffffu;
  for (uint32_t ty = 0; ty < H; ty++) {
    for (uint32_t x = 0; x < W; x++) {
      /* fragment centre (x + .5, ty + .5); y is inverted (:60-63) */
      const float fy = (float)H - ((float)ty + 0.5f);
      const uint32_t py = ou32(fy);
      const uint32_t tileX = x / tile, tileY = py / tile;
      float c[4];
      if (mode == 1u || mode == 2u) {
        uint32_t segments = 0;
        if (mode == 1u) {
          segments = countPerTile[tileY * tilesX + tileX];
        } else {
          for (uint32_t b = 0; b < bins; b++) {
            uint32_t p = tiles[(tileY * tilesX + tileX) * bins + b].tileSegmentPtr;
            p = p == 0u ? INVALID_PTR : p - 1u; /* _translateHeadPointer */
            while (segments < extra) {
              if (p >= u->maxDrawnHairSegments || p == INVALID_PTR) break;
              segments++;
              p = segs[p].next;
            }
          }
        }
        if (segments == 0u) continue; /* discard */
        c[0] = (float)segments / (float)extra;
        c[1] = 1.0f - c[0];
        c[2] = 0.0f;
        c[3] = 1.0f;
      } else {
        const uint64_t i = (uint64_t)W * py + x;
        for (int k = 0; k < 4; k++)
          c[k] = i < (uint64_t)W * H - 1u ? pix[i * 4 + k] : 0.0f;
        if (c[3] <= 0.0f && mode == 0u && (extra & 1u)) {
          const float dbg[4] = {0.0f, (float)(tileX % 2u) / 2.0f,
                                (float)(tileY % 2u) / 2.0f, 1.0f};
          const float a = c[3];
          for (int k = 0; k < 4; k++) c[k] = dbg[k] + c[k] * a;
        }
      }
      uint16_t* t = hdr + ((size_t)ty * W + x) * 4;
      const float a = c[3];
      for (int k = 0; k < 3; k++)
        t[k] = (uint16_t)of2h(c[k] * a + oh2f(t[k]) * (1.0f - a));
      t[3] = (uint16_t)of2h(a + oh2f(t[3]));
    }
  }
}

/* Entry (x, y) of the 8x8 Bayer matrix by the recursive construction: bit i
 * of x ^ y goes to bit 5 - 2i, bit i of y to bit 4 - 2i. Pinned:
 * test_oracle_post.py against the 64 values of dither.wgsl.ts:12-21. */
uint32_t or_bayer8(uint32_t x, uint32_t y) {
  const uint32_t v = (x ^ y) & 7u;
  uint32_t r = 0;
  for (uint32_t i = 0; i < 3u; i++) {
    r |= ((v >> i) & 1u) << (5u - 2u * i);
    r |= ((y >> i) & 1u) << (4u - 2u * i);
  }
  return r;
}

/* doACES_Tonemapping (aces.wgsl.ts:5-13). Pinned: closed forms at 0, 1 and
 * the large-x limit a / c. */
float or_aces(float x) {
  return osat((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f));
}

/* Store of a linear value into an rgba8unorm-srgb texel (IEC 61966-2-1),
 * evaluated in double and rounded to nearest. Pinned: every 8-bit code
 * decoded by the inverse transfer function encodes back to itself. */
uint32_t or_srgb8(float c) {
  double v = (double)osat(c);
  double s = v <= 0.0031308 ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
  return (uint32_t)floor(s * 255.0 + 0.5);
}

/* PresentPass fragment shader (presentPass.wgsl.ts:58-103) for every texel:
 * FINAL / HW_RENDER -> ditherColor (dither.wgsl.ts:23-49), exposure, ACES,
 * saturate, gamma; DEPTH -> linearizeDepth_0_1 (linearDepth.wgsl.ts:9-36)
 * mapped from [0.002, 0.01]; NORMALS -> |decodeOctahedronNormal|
 * (shaderSnippets.wgls.ts:49-58); AO -> sampleAo (textureAo.wgsl.ts:7-12);
 * other modes pass the HDR colour through. Output rgba8 sRGB, alpha 255.
 * Pinned: test_oracle_post.py (gamma 1 / exposure / dither closed forms,
 * octahedral decode of the six axes, depth map end points). */
void or_present(uint32_t W, uint32_t H, const fh_render_uniforms* u,
                const uint16_t* hdr, const float* depth, const float* ao,
                uint32_t aoW, uint32_t aoH, const uint32_t* normals,
                int ditherPerPixel, uint8_t* out, uint32_t pitch) {
  const float zNear = 0.01f, zFar = 100.0f;
  const float gamma = u->colorMgmt[0], exposure = u->colorMgmt[1];
  const float ditherStr = u->colorMgmt[2];
  const uint32_t mode = u->displayMode & 0xffu;
  for (uint32_t y = 0; y < H; y++) {
    for (uint32_t x = 0; x < W; x++) {
      const uint16_t* t = hdr + ((size_t)y * W + x) * 4;
      float c[3] = {oh2f(t[0]), oh2f(t[1]), oh2f(t[2])};
      if (mode == 0u || mode == 3u) {
        const uint32_t m = ditherPerPixel ? or_bayer8(x % 8u, y % 8u) : or_bayer8(0, 0);
        const float dm = (float)m / 63.0f * ditherStr / 256.0f;
        for (int k = 0; k < 3; k++) {
          float v = c[k] + dm;
          v = v * exposure;
          v = osat(or_aces(v));
          c[k] = powf(v, 1.0f / gamma);
        }
      } else if (mode == 5u) {
        const float d = depth[(size_t)y * W + x];
        const float lin = zNear * zFar / (zFar + (zNear - zFar) * d) / (zFar - zNear);
        const float tt = osat((lin - 0.002f) / (0.01f - 0.002f));
        c[0] = c[1] = c[2] = omix(0.0f, 1.0f, tt);
      } else if (mode == 6u) {
        const uint32_t nn = normals ? normals[(size_t)y * W + x] : 0u;
        float n[3];
        n[0] = oh2f(nn & 0xffffu) * 2.0f - 1.0f;
        n[1] = oh2f(nn >> 16) * 2.0f - 1.0f;
        n[2] = 1.0f - fabsf(n[0]) - fabsf(n[1]);
        const float tt = osat(-n[2]);
        n[0] = n[0] >= 0.0f ? n[0] - tt : n[0] + tt;
        n[1] = n[1] >= 0.0f ? n[1] - tt : n[1] + tt;
        const ov3 nv = norm3(o3(n[0], n[1], n[2]));
        c[0] = fabsf(nv.x); c[1] = fabsf(nv.y); c[2] = fabsf(nv.z);
      } else if (mode == 7u) {
        const float tx = ((float)x + 0.5f) / u->viewport[0];
        const float ty = ((float)y + 0.5f) / u->viewport[1];
        const int32_t sx = oi32((float)aoW * tx), sy = oi32((float)aoH * ty);
        float v = 0.0f;
        if (sx >= 0 && sy >= 0 && (uint32_t)sx < aoW && (uint32_t)sy < aoH)
          v = ao[(size_t)sy * aoW + (uint32_t)sx];
        c[0] = c[1] = c[2] = v;
      }
      uint8_t* o = out + (size_t)y * pitch + 4u * x;
      for (int k = 0; k < 3; k++) o[k] = (uint8_t)or_srgb8(c[k]);
      o[3] = 255;
    }
  }
}

/* ======================================================================== */
/* Hair depth / normals, shadow-map hair, GTAO                               */
/* ======================================================================== */

/* The reference draws the hair's depth and normals (HwHairPass) and the hair
 * in the shadow map with the hardware rasterizer. This restatement fixes one
 * rasterizer: vertices snapped to 1/256 px, integer edge functions at pixel
 * centres with the top-left rule, screen-space interpolation of NDC z, depth
 * clip to [0, 1], no triangle with a vertex at w <= 0 or beyond 2^20 px, and
 * the pipelines' 'less' test in primitive order (triangles run in index
 * order; a later fragment must be strictly nearer). Pinned:
 * test_oracle_hwraster.py (watertight coverage of shared edges, exact pixel
 * sets of pixel-aligned quads, constant and planar depth, order of equal
 * depths, occlusion by the meshes, octahedral normals of the axes). */
static ov4 hw_edge_vertex(const fh_render_uniforms* u, int shadow, const fh_vec4* pos,
                          const fh_vec4* tan, uint32_t idx, uint32_t side) {
  const float* mv = shadow ? u->shadows.sourceModelViewMat : u->modelViewMat;
  const float* proj = shadow ? u->shadows.sourceProjMatrix : u->projMatrix;
  const float radius = shadow ? u->shadows.sourcePosition[3] : u->fiberRadius;
  const ov4 pv = mulmv(mv, o4(pos[idx].x, pos[idx].y, pos[idx].z, 1.0f));
  const ov4 tv = mulmv(mv, o4(tan[idx].x, tan[idx].y, tan[idx].z, 1.0f));
  ov3 right = cross3(xyz(tv), o3(0.0f, 0.0f, 1.0f));
  right = dot3(right, right) >= FLOAT_EPSILON ? norm3(right) : o3(0.0f, 0.0f, 0.0f);
  const ov3 th = muls3(right, radius);
  const ov3 e = side ? sub3(xyz(pv), th) : add3(xyz(pv), th);
  return mulmv(proj, o4(e.x, e.y, e.z, 1.0f));
}

typedef struct {
  ov4 c[3];
  long long X[3], Y[3], area;
  uint32_t vi[3];
  int ok;
} hw_tri;

static void hw_tri_vertices(uint32_t t, uint32_t P, uint32_t* idx, uint32_t* side) {
  const uint32_t seg = t >> 1, strand = seg / (P - 1u), j = seg - strand * (P - 1u);
  const uint32_t id = strand * P + j;
  if ((t & 1u) == 0u) {
    idx[0] = id; side[0] = 0; idx[1] = id; side[1] = 1; idx[2] = id + 1; side[2] = 0;
  } else {
    idx[0] = id + 1; side[0] = 0; idx[1] = id; side[1] = 1; idx[2] = id + 1; side[2] = 1;
  }
}

static hw_tri hw_setup(const fh_render_uniforms* u, int shadow, const fh_vec4* pos,
                       const fh_vec4* tan, uint32_t P, uint32_t W, uint32_t H, uint32_t t) {
  hw_tri s;
  uint32_t idx[3], side[3];
  hw_tri_vertices(t, P, idx, side);
  s.ok = 1;
  for (int k = 0; k < 3; k++) {
    s.c[k] = hw_edge_vertex(u, shadow, pos, tan, idx[k], side[k]);
    s.vi[k] = (uint32_t)k;
    s.X[k] = s.Y[k] = 0;
    if (!(s.c[k].w > 0.0f)) { s.ok = 0; continue; }
    const float nx = s.c[k].x / s.c[k].w, ny = s.c[k].y / s.c[k].w;
    const float xs = (float)W * (nx * 0.5f + 0.5f);
    const float ys = (float)H * (1.0f - (ny * 0.5f + 0.5f));
    if (!(fabsf(xs) < 1048576.0f && fabsf(ys) < 1048576.0f)) { s.ok = 0; continue; }
    s.X[k] = (long long)rintf(xs * 256.0f);
    s.Y[k] = (long long)rintf(ys * 256.0f);
  }
  s.area = 0;
  if (!s.ok) return s;
  long long area = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
  if (area < 0) {
    ov4 c = s.c[1]; s.c[1] = s.c[2]; s.c[2] = c;
    long long q = s.X[1]; s.X[1] = s.X[2]; s.X[2] = q;
    q = s.Y[1]; s.Y[1] = s.Y[2]; s.Y[2] = q;
    uint32_t v = s.vi[1]; s.vi[1] = s.vi[2]; s.vi[2] = v;
    area = -area;
  }
  s.area = area;
  s.ok = area > 0;
  return s;
}

static long long hw_edge(long long ax, long long ay, long long bx, long long by, long long px, long long py) {
  return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}
static long long hw_floor_div256(long long v) { return v >= 0 ? v / 256 : -((-v + 255) / 256); }
static long long hw_ceil_div256(long long v) { return -hw_floor_div256(-v); }
static int hw_top_left(long long ax, long long ay, long long bx, long long by) {
  const long long dx = bx - ax, dy = by - ay;
  return (dy == 0 && dx > 0) || dy < 0;
}

/* encodeOctahedronNormal (shaderSnippets.wgls.ts:25-46) as rg16float bits */
uint32_t or_encode_oct(float x, float y, float z) {
  const float s = fabsf(x) + fabsf(y) + fabsf(z);
  float nx = x / s, ny = y / s, nz = z / s;
  if (nz < 0.0f) {
    const float sx = nx >= 0.0f ? 1.0f : -1.0f, sy = ny >= 0.0f ? 1.0f : -1.0f;
    const float ax = (1.0f - fabsf(ny)) * sx, ay = (1.0f - fabsf(nx)) * sy;
    nx = ax;
    ny = ay;
  }
  return opack2h(nx * 0.5f + 0.5f, ny * 0.5f + 0.5f);
}

static ov3 hw_decode_oct(uint32_t packed) {
  const float fx = oh2f(packed & 0xffffu) * 2.0f - 1.0f, fy = oh2f(packed >> 16) * 2.0f - 1.0f;
  ov3 n = o3(fx, fy, 1.0f - fabsf(fx) - fabsf(fy));
  const float t = osat(-n.z);
  n.x = n.x >= 0.0f ? n.x - t : n.x + t;
  n.y = n.y >= 0.0f ? n.y - t : n.y + t;
  return norm3(n);
}

/* HwHairPass.cmdDrawHair (shadow = 0: depth + normals of the camera view) or
 * the hair draw of ShadowMapPass (shadow = 1: depth only, shadow source
 * matrices and fiber radius). depth is read and written in place; normals
 * (rg16float bits) may be NULL. */
void or_hw_hair(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u,
                const fh_vec4* pos, const fh_vec4* tan, float* depth, uint32_t* normals,
                uint32_t S, uint32_t P) {
  const uint64_t nTri = 2ull * S * (P - 1u);
  for (uint64_t t = 0; t < nTri; t++) {
    const hw_tri s = hw_setup(u, shadow, pos, tan, P, W, H, (uint32_t)t);
    if (!s.ok) continue;
    long long minX = s.X[0], maxX = s.X[0], minY = s.Y[0], maxY = s.Y[0];
    for (int k = 1; k < 3; k++) {
      if (s.X[k] < minX) minX = s.X[k];
      if (s.X[k] > maxX) maxX = s.X[k];
      if (s.Y[k] < minY) minY = s.Y[k];
      if (s.Y[k] > maxY) maxY = s.Y[k];
    }
    const int tl0 = hw_top_left(s.X[1], s.Y[1], s.X[2], s.Y[2]);
    const int tl1 = hw_top_left(s.X[2], s.Y[2], s.X[0], s.Y[0]);
    const int tl2 = hw_top_left(s.X[0], s.Y[0], s.X[1], s.Y[1]);
    /* pixel centres 256 p + 128 within [min, max], clipped to the target */
    long long x0 = hw_ceil_div256(minX - 128), x1 = hw_floor_div256(maxX - 128);
    long long y0 = hw_ceil_div256(minY - 128), y1 = hw_floor_div256(maxY - 128);
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > (long long)W - 1) x1 = (long long)W - 1;
    if (y1 > (long long)H - 1) y1 = (long long)H - 1;
    for (long long py = y0; py <= y1; py++) {
      const long long cy = py * 256 + 128;
      for (long long px = x0; px <= x1; px++) {
        const long long cx = px * 256 + 128;
        const long long e0 = hw_edge(s.X[1], s.Y[1], s.X[2], s.Y[2], cx, cy);
        const long long e1 = hw_edge(s.X[2], s.Y[2], s.X[0], s.Y[0], cx, cy);
        const long long e2 = hw_edge(s.X[0], s.Y[0], s.X[1], s.Y[1], cx, cy);
        if (!((e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) &&
              (e2 > 0 || (e2 == 0 && tl2))))
          continue;
        const float ar = (float)s.area;
        const float l0 = (float)e0 / ar, l1 = (float)e1 / ar, l2 = (float)e2 / ar;
        const float z0 = s.c[0].z / s.c[0].w, z1 = s.c[1].z / s.c[1].w, z2 = s.c[2].z / s.c[2].w;
        float z = l0 * z0 + l1 * z1 + l2 * z2;
        if (!(z >= 0.0f && z <= 1.0f)) continue;
        z = z + 0.0f;
        const size_t p = (size_t)py * W + (size_t)px;
        if (!(z < depth[p])) continue;
        depth[p] = z;
        if (normals) {
          uint32_t idx[3], side[3];
          hw_tri_vertices((uint32_t)t, P, idx, side);
          const float l[3] = {l0, l1, l2};
          ov3 num = o3(0.0f, 0.0f, 0.0f);
          float den = 0.0f;
          for (int q = 0; q < 3; q++) {
            const fh_vec4 tg = tan[idx[s.vi[q]]];
            const ov4 tw = mulmv(u->modelMatrix, o4(tg.x, tg.y, tg.z, 1.0f));
            const float wq = l[q] / s.c[q].w;
            num = add3(num, muls3(xyz(tw), wq));
            den = den + wq;
          }
          const ov3 n = norm3(divs3(num, den));
          normals[p] = or_encode_oct(n.x, n.y, n.z);
        }
      }
    }
  }
}

/* Test hook: the number of fragments of the hair triangles at each pixel of
 * a W x H target (no depth test, no clipping of z): checks that triangles
 * sharing an edge cover each pixel centre once. */
void or_hw_coverage(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u,
                    const fh_vec4* pos, const fh_vec4* tan, uint32_t* counts, uint32_t S, uint32_t P) {
  const uint64_t nTri = 2ull * S * (P - 1u);
  for (uint64_t t = 0; t < nTri; t++) {
    const hw_tri s = hw_setup(u, shadow, pos, tan, P, W, H, (uint32_t)t);
    if (!s.ok) continue;
    const int tl0 = hw_top_left(s.X[1], s.Y[1], s.X[2], s.Y[2]);
    const int tl1 = hw_top_left(s.X[2], s.Y[2], s.X[0], s.Y[0]);
    const int tl2 = hw_top_left(s.X[0], s.Y[0], s.X[1], s.Y[1]);
    for (long long py = 0; py < (long long)H; py++)
      for (long long px = 0; px < (long long)W; px++) {
        const long long cx = px * 256 + 128, cy = py * 256 + 128;
        const long long e0 = hw_edge(s.X[1], s.Y[1], s.X[2], s.Y[2], cx, cy);
        const long long e1 = hw_edge(s.X[2], s.Y[2], s.X[0], s.Y[0], cx, cy);
        const long long e2 = hw_edge(s.X[0], s.Y[0], s.X[1], s.Y[1], cx, cy);
        if ((e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) && (e2 > 0 || (e2 == 0 && tl2)))
          counts[(size_t)py * W + (size_t)px]++;
      }
  }
}

/* ======================================================================== */
/* Background and meshes                                                     */
/* ======================================================================== */

/* drawBackgroundGradientPass.wgsl.ts:33-48 with hash / noise / fractalNoise
 * (noise.wgsl.ts:14-42), per pixel centre, stored as rgba16float bits.
 * Pinned: test_oracle_meshes.py (hash at lattice points against float64,
 * the noise's interpolation at cell corners and centres, the gradient's
 * closed form with the noise scaled away). */
static float bg_hash(ov2 p0) {
  const ov2 q = o2(p0.x * 0.3183099f + 0.71f, p0.y * 0.3183099f + 0.113f);
  const ov2 p = o2(50.0f * ofract(q.x), 50.0f * ofract(q.y));
  return -1.0f + 2.0f * ofract(p.x * p.y * (p.x + p.y));
}

static float bg_noise(ov2 p) {
  const ov2 i = o2(floorf(p.x), floorf(p.y));
  const ov2 f = o2(ofract(p.x), ofract(p.y));
  const ov2 u = o2(f.x * f.x * f.x * (f.x * (f.x * 6.0f - 15.0f) + 10.0f),
                   f.y * f.y * f.y * (f.y * (f.y * 6.0f - 15.0f) + 10.0f));
  const float a = omix(bg_hash(add2(i, o2(0.0f, 0.0f))), bg_hash(add2(i, o2(1.0f, 0.0f))), u.x);
  const float b = omix(bg_hash(add2(i, o2(0.0f, 1.0f))), bg_hash(add2(i, o2(1.0f, 1.0f))), u.x);
  return omix(a, b, u.y);
}

float or_bg_hash(float x, float y) { return bg_hash(o2(x, y)); }
float or_bg_noise(float x, float y) { return bg_noise(o2(x, y)); }

static float fractal_noise(ov2 p, float scale) {
  ov2 uv = muls2(p, scale);
  /* uv = mat2x2f(1.6, 1.2, -1.2, 1.6) * uv */
#define ROT(v) o2(1.6f * (v).x + -1.2f * (v).y, 1.2f * (v).x + 1.6f * (v).y)
  float f = 0.5000f * bg_noise(uv); uv = ROT(uv);
  f = f + 0.2500f * bg_noise(uv); uv = ROT(uv);
  f = f + 0.1250f * bg_noise(uv); uv = ROT(uv);
  f = f + 0.0625f * bg_noise(uv);
#undef ROT
  return f;
}

float or_bg_fractal(float x, float y, float scale) { return fractal_noise(o2(x, y), scale); }

static void store_rgba16f(uint16_t* hdr, size_t p, ov3 c, float a) {
  hdr[4 * p + 0] = (uint16_t)of2h(c.x);
  hdr[4 * p + 1] = (uint16_t)of2h(c.y);
  hdr[4 * p + 2] = (uint16_t)of2h(c.z);
  hdr[4 * p + 3] = (uint16_t)of2h(a);
}

void or_background(uint32_t W, uint32_t H, const fh_render_uniforms* u, uint16_t* hdr) {
  for (uint32_t py = 0; py < H; py++)
    for (uint32_t px = 0; px < W; px++) {
      const ov2 uv = o2(((float)px + 0.5f) / u->viewport[0], ((float)py + 0.5f) / u->viewport[1]);
      const float c = fractal_noise(uv, u->backgroundNoiseScale) * 0.5f + 0.5f;
      ov3 color = mix3(o3(u->backgroundColor0[0], u->backgroundColor0[1], u->backgroundColor0[2]),
                       o3(u->backgroundColor1[0], u->backgroundColor1[1], u->backgroundColor1[2]), c);
      const float g = 1.0f - uv.y;
      color = mix3(color, o3(g, g, g), u->backgroundGradientStrength);
      store_rgba16f(hdr, (size_t)py * W + px, color, 1.0f);
    }
}

/* DrawMeshesPass (shadow = 0; drawMeshesPass.ts:58-115, drawMeshesPass.wgsl.ts
 * :46-168) and the meshes' draw of ShadowMapPass (shadow = 1;
 * shadowMapPass.ts:83-133, shadowMapPass.wgsl.ts), with the rasterizer of
 * or_hw_hair. Triangles run in scene order (mesh, then index order), back
 * faces culled in the camera pass (counter-clockwise in NDC is front), none
 * in the shadow pass, collider previews only in the camera pass. Each
 * fragment that passes the 'less' test writes its depth, normal and colour
 * at once (forward shading in primitive order). Pinned: test_oracle_meshes.py
 * (a screen-aligned quad's exact pixel set and planar depth, back faces
 * culled only in the camera pass, a nearer later mesh occluding an earlier
 * one and ties kept by the earlier, the preview's flat colour, PBR against a
 * float64 restatement of the Karis13 formulas, a step shadow map whose PCF
 * hit count is known, AO strength). */
static void mat_mul4(const float* a, const float* b, float* out) {
  for (int c = 0; c < 4; c++) {
    const ov4 r = mulmv(a, o4(b[4 * c], b[4 * c + 1], b[4 * c + 2], b[4 * c + 3]));
    out[4 * c] = r.x; out[4 * c + 1] = r.y; out[4 * c + 2] = r.z; out[4 * c + 3] = r.w;
  }
}

static ov3 vtx3(const float* v, uint32_t i) { return o3(v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]); }

static float dot_max0(ov3 a, ov3 b) { return omax(0.0f, dot3(a, b)); }

typedef struct {
  ov3 positionWS, normal, toEye, albedo;
  float roughness, isMetallic, ao, shadow;
} pbr_material;

/* disneyPBR — pbr.wgsl.ts:140-160, pbr_CookTorrance :112-131,
 * FresnelSchlick :18-20, DistributionGGX :35-44, GeometrySchlickGGX :54-59,
 * GeometrySmith :71-77, pbr_mixDiffuseAndSpecular :133-138 */
static ov3 disney_pbr(const pbr_material* m, const fh_light* light) {
  const ov3 N = m->normal, V = m->toEye;
  const ov3 L = norm3(sub3(o3(light->position[0], light->position[1], light->position[2]), m->positionWS));
  const ov3 H = norm3(add3(V, L));
  const ov3 F0 = mix3(o3(0.04f, 0.04f, 0.04f), m->albedo, m->isMetallic);
  const float fk = powf(1.0f - dot_max0(H, V), 5.0f);
  const ov3 F = o3(F0.x + (1.0f - F0.x) * fk, F0.y + (1.0f - F0.y) * fk, F0.z + (1.0f - F0.z) * fk);
  const float NdotV = dot_max0(N, V), NdotL = dot_max0(N, L);
  const float r = m->roughness + 1.0f, k = (r * r) / 8.0f;
  const float g1 = NdotL / (NdotL * (1.0f - k) + k), g2 = NdotV / (NdotV * (1.0f - k) + k);
  const float G = g1 * g2;
  const float a = m->roughness * m->roughness, a2 = a * a;
  const float NdotH = dot_max0(N, H), NdotH2 = NdotH * NdotH;
  float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
  denom = O_PI * denom * denom;
  const float NDF = a2 / denom;
  const ov3 numerator = muls3(F, NDF * G);
  const float denominator = 4.0f * dot_max0(N, V) * dot_max0(N, L);
  const ov3 specular = divs3(numerator, omax(denominator, 0.001f));
  const ov3 kD = mix3(sub3(o3(1.0f, 1.0f, 1.0f), F), o3(0.0f, 0.0f, 0.0f), m->isMetallic);
  const ov3 brdf = add3(mul3(kD, m->albedo), specular);
  const ov3 radiance = muls3(o3(light->colorAndEnergy[0], light->colorAndEnergy[1], light->colorAndEnergy[2]),
                             light->colorAndEnergy[3]);
  return muls3(mul3(brdf, radiance), NdotL);
}

/* calculateDirectionalShadow — sampleShadows.wgsl.ts:47-84, _sampleShadowMap
 * :95-131, _isValidShadowSample :136-150 */
static float directional_shadow(const fh_render_uniforms* u, const float* sm, uint32_t smSize, ov3 posWS,
                                ov3 normalWS, ov3 pss) {
  const fh_shadows* sh = &u->shadows;
  const ov3 toSrc = norm3(sub3(o3(sh->sourcePosition[0], sh->sourcePosition[1], sh->sourcePosition[2]), posWS));
  const float bias = omax(fabsf(sh->bias) * (1.0f - dot3(normalWS, toSrc)), 0.0005f);
  int32_t radius = (int32_t)sh->PCF_Radius;
  if (sh->usePCSS > 0u) {
    const float d = gather_avg(sm, smSize, smSize, o2(pss.x, pss.y));
    const float dd = omax(pss.z - d, 0.0f);
    radius = 1 + oi32(dd / d * 10.0f);
  }
  const float texel = 1.0f / (float)smSize;
  float shadow = 0.0f;
  for (int32_t x = -radius; x <= radius; x++)
    for (int32_t y = -radius; y <= radius; y++) {
      const ov2 uv = add2(o2(pss.x, pss.y), o2((float)x * texel, (float)y * texel));
      shadow += (pss.z - bias > gather_avg(sm, smSize, smSize, uv)) ? 1.0f : 0.0f;
    }
  const float pcf = (float)(radius * 2 + 1);
  const float res = shadow / (pcf * pcf);
  if (pss.z > 1.0f) return 0.0f;
  if (pss.x < 0.0f || pss.x > 1.0f || pss.y < 0.0f || pss.y > 1.0f) return 0.0f;
  return res;
}

/* hw_setup's snapping and orientation for given clip positions; *swapped
 * reports that the snapped vertices ran clockwise on screen (y down), i.e.
 * counter-clockwise in NDC (y up). */
static hw_tri hw_setup_clip(const ov4* c, uint32_t W, uint32_t H, int* swapped) {
  hw_tri s;
  s.ok = 1;
  *swapped = 0;
  for (int k = 0; k < 3; k++) {
    s.c[k] = c[k];
    s.vi[k] = (uint32_t)k;
    s.X[k] = s.Y[k] = 0;
    if (!(s.c[k].w > 0.0f)) { s.ok = 0; continue; }
    const float nx = s.c[k].x / s.c[k].w, ny = s.c[k].y / s.c[k].w;
    const float xs = (float)W * (nx * 0.5f + 0.5f);
    const float ys = (float)H * (1.0f - (ny * 0.5f + 0.5f));
    if (!(fabsf(xs) < 1048576.0f && fabsf(ys) < 1048576.0f)) { s.ok = 0; continue; }
    s.X[k] = (long long)rintf(xs * 256.0f);
    s.Y[k] = (long long)rintf(ys * 256.0f);
  }
  s.area = 0;
  if (!s.ok) return s;
  long long area = (s.X[1] - s.X[0]) * (s.Y[2] - s.Y[0]) - (s.Y[1] - s.Y[0]) * (s.X[2] - s.X[0]);
  if (area < 0) {
    ov4 t = s.c[1]; s.c[1] = s.c[2]; s.c[2] = t;
    long long q = s.X[1]; s.X[1] = s.X[2]; s.X[2] = q;
    q = s.Y[1]; s.Y[1] = s.Y[2]; s.Y[2] = q;
    uint32_t v = s.vi[1]; s.vi[1] = s.vi[2]; s.vi[2] = v;
    area = -area;
    *swapped = 1;
  }
  s.area = area;
  s.ok = area > 0;
  return s;
}

/* main_fs (drawMeshesPass.wgsl.ts:88-130) and doShading (:133-146) of one
 * fragment at pixel (px, py) of triangle s (vertex indices vi, in the
 * original order), returning the colour and the normal. */
static ov3 mesh_fragment(const fh_render_uniforms* u, const fh_mesh* m, const hw_tri* s, const uint32_t* vi,
                         long long px, long long py, const float* sm, uint32_t smSize, const float* ao,
                         uint32_t aoW, uint32_t aoH, ov3* normalOut) {
  const long long cx = px * 256 + 128, cy = py * 256 + 128;
  const float ar = (float)s->area;
  const float l[3] = {(float)hw_edge(s->X[1], s->Y[1], s->X[2], s->Y[2], cx, cy) / ar,
                      (float)hw_edge(s->X[2], s->Y[2], s->X[0], s->Y[0], cx, cy) / ar,
                      (float)hw_edge(s->X[0], s->Y[0], s->X[1], s->Y[1], cx, cy) / ar};
  const float* model = m->is_collider_preview ? u->collisionSphereModelMatrix : u->modelMatrix;
  ov3 pos = o3(0.0f, 0.0f, 0.0f), nrm = o3(0.0f, 0.0f, 0.0f), pss = o3(0.0f, 0.0f, 0.0f);
  float den = 0.0f;
  float wq[3];
  for (int q = 0; q < 3; q++) {
    wq[q] = l[q] / s->c[q].w;
    den = den + wq[q];
  }
  for (int q = 0; q < 3; q++) {
    const uint32_t v = vi[s->vi[q]];
    const ov3 pv = vtx3(m->positions, v), nv = vtx3(m->normals, v);
    /* transformNormalToWorldSpace (shaderSnippets.wgls.ts:15-19) */
    const ov3 nws = norm3(xyz(mulmv(model, o4(nv.x, nv.y, nv.z, 0.0f))));
    /* projectToShadowSpace (sampleShadows.wgsl.ts:21-34); the preview's
     * matrix is mat4x4<f32>() (zeros) */
    ov4 lp = o4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!m->is_collider_preview) lp = mulmv(u->shadows.sourceMVP_Matrix, o4(pv.x, pv.y, pv.z, 1.0f));
    const ov3 sp = o3(lp.x * 0.5f + 0.5f, 1.0f - (lp.y * 0.5f + 0.5f), lp.z);
    pos = add3(pos, muls3(pv, wq[q]));
    nrm = add3(nrm, muls3(nws, wq[q]));
    pss = add3(pss, muls3(sp, wq[q]));
  }
  pos = divs3(pos, den);
  nrm = divs3(nrm, den);
  pss = divs3(pss, den);
  pbr_material mt;
  mt.positionWS = pos;
  mt.normal = norm3(nrm);
  mt.toEye = norm3(sub3(o3(u->cameraPosition[0], u->cameraPosition[1], u->cameraPosition[2]), pos));
  mt.albedo = o3(0.9f, 0.9f, 0.9f);
  mt.roughness = 0.8f;
  mt.isMetallic = 0.0f;
  mt.shadow = 1.0f - directional_shadow(u, sm, smSize, pos, mt.normal, pss);
  /* sampleAo (textureAo.wgsl.ts:7-12) at the fragment position (pixel centre) */
  const ov2 at = o2(((float)px + 0.5f) / u->viewport[0], ((float)py + 0.5f) / u->viewport[1]);
  const float aov = tex_load(ao, aoW, aoH, oi32((float)aoW * at.x), oi32((float)aoH * at.y));
  mt.ao = omix(1.0f, aov, u->ao.strength);
  if (m->is_collider_preview) {
    mt.albedo = o3(1.0f, 1.0f, 1.0f);
    mt.shadow = 1.0f;
    mt.ao = 1.0f;
    mt.roughness = 1.0f;
    mt.isMetallic = 0.4f;
  }
  const ov3 ambient = muls3(o3(u->lightAmbient[0], u->lightAmbient[1], u->lightAmbient[2]), u->lightAmbient[3]);
  ov3 rs = o3(0.0f, 0.0f, 0.0f);
  rs = add3(rs, disney_pbr(&mt, &u->light0));
  rs = add3(rs, disney_pbr(&mt, &u->light1));
  rs = add3(rs, disney_pbr(&mt, &u->light2));
  rs = muls3(rs, osat(mt.ao));
  rs = muls3(rs, oclamp(mt.shadow, 1.0f - u->shadows.strength, 1.0f));
  ov3 color = add3(ambient, rs);
  if (m->is_collider_preview) color = mix3(color, o3(1.0f, 1.0f, 0.0f), 0.5f);
  *normalOut = mt.normal;
  return color;
}

void or_meshes(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u, const fh_mesh* meshes,
               uint32_t count, const float* shadowMap, uint32_t shadowSize, const float* ao, uint32_t aoW,
               uint32_t aoH, float* depth, uint32_t* normals, uint16_t* hdr) {
  for (size_t p = 0; p < (size_t)W * H; p++) {
    depth[p] = 1.0f;  /* the passes' depth clear */
    if (!shadow) normals[p] = 0u;
  }
  float previewMvp[16], vm[16];
  mat_mul4(u->viewMatrix, u->collisionSphereModelMatrix, vm);  /* getMVP_Mat :5-8 */
  mat_mul4(u->projMatrix, vm, previewMvp);
  for (uint32_t mi = 0; mi < count; mi++) {
    const fh_mesh* m = &meshes[mi];
    if (shadow && m->is_collider_preview) continue;
    const float* mvp = shadow ? u->shadows.sourceMVP_Matrix : (m->is_collider_preview ? previewMvp : u->mvpMatrix);
    for (uint32_t t = 0; t < m->triangle_count; t++) {
      uint32_t vi[3];
      ov4 c[3];
      for (int q = 0; q < 3; q++) {
        vi[q] = m->indices[3 * (size_t)t + q];
        const ov3 p = vtx3(m->positions, vi[q]);
        c[q] = mulmv(mvp, o4(p.x, p.y, p.z, 1.0f));
      }
      int swapped;
      const hw_tri s = hw_setup_clip(c, W, H, &swapped);
      if (!s.ok) continue;
      if (!shadow && !swapped) continue;  /* back face */
      long long minX = s.X[0], maxX = s.X[0], minY = s.Y[0], maxY = s.Y[0];
      for (int k = 1; k < 3; k++) {
        if (s.X[k] < minX) minX = s.X[k];
        if (s.X[k] > maxX) maxX = s.X[k];
        if (s.Y[k] < minY) minY = s.Y[k];
        if (s.Y[k] > maxY) maxY = s.Y[k];
      }
      const int tl0 = hw_top_left(s.X[1], s.Y[1], s.X[2], s.Y[2]);
      const int tl1 = hw_top_left(s.X[2], s.Y[2], s.X[0], s.Y[0]);
      const int tl2 = hw_top_left(s.X[0], s.Y[0], s.X[1], s.Y[1]);
      long long x0 = hw_ceil_div256(minX - 128), x1 = hw_floor_div256(maxX - 128);
      long long y0 = hw_ceil_div256(minY - 128), y1 = hw_floor_div256(maxY - 128);
      if (x0 < 0) x0 = 0;
      if (y0 < 0) y0 = 0;
      if (x1 > (long long)W - 1) x1 = (long long)W - 1;
      if (y1 > (long long)H - 1) y1 = (long long)H - 1;
      for (long long py = y0; py <= y1; py++) {
        const long long cy = py * 256 + 128;
        for (long long px = x0; px <= x1; px++) {
          const long long cx = px * 256 + 128;
          const long long e0 = hw_edge(s.X[1], s.Y[1], s.X[2], s.Y[2], cx, cy);
          const long long e1 = hw_edge(s.X[2], s.Y[2], s.X[0], s.Y[0], cx, cy);
          const long long e2 = hw_edge(s.X[0], s.Y[0], s.X[1], s.Y[1], cx, cy);
          if (!((e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) &&
                (e2 > 0 || (e2 == 0 && tl2))))
            continue;
          const float ar = (float)s.area;
          const float l0 = (float)e0 / ar, l1 = (float)e1 / ar, l2 = (float)e2 / ar;
          const float z0 = s.c[0].z / s.c[0].w, z1 = s.c[1].z / s.c[1].w, z2 = s.c[2].z / s.c[2].w;
          float z = l0 * z0 + l1 * z1 + l2 * z2;
          if (!(z >= 0.0f && z <= 1.0f)) continue;
          z = z + 0.0f;
          const size_t p = (size_t)py * W + (size_t)px;
          if (!(z < depth[p])) continue;  /* 'less' in primitive order */
          depth[p] = z;
          if (shadow) continue;
          ov3 n;
          const ov3 col = mesh_fragment(u, m, &s, vi, px, py, shadowMap, shadowSize, ao, aoW, aoH, &n);
          normals[p] = or_encode_oct(n.x, n.y, n.z);
          store_rgba16f(hdr, p, col, 1.0f);
        }
      }
    }
  }
}

/* Test hooks: disneyPBR of one light and calculateDirectionalShadow. */
void or_disney_pbr(const float* positionWS, const float* normal, const float* toEye, const float* albedo,
                   float roughness, float isMetallic, const fh_light* light, float* out) {
  pbr_material m;
  m.positionWS = o3(positionWS[0], positionWS[1], positionWS[2]);
  m.normal = o3(normal[0], normal[1], normal[2]);
  m.toEye = o3(toEye[0], toEye[1], toEye[2]);
  m.albedo = o3(albedo[0], albedo[1], albedo[2]);
  m.roughness = roughness;
  m.isMetallic = isMetallic;
  m.ao = 1.0f;
  m.shadow = 1.0f;
  const ov3 r = disney_pbr(&m, light);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

float or_directional_shadow(const fh_render_uniforms* u, const float* sm, uint32_t smSize, const float* posWS,
                            const float* normalWS, const float* pss) {
  return directional_shadow(u, sm, smSize, o3(posWS[0], posWS[1], posWS[2]), o3(normalWS[0], normalWS[1], normalWS[2]),
                            o3(pss[0], pss[1], pss[2]));
}

/* AoPass (aoPass.wgsl.ts:48-79) with gtao (gtao.wgsl.ts:26-136) per texel of
 * the AO texture (aoW x aoH; AO_TEXTURE_SCALE 0.5), stored as r16float
 * (rounded to half, widened). Pinned: test_oracle_hwraster.py (the
 * background returns 0; a flat wall facing the camera is unoccluded and
 * gives the closed form of the integral; a pit is darker than the wall). */
void or_ao(const fh_render_uniforms* u, const float* depth, uint32_t W, uint32_t H,
           const uint32_t* normals, float* out, uint32_t aoW, uint32_t aoH) {
  const float PI_F = 3.14159265358979323846f, HALF_PI_F = 1.57079632679489661923f;
  const float zNear = 0.01f, zFar = 100.0f;
  const float vpx = u->viewport[0], vpy = u->viewport[1];
  const fh_ao* prm = &u->ao;
#define AO_DEPTH(px_, py_) \
  ((ou32(px_) < W && ou32(py_) < H) ? depth[(size_t)ou32(py_) * W + ou32(px_)] : 0.0f)
  for (uint32_t y = 0; y < aoH; y++) {
    for (uint32_t x = 0; x < aoW; x++) {
      const float fxp = ((float)x + 0.5f) / 0.5f, fyp = ((float)y + 0.5f) / 0.5f;
      const uint32_t ix = ou32(fxp), iy = ou32(fyp);
      const uint32_t nn = (ix < W && iy < H) ? normals[(size_t)iy * W + ix] : 0u;
      const ov3 normalWS = hw_decode_oct(nn);
      /* reprojectPositionToVS */
      float d = AO_DEPTH(fxp, fyp);
      ov3 posVS = oproject(u->projMatrixInv, o4(fxp / vpx * 2.0f - 1.0f, fyp / vpy * 2.0f - 1.0f, d, 1.0f));
      posVS.z = zNear * zFar / (zFar + (zNear - zFar) * d);
      const ov4 nv = mulmv(u->viewMatrix, o4(normalWS.x, normalWS.y, normalWS.z, 1.0f));
      const ov3 normalVS = o3(nv.x, nv.y, -nv.z);
      const ov3 toCamera = norm3(o3(-posVS.x, -posVS.y, -posVS.z));
      float ao = 0.0f;
      if (posVS.z >= 0.999f * zFar) {
        out[(size_t)y * aoW + x] = 0.0f;
        continue;
      }
      const float radius = omax((float)prm->numSteps, prm->radius * (posVS.z / zFar));
      const float stepSize = radius / (float)prm->numSteps;
      const float rotJitter = PI_F * ofract(52.9829189f * ofract(fxp * 0.06711056f + fyp * 0.00583715f));
      const float jitterMod = (fxp + fyp) * 0.25f;
      const float radiusJitter = ofract(jitterMod) * stepSize * 0.25f;
      for (uint32_t i = 0; i < prm->numDirections; i++) {
        float phi = (float)i * (PI_F / (float)prm->numDirections) + prm->directionOffset * PI_F;
        phi += rotJitter;
        float currStep = 1.0f + 0.25f * stepSize;
        currStep += radiusJitter;
        const ov3 dir = o3(cosf(phi), sinf(phi), 0.0f);
        float hx = -1.0f, hy = -1.0f;
        for (uint32_t j = 0; j < prm->numSteps; j++) {
          const float vx = dir.x * currStep, vy = dir.y * currStep;
          const float ofx = vx < 0.5f ? floorf(vx) : ceilf(vx), ofy = vy < 0.5f ? floorf(vy) : ceilf(vy);
          for (int sgn = 0; sgn < 2; sgn++) {
            const float sx = sgn ? fxp - ofx : fxp + ofx, sy = sgn ? fyp - ofy : fyp + ofy;
            const float sd = AO_DEPTH(sx, sy);
            ov3 sv = oproject(u->projMatrixInv, o4(sx / vpx * 2.0f - 1.0f, sy / vpy * 2.0f - 1.0f, sd, 1.0f));
            sv.z = zNear * zFar / (zFar + (zNear - zFar) * sd);
            const ov3 ws = sub3(sv, posVS);
            const float dist2 = len3(ws);
            const float cosh_ = dot3(ws, toCamera) * (1.0f / sqrtf(dist2));
            const float fall = 2.0f * osat((dist2 - prm->falloffStart2) / (prm->falloffEnd2 - prm->falloffStart2));
            if (sgn) hy = omax(hy, cosh_ - fall);
            else hx = omax(hx, cosh_ - fall);
          }
          currStep += stepSize;
        }
        hx = acosf(hx);
        hy = acosf(hy);
        const ov3 bit = norm3(cross3(dir, toCamera));
        const ov3 tng = cross3(toCamera, bit);
        const ov3 nx = sub3(normalVS, muls3(bit, dot3(normalVS, bit)));
        const float nnx = len3(nx);
        const float nnxInv = 1.0f / (nnx + 1e-6f);
        const float cosxi = dot3(nx, tng) * nnxInv;
        const float gamma = acosf(cosxi) - HALF_PI_F;
        const float cosgamma = dot3(nx, toCamera) * nnxInv;
        const float singamma2 = -2.0f * cosxi;
        hx = gamma + omax(-hx - gamma, -HALF_PI_F);
        hy = gamma + omin(hy - gamma, HALF_PI_F);
        ao += nnx * 0.25f * ((hx * singamma2 + cosgamma - cosf(2.0f * hx - gamma)) +
                             (hy * singamma2 + cosgamma - cosf(2.0f * hy - gamma)));
      }
      out[(size_t)y * aoW + x] = oh2f(of2h(ao / (float)prm->numDirections));
    }
  }
#undef AO_DEPTH
}
