// Simulation step of the hair chain on gfx950: grid pre-pass (K1), strand
// integration with constraints and collisions (K2) and the particle-to-grid
// scatter (K3). Reference: src/passes/simulation/*.
#include "fh_device.hpp"
#include "fh_launch.hpp"

#include <mutex>
#include <string>

namespace fh {
int fill_u32(fh_ctx* ctx, void* dst, uint64_t n_words, uint32_t value);
}

namespace {

using namespace fh;

constexpr float kGridMul = 1000000.0f;  // GRID_FLOAT_TO_U32_MUL, grids.wgsl.ts:3
constexpr float kEps = 1e-7f;           // FLOAT_EPSILON, shaderSnippets.wgls.ts:67
// SDF_SAMPLE_GRAD_OFFSET = SDF_CELL_SIZE * 0.1 (collisions.wgsl.ts:23-24),
// evaluated in double as a WGSL abstract-float constant, then rounded to f32.
constexpr float kSdfGradOffset = (float)(0.0033918858971446753 * 0.1);

__device__ __forceinline__ uint32_t grid_idx(uint32_t d, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t m = d - 1u;
  return min(z, m) * d * d + min(y, m) * d + min(x, m);
}

struct Cell {
  uint32_t mn[3], mx[3];
  f3 pin;
};

// _getGridCell — grids.wgsl.ts:61-74
__device__ __forceinline__ Cell grid_cell(uint32_t d, f3 bmin, f3 bmax, f3 p) {
  Cell c;
  f3 t = sat((p - bmin) / (bmax - bmin));
  const float dm1 = (float)(d - 1u);
  c.pin = t * dm1;
  c.mn[0] = f2u(floorf(t.x * dm1)); c.mn[1] = f2u(floorf(t.y * dm1)); c.mn[2] = f2u(floorf(t.z * dm1));
  c.mx[0] = f2u(ceilf(t.x * dm1));  c.mx[1] = f2u(ceilf(t.y * dm1));  c.mx[2] = f2u(ceilf(t.z * dm1));
  return c;
}

// _getGridCellWeights — grids.wgsl.ts:76-84
__device__ __forceinline__ f3 cell_w(uint32_t x, uint32_t y, uint32_t z, f3 pin) {
  return {clampf(1.0f - fabsf(pin.x - (float)x), 0.0f, 1.0f),
          clampf(1.0f - fabsf(pin.y - (float)y), 0.0f, 1.0f),
          clampf(1.0f - fabsf(pin.z - (float)z), 0.0f, 1.0f)};
}

__device__ __forceinline__ float grid_decode(int32_t v) { return (float)v / kGridMul; }

struct SdfView {
  const float* __restrict__ vol;
  uint32_t dims;
  f3 bmin, bmax;
};

// sampleSDFCollider — sdfCollider.ts:36-44
__device__ __forceinline__ float sample_sdf(const SdfView& s, f3 p) {
  f3 t = sat((p - s.bmin) / (s.bmax - s.bmin));
  return sample_volume(s.vol, s.dims, t);
}

// The constraint sweeps' and collisions' divisions and square roots on the
// hardware reciprocal, square root and reciprocal square root (about 1 ulp
// each; WGSL allows 2.5 ulp for division and inherits the platform's square
// root). The integration's positions are compared with the oracle within a
// relative tolerance, not bit for bit, and the sweeps are chains of dependent
// steps whose length these set (1080p: 151 -> ~105 us). The grid forces keep
// correctly rounded arithmetic: the grid velocity is vel / density, which
// does not vanish as the interpolation weights do, so a weight that rounds to
// 1e-7 instead of 0 next to an empty node moves a point by a whole grid
// displacement. The grid passes keep it too: their grids are compared bit for
// bit.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float flength(f3 a) { return __builtin_amdgcn_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 fnormalize(f3 a) { return a * frsq(dot(a, a)); }

__device__ __forceinline__ float sample_sdf_fast(const SdfView& s, f3 p) {
  const f3 e = s.bmax - s.bmin;
  f3 t = sat((p - s.bmin) * mk3(frcp(e.x), frcp(e.y), frcp(e.z)));
  return sample_volume(s.vol, s.dims, t);
}

// ---------------------------------------------------------------------------
// K1 — gridPreSimPass.wgsl.ts:38-159. One thread per grid point.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void grid_pre_sim_body(
    uint32_t d, const fh_sim_uniforms* __restrict__ u,
    const fh_density_velocity* __restrict__ dv, fh_density_grad_wind* __restrict__ out,
    const float* __restrict__ sdf, uint32_t sdf_dims, uint32_t block, uint32_t blocks) {
  const uint32_t id = block * blockDim.x + threadIdx.x;
  if (id >= d * d * d) return;
  const f3 gmin = mk3(u->gridBoundsMin[0], u->gridBoundsMin[1], u->gridBoundsMin[2]);
  const f3 gmax = mk3(u->gridBoundsMax[0], u->gridBoundsMax[1], u->gridBoundsMax[2]);
  SdfView sv{sdf, sdf_dims, mk3(u->sdfBoundsMin[0], u->sdfBoundsMin[1], u->sdfBoundsMin[2]),
             mk3(u->sdfBoundsMax[0], u->sdfBoundsMax[1], u->sdfBoundsMax[2])};
  const uint32_t z = id / (d * d);
  const uint32_t xy = id - z * d * d;
  const uint32_t y = xy / d, x = xy % d;
  // getGridPointPositionWS — grids.wgsl.ts:23-36
  const f3 cell = (gmax - gmin) / (float)(d - 1u);
  const f3 pos = gmin + cell * mk3((float)x, (float)y, (float)z);
  // getWindStrength :89-108
  float wind;
  const float sd = sample_sdf(sv, pos);
  if (sd < 0.0f) {
    wind = 0.0f;
  } else {
    const f3 toward = normalize(mk3(u->wind[0], u->wind[1], u->wind[2])) * -1.0f;
    const f3 rp = pos + (toward * sd) * u->windColisionTraceOffset;
    wind = sample_sdf(sv, rp) < 0.0f ? 0.5f : 1.0f;
  }
  // getDensityGradient :111-159
  const uint32_t idx0 = grid_idx(d, x, y, z);
  const float dens0 = grid_decode(dv[idx0].density);
  f3 res = mk3(0.0f, 0.0f, 0.0f);
  const int offs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int ox = (int)x + offs[k][0], oy = (int)y + offs[k][1], oz = (int)z + offs[k][2];
    f3 g = mk3(0.0f, 0.0f, 0.0f);
    if (!(ox < 0 || oy < 0 || oz < 0)) {
      const uint32_t io = grid_idx(d, (uint32_t)ox, (uint32_t)oy, (uint32_t)oz);
      if (io != idx0) {
        const float m = fmx(dens0 - grid_decode(dv[io].density), 0.0f);
        g = mk3(m * (float)offs[k][0], m * (float)offs[k][1], m * (float)offs[k][2]);
      }
    }
    res = res + g;
  }
  *reinterpret_cast<float4*>(&out[idx0]) = make_float4(res.x, res.y, res.z, wind);
}

struct GridPreSim {
  struct Args {
    uint32_t d;
    const fh_sim_uniforms* u;
    const fh_density_velocity* dv;
    fh_density_grad_wind* out;
    const float* sdf;
    uint32_t sdf_dims;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "grid_pre_sim";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { grid_pre_sim_body(a.d, a.u, a.dv, a.out, a.sdf, a.sdf_dims, block, blocks); }
};


// ---------------------------------------------------------------------------
// K2 — hairSimIntegrationPass.wgsl.ts:79-273. A workgroup of kSimThreads
// threads owns L strands (32, or 16 when there are too few strands to give
// every CU a couple of workgroups), staged in LDS lane-interleaved (point j
// of strand l at j * L + l). The two Gauss-Seidel sweeps along the
// strand (length constraint; local shape with the tip-tangent tail) run one
// strand per lane on the first wave, so each wave instruction advances 32
// strands; the point-local phases (Verlet integration with both grid
// samples, global shape, sphere and SDF collisions) are memory-latency bound
// and spread over all threads, one (point, strand) per thread at a time, in
// LDS order so that neighbouring threads touch neighbouring LDS words.
// ---------------------------------------------------------------------------
constexpr uint32_t kSimThreads = 256;

template <uint32_t L>
__device__ __forceinline__ void sim_integrate_body(
    uint32_t d, const fh_sim_uniforms* __restrict__ u, float4* __restrict__ prev,
    const float4* __restrict__ now, const float* __restrict__ seglen,
    const float* __restrict__ sdf, uint32_t sdf_dims,
    const fh_density_velocity* __restrict__ dv, const fh_density_grad_wind* __restrict__ dgw,
    const float4* __restrict__ initial, uint32_t S, uint32_t P, uint32_t block, uint32_t blocks) {
  extern __shared__ float4 lds[];
  float4* A = lds;                       // positions being solved (xyz, movable)
  float4* Bi = lds + (size_t)P * L;      // initial positions
  float4* K = Bi + (size_t)P * L;        // local-shape constants (h0, w0, w1, w2)
  float* Len = reinterpret_cast<float*>(K + (size_t)P * L);  // rest lengths
  const uint32_t tid = threadIdx.x;
  const uint32_t s0 = block * L;
  const uint32_t nStr = min(L, S - s0);
  const uint32_t nElem = nStr * P;
  const size_t gbase = (size_t)s0 * P;

  const f3 gmin = mk3(u->gridBoundsMin[0], u->gridBoundsMin[1], u->gridBoundsMin[2]);
  const f3 gmax = mk3(u->gridBoundsMax[0], u->gridBoundsMax[1], u->gridBoundsMax[2]);
  const f3 gravity = mk3(0.0f, -u->gravity, 0.0f);
  const f3 windDir = mk3(u->wind[0], u->wind[1], u->wind[2]);
  const float dt = u->deltaTime;

  // Staging in global order (coalesced loads); PREV goes to K until the
  // local-shape constants replace it.
  for (uint32_t e = tid; e < nElem; e += kSimThreads) {
    const uint32_t ls = e / P, pt = e - ls * P;
    const uint32_t o = pt * L + ls;
    A[o] = now[gbase + e];
    K[o] = prev[gbase + e];
    Bi[o] = initial[gbase + e];
    Len[o] = seglen[gbase + e];
  }
  __syncthreads();
  // The point-local phases walk LDS in its own order, o = point * L + strand,
  // so neighbouring threads touch neighbouring words.
  const uint32_t nO = P * L;
  // Verlet integration with grid forces, points 1..P-1 (:125-162)
  for (uint32_t o = tid + L; o < nO; o += kSimThreads) {
    const uint32_t ls = o % L;
    if (ls >= nStr) continue;
    const float4 pn4 = A[o], pp4 = K[o];
    // wind jitter is per strand (:134-139)
    const float timer = (float)u->frameIdx * 0.73f + u->windPhaseOffset * (float)(s0 + ls);
    const float jitter = fractf(timer * u->windStrengthFrequency);
    const float jitterDelta = mixf(-0.5f * u->windStrengthJitter, 0.5f * u->windStrengthJitter, jitter);
    const float jitterStr = 1.0f + jitterDelta;
    const f3 pn = mk3(pn4.x, pn4.y, pn4.z), pp = mk3(pp4.x, pp4.y, pp4.z);
    const Cell co = grid_cell(d, gmin, gmax, pn);
    f3 grad = mk3(0.0f, 0.0f, 0.0f), vel = mk3(0.0f, 0.0f, 0.0f);
    float wstr = 0.0f, dens = 0.0f;
    // _getGridDensityGradAndWind (densityGradAndWindGrid.wgsl.ts:20-44)
    for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
      for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
        for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
          const f3 w = cell_w(x, y, z, co.pin);
          const float4 gw = *reinterpret_cast<const float4*>(&dgw[grid_idx(d, x, y, z)]);
          grad = grad + w * mk3(gw.x, gw.y, gw.z);
          wstr = wstr + length(w) * gw.w;
        }
    // _getGridDensityVelocity (densityVelocityGrid.wgsl.ts:30-68)
    for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
      for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
        for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
          const f3 w = cell_w(x, y, z, co.pin);
          const int4 gv = *reinterpret_cast<const int4*>(&dv[grid_idx(d, x, y, z)]);
          vel = vel + w * mk3(grid_decode(gv.x), grid_decode(gv.y), grid_decode(gv.z));
          dens = dens + length(w) * grid_decode(gv.w);
        }
    f3 force = gravity;
    const float windCell = satf(mixf(u->windStrengthLull, 1.0f, wstr));
    force = force + windDir * fabsf(u->wind[3] * jitterStr * windCell);
    force = force + grad * u->volumePreservation;
    const f3 gridDisp = dens > 0.0f ? vel / dens : mk3(0.0f, 0.0f, 0.0f);
    // verletIntegration — integration.wgsl.ts:8-28
    const f3 fdisp = mix(pn - pp, gridDisp, u->friction);
    const f3 next = (pn + fdisp) + (force * dt) * dt;
    const f3 rr = mix(pp, next, pn4.w);
    A[o] = make_float4(rr.x, rr.y, rr.z, pn4.w);
  }
  __syncthreads();
  // Local-shape constants depend only on the initial pose (constraints.wgsl.ts:65-82).
  for (uint32_t o = tid; o + 3u * L < nO; o += kSimThreads) {
    if (o % L >= nStr) continue;
    const float4 i0 = Bi[o], i1 = Bi[o + L], i2 = Bi[o + 2 * L];
    const f3 ci = ((mk3(i0.x, i0.y, i0.z) + mk3(i1.x, i1.y, i1.z)) + mk3(i2.x, i2.y, i2.z)) / 3.0f;
    const float h0 = length(ci - mk3(i1.x, i1.y, i1.z));
    const float wT = i0.w + 2.0f * i1.w + i2.w;
    K[o] = make_float4(h0, i0.w / wT * 2.0f, i1.w / wT * -4.0f, i2.w / wT * 2.0f);
  }

  const float iters = (float)u->constraintIterations;
  const float kLen = u->stiffnessLengthConstr;
  const float kGlobal = u->stiffnessGlobalConstr / iters;
  const float kLocal = u->stiffnessLocalConstr / iters;
  const float kColl = u->stiffnessCollisions / iters;
  const float kSdf = u->stiffnessSDF / iters;
  const f4 sphere = mk4(u->collisionSphere[0], u->collisionSphere[1], u->collisionSphere[2], u->collisionSphere[3]);
  const float sdfOffset = u->sdfBoundsMax[3];
  SdfView sv{sdf, sdf_dims, mk3(u->sdfBoundsMin[0], u->sdfBoundsMin[1], u->sdfBoundsMin[2]),
             mk3(u->sdfBoundsMax[0], u->sdfBoundsMax[1], u->sdfBoundsMax[2])};
  const float invPm1 = (float)(P - 1u);
  const bool chain = tid < nStr;  // the sweeps: one strand per lane of wave 0
  const uint32_t lane = tid;

  for (uint32_t it = 0; it < u->constraintIterations; it++) {
    __syncthreads();
    // length constraint — constraints.wgsl.ts:5-28, hairSimIntegrationPass.wgsl.ts:176-190
    if (chain) {
      float4 p0 = A[lane];
      for (uint32_t j = 0; j + 1 < P; j++) {
        float4 p1 = A[(j + 1) * L + lane];
        const float expected = Len[j * L + lane];
        const f3 tg = mk3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z);
        const float corr = 1.0f - expected * frsq(dot(tg, tg));
        const float df = corr * kLen * frcp(p0.w + p1.w + kEps);
        const f3 delta = tg * df;
        p0 = make_float4(p0.x + delta.x * p0.w, p0.y + delta.y * p0.w, p0.z + delta.z * p0.w, p0.w);
        p1 = make_float4(p1.x - delta.x * p1.w, p1.y - delta.y * p1.w, p1.z - delta.z * p1.w, p1.w);
        A[j * L + lane] = p0;
        p0 = p1;
      }
      A[(P - 1) * L + lane] = p0;
    }
    __syncthreads();
    // global shape — constraints.wgsl.ts:31-49
    for (uint32_t o = tid + L; o < nO; o += kSimThreads) {
      if (o % L >= nStr) continue;
      const uint32_t pt = o / L;
      const float x = (float)pt / invPm1;
      const float att = 1.0f - satf((x - u->globalConstrExtent) / u->globalConstrFade);
      const float k = kGlobal * att;
      const float4 p = A[o];
      const float4 ini = Bi[o];
      const f3 dvv = mk3(ini.x - p.x, ini.y - p.y, ini.z - p.z);
      A[o] = make_float4(p.x + (dvv.x * p.w) * k, p.y + (dvv.y * p.w) * k, p.z + (dvv.z * p.w) * k, p.w);
    }
    __syncthreads();
    if (chain) {
      // local shape (Kelager10 triangle bending) — constraints.wgsl.ts:54-87
      float4 a = A[lane], b = A[L + lane];
      uint32_t jj = 0;
      for (; jj + 3 < P; jj++) {
        const float4 c = A[(jj + 2) * L + lane];
        const float4 Kc = K[jj * L + lane];
        const f3 cc = ((mk3(a.x, a.y, a.z) + mk3(b.x, b.y, b.z)) + mk3(c.x, c.y, c.z)) * (1.0f / 3.0f);
        const f3 hv = mk3(b.x, b.y, b.z) - cc;
        const f3 delta = hv * (1.0f - Kc.x * frsq(dot(hv, hv)));
        const float k0 = kLocal * Kc.y, k1 = kLocal * Kc.z, k2 = kLocal * Kc.w;
        A[jj * L + lane] = make_float4(a.x + delta.x * k0, a.y + delta.y * k0, a.z + delta.z * k0, a.w);
        a = make_float4(b.x + delta.x * k1, b.y + delta.y * k1, b.z + delta.z * k1, b.w);
        b = make_float4(c.x + delta.x * k2, c.y + delta.y * k2, c.z + delta.z * k2, c.w);
      }
      A[jj * L + lane] = a;
      A[(jj + 1) * L + lane] = b;
      // tip tangent match — constraints.wgsl.ts:90-120
      float4 pa = A[(jj - 1) * L + lane];
      for (; jj < P; jj++) {
        const uint32_t o = jj * L + lane;
        const float4 pb = A[o];
        const float4 i0 = Bi[o - L], i1 = Bi[o];
        const f3 tgi = mk3(i1.x - i0.x, i1.y - i0.y, i1.z - i0.z);
        const f3 proj = mk3(pa.x, pa.y, pa.z) + tgi;
        const f3 delta = proj - mk3(pb.x, pb.y, pb.z);
        const float k = kLocal * pb.w;
        pa = make_float4(pb.x + delta.x * k, pb.y + delta.y * k, pb.z + delta.z * k, pb.w);
        A[o] = pa;
      }
    }
    __syncthreads();
    // sphere + SDF collisions — collisions.wgsl.ts:4-61, :237-261
    for (uint32_t o = tid + L; o < nO; o += kSimThreads) {
      if (o % L >= nStr) continue;
      const uint32_t pt = o / L;
      const float4 p4 = A[o];
      f3 p = mk3(p4.x, p4.y, p4.z);
      const f3 v = p - xyz(sphere);
      const float toShell = fmx(0.0f, -(flength(v) - sphere.w));
      p = p + (fnormalize(v) * toShell) * kColl;
      if (pt > 2u) {
        const float sd = sample_sdf_fast(sv, p);
        const float sdo = sd + sdfOffset;
        if (!(sdo >= 0.0f)) {
          f3 cd;
          cd.x = sample_sdf_fast(sv, p + mk3(kSdfGradOffset, 0.0f, 0.0f));
          cd.y = sample_sdf_fast(sv, p + mk3(0.0f, kSdfGradOffset, 0.0f));
          cd.z = sample_sdf_fast(sv, p + mk3(0.0f, 0.0f, kSdfGradOffset));
          cd = cd - mk3(sd, sd, sd);
          p = p + (fnormalize(cd) * -sdo) * kSdf;
        }
      }
      A[o] = make_float4(p.x, p.y, p.z, p4.w);
    }
  }
  __syncthreads();
  // coalesced write-back into the PREV buffer (:268-272)
  for (uint32_t e = tid; e < nElem; e += kSimThreads) {
    const uint32_t ls = e / P, pt = e - ls * P;
    prev[gbase + e] = A[pt * L + ls];
  }
}

struct SimArgs {
  uint32_t d;
  const fh_sim_uniforms* u;
  float4* prev;
  const float4* now;
  const float* seglen;
  const float* sdf;
  uint32_t sdf_dims;
  const fh_density_velocity* dv;
  const fh_density_grad_wind* dgw;
  const float4* initial;
  uint32_t S;
  uint32_t P;
};

template <uint32_t L>
struct SimIntegrate {
  using Args = SimArgs;
  static constexpr uint32_t kThreads = kSimThreads;
  static constexpr const char* kName = "sim_integrate";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    sim_integrate_body<L>(a.d, a.u, a.prev, a.now, a.seglen, a.sdf, a.sdf_dims, a.dv, a.dgw, a.initial, a.S, a.P,
                          xcd_order(block, blocks), blocks);
  }
};


// ---------------------------------------------------------------------------
// K3 — gridPostSimPass.wgsl.ts:45-72 with _addGridDensityVelocity
// (densityVelocityGrid.wgsl.ts:93-124). Every point adds fixed-point
// velocity and density to the 8 grid nodes around it. A head of hair piles
// thousands of contributions onto the same nodes, so one global atomic per
// contribution serializes. Here the grid is cut into blocks of 8^3 nodes:
//   1. every point counts itself into the blocks its corners fall in (one
//      atomic per wave per block: a wave is two runs along strands);
//   2. one workgroup scans the counts into per-block ranges and cuts each
//      range into chunks of `chunk` (point, block) pairs;
//   3. the points are scattered into their blocks' ranges;
//   4. one workgroup per chunk adds the chunk's contributions to an LDS copy
//      of its block, then adds each non-zero node component to the grid.
// i32 addition is associative and commutative (mod 2^32), so the grid is
// bit-identical to one atomic per contribution in any order.
// ---------------------------------------------------------------------------
constexpr uint32_t kBlk = 8;
constexpr uint32_t kBlkNodes = kBlk * kBlk * kBlk;

struct P2GArgs {
  uint32_t d, nb;  // grid dims, blocks per axis
  const fh_sim_uniforms* u;
  const float4* prev;
  const float4* now;
  fh_density_velocity* dv;
  uint32_t nPoints;
  uint32_t chunk;        // pairs per accumulation workgroup
  uint32_t* counts;      // nb^3
  uint32_t* cursors;     // nb^3
  uint32_t* offsets;     // nb^3: first pair of each block
  uint32_t* chunkStart;  // nb^3 + 1: first chunk of each block, total last
  uint32_t* pairs;       // point index per (point, block) pair
};

// Blocks touched by the corner nodes of point e: [b0, b1] per axis.
__device__ __forceinline__ void point_blocks(const P2GArgs& a, f3 gmin, f3 gmax, uint32_t e, uint32_t b0[3],
                                             uint32_t b1[3]) {
  const float4 pn = a.now[e];
  const Cell co = grid_cell(a.d, gmin, gmax, mk3(pn.x, pn.y, pn.z));
#pragma unroll
  for (int i = 0; i < 3; i++) { b0[i] = co.mn[i] / kBlk; b1[i] = co.mx[i] / kBlk; }
}

__device__ __forceinline__ uint32_t blk_idx(const P2GArgs& a, uint32_t x, uint32_t y, uint32_t z) {
  return (z * a.nb + y) * a.nb + x;
}

// Adds `add` to counter[key] for every active lane, one atomic per distinct
// key in the wave; returns the lane's position among the lanes sharing its
// key, on top of the counter's old value.
__device__ __forceinline__ uint32_t wave_keyed_add(uint32_t* counter, uint32_t key, bool active) {
  const uint32_t lane = __lane_id();
  uint32_t result = 0;
  bool pend = active;
  while (true) {
    const uint64_t pm = __ballot(pend);
    if (pm == 0) break;
    const uint32_t leader = (uint32_t)__builtin_ctzll(pm);
    const uint32_t lk = (uint32_t)__shfl((int)key, (int)leader);
    const uint64_t m = __ballot(pend && key == lk);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&counter[lk], (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader);
    if (pend && key == lk) {
      result = base + mask_prefix(m);
      pend = false;
    }
  }
  return result;
}

// Steps 1 and 3. A workgroup takes `span` consecutive points. Each point
// visits the up to 8 blocks its corner nodes fall in, one block offset
// (dx, dy, dz) at a time so that a wave counts equal blocks with one atomic.
// With a block count that fits LDS the counts go to a workgroup-private
// histogram first, so hot blocks see one global atomic per workgroup:
//   count:   histogram, then counts[b] += hist[b];
//   scatter: histogram, then hist[b] = offsets[b] + (cursors[b] += hist[b]),
//            then a second walk takes each pair's slot from hist[b].
// Without LDS (a grid too large) the same walk goes to the global counters.
constexpr uint32_t kP2GBinThreads = 256;
constexpr uint32_t kP2GHistMax = 8192;  // blocks held in LDS: grids up to 160^3

template <bool SCATTER>
__device__ __forceinline__ void p2g_walk(const P2GArgs& a, f3 gmin, f3 gmax, uint32_t first, uint32_t last,
                                         uint32_t* ctr, bool place) {
  for (uint32_t base = first; base < last; base += kP2GBinThreads) {
    const uint32_t e = base + threadIdx.x;
    const bool valid = e < last;
    uint32_t b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
    if (valid) point_blocks(a, gmin, gmax, e, b0, b1);
#pragma unroll 1
    for (uint32_t o = 0; o < 8u; o++) {
      const uint32_t x = b0[0] + (o & 1u), y = b0[1] + ((o >> 1) & 1u), z = b0[2] + (o >> 2);
      const bool on = valid && x <= b1[0] && y <= b1[1] && z <= b1[2];
      const uint32_t blk = on ? blk_idx(a, x, y, z) : 0u;
      const uint32_t r = wave_keyed_add(ctr, blk, on);
      if (SCATTER && place && on) a.pairs[r] = e;
    }
  }
}

template <bool SCATTER, bool LDS>
__device__ __forceinline__ void p2g_bin_body(P2GArgs a, uint32_t nBlocks, uint32_t span, uint32_t block) {
  __shared__ uint32_t hist[LDS ? kP2GHistMax : 1];
  const f3 gmin = mk3(a.u->gridBoundsMin[0], a.u->gridBoundsMin[1], a.u->gridBoundsMin[2]);
  const f3 gmax = mk3(a.u->gridBoundsMax[0], a.u->gridBoundsMax[1], a.u->gridBoundsMax[2]);
  const uint32_t first = block * span;
  const uint32_t last = min(first + span, a.nPoints);
  if (!LDS) {
    // global counters; the scatter turns each cursor into a slot through
    // offsets, so the cursors start at each block's first pair
    p2g_walk<SCATTER>(a, gmin, gmax, first, last, SCATTER ? a.cursors : a.counts, SCATTER);
    return;
  }
  for (uint32_t i = threadIdx.x; i < nBlocks; i += kP2GBinThreads) hist[i] = 0;
  __syncthreads();
  p2g_walk<false>(a, gmin, gmax, first, last, hist, false);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nBlocks; i += kP2GBinThreads) {
    const uint32_t c = hist[i];
    if (!c) continue;
    if (SCATTER) hist[i] = atomicAdd(&a.cursors[i], c);
    else atomicAdd(&a.counts[i], c);
  }
  if (!SCATTER) return;
  __syncthreads();
  p2g_walk<true>(a, gmin, gmax, first, last, hist, true);
}

template <bool SCATTER, bool LDS>
struct P2GBin {
  struct Args {
    P2GArgs a;
    uint32_t nBlocks, span;
  };
  static constexpr uint32_t kThreads = kP2GBinThreads;
  static constexpr const char* kName = SCATTER ? "p2g_scatter" : "p2g_count";
  static __device__ void run(const Args& x, uint32_t block, uint32_t) {
    p2g_bin_body<SCATTER, LDS>(x.a, x.nBlocks, x.span, block);
  }
};

// Step 2, one workgroup: exclusive scans of the counts (-> offsets) and of
// the chunks per block (-> chunkStart, with the total at the end).
__device__ __forceinline__ void p2g_scan_body(P2GArgs a, uint32_t nBlocks, uint32_t block, uint32_t blocks) {
  __shared__ uint32_t warpSums[2][16];
  uint32_t carryP = 0, carryC = 0;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  for (uint32_t base = 0; base < nBlocks; base += 1024u) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t c = i < nBlocks ? a.counts[i] : 0u;
    const uint32_t k = (c + a.chunk - 1u) / a.chunk;
    const uint32_t ip = wave_incl_scan(c), ik = wave_incl_scan(k);
    if (lane == 63u) { warpSums[0][wid] = ip; warpSums[1][wid] = ik; }
    __syncthreads();
    uint32_t pp = 0, pk = 0, tp = 0, tk = 0;
    for (uint32_t w = 0; w < 16u; w++) {
      if (w < wid) { pp += warpSums[0][w]; pk += warpSums[1][w]; }
      tp += warpSums[0][w];
      tk += warpSums[1][w];
    }
    if (i < nBlocks) {
      a.offsets[i] = carryP + pp + ip - c;
      a.cursors[i] = carryP + pp + ip - c;  // the scatter's running slot
      a.chunkStart[i] = carryC + pk + ik - k;
    }
    carryP += tp;
    carryC += tk;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.chunkStart[nBlocks] = carryC;
}

struct P2GScan {
  struct Args {
    P2GArgs a;
    uint32_t nBlocks;
  };
  static constexpr uint32_t kThreads = 1024;
  static constexpr const char* kName = "p2g_scan";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { p2g_scan_body(a.a, a.nBlocks, block, blocks); }
};


// The accumulator of a block's nodes in LDS, component-major, one pad word
// per 8 nodes so that nodes 8 apart (one step in y) fall in different banks.
constexpr uint32_t kAccPitch = kBlkNodes + kBlkNodes / 8;
__device__ __forceinline__ uint32_t acc_idx(uint32_t l, int k) { return (uint32_t)k * kAccPitch + l + (l >> 3); }

// Step 4: one workgroup per chunk of a block's pairs.
__device__ __forceinline__ void p2g_accum_body(P2GArgs a, uint32_t nBlocks, uint32_t block, uint32_t blocks) {
  __shared__ int32_t acc[4 * kAccPitch];
  __shared__ uint32_t s_blk;
  const uint32_t total = a.chunkStart[nBlocks];
  const f3 gmin = mk3(a.u->gridBoundsMin[0], a.u->gridBoundsMin[1], a.u->gridBoundsMin[2]);
  const f3 gmax = mk3(a.u->gridBoundsMax[0], a.u->gridBoundsMax[1], a.u->gridBoundsMax[2]);
  for (uint32_t c = block; c < total; c += blocks) {
    if (threadIdx.x == 0) {
      // last block whose first chunk is <= c
      uint32_t lo = 0, hi = nBlocks;
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.chunkStart[mid] <= c) lo = mid; else hi = mid;
      }
      s_blk = lo;
    }
    for (uint32_t i = threadIdx.x; i < 4 * kAccPitch; i += blockDim.x) acc[i] = 0;
    __syncthreads();
    const uint32_t b = s_blk;
    const uint32_t bx = b % a.nb, by = (b / a.nb) % a.nb, bz = b / (a.nb * a.nb);
    const uint32_t x0 = bx * kBlk, y0 = by * kBlk, z0 = bz * kBlk;
    const uint32_t first = a.offsets[b] + (c - a.chunkStart[b]) * a.chunk;
    const uint32_t end = min(first + a.chunk, a.offsets[b] + a.counts[b]);
    for (uint32_t q = first + threadIdx.x; q < end; q += blockDim.x) {
      const uint32_t e = a.pairs[q];
      const float4 pp = a.prev[e], pn = a.now[e];
      const f3 p = mk3(pn.x, pn.y, pn.z);
      const f3 vel = p - mk3(pp.x, pp.y, pp.z);
      const Cell co = grid_cell(a.d, gmin, gmax, p);
      for (uint32_t z = co.mn[2]; z <= co.mx[2]; z++)
        for (uint32_t y = co.mn[1]; y <= co.mx[1]; y++)
          for (uint32_t x = co.mn[0]; x <= co.mx[0]; x++) {
            if (x - x0 >= kBlk || y - y0 >= kBlk || z - z0 >= kBlk) continue;  // another block's node
            const f3 w = cell_w(x, y, z, co.pin);
            const f3 cv = vel * w;
            const int32_t v[4] = {f2i(cv.x * kGridMul), f2i(cv.y * kGridMul), f2i(cv.z * kGridMul),
                                  f2i(length(w) * kGridMul)};
            const uint32_t l = ((z - z0) * kBlk + (y - y0)) * kBlk + (x - x0);
#pragma unroll
            for (int k = 0; k < 4; k++)
              if (v[k]) atomicAdd(&acc[acc_idx(l, k)], v[k]);
          }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < kBlkNodes; l += blockDim.x) {
      const uint32_t x = x0 + l % kBlk, y = y0 + (l / kBlk) % kBlk, z = z0 + l / (kBlk * kBlk);
      if (x >= a.d || y >= a.d || z >= a.d) continue;
      int32_t* cell = reinterpret_cast<int32_t*>(&a.dv[grid_idx(a.d, x, y, z)]);
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (acc[acc_idx(l, k)]) atomicAdd(cell + k, acc[acc_idx(l, k)]);
    }
    __syncthreads();
  }
}

struct P2GAccum {
  struct Args {
    P2GArgs a;
    uint32_t nBlocks;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "grid_post_sim";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) { p2g_accum_body(a.a, a.nBlocks, block, blocks); }
};


// Scratch of fh_grid_post_sim_noalloc: a fixed device array, so that the
// call allocates nothing (callers that may not allocate: code under HIP
// graph capture, serving loops). Every call on one GPU shares it: calls
// that overlap on one GPU race, and each call clears the block counts it
// needs (the kernels write everything else before reading it).
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noallocMu;
uint32_t* g_noallocPtr[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device): its
// host-side name is not a device address, so every kernel argument and fill
// uses the address looked up here, once per device.
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fh::fail(FH_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_noallocMu);
  if (g_noallocPtr[device]) return FH_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fh::fail(FH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_noallocPtr[device] = static_cast<uint32_t*>(p);
  return FH_OK;
}

int noalloc_scratch(int device, uint32_t** out) {
  std::lock_guard<std::mutex> lock(g_noallocMu);
  *out = device >= 0 && device < kMaxDevices ? g_noallocPtr[device] : nullptr;
  if (!*out) return fh::fail(FH_ERR_UNSUPPORTED, "fh_grid_post_sim_noalloc: no scratch address for this device");
  return FH_OK;
}

// fh_grid_post_sim and fh_grid_post_sim_noalloc: the same kernels, with
// their scratch in the context's grow-on-demand scratch or in `scratch`.
int grid_post_sim(fh_ctx* ctx, uint32_t grid_dims, const fh_sim_uniforms* sim_uniforms,
                  const fh_hair_data* hair_data, const fh_vec4* positions_prev,
                  const fh_vec4* positions_now, fh_density_velocity* density_velocity,
                  uint32_t strands_count, uint32_t points_per_strand, bool noalloc) {
  FH_CHECK_CTX(ctx);
  (void)hair_data;
  FH_REQUIRE(grid_dims >= 2 && grid_dims <= 1024, "fh_grid_post_sim: grid_dims must be in [2, 1024]");
  FH_REQUIRE(sim_uniforms && positions_prev && positions_now && density_velocity,
             "fh_grid_post_sim: null buffer");
  const uint64_t cells = (uint64_t)grid_dims * grid_dims * grid_dims;
  const uint64_t n = (uint64_t)strands_count * points_per_strand;
  FH_REQUIRE(n < (1ull << 28), "fh_grid_post_sim: too many points");
  P2GArgs a;
  a.d = grid_dims;
  a.nb = fh::div_ceil(grid_dims, kBlk);
  a.u = sim_uniforms;
  a.prev = reinterpret_cast<const float4*>(positions_prev);
  a.now = reinterpret_cast<const float4*>(positions_now);
  a.dv = density_velocity;
  a.nPoints = (uint32_t)n;
  // about two chunks per CU for the whole grid, more for hot blocks
  a.chunk = std::min(8192u, std::max(512u, fh::div_ceil(fh::div_ceil((uint32_t)n, (uint32_t)std::max(1, ctx->num_cus) * 2u), 256u) * 256u));
  const uint32_t nBlocks = a.nb * a.nb * a.nb;
  // scratch: counts, cursors, offsets, chunk starts (+1), up to 8 pairs per point
  const size_t words = (size_t)nBlocks * 4 + 1 + 8 * n;
  uint32_t* base = nullptr;
  int rc = FH_OK;
  if (noalloc) {
    if (words * 4 + 256 > kNoAllocScratchBytes)
      return fh::fail(FH_ERR_CAPACITY, "fh_grid_post_sim_noalloc: " + std::to_string(words * 4 + 256) +
                                           " bytes of scratch needed, more than its fixed " +
                                           std::to_string(kNoAllocScratchBytes) +
                                           " (about 500,000 points); use fh_grid_post_sim");
    if ((rc = noalloc_scratch(ctx->device, &base))) return rc;
  } else {
    if ((rc = fh::ensure_scratch(ctx, words * 4 + 256))) return rc;
    base = reinterpret_cast<uint32_t*>(ctx->scratch);
  }
  a.counts = base;
  a.cursors = base + nBlocks;
  a.offsets = base + 2 * (size_t)nBlocks;
  a.chunkStart = base + 3 * (size_t)nBlocks;
  a.pairs = base + 4 * (size_t)nBlocks + 1;
  void* dst[2] = {density_velocity, base};
  const uint64_t nw[2] = {cells * 4, nBlocks};  // the grid (gridPostSimPass.ts:36), block counts
  if ((rc = fh::zero_many(ctx, 2, dst, nw, "p2g_clear"))) return rc;
  if (n == 0) return FH_OK;
  // points per binning workgroup: about 4 workgroups per CU, at least 1024
  const uint32_t cus = (uint32_t)std::max(1, ctx->num_cus);
  const uint32_t span = std::min(1u << 16, std::max(1024u, fh::div_ceil(fh::div_ceil((uint32_t)n, cus * 4u), 1024u) * 1024u));
  const uint32_t pblocks = fh::div_ceil((uint32_t)n, span);
  const bool lds = nBlocks <= kP2GHistMax;
  if (lds) rc = fh::launch<P2GBin<false, true>>(ctx, pblocks, 0, {a, nBlocks, span});
  else rc = fh::launch<P2GBin<false, false>>(ctx, pblocks, 0, {a, nBlocks, span});
  if (rc) return rc;
  if ((rc = fh::launch<P2GScan>(ctx, 1, 0, {a, nBlocks}))) return rc;
  if (lds) rc = fh::launch<P2GBin<true, true>>(ctx, pblocks, 0, {a, nBlocks, span});
  else rc = fh::launch<P2GBin<true, false>>(ctx, pblocks, 0, {a, nBlocks, span});
  if (rc) return rc;
  // chunks: about 2n / chunk for the pairs, plus a partial one per non-empty block
  const uint32_t accumBlocks = std::min(fh::persistent_blocks(ctx), fh::div_ceil(2u * (uint32_t)n, a.chunk) + 64u);
  return fh::launch<P2GAccum>(ctx, accumBlocks, 0, {a, nBlocks});
}

}  // namespace

namespace fh {
int setup_noalloc_scratch(int device) { return noalloc_scratch_setup(device); }
}  // namespace fh

extern "C" {

int fh_grid_pre_sim(fh_ctx* ctx, uint32_t grid_dims, const fh_sim_uniforms* sim_uniforms,
                    const fh_density_velocity* density_velocity,
                    fh_density_grad_wind* density_grad_wind, const float* sdf_texture,
                    uint32_t sdf_dims) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(grid_dims >= 2 && grid_dims <= 1024, "fh_grid_pre_sim: grid_dims must be in [2, 1024]");
  FH_REQUIRE(sdf_dims >= 1, "fh_grid_pre_sim: sdf_dims must be >= 1");
  FH_REQUIRE(sim_uniforms && density_velocity && density_grad_wind && sdf_texture,
             "fh_grid_pre_sim: null buffer");
  const uint32_t n = grid_dims * grid_dims * grid_dims;
  return fh::launch<GridPreSim>(ctx, fh::div_ceil(n, 256), 0,
                                {grid_dims, sim_uniforms, density_velocity, density_grad_wind, sdf_texture, sdf_dims});
}

int fh_sim_integrate(fh_ctx* ctx, uint32_t grid_dims, const fh_sim_uniforms* sim_uniforms,
                     const fh_hair_data* hair_data, fh_vec4* positions_prev,
                     const fh_vec4* positions_now, const float* segment_lengths,
                     const float* sdf_texture, uint32_t sdf_dims,
                     const fh_density_velocity* density_velocity,
                     const fh_density_grad_wind* density_grad_wind,
                     const fh_vec4* positions_initial, uint32_t strands_count,
                     uint32_t points_per_strand) {
  FH_CHECK_CTX(ctx);
  (void)hair_data;
  FH_REQUIRE(grid_dims >= 2 && grid_dims <= 1024, "fh_sim_integrate: grid_dims must be in [2, 1024]");
  FH_REQUIRE(points_per_strand >= 4 && points_per_strand <= 64,
             "fh_sim_integrate: points_per_strand must be in [4, 64] (tfxFileLoader.ts:164)");
  FH_REQUIRE(sdf_dims >= 1, "fh_sim_integrate: sdf_dims must be >= 1");
  FH_REQUIRE(sim_uniforms && positions_prev && positions_now && segment_lengths && sdf_texture &&
                 density_velocity && density_grad_wind && positions_initial,
             "fh_sim_integrate: null buffer");
  if (strands_count == 0) return FH_OK;
  // 16 strands per workgroup when 32 would leave CUs with fewer than two
  // workgroups (the sweeps then run on twice as many waves)
  const uint32_t L = strands_count < 64u * fh::persistent_blocks(ctx) / 8u ? 16u : 32u;
  const size_t lds = (size_t)points_per_strand * L * (3 * sizeof(float4) + sizeof(float));
  FH_REQUIRE(lds <= ctx->max_lds, "fh_sim_integrate: points_per_strand too large for this device's LDS");
  const SimArgs args{grid_dims, sim_uniforms, reinterpret_cast<float4*>(positions_prev),
                     reinterpret_cast<const float4*>(positions_now), segment_lengths, sdf_texture, sdf_dims,
                     density_velocity, density_grad_wind, reinterpret_cast<const float4*>(positions_initial),
                     strands_count, points_per_strand};
  if (L == 16u) return fh::launch<SimIntegrate<16>>(ctx, fh::div_ceil(strands_count, 16u), (uint32_t)lds, args);
  return fh::launch<SimIntegrate<32>>(ctx, fh::div_ceil(strands_count, 32u), (uint32_t)lds, args);
}

int fh_grid_post_sim(fh_ctx* ctx, uint32_t grid_dims, const fh_sim_uniforms* sim_uniforms,
                     const fh_hair_data* hair_data, const fh_vec4* positions_prev,
                     const fh_vec4* positions_now, fh_density_velocity* density_velocity,
                     uint32_t strands_count, uint32_t points_per_strand) {
  return grid_post_sim(ctx, grid_dims, sim_uniforms, hair_data, positions_prev, positions_now, density_velocity,
                       strands_count, points_per_strand, false);
}

int fh_grid_post_sim_noalloc(fh_ctx* ctx, uint32_t grid_dims, const fh_sim_uniforms* sim_uniforms,
                             const fh_hair_data* hair_data, const fh_vec4* positions_prev,
                             const fh_vec4* positions_now, fh_density_velocity* density_velocity,
                             uint32_t strands_count, uint32_t points_per_strand) {
  return grid_post_sim(ctx, grid_dims, sim_uniforms, hair_data, positions_prev, positions_now, density_velocity,
                       strands_count, points_per_strand, true);
}

}  // extern "C"
