/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY (see fh_oracle.c).
 * Sequential restatement of the reference's hair compute chain, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. Never linked into or called by the product library.
 */
#ifndef FH_ORACLE_H
#define FH_ORACLE_H

#include "../include/fh_hair.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host threads for the independent loops (default 1); results do not depend on it. */
void or_set_threads(int n);
int or_get_threads(void);

float or_sample_sdf(const float* sdf, uint32_t d, const float* bmin,
                    const float* bmax, const float* p);

void or_grid_pre_sim(uint32_t d, const fh_sim_uniforms* u,
                     const fh_density_velocity* dv, fh_density_grad_wind* out,
                     const float* sdf, uint32_t sdf_dims);

void or_sim_integrate(uint32_t d, const fh_sim_uniforms* u, fh_vec4* prev,
                      const fh_vec4* now, const float* seglen,
                      const float* sdf, uint32_t sdf_dims,
                      const fh_density_velocity* dv,
                      const fh_density_grad_wind* dgw, const fh_vec4* initial,
                      uint32_t S, uint32_t P);

void or_grid_post_sim(uint32_t d, const fh_sim_uniforms* u, const fh_vec4* prev,
                      const fh_vec4* now, fh_density_velocity* dv, uint32_t S,
                      uint32_t P);

void or_debug_raster(const fh_render_uniforms* u, const fh_vec4* positions,
                     const fh_vec4* tangents, uint32_t S, uint32_t P,
                     uint32_t W, uint32_t H, uint32_t* out);

void or_tiles_draw(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                   const fh_render_uniforms* u, const fh_hair_data* hd,
                   const fh_vec4* positions, const fh_vec4* tangents,
                   fh_tile_result* tiles, const float* depth, void* segbuf,
                   uint32_t capacity, uint32_t* countPerTile,
                   uint32_t* slotKey, uint32_t S, uint32_t P);

void or_tiles_sort(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                   const uint32_t* countPerTile, uint32_t* tileList,
                   fh_sort_bucket* buckets);

void or_fine_raster(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                    const fh_render_uniforms* u, const fh_vec4* positions,
                    const fh_tile_result* tiles, const void* segbuf,
                    uint32_t capacity, void* rasterResult, const float* depth,
                    const fh_vec4* shading, const fh_vec4* tangents,
                    const uint32_t* tileList, uint32_t S, uint32_t P,
                    int canonical);

void or_shading(const fh_render_cfg* cfg, const fh_render_uniforms* u,
                const fh_vec4* positions, const fh_vec4* tangents,
                fh_vec4* out, const float* shadowTex, uint32_t shadowW,
                uint32_t shadowH, const float* aoTex, uint32_t aoW,
                uint32_t aoH, const float* depth, uint32_t dW, uint32_t dH,
                uint32_t S, uint32_t P);

uint32_t or_f2h(float f);
float or_h2f(uint32_t h);
uint32_t or_bayer8(uint32_t x, uint32_t y);
float or_aces(float x);
uint32_t or_srgb8(float c);

void or_hair_combine(const fh_render_cfg* cfg, uint32_t W, uint32_t H,
                     const fh_render_uniforms* u, const fh_tile_result* tiles,
                     const void* segbuf, const void* rasterResult,
                     const uint32_t* countPerTile, uint16_t* hdr);

void or_present(uint32_t W, uint32_t H, const fh_render_uniforms* u,
                const uint16_t* hdr, const float* depth, const float* ao,
                uint32_t aoW, uint32_t aoH, const uint32_t* normals,
                int ditherPerPixel, uint8_t* out, uint32_t pitch);

uint32_t or_encode_oct(float x, float y, float z);

void or_hw_hair(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u,
                const fh_vec4* pos, const fh_vec4* tan, float* depth, uint32_t* normals,
                uint32_t S, uint32_t P);

void or_hw_coverage(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u,
                    const fh_vec4* pos, const fh_vec4* tan, uint32_t* counts, uint32_t S, uint32_t P);

void or_ao(const fh_render_uniforms* u, const float* depth, uint32_t W, uint32_t H,
           const uint32_t* normals, float* out, uint32_t aoW, uint32_t aoH);

float or_bg_hash(float x, float y);
float or_bg_noise(float x, float y);
float or_bg_fractal(float x, float y, float scale);
void or_background(uint32_t W, uint32_t H, const fh_render_uniforms* u, uint16_t* hdr);
void or_meshes(int shadow, uint32_t W, uint32_t H, const fh_render_uniforms* u, const fh_mesh* meshes,
               uint32_t count, const float* shadowMap, uint32_t shadowSize, const float* ao, uint32_t aoW,
               uint32_t aoH, float* depth, uint32_t* normals, uint16_t* hdr);
void or_disney_pbr(const float* positionWS, const float* normal, const float* toEye, const float* albedo,
                   float roughness, float isMetallic, const fh_light* light, float* out);
float or_directional_shadow(const fh_render_uniforms* u, const float* sm, uint32_t smSize, const float* posWS,
                            const float* normalWS, const float* pss);

#ifdef __cplusplus
}
#endif

#endif
