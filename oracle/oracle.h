/*
 * oracle.h — CPU restatement of the reference's trace path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker or the timed CPU baseline, never as the product path. It follows the
 * reference's TypeScript line by line in scalar FP64 (src/bvh.ts, src/random.ts, src/brdf.ts,
 * src/color.ts, src/tracer.ts, src/progressive.ts, src/raytrace.ts); each function cites the
 * lines it restates in oracle.cpp.
 *
 * Two builds: liboracle.so evaluates pow/exp/log/sin/cos with the portable routines of
 * meercadraytracer_amd/csrc/jsmath.h (bit-identical to the device build, so GPU output can be
 * compared bitwise); liboracle_libm.so uses glibc's libm, an independent implementation, so the
 * two differ only by transcendental rounding, as V8's Math functions differ from both.
 */
#ifndef MCR_ORACLE_H
#define MCR_ORACLE_H

#include <stdint.h>
#include "../include/mcr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bvh.ts:16-178 */
int orc_build_bvh(const float *triangles, uint32_t count, float *bounds, int32_t *nodes,
                  uint32_t *order, uint32_t *node_count);

/* bvh.ts:202-282, one query per ray; rays n*6 doubles. */
void orc_intersect_closest(const mcr_geometry *g, const double *rays, uint32_t n, double t_min,
                           double t_max, int cull, int32_t *tri, double *t, double *u, double *v,
                           uint8_t *front, mcr_counters *counters);
/* bvh.ts:289-332 with a visitor that counts and continues. */
void orc_intersect_any_count(const mcr_geometry *g, const double *rays, uint32_t n, double t_min,
                             double t_max, uint32_t *count, mcr_counters *counters);
/* Brute force over every triangle (tests/bvh.test.ts:28-52). */
void orc_brute_force(const float *triangles, uint32_t count, const double *rays, uint32_t n,
                     int32_t *tri, double *t, uint32_t *all);

/* Precision-form checker (not a reference restatement): closest hit over every triangle under
 * the FP32 watertight test of Woop, Benthin and Wald (2013) on the precision form's FP32 vertices.
 * rays: n*6 floats. threads: worker threads to use. */
void orc_wide_brute(const float *triangles, const uint8_t *sides, uint32_t count, const float *rays,
                    uint32_t n, float t_min, float t_max, int cull, int32_t *tri, float *t, float *u,
                    float *v, uint8_t *front, uint32_t threads);

/* textures.ts:247-253 sampleTexture and :287-317 sampleEnvironment, for known-answer tests. */
void orc_sample_texture(const mcr_texture *texture, double u, double v, double *out4);
void orc_sample_environment(const mcr_environment *env, const mcr_texture *textures, double dx, double dy,
                            double dz, double *out3);
double orc_atan2(double y, double x);
double orc_asin(double x);

/* random.ts */
uint32_t orc_hash(uint32_t x);
uint32_t orc_seed(int32_t a, int32_t b, int32_t c, int32_t d);
void orc_random_stream(uint32_t state, uint32_t n, double *out);

/* brdf.ts / color.ts, scalar entry points for known-answer tests */
void orc_dfg(double roughness, double dot_nv, double *out2);
double orc_fresnel_weight(double dot_vh);
double orc_visibility_ggx(double alpha, double dot_nl, double dot_nv);
double orc_distribution_ggx(double alpha, double dot_nh);
double orc_blinn_phong(double shininess, double dot_nh);
double orc_smoothstep(double e0, double e1, double x);
double orc_distance_attenuation(double distance, double cutoff, double decay);
double orc_half_to_float(uint32_t h);
double orc_linear_to_srgb(double c);
double orc_srgb_to_linear(double c);
void orc_tone_map(double *c3, int mapping, double exposure);
double orc_math(int fn, double x, double y); /* 0 exp 1 log 2 pow 3 sin 4 cos 5 log2 6 hypot(x,y,0) */

/* tracer.ts:705-727 */
void orc_camera_ray(const mcr_camera *camera, double x, double y, double *out6);

/* tracer.ts:62-89: returns 0, or MCR_E_UNSUPPORTED for scene data outside this build. */
int orc_render_tile(const mcr_geometry *g, const mcr_shading *s, const mcr_camera *camera,
                    const mcr_settings *settings, mcr_tile tile, uint32_t width, uint32_t height,
                    uint32_t pass, uint32_t block, float *out, mcr_counters *counters);

/* progressive.ts:288-313 */
void orc_encode_tile(uint8_t *image, uint32_t width, mcr_tile tile, const float *data, float *sum,
                     uint32_t count, int srgb);
/* progressive.ts:160-170; returns the tile count (out may be NULL to size). */
uint32_t orc_make_tiles(uint32_t width, uint32_t height, uint32_t size, mcr_tile *out);
/* raytrace.ts:69-82 */
int orc_raytrace(const mcr_geometry *g, const mcr_shading *s, const mcr_camera *camera,
                 const mcr_settings *settings, uint32_t width, uint32_t height, uint32_t samples,
                 int srgb, uint8_t *out, mcr_counters *counters);

/* CPU baseline: the listed (tile, pass) items traced by `threads` std::threads that pull items in
 * order, as ProgressiveRenderer's runners pull its work list (progressive.ts:227-271). Outputs are
 * written to out (one RGBA tile per item at item offsets) if out != NULL. Returns wall seconds. */
double orc_render_items_mt(const mcr_geometry *g, const mcr_shading *s, const mcr_camera *camera,
                           const mcr_settings *settings, const mcr_tile_job *jobs,
                           uint32_t job_count, uint32_t threads, float *out,
                           mcr_counters *counters);

#ifdef __cplusplus
}
#endif

#endif
