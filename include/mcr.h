/*
 * mcr.h — C-ABI of the MI355X trace path for MeerCADRaytracer scenes.
 *
 * Plain pointers and sizes only: no torch or HIP types cross this boundary. Every entry point
 * returns MCR_OK (0) or a negative MCR_E_* code; the message for the last failure on the
 * calling thread is returned by mcr_last_error().
 *
 * The data structs restate the reference's plain-data scene format field for field
 * (src/types.ts:43-211). Arrays are caller-owned; the library copies what it keeps.
 */
#ifndef MCR_H
#define MCR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define MCR_OK 0
#define MCR_E_INVALID (-1)     /* a null pointer, a size out of range, an unknown enum value */
#define MCR_E_UNSUPPORTED (-2) /* scene data the path cannot trace (an index out of range, too many lights) */
#define MCR_E_DEVICE (-3)      /* a HIP call failed */
#define MCR_E_NO_SCENE (-4)    /* protocol.ts:77-79: the tile refers to a scene not received */
#define MCR_E_OVERFLOW (-5)    /* a fixed device capacity was exceeded */
#define MCR_E_CANCELLED (-6)   /* the job's generation was cancelled before it ran */
#define MCR_E_TIMEOUT (-7)     /* a wait ended before the request finished */
#define MCR_E_INTERNAL (-8)    /* an internal consistency check failed (a bug, not bad input) */

/* ---- enums restated from src/constants.ts and src/types.ts ------------------------------ */
enum { MCR_FRONT_SIDE = 0, MCR_BACK_SIDE = 1, MCR_DOUBLE_SIDE = 2 };           /* constants.ts:5-7 */
enum { MCR_MULTIPLY = 0, MCR_MIX = 1, MCR_ADD = 2 };                           /* constants.ts:9-11 */
enum {                                                                         /* constants.ts:13-20 */
  MCR_NO_TONE_MAPPING = 0, MCR_LINEAR_TONE_MAPPING = 1, MCR_REINHARD_TONE_MAPPING = 2,
  MCR_CINEON_TONE_MAPPING = 3, MCR_ACES_FILMIC_TONE_MAPPING = 4, MCR_CUSTOM_TONE_MAPPING = 5,
  MCR_AGX_TONE_MAPPING = 6, MCR_NEUTRAL_TONE_MAPPING = 7
};
enum { MCR_TRI_OPAQUE = 1, MCR_TRI_CAST_SHADOW = 2, MCR_TRI_RECEIVE_SHADOW = 4 }; /* types.ts:133-135 */
enum {                                                                          /* types.ts:41 */
  MCR_MODEL_PHYSICAL = 0, MCR_MODEL_PHONG = 1, MCR_MODEL_LAMBERT = 2, MCR_MODEL_TOON = 3,
  MCR_MODEL_BASIC = 4
};
enum {                                                                          /* types.ts:91-113 */
  MCR_LIGHT_AMBIENT = 0, MCR_LIGHT_HEMISPHERE = 1, MCR_LIGHT_DIRECTIONAL = 2, MCR_LIGHT_POINT = 3,
  MCR_LIGHT_SPOT = 4
};
enum { MCR_SHADOWS_ALL = 0, MCR_SHADOWS_CAST = 1, MCR_SHADOWS_NONE = 2 };         /* types.ts:195 */
enum { MCR_FOG_NONE = 0, MCR_FOG_LINEAR = 1, MCR_FOG_EXP2 = 2 };                 /* types.ts:115-121 */
enum { MCR_BACKGROUND_CLEAR = 0, MCR_BACKGROUND_COLOR = 1, MCR_BACKGROUND_SCREEN = 2,
       MCR_BACKGROUND_ENVIRONMENT = 3 };                                         /* types.ts:123-127 */
enum { MCR_WRAP_REPEAT = 1000, MCR_WRAP_CLAMP = 1001, MCR_WRAP_MIRROR = 1002 };  /* constants.ts:29-31 */
enum { MCR_COMBINE_MULTIPLY = 0, MCR_COMBINE_MIX = 1, MCR_COMBINE_ADD = 2 };     /* constants.ts:9-11 */
enum { MCR_ENV_EQUIRECT = 0, MCR_ENV_CUBE = 1 };                                 /* types.ts:23-39 */

/* CompiledTexture (types.ts:8-21): RGBA texels, first row first, 8-bit (is_float 0; srgb marks
 * sRGB-encoded colour channels) or float32 (is_float 1; half floats decoded as textures.ts:63-105
 * does). matrix is three.js Texture.matrix, column-major 3x3. */
typedef struct mcr_texture {
  uint32_t width, height;
  const void *data;
  int32_t is_float, srgb, flip_y, wrap_s, wrap_t, pad0;
  double matrix[9];
} mcr_texture;

/* CompiledEnvironment (types.ts:23-39): equirect (texture) or cube (faces +x -x +y -y +z -z,
 * flip_x). rotation is column-major 3x3, applied to directions before the lookup. */
typedef struct mcr_environment {
  int32_t kind;
  int32_t texture;
  int32_t faces[6];
  int32_t flip_x, pad0;
  double rotation[9];
  double intensity;
} mcr_environment;

/* ---- plain-data scene (types.ts) -------------------------------------------------------- */

/* CompiledGeometry (types.ts:137-156) with its CompiledBvh (types.ts:158-167). Per-triangle
 * arrays are in BVH leaf order (compile.ts:363-387). colors may be NULL (no vertex colours). */
typedef struct mcr_geometry {
  uint32_t triangle_count;
  uint32_t node_count;
  const float *triangles;   /* n*9: v0, e1 = v1-v0, e2 = v2-v0 */
  const float *normals;     /* n*9 */
  const float *uvs;         /* n*6 */
  const float *tangents;    /* n*4 */
  const float *colors;      /* n*12 or NULL */
  const uint32_t *materials;/* n */
  const uint8_t *flags;     /* n, MCR_TRI_* bits */
  const uint8_t *sides;     /* n, MCR_*_SIDE */
  const float *bvh_bounds;  /* node_count*6: min xyz, max xyz */
  const int32_t *bvh_nodes; /* node_count*3: first|left, count, axis */
  double epsilon;
} mcr_geometry;

/* CompiledMaterial (types.ts:43-89). Texture and environment indices are -1 when absent. */
typedef struct mcr_material {
  int32_t model, side;
  double color[3];
  double opacity;
  int32_t transparent;
  int32_t vertex_colors;
  double alpha_test;
  int32_t flat_shading, tone_mapped;
  int32_t map, alpha_map, emissive_map, normal_map, ao_map, roughness_map, metalness_map;
  int32_t env_map;
  double emissive[3];
  double normal_scale[2];
  double ao_map_intensity;
  double roughness, metalness;
  int32_t physical, combine;
  double ior, specular_intensity;
  double specular_color[3];
  double transmission, thickness;
  double attenuation[3];
  double shininess;
  double specular[3];
  double env_map_intensity, reflectivity;
} mcr_material;

/* CompiledLight (types.ts:91-113): one struct for all kinds; fields a kind lacks are ignored. */
typedef struct mcr_light {
  int32_t type, cast_shadow;
  double color[3];
  double ground_color[3];
  double direction[3];
  double position[3];
  double distance, decay, cone_cos, penumbra_cos;
} mcr_light;

/* Everything of CompiledScene but the geometry (protocol.ts:19-27 ShadingPayload), with the
 * textures message's arrays (protocol.ts:31-35). Texture and environment indices are checked
 * against the counts; at most MCR_MAX_LIGHTS_PER_SCENE lights (the lights array, ambient and hemisphere
 * entries included), one fewer when environments are present (one light slot carries the
 * environment-diffuse ray of tracer.ts:343-353). */
#define MCR_MAX_LIGHTS_PER_SCENE 64
typedef struct mcr_shading {
  uint32_t material_count, light_count;
  const mcr_material *materials;
  const mcr_light *lights;
  uint32_t texture_count, environment_count;
  int32_t background_kind;
  int32_t environment;       /* scene.environment index or -1 */
  double background_color[3];
  int32_t fog_kind, pad0;    /* MCR_FOG_* */
  double fog_color[3];
  double fog_near, fog_far, fog_density;
  const mcr_texture *textures;           /* texture_count */
  const mcr_environment *environments;   /* environment_count */
  int32_t background_texture;            /* MCR_BACKGROUND_SCREEN */
  int32_t background_environment;        /* MCR_BACKGROUND_ENVIRONMENT */
  double background_intensity;           /* scene.backgroundIntensity of a screen background */
} mcr_shading;

/* CompiledCamera (types.ts:181-190), matrices column-major. */
typedef struct mcr_camera {
  double projection_inverse[16];
  double world[16];
  double forward[3];
  double position[3];
  int32_t orthographic, pad0;
} mcr_camera;

/* TraceSettings (types.ts:193-204). seed is ToInt32(settings.seed) as random.ts:7 mixes it. */
typedef struct mcr_settings {
  int32_t max_bounces, shadows;
  int32_t global_illumination, tone_mapping;
  double tone_mapping_exposure;
  double clear_color[3];
  double clear_alpha;
  uint32_t seed, pad0;
} mcr_settings;

/* Tile (types.ts:206-211). */
typedef struct mcr_tile {
  int32_t x, y, width, height;
} mcr_tile;

/* One item of a batched render: a tile of an image for one pass. */
typedef struct mcr_tile_job {
  mcr_tile tile;
  uint32_t width, height; /* image size */
  uint32_t pass, block;
  uint64_t out_offset;    /* float offset of this tile's RGBA in the batch output buffer */
} mcr_tile_job;

/* Per-call ray and work counters (kernel counter mode). Mrays/s counts closest + any queries. */
typedef struct mcr_counters {
  uint64_t closest_queries;  /* bvh.ts:202 intersectClosest calls */
  uint64_t any_queries;      /* bvh.ts:289 intersectAny calls */
  uint64_t node_visits;      /* inner-node visits (closest: 64-byte child-pair reads) */
  uint64_t leaf_visits;
  uint64_t triangle_tests;   /* triangle records read */
  uint64_t paths;            /* camera samples traced */
  uint64_t closest_node_visits, closest_triangle_tests;  /* split by query kind, for the */
  uint64_t any_node_visits, any_triangle_tests;          /* per-kernel algorithmic bytes   */
  uint64_t shade_waves, mixed_view_waves;  /* batched runs: shade waves, and those mixing views */
  /* material-sorted shading (mcr_set_shade_sort): paths shaded per shade class (0 miss, 1 basic,
   * 2 lambert/toon, 3 phong, 4 physical, 5 physical with maps, 6 transmissive, 7 see-through), and
   * the shade waves that would mix classes in the unsorted queue order */
  uint64_t shade_class_lanes[8], class_mixed_waves;
} mcr_counters;

typedef struct mcr_ctx mcr_ctx;
typedef struct mcr_scene mcr_scene;

/* ---- context ------------------------------------------------------------------------------ */
int mcr_create(int device, mcr_ctx **out);
void mcr_destroy(mcr_ctx *ctx);
const char *mcr_last_error(void);
const char *mcr_version(void);

/* ---- builder: replaces buildBvh (src/bvh.ts:16-178), called at src/compile.ts:361 -------- */
/* Builds the binned-SAH BVH of n triangles (v0, e1, e2; n*9 floats) on the GPU. The caller
 * allocates bounds_out[(2n-1)*6], nodes_out[(2n-1)*3] and order_out[n]; *node_count_out gets
 * the node count. Arrays are bit-identical to buildBvh's. n = 0 gives 0 nodes (bvh.ts:18). */
int mcr_build_bvh(mcr_ctx *ctx, const float *triangles, uint32_t n, float *bounds_out,
                  int32_t *nodes_out, uint32_t *order_out, uint32_t *node_count_out);
/* As mcr_build_bvh, with triangles and outputs already in device memory (no host copies). */
int mcr_build_bvh_device(mcr_ctx *ctx, const float *d_triangles, uint32_t n, float *d_bounds,
                         int32_t *d_nodes, uint32_t *d_order, uint32_t *node_count_out);

/* ---- scene residency: the geometry / shading messages of src/protocol.ts:30-32,64-74 ------ */
/* Uploads a compiled geometry (BVH order, with its BVH) once; it stays in HBM until released.
 * version is the geometryVersion tile requests must name (protocol.ts:77). */
int mcr_scene_create(mcr_ctx *ctx, mcr_scene **out);
int mcr_scene_set_geometry(mcr_scene *scene, const mcr_geometry *geometry, uint32_t version);
int mcr_scene_set_shading(mcr_scene *scene, const mcr_shading *shading, uint32_t version);
void mcr_scene_release(mcr_scene *scene);
/* Device bytes the scene holds (geometry + BVH + shading), for admission by memory budget. */
uint64_t mcr_scene_bytes(const mcr_scene *scene);

/* ---- tile render: replaces Tracer.renderTile (src/tracer.ts:62-89) ------------------------ */
/* Traces one sample per pixel (or per block) of the tile and writes premultiplied RGBA floats,
 * row by row over the tile (tile.width*tile.height*4), exactly as renderTile returns them.
 * geometry_version/shading_version must match what the scene holds (protocol.ts:77-79). */
int mcr_render_tile(mcr_ctx *ctx, mcr_scene *scene, uint32_t geometry_version,
                    uint32_t shading_version, const mcr_camera *camera,
                    const mcr_settings *settings, mcr_tile tile, uint32_t width, uint32_t height,
                    uint32_t pass, uint32_t block, float *rgba_out, mcr_counters *counters);
/* Many tile jobs of one scene, camera and settings in one device launch sequence. out receives
 * each job's RGBA at job.out_offset. out_device != 0 means out is a device pointer. */
int mcr_render_tiles(mcr_ctx *ctx, mcr_scene *scene, const mcr_camera *camera,
                     const mcr_settings *settings, const mcr_tile_job *jobs, uint32_t job_count,
                     float *out, int out_device, mcr_counters *counters);

/* ---- batched form: tile jobs of many scenes, cameras and settings in one wavefront run ---- */
/* No counterpart call in the reference, which renders one tile per Tracer.renderTile call
 * (src/tracer.ts:62): this is that call for a queue of tiles from many requests at once
 * (src/progressive.ts:227-271 spreads them over workers instead). Every job names a view;
 * job_views may be NULL (every job view 0). Results equal one mcr_render_tiles call per view. */
typedef struct mcr_view {
  mcr_scene *scene;
  const mcr_camera *camera;
  const mcr_settings *settings;
} mcr_view;
int mcr_render_batch(mcr_ctx *ctx, const mcr_view *views, uint32_t view_count,
                     const mcr_tile_job *jobs, const uint32_t *job_views, uint32_t job_count,
                     float *out, int out_device, mcr_counters *counters);

/* ---- served form: a request loop over many resident scenes ---------------------------------- */
/* Replaces ProgressiveRenderer.render / stop (src/progressive.ts:186-275) for many renders at
 * once: each request is one render() call (camera, settings, size, samples, tile size, preview
 * block, output colour space). A server thread keeps admitted requests' tiles in one wavefront:
 * tiles of every live request share the device (continuous batching), newly admitted requests'
 * previews go first, tiles are accumulated and encoded in pass order into the request's 8-bit
 * image in HBM (writeTile), and finished results are the bytes raytrace() gives for the same
 * input. Admission is by an HBM budget that covers the scenes added to the server and every
 * admitted request's image and running sum; requests wait, first in first out, until both
 * memory and a request slot are free. mcr_server_cancel is stop(): queued tiles of the request
 * are dropped and its buffers are freed once its samples in flight have drained. */
typedef struct mcr_server mcr_server;
typedef struct mcr_server_config {
  uint64_t hbm_budget;      /* bytes; 0 = 200 GB */
  uint32_t slots;           /* path slots of the lane that takes the passes; 0 = 2^21 with one
                               lane, 2^22 beside a preview lane */
  uint32_t window;          /* unstarted samples kept queued ahead of the slots; 0 = slots */
  uint32_t max_requests;    /* admitted requests at once; 0 = 64, at most 1024 */
  int32_t use_graph;        /* replay captured bounce iterations (1) or launch directly (0) */
  uint32_t max_lights;      /* light slots per path: the most lights (plus 1 with environments) a
                               scene added to the server may have; 0 = 16, at most MCR_MAX_LIGHTS_PER_SCENE */
  uint32_t lanes;           /* 1: one wavefront takes every work item; 2 (0 = default): a preview lane
                               takes the previews on preview_cus CUs kept for it, beside a pass lane
                               of `slots` slots on the other CUs, each on its own stream */
  uint32_t preview_slots;   /* the preview lane's path slots; 0 = 2^18 */
  uint32_t preview_cus;     /* CUs kept for the preview lane; 0 = 3/32 of the device's */
  int32_t preview_passes;   /* beside a pass lane: 1 = the preview lane also runs passes (at most an eighth
                               of its window) while no preview waits; 0 = previews only */
} mcr_server_config;
typedef struct mcr_request {
  mcr_scene *scene;         /* added with mcr_server_add_scene */
  mcr_camera camera;
  mcr_settings settings;
  uint32_t width, height;
  uint32_t samples;         /* passes (ProgressiveOptions.samples) */
  uint32_t tile_size;       /* 0 = 64 (RaytracingRenderer.ts:105) */
  uint32_t preview_block;   /* <= 1: no preview; 0 = 4 */
  int32_t srgb;             /* outputColorSpace 'srgb' (1) or 'srgb-linear' (0) */
  uint32_t geometry_version, shading_version;  /* the scene versions the request is for, as a tile
                                                  message names them (protocol.ts:37-45); 0 = any */
} mcr_request;
enum { MCR_REQ_QUEUED = 0, MCR_REQ_RUNNING = 1, MCR_REQ_DONE = 2, MCR_REQ_CANCELLED = 3, MCR_REQ_FAILED = 4 };
typedef struct mcr_request_status {
  int32_t state;
  uint32_t samples, max_samples;   /* frame samples = min over tiles (progressive.ts:214-218) */
  int32_t preview_done;            /* every tile has its preview or a pass written */
  double submit_ms, admit_ms, first_preview_ms, done_ms;  /* server clock; -1 = not yet */
  uint64_t paths;                  /* samples traced for the request */
  uint32_t preview_iterations;     /* bounce iterations from admission to the complete preview */
  uint32_t pad0;
} mcr_request_status;
typedef struct mcr_server_stats {
  uint64_t closest_queries, any_queries, paths, iterations;
  uint64_t shade_waves, mixed_view_waves;  /* shade waves, and those holding several requests */
  uint64_t bytes_in_use, requests_live, requests_queued;
  double device_ms;                /* time with at least one lane's step in flight */
  double uptime_ms;
  double host_ms;                  /* the server thread's own work: admission, planning, encodes, launches */
  double plan_ms, deliver_ms;      /* of which job-window planning, and finished-tile delivery and encodes */
  double pass_busy_ms;             /* time the lane taking the passes has a step in flight */
} mcr_server_stats;
int mcr_server_create(mcr_ctx *ctx, const mcr_server_config *config, mcr_server **out);
void mcr_server_destroy(mcr_server *server);
/* Makes a resident scene available to requests. While a server holds a scene, its geometry,
 * shading and wide BVH cannot change (those calls return MCR_E_INVALID) and mcr_scene_release
 * is deferred until the last server lets it go: requests render the arrays admitted with them. */
int mcr_server_add_scene(mcr_server *server, mcr_scene *scene);
/* Lets a scene go (SceneSync sending a new version starts here: remove, change, add again).
 * MCR_E_INVALID while a request on the scene is queued, running or still draining. */
int mcr_server_remove_scene(mcr_server *server, mcr_scene *scene);
int mcr_server_submit(mcr_server *server, const mcr_request *request, uint64_t *id_out);
int mcr_server_cancel(mcr_server *server, uint64_t id);
int mcr_server_status(mcr_server *server, uint64_t id, mcr_request_status *out);
/* Waits until the request is done, cancelled or failed; MCR_E_TIMEOUT after timeout_ms (< 0:
 * no limit). A failed request returns its own error: MCR_E_OVERFLOW when one of its shadow rays
 * crossed more counted see-through surfaces than the device keeps (that request alone fails;
 * the server and its other requests go on), or the device error that ended every request. */
int mcr_server_wait(mcr_server *server, uint64_t id, double timeout_ms);
/* The request's current frame, width*height*4 bytes, straight alpha (the frame image). */
int mcr_server_read(mcr_server *server, uint64_t id, uint8_t *rgba_out);
/* Forgets a finished or cancelled request (its id is no longer valid). */
int mcr_server_release(mcr_server *server, uint64_t id);
int mcr_server_stats_get(mcr_server *server, mcr_server_stats *out);

/* ---- accumulate + encode: replaces writeTile/encodeTile (src/progressive.ts:288-313) ------- */
/* sum may be NULL (preview tiles). image is width*height*4 bytes, straight alpha. */
int mcr_encode_tile(mcr_ctx *ctx, uint8_t *image, uint32_t width, mcr_tile tile,
                    const float *data, float *sum, uint32_t count, int srgb, int device_ptrs);
/* As mcr_encode_tile (device pointers) with the running sum kept in FP64 (width*height*4
 * doubles). No counterpart in the reference, whose Float32Array sum (progressive.ts:200) stops
 * resolving late samples of very long renders: used for renders beyond 1024 samples. */
int mcr_encode_tile64(mcr_ctx *ctx, uint8_t *image, uint32_t width, mcr_tile tile,
                      const float *data, double *sum, uint32_t count, int srgb);

/* ---- ray queries: intersectClosest / intersectAny (src/bvh.ts:202-332) over ray batches --- */
/* rays: n*6 doubles (origin, direction). Outputs per ray: triangle (-1 = none), t, u, v, front.
 * Any output pointer may be NULL. device_ptrs != 0: every pointer is device memory. */
int mcr_intersect_closest(mcr_ctx *ctx, mcr_scene *scene, const double *rays, uint32_t n,
                          double t_min, double t_max, int cull, int32_t *tri_out, double *t_out,
                          double *u_out, double *v_out, uint8_t *front_out, int device_ptrs,
                          mcr_counters *counters);
/* Number of hits intersectAny visits per ray with a visitor that always continues
 * (tests/bvh.test.ts:100-103 counts them so). */
int mcr_intersect_any_count(mcr_ctx *ctx, mcr_scene *scene, const double *rays, uint32_t n,
                            double t_min, double t_max, uint32_t *count_out, int device_ptrs,
                            mcr_counters *counters);

/* ---- the one entry point that allocates nothing --------------------------------------------- */
/* No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. It is
 * mcr_intersect_closest (intersectClosest, src/bvh.ts:202-282, over a ray batch) on device
 * pointers with the same results, except that it allocates nothing, waits for nothing and reports
 * no counters: the kernels are queued on `stream` (a hipStream_t of the scene's device; NULL =
 * the default stream) and the outputs are ready once that stream has run them. Its scratch (the
 * traversal's spilled stack entries for a fixed grid, and its counters) is one static device
 * array of MCR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and located when
 * mcr_create first runs on that device. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two threads) race and must be serialised by the caller, for
 * example by issuing them on one stream. n is at most MCR_NOALLOC_MAX_RAYS; larger batches are
 * refused with MCR_E_INVALID. */
#define MCR_NOALLOC_SCRATCH_BYTES 58737664u   /* 512 x 256 threads x 112 stack words + counters */
#define MCR_NOALLOC_MAX_RAYS 0x80000000u
int mcr_intersect_closest_noalloc(mcr_scene *scene, const double *rays, uint32_t n, double t_min,
                                  double t_max, int cull, int32_t *tri_out, double *t_out,
                                  double *u_out, double *v_out, uint8_t *front_out, void *stream);

/* ---- precision form: compressed 8-wide BVH, FP32 watertight closest hit ------------------- */
/* A second layout of a scene's geometry, built on the device from its binary BVH: 80-byte nodes
 * of eight children whose boxes are 8-bit grid coordinates over the node, rounded outward, and
 * FP32 triangles (v0, v0 + e1, v0 + e2 rounded once). Queries test boxes in FP32 with slabs
 * widened by their rounding bound (never missing a box the ray meets) and triangles with the
 * watertight test of Woop, Benthin and Wald (2013). It stands beside intersectClosest
 * (src/bvh.ts:202-282): the same arguments and hit record in FP32, with results that can differ
 * from the FP64 query where FP32 rounding decides a hit. Rebuilt after each geometry change
 * (mcr_scene_set_geometry drops it). */
typedef struct mcr_wide_info {
  uint32_t node_count, triangle_count, depth, pad0;
  uint64_t bytes;   /* nodes and triangles */
  double build_ms;  /* device time of the build */
} mcr_wide_info;
int mcr_scene_build_wide(mcr_scene *scene, mcr_wide_info *info_out);
int mcr_scene_wide_info(const mcr_scene *scene, mcr_wide_info *out);
/* rays: n*6 floats (origin, direction). Outputs as mcr_intersect_closest, in FP32; triangle ids
 * are BVH-order indices, as there. The winner is found in FP32; its t, u and v are then the
 * reference's FP64 test on that triangle and the widened ray (bvh.ts:228-252), rounded to FP32,
 * except where that test rejects it (a ray within FP32 rounding of an edge). */
#define MCR_ISECT_FP64 0       /* the reference's FP64 arithmetic on its binary BVH (default) */
#define MCR_ISECT_WIDE_FP32 1  /* the precision form: FP32 on the 8-wide BVH */
/* How renders on this context (mcr_render_tile/tiles/batch) intersect rays: closest hits
 * (tracer.ts:113, bvh.ts:202) and shadow rays (tracer.ts:529-542, bvh.ts:289). With
 * MCR_ISECT_WIDE_FP32 every scene rendered needs mcr_scene_build_wide; shading, accumulation and
 * sampling stay as in MCR_ISECT_FP64, so images differ only where FP32 intersection does. No
 * counterpart in the reference, which has one arithmetic. */
int mcr_set_intersection(mcr_ctx *ctx, int mode);
int mcr_intersect_closest_wide(mcr_ctx *ctx, mcr_scene *scene, const float *rays, uint32_t n,
                               float t_min, float t_max, int cull, int32_t *tri_out, float *t_out,
                               float *u_out, float *v_out, uint8_t *front_out, int device_ptrs,
                               mcr_counters *counters);

/* ---- scene flattening: compileGeometry (src/compile.ts:212-306, 327-432) on the device ---- */
/* One drawn group of a mesh (geometry.groups, or one group covering everything): what addMesh
 * (compile.ts:237-252) derives from its material. Invisible and wireframe groups are left out by
 * the caller, as addMesh skips them. */
typedef struct mcr_mesh_group {
  uint32_t start, count;     /* index range; count 0xffffffff = Infinity */
  uint32_t material;         /* materialIndex(material) */
  uint8_t flags;             /* MCR_TRI_*: cast/receive of the mesh, opaque of the material */
  uint8_t side;              /* material.side */
  uint8_t flat_shading;      /* material.flatShading */
  uint8_t vertex_colors;     /* material.vertexColors === true and the geometry has colours */
} mcr_mesh_group;

/* One mesh, or one instance of an InstancedMesh (world = matrixWorld x instanceMatrix, as
 * multiply4 forms it, compile.ts:180-189). Positions are the ones getVertexPosition gives for
 * morphed and skinned meshes. */
typedef struct mcr_mesh {
  double world[16];          /* column-major, matrixWorld.elements */
  const float *positions;    /* vertex_count*3 */
  const float *normals;      /* vertex_count*3, or NULL */
  const float *uvs;          /* vertex_count*2, or NULL */
  const float *colors;       /* vertex_count*color_size, or NULL */
  const uint32_t *index;     /* index_count entries, or NULL (non-indexed) */
  const mcr_mesh_group *groups;
  uint32_t vertex_count, index_count, color_size, group_count;
  uint32_t draw_start, draw_count;  /* geometry.drawRange; count 0xffffffff = Infinity */
} mcr_mesh;

/* compileGeometry's output arrays, caller-allocated for triangle_count triangles (from
 * mcr_compile_count) and 2*triangle_count-1 BVH nodes; colors only when any_colors. */
typedef struct mcr_geometry_out {
  uint32_t triangle_count;   /* in: the count mcr_compile_count gave */
  uint32_t node_count;       /* out */
  float *triangles, *normals, *uvs, *tangents, *colors;
  uint32_t *materials;
  uint8_t *flags, *sides;
  float *bvh_bounds;
  int32_t *bvh_nodes;
  double epsilon;            /* out */
} mcr_geometry_out;

/* Triangles the meshes produce (addMesh's loops, compile.ts:235-251) and whether any group has
 * vertex colours (the colors array is then present, white where a group has none). Host only. */
int mcr_compile_count(const mcr_mesh *meshes, uint32_t mesh_count, uint32_t *triangle_count_out,
                      int *any_colors_out);
/* compileGeometry (compile.ts:327-403): flattens the meshes on the device, builds the BVH there
 * (as mcr_build_bvh), stores every array in BVH order with tangents (writeTangent, :406-432) and
 * the epsilon. Replaces the main-thread compile (compile.ts:361 calls buildBvh inside it). */
int mcr_compile_geometry(mcr_ctx *ctx, const mcr_mesh *meshes, uint32_t mesh_count, mcr_geometry_out *out);

typedef struct mcr_compile_info {
  uint32_t triangle_count, node_count;
  uint32_t host_layout;  /* 1: leaves over 31 triangles (buildBvh's depth limit) made the scene fall
                            back to the layout mcr_scene_set_geometry builds from host arrays */
  double epsilon;
} mcr_compile_info;
/* compileGeometry straight into a resident scene (compile.ts:327-432 followed by the worker's
 * stored geometry, protocol.ts:64-67): the same flatten, build and BVH-order arrays as
 * mcr_compile_geometry, laid out as the scene's device records without leaving the device, then
 * stored under `version` as mcr_scene_set_geometry would. Renders are identical to those of a scene
 * given the mcr_compile_geometry arrays. info may be NULL. */
int mcr_scene_compile_geometry(mcr_scene *scene, const mcr_mesh *meshes, uint32_t mesh_count, uint32_t version,
                               mcr_compile_info *info);

/* ---- device buffers, for hosts without a GPU allocator of their own ----------------------- */
int mcr_device_alloc(mcr_ctx *ctx, uint64_t bytes, void **out);
int mcr_device_free(mcr_ctx *ctx, void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. Synchronous. */
int mcr_copy(mcr_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind);
int mcr_memset(mcr_ctx *ctx, void *dst, int value, uint64_t bytes);

/* ---- timing: device time of the last call's kernels, by kernel family (milliseconds) ------ */
typedef struct mcr_timing {
  double closest_ms, any_ms, shade_ms, other_ms, total_ms;
  uint64_t closest_launches, any_launches;
} mcr_timing;
int mcr_last_timing(mcr_ctx *ctx, mcr_timing *out);
int mcr_set_timing(mcr_ctx *ctx, int enabled);
/* Path slots of the wavefront (samples in flight). Default 2^23 (17.5 GB); at least 256, at most
 * 2^26; a run allocates no more slots than it has samples. Samples beyond the
 * capacity start as earlier ones finish, so the capacity changes memory use, not results. */
int mcr_set_wavefront_capacity(mcr_ctx *ctx, uint32_t slots);
/* Replays bounce iterations from a captured HIP graph (default on; off launches each kernel with
 * grids sized to the live paths). Timing mode always launches directly. Results are the same. */
int mcr_set_graph(mcr_ctx *ctx, int enabled);
/* Material-sorted shading (default on): before each shade launch the live paths are binned by the
 * branch their vertex takes (a miss, the material model, maps, transmission, see-through), so the
 * lanes of a wave shade alike, and misses are shaded by a light kernel of their own. Each sample
 * keeps its own random stream, so results are the same either way. No counterpart in the
 * reference (tracer.ts shades one sample at a time). */
int mcr_set_shade_sort(mcr_ctx *ctx, int enabled);

#ifdef __cplusplus
}
#endif

#endif /* MCR_H */
