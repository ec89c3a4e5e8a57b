// context.h — host-side state shared by the C-ABI entry points: the error slot, the wavefront
// engine (path state, views, jobs, scratch, captured iterations) and the context / scene objects.
#ifndef MCR_CONTEXT_H
#define MCR_CONTEXT_H

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcr.h"
#include "builder.h"
#include "device.h"
#include "wavefront.h"
#include "wide.h"

// Records msg as this thread's last error (mcr_last_error) and returns code.
int mcr_fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return mcr_fail(MCR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define MCR_TRY(expr)      \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

template <class T>
hipError_t dalloc(T **p, size_t count) {
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}

// Grid of persistent traversal blocks: enough waves to fill 256 CUs several times over.
const uint32_t TRAVERSE_GRID = 256 * 5;
const uint32_t TRAVERSE_THREADS = TRAVERSE_GRID * 256;
// Path slots (337 bytes each plus 109 bytes per light slot: 2^23 slots of a five-light scene hold
// 7.4 GB). More slots keep more waves per kernel busy: a 1080p, 64 spp render of 1.38M triangles
// runs 1,604, 1,688, 1,798 and 1,812 Mrays/s at 2M, 3M, 8M and 12M slots.
const uint32_t DEFAULT_CAPACITY = 1u << 23;
// Light slots take 109 bytes each per path slot (unshadowed term, shadow ray and colour, status,
// queue entry); scenes with many lights get fewer path slots so that they stay within 24 GB.
const uint64_t LIGHT_SLOT_BYTES = 109;
const uint64_t LIGHT_SLOT_BUDGET = 24ull << 30;
// A server holds all its slots from the start and its latency grows with them: 2^21 by default.
const uint32_t SERVER_DEFAULT_SLOTS = 1u << 21;
const uint32_t SERVER_DEFAULT_PREVIEW_SLOTS = 1u << 18;
const uint32_t SERVER_DEFAULT_PASS_SLOTS = 1u << 22;

// One wavefront engine: the path slots and everything its kernels read and write. A context owns
// one for its synchronous calls; a server owns another for its request loop.
struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Wavefront w = {};
  void *w_block = nullptr;
  uint32_t capacity = 0;
  uint32_t lights = 0;             // light slots per path slot of the allocated wavefront
  uint32_t capacity_limit = DEFAULT_CAPACITY;
  int32_t *spill = nullptr;        // traversal stack overflow, per traversal thread
  double *seen = nullptr;          // see-through shadow hits, per traversal thread
  int2 *seen_meta = nullptr;
  double2 *seen_uv = nullptr;      // their barycentrics, for textured see-through materials
  DevCounters *counters = nullptr;
  uint32_t *fetch = nullptr;       // work counters of the ray-query kernels (WaveFetch)
  ViewDev *views = nullptr;
  uint32_t views_cap = 0;
  JobDev *jobs = nullptr;
  uint32_t jobs_cap = 0;
  uint32_t *job_done = nullptr;
  uint32_t job_done_cap = 0;
  Control *ctl = nullptr;          // device
  Control *ctl_host = nullptr;     // pinned staging
  uint32_t *counts_host = nullptr; // pinned
  bool mix = false;                // the run holds more than one view
  bool wide = false;               // the run traces with the precision form's FP32 wide BVH
  bool tex = false;                // a view of the run has textured see-through materials
  // Two captured bounce iterations (queue 0 then queue 1), replayed while paths are live; one
  // graph per kernel variant (single or mixed views, FP64 binary or FP32 wide traversal).
  bool use_graph = true;
  // Bins the live queue by shade class (miss, material model, transmissive, see-through) before
  // shade, so a wave's lanes take the same branch and misses are shaded by a light kernel of
  // their own; results do not depend on it.
  bool sort_shade = true;
  hipStream_t capture_stream = nullptr;  // captures the graphs; lives as long as the engine
  hipGraphExec_t graph[8] = {};
  const void *graph_key[8][4] = {};
  // timing
  bool timing = false;
  mcr_timing last_timing = {};
  hipEvent_t ev[8] = {};
};

int engine_init(Engine &e, int device, hipStream_t stream);
void engine_free(Engine &e);
// Path slots, light slots, views and jobs the next run needs (buffers only grow; the capacity is
// clamped to capacity_limit and to the light-slot budget).
int engine_reserve(Engine &e, uint64_t samples, uint32_t nviews, uint32_t njobs, bool track_done, uint32_t lights);
// The path slots a wavefront with `lights` light slots per path slot may hold.
uint32_t engine_capacity_for(const Engine &e, uint32_t lights);
// Host arrays must stay alive and unchanged until the next engine_sync.
int engine_upload_views(Engine &e, const ViewDev *views, uint32_t n);
int engine_upload_jobs(Engine &e, const JobDev *jobs, uint32_t first, uint32_t n);
// The caller must have waited for the stream since the previous call (the staging is reused).
int engine_set_control(Engine &e, float *out, uint32_t njobs, uint32_t total, bool track_done);
int engine_zero_done(Engine &e, uint32_t first, uint32_t n);
// Every slot free, every counter zero.
int engine_reset(Engine &e);
// Starts `count` samples from `next` in free slots; see k_refill.
int engine_refill(Engine &e, int q, uint32_t count, uint32_t free_top, uint32_t next, uint32_t queued);
// Puts the live slots of queue q (at most `live`) in ascending slot order; between iterations only.
int engine_sort_queue(Engine &e, int q, uint32_t live);
// One bounce iteration over the `live` paths of queue q (grids sized from live).
int engine_iteration(Engine &e, int q, uint32_t live, uint32_t max_lights, double *t_closest,
                     double *t_shade, double *t_any);
// Two iterations, q = 0 then 1, from the captured graph (grids sized for every slot).
int engine_iteration_pair(Engine &e, uint32_t max_lights);
void engine_drop_graphs(Engine &e);
// Copies the counter words to counts_host and waits for the stream.
int engine_sync(Engine &e);
// The copy alone (the caller waits), and the overflow check on the copied words.
int engine_fetch_counts(Engine &e);
int engine_check(Engine &e);
// A whole call: every job's samples traced to completion, results in out.
int engine_run(Engine &e, const std::vector<ViewDev> &views, const std::vector<JobDev> &jobs,
               uint32_t total, float *out, mcr_counters *counters);
void engine_add_counters(const DevCounters &d, uint64_t paths, mcr_counters *out);
// The MCR_COUNTER_COPIES device copies of the counters: bytes, clear, and their sum on the host.
const size_t COUNTER_BYTES = (size_t)MCR_COUNTER_COPIES * MCR_COUNTER_STRIDE;
const size_t FETCH_BYTES = sizeof(uint32_t) * MCR_FETCH_COUNTERS * MCR_COUNT_STRIDE;
hipError_t counters_clear(DevCounters *d, hipStream_t s);
hipError_t counters_read(const DevCounters *d, DevCounters &sum);
DevCounters counters_sum(const void *copies);

struct mcr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Engine eng;
  float *out_tmp = nullptr;
  size_t out_tmp_cap = 0;
  BuilderWorkspace builder;
  int intersection = MCR_ISECT_FP64;  // mcr_set_intersection
  std::mutex mu;
};

struct mcr_scene {
  mcr_ctx *ctx = nullptr;
  // Servers holding the scene (mcr_server_add_scene): while any does, its device arrays are in use
  // by requests it may admit, so geometry, shading and wide-BVH changes are refused and a release
  // is deferred until the last server lets the scene go.
  int server_refs = 0;
  bool release_pending = false;
  DevScene d = {};
  uint32_t geometry_version = 0, shading_version = 0;
  bool has_geometry = false, has_shading = false;
  float4 *tri = nullptr;
  float4 *inner = nullptr;
  int2 *longleaf = nullptr;
  float *normals = nullptr;
  float *colors = nullptr;
  float *tangents = nullptr;
  float *uvs = nullptr;
  void *texels = nullptr;              // every texture's texels, one allocation
  DevTexture *textures = nullptr;
  DevEnvironment *environments = nullptr;
  DevMaterial *materials = nullptr;
  DevLight *lights = nullptr;
  uint64_t bytes_geometry = 0, bytes_shading = 0;
  // precision form (wide.hip)
  bool has_wide = false;
  bool shadow_tex = false;         // a material with map or alphaMap: shadow rays sample textures
  uint4 *wide_nodes = nullptr;
  float4 *wide_tri = nullptr;
  WideScene wide = {};
  mcr_wide_info wide_info = {};
};

// Server holds on a scene (under one library-wide lock). scene_held: refused changes report it.
void scene_hold(mcr_scene *s);
// Drops one hold; frees the scene if it was released while held and this was the last hold.
void scene_drop(mcr_scene *s);
bool scene_held(const mcr_scene *s);
int scene_held_error();

// Frees the scene's wide BVH (wide.hip).
void wide_free(mcr_scene *s);
void free_geometry(mcr_scene *s);

// The view of (scene, camera, settings) as the kernels read it.
ViewDev make_view(const mcr_scene *s, const mcr_camera &cam, const mcr_settings &st);
// Checks a tile job and lays it out at sample `start`; returns its sample count or -1 (error set).
int64_t layout_job(const mcr_tile_job &j, uint32_t index, uint32_t view, uint64_t start, JobDev &out);

#endif  // MCR_CONTEXT_H
