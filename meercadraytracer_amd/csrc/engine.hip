// engine.hip — the host side of the wavefront: buffers, the bounce iteration (direct launches or a
// captured graph of two iterations), slot refills and the synchronous whole-call loop.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>

#include "context.h"
#include "kernels.h"

namespace {

uint32_t blocks_of(uint64_t n) { return (uint32_t)((n + 255) / 256); }

int alloc_wavefront(Engine &e, uint32_t capacity, uint32_t lights) {
  hipFree(e.w_block);
  e.w_block = nullptr;
  e.capacity = 0;
  e.lights = 0;
  engine_drop_graphs(e);
  size_t P = capacity;
  size_t L = (size_t)lights * P;
  uint32_t regions = std::min<uint32_t>(lights, MCR_SHADOW_REGIONS);
  uint32_t region_cap = (uint32_t)(((lights + MCR_SHADOW_REGIONS - 1) / MCR_SHADOW_REGIONS) * P);
  size_t bytes = 0;
  auto take = [&](size_t n) {
    size_t off = (bytes + 255) / 256 * 256;
    bytes = off + n;
    return off;
  };
  size_t o_f = take(sizeof(double) * F_COUNT * P);
  size_t o_i = take(sizeof(int32_t) * I_COUNT * P);
  size_t o_mode = take(P);
  size_t o_contrib = take(sizeof(double) * 3 * L);
  size_t o_shadow = take(sizeof(double) * 3 * L);
  size_t o_sray = take(sizeof(double) * 7 * L);
  size_t o_status = take(L);
  size_t o_q0 = take(sizeof(int32_t) * P);
  size_t o_q1 = take(sizeof(int32_t) * P);
  size_t o_counts = take(sizeof(uint32_t) * COUNT_WORDS_ALLOC);
  size_t o_sref = take(sizeof(uint32_t) * regions * (size_t)region_cap);
  size_t o_free = take(sizeof(int32_t) * P);
  size_t o_deep = take(sizeof(double) * (size_t)MCR_DEEP_HITS * MCR_DEEP_THREADS);
  size_t o_bins = take(sizeof(int32_t) * MCR_SHADE_CLASSES * P);
  HIP_TRY(hipMalloc(&e.w_block, bytes));
  char *b = (char *)e.w_block;
  Wavefront &w = e.w;
  w.capacity = capacity;
  w.lights = lights;
  w.region_cap = region_cap;
  w.f = (double *)(b + o_f);
  w.i = (int32_t *)(b + o_i);
  w.mode = (uint8_t *)(b + o_mode);
  w.contrib = (double *)(b + o_contrib);
  w.shadow = (double *)(b + o_shadow);
  w.sray = (double *)(b + o_sray);
  w.status = (uint8_t *)(b + o_status);
  w.queue[0] = (int32_t *)(b + o_q0);
  w.queue[1] = (int32_t *)(b + o_q1);
  w.counts = (uint32_t *)(b + o_counts);
  w.sref = (uint32_t *)(b + o_sref);
  w.free_list = (int32_t *)(b + o_free);
  w.deep_seen = (double *)(b + o_deep);
  w.bins = (int32_t *)(b + o_bins);
  w.ctl = e.ctl;
  e.capacity = capacity;
  e.lights = lights;
  return MCR_OK;
}

template <class T>
int grow(T *&p, uint32_t &cap, uint32_t n) {
  if (cap >= n) return MCR_OK;
  uint32_t want = std::max<uint32_t>(n, cap + cap / 2);
  hipFree(p);
  p = nullptr;
  cap = 0;
  HIP_TRY(dalloc(&p, want));
  cap = want;
  return MCR_OK;
}

}  // namespace

void engine_drop_graphs(Engine &e) {
  // A launch of the graph may still be running.
  bool any = false;
  for (auto &g : e.graph) any |= g != nullptr;
  if (e.stream && any) hipStreamSynchronize(e.stream);
  for (auto &g : e.graph) {
    if (g) hipGraphExecDestroy(g);
    g = nullptr;
  }
}

int engine_init(Engine &e, int device, hipStream_t stream) {
  e.device = device;
  e.stream = stream;
  HIP_TRY(hipMalloc(&e.spill, sizeof(int32_t) * (size_t)(MCR_STACK_MAX - MCR_STACK_LDS) * TRAVERSE_THREADS));
  HIP_TRY(hipMalloc(&e.seen, sizeof(double) * (size_t)MCR_SHADOW_HIT_CAP * TRAVERSE_THREADS));
  HIP_TRY(hipMalloc(&e.seen_meta, sizeof(int2) * (size_t)MCR_SHADOW_HIT_CAP * TRAVERSE_THREADS));
  HIP_TRY(hipMalloc(&e.seen_uv, sizeof(double2) * (size_t)MCR_SHADOW_HIT_CAP * TRAVERSE_THREADS));
  HIP_TRY(hipMalloc(&e.counters, COUNTER_BYTES));
  HIP_TRY(hipMalloc(&e.fetch, FETCH_BYTES));
  HIP_TRY(hipMalloc(&e.ctl, sizeof(Control)));
  HIP_TRY(hipHostMalloc(&e.ctl_host, sizeof(Control), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(&e.counts_host, sizeof(uint32_t) * COUNT_WORDS_ALLOC, hipHostMallocDefault));
  memset(e.ctl_host, 0, sizeof(Control));
  memset(e.counts_host, 0, sizeof(uint32_t) * COUNT_WORDS_ALLOC);
  for (auto &ev : e.ev) HIP_TRY(hipEventCreate(&ev));
  return MCR_OK;
}

void engine_free(Engine &e) {
  if (e.stream) hipStreamSynchronize(e.stream);
  engine_drop_graphs(e);
  if (e.capture_stream) hipStreamDestroy(e.capture_stream);
  hipFree(e.w_block);
  hipFree(e.spill);
  hipFree(e.seen);
  hipFree(e.seen_meta);
  hipFree(e.seen_uv);
  hipFree(e.counters);
  hipFree(e.fetch);
  hipFree(e.views);
  hipFree(e.jobs);
  hipFree(e.job_done);
  hipFree(e.ctl);
  if (e.ctl_host) hipHostFree(e.ctl_host);
  if (e.counts_host) hipHostFree(e.counts_host);
  for (auto &ev : e.ev)
    if (ev) hipEventDestroy(ev);
  e = Engine();
}

uint32_t engine_capacity_for(const Engine &e, uint32_t lights) {
  uint64_t by_lights = LIGHT_SLOT_BUDGET / ((uint64_t)std::max<uint32_t>(lights, 1) * LIGHT_SLOT_BYTES);
  return (uint32_t)std::max<uint64_t>(256, std::min<uint64_t>(e.capacity_limit, by_lights));
}

int engine_reserve(Engine &e, uint64_t samples, uint32_t nviews, uint32_t njobs, bool track_done, uint32_t lights) {
  lights = std::max<uint32_t>(lights, 1);
  uint32_t limit = engine_capacity_for(e, std::max(lights, e.lights));
  uint32_t cap = (uint32_t)std::min<uint64_t>(samples, limit);
  cap = std::max<uint32_t>(cap, 256);
  cap = std::min(std::max(cap, e.capacity), limit);
  if (e.capacity < cap || e.lights < lights) MCR_TRY(alloc_wavefront(e, cap, std::max(lights, e.lights)));
  const void *old_views = e.views;
  MCR_TRY(grow(e.views, e.views_cap, std::max<uint32_t>(nviews, 1)));
  if (old_views != e.views) engine_drop_graphs(e);
  MCR_TRY(grow(e.jobs, e.jobs_cap, std::max<uint32_t>(njobs, 1)));
  if (track_done) MCR_TRY(grow(e.job_done, e.job_done_cap, std::max<uint32_t>(njobs, 1)));
  return MCR_OK;
}

int engine_upload_views(Engine &e, const ViewDev *views, uint32_t n) {
  if (n) HIP_TRY(hipMemcpyAsync(e.views, views, n * sizeof(ViewDev), hipMemcpyHostToDevice, e.stream));
  return MCR_OK;
}

int engine_upload_jobs(Engine &e, const JobDev *jobs, uint32_t first, uint32_t n) {
  if (n)
    HIP_TRY(hipMemcpyAsync(e.jobs + first, jobs, n * sizeof(JobDev), hipMemcpyHostToDevice, e.stream));
  return MCR_OK;
}

int engine_set_control(Engine &e, float *out, uint32_t njobs, uint32_t total, bool track_done) {
  // Callers have waited for the stream since the last call, so the staging struct is free.
  *e.ctl_host = Control{e.jobs, out, track_done ? e.job_done : nullptr, njobs, total};
  HIP_TRY(hipMemcpyAsync(e.ctl, e.ctl_host, sizeof(Control), hipMemcpyHostToDevice, e.stream));
  return MCR_OK;
}

int engine_zero_done(Engine &e, uint32_t first, uint32_t n) {
  if (n) HIP_TRY(hipMemsetAsync(e.job_done + first, 0, n * sizeof(uint32_t), e.stream));
  return MCR_OK;
}

int engine_reset(Engine &e) {
  launch_reset_slots(e.stream, e.w);
  HIP_TRY(hipGetLastError());
  return MCR_OK;
}

int engine_refill(Engine &e, int q, uint32_t count, uint32_t free_top, uint32_t next, uint32_t queued) {
  launch_refill(e.stream, e.mix, e.views, e.w, q, count, free_top, next, queued);
  HIP_TRY(hipGetLastError());
  return MCR_OK;
}

int engine_sort_queue(Engine &e, int q, uint32_t live) {
  launch_sort_queue(e.stream, e.w, q, live);
  HIP_TRY(hipGetLastError());
  return MCR_OK;
}

int engine_iteration(Engine &e, int q, uint32_t live, uint32_t max_lights, double *t_closest,
                     double *t_shade, double *t_any) {
  hipStream_t sm = e.stream;
  const Wavefront &w = e.w;
  bool timed = e.timing && t_closest;
  if (timed) HIP_TRY(hipEventRecord(e.ev[0], sm));
  launch_extend(sm, e.mix, e.wide, e.views, w, q, std::min<uint32_t>(TRAVERSE_GRID, blocks_of(live)), e.spill,
                TRAVERSE_THREADS, e.counters);
  if (timed) HIP_TRY(hipEventRecord(e.ev[1], sm));
  if (e.sort_shade) launch_shade_bin(sm, e.mix, e.views, w, q, live, e.counters);
  launch_shade(sm, e.mix, e.sort_shade, e.views, w, q, live, e.counters);
  if (timed) HIP_TRY(hipEventRecord(e.ev[2], sm));
  uint64_t max_shadow = (uint64_t)live * max_lights;
  if (max_shadow) {
    launch_shadow(sm, e.mix, e.wide, e.tex, e.views, w, std::min<uint32_t>(TRAVERSE_GRID, blocks_of(max_shadow)), e.spill,
                  TRAVERSE_THREADS, e.seen, e.seen_meta, e.seen_uv, e.counters);
    launch_shadow_deep(sm, e.mix, e.tex, e.views, w, e.spill, TRAVERSE_THREADS, e.counters);
  }
  if (timed) HIP_TRY(hipEventRecord(e.ev[3], sm));
  launch_resolve(sm, e.mix, e.views, w, q, live);
  HIP_TRY(hipGetLastError());
  if (timed) {
    HIP_TRY(hipEventRecord(e.ev[4], sm));
    HIP_TRY(hipEventSynchronize(e.ev[4]));
    float a, b, d;
    hipEventElapsedTime(&a, e.ev[0], e.ev[1]);
    hipEventElapsedTime(&b, e.ev[1], e.ev[2]);
    hipEventElapsedTime(&d, e.ev[2], e.ev[3]);
    *t_closest += a;
    *t_shade += b;
    *t_any += d;
  }
  return MCR_OK;
}

int engine_iteration_pair(Engine &e, uint32_t max_lights) {
  int v = (e.mix ? 1 : 0) + (e.wide ? 2 : 0) + (e.tex ? 4 : 0);
  const void *key[4] = {e.w_block, e.views, e.counters,
                        (const void *)(uintptr_t)(((uint64_t)e.capacity * 64u + max_lights) * 2u + e.sort_shade)};
  hipGraphExec_t &graph = e.graph[v];
  if (graph && memcmp(key, e.graph_key[v], sizeof(key)) != 0) {
    HIP_TRY(hipStreamSynchronize(e.stream));  // a launch of the old graph may still be running
    hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  if (!graph) {
    // Grids for every slot: a replay does not know how many paths are live. Blocks past the live
    // count return at once, and traversal waves leave as soon as the shared counter runs out.
    if (!e.capture_stream) HIP_TRY(hipStreamCreateWithFlags(&e.capture_stream, hipStreamNonBlocking));
    hipStream_t cs = e.capture_stream;
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (err == hipSuccess) {
      uint32_t cap = e.capacity;
      for (int q = 0; q < 2; q++) {
        launch_extend(cs, e.mix, e.wide, e.views, e.w, q, std::min<uint32_t>(TRAVERSE_GRID, blocks_of(cap)), e.spill,
                      TRAVERSE_THREADS, e.counters);
        if (e.sort_shade) launch_shade_bin(cs, e.mix, e.views, e.w, q, cap, e.counters);
        launch_shade(cs, e.mix, e.sort_shade, e.views, e.w, q, cap, e.counters);
        if (max_lights) {
          launch_shadow(cs, e.mix, e.wide, e.tex, e.views, e.w, std::min<uint32_t>(TRAVERSE_GRID, blocks_of((uint64_t)cap * max_lights)),
                        e.spill, TRAVERSE_THREADS, e.seen, e.seen_meta, e.seen_uv, e.counters);
          launch_shadow_deep(cs, e.mix, e.tex, e.views, e.w, e.spill, TRAVERSE_THREADS, e.counters);
        }
        launch_resolve(cs, e.mix, e.views, e.w, q, cap);
      }
      err = hipStreamEndCapture(cs, &g);
    }
    if (err == hipSuccess) err = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    if (err != hipSuccess) {
      graph = nullptr;
      return mcr_fail(MCR_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(err));
    }
    memcpy(e.graph_key[v], key, sizeof(key));
  }
  HIP_TRY(hipGraphLaunch(graph, e.stream));
  return MCR_OK;
}

int engine_fetch_counts(Engine &e) {
  HIP_TRY(hipMemcpyAsync(e.counts_host, e.w.counts, sizeof(uint32_t) * COUNT_WORDS_ALLOC, hipMemcpyDeviceToHost,
                         e.stream));
  return MCR_OK;
}

int engine_check(Engine &e) {
  uint32_t o = e.counts_host[cw(C_OVERFLOW)];
  if (o & OVERFLOW_WIDE)
    return mcr_fail(MCR_E_INTERNAL, "a wide-BVH traversal met a corrupt node or overflowed its stack");
  if (o)
    return mcr_fail(MCR_E_OVERFLOW, "a shadow ray crossed more than " + std::to_string(MCR_DEEP_HITS) +
                                        " counted see-through surfaces");
  return MCR_OK;
}

int engine_sync(Engine &e) {
  MCR_TRY(engine_fetch_counts(e));
  HIP_TRY(hipStreamSynchronize(e.stream));
  return engine_check(e);
}

DevCounters counters_sum(const void *copies) {
  DevCounters s = {};
  const size_t words = sizeof(DevCounters) / sizeof(unsigned long long);
  unsigned long long *o = reinterpret_cast<unsigned long long *>(&s);
  for (int c = 0; c < MCR_COUNTER_COPIES; c++) {
    const unsigned long long *in =
        reinterpret_cast<const unsigned long long *>(static_cast<const char *>(copies) + (size_t)c * MCR_COUNTER_STRIDE);
    for (size_t k = 0; k < words; k++) o[k] += in[k];
  }
  return s;
}

hipError_t counters_clear(DevCounters *d, hipStream_t s) { return hipMemsetAsync(d, 0, COUNTER_BYTES, s); }

hipError_t counters_read(const DevCounters *d, DevCounters &sum) {
  std::vector<char> buf(COUNTER_BYTES);
  hipError_t err = hipMemcpy(buf.data(), d, COUNTER_BYTES, hipMemcpyDeviceToHost);
  if (err == hipSuccess) sum = counters_sum(buf.data());
  return err;
}

void engine_add_counters(const DevCounters &hc, uint64_t paths, mcr_counters *counters) {
  counters->closest_queries += hc.closest_queries;
  counters->any_queries += hc.any_queries;
  counters->node_visits += hc.node_visits;
  counters->leaf_visits += hc.leaf_visits;
  counters->triangle_tests += hc.triangle_tests;
  counters->paths += paths;
  counters->closest_node_visits += hc.closest_node_visits;
  counters->closest_triangle_tests += hc.closest_triangle_tests;
  counters->any_node_visits += hc.any_node_visits;
  counters->any_triangle_tests += hc.any_triangle_tests;
  counters->shade_waves += hc.shade_waves;
  counters->mixed_view_waves += hc.mixed_view_waves;
  for (int c = 0; c < 8; c++) counters->shade_class_lanes[c] += hc.shade_class_lanes[c];
  counters->class_mixed_waves += hc.class_mixed_waves;
}

int engine_run(Engine &e, const std::vector<ViewDev> &views, const std::vector<JobDev> &jobs,
               uint32_t total, float *out, mcr_counters *counters) {
  if (jobs.empty() || total == 0) return MCR_OK;
  uint32_t nviews = (uint32_t)views.size(), njobs = (uint32_t)jobs.size();
  uint32_t max_lights = 0;
  e.tex = false;
  for (const ViewDev &v : views) {
    max_lights = std::max(max_lights, v.shadow_slots);
    e.tex |= v.shadow_tex != 0;
  }
  MCR_TRY(engine_reserve(e, total, nviews, njobs, false, max_lights));
  e.mix = nviews > 1;
  hipStream_t sm = e.stream;
  MCR_TRY(engine_set_control(e, out, njobs, total, false));
  MCR_TRY(engine_upload_views(e, views.data(), nviews));
  MCR_TRY(engine_upload_jobs(e, jobs.data(), 0, njobs));
  HIP_TRY(counters_clear(e.counters, sm));
  if (e.timing) HIP_TRY(hipEventRecord(e.ev[6], sm));
  // Path regeneration: resolve starts the call's next sample in every slot whose sample finished,
  // so the wavefront stays full across passes and tiles and only the very end of the call runs
  // with few live paths.
  MCR_TRY(engine_reset(e));
  uint32_t first = std::min<uint32_t>(e.capacity, total);
  MCR_TRY(engine_refill(e, 0, first, e.capacity, 0, 0));
  double t_closest = 0, t_any = 0, t_shade = 0;
  uint64_t n_closest = 0, n_any = 0;
  uint32_t live = first;
  int q = 0;
  bool graph = e.use_graph && !e.timing;
  while (live > 0) {
    if (graph) {
      MCR_TRY(engine_iteration_pair(e, max_lights));
      MCR_TRY(engine_sync(e));
      live = e.counts_host[cw(C_QUEUE0)];
    } else {
      MCR_TRY(engine_iteration(e, q, live, max_lights, &t_closest, &t_shade, &t_any));
      n_closest++;
      if (max_lights) n_any++;
      MCR_TRY(engine_sync(e));
      live = e.counts_host[cw(q ^ 1)];
      q ^= 1;
    }
  }
  if (e.timing) {
    HIP_TRY(hipEventRecord(e.ev[7], sm));
    HIP_TRY(hipEventSynchronize(e.ev[7]));
    float tot;
    hipEventElapsedTime(&tot, e.ev[6], e.ev[7]);
    e.last_timing = mcr_timing{t_closest, t_any, t_shade, tot - t_closest - t_any - t_shade, tot, n_closest, n_any};
  }
  if (counters) {
    DevCounters hc;
    HIP_TRY(counters_read(e.counters, hc));
    engine_add_counters(hc, total, counters);
  }
  return MCR_OK;
}
