// server.hip — the served form: a request loop over scenes resident in HBM.
//
// Each request is one ProgressiveRenderer.render call (src/progressive.ts:194-275): its tiles
// (makeTiles, centre first), a work list of preview items then pass-major items, and an 8-bit
// image plus an FP32 running sum written tile by tile (writeTile, progressive.ts:288-313). A
// server thread owns its own wavefront engine and, every host step:
//   1. frees what released or cancelled requests held, admits queued requests while the HBM
//      budget and a request slot allow (first in, first out);
//   2. rebuilds the job window: jobs whose samples have started stay where they are; every
//      unstarted job is chosen afresh — the previews of newly admitted requests first, then the
//      passes of the oldest running request — so a new request's preview overtakes older passes,
//      the wavefront works on few scenes at a time, and a cancelled request's unstarted tiles
//      simply disappear;
//   3. refills idle path slots and runs bounce iterations (a captured pair, or one direct);
//   4. reads each job's finished-sample count, accumulates and encodes finished tiles in pass
//      order (so a finished request holds the bytes raytrace() gives), and stamps the time of
//      the first complete preview and of completion.
// Samples of every live request share the path slots, so requests that arrive while others run
// start as soon as slots free up (continuous batching), and a result does not depend on what it
// shared the device with: every sample carries its own seed and output position.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "context.h"
#include "kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Item {
  uint32_t tile, pass, block;
};

struct SrvTile {
  mcr_tile t;
  uint32_t samples = 0;             // passes accumulated (tileSamples, progressive.ts:203)
  bool covered = false;             // preview or a pass written
  std::map<uint32_t, uint64_t> ready;  // finished passes waiting for their turn: pass -> arena offset
  uint64_t preview = UINT64_MAX;    // arena offset of a finished preview not yet written
};

struct SrvRequest {
  uint64_t id = 0;
  mcr_request r = {};
  int state = MCR_REQ_QUEUED;
  int view = -1;
  uint8_t *image = nullptr;
  float *sum = nullptr;
  uint64_t bytes = 0;
  std::vector<SrvTile> tiles;
  std::deque<Item> pending;         // work items not in the job window, in work-list order
  uint32_t jobs_live = 0;           // jobs in the window or waiting to be encoded
  uint32_t covered = 0, finished_tiles = 0;
  bool cancel = false, release = false;
  double t_submit = -1, t_admit = -1, t_preview = -1, t_done = -1;
  uint64_t paths = 0;
  uint64_t it_admit = 0, it_preview = 0;  // bounce iterations run by the server so far
};

struct SrvJob {
  JobDev dev;
  uint64_t req;
  Item item;
  uint64_t arena;    // float offset of the tile result
  uint64_t floats;
  uint32_t size;     // samples
  uint32_t done = 0;
};

// First-fit allocator over the result arena (offsets and sizes in floats, multiples of 4).
struct Arena {
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void reset(uint64_t n) {
    free_.clear();
    free_[0] = n;
  }
  uint64_t alloc(uint64_t n) {
    n = (n + 3) & ~3ull;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      uint64_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      return off;
    }
    return UINT64_MAX;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 3) & ~3ull;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// makeTiles (progressive.ts:160-170): row-major tiles, stable-sorted by the distance of the tile
// centre to the image centre as Math.hypot gives it.
std::vector<mcr_tile> make_tiles(uint32_t width, uint32_t height, uint32_t size) {
  std::vector<mcr_tile> tiles;
  for (uint32_t y = 0; y < height; y += size)
    for (uint32_t x = 0; x < width; x += size)
      tiles.push_back(mcr_tile{(int32_t)x, (int32_t)y, (int32_t)std::min(size, width - x),
                               (int32_t)std::min(size, height - y)});
  double cx = width / 2.0, cy = height / 2.0;
  std::vector<double> d(tiles.size());
  for (size_t i = 0; i < tiles.size(); i++)
    d[i] = js_hypot2(tiles[i].x + tiles[i].width / 2.0 - cx, tiles[i].y + tiles[i].height / 2.0 - cy);
  std::vector<size_t> idx(tiles.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return d[a] < d[b]; });
  std::vector<mcr_tile> out;
  for (size_t i : idx) out.push_back(tiles[i]);
  return out;
}

}  // namespace

struct mcr_server {
  mcr_ctx *ctx = nullptr;
  hipStream_t stream = nullptr, copy_stream = nullptr;
  hipEvent_t encoded = nullptr;    // recorded after every encode launch
  Engine eng;
  mcr_server_config cfg = {};
  float *arena_dev = nullptr;
  uint64_t arena_floats = 0;
  Arena arena;
  std::vector<ViewDev> views;
  std::vector<int> view_free;
  bool views_dirty = true;
  std::unordered_map<uint64_t, std::unique_ptr<SrvRequest>> reqs;
  std::deque<uint64_t> admit_queue;
  std::vector<uint64_t> running;   // admitted, in admission order
  std::vector<SrvJob> table;       // the device job list, by sample_start
  uint32_t total = 0;              // samples [0, total) may be started
  int q = 0;                       // queue holding the live paths
  uint64_t bytes_used = 0;
  std::vector<mcr_scene *> scenes;  // held (scene_hold) until removed or the server is destroyed
  std::vector<uint64_t> scene_bytes;  // what each counts against the HBM budget
  EncodeItem *enc_host = nullptr, *enc_dev = nullptr;
  uint32_t enc_cap = 0;
  JobDev *jobs_host = nullptr;     // pinned staging of the job list
  uint32_t *done_host = nullptr;   // pinned
  uint32_t host_cap = 0;
  void *counters_host = nullptr;  // pinned: the device counter copies
  // stats
  uint64_t iterations = 0, paths = 0;
  double device_ms = 0;
  DevCounters counters = {};
  // thread
  std::thread thread;
  std::mutex mu;
  std::condition_variable wake, changed;
  bool stop = false;
  uint64_t next_id = 1;
  Clock::time_point t0;
  int fatal = MCR_OK;
  std::string fatal_msg;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

// Renders beyond this many samples keep their running sum in FP64.
const uint32_t WIDE_SUM_SAMPLES = 1024;

uint64_t sum_bytes(const mcr_request &r) {
  return (uint64_t)r.width * r.height * (r.samples > WIDE_SUM_SAMPLES ? 32 : 16);
}
uint64_t request_bytes(const mcr_request &r) { return (uint64_t)r.width * r.height * 4 + sum_bytes(r); }

// Frees a terminal request's device memory and slot once nothing of it is in flight.
void retire(mcr_server *s, SrvRequest &r) {
  if (r.jobs_live || r.view < 0) return;
  if (r.state != MCR_REQ_DONE && r.state != MCR_REQ_CANCELLED && r.state != MCR_REQ_FAILED) return;
  if (r.state == MCR_REQ_CANCELLED || r.state == MCR_REQ_FAILED) {
    hipFree(r.image);
    hipFree(r.sum);
    r.image = nullptr;
    r.sum = nullptr;
    s->bytes_used -= r.bytes;
  } else {
    hipFree(r.sum);  // the image stays readable until released
    r.sum = nullptr;
    s->bytes_used -= sum_bytes(r.r);
    r.bytes = (uint64_t)r.r.width * r.r.height * 4;
  }
  s->view_free.push_back(r.view);
  r.view = -1;
  s->running.erase(std::remove(s->running.begin(), s->running.end(), r.id), s->running.end());
}

int admit(mcr_server *s, SrvRequest &r) {
  uint64_t npx = (uint64_t)r.r.width * r.r.height;
  HIP_TRY(hipMalloc(&r.image, npx * 4));
  uint64_t sb = sum_bytes(r.r);
  hipError_t e = hipMalloc(&r.sum, sb);
  if (e != hipSuccess) {
    hipFree(r.image);
    r.image = nullptr;
    return mcr_fail(MCR_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  HIP_TRY(hipMemsetAsync(r.image, 0, npx * 4, s->stream));
  HIP_TRY(hipMemsetAsync(r.sum, 0, sb, s->stream));
  r.view = s->view_free.back();
  s->view_free.pop_back();
  s->views[r.view] = make_view(r.r.scene, r.r.camera, r.r.settings);
  s->views_dirty = true;
  s->bytes_used += r.bytes;
  auto tiles = make_tiles(r.r.width, r.r.height, r.r.tile_size);
  r.tiles.resize(tiles.size());
  for (size_t i = 0; i < tiles.size(); i++) r.tiles[i].t = tiles[i];
  // progressive.ts:207-210: preview items for every tile, then pass-major items.
  if (r.r.preview_block > 1)
    for (uint32_t i = 0; i < tiles.size(); i++) r.pending.push_back(Item{i, 0, r.r.preview_block});
  for (uint32_t p = 0; p < r.r.samples; p++)
    for (uint32_t i = 0; i < tiles.size(); i++) r.pending.push_back(Item{i, p, 1});
  r.state = MCR_REQ_RUNNING;
  r.t_admit = s->now();
  r.it_admit = s->iterations;
  s->running.push_back(r.id);
  return MCR_OK;
}

// Admissions, cancellations and releases (under the lock).
int housekeeping(mcr_server *s) {
  // Encodes of the last step may still read or write what is freed below.
  bool frees = false;
  for (auto &kv : s->reqs) {
    const SrvRequest &r = *kv.second;
    frees |= r.cancel || r.release ||
             (r.jobs_live == 0 && r.view >= 0 && r.state != MCR_REQ_RUNNING && r.state != MCR_REQ_QUEUED);
  }
  if (frees) HIP_TRY(hipStreamSynchronize(s->stream));
  for (auto it = s->reqs.begin(); it != s->reqs.end();) {
    SrvRequest &r = *it->second;
    if (r.cancel && (r.state == MCR_REQ_QUEUED || r.state == MCR_REQ_RUNNING)) {
      if (r.state == MCR_REQ_QUEUED)
        s->admit_queue.erase(std::remove(s->admit_queue.begin(), s->admit_queue.end(), r.id), s->admit_queue.end());
      r.state = MCR_REQ_CANCELLED;
      r.pending.clear();
      for (SrvTile &t : r.tiles) {
        uint64_t floats = (uint64_t)t.t.width * t.t.height * 4;
        for (auto &kv : t.ready) s->arena.release(kv.second, floats);
        t.ready.clear();
        if (t.preview != UINT64_MAX) s->arena.release(t.preview, floats);
        t.preview = UINT64_MAX;
      }
      r.t_done = s->now();
      s->changed.notify_all();
    }
    retire(s, r);
    if (r.release && r.view < 0 && r.jobs_live == 0) {
      hipFree(r.image);
      hipFree(r.sum);
      s->bytes_used -= r.image ? r.bytes : 0;
      it = s->reqs.erase(it);
      continue;
    }
    ++it;
  }
  while (!s->admit_queue.empty() && !s->view_free.empty()) {
    SrvRequest &r = *s->reqs[s->admit_queue.front()];
    if (s->bytes_used + r.bytes > s->cfg.hbm_budget) break;
    s->admit_queue.pop_front();
    MCR_TRY(admit(s, r));
  }
  return MCR_OK;
}

// Step 2: the job window. `next` is the first sample not yet started; `idle` slots are free now
// and take their samples from the window first, so the window holds that many samples more.
void plan_window(mcr_server *s, uint32_t next, uint32_t idle) {
  std::vector<SrvJob> keep;
  std::unordered_map<uint64_t, std::vector<Item>> returned;
  uint32_t base = next;
  for (SrvJob &j : s->table) {
    if (j.done == j.size) continue;  // finished: handled by deliver()
    if (j.dev.sample_start < next) {
      keep.push_back(j);
      base = std::max(base, j.dev.sample_start + j.size);
    } else {
      returned[j.req].push_back(j.item);
      s->arena.release(j.arena, j.floats);
      SrvRequest &r = *s->reqs[j.req];
      r.jobs_live--;
    }
  }
  for (auto &kv : returned) {
    SrvRequest &r = *s->reqs[kv.first];
    if (r.state != MCR_REQ_RUNNING) continue;
    for (auto it = kv.second.rbegin(); it != kv.second.rend(); ++it) r.pending.push_front(*it);
  }
  s->table.swap(keep);
  // A replayed pair runs two iterations before the next step, so it needs two windows.
  uint64_t target = (uint64_t)s->cfg.window * (s->eng.use_graph ? 2 : 1) + idle, queued = 0;
  uint32_t start = base;
  auto place = [&](SrvRequest &r) -> bool {
    Item it = r.pending.front();
    const mcr_tile &t = r.tiles[it.tile].t;
    // Sample indices stay below 2^31; past that the window drains and numbering restarts.
    if ((uint64_t)start + (uint64_t)t.width * t.height > (1ull << 31)) return false;
    uint64_t floats = (uint64_t)t.width * t.height * 4;
    uint64_t off = s->arena.alloc(floats);
    if (off == UINT64_MAX) return false;
    mcr_tile_job tj{t, r.r.width, r.r.height, it.pass, it.block, off};
    SrvJob j;
    int64_t n = layout_job(tj, 0, (uint32_t)r.view, start, j.dev);
    j.req = r.id;
    j.item = it;
    j.arena = off;
    j.floats = floats;
    j.size = (uint32_t)n;
    j.done = 0;
    s->table.push_back(j);
    r.pending.pop_front();
    r.jobs_live++;
    start += (uint32_t)n;
    queued += (uint64_t)n;
    return true;
  };
  bool room = true;
  // Previews of requests whose preview has not been queued completely go first.
  for (uint64_t id : s->running) {
    SrvRequest &r = *s->reqs[id];
    while (room && r.state == MCR_REQ_RUNNING && !r.pending.empty() && r.pending.front().block > 1)
      room = place(r);
  }
  // Then the passes of the oldest request first: the wavefront keeps working on few scenes at a
  // time (their BVHs stay in cache), and requests finish in arrival order.
  for (uint64_t id : s->running) {
    SrvRequest &r = *s->reqs[id];
    while (room && queued < target && r.state == MCR_REQ_RUNNING && !r.pending.empty()) room = place(r);
    if (!room || queued >= target) break;
  }
  s->total = start;
}

// Step 4: finished jobs go to their requests; tiles are written in pass order. Returns whether
// anything was written.
int deliver(mcr_server *s, bool &progress) {
  std::vector<SrvJob> still;
  for (SrvJob &j : s->table) {
    if (j.done != j.size) {
      still.push_back(j);
      continue;
    }
    progress = true;
    SrvRequest &r = *s->reqs[j.req];
    r.paths += j.size;
    s->paths += j.size;
    r.jobs_live--;
    SrvTile *t = r.state == MCR_REQ_RUNNING ? &r.tiles[j.item.tile] : nullptr;
    if (!t) s->arena.release(j.arena, j.floats);
    else if (j.item.block > 1) t->preview = j.arena;
    else t->ready[j.item.pass] = j.arena;
  }
  s->table.swap(still);
  // Rounds of one launch each; a tile is staged at most once per round, so no two items of a
  // launch write the same pixel.
  uint32_t nenc = 0, launched = 0;
  std::vector<std::pair<uint64_t, uint64_t>> to_free;  // arena ranges, freed after the launches
  for (;;) {
    uint32_t first = nenc;
    for (uint64_t id : s->running) {
      SrvRequest &r = *s->reqs[id];
      if (r.state != MCR_REQ_RUNNING) continue;
      for (uint32_t i = 0; i < r.tiles.size() && nenc < s->enc_cap; i++) {
        SrvTile &t = r.tiles[i];
        uint64_t floats = (uint64_t)t.t.width * t.t.height * 4;
        uint64_t data;
        bool preview = false;
        if (t.preview != UINT64_MAX) {
          // progressive.ts:257-258: a preview is written only over a tile with no samples yet.
          data = t.preview;
          t.preview = UINT64_MAX;
          to_free.emplace_back(data, floats);
          if (t.samples != 0 || t.ready.count(0)) continue;
          preview = true;
        } else {
          auto it = t.ready.find(t.samples);
          if (it == t.ready.end()) continue;
          data = it->second;
          t.ready.erase(it);
          to_free.emplace_back(data, floats);
          t.samples++;
          if (t.samples == r.r.samples) r.finished_tiles++;
        }
        bool wide = r.r.samples > WIDE_SUM_SAMPLES;
        s->enc_host[nenc++] = EncodeItem{r.image, preview || wide ? nullptr : r.sum,
                                         preview || !wide ? nullptr : reinterpret_cast<double *>(r.sum),
                                         s->arena_dev + data, r.r.width,
                                         t.t.x, t.t.y, t.t.width, t.t.height, preview ? 1u : t.samples,
                                         r.r.srgb, 0};
        if (!t.covered) {
          t.covered = true;
          r.covered++;
        }
      }
    }
    if (nenc == first) break;
    uint32_t blocks = 0;
    for (uint32_t k = first; k < nenc; k++) {
      s->enc_host[k].block_start = blocks;
      blocks += (uint32_t)(((uint64_t)s->enc_host[k].tw * s->enc_host[k].th + 255) / 256);
    }
    HIP_TRY(hipMemcpyAsync(s->enc_dev + first, s->enc_host + first, (nenc - first) * sizeof(EncodeItem),
                           hipMemcpyHostToDevice, s->stream));
    launch_encode_batch(s->stream, s->enc_dev + first, nenc - first, blocks);
    HIP_TRY(hipGetLastError());
    launched = nenc;
    progress = true;
    if (nenc == s->enc_cap) break;  // the rest waits for the next step
  }
  if (launched) HIP_TRY(hipEventRecord(s->encoded, s->stream));
  // Ranges are reused only by later kernels on this stream, after the encodes that read them.
  for (auto &f : to_free) s->arena.release(f.first, f.second);
  // Status: first complete preview, completion (stamped once the writes have finished).
  bool need_sync = false;
  for (uint64_t id : s->running) {
    SrvRequest &r = *s->reqs[id];
    if (r.state != MCR_REQ_RUNNING) continue;
    if ((r.t_preview < 0 && r.covered == r.tiles.size()) || r.finished_tiles == r.tiles.size()) need_sync = true;
  }
  if (need_sync) HIP_TRY(hipStreamSynchronize(s->stream));
  double now = s->now();
  std::vector<uint64_t> ids(s->running);
  for (uint64_t id : ids) {
    SrvRequest &r = *s->reqs[id];
    if (r.state != MCR_REQ_RUNNING) continue;
    if (r.t_preview < 0 && r.covered == r.tiles.size()) {
      r.t_preview = now;
      r.it_preview = s->iterations;
      s->changed.notify_all();
    }
    if (r.finished_tiles == r.tiles.size() && r.pending.empty()) {
      r.state = MCR_REQ_DONE;
      r.t_done = now;
      retire(s, r);
      s->changed.notify_all();
    }
  }
  return MCR_OK;
}

int ensure_host(mcr_server *s, uint32_t n) {
  if (s->host_cap >= n) return MCR_OK;
  uint32_t want = std::max<uint32_t>(n, s->host_cap * 2);
  if (s->jobs_host) hipHostFree(s->jobs_host);
  if (s->done_host) hipHostFree(s->done_host);
  s->jobs_host = nullptr;
  s->done_host = nullptr;
  s->host_cap = 0;
  HIP_TRY(hipHostMalloc(&s->jobs_host, want * sizeof(JobDev), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(&s->done_host, want * sizeof(uint32_t), hipHostMallocDefault));
  s->host_cap = want;
  return MCR_OK;
}

// One host step with the lock held across planning and delivery but not across device waits.
int step(mcr_server *s, std::unique_lock<std::mutex> &lk, bool &progress) {
  Engine &e = s->eng;
  MCR_TRY(housekeeping(s));
  // With no path live every started sample has finished and no job in the table has started,
  // so sample numbering starts again from 0.
  uint32_t next = e.counts_host[cw(s->q)] == 0 ? 0 : std::min(e.counts_host[cw(C_NEXT)], s->total);
  plan_window(s, next, e.counts_host[cw(C_FREE)]);
  uint32_t n = (uint32_t)s->table.size();
  if (n == 0 && e.counts_host[cw(s->q)] == 0) return deliver(s, progress);
  progress = true;
  MCR_TRY(engine_reserve(e, e.capacity_limit, (uint32_t)s->views.size(), n, true, s->cfg.max_lights));
  MCR_TRY(ensure_host(s, std::max<uint32_t>(n, 1)));
  bool mixed = false;
  for (const SrvJob &j : s->table) mixed |= j.dev.view != 0;
  // Samples of other views may still be in flight with no job left in the table.
  mixed |= s->running.size() > 1;
  e.mix = mixed;
  // The previous step waited for its copies, so the pinned staging is free.
  for (uint32_t i = 0; i < n; i++) {
    s->jobs_host[i] = s->table[i].dev;
    s->done_host[i] = s->table[i].done;
  }
  if (s->views_dirty) {
    MCR_TRY(engine_upload_views(e, s->views.data(), (uint32_t)s->views.size()));
    s->views_dirty = false;
  }
  if (n) {
    HIP_TRY(hipMemcpyAsync(e.jobs, s->jobs_host, n * sizeof(JobDev), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(e.job_done, s->done_host, n * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
  }
  MCR_TRY(engine_set_control(e, s->arena_dev, n, s->total, true));
  uint32_t free_slots = e.counts_host[cw(C_FREE)], live = e.counts_host[cw(s->q)];
  uint32_t k = std::min<uint32_t>(free_slots, s->total - next);
  MCR_TRY(engine_refill(e, s->q, k, free_slots, next, live));
  live += k;
  // Refilled slots come from the free list in the order they went idle (after a drain, every
  // slot does), and in-place regeneration would keep them scattered for the rest of the run;
  // re-sorting the queue after a refill puts every wave back on neighbouring slots (coalesced
  // path-state accesses). Steps without a refill keep the queue's own order: waves requeue in
  // the order they finish, which groups paths of similar cost.
  if (k) MCR_TRY(engine_sort_queue(e, s->q, live));
  uint32_t max_lights = 0;
  e.tex = false;
  for (const ViewDev &v : s->views) {
    max_lights = std::max(max_lights, v.shadow_slots);
    e.tex |= v.shadow_tex != 0;
  }
  lk.unlock();
  auto t_dev = Clock::now();
  int rc = MCR_OK;
  if (live) {
    if (e.use_graph) {
      if (s->q != 0) {  // the captured pair starts on queue 0
        rc = engine_iteration(e, s->q, live, max_lights, nullptr, nullptr, nullptr);
        s->q ^= 1;
        s->iterations++;
      }
      if (!rc) rc = engine_iteration_pair(e, max_lights);
      s->iterations += 2;
    } else {
      rc = engine_iteration(e, s->q, live, max_lights, nullptr, nullptr, nullptr);
      s->q ^= 1;
      s->iterations++;
    }
  }
  // One wait per step: counter words, finished samples per job and ray counters together.
  if (!rc) rc = engine_fetch_counts(e);
  if (!rc) {
    hipError_t err = hipSuccess;
    if (n) err = hipMemcpyAsync(s->done_host, e.job_done, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
    if (err == hipSuccess) err = hipMemcpyAsync(s->counters_host, e.counters, COUNTER_BYTES, hipMemcpyDeviceToHost, s->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(s->stream);
    if (err != hipSuccess) rc = mcr_fail(MCR_E_DEVICE, std::string("job counts: ") + hipGetErrorString(err));
  }
  if (!rc) rc = engine_check(e);
  if (!rc) s->counters = counters_sum(s->counters_host);
  lk.lock();
  s->device_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_dev).count();
  if (rc) return rc;
  for (uint32_t i = 0; i < n; i++) s->table[i].done = s->done_host[i];
  return deliver(s, progress);
}

void run(mcr_server *s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    Engine &e = s->eng;
    bool busy = !s->running.empty() || !s->admit_queue.empty() || !s->table.empty() ||
                e.counts_host[cw(s->q)] != 0;
    bool chores = false;
    for (auto &kv : s->reqs) chores |= kv.second->cancel || kv.second->release;
    if (!busy && !chores) {
      s->wake.wait(lk);
      continue;
    }
    bool progress = false;
    int rc = step(s, lk, progress);
    // Nothing to trace (e.g. every queued request waits for memory): sleep until woken.
    if (!rc && !progress) s->wake.wait_for(lk, std::chrono::milliseconds(2));
    if (rc) {
      // A device failure ends every request; the server keeps answering status calls.
      s->fatal = rc;
      s->fatal_msg = mcr_last_error();
      for (auto &kv : s->reqs) {
        SrvRequest &r = *kv.second;
        if (r.state == MCR_REQ_QUEUED || r.state == MCR_REQ_RUNNING) {
          r.state = MCR_REQ_FAILED;
          r.t_done = s->now();
        }
      }
      s->admit_queue.clear();
      s->running.clear();
      s->table.clear();
      s->changed.notify_all();
      s->wake.wait(lk, [&] { return s->stop; });
    }
  }
}

}  // namespace

extern "C" {

int mcr_server_create(mcr_ctx *c, const mcr_server_config *config, mcr_server **out) {
  if (!c || !out) return mcr_fail(MCR_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  std::unique_ptr<mcr_server> s(new mcr_server());
  s->ctx = c;
  if (config) s->cfg = *config;
  else s->cfg.use_graph = 1;
  if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 200ull << 30;
  if (!s->cfg.slots) s->cfg.slots = SERVER_DEFAULT_SLOTS;
  if (s->cfg.slots < 256 || s->cfg.slots > (1u << 26)) return mcr_fail(MCR_E_INVALID, "slots must be in [256, 2^26]");
  if (!s->cfg.window) s->cfg.window = s->cfg.slots;
  if (!s->cfg.max_requests) s->cfg.max_requests = 64;
  if (!s->cfg.max_lights) s->cfg.max_lights = 16;
  if (s->cfg.max_lights > MCR_MAX_LIGHTS) return mcr_fail(MCR_E_INVALID, "max_lights beyond MCR_MAX_LIGHTS_PER_SCENE");
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&s->encoded, hipEventDisableTiming));
  MCR_TRY(engine_init(s->eng, c->device, s->stream));
  s->eng.capacity_limit = s->cfg.slots;
  s->eng.use_graph = s->cfg.use_graph != 0;
  s->views.assign(s->cfg.max_requests, ViewDev());
  memset(s->views.data(), 0, s->views.size() * sizeof(ViewDev));
  for (int i = (int)s->cfg.max_requests - 1; i >= 0; i--) s->view_free.push_back(i);
  // Tile results wait here until encoded: the window plus tiles waiting for an earlier pass.
  s->arena_floats = std::max<uint64_t>((uint64_t)s->cfg.window * 4 * 16, 1ull << 26);
  HIP_TRY(hipMalloc(&s->arena_dev, s->arena_floats * sizeof(float)));
  s->arena.reset(s->arena_floats);
  HIP_TRY(hipHostMalloc(&s->counters_host, COUNTER_BYTES, hipHostMallocDefault));
  memset(s->counters_host, 0, COUNTER_BYTES);
  s->enc_cap = 1u << 16;
  HIP_TRY(hipHostMalloc(&s->enc_host, s->enc_cap * sizeof(EncodeItem), hipHostMallocDefault));
  HIP_TRY(hipMalloc(&s->enc_dev, s->enc_cap * sizeof(EncodeItem)));
  MCR_TRY(engine_reserve(s->eng, s->cfg.slots, s->cfg.max_requests, 1024, true, s->cfg.max_lights));
  if (engine_capacity_for(s->eng, s->cfg.max_lights) < s->cfg.slots)
    return mcr_fail(MCR_E_INVALID, "slots x max_lights exceeds the light-slot memory budget; lower one of them");
  MCR_TRY(engine_reset(s->eng));
  HIP_TRY(counters_clear(s->eng.counters, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  memset(s->eng.counts_host, 0, sizeof(uint32_t) * COUNT_WORDS_ALLOC);
  s->eng.counts_host[cw(C_FREE)] = s->eng.capacity;
  uint64_t fixed = (uint64_t)s->arena_floats * 4;
  s->bytes_used = fixed;
  if (s->bytes_used > s->cfg.hbm_budget) return mcr_fail(MCR_E_OVERFLOW, "the HBM budget does not cover the server's own buffers");
  s->t0 = Clock::now();
  mcr_server *raw = s.release();
  raw->thread = std::thread(run, raw);
  *out = raw;
  return MCR_OK;
}

void mcr_server_destroy(mcr_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->wake.notify_all();
  if (s->thread.joinable()) s->thread.join();
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->stream);
  for (auto &kv : s->reqs) {
    hipFree(kv.second->image);
    hipFree(kv.second->sum);
  }
  engine_free(s->eng);
  for (mcr_scene *scene : s->scenes) scene_drop(scene);
  hipFree(s->arena_dev);
  hipFree(s->enc_dev);
  if (s->enc_host) hipHostFree(s->enc_host);
  if (s->jobs_host) hipHostFree(s->jobs_host);
  if (s->done_host) hipHostFree(s->done_host);
  if (s->counters_host) hipHostFree(s->counters_host);
  if (s->encoded) hipEventDestroy(s->encoded);
  if (s->copy_stream) hipStreamDestroy(s->copy_stream);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

int mcr_server_add_scene(mcr_server *s, mcr_scene *scene) {
  if (!s || !scene) return mcr_fail(MCR_E_INVALID, "null argument");
  if (!scene->has_geometry || !scene->has_shading) return mcr_fail(MCR_E_NO_SCENE, "the scene has no geometry or shading");
  if (scene->ctx != s->ctx) return mcr_fail(MCR_E_INVALID, "the scene belongs to another context");
  uint32_t slots_needed = scene->d.light_count + (scene->d.environment_count ? 1 : 0);
  if (slots_needed > s->cfg.max_lights)
    return mcr_fail(MCR_E_UNSUPPORTED, "the scene needs " + std::to_string(slots_needed) +
                                           " light slots; the server was created with max_lights " +
                                           std::to_string(s->cfg.max_lights));
  std::lock_guard<std::mutex> lk(s->mu);
  if (std::find(s->scenes.begin(), s->scenes.end(), scene) != s->scenes.end()) return MCR_OK;
  uint64_t b = mcr_scene_bytes(scene);
  if (s->bytes_used + b > s->cfg.hbm_budget) return mcr_fail(MCR_E_OVERFLOW, "the scene does not fit the HBM budget");
  s->bytes_used += b;
  s->scenes.push_back(scene);
  s->scene_bytes.push_back(b);
  scene_hold(scene);
  return MCR_OK;
}

int mcr_server_remove_scene(mcr_server *s, mcr_scene *scene) {
  if (!s || !scene) return mcr_fail(MCR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = std::find(s->scenes.begin(), s->scenes.end(), scene);
  if (it == s->scenes.end()) return mcr_fail(MCR_E_INVALID, "the scene was not added to the server");
  // A request holds the scene from submission until its last sample has drained (retire()).
  for (auto &kv : s->reqs) {
    const SrvRequest &r = *kv.second;
    bool live = r.state == MCR_REQ_QUEUED || r.state == MCR_REQ_RUNNING || r.view >= 0;
    if (live && r.r.scene == scene)
      return mcr_fail(MCR_E_INVALID, "requests on the scene are still live; cancel them or wait for them first");
  }
  size_t k = (size_t)(it - s->scenes.begin());
  s->bytes_used -= s->scene_bytes[k];
  s->scenes.erase(it);
  s->scene_bytes.erase(s->scene_bytes.begin() + (long)k);
  scene_drop(scene);  // may free it, if it was released while held
  return MCR_OK;
}

int mcr_server_submit(mcr_server *s, const mcr_request *req, uint64_t *id_out) {
  if (!s || !req || !id_out) return mcr_fail(MCR_E_INVALID, "null argument");
  mcr_request r = *req;
  if (!r.tile_size) r.tile_size = 64;
  if (!r.preview_block) r.preview_block = 4;
  if (r.width == 0 || r.height == 0 || r.samples == 0) return mcr_fail(MCR_E_INVALID, "empty request");
  if ((uint64_t)r.width * r.height > (1u << 28)) return mcr_fail(MCR_E_INVALID, "image too large");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fatal) return mcr_fail(s->fatal, s->fatal_msg);
  if (std::find(s->scenes.begin(), s->scenes.end(), r.scene) == s->scenes.end())
    return mcr_fail(MCR_E_INVALID, "the scene was not added to the server");
  // protocol.ts:76-78: the request names the scene versions it was made for (0: any). A held scene
  // cannot change, so what matches now matches until the request is done.
  const mcr_scene *sc = r.scene;
  if (!sc->has_geometry || !sc->has_shading || (r.geometry_version && r.geometry_version != sc->geometry_version) ||
      (r.shading_version && r.shading_version != sc->shading_version))
    return mcr_fail(MCR_E_NO_SCENE, "the tile refers to a scene this worker has not received");
  uint64_t need = request_bytes(r);
  if (need + (uint64_t)s->arena_floats * 4 > s->cfg.hbm_budget)
    return mcr_fail(MCR_E_OVERFLOW, "the request can never fit the HBM budget");
  auto q = std::unique_ptr<SrvRequest>(new SrvRequest());
  q->id = s->next_id++;
  q->r = r;
  q->bytes = need;
  q->t_submit = s->now();
  *id_out = q->id;
  s->admit_queue.push_back(q->id);
  s->reqs[q->id] = std::move(q);
  s->wake.notify_all();
  return MCR_OK;
}

int mcr_server_cancel(mcr_server *s, uint64_t id) {
  if (!s) return mcr_fail(MCR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return mcr_fail(MCR_E_INVALID, "no such request");
  it->second->cancel = true;
  s->wake.notify_all();
  return MCR_OK;
}

static void fill_status(const SrvRequest &r, mcr_request_status *o) {
  o->state = r.state;
  uint32_t m = r.tiles.empty() ? 0 : UINT32_MAX;
  for (const SrvTile &t : r.tiles) m = std::min(m, t.samples);
  o->samples = m;
  o->max_samples = r.r.samples;
  o->preview_done = r.t_preview >= 0;
  o->submit_ms = r.t_submit;
  o->admit_ms = r.t_admit;
  o->first_preview_ms = r.t_preview;
  o->done_ms = r.t_done;
  o->paths = r.paths;
  o->preview_iterations = r.t_preview >= 0 ? (uint32_t)(r.it_preview - r.it_admit) : 0;
}

int mcr_server_status(mcr_server *s, uint64_t id, mcr_request_status *out) {
  if (!s || !out) return mcr_fail(MCR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return mcr_fail(MCR_E_INVALID, "no such request");
  fill_status(*it->second, out);
  return MCR_OK;
}

int mcr_server_wait(mcr_server *s, uint64_t id, double timeout_ms) {
  if (!s) return mcr_fail(MCR_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto terminal = [&] {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return true;
    int st = it->second->state;
    return st == MCR_REQ_DONE || st == MCR_REQ_CANCELLED || st == MCR_REQ_FAILED;
  };
  if (timeout_ms < 0) s->changed.wait(lk, terminal);
  else if (!s->changed.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), terminal))
    return mcr_fail(MCR_E_TIMEOUT, "the request is still running");
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return mcr_fail(MCR_E_INVALID, "no such request");
  if (it->second->state == MCR_REQ_FAILED) return mcr_fail(s->fatal ? s->fatal : MCR_E_DEVICE, s->fatal_msg);
  if (it->second->state == MCR_REQ_CANCELLED) return mcr_fail(MCR_E_CANCELLED, "the request was cancelled");
  return MCR_OK;
}

int mcr_server_read(mcr_server *s, uint64_t id, uint8_t *rgba) {
  if (!s || !rgba) return mcr_fail(MCR_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return mcr_fail(MCR_E_INVALID, "no such request");
  SrvRequest &r = *it->second;
  if (!r.image) {
    memset(rgba, 0, (size_t)r.r.width * r.r.height * 4);
    return MCR_OK;
  }
  HIP_TRY(hipSetDevice(s->ctx->device));
  // After every encode launched so far; the image stays allocated while the lock is held.
  HIP_TRY(hipStreamWaitEvent(s->copy_stream, s->encoded, 0));
  HIP_TRY(hipMemcpyAsync(rgba, r.image, (size_t)r.r.width * r.r.height * 4, hipMemcpyDeviceToHost, s->copy_stream));
  HIP_TRY(hipStreamSynchronize(s->copy_stream));
  return MCR_OK;
}

int mcr_server_release(mcr_server *s, uint64_t id) {
  if (!s) return mcr_fail(MCR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return mcr_fail(MCR_E_INVALID, "no such request");
  int st = it->second->state;
  if (st == MCR_REQ_QUEUED || st == MCR_REQ_RUNNING) return mcr_fail(MCR_E_INVALID, "the request is still live; cancel it first");
  it->second->release = true;
  s->wake.notify_all();
  return MCR_OK;
}

int mcr_server_stats_get(mcr_server *s, mcr_server_stats *out) {
  if (!s || !out) return mcr_fail(MCR_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out->closest_queries = s->counters.closest_queries;
  out->any_queries = s->counters.any_queries;
  out->paths = s->paths;
  out->iterations = s->iterations;
  out->shade_waves = s->counters.shade_waves;
  out->mixed_view_waves = s->counters.mixed_view_waves;
  out->bytes_in_use = s->bytes_used;
  out->requests_live = s->running.size();
  out->requests_queued = s->admit_queue.size();
  out->device_ms = s->device_ms;
  out->uptime_ms = s->now();
  return MCR_OK;
}

}  // extern "C"
