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
#include <set>
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
  uint8_t listed = 0;               // lanes (by bit) whose encode list holds this tile
};

struct TileRef {
  uint64_t req;
  uint32_t tile;
};

struct SrvRequest {
  uint64_t id = 0;
  mcr_request r = {};
  int state = MCR_REQ_QUEUED;
  int error = MCR_OK;      // a failed request's own error (MCR_OK: the server's failure ended it)
  std::string error_msg;
  int view = -1;
  uint8_t *image = nullptr;
  float *sum = nullptr;
  uint64_t bytes = 0;
  std::vector<SrvTile> tiles;
  // Work items not in a job window, in work-list order: previews, and the pass-major items.
  std::deque<Item> pending_preview, pending_pass;
  uint32_t jobs_live = 0;           // jobs in the window or waiting to be encoded
  uint32_t covered = 0, finished_tiles = 0;
  bool cancel = false, release = false;
  double t_submit = -1, t_admit = -1, t_preview = -1, t_done = -1;
  uint64_t paths = 0;
  uint64_t it_admit = 0, it_preview = 0;  // bounce iterations of the preview lane so far
  // The first complete preview and the completion are stamped once the encodes that wrote them have
  // finished: the encode launches (by lane) they wait for, set when the host staged the last write.
  bool preview_staged = false, done_staged = false;
  uint64_t preview_fence[2] = {0, 0}, done_fence[2] = {0, 0};
  uint64_t staged_seq[2] = {0, 0};  // per lane, the fence of its last encode call holding an item of this request
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

// Best-fit allocator over the result arena (offsets and sizes in floats, multiples of 4): free
// ranges indexed by offset (to merge neighbours on release) and by (length, offset) (to find the
// smallest range that fits), so both operations cost O(log ranges) however fragmented it gets.
struct Arena {
  std::map<uint64_t, uint64_t> by_off;        // offset -> length
  std::set<std::pair<uint64_t, uint64_t>> by_len;  // (length, offset)
  void reset(uint64_t base, uint64_t n) {
    by_off.clear();
    by_len.clear();
    put(base, n);
  }
  void put(uint64_t off, uint64_t len) {
    by_off[off] = len;
    by_len.emplace(len, off);
  }
  void take(std::map<uint64_t, uint64_t>::iterator it) {
    by_len.erase({it->second, it->first});
    by_off.erase(it);
  }
  uint64_t alloc(uint64_t n) {
    n = (n + 3) & ~3ull;
    auto it = by_len.lower_bound({n, 0});
    if (it == by_len.end()) return UINT64_MAX;
    uint64_t len = it->first, off = it->second;
    by_len.erase(it);
    by_off.erase(off);
    if (len > n) put(off + n, len - n);
    return off;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 3) & ~3ull;
    auto next = by_off.lower_bound(off);
    if (next != by_off.end() && off + n == next->first) {
      n += next->second;
      take(next);
    }
    auto after = by_off.lower_bound(off);
    if (after != by_off.begin()) {
      auto prev = std::prev(after);
      if (prev->first + prev->second == off) {
        off = prev->first;
        n += prev->second;
        take(prev);
      }
    }
    put(off, n);
  }
};

// One wavefront with its own stream and job window. A server runs one lane that takes every work
// item, or a preview lane (the previews, on a few CUs kept for it) beside a pass lane (the passes,
// on the other CUs): a preview then waits for neither the pass lane's long bounce iterations nor
// its slots, and the pass lane can run as many slots as throughput wants.
struct Lane {
  Engine eng;
  hipStream_t stream = nullptr;    // the lane's steps, on the lane's CUs
  hipEvent_t done = nullptr;       // recorded after a launched step's count copies
  bool previews = true, passes = true;  // the items the lane takes
  uint32_t window = 0;             // unstarted samples queued ahead of the slots
  std::vector<SrvJob> table;       // the lane's device job list, by sample_start
  uint32_t total = 0;              // samples [0, total) may be started
  int q = 0;                       // queue holding the live paths
  bool inflight = false, views_dirty = true;
  uint32_t launched = 0;           // jobs of the step in flight
  uint64_t iterations = 0;
  Clock::time_point launched_at;   // when the step in flight was launched
  double busy_ms = 0;              // launch to seen-complete, summed over the lane's steps
  // Bounce iterations per step: a captured pair, or one direct launch; the pass lane of two lanes
  // runs two of those per step (its steps are long; fewer host round trips between them).
  uint32_t launches_per_step = 1;
  uint32_t iterations_per_step() const { return launches_per_step * (eng.use_graph ? 2u : 1u); }
  JobDev *jobs_host = nullptr;     // pinned staging of the job list
  uint32_t *done_host = nullptr;   // pinned: finished samples per job
  uint32_t host_cap = 0;
  void *counters_host = nullptr;   // pinned: the device counter copies
  DevCounters counters = {};
  // The lane's tile results wait in its own range of the server's arena until encoded, and the
  // lane encodes them on its own encode stream: a range read by an encode returns to the allocator
  // once the host has seen that encode's fence complete, so no step waits for encodes.
  Arena arena;
  EncodeItem *enc_host = nullptr, *enc_dev = nullptr;  // staging of the encode items
  hipStream_t enc_stream = nullptr;  // the lane's encodes, on the lane's CUs
  hipEvent_t encoded = nullptr;    // recorded after every encode launch of the lane
  // One fence per deliver call that launches encodes, polled (never waited for) to learn which
  // writes have finished: fences [1, enc_done] have completed, `fences` holds the rest in order.
  uint64_t enc_seq = 0, enc_done = 0;
  std::deque<std::pair<uint64_t, hipEvent_t>> fences;
  std::vector<hipEvent_t> spare_fences;
  // The encode items are staged in two halves used by turns; a half is rewritten once the fence of
  // its last use has completed (two deliver calls back, so normally without a wait).
  uint32_t half = 0;
  uint64_t half_fence[2] = {0, 0};
  // Arena ranges read by encodes of this lane, freed once the fence `seq` has completed.
  struct Deferred {
    uint64_t seq, off, len;
  };
  std::deque<Deferred> deferred;
  std::vector<TileRef> dirty;      // tiles holding a result of this lane that waits for its encode
  uint8_t bit = 1;                 // this lane's bit in SrvTile::listed
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
  hipStream_t stream = nullptr, copy_stream = nullptr;  // stream: admissions and frees
  hipEvent_t admitted = nullptr;   // recorded on `stream` after each admission's clears
  hipEvent_t read_ready = nullptr; // recorded on `stream` by mcr_server_read
  std::vector<std::unique_ptr<Lane>> lanes;  // lanes[0] takes the previews
  mcr_server_config cfg = {};
  float *arena_dev = nullptr;
  uint64_t arena_floats = 0;
  std::vector<ViewDev> views;
  std::vector<int> view_free;
  // The most light slots of a scene the server has held: the captured iterations' shadow grids are
  // sized by it, so they are captured once, not again each time a request with more lights comes.
  uint32_t scene_slots = 0;
  std::unordered_map<uint64_t, std::unique_ptr<SrvRequest>> reqs;
  std::deque<uint64_t> admit_queue;
  std::vector<uint64_t> running;   // admitted, in admission order
  uint64_t bytes_used = 0;
  std::vector<mcr_scene *> scenes;  // held (scene_hold) until removed or the server is destroyed
  std::vector<uint64_t> scene_bytes;  // what each counts against the HBM budget
  uint32_t enc_cap = 0;             // encode items staged per launch
  // stats
  uint64_t paths = 0;
  double device_ms = 0;            // time with at least one lane's step in flight
  double host_ms = 0, plan_ms = 0, deliver_ms = 0;  // the server thread's own work
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
  uint64_t preview_iterations() const { return lanes[0]->iterations; }
  DevCounters counters() const {
    DevCounters c = {};
    for (const auto &l : lanes) {
      const unsigned long long *a = reinterpret_cast<const unsigned long long *>(&l->counters);
      unsigned long long *o = reinterpret_cast<unsigned long long *>(&c);
      for (size_t k = 0; k < sizeof(DevCounters) / sizeof(unsigned long long); k++) o[k] += a[k];
    }
    return c;
  }
};

namespace {

// The lane whose arena holds preview (or pass) results.
Lane &result_lane(mcr_server *s, bool preview) { return *s->lanes[preview || s->lanes.size() == 1 ? 0 : 1]; }

// Orders `stream` after every encode launched so far (request buffers are written by encodes).
void after_encodes(mcr_server *s, hipStream_t stream) {
  for (auto &l : s->lanes) (void)hipStreamWaitEvent(stream, l->encoded, 0);
}

// Renders beyond this many samples keep their running sum in FP64.
const uint32_t WIDE_SUM_SAMPLES = 1024;

uint64_t sum_bytes(const mcr_request &r) {
  return (uint64_t)r.width * r.height * (r.samples > WIDE_SUM_SAMPLES ? 32 : 16);
}
uint64_t request_bytes(const mcr_request &r) { return (uint64_t)r.width * r.height * 4 + sum_bytes(r); }

// Frees a terminal request's device memory and slot once nothing of it is in flight. Request
// buffers are written only by encodes (on the lanes' streams), so the frees on s->stream are
// ordered after every encode launched so far.
void retire(mcr_server *s, SrvRequest &r) {
  if (r.jobs_live || r.view < 0) return;
  if (r.state != MCR_REQ_DONE && r.state != MCR_REQ_CANCELLED && r.state != MCR_REQ_FAILED) return;
  after_encodes(s, s->stream);
  if (r.state == MCR_REQ_CANCELLED || r.state == MCR_REQ_FAILED) {
    hipFreeAsync(r.image, s->stream);
    hipFreeAsync(r.sum, s->stream);
    r.image = nullptr;
    r.sum = nullptr;
    s->bytes_used -= r.bytes;
  } else {
    hipFreeAsync(r.sum, s->stream);  // the image stays readable until released
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
  HIP_TRY(hipMallocAsync((void **)&r.image, npx * 4, s->stream));
  uint64_t sb = sum_bytes(r.r);
  hipError_t e = hipMallocAsync((void **)&r.sum, sb, s->stream);
  if (e != hipSuccess) {
    hipFreeAsync(r.image, s->stream);
    r.image = nullptr;
    return mcr_fail(MCR_E_DEVICE, std::string("hipMallocAsync: ") + hipGetErrorString(e));
  }
  HIP_TRY(hipMemsetAsync(r.image, 0, npx * 4, s->stream));
  HIP_TRY(hipMemsetAsync(r.sum, 0, sb, s->stream));
  HIP_TRY(hipEventRecord(s->admitted, s->stream));  // encodes into the buffers wait for this
  r.view = s->view_free.back();
  s->view_free.pop_back();
  s->views[r.view] = make_view(r.r.scene, r.r.camera, r.r.settings);
  for (auto &l : s->lanes) l->views_dirty = true;
  s->bytes_used += r.bytes;
  auto tiles = make_tiles(r.r.width, r.r.height, r.r.tile_size);
  r.tiles.resize(tiles.size());
  for (size_t i = 0; i < tiles.size(); i++) r.tiles[i].t = tiles[i];
  // progressive.ts:207-210: preview items for every tile, then pass-major items.
  if (r.r.preview_block > 1)
    for (uint32_t i = 0; i < tiles.size(); i++) r.pending_preview.push_back(Item{i, 0, r.r.preview_block});
  for (uint32_t p = 0; p < r.r.samples; p++)
    for (uint32_t i = 0; i < tiles.size(); i++) r.pending_pass.push_back(Item{i, p, 1});
  r.state = MCR_REQ_RUNNING;
  r.t_admit = s->now();
  r.it_admit = s->preview_iterations();
  s->running.push_back(r.id);
  return MCR_OK;
}

// Ends a queued or running request as cancelled or failed: its unstarted items are dropped and
// the tile results not yet encoded return to their arenas. Its samples in flight finish and are
// discarded (collect); retire() frees its buffers once none is left.
void end_request(mcr_server *s, SrvRequest &r, int state) {
  if (r.state == MCR_REQ_QUEUED)
    s->admit_queue.erase(std::remove(s->admit_queue.begin(), s->admit_queue.end(), r.id), s->admit_queue.end());
  r.state = state;
  r.pending_preview.clear();
  r.pending_pass.clear();
  for (SrvTile &t : r.tiles) {
    uint64_t floats = (uint64_t)t.t.width * t.t.height * 4;
    for (auto &kv : t.ready) result_lane(s, false).arena.release(kv.second, floats);
    t.ready.clear();
    if (t.preview != UINT64_MAX) result_lane(s, true).arena.release(t.preview, floats);
    t.preview = UINT64_MAX;
  }
  r.t_done = s->now();
  s->changed.notify_all();
}

// Admissions, cancellations and releases (under the lock).
int housekeeping(mcr_server *s) {
  for (auto it = s->reqs.begin(); it != s->reqs.end();) {
    SrvRequest &r = *it->second;
    if (r.cancel && (r.state == MCR_REQ_QUEUED || r.state == MCR_REQ_RUNNING)) end_request(s, r, MCR_REQ_CANCELLED);
    retire(s, r);
    if (r.release && r.view < 0 && r.jobs_live == 0) {
      after_encodes(s, s->stream);
      if (r.image) hipFreeAsync(r.image, s->stream);
      if (r.sum) hipFreeAsync(r.sum, s->stream);
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

// The job window of lane L. `next` is the lane's first sample not yet started; `idle` slots are
// free now and take their samples from the window first, so the window holds that many samples
// more. A lane that takes both kinds of items returns its unstarted jobs to their requests and
// chooses the window afresh, so a new request's preview overtakes queued passes. A lane of one
// kind keeps its unstarted jobs where they are (renumbered to follow the started ones) and only
// drops those of requests that have ended, so a step's planning costs what it adds.
void plan_window(mcr_server *s, Lane &L, uint32_t next, uint32_t idle) {
  std::vector<SrvJob> keep;
  std::unordered_map<uint64_t, std::vector<Item>> returned;
  const bool afresh = L.previews && L.passes;
  uint32_t base = next;
  for (SrvJob &j : L.table)  // started jobs come first (the table is in sample order)
    if (j.done != j.size && j.dev.sample_start < next) base = std::max(base, j.dev.sample_start + j.size);
  uint32_t cursor = base;
  keep.reserve(L.table.size());
  for (SrvJob &j : L.table) {
    if (j.done == j.size) continue;  // finished: taken out by collect() before this runs
    if (j.dev.sample_start < next) {
      keep.push_back(j);
      continue;
    }
    SrvRequest &r = *s->reqs[j.req];
    if (!afresh && r.state == MCR_REQ_RUNNING) {
      j.dev.sample_start = cursor;
      cursor += j.size;
      keep.push_back(j);
      continue;
    }
    if (r.state == MCR_REQ_RUNNING) returned[j.req].push_back(j.item);
    result_lane(s, j.item.block > 1).arena.release(j.arena, j.floats);
    r.jobs_live--;
  }
  for (auto &kv : returned) {
    SrvRequest &r = *s->reqs[kv.first];
    for (auto it = kv.second.rbegin(); it != kv.second.rend(); ++it)
      (it->block > 1 ? r.pending_preview : r.pending_pass).push_front(*it);
  }
  L.table.swap(keep);
  base = cursor;
  // A step runs L.iterations_per_step bounce iterations, each of which may start a window's worth.
  uint64_t target = (uint64_t)L.window * L.iterations_per_step() + idle, queued = cursor - next;
  uint32_t start = base;
  auto place = [&](SrvRequest &r, std::deque<Item> &pending) -> bool {
    Item it = pending.front();
    const mcr_tile &t = r.tiles[it.tile].t;
    // Sample indices stay below 2^31; past that the window drains and numbering restarts.
    if ((uint64_t)start + (uint64_t)t.width * t.height > (1ull << 31)) return false;
    uint64_t floats = (uint64_t)t.width * t.height * 4;
    uint64_t off = result_lane(s, it.block > 1).arena.alloc(floats);
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
    L.table.push_back(j);
    pending.pop_front();
    r.jobs_live++;
    start += (uint32_t)n;
    queued += (uint64_t)n;
    return true;
  };
  bool room = true;
  // Previews of requests whose preview has not been queued completely go first.
  if (L.previews)
    for (uint64_t id : s->running) {
      SrvRequest &r = *s->reqs[id];
      while (room && r.state == MCR_REQ_RUNNING && !r.pending_preview.empty()) room = place(r, r.pending_preview);
    }
  // Then the passes of the oldest request first: the wavefront keeps working on few scenes at a
  // time (their BVHs stay in cache), and requests finish in arrival order.
  // A preview lane beside a pass lane fills only an eighth of its window with passes: a preview's
  // latency grows with the paths its lane carries, so passes there only take up CUs that would
  // otherwise idle between previews.
  uint64_t pass_target = target;
  if (L.previews && s->lanes.size() > 1)
    pass_target = std::min<uint64_t>(target, (uint64_t)(L.window / 8) * L.iterations_per_step());
  if (L.passes)
    for (uint64_t id : s->running) {
      SrvRequest &r = *s->reqs[id];
      // One lane: a request's passes follow its preview (progressive.ts:207-210).
      if (L.previews && !r.pending_preview.empty()) continue;
      while (room && queued < pass_target && r.state == MCR_REQ_RUNNING && !r.pending_pass.empty())
        room = place(r, r.pending_pass);
      if (!room || queued >= pass_target) break;
    }
  L.total = start;
}

// Finished jobs of lane L (its step has been seen complete) go to their requests: tile results
// wait in the arena, listed for the lane's encodes. Sets progress if anything moved.
void collect(mcr_server *s, Lane &L, bool &progress) {
  std::vector<SrvJob> still;
  for (SrvJob &j : L.table) {
    if (j.done != j.size) {
      still.push_back(j);
      continue;
    }
    progress = true;
    SrvRequest &r = *s->reqs[j.req];
    r.paths += j.size;
    s->paths += j.size;
    r.jobs_live--;
    Lane &E = result_lane(s, j.item.block > 1);  // holds the result and encodes it
    SrvTile *t = r.state == MCR_REQ_RUNNING ? &r.tiles[j.item.tile] : nullptr;
    if (!t) {
      E.arena.release(j.arena, j.floats);
      continue;
    }
    if (j.item.block > 1) t->preview = j.arena;
    else t->ready[j.item.pass] = j.arena;
    if (!(t->listed & E.bit)) {
      t->listed |= E.bit;
      E.dirty.push_back(TileRef{r.id, j.item.tile});
    }
  }
  L.table.swap(still);
}

// Takes the lane's completed encode fences (no wait).
int fences_poll(Lane &L) {
  while (!L.fences.empty()) {
    hipError_t q = hipEventQuery(L.fences.front().second);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return mcr_fail(MCR_E_DEVICE, std::string("encode: ") + hipGetErrorString(q));
    L.enc_done = L.fences.front().first;
    L.spare_fences.push_back(L.fences.front().second);
    L.fences.pop_front();
  }
  // Arena ranges whose encodes have finished are free on the device: back to the allocator.
  while (!L.deferred.empty() && L.deferred.front().seq <= L.enc_done) {
    L.arena.release(L.deferred.front().off, L.deferred.front().len);
    L.deferred.pop_front();
  }
  return MCR_OK;
}

// Waits until the lane's fence `seq` (0: none) has completed.
int fence_wait(Lane &L, uint64_t seq) {
  MCR_TRY(fences_poll(L));
  if (seq <= L.enc_done) return MCR_OK;
  for (auto &f : L.fences)
    if (f.first == seq) HIP_TRY(hipEventSynchronize(f.second));
  return fences_poll(L);
}

// L's tile results are written in pass order by encodes on the lane's encode stream, which run
// beside the lane's next step (launched first, so the device does not wait for this host work).
// The step whose results these are has been seen complete by the host, so the encodes need no
// wait for it; the arena ranges they read return to the allocator once their fence has
// completed. Sets progress if anything moved.
int deliver(mcr_server *s, Lane &L, bool &progress) {
  if (L.dirty.empty()) return MCR_OK;  // no tile result of the lane waits for its encode
  // Rounds of one launch each over the tiles holding a result of this lane; a tile is staged at
  // most once per round, so no two items of a launch write the same pixel. A tile stays listed
  // while it may hold another item for a later round.
  uint32_t nenc = 0, launched = 0;
  std::vector<std::pair<uint64_t, uint64_t>> to_free;  // arena ranges, freed after the launches
  // This call stages into half L.half, whose last encodes (and their copy) must have finished.
  MCR_TRY(fence_wait(L, L.half_fence[L.half]));
  const uint32_t cap = s->enc_cap / 2;
  EncodeItem *enc_host = L.enc_host + (size_t)L.half * cap, *enc_dev = L.enc_dev + (size_t)L.half * cap;
  std::vector<TileRef> cur, again;
  std::vector<SrvRequest *> touched;  // requests with an item staged by this call
  cur.swap(L.dirty);
  for (;;) {
    uint32_t first = nenc;
    again.clear();
    for (const TileRef &d : cur) {
      auto f = s->reqs.find(d.req);
      if (f == s->reqs.end()) continue;
      SrvRequest &r = *f->second;
      SrvTile &t = r.tiles[d.tile];
      if (r.state != MCR_REQ_RUNNING) {
        t.listed &= (uint8_t)~L.bit;
        continue;
      }
      if (nenc == cap) {  // waits for the next call
        L.dirty.push_back(d);
        continue;
      }
      uint64_t floats = (uint64_t)t.t.width * t.t.height * 4;
      uint64_t data;
      bool preview = false;
      if (L.previews && t.preview != UINT64_MAX) {
        // progressive.ts:257-258: a preview is written only over a tile with no samples yet.
        data = t.preview;
        t.preview = UINT64_MAX;
        to_free.emplace_back(data, floats);
        again.push_back(d);  // its passes, in a later round
        if (t.samples != 0 || t.ready.count(0)) continue;
        preview = true;
      } else if (&L == &result_lane(s, false)) {
        auto it = t.ready.find(t.samples);
        if (it == t.ready.end()) {
          t.listed &= (uint8_t)~L.bit;
          continue;
        }
        data = it->second;
        t.ready.erase(it);
        to_free.emplace_back(data, floats);
        t.samples++;
        if (t.samples == r.r.samples) r.finished_tiles++;
        again.push_back(d);
      } else {
        t.listed &= (uint8_t)~L.bit;
        continue;
      }
      bool wide = r.r.samples > WIDE_SUM_SAMPLES;
      touched.push_back(&r);
      enc_host[nenc++] = EncodeItem{r.image, preview || wide ? nullptr : r.sum,
                                      preview || !wide ? nullptr : reinterpret_cast<double *>(r.sum),
                                      s->arena_dev + data, r.r.width,
                                      t.t.x, t.t.y, t.t.width, t.t.height, preview ? 1u : t.samples,
                                      r.r.srgb, 0};
      if (!t.covered) {
        t.covered = true;
        r.covered++;
      }
    }
    if (nenc > first) {
      uint32_t blocks = 0;
      for (uint32_t k = first; k < nenc; k++) {
        enc_host[k].block_start = blocks;
        blocks += (uint32_t)(((uint64_t)enc_host[k].tw * enc_host[k].th + 255) / 256);
      }
      if (first == 0) {
        // After the buffers' clears at admission; a pass lane also after every preview encode
        // launched so far (a preview is written only before a tile's first pass).
        HIP_TRY(hipStreamWaitEvent(L.enc_stream, s->admitted, 0));
        if (!L.previews) HIP_TRY(hipStreamWaitEvent(L.enc_stream, s->lanes[0]->encoded, 0));
      }
      HIP_TRY(hipMemcpyAsync(enc_dev + first, enc_host + first, (nenc - first) * sizeof(EncodeItem),
                             hipMemcpyHostToDevice, L.enc_stream));
      launch_encode_batch(L.enc_stream, enc_dev + first, nenc - first, blocks);
      HIP_TRY(hipGetLastError());
      launched = nenc;
      progress = true;
    }
    if (again.empty()) break;
    if (nenc == cap) {  // the rest waits for the next call
      L.dirty.insert(L.dirty.end(), again.begin(), again.end());
      break;
    }
    cur.swap(again);
  }
  if (launched) {
    HIP_TRY(hipEventRecord(L.encoded, L.enc_stream));
    hipEvent_t ev = nullptr;
    if (!L.spare_fences.empty()) {
      ev = L.spare_fences.back();
      L.spare_fences.pop_back();
    } else {
      HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(ev, L.enc_stream));
    L.fences.emplace_back(++L.enc_seq, ev);
    L.half_fence[L.half] = L.enc_seq;
    L.half ^= 1;
    const int li = __builtin_ctz(L.bit);
    for (SrvRequest *r : touched) r->staged_seq[li] = L.enc_seq;
  }
  // The ranges go back to the allocator once this call's fence has completed (fences_poll): a step
  // of any lane may reuse them then without waiting for the encodes on the device.
  for (auto &f : to_free) L.deferred.push_back(Lane::Deferred{L.enc_seq, f.first, f.second});
  // A request whose last preview tile or last pass has been staged waits for the encodes that hold
  // its items on each lane (the lane's last encode call with one of them; stream order on the
  // lane's encode stream covers the earlier ones); stamp() marks it once they have finished. An
  // encode of another request does not hold it back.
  for (uint64_t id : s->running) {
    SrvRequest &r = *s->reqs[id];
    if (r.state != MCR_REQ_RUNNING) continue;
    if (!r.preview_staged && r.covered == r.tiles.size()) {
      r.preview_staged = true;
      for (size_t k = 0; k < s->lanes.size(); k++) r.preview_fence[k] = r.staged_seq[k];
    }
    if (!r.done_staged && r.finished_tiles == r.tiles.size() && r.pending_pass.empty() && r.pending_preview.empty()) {
      r.done_staged = true;
      for (size_t k = 0; k < s->lanes.size(); k++) r.done_fence[k] = r.staged_seq[k];
    }
  }
  return MCR_OK;
}

// Stamps the first complete preview and the completion of requests whose encodes have finished.
// Sets progress if anything changed.
void stamp(mcr_server *s, bool &progress) {
  auto finished = [&](const uint64_t *fence) {
    for (size_t k = 0; k < s->lanes.size(); k++)
      if (s->lanes[k]->enc_done < fence[k]) return false;
    return true;
  };
  double now = s->now();
  std::vector<uint64_t> ids(s->running);
  for (uint64_t id : ids) {
    SrvRequest &r = *s->reqs[id];
    if (r.state != MCR_REQ_RUNNING) continue;
    if (r.t_preview < 0 && r.preview_staged && finished(r.preview_fence)) {
      r.t_preview = now;
      r.it_preview = s->preview_iterations();
      s->changed.notify_all();
      progress = true;
    }
    if (r.done_staged && finished(r.done_fence)) {
      r.state = MCR_REQ_DONE;
      r.t_done = now;
      retire(s, r);
      s->changed.notify_all();
      progress = true;
    }
  }
}

int ensure_host(Lane &L, uint32_t n) {
  if (L.host_cap >= n) return MCR_OK;
  uint32_t want = std::max<uint32_t>(n, L.host_cap * 2);
  if (L.jobs_host) hipHostFree(L.jobs_host);
  if (L.done_host) hipHostFree(L.done_host);
  L.jobs_host = nullptr;
  L.done_host = nullptr;
  L.host_cap = 0;
  HIP_TRY(hipHostMalloc(&L.jobs_host, want * sizeof(JobDev), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(&L.done_host, want * sizeof(uint32_t), hipHostMallocDefault));
  L.host_cap = want;
  return MCR_OK;
}

// A shadow ray crossed more counted see-through surfaces than the reference-order pass keeps
// (the reference's list has no bound, so the image would differ from its raytrace()): the step's
// counts name the views of such rays, and only the requests holding those views fail. Every
// request's samples of the step have finished, so a flagged view still belongs to the request
// whose samples set it (views return to the pool only once a request has no job left).
int fail_overflowed(mcr_server *s, Lane &L) {
  uint32_t *c = L.eng.counts_host;
  for (uint64_t id : s->running) {
    SrvRequest &r = *s->reqs[id];
    if (r.view < 0 || r.state != MCR_REQ_RUNNING) continue;
    if (!((c[C_FAIL_BITS + r.view / 32] >> (r.view % 32)) & 1u)) continue;
    r.error = MCR_E_OVERFLOW;
    r.error_msg = "a shadow ray crossed more than " + std::to_string(MCR_DEEP_HITS) + " counted see-through surfaces";
    end_request(s, r, MCR_REQ_FAILED);
  }
  // The lane's next step is ordered after these clears on its stream (not a captured graph).
  HIP_TRY(hipMemsetAsync(L.eng.w.counts + cw(C_OVERFLOW), 0, sizeof(uint32_t), L.stream));
  HIP_TRY(hipMemsetAsync(L.eng.w.counts + C_FAIL_BITS, 0, sizeof(uint32_t) * (MCR_FAIL_VIEWS / 32), L.stream));
  c[cw(C_OVERFLOW)] = 0;
  memset(c + C_FAIL_BITS, 0, sizeof(uint32_t) * (MCR_FAIL_VIEWS / 32));
  return MCR_OK;
}

// Lane L's step in flight: if it has finished, takes its counts (no wait otherwise).
int lane_poll(mcr_server *s, Lane &L, bool &progress) {
  if (!L.inflight) return MCR_OK;
  hipError_t q = hipEventQuery(L.done);
  if (q == hipErrorNotReady) return MCR_OK;
  if (q != hipSuccess) return mcr_fail(MCR_E_DEVICE, std::string("lane step: ") + hipGetErrorString(q));
  L.inflight = false;
  L.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - L.launched_at).count();
  progress = true;
  if (L.eng.counts_host[cw(C_OVERFLOW)] == OVERFLOW_SEEN) MCR_TRY(fail_overflowed(s, L));
  MCR_TRY(engine_check(L.eng));
  for (uint32_t i = 0; i < L.launched; i++) L.table[i].done = L.done_host[i];
  L.counters = counters_sum(L.counters_host);
  return MCR_OK;
}

// Plans lane L's window and launches its next step (a refill, bounce iterations and the copies of
// its counts), recording L.done after them. Waits for nothing.
int lane_launch(mcr_server *s, Lane &L, bool &progress) {
  Engine &e = L.eng;
  // With no path live every started sample has finished and no job in the table has started,
  // so sample numbering starts again from 0.
  uint32_t next = e.counts_host[cw(L.q)] == 0 ? 0 : std::min(e.counts_host[cw(C_NEXT)], L.total);
  Clock::time_point t_plan = Clock::now();
  plan_window(s, L, next, e.counts_host[cw(C_FREE)]);
  s->plan_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_plan).count();
  uint32_t n = (uint32_t)L.table.size();
  if (n == 0 && e.counts_host[cw(L.q)] == 0) return MCR_OK;
  progress = true;
  MCR_TRY(engine_reserve(e, e.capacity_limit, (uint32_t)s->views.size(), n, true, s->cfg.max_lights));
  MCR_TRY(ensure_host(L, std::max<uint32_t>(n, 1)));
  bool mixed = s->running.size() > 1;  // samples of other views may be in flight with no job left
  for (const SrvJob &j : L.table) mixed |= j.dev.view != 0;
  e.mix = mixed;
  // The lane's last step has completed, so its pinned staging is free.
  for (uint32_t i = 0; i < n; i++) {
    L.jobs_host[i] = L.table[i].dev;
    L.done_host[i] = L.table[i].done;
  }
  if (L.views_dirty) {
    MCR_TRY(engine_upload_views(e, s->views.data(), (uint32_t)s->views.size()));
    L.views_dirty = false;
  }
  if (n) {
    HIP_TRY(hipMemcpyAsync(e.jobs, L.jobs_host, n * sizeof(JobDev), hipMemcpyHostToDevice, L.stream));
    HIP_TRY(hipMemcpyAsync(e.job_done, L.done_host, n * sizeof(uint32_t), hipMemcpyHostToDevice, L.stream));
  }
  MCR_TRY(engine_set_control(e, s->arena_dev, n, L.total, true));
  uint32_t free_slots = e.counts_host[cw(C_FREE)], live = e.counts_host[cw(L.q)];
  uint32_t k = std::min<uint32_t>(free_slots, L.total - next);
  MCR_TRY(engine_refill(e, L.q, k, free_slots, next, live));
  live += k;
  // Refilled slots come from the free list in the order they went idle (after a drain, every
  // slot does), and in-place regeneration would keep them scattered for the rest of the run;
  // re-sorting the queue after a refill puts every wave back on neighbouring slots (coalesced
  // path-state accesses). Steps without a refill keep the queue's own order: waves requeue in
  // the order they finish, which groups paths of similar cost.
  if (k) MCR_TRY(engine_sort_queue(e, L.q, live));
  uint32_t max_lights = s->scene_slots;
  e.tex = false;
  for (const ViewDev &v : s->views) {
    max_lights = std::max(max_lights, v.shadow_slots);
    e.tex |= v.shadow_tex != 0;
  }
  for (uint32_t k = 0; live && k < L.launches_per_step; k++) {
    if (e.use_graph) {
      if (L.q != 0) {  // the captured pair starts on queue 0
        MCR_TRY(engine_iteration(e, L.q, live, max_lights, nullptr, nullptr, nullptr));
        L.q ^= 1;
        L.iterations++;
      }
      MCR_TRY(engine_iteration_pair(e, max_lights));
      L.iterations += 2;
    } else {
      MCR_TRY(engine_iteration(e, L.q, live, max_lights, nullptr, nullptr, nullptr));
      L.q ^= 1;
      L.iterations++;
    }
  }
  // Counter words, finished samples per job and ray counters, behind one event.
  MCR_TRY(engine_fetch_counts(e));
  if (n) HIP_TRY(hipMemcpyAsync(L.done_host, e.job_done, n * sizeof(uint32_t), hipMemcpyDeviceToHost, L.stream));
  HIP_TRY(hipMemcpyAsync(L.counters_host, e.counters, COUNTER_BYTES, hipMemcpyDeviceToHost, L.stream));
  HIP_TRY(hipEventRecord(L.done, L.stream));
  L.launched = n;
  L.inflight = true;
  L.launched_at = Clock::now();
  return MCR_OK;
}

bool lanes_busy(const mcr_server *s) {
  for (const auto &l : s->lanes)
    if (l->inflight || !l->table.empty() || l->eng.counts_host[cw(l->q)] != 0) return true;
  return false;
}

void run(mcr_server *s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  Clock::time_point busy_since;
  bool was_busy = false;
  while (!s->stop) {
    bool busy = !s->running.empty() || !s->admit_queue.empty() || lanes_busy(s);
    bool chores = false;
    for (auto &kv : s->reqs) chores |= kv.second->cancel || kv.second->release;
    if (!busy && !chores) {
      s->wake.wait(lk);
      continue;
    }
    bool progress = false;
    Clock::time_point t_work = Clock::now();
    int rc = housekeeping(s);
    // Each lane: take a finished step, launch its next step, encode the finished tiles. The lanes
    // run side by side on their own streams; this thread polls and never waits for a step.
    for (auto &l : s->lanes) {
      if (!rc) rc = lane_poll(s, *l, progress);
      if (rc || l->inflight) continue;
      collect(s, *l, progress);
      rc = lane_launch(s, *l, progress);
      Clock::time_point t_deliver = Clock::now();
      if (!rc) rc = deliver(s, *l, progress);
      s->deliver_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_deliver).count();
    }
    for (auto &l : s->lanes)
      if (!rc) rc = fences_poll(*l);
    if (!rc) stamp(s, progress);
    s->host_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_work).count();
    bool inflight = false;
    for (auto &l : s->lanes) inflight |= l->inflight;
    auto now = Clock::now();
    if (inflight && !was_busy) busy_since = now;
    if (!inflight && was_busy) s->device_ms += std::chrono::duration<double, std::milli>(now - busy_since).count();
    was_busy = inflight;
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
      for (auto &l : s->lanes) l->table.clear();
      s->changed.notify_all();
      s->wake.wait(lk, [&] { return s->stop; });
      continue;
    }
    if (progress) continue;
    bool encoding = false;
    for (auto &l : s->lanes) encoding |= !l->fences.empty();
    if (inflight || encoding) {
      // Steps or encodes in flight: look again shortly (API calls take the lock meanwhile).
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(40));
      lk.lock();
    } else {
      // Nothing to trace (e.g. every queued request waits for memory): sleep until woken.
      s->wake.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
  if (was_busy) s->device_ms += std::chrono::duration<double, std::milli>(Clock::now() - busy_since).count();
}

}  // namespace

extern "C" {

static void server_free(mcr_server *s) {
  hipSetDevice(s->ctx->device);
  for (auto &l : s->lanes) {
    if (l->stream) hipStreamSynchronize(l->stream);
    if (l->enc_stream) hipStreamSynchronize(l->enc_stream);
  }
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto &kv : s->reqs) {
    if (kv.second->image) hipFreeAsync(kv.second->image, s->stream);
    if (kv.second->sum) hipFreeAsync(kv.second->sum, s->stream);
  }
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto &l : s->lanes) {
    engine_free(l->eng);
    hipFree(l->enc_dev);
    if (l->enc_host) hipHostFree(l->enc_host);
    if (l->encoded) hipEventDestroy(l->encoded);
    if (l->jobs_host) hipHostFree(l->jobs_host);
    if (l->done_host) hipHostFree(l->done_host);
    if (l->counters_host) hipHostFree(l->counters_host);
    if (l->done) hipEventDestroy(l->done);
    if (l->enc_stream) hipStreamDestroy(l->enc_stream);
    for (auto &f : l->fences) hipEventDestroy(f.second);
    for (hipEvent_t ev : l->spare_fences) hipEventDestroy(ev);
    if (l->stream) hipStreamDestroy(l->stream);
  }
  s->lanes.clear();
  for (mcr_scene *scene : s->scenes) scene_drop(scene);
  s->scenes.clear();
  hipFree(s->arena_dev);
  if (s->admitted) hipEventDestroy(s->admitted);
  if (s->read_ready) hipEventDestroy(s->read_ready);
  if (s->copy_stream) hipStreamDestroy(s->copy_stream);
  if (s->stream) hipStreamDestroy(s->stream);
}

const uint32_t LANE_JOBS = 1u << 16;  // job-list entries a lane starts with

// A lane on CUs [cu_first, cu_first + cu_count) of the device (all of them when cu_count is 0 or
// a CU-masked stream cannot be made), with `slots` path slots.
static int lane_init(mcr_server *s, Lane &L, uint32_t slots, uint32_t window, bool previews, bool passes,
                     int cu_first, int cu_count, int cus) {
  L.previews = previews;
  L.passes = passes;
  L.window = window;
  if (cu_count > 0 && cu_count < cus) {
    uint32_t words = (uint32_t)(cus + 31) / 32;
    std::vector<uint32_t> mask(words, 0u);
    for (int k = cu_first; k < cu_first + cu_count && k < cus; k++) mask[(size_t)k / 32] |= 1u << (k % 32);
    if (hipExtStreamCreateWithCUMask(&L.stream, words, mask.data()) != hipSuccess ||
        hipExtStreamCreateWithCUMask(&L.enc_stream, words, mask.data()) != hipSuccess) {
      (void)hipGetLastError();
      if (L.stream) hipStreamDestroy(L.stream);
      if (L.enc_stream) hipStreamDestroy(L.enc_stream);
      L.stream = L.enc_stream = nullptr;
    }
  }
  if (!L.stream) HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  if (!L.enc_stream) HIP_TRY(hipStreamCreateWithFlags(&L.enc_stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&L.encoded, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc(&L.enc_host, s->enc_cap * sizeof(EncodeItem), hipHostMallocDefault));
  HIP_TRY(hipMalloc(&L.enc_dev, s->enc_cap * sizeof(EncodeItem)));
  MCR_TRY(engine_init(L.eng, s->ctx->device, L.stream));
  L.eng.capacity_limit = slots;
  L.eng.use_graph = s->cfg.use_graph != 0;
  HIP_TRY(hipHostMalloc(&L.counters_host, COUNTER_BYTES, hipHostMallocDefault));
  memset(L.counters_host, 0, COUNTER_BYTES);
  // Job lists sized up front: growing one later frees device memory, which waits for every lane.
  MCR_TRY(engine_reserve(L.eng, slots, s->cfg.max_requests, LANE_JOBS, true, s->cfg.max_lights));
  MCR_TRY(ensure_host(L, LANE_JOBS));
  if (engine_capacity_for(L.eng, s->cfg.max_lights) < slots)
    return mcr_fail(MCR_E_INVALID, "slots x max_lights exceeds the light-slot memory budget; lower one of them");
  MCR_TRY(engine_reset(L.eng));
  HIP_TRY(counters_clear(L.eng.counters, L.stream));
  HIP_TRY(hipStreamSynchronize(L.stream));
  memset(L.eng.counts_host, 0, sizeof(uint32_t) * COUNT_WORDS_ALLOC);
  L.eng.counts_host[cw(C_FREE)] = L.eng.capacity;
  return MCR_OK;
}

int mcr_server_create(mcr_ctx *c, const mcr_server_config *config, mcr_server **out) {
  if (!c || !out) return mcr_fail(MCR_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  std::unique_ptr<mcr_server> s(new mcr_server());
  s->ctx = c;
  if (config) s->cfg = *config;
  else s->cfg.use_graph = 1;
  if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 200ull << 30;
  if (!s->cfg.lanes) s->cfg.lanes = 2;
  if (s->cfg.lanes > 2) return mcr_fail(MCR_E_INVALID, "lanes must be 1 or 2");
  // One lane bounds its slots for short previews; a pass lane can take more for throughput.
  if (!s->cfg.slots) s->cfg.slots = s->cfg.lanes == 1 ? SERVER_DEFAULT_SLOTS : SERVER_DEFAULT_PASS_SLOTS;
  if (s->cfg.slots < 256 || s->cfg.slots > (1u << 26)) return mcr_fail(MCR_E_INVALID, "slots must be in [256, 2^26]");
  if (!s->cfg.window) s->cfg.window = s->cfg.slots;
  if (!s->cfg.max_requests) s->cfg.max_requests = 64;
  if (s->cfg.max_requests > MCR_FAIL_VIEWS)
    return mcr_fail(MCR_E_INVALID, "max_requests must be at most " + std::to_string(MCR_FAIL_VIEWS));
  if (!s->cfg.max_lights) s->cfg.max_lights = 16;
  if (s->cfg.max_lights > MCR_MAX_LIGHTS) return mcr_fail(MCR_E_INVALID, "max_lights beyond MCR_MAX_LIGHTS_PER_SCENE");
  if (!s->cfg.preview_slots) s->cfg.preview_slots = SERVER_DEFAULT_PREVIEW_SLOTS;
  if (s->cfg.preview_slots < 256 || s->cfg.preview_slots > (1u << 26))
    return mcr_fail(MCR_E_INVALID, "preview_slots must be in [256, 2^26]");
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  if (!s->cfg.preview_cus) s->cfg.preview_cus = (uint32_t)std::max(1, cus * 3 / 32);
  if (s->cfg.lanes == 2 && (int)s->cfg.preview_cus >= cus)
    return mcr_fail(MCR_E_INVALID, "preview_cus must leave CUs for the pass lane");
  int rc = MCR_OK;
  auto fail_free = [&](int code) {
    server_free(s.get());
    return code;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->admitted, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->read_ready, hipEventDisableTiming) != hipSuccess)
    return fail_free(mcr_fail(MCR_E_DEVICE, "server streams"));
  s->enc_cap = 1u << 16;
  s->views.assign(s->cfg.max_requests, ViewDev());
  memset(s->views.data(), 0, s->views.size() * sizeof(ViewDev));
  for (int i = (int)s->cfg.max_requests - 1; i >= 0; i--) s->view_free.push_back(i);
  if (s->cfg.lanes == 1) {
    s->lanes.emplace_back(new Lane());
    rc = lane_init(s.get(), *s->lanes[0], s->cfg.slots, s->cfg.window, true, true, 0, 0, cus);
  } else {
    // The preview lane on the last preview_cus CUs, the pass lane on the others.
    int pc = (int)s->cfg.preview_cus;
    s->lanes.emplace_back(new Lane());
    s->lanes.emplace_back(new Lane());
    // With preview_passes the preview lane also runs passes while no preview waits (its window
    // puts previews first), so its CUs do not idle between previews.
    rc = lane_init(s.get(), *s->lanes[0], s->cfg.preview_slots, s->cfg.preview_slots, true,
                   s->cfg.preview_passes > 0, cus - pc, pc, cus);
    if (!rc) rc = lane_init(s.get(), *s->lanes[1], s->cfg.slots, s->cfg.window, false, true, 0, cus - pc, cus);
    s->lanes[1]->launches_per_step = 2;
  }
  if (rc) return fail_free(rc);
  for (size_t i = 0; i < s->lanes.size(); i++) s->lanes[i]->bit = (uint8_t)(1u << i);
  // Tile results wait in the arena until encoded: each lane's windows plus tiles waiting for an
  // earlier pass, in one range per lane.
  std::vector<uint64_t> share;
  for (auto &l : s->lanes)
    share.push_back(std::max<uint64_t>((uint64_t)l->window * (l->iterations_per_step() + 2) * 4 * 2, 1ull << 24));
  s->arena_floats = 0;
  for (uint64_t n : share) s->arena_floats += n;
  if (hipMalloc(&s->arena_dev, s->arena_floats * sizeof(float)) != hipSuccess)
    return fail_free(mcr_fail(MCR_E_DEVICE, "server arena"));
  uint64_t at = 0;
  for (size_t i = 0; i < s->lanes.size(); i++) {
    s->lanes[i]->arena.reset(at, share[i]);
    at += share[i];
  }
  uint64_t fixed = (uint64_t)s->arena_floats * 4;
  s->bytes_used = fixed;
  if (s->bytes_used > s->cfg.hbm_budget)
    return fail_free(mcr_fail(MCR_E_OVERFLOW, "the HBM budget does not cover the server's own buffers"));
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
  server_free(s);
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
  s->scene_slots = std::max(s->scene_slots, slots_needed);
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
  const SrvRequest &r = *it->second;
  if (r.state == MCR_REQ_FAILED)
    return r.error ? mcr_fail(r.error, r.error_msg) : mcr_fail(s->fatal ? s->fatal : MCR_E_DEVICE, s->fatal_msg);
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
  // After everything queued on the server stream so far (the image's allocation and clearing)
  // and every encode launched; the image stays allocated while the lock is held.
  HIP_TRY(hipEventRecord(s->read_ready, s->stream));
  HIP_TRY(hipStreamWaitEvent(s->copy_stream, s->read_ready, 0));
  after_encodes(s, s->copy_stream);
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
  DevCounters c = s->counters();
  out->closest_queries = c.closest_queries;
  out->any_queries = c.any_queries;
  out->paths = s->paths;
  out->iterations = 0;
  for (const auto &l : s->lanes) out->iterations += l->iterations;
  out->shade_waves = c.shade_waves;
  out->mixed_view_waves = c.mixed_view_waves;
  out->bytes_in_use = s->bytes_used;
  out->requests_live = s->running.size();
  out->requests_queued = s->admit_queue.size();
  out->device_ms = s->device_ms;
  out->uptime_ms = s->now();
  out->host_ms = s->host_ms;
  out->plan_ms = s->plan_ms;
  out->deliver_ms = s->deliver_ms;
  out->pass_busy_ms = s->lanes.back()->busy_ms;
  return MCR_OK;
}

}  // extern "C"
