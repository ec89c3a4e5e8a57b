// trace.hip — wavefront path-tracing kernels for gfx950: ray generation, extend (closest hit),
// shade, shadow (any hit) and resolve, plus the ray-query and tile-encode kernels.
//
// The kernels restate src/tracer.ts (samplePixel, lightAt, visibleLight, transmittance,
// attenuateShadow, cosineDirection, reflection, refract, fog, background), src/bvh.ts
// (traversals) and src/progressive.ts:288-313 (writeTile) in FP64 with -ffp-contract=off, so
// each sample reproduces the reference's arithmetic; the order in which light terms are summed
// is kept by deferring the sum to the resolve kernel, after the shadow rays have run.
#include <hip/hip_runtime.h>

#include "../../include/mcr.h"
#include "device.h"
#include "kernels.h"
#include "shading.h"
#include "texture.h"
#include "wavefront.h"

__constant__ float g_dfg[512];
__constant__ float g_srgb8[256];

// ---- helpers ----------------------------------------------------------------------------------
// The light-slot arrays (unshadowed terms, shadow rays, shadow colours) are written by one kernel
// and read once by the next: streamed with non-temporal accesses, so they pass the caches by and
// leave them to the BVH and triangles the traversals re-read.
__device__ inline void nt_st(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ inline double nt_ld(const double *p) { return __builtin_nontemporal_load(p); }


// Block-wide queue appends. A same-address atomic per wave serialises thousands of waves on the
// counter's cache line (the resolve pass spent most of its time there); one atomic per block per
// counter cuts that by the block's wave count (WAVES: 4 for 256 threads, 16 for 1,024). Every
// thread of the block must call these; `lds` is shared memory of WAVES words per counter.
//
// block_reserve: the base of this wave's `total` entries (wave-uniform) on each of K counters.
template <int K, int WAVES = 4>
__device__ inline void block_reserve(uint32_t *const (&counter)[K], const uint32_t (&total)[K], uint32_t *lds,
                                     uint32_t (&base)[K]) {
  uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < K; k++) lds[WAVES * k + wv] = total[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t got[K];
    for (int k = 0; k < K; k++) {  // all K atomics in flight before any result is used
      uint32_t t = 0;
      for (int v = 0; v < WAVES; v++) t += lds[WAVES * k + v];
      got[k] = t ? atomicAdd(counter[k], t) : 0u;
    }
    for (int k = 0; k < K; k++) {
      uint32_t acc = got[k];
      for (int v = 0; v < WAVES; v++) {
        uint32_t x = lds[WAVES * k + v];
        lds[WAVES * k + v] = acc;
        acc += x;
      }
    }
  }
  __syncthreads();
  for (int k = 0; k < K; k++) base[k] = lds[WAVES * k + wv];
  __syncthreads();  // the words are reused by the next call
}

// block_append: one entry for each thread with `want`, on each of K counters; returns positions.
template <int K, int WAVES = 4>
__device__ inline void block_append(uint32_t *const (&counter)[K], const bool (&want)[K], uint32_t *lds,
                                    uint32_t (&pos)[K]) {
  uint64_t below = (1ull << (threadIdx.x & 63)) - 1ull;
  uint64_t m[K];
  uint32_t total[K];
  for (int k = 0; k < K; k++) {
    m[k] = __ballot(want[k]);
    total[k] = (uint32_t)__popcll(m[k]);
  }
  block_reserve<K, WAVES>(counter, total, lds, pos);
  for (int k = 0; k < K; k++) pos[k] += (uint32_t)__popcll(m[k] & below);
}

__device__ inline int find_job(const JobDev *jobs, uint32_t njobs, uint32_t s) {
  int lo = 0, hi = (int)njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].sample_start <= s) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// find_job with a guess: `hint` is returned when the table says it is the job holding sample s
// (the last job whose first sample is at or before s), else the table is searched. Both checks
// load at once, so a right guess costs one memory round trip instead of a search.
__device__ inline int job_of(const JobDev *jobs, uint32_t njobs, uint32_t s, int hint) {
  if (njobs <= 1) return 0;
  uint32_t h = min((uint32_t)max(hint, 0), njobs - 1), h1 = min(h + 1, njobs - 1);
  uint32_t a = jobs[h].sample_start, b = jobs[h1].sample_start;
  if (hint >= 0 && a <= s && (h1 == h || b > s)) return (int)h;
  return find_job(jobs, njobs, s);
}

// The job hint for samples handed out to a wave in one consecutive range: one search for the
// range's first sample (the first active lane's), wave-uniform, so it runs on the scalar unit.
__device__ inline int wave_job_hint(const JobDev *jobs, uint32_t njobs, uint32_t s) {
  return find_job(jobs, njobs, __builtin_amdgcn_readfirstlane(s));
}

// tracer.ts:715-727
__device__ inline void unproject(const mcr_camera &c, double x, double y, double z, double *out) {
  const double *p = c.projection_inverse;
  double w = p[3] * x + p[7] * y + p[11] * z + p[15];
  double vx = (p[0] * x + p[4] * y + p[8] * z + p[12]) / w;
  double vy = (p[1] * x + p[5] * y + p[9] * z + p[13]) / w;
  double vz = (p[2] * x + p[6] * y + p[10] * z + p[14]) / w;
  const double *m = c.world;
  out[0] = m[0] * vx + m[4] * vy + m[8] * vz + m[12];
  out[1] = m[1] * vx + m[5] * vy + m[9] * vz + m[13];
  out[2] = m[2] * vx + m[6] * vy + m[10] * vz + m[14];
}

// Runs f(view) for this lane. MIX = false: every sample of the run has view 0, which the kernel
// reads once into scalar registers. MIX = true: each lane reads its own view; the lanes of a wave
// mostly share one, so the reads hit the same cache lines.
template <bool MIX, class F>
MCR_INLINE void with_view(const ViewDev *__restrict__ views, uint32_t v, F &&f) {
  f(views[MIX ? v : 0]);
}

// Adds one finished sample per active lane to done[job], one atomic per distinct job in the wave.
// Every lane of the wave must call it.
__device__ inline void count_done(uint32_t *done, int job, bool active) {
  uint64_t pending = __ballot(active);
  while (pending) {
    int leader = __ffsll((unsigned long long)pending) - 1;
    int jl = __shfl(job, leader, 64);
    uint64_t same = __ballot(active && job == jl);
    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&done[jl], (uint32_t)__popcll(same));
    if (job == jl) active = false;
    pending &= ~same;
  }
}

// ---- ray generation: tracer.ts:62-80 + 705-713 --------------------------------------------------
// The screen position (sx, sy) of sample s of job j: its pixel, block centre or jittered position,
// and the random stream after the jitter (tracer.ts:66-78).
// Block rows per band of a job's sample order.
#ifndef MCR_SAMPLE_BAND
#define MCR_SAMPLE_BAND 4
#endif

// The pixel offset (x, y) in its tile of sample li of job j. Samples run through the tile in bands
// of MCR_SAMPLE_BAND block rows, each band column by column, so 64 consecutive samples (a wave, as
// slots are filled) cover a 16 x 4 patch of neighbouring pixels rather than a strip of one row:
// their rays, and the shadow rays from their hits, traverse more of the same nodes (bands of 2, 3,
// 4 and 8 rows: 4 was fastest on the 1080p frame). The order only decides which samples run side
// by side; a sample's seed and output position are its pixel's.
MCR_INLINE void job_pixel(const JobDev &j, uint32_t li, uint32_t &x, uint32_t &y) {
  const uint32_t B = MCR_SAMPLE_BAND;
  uint32_t band = li / (j.cols * B), r = li - band * (j.cols * B);
  uint32_t bh = min(B, j.rows - band * B);
  uint32_t cx = r / bh;
  x = cx * j.block;
  y = (band * B + (r - cx * bh)) * j.block;
}

MCR_INLINE void sample_screen(const JobDev &j, uint32_t s, uint32_t seed, double &sx, double &sy, uint32_t &rng) {
  uint32_t li = s - j.sample_start;
  uint32_t x, y;
  job_pixel(j, li, x, y);
  int px = j.tx + (int)x, py = j.ty + (int)y;
  rng = rng_seed((uint32_t)px, (uint32_t)py, j.pass, seed);
  double jx = 0.5, jy = 0.5;
  if (j.block > 1) {
    jx = js_min((double)j.block, (double)j.width - px) * 0.5;
    jy = js_min((double)j.block, (double)j.height - py) * 0.5;
  } else if (j.pass > 0) {
    jx = rng_next(rng);
    jy = rng_next(rng);
  }
  sx = (px + jx) / j.width;
  sy = (py + jy) / j.height;
}

// Starts sample s of job j (its pixel, pass and block) in path slot `slot`. A sample carries its
// own seed and output position, so which slot runs it, and when, never changes a result.
MCR_INLINE void start_sample(const Wavefront &w, uint32_t slot, uint32_t s, const JobDev &j, int job,
                                    const ViewDev &V) {
  const mcr_camera &cam = V.cam;
  double sx, sy;
  uint32_t rng;
  sample_screen(j, s, V.st.seed, sx, sy, rng);
  double nx = sx * 2 - 1, ny = 1 - sy * 2;
  double a[3], b[3];
  unproject(cam, nx, ny, -1, a);
  unproject(cam, nx, ny, 1, b);
  double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  double len = or_one(js_hypot3(dx, dy, dz));
  dx /= len;
  dy /= len;
  dz /= len;
  size_t P = w.capacity;
  double *f = w.f;
  f[F_OX * P + slot] = a[0];
  f[F_OY * P + slot] = a[1];
  f[F_OZ * P + slot] = a[2];
  f[F_DX * P + slot] = dx;
  f[F_DY * P + slot] = dy;
  f[F_DZ * P + slot] = dz;
  f[F_TR * P + slot] = 1;
  f[F_TG * P + slot] = 1;
  f[F_TB * P + slot] = 1;
  f[F_LR * P + slot] = 0;
  f[F_LG * P + slot] = 0;
  f[F_LB * P + slot] = 0;
  f[F_MR * P + slot] = 0;
  f[F_MG * P + slot] = 0;
  f[F_MB * P + slot] = 0;
  f[F_FOG * P + slot] = 0;
  f[F_ENVS * P + slot] = 1;
  int32_t *iv = w.i;
  iv[I_RNG * P + slot] = (int32_t)rng;
  iv[I_FLAGS * P + slot] = PF_TONEMAP | PF_CULL;
  iv[I_BOUNCE * P + slot] = 0;
  iv[I_PASSES * P + slot] = 0;
  iv[I_SAMPLE * P + slot] = (int32_t)s;
  iv[I_VIEW * P + slot] = (int32_t)j.view;
  iv[I_ENV * P + slot] = -1;
  iv[I_JOB * P + slot] = job;
}

// Starts sample s in `slot`: finds its job (starting from `hint`), then runs start_sample under
// the job's view.
template <bool MIX>
MCR_INLINE void start_sample_any(const ViewDev *__restrict__ views, const Wavefront &w,
                                        const Control &ctl, uint32_t slot, uint32_t s, int hint) {
  int job = job_of(ctl.jobs, ctl.njobs, s, hint);
  const JobDev &j = ctl.jobs[job];
  with_view<MIX>(views, j.view, [&](const ViewDev &V) { start_sample(w, slot, s, j, job, V); });
}


// ---- extend: closest hit for every live slot (bvh.ts:202-282) ------------------------------------
// WIDE: the precision form's FP32 query over the scene's 8-wide BVH (wide.h) instead of the FP64
// binary traversal; the ray is rounded to FP32 and the hit widened back to FP64 for shading.
template <bool MIX, bool WIDE, class Stack>
MCR_INLINE void extend_loop(const ViewDev *__restrict__ views, const Wavefront &w, int q, Stack &st,
                            DevCounters *cnt) {
  uint32_t n = w.counts[cw(q)];
  TravCount tc = {0, 0, 0};
  unsigned long long queries = 0;
  bool bad = false;
  size_t P = w.capacity;
  const double *f = w.f;
  // Waves take 64 rays at a time from a shared counter, so no wave idles while others still
  // hold long traversals (a static split leaves the slowest waves running alone).
  WaveFetch fetch(&w.counts[cw(C_EXTEND_FETCH)], n);
  for (uint32_t b = fetch.next(); b < n; b = fetch.next()) {
    uint32_t i = b + (threadIdx.x & 63);
    if (i >= n) continue;
    int slot = w.queue[q][i];
    double ox = f[F_OX * P + slot], oy = f[F_OY * P + slot], oz = f[F_OZ * P + slot];
    double dx = f[F_DX * P + slot], dy = f[F_DY * P + slot], dz = f[F_DZ * P + slot];
    int flags = w.i[I_FLAGS * P + slot];
    TriHit h;
    with_view<MIX>(views, (uint32_t)w.i[I_VIEW * P + slot], [&](const ViewDev &V) {
      if constexpr (WIDE) {
        WideRay wr;
        WideHit wh;
        wh.tri = -1;
        wh.t = INFINITY;
        wh.u = wh.v = 0;
        wh.front = true;
        if (wide_ray((float)ox, (float)oy, (float)oz, (float)dx, (float)dy, (float)dz, wr))
          wide_closest(V.wide, wr, 0.0f, INFINITY, (flags & PF_CULL) != 0, wh, st, tc, bad);
        h.tri = wh.tri;
        h.t = wh.tri >= 0 ? (double)wh.t : JM_INF;
        h.u = wh.u;
        h.v = wh.v;
        h.front = wh.front;
        if (wh.tri >= 0) refine_hit(V.sc.tri, wh.tri, ox, oy, oz, dx, dy, dz, h.t, h.u, h.v);
      } else {
        intersect_closest(V.sc, ox, oy, oz, dx, dy, dz, 0, JM_INF, (flags & PF_CULL) != 0, h, st, tc);
      }
    });
    queries++;
    w.i[I_HTRI * P + slot] = h.tri;
    w.f[F_HT * P + slot] = h.t;
    w.f[F_HU * P + slot] = h.u;
    w.f[F_HV * P + slot] = h.v;
    w.i[I_FLAGS * P + slot] = h.front ? (flags | PF_FRONT) : (flags & ~PF_FRONT);
  }
  if (bad) atomicOr(&w.counts[cw(C_OVERFLOW)], OVERFLOW_WIDE);
  flush_counts(cnt, tc, queries, 0);
}

template <bool MIX>
__global__ __launch_bounds__(256, MIX ? 4 : 5) void k_extend(const ViewDev *__restrict__ views, Wavefront w, int q,
                                                int32_t *spill, uint32_t spill_stride,
                                                DevCounters *cnt) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < 2 * MCR_SHADOW_REGIONS) w.counts[cw(C_SHADOW_SEG + gid)] = 0;  // shade of this iteration appends
  if (gid < MCR_SHADE_CLASSES) w.counts[cw(C_BIN + gid)] = 0;                  // its shade bins
  if (gid == 0) w.counts[cw(C_DEEP)] = 0;                                       // shadow of this iteration marks
  MCR_LANE_STACK(st, spill, spill_stride);
  extend_loop<MIX, false>(views, w, q, st, cnt);
}

template <bool MIX>
__global__ __launch_bounds__(256) void k_extend_wide(const ViewDev *__restrict__ views, Wavefront w, int q,
                                                        int32_t *spill, uint32_t spill_stride,
                                                        DevCounters *cnt) {
  __shared__ uint32_t lds_base[MCR_WIDE_STACK_LDS * 256], lds_bits[MCR_WIDE_STACK_LDS * 256];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < 2 * MCR_SHADOW_REGIONS) w.counts[cw(C_SHADOW_SEG + gid)] = 0;
  if (gid < MCR_SHADE_CLASSES) w.counts[cw(C_BIN + gid)] = 0;
  if (gid == 0) w.counts[cw(C_DEEP)] = 0;
  WideStack st{lds_base + threadIdx.x, lds_bits + threadIdx.x, spill + gid, (int)spill_stride, 256, 0};
  extend_loop<MIX, true>(views, w, q, st, cnt);
}

// Writes one finished sample of job j (RGBA, doubles rounded to float as Float32Array.set does)
// over its block of the tile (tracer.ts:81-85).
MCR_INLINE void write_sample(const JobDev &j, float *out, uint32_t s, const double *c) {
  uint32_t li = s - j.sample_start;
  uint32_t x, y;
  job_pixel(j, li, x, y);
  uint32_t ymax = min(y + j.block, (uint32_t)j.th), xmax = min(x + j.block, (uint32_t)j.tw);
  float4 v = make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[3]);
  for (uint32_t by = y; by < ymax; by++)
    for (uint32_t bx = x; bx < xmax; bx++)
      *reinterpret_cast<float4 *>(out + j.out_offset + ((size_t)by * j.tw + bx) * 4) = v;
}

struct LightSample {
  double lx, ly, lz, lr, lg, lb, dist;
};

// tracer.ts:486-512
MCR_INLINE bool light_at(const DevLight &l, double px, double py, double pz, LightSample &o) {
  if (l.type == MCR_LIGHT_DIRECTIONAL) {
    o.lx = l.direction[0];
    o.ly = l.direction[1];
    o.lz = l.direction[2];
    o.lr = l.color[0];
    o.lg = l.color[1];
    o.lb = l.color[2];
    o.dist = JM_INF;
    return true;
  }
  if (l.type == MCR_LIGHT_POINT || l.type == MCR_LIGHT_SPOT) {
    double x = l.position[0] - px, y = l.position[1] - py, z = l.position[2] - pz;
    double distance = js_hypot3(x, y, z);
    if (distance == 0) return false;
    o.lx = x / distance;
    o.ly = y / distance;
    o.lz = z / distance;
    double scale = distance_attenuation(distance, l.distance, l.decay);
    if (l.type == MCR_LIGHT_SPOT) {
      double ac = o.lx * l.direction[0] + o.ly * l.direction[1] + o.lz * l.direction[2];
      scale *= smoothstep_js(l.cone_cos, l.penumbra_cos, ac);
    }
    if (scale <= 0) return false;
    o.lr = l.color[0] * scale;
    o.lg = l.color[1] * scale;
    o.lb = l.color[2] * scale;
    o.dist = distance;
    return true;
  }
  return false;
}

// background (tracer.ts:659-691) for a primary ray that met no surface: colour, screen texture,
// environment or the clear colour. Screen and environment backgrounds are tone mapped unless their
// texture holds sRGB-encoded data.
MCR_INLINE void background(const ViewDev &V, const Control &ctl, uint32_t smp, int job, double dx, double dy,
                           double dz, double *c) {
  const DevScene &sc = V.sc;
  const mcr_settings &st = V.st;
  if (sc.background_kind == MCR_BACKGROUND_COLOR) {
    c[0] = sc.background_color[0];
    c[1] = sc.background_color[1];
    c[2] = sc.background_color[2];
    c[3] = 1;
  } else if (sc.background_kind == MCR_BACKGROUND_SCREEN || sc.background_kind == MCR_BACKGROUND_ENVIRONMENT) {
    bool srgb;
    if (sc.background_kind == MCR_BACKGROUND_SCREEN) {
      const DevTexture &tx = sc.textures[sc.background_texture];
      double sx, sy, t4[4];
      uint32_t rng;
      sample_screen(ctl.jobs[job], smp, st.seed, sx, sy, rng);
      tex_sample(tx, sx, 1 - sy, t4);
      for (int i = 0; i < 3; i++) c[i] = t4[i] * sc.background_intensity;
      srgb = tx.srgb != 0;
    } else {
      const DevEnvironment &env = sc.environments[sc.background_environment];
      env_sample(env, sc.textures, dx, dy, dz, c);
      srgb = sc.textures[env.kind == MCR_ENV_CUBE ? env.faces[0] : env.texture].srgb != 0;
    }
    if (!srgb && st.tone_mapping != MCR_NO_TONE_MAPPING) tone_map(c, st.tone_mapping, st.tone_mapping_exposure);
    c[3] = 1;
  } else {
    double a = st.clear_alpha;
    c[0] = st.clear_color[0] * a;
    c[1] = st.clear_color[1] * a;
    c[2] = st.clear_color[2] * a;
    c[3] = a;
  }
}

// ---- shade: one path vertex of samplePixel (tracer.ts:113-469) ----------------------------------
// Writes the vertex's unshadowed light terms and shadow rays into the light slots, the values
// resolve needs to finish `lr += tr * (outR + er)`, and the continuation ray (the continuation
// uses only the random stream and the material, never the shadow results, so it runs here).
// A path whose ray met nothing: the camera ray's background, written as the sample, or the last
// surface's environment added to the radiance before the sample is written (tracer.ts:119-124).
// Returns the slot's PathMode.
MCR_INLINE int shade_miss(const ViewDev &V, const Wavefront &w, const Control &ctl, int slot, int flags,
                          double dx, double dy, double dz, double tr, double tg, double tb) {
  size_t P = w.capacity;
  if (!(flags & PF_SURFACE)) {
    double c[4];
    uint32_t smp = (uint32_t)w.i[I_SAMPLE * P + slot];
    int job = job_of(ctl.jobs, ctl.njobs, smp, w.i[I_JOB * P + slot]);
    background(V, ctl, smp, job, dx, dy, dz, c);
    write_sample(ctl.jobs[job], ctl.out, smp, c);
    return MODE_DONE;
  }
  int le = w.i[I_ENV * P + slot];
  if (le >= 0) {
    double env[3];
    double scale = w.f[F_ENVS * P + slot];
    env_sample(V.sc.environments[le], V.sc.textures, dx, dy, dz, env);
    w.f[F_LR * P + slot] += tr * env[0] * scale;
    w.f[F_LG * P + slot] += tg * env[1] * scale;
    w.f[F_LB * P + slot] += tb * env[2] * scale;
  }
  return MODE_FINISH;
}

MCR_INLINE uint32_t shade_slot(const ViewDev &V, const Wavefront &w, const Control &ctl,
                                     int slot) {
  const DevScene &sc = V.sc;
  const mcr_settings &st = V.st;
  const mcr_camera &cam = V.cam;
  size_t P = w.capacity;
  double *f = w.f;
  int32_t *iv = w.i;
  uint32_t nshadow = 0;
  const uint32_t nl = V.light_slots;
  {
    double eps = sc.epsilon;
    int flags = iv[I_FLAGS * P + slot];
    uint32_t rng = (uint32_t)iv[I_RNG * P + slot];
    int bounce = iv[I_BOUNCE * P + slot];
    int passes = iv[I_PASSES * P + slot];
    double ox = f[F_OX * P + slot], oy = f[F_OY * P + slot], oz = f[F_OZ * P + slot];
    double dx = f[F_DX * P + slot], dy = f[F_DY * P + slot], dz = f[F_DZ * P + slot];
    double tr = f[F_TR * P + slot], tg = f[F_TG * P + slot], tb = f[F_TB * P + slot];
    int k = iv[I_HTRI * P + slot];
    int mode;
    for (uint32_t l = 0; l < V.shadow_slots; l++) w.status[(size_t)l * P + slot] = LS_NONE;
    if (k < 0) {
      mode = shade_miss(V, w, ctl, slot, flags, dx, dy, dz, tr, tg, tb);
    } else {
      double t = f[F_HT * P + slot];
      double mr = f[F_MR * P + slot], mg = f[F_MG * P + slot], mb = f[F_MB * P + slot];
      if (mr > 0 || mg > 0 || mb > 0) {  // Beer-Lambert (tracer.ts:128-131)
        tr *= jm_exp(-mr * t);
        tg *= jm_exp(-mg * t);
        tb *= jm_exp(-mb * t);
      }
      bool front = (flags & PF_FRONT) != 0;
      const float4 *rec = sc.tri + (size_t)k * 3;
      float4 ra = rec[0], rb = rec[1], rc = rec[2];
      int tflags = (int)((__float_as_uint(rc.y) >> 8) & 0xffu);
      const DevMaterial &mat = sc.materials[__float_as_uint(rc.z)];
      double px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t;
      double w1 = f[F_HU * P + slot], w2 = f[F_HV * P + slot], w0 = 1 - w1 - w2;
      double e1x = ra.w, e1y = rb.x, e1z = rb.y, e2x = rb.z, e2y = rb.w, e2z = rc.x;
      double gx = e1y * e2z - e1z * e2y;
      double gy = e1z * e2x - e1x * e2z;
      double gz = e1x * e2y - e1y * e2x;
      double gl = or_one(js_hypot3(gx, gy, gz));
      double face = front ? 1 : -1;
      gx = (gx / gl) * face;
      gy = (gy / gl) * face;
      gz = (gz / gl) * face;
      // Texture coordinates (tracer.ts:151-153), read only by textured materials.
      double tu = 0, tv = 0, texel[4];
      if (mat.textured) {
        const float *uvs = sc.uvs + (size_t)k * 6;
        tu = uvs[0] * w0 + uvs[2] * w1 + uvs[4] * w2;
        tv = uvs[1] * w0 + uvs[3] * w1 + uvs[5] * w2;
      }
      double cr = mat.color[0], cg = mat.color[1], cb = mat.color[2];
      double alpha = mat.opacity;
      if (mat.map >= 0) {
        tex_sample(sc.textures[mat.map], tu, tv, texel);
        cr *= texel[0];
        cg *= texel[1];
        cb *= texel[2];
        alpha *= texel[3];
      }
      if (mat.vertex_colors && sc.colors) {
        const float *col = sc.colors + (size_t)k * 12;
        cr *= col[0] * w0 + col[4] * w1 + col[8] * w2;
        cg *= col[1] * w0 + col[5] * w1 + col[9] * w2;
        cb *= col[2] * w0 + col[6] * w1 + col[10] * w2;
        alpha *= col[3] * w0 + col[7] * w1 + col[11] * w2;
      }
      if (mat.alpha_map >= 0) {
        tex_sample(sc.textures[mat.alpha_map], tu, tv, texel);
        alpha *= texel[1];
      }
      bool cutOut = mat.alpha_test > 0 && alpha < mat.alpha_test;
      if (cutOut || (mat.transparent && rng_next(rng) >= alpha)) {
        // pass straight through (tracer.ts:170-176)
        if (++passes > 64) {
          mode = MODE_FINISH;
        } else {
          double side = gx * dx + gy * dy + gz * dz > 0 ? eps : -eps;
          f[F_OX * P + slot] = px + gx * side;
          f[F_OY * P + slot] = py + gy * side;
          f[F_OZ * P + slot] = pz + gz * side;
          mode = MODE_EXTEND;
        }
      } else {
        if (!(flags & PF_SURFACE)) {
          flags |= PF_SURFACE;
          flags = mat.tone_mapped ? (flags | PF_TONEMAP) : (flags & ~PF_TONEMAP);
          double fogf = 0;
          if (sc.fog_kind != MCR_FOG_NONE) {  // tracer.ts:649-656
            const double *fw = cam.forward, *cp = cam.position;
            double depth = (px - cp[0]) * fw[0] + (py - cp[1]) * fw[1] + (pz - cp[2]) * fw[2];
            if (sc.fog_kind == MCR_FOG_LINEAR) fogf = smoothstep_js(sc.fog_near, sc.fog_far, depth);
            else fogf = 1 - jm_exp(-sc.fog_density * sc.fog_density * depth * depth);
          }
          f[F_FOG * P + slot] = fogf;
        }
        const float *nrm = sc.normals + (size_t)k * 9;
        double nx = (nrm[0] * w0 + nrm[3] * w1 + nrm[6] * w2) * face;
        double ny = (nrm[1] * w0 + nrm[4] * w1 + nrm[7] * w2) * face;
        double nz = (nrm[2] * w0 + nrm[5] * w1 + nrm[8] * w2) * face;
        double nlen = or_one(js_hypot3(nx, ny, nz));
        nx /= nlen;
        ny /= nlen;
        nz /= nlen;
        if (mat.normal_map >= 0) {  // tangent-space normal map (tracer.ts:191-209)
          tex_sample(sc.textures[mat.normal_map], tu, tv, texel);
          double mx = (texel[0] * 2 - 1) * mat.normal_scale[0];
          double my = (texel[1] * 2 - 1) * mat.normal_scale[1];
          double mz = texel[2] * 2 - 1;
          const float *tg4 = sc.tangents + (size_t)k * 4;
          double ax = tg4[0], ay = tg4[1], az = tg4[2];
          double along = ax * nx + ay * ny + az * nz;
          ax -= nx * along;
          ay -= ny * along;
          az -= nz * along;
          double al = or_one(js_hypot3(ax, ay, az));
          ax = (ax / al) * face;
          ay = (ay / al) * face;
          az = (az / al) * face;
          double sign = tg4[3];
          double bx = (ny * az - nz * ay) * sign, by = (nz * ax - nx * az) * sign, bz = (nx * ay - ny * ax) * sign;
          double qx = ax * mx + bx * my + nx * mz;
          double qy = ay * mx + by * my + ny * mz;
          double qz = az * mx + bz * my + nz * mz;
          nlen = or_one(js_hypot3(qx, qy, qz));
          nx = qx / nlen;
          ny = qy / nlen;
          nz = qz / nlen;
        }
        double vx = -dx, vy = -dy, vz = -dz;
        double dotNV = js_min(js_max(nx * vx + ny * vy + nz * vz, 0), 1);
        double aR = 0, aG = 0, aB = 0;  // out* before the light sum (basic: the colour)
        double iR = 0, iG = 0, iB = 0;  // indirect term added after it
        bool has_ind = false;
        bool receive = (tflags & MCR_TRI_RECEIVE_SHADOW) != 0;
        bool canContinue = bounce < st.max_bounces;
        double specR = 0, specG = 0, specB = 0, diffR = 0, diffG = 0, diffB = 0;
        double transR = 0, transG = 0, transB = 0;
        double roughness = 1;
        bool castOnly = st.shadows == MCR_SHADOWS_CAST;
        int environment = -1;  // physical materials: envMap, else the scene's (tracer.ts:333-340)
        double environmentScale = 1;
        int combine = 0;
        if (mat.model == MCR_MODEL_BASIC) {
          aR = cr;
          aG = cg;
          aB = cb;
        } else {
          double ir = 0, ig = 0, ib = 0;
          for (uint32_t li = 0; li < sc.light_count; li++) {
            const DevLight &L = sc.lights[li];
            if (L.type == MCR_LIGHT_AMBIENT) {
              ir += L.color[0];
              ig += L.color[1];
              ib += L.color[2];
            } else if (L.type == MCR_LIGHT_HEMISPHERE) {
              double hw = 0.5 * (nx * L.direction[0] + ny * L.direction[1] + nz * L.direction[2]) + 0.5;
              ir += L.ground_color[0] + (L.color[0] - L.ground_color[0]) * hw;
              ig += L.ground_color[1] + (L.color[1] - L.ground_color[1]) * hw;
              ib += L.ground_color[2] + (L.color[2] - L.ground_color[2]) * hw;
            }
          }
          double ao = 1;
          if (mat.ao_map >= 0) {
            tex_sample(sc.textures[mat.ao_map], tu, tv, texel);
            ao = (texel[0] - 1) * mat.ao_map_intensity + 1;
          }
          bool physicalModel = mat.model == MCR_MODEL_PHYSICAL;
          double metalness = 0, sr = 0.04, sg = 0.04, sb = 0.04, f90 = 1, br = 0, bg = 0, bbl = 0;
          double drr = 0, dgg = 0, dbb = 0, fa = 0, fb = 0, ems = 0, compR = 0, compG = 0, compB = 0;
          double ggxAlpha = 0, keepDiffuse = 1;
          if (physicalModel) {
            metalness = mat.metalness;
            roughness = mat.roughness;
            if (mat.roughness_map >= 0) {
              tex_sample(sc.textures[mat.roughness_map], tu, tv, texel);
              roughness *= texel[1];
            }
            if (mat.metalness_map >= 0) {
              tex_sample(sc.textures[mat.metalness_map], tu, tv, texel);
              metalness *= texel[2];
            }
            double shadeRoughness = js_min(js_max(roughness, 0.045), 1);
            if (mat.physical) {
              double fr = jm_pow((mat.ior - 1) / (mat.ior + 1), 2);
              sr = js_min(fr * mat.specular_color[0], 1) * mat.specular_intensity;
              sg = js_min(fr * mat.specular_color[1], 1) * mat.specular_intensity;
              sb = js_min(fr * mat.specular_color[2], 1) * mat.specular_intensity;
              f90 = mat.specular_intensity + (1 - mat.specular_intensity) * metalness;
            }
            br = sr + (cr - sr) * metalness;
            bg = sg + (cg - sg) * metalness;
            bbl = sb + (cb - sb) * metalness;
            drr = cr * (1 - metalness);
            dgg = cg * (1 - metalness);
            dbb = cb * (1 - metalness);
            dfg(g_dfg, shadeRoughness, dotNV, fa, fb);
            double essMs = fa + fb;
            ems = 1 - essMs;
            compR = 1 + br * (1 / essMs - 1);
            compG = 1 + bg * (1 / essMs - 1);
            compB = 1 + bbl * (1 / essMs - 1);
            ggxAlpha = shadeRoughness * shadeRoughness;
            keepDiffuse = 1 - mat.transmission;
          }
          // Direct lights (tracer.ts:286-312 physical, 357-382 lambert / phong / toon).
          for (uint32_t li = 0; li < nl; li++) {
            const DevLight &L = sc.lights[li];
            LightSample ls;
            if (!light_at(L, px, py, pz, ls)) continue;
            double r, g, b;
            if (physicalModel) {
              double dotNL = nx * ls.lx + ny * ls.ly + nz * ls.lz;
              if (dotNL <= 0) continue;
              double hx = ls.lx + vx, hy = ls.ly + vy, hz = ls.lz + vz;
              double hl = or_one(js_hypot3(hx, hy, hz));
              hx /= hl;
              hy /= hl;
              hz /= hl;
              double dotNH = js_min(js_max(nx * hx + ny * hy + nz * hz, 0), 1);
              double dotVH = js_min(js_max(vx * hx + vy * hy + vz * hz, 0), 1);
              double fw = fresnel_weight(dotVH);
              double vd = visibility_ggx(ggxAlpha, dotNL, dotNV) * distribution_ggx(ggxAlpha, dotNH);
              double specularR = (br * (1 - fw) + f90 * fw) * vd * compR;
              double specularG = (bg * (1 - fw) + f90 * fw) * vd * compG;
              double specularB = (bbl * (1 - fw) + f90 * fw) * vd * compB;
              double diffuseR = drr * MCR_INV_PI * (1 - (sr * (1 - fw) + f90 * fw)) * keepDiffuse;
              double diffuseG = dgg * MCR_INV_PI * (1 - (sg * (1 - fw) + f90 * fw)) * keepDiffuse;
              double diffuseB = dbb * MCR_INV_PI * (1 - (sb * (1 - fw) + f90 * fw)) * keepDiffuse;
              double scale = dotNL;
              r = (specularR + diffuseR) * scale * ls.lr;
              g = (specularG + diffuseG) * scale * ls.lg;
              b = (specularB + diffuseB) * scale * ls.lb;
            } else {
              double cosine = nx * ls.lx + ny * ls.ly + nz * ls.lz;
              double irradiance;
              if (mat.model == MCR_MODEL_TOON) irradiance = cosine * 0.5 + 0.5 < 0.7 ? 0.7 : 1;
              else if (cosine <= 0) continue;
              else irradiance = cosine;
              r = cr * MCR_INV_PI;
              g = cg * MCR_INV_PI;
              b = cb * MCR_INV_PI;
              if (mat.model == MCR_MODEL_PHONG && cosine > 0) {
                double hx = ls.lx + vx, hy = ls.ly + vy, hz = ls.lz + vz;
                double hl = or_one(js_hypot3(hx, hy, hz));
                hx /= hl;
                hy /= hl;
                hz /= hl;
                double dotNH = js_min(js_max(nx * hx + ny * hy + nz * hz, 0), 1);
                double dotVH = js_min(js_max(vx * hx + vy * hy + vz * hz, 0), 1);
                double fw = fresnel_weight(dotVH);
                double d = blinn_phong(mat.shininess, dotNH);
                r += (mat.specular[0] * (1 - fw) + fw) * d;
                g += (mat.specular[1] * (1 - fw) + fw) * d;
                b += (mat.specular[2] * (1 - fw) + fw) * d;
              }
              r *= irradiance * ls.lr;
              g *= irradiance * ls.lg;
              b *= irradiance * ls.lb;
            }
            if (r + g + b <= 0) continue;
            size_t ls_idx = (size_t)li * P + slot;
            const size_t RS = (size_t)w.lights * P;  // one row of light slots
            nt_st(&w.contrib[0 * RS + ls_idx], r);
            nt_st(&w.contrib[1 * RS + ls_idx], g);
            nt_st(&w.contrib[2 * RS + ls_idx], b);
            // visibleLight (tracer.ts:515-526); 'castShadow' in light holds for these three types.
            if (st.shadows == MCR_SHADOWS_NONE || (castOnly && (!L.cast_shadow || !receive))) {
              nt_st(&w.shadow[0 * RS + ls_idx], 1);
              nt_st(&w.shadow[1 * RS + ls_idx], 1);
              nt_st(&w.shadow[2 * RS + ls_idx], 1);
              w.status[ls_idx] = LS_VISIBLE;
            } else {
              double side = gx * ls.lx + gy * ls.ly + gz * ls.lz > 0 ? eps : -eps;
              nt_st(&w.sray[0 * RS + ls_idx], px + gx * side);
              nt_st(&w.sray[1 * RS + ls_idx], py + gy * side);
              nt_st(&w.sray[2 * RS + ls_idx], pz + gz * side);
              nt_st(&w.sray[3 * RS + ls_idx], ls.lx);
              nt_st(&w.sray[4 * RS + ls_idx], ls.ly);
              nt_st(&w.sray[5 * RS + ls_idx], ls.lz);
              nt_st(&w.sray[6 * RS + ls_idx], ls.dist == JM_INF ? JM_INF : ls.dist - 2 * eps);
              w.status[ls_idx] = LS_RAY;
              nshadow++;
            }
          }
          if (physicalModel) {
            double fssR = sr * fa + f90 * fb, fssG = sg * fa + f90 * fb, fssB = sb * fa + f90 * fb;
            double scatterR = fssR + multi_scatter(fssR, sr, ems);
            double scatterG = fssG + multi_scatter(fssG, sg, ems);
            double scatterB = fssB + multi_scatter(fssB, sb, ems);
            iR = ir * MCR_INV_PI * drr * (1 - scatterR) * ao * keepDiffuse;
            iG = ig * MCR_INV_PI * dgg * (1 - scatterG) * ao * keepDiffuse;
            iB = ib * MCR_INV_PI * dbb * (1 - scatterB) * ao * keepDiffuse;
            double mfR = cr * fa + f90 * fb, mfG = cg * fa + f90 * fb, mfB = cb * fa + f90 * fb;
            specR = fssR + (mfR - fssR) * metalness;
            specG = fssG + (mfG - fssG) * metalness;
            specB = fssB + (mfB - fssB) * metalness;
            diffR = drr * (1 - scatterR) * keepDiffuse;
            diffG = dgg * (1 - scatterG) * keepDiffuse;
            diffB = dbb * (1 - scatterB) * keepDiffuse;
            double transmission = mat.transmission;
            if (transmission > 0) {
              double fr = sr * fa + f90 * fb, fg = sg * fa + f90 * fb, fbl = sb * fa + f90 * fb;
              transR = transmission * (1 - fr) * cr;
              transG = transmission * (1 - fg) * cg;
              transB = transmission * (1 - fbl) * cb;
            }
            if (mat.env_map >= 0) {
              environment = mat.env_map;
            } else if (sc.environment >= 0) {
              environment = sc.environment;
              environmentScale = mat.env_map_intensity;
            }
            // Environment light on the diffuse layer, one cosine-weighted direction traced with
            // transmittance(..., castOnly false) in the slot after the lights (tracer.ts:341-353).
            if (environment >= 0 && !st.global_illumination) {
              double r1 = rng_next(rng), r2 = rng_next(rng);
              double phi = 2 * MCR_PI * r1;
              double radius = jm_sqrt(r2);
              double ex, ey, ez;
              to_world(nx, ny, nz, radius * jm_cos(phi), radius * jm_sin(phi), jm_sqrt(js_max(0, 1 - r2)), ex, ey, ez);
              if (ex * gx + ey * gy + ez * gz > 0) {
                double env[3];
                env_sample(sc.environments[environment], sc.textures, ex, ey, ez, env);
                const size_t RS = (size_t)w.lights * P;
                size_t ls_idx = (size_t)nl * P + slot;
                nt_st(&w.contrib[0 * RS + ls_idx], env[0] * environmentScale * diffR);
                nt_st(&w.contrib[1 * RS + ls_idx], env[1] * environmentScale * diffG);
                nt_st(&w.contrib[2 * RS + ls_idx], env[2] * environmentScale * diffB);
                nt_st(&w.sray[0 * RS + ls_idx], px + gx * eps);
                nt_st(&w.sray[1 * RS + ls_idx], py + gy * eps);
                nt_st(&w.sray[2 * RS + ls_idx], pz + gz * eps);
                nt_st(&w.sray[3 * RS + ls_idx], ex);
                nt_st(&w.sray[4 * RS + ls_idx], ey);
                nt_st(&w.sray[5 * RS + ls_idx], ez);
                nt_st(&w.sray[6 * RS + ls_idx], JM_INF);
                w.status[ls_idx] = LS_RAY;
                nshadow++;
              }
            }
            if (!st.global_illumination) diffR = diffG = diffB = 0;
          } else {
            iR = ir * MCR_INV_PI * cr * ao;
            iG = ig * MCR_INV_PI * cg * ao;
            iB = ib * MCR_INV_PI * cb * ao;
            if (st.global_illumination) {
              diffR = cr;
              diffG = cg;
              diffB = cb;
            }
          }
          has_ind = true;
        }
        // envMap on lambert, phong, toon and basic materials (tracer.ts:392-406): the reflected
        // sample and reflectivity, combined with out* in resolve after the light sum.
        if (mat.model != MCR_MODEL_PHYSICAL && mat.env_map >= 0 && mat.combine >= MCR_COMBINE_MULTIPLY &&
            mat.combine <= MCR_COMBINE_ADD) {
          double dn = dx * nx + dy * ny + dz * nz;
          double e[3];
          env_sample(sc.environments[mat.env_map], sc.textures, dx - 2 * dn * nx, dy - 2 * dn * ny, dz - 2 * dn * nz,
                     e);
          f[F_XR * P + slot] = e[0];
          f[F_XG * P + slot] = e[1];
          f[F_XB * P + slot] = e[2];
          f[F_XS * P + slot] = mat.reflectivity;
          combine = mat.combine - MCR_COMBINE_MULTIPLY + 1;
        }
        flags = (flags & ~PF_COMBINE_MASK) | (combine << PF_COMBINE_SHIFT);
        double er = mat.emissive[0], eg = mat.emissive[1], eb = mat.emissive[2];
        if (mat.emissive_map >= 0 && (er > 0 || eg > 0 || eb > 0)) {
          tex_sample(sc.textures[mat.emissive_map], tu, tv, texel);
          er *= texel[0];
          eg *= texel[1];
          eb *= texel[2];
        }
        f[F_QR * P + slot] = tr;
        f[F_QG * P + slot] = tg;
        f[F_QB * P + slot] = tb;
        f[F_AR * P + slot] = aR;
        f[F_AG * P + slot] = aG;
        f[F_AB * P + slot] = aB;
        f[F_IR * P + slot] = iR;
        f[F_IG * P + slot] = iG;
        f[F_IB * P + slot] = iB;
        f[F_ER * P + slot] = er;
        f[F_EG * P + slot] = eg;
        f[F_EB * P + slot] = eb;
        flags = has_ind ? (flags | PF_INDIRECT) : (flags & ~PF_INDIRECT);
        // Continuation (tracer.ts:412-469).
        mode = MODE_ADD_FINISH;
        if (canContinue) {
          double ws = luminance(specR, specG, specB);
          double wt = luminance(transR, transG, transB);
          double wd = luminance(diffR, diffG, diffB);
          double total = ws + wt + wd;
          if (total > 1e-6) {
            double pick = rng_next(rng) * total;
            double wr, wg, wb, p, sx = 0, sy = 0, sz = 0;
            bool stop = false;
            bool cull = (flags & PF_CULL) != 0;
            if (pick < ws) {
              p = ws / total;
              wr = specR;
              wg = specG;
              wb = specB;
              double hx = nx, hy = ny, hz = nz;  // reflection (tracer.ts:608-621)
              if (roughness > 0.02) {
                double a = roughness * roughness;
                double r1 = rng_next(rng), r2 = rng_next(rng);
                double phi = 2 * MCR_PI * r1;
                double cosT = jm_sqrt((1 - r2) / (1 + (a * a - 1) * r2));
                double sinT = jm_sqrt(js_max(0, 1 - cosT * cosT));
                to_world(nx, ny, nz, sinT * jm_cos(phi), sinT * jm_sin(phi), cosT, hx, hy, hz);
              }
              double dn = dx * hx + dy * hy + dz * hz;
              sx = dx - 2 * dn * hx;
              sy = dy - 2 * dn * hy;
              sz = dz - 2 * dn * hz;
              if (sx * gx + sy * gy + sz * gz <= 0) stop = true;
            } else if (pick < ws + wt) {
              p = wt / total;
              wr = transR;
              wg = transG;
              wb = transB;
              if (mat.thickness > 0) {
                double eta = front ? 1 / mat.ior : mat.ior;
                double cosi = -(dx * nx + dy * ny + dz * nz);  // refract (tracer.ts:624-633)
                double kk = 1 - eta * eta * (1 - cosi * cosi);
                if (!(kk < 0)) {
                  double s = eta * cosi - jm_sqrt(kk);
                  sx = eta * dx + s * nx;
                  sy = eta * dy + s * ny;
                  sz = eta * dz + s * nz;
                  double length = or_one(js_hypot3(sx, sy, sz));
                  sx /= length;
                  sy /= length;
                  sz /= length;
                  cull = !front;
                  f[F_MR * P + slot] = front ? mat.attenuation[0] : 0;
                  f[F_MG * P + slot] = front ? mat.attenuation[1] : 0;
                  f[F_MB * P + slot] = front ? mat.attenuation[2] : 0;
                } else {
                  double dn = dx * nx + dy * ny + dz * nz;  // total internal reflection
                  sx = dx - 2 * dn * nx;
                  sy = dy - 2 * dn * ny;
                  sz = dz - 2 * dn * nz;
                }
              } else {
                sx = dx;
                sy = dy;
                sz = dz;
              }
            } else {
              p = wd / total;
              wr = diffR;
              wg = diffG;
              wb = diffB;
              double r1 = rng_next(rng), r2 = rng_next(rng);  // cosineDirection (tracer.ts:600-605)
              double phi = 2 * MCR_PI * r1;
              double radius = jm_sqrt(r2);
              to_world(nx, ny, nz, radius * jm_cos(phi), radius * jm_sin(phi), jm_sqrt(js_max(0, 1 - r2)),
                       sx, sy, sz);
              if (sx * gx + sy * gy + sz * gz <= 0) stop = true;
            }
            if (!stop) {
              iv[I_ENV * P + slot] = environment;  // lastEnvironment (tracer.ts:427-458)
              f[F_ENVS * P + slot] = environmentScale;
              tr *= wr / p;
              tg *= wg / p;
              tb *= wb / p;
              double side = sx * gx + sy * gy + sz * gz > 0 ? eps : -eps;
              f[F_OX * P + slot] = px + gx * side;
              f[F_OY * P + slot] = py + gy * side;
              f[F_OZ * P + slot] = pz + gz * side;
              f[F_DX * P + slot] = sx;
              f[F_DY * P + slot] = sy;
              f[F_DZ * P + slot] = sz;
              bounce++;
              bool alive = true;
              if (bounce >= 3) {  // Russian roulette
                double qq = js_min(0.95, js_max3(tr, tg, tb));
                if (rng_next(rng) >= qq) {
                  alive = false;
                } else {
                  tr /= qq;
                  tg /= qq;
                  tb /= qq;
                }
              }
              flags = cull ? (flags | PF_CULL) : (flags & ~PF_CULL);
              if (alive) mode = MODE_ADD_EXTEND;
            }
          }
        }
      }
    }
    f[F_TR * P + slot] = tr;
    f[F_TG * P + slot] = tg;
    f[F_TB * P + slot] = tb;
    iv[I_FLAGS * P + slot] = flags;
    iv[I_RNG * P + slot] = (int32_t)rng;
    iv[I_BOUNCE * P + slot] = bounce;
    iv[I_PASSES * P + slot] = passes;
    w.mode[slot] = (uint8_t)mode;
  }
  return nshadow;
}

// ---- shade classes --------------------------------------------------------------------------------
// The branch of shade_slot a path takes at its vertex: 0 a miss (background or environment,
// tracer.ts:119-124), 1 basic, 2 lambert or toon, 3 phong, 4 physical, 5 physical with maps,
// 6 physical with transmission (tracer.ts:254-389), 7 a material the path may pass straight
// through (alphaTest, transparent: tracer.ts:170-176).
__device__ inline uint32_t shade_class(const DevScene &sc, int k) {
  if (k < 0) return 0;
  const DevMaterial &m = sc.materials[__float_as_uint(sc.tri[(size_t)k * 3 + 2].z)];
  if (m.alpha_test > 0 || m.transparent) return 7;
  if (m.model == MCR_MODEL_BASIC) return 1;
  if (m.model == MCR_MODEL_LAMBERT || m.model == MCR_MODEL_TOON) return 2;
  if (m.model == MCR_MODEL_PHONG) return 3;
  return m.transmission > 0 ? 6 : m.textured ? 5 : 4;
}

// Bins the live slots of queue q by shade class: class c's slots go to region c of w.bins, in
// queue order within each block, and word C_BIN + c counts them. Shade then walks the regions one
// after another, so all but at most MCR_SHADE_CLASSES - 1 of its waves hold one class. Every
// sample keeps its own random stream, so the order shade visits slots in changes no result.
template <bool MIX>
// Queue entries per thread of k_shade_bin: a wave takes SB_RUNS runs of 64 consecutive entries, so
// a block reserves its bins' space for 1,024 entries with one returning atomic per class (the
// blocks' same-address atomics on the class counters bound the kernel at 256 entries per block).
#define SB_RUNS 4
__global__ __launch_bounds__(256) void k_shade_bin(const ViewDev *__restrict__ views, Wavefront w, int q,
                                                   DevCounters *cnt) {
  uint32_t n = w.counts[cw(q)];
  if (blockIdx.x * blockDim.x * SB_RUNS >= n) return;  // every thread of a block that goes on reaches the append
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t first = (blockIdx.x * 4 + wv) * 64 * SB_RUNS + lane;  // run r: first + 64 r
  size_t P = w.capacity;
  int slot[SB_RUNS];
  uint32_t cls[SB_RUNS];
#pragma unroll
  for (int r = 0; r < SB_RUNS; r++) slot[r] = first + 64 * r < n ? w.queue[q][first + 64 * r] : -1;
#pragma unroll
  for (int r = 0; r < SB_RUNS; r++) {
    cls[r] = MCR_SHADE_CLASSES;  // none
    if (slot[r] >= 0) {
      int k = w.i[I_HTRI * P + slot[r]];
      uint32_t view = MIX ? (uint32_t)w.i[I_VIEW * P + slot[r]] : 0;
      with_view<MIX>(views, view, [&](const ViewDev &V) { cls[r] = shade_class(V.sc, k); });
    }
  }
  // Positions in queue order within each class: a class's entries of run r follow those of runs
  // before it in this wave, lanes in order within a run.
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t *ctr[MCR_SHADE_CLASSES];
  uint32_t total[MCR_SHADE_CLASSES], pos[MCR_SHADE_CLASSES], at = 0;
  uint32_t kinds_mixed = 0;  // runs (64 consecutive entries, shade's waves) holding more than one class
#pragma unroll
  for (int c = 0; c < MCR_SHADE_CLASSES; c++) {
    ctr[c] = &w.counts[cw(C_BIN + c)];
    total[c] = 0;
  }
  uint32_t off[SB_RUNS];
#pragma unroll
  for (int r = 0; r < SB_RUNS; r++) {
    uint32_t kinds = 0;
    off[r] = 0;
#pragma unroll
    for (int c = 0; c < MCR_SHADE_CLASSES; c++) {
      uint64_t m = __ballot(cls[r] == (uint32_t)c);
      kinds += m != 0;
      if (cls[r] == (uint32_t)c) off[r] = total[c] + (uint32_t)__popcll(m & below);
      total[c] += (uint32_t)__popcll(m);
    }
    kinds_mixed += kinds > 1;
  }
  __shared__ uint32_t lds[4 * MCR_SHADE_CLASSES];
  block_reserve<MCR_SHADE_CLASSES>(ctr, total, lds, pos);
  // Statistics: lanes per class and mixed runs, one atomic per class per block.
  {
    __shared__ uint32_t mixed_runs, lanes[MCR_SHADE_CLASSES];
    if (threadIdx.x < MCR_SHADE_CLASSES) lanes[threadIdx.x] = 0;
    if (threadIdx.x == 0) mixed_runs = 0;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < MCR_SHADE_CLASSES; c++)
        if (total[c]) atomicAdd(&lanes[c], total[c]);
      if (kinds_mixed) atomicAdd(&mixed_runs, kinds_mixed);
    }
    __syncthreads();
    DevCounters *cc = counter_copy(cnt);
    if (threadIdx.x < MCR_SHADE_CLASSES && lanes[threadIdx.x])
      atomicAdd(&cc->shade_class_lanes[threadIdx.x], (unsigned long long)lanes[threadIdx.x]);
    if (threadIdx.x == 0 && mixed_runs) atomicAdd(&cc->class_mixed_waves, (unsigned long long)mixed_runs);
  }
#pragma unroll
  for (int r = 0; r < SB_RUNS; r++) {
#pragma unroll
    for (int c = 0; c < MCR_SHADE_CLASSES; c++)
      if (cls[r] == (uint32_t)c) at = pos[c];
    if (slot[r] >= 0) w.bins[(size_t)cls[r] * P + at + off[r]] = slot[r];
  }
}

// Entry i of the binned queue: the regions of w.bins one after another.
__device__ inline int binned_slot(const Wavefront &w, uint32_t i) {
  uint32_t base = 0, c = 0;
  for (; c + 1 < MCR_SHADE_CLASSES; c++) {
    uint32_t k = w.counts[cw(C_BIN + c)];
    if (i < base + k) break;
    base += k;
  }
  return w.bins[(size_t)c * w.capacity + (i - base)];
}

// The misses of a binned queue (class 0, region 0 of w.bins): the miss branch of shade_slot on its
// own, in a kernel light enough to run many waves per SIMD. The light-slot statuses of these slots
// are left as they are: resolve reads them only for vertices that add light, and the shadow kernels
// only for rays shade queued in this iteration.
template <bool MIX>
__global__ __launch_bounds__(256) void k_shade_miss(const ViewDev *__restrict__ views, Wavefront w,
                                                    DevCounters *cnt) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n = w.counts[cw(C_BIN)];
  if (gid >= n) return;
  size_t P = w.capacity;
  int slot = w.bins[gid];
  uint32_t view = MIX ? (uint32_t)w.i[I_VIEW * P + slot] : 0;
  if (MIX) {  // shade waves, as k_shade counts them
    uint64_t lanes = __ballot(true);
    int leader = __ffsll((unsigned long long)lanes) - 1;
    uint32_t v0 = __shfl(view, leader, 64);
    bool mixed = __ballot(view != v0) != 0;
    if ((int)(threadIdx.x & 63) == leader) {
      DevCounters *cc = counter_copy(cnt);
      atomicAdd(&cc->shade_waves, 1ull);
      if (mixed) atomicAdd(&cc->mixed_view_waves, 1ull);
    }
  }
  int flags = w.i[I_FLAGS * P + slot];
  double dx = w.f[F_DX * P + slot], dy = w.f[F_DY * P + slot], dz = w.f[F_DZ * P + slot];
  double tr = w.f[F_TR * P + slot], tg = w.f[F_TG * P + slot], tb = w.f[F_TB * P + slot];
  const Control &ctl = *w.ctl;
  int mode = MODE_DONE;
  with_view<MIX>(views, view, [&](const ViewDev &V) { mode = shade_miss(V, w, ctl, slot, flags, dx, dy, dz, tr, tg, tb); });
  w.mode[slot] = (uint8_t)mode;
}

// BINNED: the slots come from k_shade_bin's regions instead of queue q, past the misses of region
// 0 (k_shade_miss shades those).
template <bool MIX, bool BINNED>
__global__ __launch_bounds__(256) void k_shade(const ViewDev *__restrict__ views, Wavefront w, int q,
                                               DevCounters *cnt) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t skip = BINNED ? w.counts[cw(C_BIN)] : 0;
  uint32_t n = w.counts[cw(q)] - skip;
  if (gid == 0) {
    w.counts[cw(q ^ 1)] = 0;  // resolve of this iteration queues into it
    for (int k = 0; k < MCR_FETCH_COUNTERS; k++) {
      w.counts[cw(C_EXTEND_FETCH + k)] = 0;
      w.counts[cw(C_SHADOW_FETCH + k)] = 0;
    }
  }
  if (blockIdx.x * blockDim.x >= n) return;
  // Every thread of the block reaches the queue append below, so the bound check guards the work only.
  bool live = gid < n;
  size_t P = w.capacity;
  int slot = !live ? 0 : BINNED ? binned_slot(w, gid + skip) : w.queue[q][gid];
  uint32_t nshadow = 0, nl = 0;
  bool first_hit = live && w.i[I_BOUNCE * P + slot] == 0;  // the vertex a camera ray found
  uint32_t view = live ? (uint32_t)w.i[I_VIEW * P + slot] : 0;
  if (MIX) {  // how often a wave holds samples of several views
    uint64_t lanes = __ballot(live);
    if (lanes) {
      int leader = __ffsll((unsigned long long)lanes) - 1;
      uint32_t v0 = __shfl(view, leader, 64);
      bool mixed = __ballot(live && view != v0) != 0;
      if ((int)(threadIdx.x & 63) == leader) {
        DevCounters *cc = counter_copy(cnt);
        atomicAdd(&cc->shade_waves, 1ull);
        if (mixed) atomicAdd(&cc->mixed_view_waves, 1ull);
      }
    }
  }
  if (live) {
    const Control &ctl = *w.ctl;
    with_view<MIX>(views, view, [&](const ViewDev &V) {
      nshadow = shade_slot(V, w, ctl, slot);
      nl = V.shadow_slots;
    });
  }
  // Enqueue the wave's shadow rays (one atomic per counter per block). The queue has a region per
  // light (light li in region li % 16 when there are more than 16), `region_cap` entries each: rays
  // of first-hit vertices (neighbouring pixels) fill it from the front and rays of later vertices
  // from the back, so a traversal chunk mostly holds rays towards one light from neighbouring points
  // rather than a mix diluted by scattered later bounces. The order of the queue does not change
  // any result: each ray writes its own (slot, light) entry.
  uint64_t mask = 0;
  if (live && nshadow)
    for (uint32_t li = 0; li < nl; li++)
      if (w.status[(size_t)li * P + slot] == LS_RAY) mask |= 1ull << li;
  constexpr int NSEG = 2 * MCR_SHADOW_REGIONS;
  uint32_t tot[NSEG], got[NSEG];
  uint32_t *ctr[NSEG];
  const uint32_t L = w.lights;  // uniform bound of every lane's light loop
  // Regions in an unrolled loop (register arrays stay indexed by constants), the lights of a
  // region (li = r, r + 16, ...) in a loop of their own.
#pragma unroll
  for (int r = 0; r < MCR_SHADOW_REGIONS; r++) {
    uint32_t nf = 0, nb = 0;
    for (uint32_t li = r; li < L; li += MCR_SHADOW_REGIONS) {
      bool want = (mask >> li) & 1ull;
      nf += (uint32_t)__popcll(__ballot(want && first_hit));
      nb += (uint32_t)__popcll(__ballot(want && !first_hit));
    }
    tot[2 * r] = nf;
    tot[2 * r + 1] = nb;
    ctr[2 * r] = &w.counts[cw(C_SHADOW_SEG + 2 * r)];
    ctr[2 * r + 1] = &w.counts[cw(C_SHADOW_SEG + 2 * r + 1)];
  }
  __shared__ uint32_t lds[4 * NSEG];
  block_reserve<NSEG>(ctr, tot, lds, got);
  uint64_t below = (1ull << (threadIdx.x & 63)) - 1ull;
  const size_t RC = w.region_cap;
#pragma unroll
  for (int r = 0; r < MCR_SHADOW_REGIONS; r++) {
    uint32_t front = got[2 * r], back = got[2 * r + 1];
    for (uint32_t li = r; li < L; li += MCR_SHADOW_REGIONS) {
      bool want = (mask >> li) & 1ull;
      uint64_t bf = __ballot(want && first_hit), bl = __ballot(want && !first_hit);
      uint32_t entry = sref_entry((uint32_t)slot, li);
      if (want && first_hit) w.sref[(size_t)r * RC + front + (uint32_t)__popcll(bf & below)] = entry;
      if (want && !first_hit) w.sref[(size_t)(r + 1) * RC - 1 - (back + (uint32_t)__popcll(bl & below))] = entry;
      front += (uint32_t)__popcll(bf);
      back += (uint32_t)__popcll(bl);
    }
  }
}

// ---- shadow: transmittance + attenuateShadow for every queued shadow ray (tracer.ts:529-594) ----
// attenuateShadow's texture lookups (tracer.ts:555-567): alpha and colour at the hit's uv.
MCR_INLINE void shadow_texel(const DevScene &sc, const DevMaterial &m, int k, double u, double v, double &alpha,
                             double &cr, double &cg, double &cb) {
  const float *uvs = sc.uvs + (size_t)k * 6;
  double w0 = 1 - u - v;
  double tu = uvs[0] * w0 + uvs[2] * u + uvs[4] * v;
  double tv = uvs[1] * w0 + uvs[3] * u + uvs[5] * v;
  double tx[4];
  cr = m.color[0];
  cg = m.color[1];
  cb = m.color[2];
  if (m.map >= 0) {
    tex_sample(sc.textures[m.map], tu, tv, tx);
    alpha *= tx[3];
    cr *= tx[0];
    cg *= tx[1];
    cb *= tx[2];
  }
  if (m.alpha_map >= 0) {
    tex_sample(sc.textures[m.alpha_map], tu, tv, tx);
    alpha *= tx[1];
  }
}

// Collects the see-through hits of a shadow ray during a near-first traversal that stops at the
// first opaque, shadow-casting hit (see occluded() in device.h).
// TEX: materials may carry map or alphaMap; without them the hit's u, v are never kept.
template <bool TEX>
struct ShadowCollector {
  const DevScene *sc;
  bool castOnly;
  double *t;        // records of every traversal thread, this thread's at lane + i * stride
  int2 *meta;       // (leaf first triangle, triangle index | front << 31)
  double2 *uv;      // the hit's barycentrics, kept for textured materials (null without textures)
  uint32_t lane, stride;
  int n;
  bool overflow;
  __device__ bool operator()(int k, int leaf_first, double tk, double u, double v, bool front, const float4 &c) {
    int flags = (int)((__float_as_uint(c.y) >> 8) & 0xffu);
    if (castOnly && (flags & MCR_TRI_CAST_SHADOW) == 0) return false;  // not in the way
    if (flags & MCR_TRI_OPAQUE) return true;
    const DevMaterial &m = sc->materials[__float_as_uint(c.z)];
    bool textured = TEX && (m.map >= 0 || m.alpha_map >= 0);
    // Cut out (tracer.ts:567). With a texture the coverage is known only after the lookup, which
    // the replay does: the hit is kept here and dropped there, outside the traversal loop.
    if (!textured && m.alpha_test > 0 && m.opacity < m.alpha_test) return false;
    if (n < MCR_SHADOW_HIT_CAP) {
      size_t at = lane + (size_t)n * stride;
      t[at] = tk;
      meta[at] = make_int2(leaf_first, (int)((uint32_t)k | (front ? 0x80000000u : 0u)));
      if (textured) uv[at] = make_double2(u, v);
      n++;
    } else {
      overflow = true;
    }
    return false;
  }
};

// Reference visit order of two hits: leaves by descending first triangle (right subtrees first),
// then triangles by ascending index within a leaf.
__device__ inline bool visits_before(int2 a, int2 b) {
  if (a.x != b.x) return a.x > b.x;
  return (a.y & 0x7fffffff) < (b.y & 0x7fffffff);
}

// Entry e of the shadow queue, whose segments end at seg_end (two per region: first-hit rays
// from the front of the region, later ones from its back).
MCR_INLINE uint32_t shadow_queue_entry(const Wavefront &w, const uint32_t *seg_end, uint32_t e) {
  uint32_t seg = 0, start = 0;
  for (int k = 0; k < 2 * MCR_SHADOW_REGIONS; k++)
    if (e >= seg_end[k]) {
      seg = (uint32_t)k + 1;
      start = seg_end[k];
    }
  size_t region = (size_t)(seg >> 1) * w.region_cap;
  return (seg & 1) ? w.sref[region + w.region_cap - 1 - (e - start)] : w.sref[region + (e - start)];
}

// attenuateShadow (tracer.ts:544-594) over the see-through hits a shadow traversal kept (n of
// them, this thread's records at T/M/UV[i * S]): the hits are put in the reference's visit order
// and applied one by one; c is the light's colour factor (1, 1, 1 on entry).
template <bool TEX>
MCR_INLINE void shadow_replay(const DevScene &sc, double *T, int2 *M, double2 *UV, size_t S, int n, double c[3]) {
  for (int i = 1; i < n; i++) {  // insertion sort, few hits
    double ti = T[i * S];
    int2 mi = M[i * S];
    double2 ui = TEX ? UV[i * S] : make_double2(0, 0);
    int j = i - 1;
    while (j >= 0 && visits_before(mi, M[j * S])) {
      T[(j + 1) * S] = T[j * S];
      M[(j + 1) * S] = M[j * S];
      if (TEX) UV[(j + 1) * S] = UV[j * S];
      j--;
    }
    T[(j + 1) * S] = ti;
    M[(j + 1) * S] = mi;
    if (TEX) UV[(j + 1) * S] = ui;
  }
  double depth[3] = {0, 0, 0};
  double eps = sc.epsilon;
  for (int i = 0; i < n; i++) {
    double t = T[i * S];
    int2 mi = M[i * S];
    int k = mi.y & 0x7fffffff;
    bool front = (mi.y & 0x80000000) != 0;
    bool dup = false;
    for (int j = 0; j < i; j++)  // tracer.ts:572: hits already counted (marked by leaf -1)
      if (M[j * S].x >= 0 && jm_fabs(T[j * S] - t) <= eps) dup = true;
    if (dup) {
      M[i * S] = make_int2(-1, mi.y);
      continue;
    }
    const DevMaterial &m = sc.materials[__float_as_uint(sc.tri[(size_t)k * 3 + 2].z)];
    double alpha = m.opacity, mcr = m.color[0], mcg = m.color[1], mcb = m.color[2];
    if (TEX && (m.map >= 0 || m.alpha_map >= 0)) {
      double2 hu = UV[i * S];
      shadow_texel(sc, m, k, hu.x, hu.y, alpha, mcr, mcg, mcb);
      if (m.alpha_test > 0 && alpha < m.alpha_test) {  // cut out: neither counted nor seen
        M[i * S] = make_int2(-1, mi.y);
        continue;
      }
    }
    if (!m.transparent) alpha = 1;
    if (m.thickness > 0) {
      double sgn = front ? -t : t;
      depth[0] = depth[0] + m.attenuation[0] * sgn;
      depth[1] = depth[1] + m.attenuation[1] * sgn;
      depth[2] = depth[2] + m.attenuation[2] * sgn;
    }
    double open = 1 - alpha, through = m.transmission;
    c[0] = c[0] * (open + alpha * through * mcr);
    c[1] = c[1] * (open + alpha * through * mcg);
    c[2] = c[2] * (open + alpha * through * mcb);
    if (c[0] + c[1] + c[2] < 1e-4) {
      c[0] = c[1] = c[2] = 0;
      break;
    }
  }
  for (int k = 0; k < 3; k++)
    if (depth[k] > 0) c[k] = c[k] * jm_exp(-depth[k]);
}

template <bool MIX, bool WIDE, bool TEX, class Stack>
MCR_INLINE void shadow_loop(const ViewDev *__restrict__ views, const Wavefront &w, Stack &st,
                            double *seen, int2 *seen_meta, double2 *seen_uv, DevCounters *cnt) {
  // The queue is the concatenation of its segments: per region, first-hit rays (front of the
  // region of sref), then later ones (back of the region, in reverse).
  uint32_t seg_end[2 * MCR_SHADOW_REGIONS];
  uint32_t n = 0;
  for (int k = 0; k < 2 * MCR_SHADOW_REGIONS; k++) {
    n += w.counts[cw(C_SHADOW_SEG + k)];
    seg_end[k] = n;
  }
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nthreads = gridDim.x * blockDim.x;
  TravCount tc = {0, 0, 0};
  unsigned long long queries = 0;
  size_t P = w.capacity;
  const size_t RS = (size_t)w.lights * P;
  bool bad = false;
  unsigned long long ndeep = 0;
  double *T = seen + gid;
  int2 *M = seen_meta + gid;
  double2 *UV = TEX ? seen_uv + gid : nullptr;  // the replay's view of this thread's records
  const size_t S = nthreads;
  WaveFetch fetch(&w.counts[cw(C_SHADOW_FETCH)], n);
  for (uint32_t b = fetch.next(); b < n; b = fetch.next()) {
    uint32_t e = b + (threadIdx.x & 63);
    if (e >= n) continue;
    uint32_t ref = shadow_queue_entry(w, seg_end, e);
    uint32_t slot = sref_slot(ref), li = sref_light(ref);
    size_t idx = (size_t)li * P + slot;
    double ox = nt_ld(&w.sray[0 * RS + idx]), oy = nt_ld(&w.sray[1 * RS + idx]), oz = nt_ld(&w.sray[2 * RS + idx]);
    double dx = nt_ld(&w.sray[3 * RS + idx]), dy = nt_ld(&w.sray[4 * RS + idx]), dz = nt_ld(&w.sray[5 * RS + idx]);
    double maxT = nt_ld(&w.sray[6 * RS + idx]);
    double c[3] = {1, 1, 1};
    bool deep = false;
    with_view<MIX>(views, (uint32_t)w.i[I_VIEW * P + slot], [&](const ViewDev &V) {
    const DevScene &sc = V.sc;
    // The environment-diffuse ray (slot light_slots) is transmittance(..., false) (tracer.ts:346).
    ShadowCollector<TEX> col{&sc, V.st.shadows == MCR_SHADOWS_CAST && (uint32_t)li != V.light_slots, seen, seen_meta,
                             TEX ? seen_uv : nullptr, gid, (uint32_t)S, 0, false};
    bool blocked;
    if constexpr (WIDE) {
      WideRay wr;
      // Flags and material come from the triangle's FP64-path record.
      auto on_hit = [&](int k, int leaf, float t, float u, float v, bool front) {
        return col(k, leaf, (double)t, (double)u, (double)v, front, sc.tri[(size_t)k * 3 + 2]);
      };
      blocked = wide_ray((float)ox, (float)oy, (float)oz, (float)dx, (float)dy, (float)dz, wr) &&
                wide_occluded(V.wide, wr, 0.0f, (float)maxT, st, tc, bad, on_hit);
    } else {
      blocked = occluded(sc, ox, oy, oz, dx, dy, dz, 0, maxT, st, tc, col);
    }
    // More see-through hits than the list keeps: k_shadow_deep traces the ray again in the
    // reference's order. An opaque hit settles the ray whatever else it crosses.
    deep = !blocked && col.overflow;
    if (blocked) {
      c[0] = c[1] = c[2] = 0;
    } else if (col.n > 0 && !deep) {
      shadow_replay<TEX>(sc, T, M, UV, S, col.n, c);
    }
    });
    queries++;
    if (deep) {
      w.status[idx] = LS_DEEP;
      ndeep++;
      continue;
    }
    bool visible = c[0] + c[1] + c[2] > 0;
    nt_st(&w.shadow[0 * RS + idx], c[0]);
    nt_st(&w.shadow[1 * RS + idx], c[1]);
    nt_st(&w.shadow[2 * RS + idx], c[2]);
    w.status[idx] = visible ? LS_VISIBLE : LS_BLOCKED;
  }
  uint32_t wave_deep = (uint32_t)wave_sum(ndeep);
  if ((threadIdx.x & 63) == 0 && wave_deep) atomicAdd(&w.counts[cw(C_DEEP)], wave_deep);
  if (bad) atomicOr(&w.counts[cw(C_OVERFLOW)], OVERFLOW_WIDE);
  flush_counts(cnt, tc, 0, queries);
}

// attenuateShadow (tracer.ts:544-594) applied to each hit as intersectAny visits it, for shadow
// rays retraced in the reference's order: the distances of the surfaces already counted live in
// this thread's MCR_DEEP_HITS words of deep_seen.
template <bool TEX>
struct DeepShadow {
  const DevScene *sc;
  bool castOnly;
  double *seen;     // this thread's counted distances at lane + j * MCR_DEEP_THREADS
  uint32_t lane;
  int n;
  bool overflow;
  double c[3];
  double depth[3];
  __device__ bool operator()(int k, double t, double u, double v, bool front) {
    float4 rec = sc->tri[(size_t)k * 3 + 2];
    int flags = (int)((__float_as_uint(rec.y) >> 8) & 0xffu);
    if (castOnly && (flags & MCR_TRI_CAST_SHADOW) == 0) return true;
    if (flags & MCR_TRI_OPAQUE) {
      c[0] = c[1] = c[2] = 0;
      return false;
    }
    const DevMaterial &m = sc->materials[__float_as_uint(rec.z)];
    double alpha = m.opacity, cr = m.color[0], cg = m.color[1], cb = m.color[2];
    if (TEX && (m.map >= 0 || m.alpha_map >= 0)) shadow_texel(*sc, m, k, u, v, alpha, cr, cg, cb);
    if (m.alpha_test > 0 && alpha < m.alpha_test) return true;
    if (!m.transparent) alpha = 1;
    double eps = sc->epsilon;
    for (int j = 0; j < n; j++)
      if (jm_fabs(seen[lane + (size_t)j * MCR_DEEP_THREADS] - t) <= eps) return true;
    if (n == MCR_DEEP_HITS) {  // the one capacity left: reported, never approximated
      overflow = true;
      return false;
    }
    seen[lane + (size_t)n * MCR_DEEP_THREADS] = t;
    n++;
    if (m.thickness > 0) {
      double sgn = front ? -t : t;
      depth[0] = depth[0] + m.attenuation[0] * sgn;
      depth[1] = depth[1] + m.attenuation[1] * sgn;
      depth[2] = depth[2] + m.attenuation[2] * sgn;
    }
    double open = 1 - alpha, through = m.transmission;
    c[0] = c[0] * (open + alpha * through * cr);
    c[1] = c[1] * (open + alpha * through * cg);
    c[2] = c[2] * (open + alpha * through * cb);
    if (c[0] + c[1] + c[2] < 1e-4) {
      c[0] = c[1] = c[2] = 0;
      return false;
    }
    return true;
  }
};

// The reference-order pass over the shadow rays k_shadow marked LS_DEEP (none in most iterations:
// the kernel then returns at once). MCR_DEEP_THREADS threads stride over the queue; each marked
// ray is traced with intersectAny's depth-first order on the FP64 binary BVH and attenuateShadow
// per hit, exactly as transmittance does (tracer.ts:529-542).
template <bool MIX, bool TEX>
__global__ __launch_bounds__(256) void k_shadow_deep(const ViewDev *__restrict__ views, Wavefront w,
                                                     int32_t *spill, uint32_t spill_stride, DevCounters *cnt) {
  if (w.counts[cw(C_DEEP)] == 0) return;
  uint32_t seg_end[2 * MCR_SHADOW_REGIONS];
  uint32_t n = 0;
  for (int k = 0; k < 2 * MCR_SHADOW_REGIONS; k++) {
    n += w.counts[cw(C_SHADOW_SEG + k)];
    seg_end[k] = n;
  }
  MCR_LANE_STACK(st, spill, spill_stride);
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  size_t P = w.capacity;
  const size_t RS = (size_t)w.lights * P;
  TravCount tc = {0, 0, 0};
  bool overflow = false;
  for (uint32_t e = gid; e < n; e += MCR_DEEP_THREADS) {
    uint32_t ref = shadow_queue_entry(w, seg_end, e);
    uint32_t slot = sref_slot(ref), li = sref_light(ref);
    size_t idx = (size_t)li * P + slot;
    if (w.status[idx] != LS_DEEP) continue;
    double ox = w.sray[0 * RS + idx], oy = w.sray[1 * RS + idx], oz = w.sray[2 * RS + idx];
    double dx = w.sray[3 * RS + idx], dy = w.sray[4 * RS + idx], dz = w.sray[5 * RS + idx];
    double maxT = w.sray[6 * RS + idx];
    double c[3];
    uint32_t view = MIX ? (uint32_t)w.i[I_VIEW * P + slot] : 0;
    bool over = false;
    with_view<MIX>(views, view, [&](const ViewDev &V) {
      const DevScene &sc = V.sc;
      DeepShadow<TEX> vis{&sc, V.st.shadows == MCR_SHADOWS_CAST && li != V.light_slots, w.deep_seen, gid, 0, false,
                          {1, 1, 1}, {0, 0, 0}};
      intersect_any(sc, ox, oy, oz, dx, dy, dz, 0, maxT, st, tc, vis);
      over = vis.overflow;
      for (int k = 0; k < 3; k++) c[k] = vis.depth[k] > 0 ? vis.c[k] * jm_exp(-vis.depth[k]) : vis.c[k];
    });
    if (over) {
      overflow = true;
      if (view < MCR_FAIL_VIEWS) atomicOr(&w.counts[C_FAIL_BITS + view / 32], 1u << (view % 32));
    }
    bool visible = c[0] + c[1] + c[2] > 0;
    w.shadow[0 * RS + idx] = c[0];
    w.shadow[1 * RS + idx] = c[1];
    w.shadow[2 * RS + idx] = c[2];
    w.status[idx] = visible ? LS_VISIBLE : LS_BLOCKED;
  }
  if (overflow) atomicOr(&w.counts[cw(C_OVERFLOW)], OVERFLOW_SEEN);
  flush_counts(cnt, tc, 0, 0);
}

template <bool MIX, bool TEX>
__global__ __launch_bounds__(256, 5) void k_shadow(const ViewDev *__restrict__ views, Wavefront w,
                                                int32_t *spill, uint32_t spill_stride,
                                                double *seen, int2 *seen_meta, double2 *seen_uv,
                                                DevCounters *cnt) {
  MCR_LANE_STACK(st, spill, spill_stride);
  shadow_loop<MIX, false, TEX>(views, w, st, seen, seen_meta, seen_uv, cnt);
}

// Five waves per SIMD: the any-hit loop is latency bound, and at the register count the compiler
// picks unasked only four fit (4.74 -> 4.44 ms per pass of the 1.38M-triangle 1080p frame). The closest-hit variant would spill.
template <bool MIX, bool TEX>
__global__ __launch_bounds__(256, 5) void k_shadow_wide(const ViewDev *__restrict__ views, Wavefront w,
                                                        int32_t *spill, uint32_t spill_stride,
                                                        double *seen, int2 *seen_meta, double2 *seen_uv,
                                                        DevCounters *cnt) {
  __shared__ uint32_t lds_base[MCR_WIDE_STACK_LDS * 256], lds_bits[MCR_WIDE_STACK_LDS * 256];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  WideStack st{lds_base + threadIdx.x, lds_bits + threadIdx.x, spill + gid, (int)spill_stride, 256, 0};
  shadow_loop<MIX, true, TEX>(views, w, st, seen, seen_meta, seen_uv, cnt);
}

// ---- resolve: finish the vertex's radiance in the reference's summation order ---------------------
MCR_INLINE void resolve_slot(const ViewDev &V, const Wavefront &w, const Control &ctl, int slot,
                                    int mode) {
  const DevScene &sc = V.sc;
  const mcr_settings &st = V.st;
  size_t P = w.capacity;
  const size_t RS = (size_t)w.lights * P;
  double *f = w.f;
  if (mode == MODE_ADD_EXTEND || mode == MODE_ADD_FINISH) {
    double oR = f[F_AR * P + slot], oG = f[F_AG * P + slot], oB = f[F_AB * P + slot];
    uint32_t nl = V.light_slots;
    for (uint32_t li = 0; li < nl; li++) {
      size_t idx = (size_t)li * P + slot;
      if (w.status[idx] != LS_VISIBLE) continue;
      oR += nt_ld(&w.contrib[0 * RS + idx]) * nt_ld(&w.shadow[0 * RS + idx]);
      oG += nt_ld(&w.contrib[1 * RS + idx]) * nt_ld(&w.shadow[1 * RS + idx]);
      oB += nt_ld(&w.contrib[2 * RS + idx]) * nt_ld(&w.shadow[2 * RS + idx]);
    }
    int flags = w.i[I_FLAGS * P + slot];
    if (flags & PF_INDIRECT) {
      oR += f[F_IR * P + slot];
      oG += f[F_IG * P + slot];
      oB += f[F_IB * P + slot];
    }
    if (V.shadow_slots > nl) {  // environment light on the diffuse layer (tracer.ts:341-353)
      size_t idx = (size_t)nl * P + slot;
      if (w.status[idx] == LS_VISIBLE) {
        oR += nt_ld(&w.contrib[0 * RS + idx]) * nt_ld(&w.shadow[0 * RS + idx]);
        oG += nt_ld(&w.contrib[1 * RS + idx]) * nt_ld(&w.shadow[1 * RS + idx]);
        oB += nt_ld(&w.contrib[2 * RS + idx]) * nt_ld(&w.shadow[2 * RS + idx]);
      }
    }
    int combine = (flags & PF_COMBINE_MASK) >> PF_COMBINE_SHIFT;
    if (combine) {  // tracer.ts:392-406
      double eR = f[F_XR * P + slot], eG = f[F_XG * P + slot], eB = f[F_XB * P + slot], s = f[F_XS * P + slot];
      if (combine == 1) {
        oR *= 1 - s + s * eR;
        oG *= 1 - s + s * eG;
        oB *= 1 - s + s * eB;
      } else if (combine == 2) {
        oR += (eR - oR) * s;
        oG += (eG - oG) * s;
        oB += (eB - oB) * s;
      } else {
        oR += eR * s;
        oG += eG * s;
        oB += eB * s;
      }
    }
    f[F_LR * P + slot] += f[F_QR * P + slot] * (oR + f[F_ER * P + slot]);
    f[F_LG * P + slot] += f[F_QG * P + slot] * (oG + f[F_EG * P + slot]);
    f[F_LB * P + slot] += f[F_QB * P + slot] * (oB + f[F_EB * P + slot]);
  }
  if (mode == MODE_ADD_FINISH || mode == MODE_FINISH) {
    // tracer.ts:472-477: tone map, then fog, alpha 1.
    double c[4] = {f[F_LR * P + slot], f[F_LG * P + slot], f[F_LB * P + slot], 1};
    int flags = w.i[I_FLAGS * P + slot];
    if ((flags & PF_TONEMAP) && st.tone_mapping != MCR_NO_TONE_MAPPING)
      tone_map(c, st.tone_mapping, st.tone_mapping_exposure);
    double fogf = f[F_FOG * P + slot];
    if (fogf > 0 && sc.fog_kind != MCR_FOG_NONE)
      for (int k = 0; k < 3; k++) c[k] = c[k] + (sc.fog_color[k] - c[k]) * fogf;
    uint32_t smp = (uint32_t)w.i[I_SAMPLE * P + slot];
    write_sample(ctl.jobs[job_of(ctl.jobs, ctl.njobs, smp, w.i[I_JOB * P + slot])], ctl.out, smp, c);
  }
}

// 1,024 threads per block: three same-address atomics per block (the sample counter, the free
// list and the next queue) for four times the slots of a 256-thread block.
#define RESOLVE_THREADS 1024
template <bool MIX>
__global__ __launch_bounds__(RESOLVE_THREADS) void k_resolve(const ViewDev *__restrict__ views, Wavefront w, int q) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n = w.counts[cw(q)];
  if (blockIdx.x * blockDim.x >= n) return;
  bool live = gid < n;
  size_t P = w.capacity;
  const Control &ctl = *w.ctl;
  int slot = live ? w.queue[q][gid] : 0;
  int mode = live ? w.mode[slot] : MODE_DONE;
  if (mode != MODE_DONE && mode != MODE_EXTEND)
    with_view<MIX>(views, (uint32_t)w.i[I_VIEW * P + slot],
              [&](const ViewDev &V) { resolve_slot(V, w, ctl, slot, mode); });
  bool requeue = mode == MODE_EXTEND || mode == MODE_ADD_EXTEND;
  bool finished = live && !requeue;
  if (ctl.job_done) {
    int job = finished ? job_of(ctl.jobs, ctl.njobs, (uint32_t)w.i[I_SAMPLE * P + slot], w.i[I_JOB * P + slot]) : 0;
    count_done(ctl.job_done, job, finished);
  }
  // A finished slot starts the call's next sample at once, so the slot keeps its place in the
  // queue and consecutive lanes keep reading consecutive slots. With no sample left it goes to
  // the free list, from which a refill restarts it when the job list grows.
  constexpr int WAVES = RESOLVE_THREADS / 64;
  __shared__ uint32_t lds[2 * WAVES];
  uint32_t next;
  {
    uint32_t *c[1] = {&w.counts[cw(C_NEXT)]};
    bool want[1] = {finished};
    uint32_t pos[1];
    block_append<1, WAVES>(c, want, lds, pos);
    next = pos[0];
  }
  bool idle = false;
  if (finished) {
    if (next < ctl.total) {
      start_sample_any<MIX>(views, w, ctl, (uint32_t)slot, next, wave_job_hint(ctl.jobs, ctl.njobs, next));
      requeue = true;
    } else {
      idle = true;
    }
  }
  uint32_t *c[2] = {&w.counts[cw(C_FREE)], &w.counts[cw(q ^ 1)]};
  bool want[2] = {idle, requeue};
  uint32_t pos[2];
  block_append<2, WAVES>(c, want, lds, pos);
  if (idle) w.free_list[pos[0]] = slot;
  if (requeue) w.queue[q ^ 1][pos[1]] = slot;
}

// Restarts `count` free slots (the top of the free list) with samples next .. next+count-1 and
// queues them on queue q after its `queued` live entries; the host passes the counter values it
// read after the last iteration and this kernel writes the updated ones.
template <bool MIX>
__global__ __launch_bounds__(256) void k_refill(const ViewDev *__restrict__ views, Wavefront w, int q,
                                                uint32_t count, uint32_t free_top, uint32_t next,
                                                uint32_t queued) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    w.counts[cw(C_FREE)] = free_top - count;
    w.counts[cw(C_NEXT)] = next + count;
    w.counts[cw(q)] = queued + count;
  }
  if (i >= count) return;
  int slot = w.free_list[free_top - 1 - i];
  w.queue[q][queued + i] = slot;
  start_sample_any<MIX>(views, w, *w.ctl, (uint32_t)slot, next + i,
                        wave_job_hint(w.ctl->jobs, w.ctl->njobs, next + i));
}

// ---- queue re-sort: live slots of queue q in ascending slot order ---------------------------------
// Slots that went idle and were refilled come back in free-list order, scattered over the slot
// range, and in-place regeneration keeps them where they are; a wave over such a queue touches
// many more lines of the per-slot arrays than one over neighbouring slots. Between a resolve and
// the next shade the per-slot mode byte is dead, so it marks the live slots here.
#define MODE_SORT_MARK 0xffu
__global__ __launch_bounds__(256) void k_sort_mark(Wavefront w, int q) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) w.counts[cw(C_SORT)] = 0;
  if (i < w.counts[cw(q)]) w.mode[w.queue[q][i]] = (uint8_t)MODE_SORT_MARK;
}

__global__ __launch_bounds__(256) void k_sort_gather(Wavefront w, int q) {
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = slot < w.capacity && w.mode[slot] == MODE_SORT_MARK;
  __shared__ uint32_t lds[4];
  uint32_t *c[1] = {&w.counts[cw(C_SORT)]};
  bool want[1] = {live};
  uint32_t pos[1];
  block_append<1>(c, want, lds, pos);
  if (live) {
    w.queue[q][pos[0]] = (int32_t)slot;
    w.mode[slot] = (uint8_t)MODE_DONE;
  }
}

// Free list = every slot, in descending order, so a refill of k slots takes slots 0 .. k-1.
__global__ __launch_bounds__(256) void k_reset_slots(Wavefront w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0)
    for (int k = 0; k < C_WORDS; k++) w.counts[cw(k)] = k == C_FREE ? w.capacity : 0;
  if (i < MCR_FAIL_VIEWS / 32) w.counts[C_FAIL_BITS + i] = 0;
  if (i < w.capacity) w.free_list[i] = (int32_t)(w.capacity - 1 - i);
}

// ---- ray queries over caller rays (bvh.ts:202-332) ---------------------------------------------
__global__ __launch_bounds__(256) void k_probe_closest(DevScene sc, const double *rays, uint32_t n,
                                                       double tmin, double tmax, int cull,
                                                       int32_t *tri, double *t, double *u, double *v,
                                                       uint8_t *front, int32_t *spill,
                                                       uint32_t spill_stride, DevCounters *cnt,
                                                       uint32_t *fetch_counters) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  MCR_LANE_STACK(st, spill, spill_stride);
  TravCount tc = {0, 0, 0};
  unsigned long long queries = 0;
  WaveFetch fetch(fetch_counters, n);
  for (uint32_t b = fetch.next(); b < n; b = fetch.next()) {
    uint32_t i = b + (threadIdx.x & 63);
    if (i >= n) continue;
    const double *r = rays + (size_t)i * 6;
    TriHit h;
    intersect_closest(sc, r[0], r[1], r[2], r[3], r[4], r[5], tmin, tmax, cull != 0, h, st, tc);
    queries++;
    if (tri) tri[i] = h.tri;
    if (t) t[i] = h.t;
    if (u) u[i] = h.u;
    if (v) v[i] = h.v;
    if (front) front[i] = h.front ? 1 : 0;
  }
  flush_counts(cnt, tc, queries, 0);
}

struct CountVisitor {
  uint32_t n;
  __device__ bool operator()(int, double, double, double, bool) {
    n++;
    return true;
  }
};

__global__ __launch_bounds__(256) void k_probe_any_count(DevScene sc, const double *rays, uint32_t n,
                                                         double tmin, double tmax, uint32_t *count,
                                                         int32_t *spill, uint32_t spill_stride,
                                                         DevCounters *cnt, uint32_t *fetch_counters) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  MCR_LANE_STACK(st, spill, spill_stride);
  TravCount tc = {0, 0, 0};
  unsigned long long queries = 0;
  WaveFetch fetch(fetch_counters, n);
  for (uint32_t b = fetch.next(); b < n; b = fetch.next()) {
    uint32_t i = b + (threadIdx.x & 63);
    if (i >= n) continue;
    const double *r = rays + (size_t)i * 6;
    CountVisitor cv{0};
    intersect_any(sc, r[0], r[1], r[2], r[3], r[4], r[5], tmin, tmax, st, tc, cv);
    queries++;
    count[i] = cv.n;
  }
  flush_counts(cnt, tc, 0, queries);
}

// ---- writeTile (progressive.ts:288-313) --------------------------------------------------------
// ECMAScript ToUint8Clamp: round half to even, clamp to [0, 255], NaN -> 0.
__device__ inline uint8_t to_uint8_clamp(double v) {
  if (!(v > 0)) return 0;
  if (v >= 255) return 255;
  double f = jm_floor(v);
  if (f + 0.5 < v) return (uint8_t)(f + 1);
  if (v < f + 0.5) return (uint8_t)f;
  int fi = (int)f;
  return (uint8_t)((fi & 1) ? fi + 1 : fi);
}

// One pixel of writeTile: sum += sample (the array keeps the float, the frame the double), mean,
// un-premultiply, sRGB encode, ToUint8Clamp(v * 255 + 0.5).
// sum64, when given instead of sum, keeps the running sum in FP64 (renders beyond 1024 samples,
// where a float sum would round away the late samples' contribution).
MCR_INLINE void encode_pixel(uint8_t *image, uint32_t width, int32_t tx, int32_t ty, int32_t tw,
                             const float *data, float *sum, double *sum64, uint32_t count, int srgb,
                             uint32_t i) {
  uint32_t x = i % (uint32_t)tw, y = i / (uint32_t)tw;
  size_t s = (size_t)i * 4;
  size_t d = (((size_t)ty + y) * width + tx + x) * 4;
  double c[4] = {data[s], data[s + 1], data[s + 2], data[s + 3]};
  if (sum64) {
    double scale = 1.0 / count;
    for (int k = 0; k < 4; k++) {
      double v = sum64[d + k] + c[k];
      sum64[d + k] = v;
      c[k] = v * scale;
    }
  } else if (sum) {
    double scale = 1.0 / count;
    for (int k = 0; k < 4; k++) {
      double v = (double)sum[d + k] + c[k];
      sum[d + k] = (float)v;
      c[k] = v;
    }
    for (int k = 0; k < 4; k++) c[k] *= scale;
  }
  if (c[3] > 0) {
    c[0] /= c[3];
    c[1] /= c[3];
    c[2] /= c[3];
  }
  if (srgb)
    for (int k = 0; k < 3; k++) c[k] = linear_to_srgb(js_max(0, c[k]));
  for (int k = 0; k < 4; k++) image[d + k] = to_uint8_clamp(c[k] * 255 + 0.5);
}

__global__ __launch_bounds__(256) void k_encode(uint8_t *image, uint32_t width, mcr_tile tile,
                                                const float *data, float *sum, double *sum64,
                                                uint32_t count, int srgb) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)tile.width * (uint32_t)tile.height) return;
  encode_pixel(image, width, tile.x, tile.y, tile.width, data, sum, sum64, count, srgb, i);
}

// Many tiles at once: block b works on the item whose block range holds b. No two items of one
// launch touch the same pixel (the host puts at most one pass of a tile in a launch).
__global__ __launch_bounds__(256) void k_encode_batch(const EncodeItem *items, uint32_t n) {
  int lo = 0, hi = (int)n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (items[mid].block_start <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const EncodeItem &it = items[lo];
  uint32_t i = (blockIdx.x - it.block_start) * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)it.tw * (uint32_t)it.th) return;
  encode_pixel(it.image, it.width, it.tx, it.ty, it.tw, it.data, it.sum, it.sum64, it.count, it.srgb, i);
}

// ---- launch wrappers ----------------------------------------------------------------------------
static inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

hipError_t mcr_upload_srgb8(const float *table256) {
  return hipMemcpyToSymbol(HIP_SYMBOL(g_srgb8), table256, 256 * sizeof(float));
}
hipError_t mcr_upload_dfg(const float *lut512) {
  return hipMemcpyToSymbol(HIP_SYMBOL(g_dfg), lut512, 512 * sizeof(float));
}

void launch_reset_slots(hipStream_t s, const Wavefront &w) {
  hipLaunchKernelGGL(k_reset_slots, dim3(blocks_for(w.capacity)), dim3(256), 0, s, w);
}
#define MCR_LAUNCH_MIX(kernel, grid, stream, ...)                                           \
  do {                                                                                      \
    if (mix) hipLaunchKernelGGL(kernel<true>, dim3(grid), dim3(256), 0, stream, __VA_ARGS__); \
    else hipLaunchKernelGGL(kernel<false>, dim3(grid), dim3(256), 0, stream, __VA_ARGS__);     \
  } while (0)

void launch_refill(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                   uint32_t count, uint32_t free_top, uint32_t next, uint32_t queued) {
  MCR_LAUNCH_MIX(k_refill, blocks_for(count ? count : 1), s, views, w, q, count, free_top, next, queued);
}
void launch_sort_queue(hipStream_t s, const Wavefront &w, int q, uint32_t live) {
  hipLaunchKernelGGL(k_sort_mark, dim3(blocks_for(live ? live : 1)), dim3(256), 0, s, w, q);
  hipLaunchKernelGGL(k_sort_gather, dim3(blocks_for(w.capacity)), dim3(256), 0, s, w, q);
}
void launch_extend(hipStream_t s, bool mix, bool wide, const ViewDev *views, const Wavefront &w, int q,
                   uint32_t grid, int32_t *spill, uint32_t spill_stride, DevCounters *cnt) {
  if (wide) MCR_LAUNCH_MIX(k_extend_wide, grid, s, views, w, q, spill, spill_stride, cnt);
  else MCR_LAUNCH_MIX(k_extend, grid, s, views, w, q, spill, spill_stride, cnt);
}
void launch_shade_bin(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                      uint32_t max_live, DevCounters *cnt) {
  MCR_LAUNCH_MIX(k_shade_bin, blocks_for(max_live ? (max_live + SB_RUNS - 1) / SB_RUNS : 1), s, views, w, q, cnt);
}
void launch_shade(hipStream_t s, bool mix, bool binned, const ViewDev *views, const Wavefront &w, int q,
                  uint32_t max_live, DevCounters *cnt) {
  uint32_t grid = blocks_for(max_live ? max_live : 1);
  if (binned) {
    MCR_LAUNCH_MIX(k_shade_miss, grid, s, views, w, cnt);
    if (mix) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<true, true>), dim3(grid), dim3(256), 0, s, views, w, q, cnt);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<false, true>), dim3(grid), dim3(256), 0, s, views, w, q, cnt);
  } else {
    if (mix) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<true, false>), dim3(grid), dim3(256), 0, s, views, w, q, cnt);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_shade<false, false>), dim3(grid), dim3(256), 0, s, views, w, q, cnt);
  }
}
template <bool TEX>
static void launch_shadow_tex(hipStream_t s, bool mix, bool wide, const ViewDev *views, const Wavefront &w,
                              uint32_t grid, int32_t *spill, uint32_t spill_stride, double *seen, int2 *seen_meta,
                              double2 *seen_uv, DevCounters *cnt) {
  if (wide) {
    if (mix)
      hipLaunchKernelGGL((k_shadow_wide<true, TEX>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, seen,
                         seen_meta, seen_uv, cnt);
    else
      hipLaunchKernelGGL((k_shadow_wide<false, TEX>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride,
                         seen, seen_meta, seen_uv, cnt);
  } else {
    if (mix)
      hipLaunchKernelGGL((k_shadow<true, TEX>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, seen,
                         seen_meta, seen_uv, cnt);
    else
      hipLaunchKernelGGL((k_shadow<false, TEX>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, seen,
                         seen_meta, seen_uv, cnt);
  }
}
void launch_shadow(hipStream_t s, bool mix, bool wide, bool tex, const ViewDev *views, const Wavefront &w,
                   uint32_t grid, int32_t *spill, uint32_t spill_stride, double *seen, int2 *seen_meta,
                   double2 *seen_uv, DevCounters *cnt) {
  if (tex) launch_shadow_tex<true>(s, mix, wide, views, w, grid, spill, spill_stride, seen, seen_meta, seen_uv, cnt);
  else launch_shadow_tex<false>(s, mix, wide, views, w, grid, spill, spill_stride, seen, seen_meta, seen_uv, cnt);
}
void launch_shadow_deep(hipStream_t s, bool mix, bool tex, const ViewDev *views, const Wavefront &w, int32_t *spill,
                        uint32_t spill_stride, DevCounters *cnt) {
  const uint32_t grid = MCR_DEEP_THREADS / 256;
  if (tex) {
    if (mix) hipLaunchKernelGGL((k_shadow_deep<true, true>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, cnt);
    else hipLaunchKernelGGL((k_shadow_deep<false, true>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, cnt);
  } else {
    if (mix) hipLaunchKernelGGL((k_shadow_deep<true, false>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, cnt);
    else hipLaunchKernelGGL((k_shadow_deep<false, false>), dim3(grid), dim3(256), 0, s, views, w, spill, spill_stride, cnt);
  }
}
void launch_resolve(hipStream_t s, bool mix, const ViewDev *views, const Wavefront &w, int q,
                    uint32_t max_live) {
  const uint32_t grid = ((max_live ? max_live : 1) + RESOLVE_THREADS - 1) / RESOLVE_THREADS;
  if (mix) hipLaunchKernelGGL(k_resolve<true>, dim3(grid), dim3(RESOLVE_THREADS), 0, s, views, w, q);
  else hipLaunchKernelGGL(k_resolve<false>, dim3(grid), dim3(RESOLVE_THREADS), 0, s, views, w, q);
}
void launch_probe_closest(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                          double tmin, double tmax, int cull, int32_t *tri, double *t, double *u,
                          double *v, uint8_t *front, uint32_t grid, int32_t *spill,
                          uint32_t spill_stride, DevCounters *cnt, uint32_t *fetch) {
  hipLaunchKernelGGL(k_probe_closest, dim3(grid), dim3(256), 0, s, sc, rays, n, tmin, tmax, cull,
                     tri, t, u, v, front, spill, spill_stride, cnt, fetch);
}
void launch_probe_any_count(hipStream_t s, const DevScene &sc, const double *rays, uint32_t n,
                            double tmin, double tmax, uint32_t *count, uint32_t grid,
                            int32_t *spill, uint32_t spill_stride, DevCounters *cnt, uint32_t *fetch) {
  hipLaunchKernelGGL(k_probe_any_count, dim3(grid), dim3(256), 0, s, sc, rays, n, tmin, tmax,
                     count, spill, spill_stride, cnt, fetch);
}
void launch_encode_batch(hipStream_t s, const EncodeItem *items, uint32_t n, uint32_t blocks) {
  if (n && blocks) hipLaunchKernelGGL(k_encode_batch, dim3(blocks), dim3(256), 0, s, items, n);
}
void launch_encode(hipStream_t s, uint8_t *image, uint32_t width, const mcr_tile &tile,
                   const float *data, float *sum, double *sum64, uint32_t count, int srgb) {
  uint64_t npx = (uint64_t)tile.width * (uint64_t)tile.height;
  hipLaunchKernelGGL(k_encode, dim3(blocks_for(npx)), dim3(256), 0, s, image, width, tile, data, sum,
                     sum64, count, srgb);
}
