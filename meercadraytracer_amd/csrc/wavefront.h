// wavefront.h — path state of the wavefront tracer, structure-of-arrays in HBM.
//
// One path slot per camera sample in flight. Each bounce iteration runs four kernels over the
// live slots: extend (closest hit), shade (surface, lights, continuation), shadow (any-hit for the
// shadow rays shade emitted) and resolve (sum the light terms in the reference's order, finish or
// requeue). Every per-sample quantity is kept in FP64, as the reference's JS numbers are.
#ifndef MCR_WAVEFRONT_H
#define MCR_WAVEFRONT_H

#include "../../include/mcr.h"
#include "wide.h"

#define MCR_MAX_LIGHTS MCR_MAX_LIGHTS_PER_SCENE  // lights per scene (mcr.h)
#define MCR_SHADOW_REGIONS 16      // regions of the shadow-ray queue; light li queues in region li % 16
#define MCR_FETCH_COUNTERS 8       // interleaved work counters of the traversal kernels
#define MCR_SHADOW_HIT_CAP 64      // see-through hits a shadow traversal keeps for its ordered replay
#define MCR_DEEP_HITS 1024         // counted see-through surfaces a replayed shadow ray may cross
#define MCR_DEEP_THREADS (64 * 256)  // threads of the shadow kernel's reference-order pass
#define MCR_SLOT_BITS 26           // a shadow-queue entry: light << 26 | slot (capacity <= 2^26)
#define MCR_SHADE_CLASSES 8        // shade classes the live queue is binned by before shade

// Double fields of a path slot.
enum PathField {
  F_OX, F_OY, F_OZ, F_DX, F_DY, F_DZ,   // current ray
  F_TR, F_TG, F_TB,                     // throughput
  F_LR, F_LG, F_LB,                     // radiance
  F_MR, F_MG, F_MB,                     // absorption of the current volume
  F_QR, F_QG, F_QB,                     // throughput at the vertex being resolved
  F_AR, F_AG, F_AB,                     // out* before the lights (basic: the colour)
  F_IR, F_IG, F_IB,                     // indirect term added after the lights
  F_ER, F_EG, F_EB,                     // emission
  F_FOG,                                // fog factor of the first surface
  F_HT, F_HU, F_HV,                     // hit t, u, v
  F_ENVS,                               // lastEnvironmentScale (tracer.ts:427-458)
  F_XR, F_XG, F_XB, F_XS,               // envMap sample and reflectivity of a non-physical vertex
  F_COUNT
};

// Integer fields. I_VIEW is the view (scene, camera, settings) of the slot's sample.
// I_ENV is lastEnvironment (tracer.ts:427-458), -1 for none. I_JOB is the index the sample's job
// had in the job table when the sample started: a hint, checked against the table before use,
// since a served run may rewrite the table while samples are in flight.
enum PathIntField { I_RNG, I_FLAGS, I_BOUNCE, I_PASSES, I_SAMPLE, I_HTRI, I_VIEW, I_ENV, I_JOB, I_COUNT };

enum PathFlags {
  PF_SURFACE = 1, PF_TONEMAP = 2, PF_CULL = 4, PF_FRONT = 8, PF_INDIRECT = 16,
  PF_COMBINE_SHIFT = 5,  // 2 bits: 0 none, else the envMap combine operation + 1 (tracer.ts:393-406)
  PF_COMBINE_MASK = 3 << 5,
};

// What resolve does with a slot after shade.
enum PathMode {
  MODE_DONE = 0,        // sample written (background), nothing more
  MODE_EXTEND = 1,      // pass-through: trace again, nothing to add
  MODE_ADD_EXTEND = 2,  // add this vertex's light, then trace the continuation
  MODE_ADD_FINISH = 3,  // add this vertex's light, then write the sample
  MODE_FINISH = 4,      // write the sample
};

// Light-slot status after shade / shadow. LS_DEEP: the shadow ray crossed more see-through surfaces
// than the traversal keeps for its ordered replay; the reference-order pass traces it again.
enum LightStatus { LS_NONE = 0, LS_VISIBLE = 1, LS_RAY = 2, LS_BLOCKED = 3, LS_DEEP = 4 };

// Counter words of Wavefront::counts.
enum CountWord {
  C_QUEUE0 = 0,     // live slots in queue 0
  C_QUEUE1 = 1,     // live slots in queue 1
  C_SPARE = 2,      // unused
  C_OVERFLOW = 3,   // OVERFLOW_* bits: a shadow ray met more see-through hits than it can keep, or a
                    // wide traversal met a corrupt node or a full stack
  C_FREE = 4,       // slots on the free list
  C_NEXT = 5,       // next sample to start (may run past Control::total; the host clamps it)
  C_SORT = 6,       // appends of the queue re-sort (k_sort_gather)
  C_DEEP = 7,       // shadow rays of this iteration marked LS_DEEP
  C_EXTEND_FETCH = 8,                            // MCR_FETCH_COUNTERS words (WaveFetch)
  C_SHADOW_FETCH = C_EXTEND_FETCH + MCR_FETCH_COUNTERS,
  // Shadow rays queued by shade, two words per region: word 2*r counts rays of first-hit vertices
  // (stored from the front of region r of sref), word 2*r+1 rays of later vertices (from the back
  // of that region).
  C_SHADOW_SEG = C_SHADOW_FETCH + MCR_FETCH_COUNTERS,
  C_SHADOW_SEG_END = C_SHADOW_SEG + 2 * MCR_SHADOW_REGIONS,
  // Live slots binned by shade class before shade (k_shade_bin), one word per class.
  C_BIN = C_SHADOW_SEG_END,
  C_BIN_END = C_BIN + MCR_SHADE_CLASSES,
  C_WORDS = C_BIN_END,
};

enum OverflowBits { OVERFLOW_SEEN = 1, OVERFLOW_WIDE = 2 };

// Shadow-queue entries: the ray of light li from path slot `slot`.
__host__ __device__ constexpr inline uint32_t sref_entry(uint32_t slot, uint32_t li) { return (li << MCR_SLOT_BITS) | slot; }
__host__ __device__ constexpr inline uint32_t sref_slot(uint32_t e) { return e & ((1u << MCR_SLOT_BITS) - 1u); }
__host__ __device__ constexpr inline uint32_t sref_light(uint32_t e) { return e >> MCR_SLOT_BITS; }

// Each counter word sits on a 128-byte cache line of its own: atomics on one line serialise, and
// the traversal kernels' work counters must not queue behind each other or behind the queue
// lengths. Wavefront::counts holds C_WORDS such lines; word x is counts[cw(x)].
#define MCR_COUNT_STRIDE 32
__host__ __device__ constexpr inline uint32_t cw(uint32_t word) { return word * MCR_COUNT_STRIDE; }
// After the counter lines, one bit per view (packed, view v at bit v % 32 of word C_FAIL_BITS +
// v / 32): a shadow ray of that view overflowed the reference-order pass. A server fails only the
// requests whose bits are set; views from MCR_FAIL_VIEWS on set OVERFLOW_SEEN alone.
#define MCR_FAIL_VIEWS 1024
const uint32_t C_FAIL_BITS = C_WORDS * MCR_COUNT_STRIDE;
const uint32_t COUNT_WORDS_ALLOC = C_FAIL_BITS + MCR_FAIL_VIEWS / 32;

// Work distribution of the traversal kernels over a queue of n items in 64-item chunks. One
// counter would serialise every wave's atomic on one cache line; MCR_FETCH_COUNTERS counters
// interleave instead: counter k hands out chunks k, k + K, k + 2K, ... (K counters), each wave
// starts on counter (its wave index mod K) and moves on to the next counter when its own runs out.
// The counters advance together, so the chunks in flight still sweep the queue front to back
// (neighbouring paths stay in flight together) while each counter sees a K-th of the atomics.
// Every lane of the wave must call next(); it returns the chunk's first item, or n when none is left.
struct WaveFetch {
  uint32_t *counters;
  uint32_t n, k, tried;  // wave-uniform
  __device__ WaveFetch(uint32_t *c, uint32_t items)
      : counters(c), n(items), k((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % MCR_FETCH_COUNTERS),
        tried(0) {}
  __device__ uint32_t next() {
    uint32_t chunks = (n + 63) / 64;
    while (tried < MCR_FETCH_COUNTERS) {
      uint32_t j = 0;
      if ((threadIdx.x & 63) == 0) j = atomicAdd(&counters[cw(k)], 1u);
      j = __builtin_amdgcn_readfirstlane(__shfl(j, 0, 64));
      uint64_t chunk = (uint64_t)j * MCR_FETCH_COUNTERS + k;
      if (chunk < chunks) return (uint32_t)chunk * 64u;
      k = k + 1 == MCR_FETCH_COUNTERS ? 0 : k + 1;
      tried++;
    }
    return n;
  }
};

struct JobDev;

// What the kernels of one call read besides the path state. It lives in device memory so a
// captured iteration can be replayed while the host grows the job list between replays.
struct Control {
  const JobDev *jobs;        // sorted by sample_start
  float *out;                // tile results: job.out_offset floats from here
  uint32_t *job_done;        // finished samples per job, or null
  uint32_t njobs;
  uint32_t total;            // samples that may be started: [0, total)
};

// A view: the scene, camera and settings a sample is traced with. Jobs of different views can
// share one wavefront run; kernels make the view wave-uniform before they read it.
struct ViewDev {
  DevScene sc;
  WideScene wide;            // the precision form's layout, when the run uses it
  mcr_camera cam;
  mcr_settings st;
  uint32_t light_slots;      // the scene's light count
  uint32_t shadow_slots;     // light_slots, plus the environment-diffuse slot when the scene has environments
  uint32_t shadow_tex;       // a material has map or alphaMap, which shadow rays sample
  uint32_t pad;
};

struct Wavefront {
  uint32_t capacity;         // path slots
  uint32_t lights;           // light slots per path slot (the largest shadow_slots of the views it runs)
  uint32_t region_cap;       // entries of each shadow-queue region: ceil(lights / 16) * capacity
  double *f;                 // F_COUNT * capacity
  int32_t *i;                // I_COUNT * capacity
  uint8_t *mode;             // capacity
  // light slots, indexed (row * lights + light) * capacity + slot
  double *contrib;           // 3 rows: the unshadowed r, g, b of the light at this vertex
  double *shadow;            // 3 rows: shadowColor after the shadow ray
  double *sray;              // 7 rows: shadow ray origin, direction, maxT
  uint8_t *status;           // 1 row: LightStatus
  // queues
  int32_t *queue[2];         // live slots
  uint32_t *counts;          // CountWord
  uint32_t *sref;            // shadow queue: min(lights, 16) regions of region_cap sref_entry values
  int32_t *free_list;        // slots with no sample
  int32_t *bins;             // MCR_SHADE_CLASSES regions of capacity entries: the live queue by shade class
  double *deep_seen;         // MCR_DEEP_HITS counted distances per thread of the reference-order pass
  const Control *ctl;
};

struct JobDev {  // one tile job, device side
  int32_t tx, ty, tw, th;
  uint32_t width, height, pass, block;
  uint64_t out_offset;
  uint32_t sample_start, cols;  // samples [sample_start, + cols * rows): cols x rows blocks
  uint32_t view, rows;
};

// One tile result to accumulate and encode (writeTile): sum may be null (a preview tile).
struct EncodeItem {
  uint8_t *image;
  float *sum;
  double *sum64;             // FP64 running sum instead of sum (renders beyond 1024 samples)
  const float *data;
  uint32_t width;
  int32_t tx, ty, tw, th;
  uint32_t count;
  int32_t srgb;
  uint32_t block_start;      // first block of the launch that works on this item
};

#endif  // MCR_WAVEFRONT_H
