// wavefront.h — path state of the wavefront tracer, structure-of-arrays in HBM.
//
// One path slot per camera sample in flight. Each bounce iteration runs four kernels over the
// live slots: extend (closest hit), shade (surface, lights, continuation), shadow (any-hit for the
// shadow rays shade emitted) and resolve (sum the light terms in the reference's order, finish or
// requeue). Every per-sample quantity is kept in FP64, as the reference's JS numbers are.
#ifndef MCR_WAVEFRONT_H
#define MCR_WAVEFRONT_H

#include "device.h"

#define MCR_MAX_LIGHTS 16          // punctual lights traced per scene
#define MCR_SHADOW_HIT_CAP 64      // see-through surfaces remembered per shadow ray (tracer.ts:48)

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
  F_COUNT
};

// Integer fields.
enum PathIntField { I_RNG, I_FLAGS, I_BOUNCE, I_PASSES, I_SAMPLE, I_HTRI, I_COUNT };

enum PathFlags {
  PF_SURFACE = 1, PF_TONEMAP = 2, PF_CULL = 4, PF_FRONT = 8, PF_INDIRECT = 16,
};

// What resolve does with a slot after shade.
enum PathMode {
  MODE_DONE = 0,        // sample written (background), nothing more
  MODE_EXTEND = 1,      // pass-through: trace again, nothing to add
  MODE_ADD_EXTEND = 2,  // add this vertex's light, then trace the continuation
  MODE_ADD_FINISH = 3,  // add this vertex's light, then write the sample
  MODE_FINISH = 4,      // write the sample
};

// Light-slot status after shade / shadow.
enum LightStatus { LS_NONE = 0, LS_VISIBLE = 1, LS_RAY = 2, LS_BLOCKED = 3 };

struct Wavefront {
  uint32_t capacity;         // path slots
  double *f;                 // F_COUNT * capacity
  int32_t *i;                // I_COUNT * capacity
  uint8_t *mode;             // capacity
  // light slots, indexed (row * MCR_MAX_LIGHTS + light) * capacity + slot
  double *contrib;           // 3 rows: the unshadowed r, g, b of the light at this vertex
  double *shadow;            // 3 rows: shadowColor after the shadow ray
  double *sray;              // 7 rows: shadow ray origin, direction, maxT
  uint8_t *status;           // 1 row: LightStatus
  // queues
  int32_t *queue[2];         // live slots
  uint32_t *counts;          // [0],[1] queue lengths, [2] shadow queue length, [3] overflow flag,
                             // [5] next sample to start, [6],[7] extend / shadow work counters
  int32_t *sref;             // shadow queue: slot * MCR_MAX_LIGHTS + light
  uint32_t seed;             // settings.seed of the call (random.ts:6 mixes it into every seed)
};

struct JobDev {  // one tile job, device side
  int32_t tx, ty, tw, th;
  uint32_t width, height, pass, block;
  uint64_t out_offset;
  uint32_t sample_start, cols;
};

#endif  // MCR_WAVEFRONT_H
