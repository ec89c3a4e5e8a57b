// device.h — HBM layout of a resident scene, and the traversal routines over it.
//
// Layout (all arrays in BVH leaf order, uploaded once per geometry version):
//   tri      3 x float4 per triangle (48 B, one 16-byte-aligned record per lane access):
//            {v0.x v0.y v0.z e1.x} {e1.y e1.z e2.x e2.y} {e2.z, side|flags<<8, material, 0}
//   inner    one 64-byte record per inner node of the reference's binary BVH, holding BOTH
//            children's boxes, so one record read serves the two slab tests an inner visit does
//            (bvh.ts:259-260): {L.min L.max.x} {L.max.yz R.min.xy} {R.min.z R.max} {refL refR axis 0}
//   A child ref is an inner-record index (>= 0) or a leaf (bit 31 set): bits 26..30 hold the count
//   (1..31; 0 = a long leaf whose (first, count) sits in `longleaf`), bits 0..25 the first triangle.
//   Node numbering, child order and the split axis are the reference's (bvh.ts:165-171), so the
//   traversals below visit nodes in exactly the reference's order.
#ifndef MCR_DEVICE_H
#define MCR_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jsmath.h"

#define MCR_LEAF_BIT 0x80000000u
#define MCR_LEAF_FIRST_BITS 26
#define MCR_LEAF_FIRST_MASK ((1u << MCR_LEAF_FIRST_BITS) - 1u)
#define MCR_MAX_TRIANGLES (1u << MCR_LEAF_FIRST_BITS)
#define MCR_STACK_LDS 16      // per-lane traversal stack entries kept in LDS
#define MCR_STACK_MAX 128     // total per-lane stack (tree depth <= 100, bvh.ts:5)

// Device functions the traversal and shading kernels must not call out of line (a call spills the
// ray and hit state to scratch and makes every field pointer per-lane).
#define MCR_INLINE __device__ __attribute__((always_inline)) inline

struct DevMaterial {
  int model, side, transparent, vertex_colors;
  int tone_mapped, physical, pad0, pad1;
  double color[3];
  double opacity, alpha_test;
  double emissive[3];
  double roughness, metalness, ior, specular_intensity;
  double specular_color[3];
  double transmission, thickness;
  double attenuation[3];
  double shininess;
  double specular[3];
  // maps and environments (textures.ts, tracer.ts:151-262, 336-406); -1 = none
  int map, alpha_map, emissive_map, normal_map;
  int ao_map, roughness_map, metalness_map, env_map;
  int combine, textured;  // textured: one of the maps above is set (uvs are read)
  double normal_scale[2];
  double ao_map_intensity, env_map_intensity, reflectivity;
};

// CompiledTexture (types.ts:8-21) in HBM: texels are uint8 RGBA (is_float 0) or float RGBA.
struct DevTexture {
  const void *data;
  int32_t width, height;
  int32_t is_float, srgb, flip_y, wrap_s, wrap_t, pad;
  double matrix[9];
};

// CompiledEnvironment (types.ts:23-39).
struct DevEnvironment {
  int32_t kind, texture;
  int32_t faces[6];
  int32_t flip_x, pad;
  double rotation[9];
  double intensity;
};

struct DevLight {
  int type, cast_shadow, pad0, pad1;
  double color[3], ground_color[3], direction[3], position[3];
  double distance, decay, cone_cos, penumbra_cos;
};

struct DevScene {
  const float4 *tri;        // 3 per triangle
  const float4 *inner;      // 4 per inner record
  const int2 *longleaf;     // (first, count) of leaves that do not pack into a ref
  const float *normals;     // 9 per triangle
  const float *colors;      // 12 per triangle or null
  const DevMaterial *materials;
  const DevLight *lights;
  const float *tangents;    // 4 per triangle (normal maps)
  const float *uvs;         // 6 per triangle
  const DevTexture *textures;
  const DevEnvironment *environments;
  uint32_t texture_count, environment_count;
  int32_t environment;      // scene.environment, or -1
  int32_t background_texture, background_environment, pad_bg;
  double background_intensity;
  uint32_t triangle_count, inner_count, material_count, light_count;
  int32_t root_ref;         // child ref of node 0, or -1 for an empty BVH
  float root_box[6];
  double epsilon;
  // shading payload scalars
  int background_kind, fog_kind;
  double background_color[3];
  double fog_color[3];
  double fog_near, fog_far, fog_density;
};

struct DevCounters {  // accumulated per wave, flushed with one atomic per field
  unsigned long long closest_queries, any_queries, node_visits, leaf_visits, triangle_tests, paths;
  unsigned long long closest_node_visits, closest_triangle_tests, any_node_visits, any_triangle_tests;
  unsigned long long shade_waves, mixed_view_waves;  // shade waves, and those holding several views
  // Material-sorted shading (k_shade_bin): live paths per shade class, and the waves of the
  // unsorted queue order whose lanes hold more than one class.
  unsigned long long shade_class_lanes[8], class_mixed_waves;
};

// The counters live in MCR_COUNTER_COPIES copies, 256 bytes apart, and block b adds to copy
// b % MCR_COUNTER_COPIES: thousands of waves flushing at a kernel's end then spread over as many
// cache lines instead of queueing on one. Readers sum the copies.
#define MCR_COUNTER_COPIES 64
#define MCR_COUNTER_STRIDE 256
static_assert(sizeof(DevCounters) <= MCR_COUNTER_STRIDE, "one counter copy per stride");

__device__ inline DevCounters *counter_copy(DevCounters *base) {
  return reinterpret_cast<DevCounters *>(reinterpret_cast<char *>(base) +
                                         (size_t)(blockIdx.x % MCR_COUNTER_COPIES) * MCR_COUNTER_STRIDE);
}

__device__ inline bool ref_is_leaf(int32_t ref) { return (uint32_t)ref & MCR_LEAF_BIT; }

MCR_INLINE void leaf_range(const DevScene &s, int32_t ref, int &first, int &count) {
  uint32_t r = (uint32_t)ref;
  uint32_t c = (r >> MCR_LEAF_FIRST_BITS) & 31u;
  if (c != 0) {
    first = (int)(r & MCR_LEAF_FIRST_MASK);
    count = (int)c;
  } else {
    int2 l = s.longleaf[r & MCR_LEAF_FIRST_MASK];
    first = l.x;
    count = l.y;
  }
}

// bvh.ts:336-338
__device__ inline double safe_inverse(double d) { return d != 0 ? 1.0 / d : 1e30; }

// bvh.ts:340-356 on one child box given as 6 floats. EXACT = false uses the hardware's min/max,
// which differs from JS Math.min/max only for NaN operands and the order of -0/+0. Neither can
// change the result here: with finite inverse directions every slab distance is a number, and
// signed zeros only meet comparisons. Traversals take the EXACT form for rays whose inverse
// direction overflows (|d| below 2^-1023), where 0 * inf = NaN can occur.
template <bool EXACT>
MCR_INLINE double box_entry_t(float bx0, float by0, float bz0, float bx1, float by1, float bz1,
                                     double ox, double oy, double oz, double ix, double iy, double iz,
                                     double tmin, double tmax) {
  double t0 = ((double)bx0 - ox) * ix, t1 = ((double)bx1 - ox) * ix;
  double lo, hi;
  if (EXACT) {
    lo = js_min(t0, t1);
    hi = js_max(t0, t1);
  } else {
    lo = __builtin_fmin(t0, t1);
    hi = __builtin_fmax(t0, t1);
  }
  t0 = ((double)by0 - oy) * iy;
  t1 = ((double)by1 - oy) * iy;
  if (EXACT) {
    lo = js_max(lo, js_min(t0, t1));
    hi = js_min(hi, js_max(t0, t1));
  } else {
    lo = __builtin_fmax(lo, __builtin_fmin(t0, t1));
    hi = __builtin_fmin(hi, __builtin_fmax(t0, t1));
  }
  t0 = ((double)bz0 - oz) * iz;
  t1 = ((double)bz1 - oz) * iz;
  if (EXACT) {
    lo = js_max(lo, js_min(t0, t1));
    hi = js_min(hi, js_max(t0, t1));
  } else {
    lo = __builtin_fmax(lo, __builtin_fmin(t0, t1));
    hi = __builtin_fmin(hi, __builtin_fmax(t0, t1));
  }
  if (hi >= lo && hi >= tmin && lo <= tmax) return EXACT ? js_max(lo, tmin) : __builtin_fmax(lo, tmin);
  return JM_INF;
}

// ---- the slab test decided in FP32 first -------------------------------------------------------
// box_entry_t is the reference's FP64 test. Most boxes are far from the ray's boundary cases, and
// for those an FP32 evaluation of the same slab distances, t' = fma(b, 1/d, -o/d), already settles
// the outcome: with |o| and every box coordinate bounded by B (the root box bounds every node box),
// |t' - t| <= 2^-22.99 (B + |o|) |1/d| per axis, where t is the FP64 value the reference computes
// ((b - o) * inv rounded twice). Entry and exit are maxima and minima of these, so each is within
// E = 2^-22.99 max over axes of (B + |o|) |1/d| of its FP64 value, and their FP32 difference
// within 2E plus its own rounding (2^-24 relative). The margin M = 2^-21 max over axes of
// (B + |o|) |1/d| is 3.97 E, so a comparison that clears it is settled: exit - entry >= M in FP32
// gives at least M (1 - 2^-24) - 2E > 0 in FP64, exit >= M gives at least M - E > 0, and the
// thresholds against tmax carry a slack of at least M. A box is called a hit or a miss in FP32
// only when each of the test's three comparisons (exit >= entry, exit >= tmin = 0,
// entry <= tmax) clears its threshold by the margin; otherwise (a ray grazing a box face within
// ~M) the FP64 test decides. So every outcome,
// and every near/far order the closest-hit traversal takes, is the reference's, while the common
// case costs a third of the FP64 test. The traversals use it for rays with tmin = 0 (every ray
// the renderer traces) whose inverse direction is finite and moderate.
struct SlabF32 {
  float ix, iy, iz;        // fl32(1/d)
  float oix, oiy, oiz;     // fl32(o * (1/d)) of the FP64 product
  float M;                 // the margin, rounded up
  bool ok;                 // the pre-test applies (otherwise the FP64 test only)
};

// FP32 thresholds for the test entry <= t: an estimate at or below `below` is <= t in FP64, one
// above `above` is > t in FP64.
MCR_INLINE void slab_threshold(double t, float M, float &below, float &above) {
  if (!(jm_fabs(t) < JM_INF)) {  // an unbounded ray: every finite estimate is below it
    below = above = (float)t;
    return;
  }
  double slack = jm_fabs(t) * 0x1p-20 + (double)M;  // the relative part absorbs the FP64 rounding of t -/+ M
  below = __double2float_rd(t - slack);
  above = __double2float_ru(t + slack);
}

MCR_INLINE SlabF32 slab_setup(const float *root_box, double ox, double oy, double oz, double ix, double iy,
                              double iz) {
  SlabF32 r;
  float B = 0;
  for (int k = 0; k < 6; k++) B = fmaxf(B, fabsf(root_box[k]));
  double ax = jm_fabs(ix), ay = jm_fabs(iy), az = jm_fabs(iz);
  double mx = ((double)B + jm_fabs(ox)) * ax, my = ((double)B + jm_fabs(oy)) * ay, mz = ((double)B + jm_fabs(oz)) * az;
  double m = __builtin_fmax(__builtin_fmax(mx, my), mz) * 0x1p-21;
  m = __builtin_fmax(m, 0x1p-100);
  r.ok = ax <= 0x1p64 && ay <= 0x1p64 && az <= 0x1p64 && m <= 0x1p80;
  r.ix = (float)ix;
  r.iy = (float)iy;
  r.iz = (float)iz;
  r.oix = (float)(ox * ix);
  r.oiy = (float)(oy * iy);
  r.oiz = (float)(oz * iz);
  r.M = __double2float_ru(m);
  return r;
}

// 1: the FP64 test (tmin = 0) hits; 0: it misses; 2: too close to call. `lo` gets the FP32 entry
// estimate (before clamping to 0). th / tm come from slab_threshold(tmax).
MCR_INLINE int slab_class(float bx0, float by0, float bz0, float bx1, float by1, float bz1, const SlabF32 &r,
                          float th, float tm, float &lo_out) {
  float t0 = __builtin_fmaf(bx0, r.ix, -r.oix), t1 = __builtin_fmaf(bx1, r.ix, -r.oix);
  float lo = __builtin_fminf(t0, t1), hi = __builtin_fmaxf(t0, t1);
  t0 = __builtin_fmaf(by0, r.iy, -r.oiy);
  t1 = __builtin_fmaf(by1, r.iy, -r.oiy);
  lo = __builtin_fmaxf(lo, __builtin_fminf(t0, t1));
  hi = __builtin_fminf(hi, __builtin_fmaxf(t0, t1));
  t0 = __builtin_fmaf(bz0, r.iz, -r.oiz);
  t1 = __builtin_fmaf(bz1, r.iz, -r.oiz);
  lo = __builtin_fmaxf(lo, __builtin_fminf(t0, t1));
  hi = __builtin_fminf(hi, __builtin_fmaxf(t0, t1));
  lo_out = lo;
  float d = hi - lo;
  if (d >= r.M && hi >= r.M && lo <= th) return 1;
  if (d < -r.M || hi < -r.M || lo > tm) return 0;
  return 2;
}

__device__ inline bool finite_inverse(double ix, double iy, double iz) {
  return jm_fabs(ix) < JM_INF && jm_fabs(iy) < JM_INF && jm_fabs(iz) < JM_INF;
}

// Per-lane traversal stack: the first MCR_STACK_LDS entries in LDS (lane-interleaved, so a
// wave's pushes hit distinct banks), the rest in a global spill area owned by this thread.
struct LaneStack {
  int32_t *lds;      // base for this lane: lds[i * stride]
  int32_t *spill;    // base for this lane: spill[(i - LDS) * spill_stride]
  int stride, spill_stride;
  int sp;
  __device__ void push(int32_t v) {
    if (sp < MCR_STACK_LDS) lds[sp * stride] = v;
    else spill[(size_t)(sp - MCR_STACK_LDS) * spill_stride] = v;
    sp++;
  }
  __device__ int32_t pop() {
    sp--;
    if (sp < MCR_STACK_LDS) return lds[sp * stride];
    return spill[(size_t)(sp - MCR_STACK_LDS) * spill_stride];
  }
};

// Declares a traversal kernel's per-lane stack: MCR_STACK_LDS entries per lane in LDS (blocks of
// 256 threads), the rest in the lane's part of the global spill area.
#define MCR_LANE_STACK(st, spill_base, spill_stride_)                     \
  __shared__ int32_t lds_stack_[MCR_STACK_LDS * 256];                     \
  LaneStack st;                                                          \
  st.lds = lds_stack_ + threadIdx.x;                                      \
  st.stride = 256;                                                        \
  st.spill = (spill_base) + (blockIdx.x * blockDim.x + threadIdx.x);      \
  st.spill_stride = (int)(spill_stride_)

struct TriHit {
  int tri;
  double t, u, v;
  bool front;
};

struct TravCount {
  uint32_t nodes, leaves, tris;
};

// Möller–Trumbore in FP64 on one triangle record (bvh.ts:228-252). Returns false if missed;
// otherwise t, u, v and det.
MCR_INLINE bool mt_test(const float4 a, const float4 b, const float4 c, double ox, double oy,
                               double oz, double dx, double dy, double dz, int cull_side,
                               double &t, double &u, double &v, double &det_out) {
  double e1x = a.w, e1y = b.x, e1z = b.y;
  double e2x = b.z, e2y = b.w, e2z = c.x;
  double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
  double det = e1x * px + e1y * py + e1z * pz;
  if (det == 0) return false;
  if (cull_side == 0 && det < 0) return false;
  if (cull_side == 1 && det > 0) return false;
  double inv = 1.0 / det;
  double sx = ox - (double)a.x, sy = oy - (double)a.y, sz = oz - (double)a.z;
  u = (sx * px + sy * py + sz * pz) * inv;
  if (u < 0 || u > 1) return false;
  double qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
  v = (dx * qx + dy * qy + dz * qz) * inv;
  if (v < 0 || u + v > 1) return false;
  t = (e2x * qx + e2y * qy + e2z * qz) * inv;
  det_out = det;
  return true;
}

// How a traversal evaluates the reference's FP64 slab test (bvh.ts:340-356). Every mode gives the
// reference's outcomes; they differ in cost and in the rays they apply to.
enum SlabMode {
  SLAB_JS = 0,    // FP64 with JS Math.min/max: rays whose inverse direction overflows (NaN possible)
  SLAB_FP64 = 1,  // FP64 with the hardware's min/max (identical without NaN)
  SLAB_FP32 = 2,  // the FP32 pre-test, FP64 where it cannot decide (slab_setup's ok holds)
};

// The FP64 inverse direction for the FP64 test: SLAB_FP32 traversals recompute it where an FP32
// outcome was too close to call (rare), which gives the same bits and keeps six registers free
// over the loop. `salt` varies with the loop (a box word): adding 0 * salt (a signed zero; every
// component of d is nonzero in this mode) leaves d unchanged but ties the divisions to the loop,
// so the compiler does not hoist them out and keep their results live after all.
MCR_INLINE void fresh_inverse(double dx, double dy, double dz, int32_t salt, double &ix, double &iy, double &iz) {
  double z = 0.0 * (double)salt;
  ix = safe_inverse(dx + z);
  iy = safe_inverse(dy + z);
  iz = safe_inverse(dz + z);
}

// The precision form's refinement of a first hit found in FP32: triangle k's FP64 Möller–Trumbore
// (bvh.ts:228-252) on the FP64 ray gives t, u and v exactly as the reference computes them for that
// triangle. Returns false (values untouched) where the FP64 test rejects the triangle that the FP32
// watertight test accepted: a ray through an edge, within FP32 rounding of it.
MCR_INLINE bool refine_hit(const float4 *tri, int k, double ox, double oy, double oz, double dx, double dy,
                           double dz, double &t, double &u, double &v) {
  const float4 *r = tri + (size_t)k * 3;
  double tt, uu, vv, det;
  if (!mt_test(r[0], r[1], r[2], ox, oy, oz, dx, dy, dz, 2, tt, uu, vv, det)) return false;
  t = tt;
  u = uu;
  v = vv;
  return true;
}

// bvh.ts:202-282 over the device layout. cull: drop back faces of FrontSide and front faces of
// BackSide triangles. Near/far orders use the FP64 entry distances (exact-t ties between
// triangles are broken by visit order, so the order must be the reference's).
template <bool EXACT>
MCR_INLINE bool closest_impl(const DevScene &s, double ox, double oy, double oz, double dx,
                                    double dy, double dz, double ix, double iy, double iz,
                                    double tmin, bool cull, TriHit &hit, LaneStack &st,
                                    TravCount &cnt) {
  bool neg[3] = {ix < 0, iy < 0, iz < 0};
  st.sp = 0;
  int32_t ref = s.root_ref;
  for (;;) {
    if (ref_is_leaf(ref)) {
      int first, count;
      leaf_range(s, ref, first, count);
      cnt.leaves++;
      for (int k = first; k < first + count; k++) {
        cnt.tris++;
        const float4 *r = s.tri + (size_t)k * 3;
        float4 a = r[0], b = r[1], c = r[2];
        int side = cull ? (int)(__float_as_uint(c.y) & 0xffu) : 2;
        double t, u, v, det;
        if (!mt_test(a, b, c, ox, oy, oz, dx, dy, dz, side, t, u, v, det)) continue;
        if (t > tmin && t < hit.t) {
          hit.t = t;
          hit.u = u;
          hit.v = v;
          hit.tri = k;
          hit.front = det > 0;
        }
      }
    } else {
      cnt.nodes++;
      const float4 *r = s.inner + (size_t)ref * 4;
      float4 p = r[0], q = r[1], w = r[2];
      int4 m = *reinterpret_cast<const int4 *>(r + 3);
      bool right_first = neg[m.z];
      double tl = box_entry_t<EXACT>(p.x, p.y, p.z, p.w, q.x, q.y, ox, oy, oz, ix, iy, iz, tmin, hit.t);
      double tr = box_entry_t<EXACT>(q.z, q.w, w.x, w.y, w.z, w.w, ox, oy, oz, ix, iy, iz, tmin, hit.t);
      int32_t near_ref = right_first ? m.y : m.x, far_ref = right_first ? m.x : m.y;
      double tn = right_first ? tr : tl, tf = right_first ? tl : tr;
      if (tn < JM_INF) {
        if (tf < JM_INF) {
          if (tf < tn) {
            st.push(near_ref);
            ref = far_ref;
          } else {
            st.push(far_ref);
            ref = near_ref;
          }
        } else {
          ref = near_ref;
        }
        continue;
      }
      if (tf < JM_INF) {
        ref = far_ref;
        continue;
      }
    }
    if (st.sp == 0) break;
    ref = st.pop();
  }
  return hit.tri >= 0;
}

MCR_INLINE bool intersect_closest(const DevScene &s, double ox, double oy, double oz,
                                         double dx, double dy, double dz, double tmin, double tmax,
                                         bool cull, TriHit &hit, LaneStack &st, TravCount &cnt) {
  hit.tri = -1;
  hit.t = tmax;
  hit.u = 0;
  hit.v = 0;
  hit.front = true;
  if (s.root_ref == -1) return false;
  double ix = safe_inverse(dx), iy = safe_inverse(dy), iz = safe_inverse(dz);
  if (finite_inverse(ix, iy, iz))
    return closest_impl<false>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin, cull, hit, st, cnt);
  return closest_impl<true>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin, cull, hit, st, cnt);
}

// The pre-test's parameters for an any-hit ray (tmin = 0), or ok = false where only FP64 applies.
MCR_INLINE SlabF32 slab_for(const DevScene &s, double ox, double oy, double oz, double ix, double iy, double iz,
                            double tmin) {
  SlabF32 sl;
  sl.ok = false;
  if (tmin == 0 && finite_inverse(ix, iy, iz)) sl = slab_setup(s.root_box, ox, oy, oz, ix, iy, iz);
  return sl;
}

MCR_INLINE int slab_mode(const SlabF32 &sl, double ix, double iy, double iz) {
  if (sl.ok) return SLAB_FP32;
  return finite_inverse(ix, iy, iz) ? SLAB_FP64 : SLAB_JS;
}

// Outcome of one child box of an any-hit traversal (fixed tmin, tmax); `entry` estimates where
// the ray enters it (used to order visits only).
template <int MODE>
MCR_INLINE bool child_hit(float bx0, float by0, float bz0, float bx1, float by1, float bz1, const SlabF32 &sl,
                          float th, float tm, double ox, double oy, double oz, double dx, double dy, double dz,
                          double ix, double iy, double iz, double tmin, double tmax, float &entry) {
  if (MODE == SLAB_FP32) {
    int c = slab_class(bx0, by0, bz0, bx1, by1, bz1, sl, th, tm, entry);
    if (c != 2) return c == 1;
    fresh_inverse(dx, dy, dz, __float_as_int(bx0), ix, iy, iz);
  }
  double t = box_entry_t<MODE == SLAB_JS>(bx0, by0, bz0, bx1, by1, bz1, ox, oy, oz, ix, iy, iz, tmin, tmax);
  entry = (float)t;
  return t < JM_INF;
}

// bvh.ts:289-332 over the device layout: every hit in (tmin, tmax), no culling, in the
// reference's depth-first order (right child first). Both child boxes are tested when their parent
// is visited; a box test does not depend on when it runs, so the visit order is unchanged.
// visit(k, t, u, v, front) returns false to stop.
template <int MODE, class Visit>
MCR_INLINE void any_impl(const DevScene &s, double ox, double oy, double oz, double dx,
                                double dy, double dz, double ix, double iy, double iz, const SlabF32 &sl,
                                double tmin, double tmax, LaneStack &st, TravCount &cnt, Visit &visit) {
  const float *rb = s.root_box;
  if (!(box_entry_t<MODE == SLAB_JS>(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], ox, oy, oz, ix, iy, iz, tmin, tmax) <
        JM_INF))
    return;
  float th = 0, tm = 0;
  if (MODE == SLAB_FP32) slab_threshold(tmax, sl.M, th, tm);
  st.sp = 0;
  int32_t ref = s.root_ref;
  for (;;) {
    if (ref_is_leaf(ref)) {
      int first, count;
      leaf_range(s, ref, first, count);
      cnt.leaves++;
      for (int k = first; k < first + count; k++) {
        cnt.tris++;
        const float4 *r = s.tri + (size_t)k * 3;
        float4 a = r[0], b = r[1], c = r[2];
        double t, u, v, det;
        if (!mt_test(a, b, c, ox, oy, oz, dx, dy, dz, 2, t, u, v, det)) continue;
        if (t > tmin && t < tmax && !visit(k, t, u, v, det > 0)) return;
      }
    } else {
      cnt.nodes++;
      const float4 *r = s.inner + (size_t)ref * 4;
      float4 p = r[0], q = r[1], w = r[2];
      int4 m = *reinterpret_cast<const int4 *>(r + 3);
      float el, er;
      bool hl = child_hit<MODE>(p.x, p.y, p.z, p.w, q.x, q.y, sl, th, tm, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin,
                                tmax, el);
      bool hr = child_hit<MODE>(q.z, q.w, w.x, w.y, w.z, w.w, sl, th, tm, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin,
                                tmax, er);
      if (hr) {
        if (hl) st.push(m.x);
        ref = m.y;
        continue;
      }
      if (hl) {
        ref = m.x;
        continue;
      }
    }
    if (st.sp == 0) break;
    ref = st.pop();
  }
}

template <class Visit>
MCR_INLINE void intersect_any(const DevScene &s, double ox, double oy, double oz, double dx,
                                     double dy, double dz, double tmin, double tmax, LaneStack &st,
                                     TravCount &cnt, Visit &visit) {
  if (s.root_ref == -1) return;
  double ix = safe_inverse(dx), iy = safe_inverse(dy), iz = safe_inverse(dz);
  SlabF32 sl = slab_for(s, ox, oy, oz, ix, iy, iz, tmin);
  switch (slab_mode(sl, ix, iy, iz)) {
    case SLAB_FP32: any_impl<SLAB_FP32>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, visit); break;
    case SLAB_FP64: any_impl<SLAB_FP64>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, visit); break;
    default: any_impl<SLAB_JS>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, visit);
  }
}

// Shadow-ray query (tracer.ts:529-594 over bvh.ts:289-332). The reference visits every hit in
// (tmin, tmax) in depth-first order, right child first, and stops at the first opaque one. The
// set of triangles it can visit depends only on box tests with the fixed (tmin, tmax), never on
// the order, and an opaque hit anywhere in that set makes the light invisible whatever came
// before it. So this traversal takes the nearer child first (by entry estimates, which only order
// the visits) and stops at the first opaque hit; see-through hits are handed to on_hit with their
// leaf's first triangle, so the caller can replay them in the reference's order (leaves by
// descending first triangle, then ascending index).
// on_hit(k, leaf_first, t, u, v, front, record) returns true for an opaque hit.
template <int MODE, class OnHit>
MCR_INLINE bool occluded_impl(const DevScene &s, double ox, double oy, double oz, double dx,
                                     double dy, double dz, double ix, double iy, double iz,
                                     const SlabF32 &sl, double tmin, double tmax,
                                     LaneStack &st, TravCount &cnt, OnHit &on_hit) {
  const float *rb = s.root_box;
  if (!(box_entry_t<MODE == SLAB_JS>(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], ox, oy, oz, ix, iy, iz, tmin, tmax) <
        JM_INF))
    return false;
  float th = 0, tm = 0;
  if (MODE == SLAB_FP32) slab_threshold(tmax, sl.M, th, tm);
  st.sp = 0;
  int32_t ref = s.root_ref;
  for (;;) {
    if (ref_is_leaf(ref)) {
      int first, count;
      leaf_range(s, ref, first, count);
      cnt.leaves++;
      for (int k = first; k < first + count; k++) {
        cnt.tris++;
        const float4 *r = s.tri + (size_t)k * 3;
        float4 a = r[0], b = r[1], c = r[2];
        double t, u, v, det;
        if (!mt_test(a, b, c, ox, oy, oz, dx, dy, dz, 2, t, u, v, det)) continue;
        if (t > tmin && t < tmax && on_hit(k, first, t, u, v, det > 0, c)) return true;
      }
    } else {
      cnt.nodes++;
      const float4 *r = s.inner + (size_t)ref * 4;
      float4 p = r[0], q = r[1], w = r[2];
      int4 m = *reinterpret_cast<const int4 *>(r + 3);
      float el, er;
      bool hl = child_hit<MODE>(p.x, p.y, p.z, p.w, q.x, q.y, sl, th, tm, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin,
                                tmax, el);
      bool hr = child_hit<MODE>(q.z, q.w, w.x, w.y, w.z, w.w, sl, th, tm, ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin,
                                tmax, er);
      if (hl && hr) {
        bool left_first = el <= er;
        st.push(left_first ? m.y : m.x);
        ref = left_first ? m.x : m.y;
        continue;
      }
      if (hl) {
        ref = m.x;
        continue;
      }
      if (hr) {
        ref = m.y;
        continue;
      }
    }
    if (st.sp == 0) break;
    ref = st.pop();
  }
  return false;
}

template <class OnHit>
MCR_INLINE bool occluded(const DevScene &s, double ox, double oy, double oz, double dx,
                                double dy, double dz, double tmin, double tmax, LaneStack &st,
                                TravCount &cnt, OnHit &on_hit) {
  if (s.root_ref == -1) return false;
  double ix = safe_inverse(dx), iy = safe_inverse(dy), iz = safe_inverse(dz);
  SlabF32 sl = slab_for(s, ox, oy, oz, ix, iy, iz, tmin);
  switch (slab_mode(sl, ix, iy, iz)) {
    case SLAB_FP32:
      return occluded_impl<SLAB_FP32>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, on_hit);
    case SLAB_FP64:
      return occluded_impl<SLAB_FP64>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, on_hit);
    default:
      return occluded_impl<SLAB_JS>(s, ox, oy, oz, dx, dy, dz, ix, iy, iz, sl, tmin, tmax, st, cnt, on_hit);
  }
}

// Sum over the 64 lanes of a wave.
__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline void flush_counts(DevCounters *c, const TravCount &t, unsigned long long closest,
                                    unsigned long long any) {
  unsigned long long n = wave_sum(t.nodes), l = wave_sum(t.leaves), k = wave_sum(t.tris);
  unsigned long long q = wave_sum(closest), a = wave_sum(any);
  c = counter_copy(c);
  if ((threadIdx.x & 63) == 0) {
    if (n) atomicAdd(&c->node_visits, n);
    if (l) atomicAdd(&c->leaf_visits, l);
    if (k) atomicAdd(&c->triangle_tests, k);
    if (q) atomicAdd(&c->closest_queries, q);
    if (a) atomicAdd(&c->any_queries, a);
    // A kernel runs one kind of query, so its node and triangle counts belong to that kind.
    if (q) {
      atomicAdd(&c->closest_node_visits, n);
      atomicAdd(&c->closest_triangle_tests, k);
    } else if (a) {
      atomicAdd(&c->any_node_visits, n);
      atomicAdd(&c->any_triangle_tests, k);
    }
  }
}

#endif  // MCR_DEVICE_H
