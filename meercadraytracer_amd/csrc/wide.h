// wide.h — the precision form's scene layout (a compressed 8-wide BVH over FP32 triangles) and
// its closest-hit query: a conservative FP32 slab test on 8-bit quantized child boxes and the
// watertight ray/triangle test of Woop, Benthin and Wald (JCGT 2(1), 2013).
//
// Node (80 B = 5 x uint4, one 16-byte load per part):
//   n0 {origin x, y, z (f32 bits), ex | ey << 8 | ez << 16 | imask << 24}
//   n1 {child_base, tri_base, meta bytes of slots 0-3, meta bytes of slots 4-7}
//   n2 {qlo.x slots 0-3, qlo.x 4-7, qlo.y 0-3, qlo.y 4-7}
//   n3 {qlo.z 0-3, qlo.z 4-7, qhi.x 0-3, qhi.x 4-7}
//   n4 {qhi.y 0-3, qhi.y 4-7, qhi.z 0-3, qhi.z 4-7}
// Slot i's box on axis a is [o_a + qlo_a[i] * 2^(e_a - 127), o_a + qhi_a[i] * 2^(e_a - 127)]: the
// FP32 box of the binary node it stands for, rounded outward to an 8-bit grid over the node. imask
// bit i marks an inner child, node child_base + (number of inner slots below i). Otherwise meta byte
// i is the slot's leaf triangle count (0 = empty slot), and its triangles start at tri_base + (sum
// of the meta bytes below i). Slot i holds the child lying towards octant i of the node (bit a set
// = positive side of axis a), so taking hit slots in the order i ^ (the ray's negative-axis bits)
// runs roughly front to back.
//
// Triangle (48 B = 3 x float4): {v0, index} {v1, side} {v2, leaf}. v1 and v2 are the reference's FP64
// vertex sums v0 + e1 and v0 + e2 (bvh.ts:26-27) rounded once to FP32; the binary node boxes are
// those sums' FP32 bounds (bvh.ts:28-31), and rounding to nearest is monotone, so every FP32
// triangle lies inside its leaf's box. index is the triangle's position in BVH order, the id the
// FP64 queries return; leaf is the first triangle of the binary leaf it came from (shadow rays
// replay see-through hits in the reference's leaf order).
#ifndef MCR_WIDE_H
#define MCR_WIDE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

#define MCR_WIDE_LEAF_MAX 16      // triangles in one leaf slot (buildBvh's SAH leaves hold <= 16)
#define MCR_WIDE_STACK_LDS 12     // per-lane node-group stack entries kept in LDS
// Entries beyond those go to the engine's per-thread spill area, two words each.
#define MCR_WIDE_STACK_MAX (MCR_WIDE_STACK_LDS + (MCR_STACK_MAX - MCR_STACK_LDS) / 2)

#define WHD __host__ __device__ __attribute__((always_inline)) inline

struct WideScene {
  const uint4 *nodes;    // 5 per node
  const float4 *tri;     // 3 per triangle
  uint32_t node_count, triangle_count;
};

// A child being placed in a wide node: a binary inner record (inner >= 0) or a triangle range
// with its box. Ranges over MCR_WIDE_LEAF_MAX (leaves buildBvh makes at its depth limit) are split
// in halves under the same box as the collapse needs slots.
struct WideChild {
  int32_t inner;
  uint32_t first, count;
  uint32_t leaf;           // first triangle of the binary leaf a range came from
  float b[6];
};

WHD bool wide_expandable(const WideChild &c) { return c.inner >= 0 || c.count > MCR_WIDE_LEAF_MAX; }

WHD float wide_half_area(const float *b) {
  float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

// Smallest e in [-126, 127] with 255 * 2^e >= hi - lo (exact in FP64).
WHD int wide_exponent(float lo, float hi) {
  double ext = (double)hi - (double)lo;
  if (!(ext > 0)) return -126;
  int e = (int)ceil(log2(ext / 255.0));
  if (e < -126) e = -126;
  if (e > 127) e = 127;
  while (e < 127 && ldexp(255.0, e) < ext) e++;
  while (e > -126 && ldexp(255.0, e - 1) >= ext) e--;
  return e;
}

// Outward 8-bit grid coordinates of [clo, chi] over origin p with step 2^e. Returns false if chi
// does not fit under 255 steps (the caller then takes the next exponent).
WHD bool wide_quantize(float p, int e, float clo, float chi, uint32_t &qlo, uint32_t &qhi) {
  double s = ldexp(1.0, e);
  double a = floor(((double)clo - (double)p) / s);
  if ((double)p + a * s > (double)clo) a -= 1;
  double b = ceil(((double)chi - (double)p) / s);
  if ((double)p + b * s < (double)chi) b += 1;
  if (a < 0) a = 0;
  if (b > 255) return false;
  if (b < a) b = a;
  qlo = (uint32_t)a;
  qhi = (uint32_t)b;
  return true;
}

WHD float wsel(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

// Per-ray constants of both tests.
struct WideRay {
  float ox, oy, oz;
  float ix, iy, iz;        // inverse direction, |d| clamped to >= 2^-64 (stays finite)
  uint32_t oct;            // bit a set: direction negative on axis a
  int kx, ky, kz;          // watertight test: kz the dominant axis, (kx, ky, kz) right-handed
  float sx, sy, sz;        // shear constants
};

WHD float wide_inverse(float d) {
  const float tiny = 5.421010862427522e-20f;  // 2^-64
  float a = fabsf(d) < tiny ? copysignf(tiny, d) : d;
  return 1.0f / a;
}

// Returns false for a zero direction (no triangle can be hit).
WHD bool wide_ray(float ox, float oy, float oz, float dx, float dy, float dz, WideRay &r) {
  r.ox = ox;
  r.oy = oy;
  r.oz = oz;
  r.ix = wide_inverse(dx);
  r.iy = wide_inverse(dy);
  r.iz = wide_inverse(dz);
  r.oct = (dx < 0 ? 1u : 0u) | (dy < 0 ? 2u : 0u) | (dz < 0 ? 4u : 0u);
  float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  int kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  float dkz = wsel(dx, dy, dz, kz);
  if (dkz < 0) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  if (dkz == 0) {
    r.sx = r.sy = r.sz = 0;
    return false;
  }
  r.sx = wsel(dx, dy, dz, kx) / dkz;
  r.sy = wsel(dx, dy, dz, ky) / dkz;
  r.sz = 1.0f / dkz;
  return true;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013): vertices relative to the origin, sheared
// so the ray runs along +z, then three 2D edge functions. Products and differences are rounded
// separately (never fused), so an edge shared by two triangles gives exactly opposite values and
// no ray slips between them; a zero edge value is recomputed in FP64. side: 0 drops det < 0, 1
// drops det > 0, 2 keeps both (bvh.ts:233-237). det has the sign of the reference's Möller–Trumbore
// determinant, u weighs v1 and v weighs v2 as there.
WHD bool wide_tri_test(const float4 &a, const float4 &b, const float4 &c, const WideRay &r, int side,
                       float &t, float &u, float &v, bool &front) {
  float Ax = a.x - r.ox, Ay = a.y - r.oy, Az = a.z - r.oz;
  float Bx = b.x - r.ox, By = b.y - r.oy, Bz = b.z - r.oz;
  float Cx = c.x - r.ox, Cy = c.y - r.oy, Cz = c.z - r.oz;
  float akz = wsel(Ax, Ay, Az, r.kz), bkz = wsel(Bx, By, Bz, r.kz), ckz = wsel(Cx, Cy, Cz, r.kz);
  float ax = wsel(Ax, Ay, Az, r.kx) - r.sx * akz, ay = wsel(Ax, Ay, Az, r.ky) - r.sy * akz;
  float bx = wsel(Bx, By, Bz, r.kx) - r.sx * bkz, by = wsel(Bx, By, Bz, r.ky) - r.sy * bkz;
  float cx = wsel(Cx, Cy, Cz, r.kx) - r.sx * ckz, cy = wsel(Cx, Cy, Cz, r.ky) - r.sy * ckz;
  float U = cx * by - cy * bx;
  float V = ax * cy - ay * cx;
  float W = bx * ay - by * ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)cx * (double)by - (double)cy * (double)bx);
    V = (float)((double)ax * (double)cy - (double)ay * (double)cx);
    W = (float)((double)bx * (double)ay - (double)by * (double)ax);
  }
  bool neg = U < 0 || V < 0 || W < 0, pos = U > 0 || V > 0 || W > 0;
  if (side == 0 ? neg : (side == 1 ? pos : (neg && pos))) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float T = U * (r.sz * akz) + V * (r.sz * bkz) + W * (r.sz * ckz);
  float rcp = 1.0f / det;
  t = T * rcp;
  u = V * rcp;
  v = W * rcp;
  front = det > 0;
  return true;
}

struct WideHit {
  int tri;
  float t, u, v;
  bool front;
};

// Relative bound, in units of FP32 rounding (2^-24), that widens each slab so the rounded
// distances can only grow the interval: for one axis, fma(q, s*i, (p-o)*i) differs from the exact
// (p + q*s - o)/d by at most about 5.2 * 2^-24 * (q*s + |p-o|) * |i|.
#define MCR_WIDE_SLAB_ULPS 8.0f

// The eight child boxes of one node against (tmin, tfar): returns hit slots as bits i ^ oct,
// split into inner and leaf children.
WHD void wide_node_hits(const uint4 &n0, const uint4 &n1, const uint4 &n2, const uint4 &n3,
                        const uint4 &n4, const WideRay &r, float tmin, float tfar, uint32_t &inner_hits,
                        uint32_t &leaf_hits) {
  const float k = MCR_WIDE_SLAB_ULPS * 5.9604644775390625e-08f;  // 2^-24
  uint32_t eb = n0.w;
  float sx = __uint_as_float((eb & 0xffu) << 23), sy = __uint_as_float(((eb >> 8) & 0xffu) << 23),
        sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
  uint32_t imask = eb >> 24;
  float px = __uint_as_float(n0.x) - r.ox, py = __uint_as_float(n0.y) - r.oy, pz = __uint_as_float(n0.z) - r.oz;
  float Ax = sx * r.ix, Ay = sy * r.iy, Az = sz * r.iz;
  float Bx = px * r.ix, By = py * r.iy, Bz = pz * r.iz;
  float Ex = k * fmaf(255.0f, sx, fabsf(px)) * fabsf(r.ix);
  float Ey = k * fmaf(255.0f, sy, fabsf(py)) * fabsf(r.iy);
  float Ez = k * fmaf(255.0f, sz, fabsf(pz)) * fabsf(r.iz);
  float nBx = Bx - Ex, fBx = Bx + Ex, nBy = By - Ey, fBy = By + Ey, nBz = Bz - Ez, fBz = Bz + Ez;
  // Near planes are the low bounds on positive axes and the high bounds on negative ones.
  bool gx = r.ix < 0, gy = r.iy < 0, gz = r.iz < 0;
  uint32_t nx0 = gx ? n3.z : n2.x, nx1 = gx ? n3.w : n2.y, fx0 = gx ? n2.x : n3.z, fx1 = gx ? n2.y : n3.w;
  uint32_t ny0 = gy ? n4.x : n2.z, ny1 = gy ? n4.y : n2.w, fy0 = gy ? n2.z : n4.x, fy1 = gy ? n2.w : n4.y;
  uint32_t nz0 = gz ? n4.z : n3.x, nz1 = gz ? n4.w : n3.y, fz0 = gz ? n3.x : n4.z, fz1 = gz ? n3.y : n4.w;
  uint64_t meta = ((uint64_t)n1.w << 32) | n1.z;
  uint32_t ih = 0, lh = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int sh = (i & 3) * 8;
    uint32_t wnx = i < 4 ? nx0 : nx1, wfx = i < 4 ? fx0 : fx1;
    uint32_t wny = i < 4 ? ny0 : ny1, wfy = i < 4 ? fy0 : fy1;
    uint32_t wnz = i < 4 ? nz0 : nz1, wfz = i < 4 ? fz0 : fz1;
    float tnx = fmaf((float)((wnx >> sh) & 0xffu), Ax, nBx);
    float tfx = fmaf((float)((wfx >> sh) & 0xffu), Ax, fBx);
    float tny = fmaf((float)((wny >> sh) & 0xffu), Ay, nBy);
    float tfy = fmaf((float)((wfy >> sh) & 0xffu), Ay, fBy);
    float tnz = fmaf((float)((wnz >> sh) & 0xffu), Az, nBz);
    float tfz = fmaf((float)((wfz >> sh) & 0xffu), Az, fBz);
    float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
    float tf = fminf(fminf(tfx, tfy), fminf(tfz, tfar));
    bool inner = (imask >> i) & 1u;
    bool valid = inner || ((meta >> (8 * i)) & 0xffu) != 0;
    bool hit = valid && tn <= tf;
    uint32_t bit = 1u << ((uint32_t)i ^ r.oct);
    ih |= (hit && inner) ? bit : 0u;
    lh |= (hit && !inner) ? bit : 0u;
  }
  inner_hits = ih;
  leaf_hits = lh;
}

// First triangle of leaf slot `slot`: tri_base + the meta bytes below the slot (inner slots hold 0).
WHD uint32_t wide_leaf_offset(const uint4 &n1, uint32_t slot) {
  uint64_t meta = ((uint64_t)n1.w << 32) | n1.z;
  uint64_t below = meta & ((1ull << (8 * slot)) - 1ull);
  return (uint32_t)((below * 0x0101010101010101ull) >> 56);
}

WHD uint32_t wide_leaf_count(const uint4 &n1, uint32_t slot) {
  return ((slot < 4 ? n1.z : n1.w) >> ((slot & 3) * 8)) & 0xffu;
}

// Closest hit in (tmin, tmax) over the wide BVH. Every hit child is visited unless its box lies
// beyond the best hit when its parent is visited, as bvh.ts:255-271 culls binary children; ties in
// t keep the first triangle tested. Stack: push(base, bits) / pop(base, bits) / sp. Returns false
// and sets `bad` on a corrupt node index or a stack overflow (the ray then reports no hit).
template <class Stack>
WHD bool wide_closest(const WideScene &ws, const WideRay &r, float tmin, float tmax, bool cull,
                      WideHit &hit, Stack &st, TravCount &cnt, bool &bad) {
  hit.tri = -1;
  hit.t = tmax;
  hit.u = 0;
  hit.v = 0;
  hit.front = true;
  if (ws.node_count == 0) return false;
  st.sp = 0;
  uint32_t node = 0;
  uint32_t g_base = 0, g_imask = 0, g_hits = 0;
  for (;;) {
    if (node >= ws.node_count) {
      bad = true;
      hit.tri = -1;
      return false;
    }
    cnt.nodes++;
    const uint4 *np = ws.nodes + (size_t)node * 5;
    uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    uint32_t ih, lh;
    wide_node_hits(n0, n1, n2, n3, n4, r, tmin, hit.t, ih, lh);
    while (lh) {
      uint32_t b = (uint32_t)__builtin_ctz(lh);
      lh &= lh - 1;
      uint32_t slot = b ^ r.oct;
      uint32_t first = n1.y + wide_leaf_offset(n1, slot), count = wide_leaf_count(n1, slot);
      if (first + count > ws.triangle_count) {
        bad = true;
        hit.tri = -1;
        return false;
      }
      cnt.leaves++;
      for (uint32_t k = first; k < first + count; k++) {
        cnt.tris++;
        const float4 *tp = ws.tri + (size_t)k * 3;
        float4 a = tp[0], b4 = tp[1], c = tp[2];
        int side = cull ? (int)(__float_as_uint(b4.w) & 0xffu) : 2;
        float t, u, v;
        bool front;
        if (!wide_tri_test(a, b4, c, r, side, t, u, v, front)) continue;
        if (t > tmin && t < hit.t) {
          hit.t = t;
          hit.u = u;
          hit.v = v;
          hit.tri = (int)__float_as_uint(a.w);
          hit.front = front;
        }
      }
    }
    if (ih) {
      if (g_hits) {
        if (st.sp >= MCR_WIDE_STACK_MAX) {
          bad = true;
          hit.tri = -1;
          return false;
        }
        st.push(g_base, g_imask | (g_hits << 8));
      }
      g_base = n1.x;
      g_imask = n0.w >> 24;
      g_hits = ih;
    }
    if (!g_hits) {
      if (st.sp == 0) break;
      uint32_t bits;
      st.pop(g_base, bits);
      g_imask = bits & 0xffu;
      g_hits = bits >> 8;
    }
    uint32_t b = (uint32_t)__builtin_ctz(g_hits);
    g_hits &= g_hits - 1;
    uint32_t slot = b ^ r.oct;
    node = g_base + (uint32_t)__builtin_popcount(g_imask & ((1u << slot) - 1u));
  }
  return hit.tri >= 0;
}

// Shadow-ray query over the wide BVH (tracer.ts:529-594 over bvh.ts:289-332): every triangle hit
// in (tmin, tmax), no culling, nearer slots first; on_hit(k, leaf, t, u, v, front) returns true for an
// opaque hit, which ends the query (the light is blocked whatever else lies on the segment).
template <class Stack, class OnHit>
WHD bool wide_occluded(const WideScene &ws, const WideRay &r, float tmin, float tmax, Stack &st,
                       TravCount &cnt, bool &bad, OnHit &on_hit) {
  if (ws.node_count == 0) return false;
  st.sp = 0;
  uint32_t node = 0;
  uint32_t g_base = 0, g_imask = 0, g_hits = 0;
  for (;;) {
    if (node >= ws.node_count) {
      bad = true;
      return false;
    }
    cnt.nodes++;
    const uint4 *np = ws.nodes + (size_t)node * 5;
    uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    uint32_t ih, lh;
    wide_node_hits(n0, n1, n2, n3, n4, r, tmin, tmax, ih, lh);
    while (lh) {
      uint32_t b = (uint32_t)__builtin_ctz(lh);
      lh &= lh - 1;
      uint32_t slot = b ^ r.oct;
      uint32_t first = n1.y + wide_leaf_offset(n1, slot), count = wide_leaf_count(n1, slot);
      if (first + count > ws.triangle_count) {
        bad = true;
        return false;
      }
      cnt.leaves++;
      for (uint32_t k = first; k < first + count; k++) {
        cnt.tris++;
        const float4 *tp = ws.tri + (size_t)k * 3;
        float4 a = tp[0], b4 = tp[1], c = tp[2];
        float t, u, v;
        bool front;
        if (!wide_tri_test(a, b4, c, r, 2, t, u, v, front)) continue;
        if (t > tmin && t < tmax &&
            on_hit((int)__float_as_uint(a.w), (int)__float_as_uint(c.w), t, u, v, front))
          return true;
      }
    }
    if (ih) {
      if (g_hits) {
        if (st.sp >= MCR_WIDE_STACK_MAX) {
          bad = true;
          return false;
        }
        st.push(g_base, g_imask | (g_hits << 8));
      }
      g_base = n1.x;
      g_imask = n0.w >> 24;
      g_hits = ih;
    }
    if (!g_hits) {
      if (st.sp == 0) break;
      uint32_t bits;
      st.pop(g_base, bits);
      g_imask = bits & 0xffu;
      g_hits = bits >> 8;
    }
    uint32_t b = (uint32_t)__builtin_ctz(g_hits);
    g_hits &= g_hits - 1;
    uint32_t slot = b ^ r.oct;
    node = g_base + (uint32_t)__builtin_popcount(g_imask & ((1u << slot) - 1u));
  }
  return false;
}

// Node-group stack of the wide traversals: base and (imask | hits << 8) words, lane-interleaved
// in LDS (`lanes` threads per block), then in the thread's spill area, two words per entry.
struct WideStack {
  uint32_t *lb, *lw;
  int32_t *spill;
  int spill_stride, lanes;
  int sp;
  __device__ void push(uint32_t base, uint32_t bits) {
    if (sp < MCR_WIDE_STACK_LDS) {
      lb[sp * lanes] = base;
      lw[sp * lanes] = bits;
    } else {
      size_t o = (size_t)(2 * (sp - MCR_WIDE_STACK_LDS)) * spill_stride;
      spill[o] = (int32_t)base;
      spill[o + spill_stride] = (int32_t)bits;
    }
    sp++;
  }
  __device__ void pop(uint32_t &base, uint32_t &bits) {
    sp--;
    if (sp < MCR_WIDE_STACK_LDS) {
      base = lb[sp * lanes];
      bits = lw[sp * lanes];
    } else {
      size_t o = (size_t)(2 * (sp - MCR_WIDE_STACK_LDS)) * spill_stride;
      base = (uint32_t)spill[o];
      bits = (uint32_t)spill[o + spill_stride];
    }
  }
};

#endif  // MCR_WIDE_H
