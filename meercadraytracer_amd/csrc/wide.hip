// wide.hip — the precision form: builds a resident scene's compressed 8-wide BVH on the device from
// its binary BVH, and answers closest-hit queries over it in FP32 (wide.h).
//
// Build: level by level from the root. Each wide node starts from the one binary node (or long
// triangle range) it stands for and repeatedly opens its largest-area openable child until it has
// eight children or none can open; children that can open become the next level's nodes, the rest
// become leaf slots whose triangles are copied, contiguously per node, into the wide triangle array.
#include <string>

#include "context.h"
#include "wide.h"

namespace {

struct WideBuild {
  const float4 *tri;     // the scene's triangle records (3 per triangle, BVH order)
  const float4 *inner;   // binary inner records
  const int2 *longleaf;
  uint4 *nodes;          // 5 per wide node
  float4 *wtri;          // 3 per triangle
  WideChild *src;        // per wide node: the child it stands for
  uint32_t *ctr;         // [0] nodes allocated, [1] triangles placed, [2] error
  uint32_t node_cap, tri_count;
};

__device__ void decode_ref(const WideBuild &a, int32_t ref, WideChild &c) {
  if (ref_is_leaf(ref)) {
    uint32_t r = (uint32_t)ref, cnt = (r >> MCR_LEAF_FIRST_BITS) & 31u;
    c.inner = -1;
    if (cnt) {
      c.first = r & MCR_LEAF_FIRST_MASK;
      c.count = cnt;
    } else {
      int2 l = a.longleaf[r & MCR_LEAF_FIRST_MASK];
      c.first = (uint32_t)l.x;
      c.count = (uint32_t)l.y;
    }
    c.leaf = c.first;
  } else {
    c.inner = ref;
    c.first = 0;
    c.count = 0;
    c.leaf = 0;
  }
}

__device__ void open_child(const WideBuild &a, const WideChild &c, WideChild &l, WideChild &r) {
  if (c.inner >= 0) {
    const float4 *p = a.inner + (size_t)c.inner * 4;
    float4 x = p[0], y = p[1], z = p[2];
    int4 m = *reinterpret_cast<const int4 *>(p + 3);
    decode_ref(a, m.x, l);
    decode_ref(a, m.y, r);
    float lb[6] = {x.x, x.y, x.z, x.w, y.x, y.y}, rb[6] = {y.z, y.w, z.x, z.y, z.z, z.w};
    for (int k = 0; k < 6; k++) {
      l.b[k] = lb[k];
      r.b[k] = rb[k];
    }
  } else {
    uint32_t h = c.count / 2;
    l = c;
    r = c;
    l.count = h;
    r.first = c.first + h;
    r.count = c.count - h;
  }
}

__global__ void k_wide_root(WideBuild a, int32_t root_ref, float b0, float b1, float b2, float b3,
                            float b4, float b5) {
  WideChild c;
  decode_ref(a, root_ref, c);
  c.b[0] = b0;
  c.b[1] = b1;
  c.b[2] = b2;
  c.b[3] = b3;
  c.b[4] = b4;
  c.b[5] = b5;
  a.src[0] = c;
  a.ctr[0] = 1;
  a.ctr[1] = 0;
  a.ctr[2] = 0;
}

__global__ __launch_bounds__(256) void k_wide_level(WideBuild a, uint32_t lo, uint32_t hi) {
  uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  WideChild ch[8];
  ch[0] = a.src[i];
  int k = 1;
  while (k < 8) {
    int best = -1;
    float best_area = -1.0f;
    for (int j = 0; j < k; j++) {
      if (!wide_expandable(ch[j])) continue;
      float ar = wide_half_area(ch[j].b);
      if (ar > best_area || best < 0) {
        best_area = ar;
        best = j;
      }
    }
    if (best < 0) break;
    WideChild l, r;
    open_child(a, ch[best], l, r);
    ch[best] = l;
    ch[k++] = r;
  }
  // Node frame: the union of the children's boxes, one power-of-two step per axis.
  float lo3[3], hi3[3];
  for (int ax = 0; ax < 3; ax++) {
    lo3[ax] = ch[0].b[ax];
    hi3[ax] = ch[0].b[ax + 3];
    for (int j = 1; j < k; j++) {
      lo3[ax] = fminf(lo3[ax], ch[j].b[ax]);
      hi3[ax] = fmaxf(hi3[ax], ch[j].b[ax + 3]);
    }
  }
  uint32_t qlo[8][3], qhi[8][3];
  int ex[3];
  for (int ax = 0; ax < 3; ax++) {
    int e = wide_exponent(lo3[ax], hi3[ax]);
    for (;;) {
      bool ok = true;
      for (int j = 0; j < k && ok; j++)
        ok = wide_quantize(lo3[ax], e, ch[j].b[ax], ch[j].b[ax + 3], qlo[j][ax], qhi[j][ax]);
      if (ok) break;
      if (e >= 127) {
        a.ctr[2] = 1;
        return;
      }
      e++;
    }
    ex[ax] = e;
  }
  // Slots by octant: greedily give the (child, slot) pair whose centroid offset points furthest
  // along the slot's diagonal.
  float cx = 0.5f * (lo3[0] + hi3[0]), cy = 0.5f * (lo3[1] + hi3[1]), cz = 0.5f * (lo3[2] + hi3[2]);
  int child_at[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  uint32_t placed = 0;
  for (int it = 0; it < k; it++) {
    float best = -INFINITY;
    int bj = -1, bs = -1;
    for (int j = 0; j < k; j++) {
      if ((placed >> j) & 1u) continue;
      float dx = 0.5f * (ch[j].b[0] + ch[j].b[3]) - cx, dy = 0.5f * (ch[j].b[1] + ch[j].b[4]) - cy,
            dz = 0.5f * (ch[j].b[2] + ch[j].b[5]) - cz;
      for (int s = 0; s < 8; s++) {
        if (child_at[s] >= 0) continue;
        float sc = ((s & 1) ? dx : -dx) + ((s & 2) ? dy : -dy) + ((s & 4) ? dz : -dz);
        if (bj < 0 || sc > best) {
          best = sc;
          bj = j;
          bs = s;
        }
      }
    }
    child_at[bs] = bj;
    placed |= 1u << bj;
  }
  uint32_t ninner = 0, nleaf = 0;
  for (int j = 0; j < k; j++) {
    if (wide_expandable(ch[j])) ninner++;
    else nleaf += ch[j].count;
  }
  uint32_t child_base = 0, tri_base = 0;
  if (ninner) {
    child_base = atomicAdd(&a.ctr[0], ninner);
    if (child_base + ninner > a.node_cap) {
      a.ctr[2] = 2;
      return;
    }
  }
  if (nleaf) {
    tri_base = atomicAdd(&a.ctr[1], nleaf);
    if (tri_base + nleaf > a.tri_count) {
      a.ctr[2] = 3;
      return;
    }
  }
  uint32_t imask = 0, meta[2] = {0, 0}, q[12] = {0};
  uint32_t rank = 0, off = 0;
  for (int s = 0; s < 8; s++) {
    int j = child_at[s];
    if (j < 0) continue;
    int w = s >> 2, sh = (s & 3) * 8;
    q[0 + w] |= qlo[j][0] << sh;
    q[2 + w] |= qlo[j][1] << sh;
    q[4 + w] |= qlo[j][2] << sh;
    q[6 + w] |= qhi[j][0] << sh;
    q[8 + w] |= qhi[j][1] << sh;
    q[10 + w] |= qhi[j][2] << sh;
    if (wide_expandable(ch[j])) {
      imask |= 1u << s;
      a.src[child_base + rank++] = ch[j];
    } else {
      meta[w] |= ch[j].count << sh;
      for (uint32_t t = 0; t < ch[j].count; t++) {
        uint32_t kk = ch[j].first + t;
        const float4 *p = a.tri + (size_t)kk * 3;
        float4 A = p[0], B = p[1], C = p[2];
        float4 *o = a.wtri + (size_t)(tri_base + off + t) * 3;
        o[0] = make_float4(A.x, A.y, A.z, __uint_as_float(kk));
        o[1] = make_float4((float)((double)A.x + (double)A.w), (float)((double)A.y + (double)B.x),
                           (float)((double)A.z + (double)B.y), __uint_as_float(__float_as_uint(C.y) & 0xffu));
        o[2] = make_float4((float)((double)A.x + (double)B.z), (float)((double)A.y + (double)B.w),
                           (float)((double)A.z + (double)C.x), __uint_as_float(ch[j].leaf));
      }
      off += ch[j].count;
    }
  }
  uint32_t ebits = (uint32_t)(ex[0] + 127) | ((uint32_t)(ex[1] + 127) << 8) | ((uint32_t)(ex[2] + 127) << 16) |
                   (imask << 24);
  uint4 *n = a.nodes + (size_t)i * 5;
  n[0] = make_uint4(__float_as_uint(lo3[0]), __float_as_uint(lo3[1]), __float_as_uint(lo3[2]), ebits);
  n[1] = make_uint4(child_base, tri_base, meta[0], meta[1]);
  n[2] = make_uint4(q[0], q[1], q[2], q[3]);
  n[3] = make_uint4(q[4], q[5], q[6], q[7]);
  n[4] = make_uint4(q[8], q[9], q[10], q[11]);
}

// Closest hits of caller rays (6 floats each) over the wide BVH. Five waves per SIMD (96 VGPRs, no
// spills; four at the compiler's own choice of 101): 2,318 -> 2,543 Mrays/s on the 1M probe rays.
__global__ __launch_bounds__(256, 5) void k_wide_closest(WideScene ws, const float4 *tri64, const float *rays, uint32_t n,
                                                      float tmin, float tmax, int cull, int32_t *tri, float *t,
                                                      float *u, float *v, uint8_t *front, int32_t *spill,
                                                      uint32_t spill_stride, DevCounters *cnt, uint32_t *err,
                                                      uint32_t *fetch_counters) {
  __shared__ uint32_t lds_base[MCR_WIDE_STACK_LDS * 256], lds_bits[MCR_WIDE_STACK_LDS * 256];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  WideStack st;
  st.lb = lds_base + threadIdx.x;
  st.lw = lds_bits + threadIdx.x;
  st.spill = spill + gid;
  st.spill_stride = (int)spill_stride;
  st.lanes = 256;
  st.sp = 0;
  TravCount tc = {0, 0, 0};
  unsigned long long queries = 0;
  bool bad = false;
  WaveFetch fetch(fetch_counters, n);
  for (uint32_t b = fetch.next(); b < n; b = fetch.next()) {
    uint32_t i = b + (threadIdx.x & 63);
    if (i >= n) continue;
    const float *r = rays + (size_t)i * 6;
    WideRay wr;
    WideHit h;
    h.tri = -1;
    h.t = tmax;
    h.u = h.v = 0;
    h.front = true;
    if (wide_ray(r[0], r[1], r[2], r[3], r[4], r[5], wr))
      wide_closest(ws, wr, tmin, tmax, cull != 0, h, st, tc, bad);
    queries++;
    // The winner's t, u, v from the reference's FP64 test on the same (widened) ray.
    double rt = h.t, ru = h.u, rv = h.v;
    if (h.tri >= 0) refine_hit(tri64, h.tri, r[0], r[1], r[2], r[3], r[4], r[5], rt, ru, rv);
    if (tri) tri[i] = h.tri;
    if (t) t[i] = (float)rt;
    if (u) u[i] = (float)ru;
    if (v) v[i] = (float)rv;
    if (front) front[i] = h.front ? 1 : 0;
  }
  if (bad) *err = 1;
  flush_counts(cnt, tc, queries, 0);
}

int fail(int code, const std::string &msg) { return mcr_fail(code, msg); }

}  // namespace

void wide_free(mcr_scene *s) {
  hipFree(s->wide_nodes);
  hipFree(s->wide_tri);
  s->wide_nodes = nullptr;
  s->wide_tri = nullptr;
  s->wide = WideScene{nullptr, nullptr, 0, 0};
  s->wide_info = mcr_wide_info{};
  s->has_wide = false;
}

extern "C" {

int mcr_scene_build_wide(mcr_scene *s, mcr_wide_info *info_out) {
  if (!s) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_geometry) return fail(MCR_E_NO_SCENE, "the scene has no geometry");
  if (scene_held(s)) return scene_held_error();
  mcr_ctx *c = s->ctx;
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  HIP_TRY(hipStreamSynchronize(sm));
  wide_free(s);
  const DevScene &d = s->d;
  uint32_t n = d.triangle_count;
  mcr_wide_info info = {};
  if (d.root_ref == -1 || n == 0) {
    s->has_wide = true;
    if (info_out) *info_out = info;
    return MCR_OK;
  }
  // Every wide node but the root stands for a distinct binary inner node or a piece of a long
  // leaf split into halves of more than MCR_WIDE_LEAF_MAX / 2 triangles.
  uint64_t cap64 = (uint64_t)d.inner_count + n / 8 + 2;
  if (cap64 >= (1ull << 31)) return fail(MCR_E_OVERFLOW, "scene too large for the wide BVH");
  uint32_t cap = (uint32_t)cap64;
  WideBuild a = {};
  a.tri = d.tri;
  a.inner = d.inner;
  a.longleaf = d.longleaf;
  a.node_cap = cap;
  a.tri_count = n;
  uint32_t *ctr_host = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    hipFree(a.src);
    hipFree(a.ctr);
    hipHostFree(ctr_host);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  };
  int rc = [&]() -> int {
    HIP_TRY(dalloc(&s->wide_nodes, (size_t)cap * 5));
    HIP_TRY(dalloc(&s->wide_tri, (size_t)n * 3));
    HIP_TRY(dalloc(&a.src, cap));
    HIP_TRY(dalloc(&a.ctr, 4));
    HIP_TRY(hipHostMalloc((void **)&ctr_host, 4 * sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    a.nodes = s->wide_nodes;
    a.wtri = s->wide_tri;
    HIP_TRY(hipEventRecord(e0, sm));
    hipLaunchKernelGGL(k_wide_root, dim3(1), dim3(1), 0, sm, a, d.root_ref, d.root_box[0], d.root_box[1],
                       d.root_box[2], d.root_box[3], d.root_box[4], d.root_box[5]);
    HIP_TRY(hipGetLastError());
    uint32_t lo = 0, hi = 1, depth = 0;
    while (lo < hi) {
      hipLaunchKernelGGL(k_wide_level, dim3((hi - lo + 255) / 256), dim3(256), 0, sm, a, lo, hi);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(ctr_host, a.ctr, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipStreamSynchronize(sm));
      if (ctr_host[2]) return fail(MCR_E_INTERNAL, "wide BVH build failed (code " + std::to_string(ctr_host[2]) + ")");
      depth++;
      lo = hi;
      hi = ctr_host[0];
    }
    HIP_TRY(hipEventRecord(e1, sm));
    HIP_TRY(hipEventSynchronize(e1));
    if (ctr_host[1] != n) return fail(MCR_E_INTERNAL, "wide BVH build placed the wrong number of triangles");
    // A node group is pushed per ancestor at most.
    if (depth > MCR_WIDE_STACK_MAX)
      return fail(MCR_E_OVERFLOW, "wide BVH deeper (" + std::to_string(depth) + ") than its traversal stack");
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    info.node_count = hi;
    info.triangle_count = n;
    info.depth = depth;
    info.bytes = (uint64_t)hi * 80 + (uint64_t)n * 48;
    info.build_ms = ms;
    return MCR_OK;
  }();
  cleanup();
  if (rc) {
    wide_free(s);
    return rc;
  }
  s->wide = WideScene{s->wide_nodes, s->wide_tri, info.node_count, n};
  s->wide_info = info;
  s->has_wide = true;
  if (info_out) *info_out = info;
  return MCR_OK;
}

int mcr_scene_wide_info(const mcr_scene *s, mcr_wide_info *out) {
  if (!s || !out) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_wide) return fail(MCR_E_NO_SCENE, "the scene has no wide BVH (mcr_scene_build_wide)");
  *out = s->wide_info;
  return MCR_OK;
}

int mcr_intersect_closest_wide(mcr_ctx *c, mcr_scene *s, const float *rays, uint32_t n, float t_min,
                               float t_max, int cull, int32_t *tri_out, float *t_out, float *u_out,
                               float *v_out, uint8_t *front_out, int device_ptrs, mcr_counters *counters) {
  if (!c || !s || (!rays && n)) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_wide) return fail(MCR_E_NO_SCENE, "the scene has no wide BVH (mcr_scene_build_wide)");
  if (n == 0) return MCR_OK;
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  const float *drays = rays;
  int32_t *dtri = tri_out;
  float *dt = t_out, *du = u_out, *dv = v_out;
  uint8_t *dfront = front_out;
  void *tmp = nullptr;
  size_t bytes = device_ptrs ? 256 : (size_t)n * (24 + 4 * 4 + 1) + 1024;
  HIP_TRY(hipMalloc(&tmp, bytes));
  char *b = (char *)tmp;
  uint32_t *derr = (uint32_t *)b;
  b += 256;
  if (!device_ptrs) {
    float *r = (float *)b;
    b += (size_t)n * 24;
    dt = t_out ? (float *)b : nullptr;
    b += (size_t)n * 4;
    du = u_out ? (float *)b : nullptr;
    b += (size_t)n * 4;
    dv = v_out ? (float *)b : nullptr;
    b += (size_t)n * 4;
    dtri = tri_out ? (int32_t *)b : nullptr;
    b += (size_t)n * 4;
    dfront = front_out ? (uint8_t *)b : nullptr;
    HIP_TRY(hipMemcpyAsync(r, rays, (size_t)n * 24, hipMemcpyHostToDevice, sm));
    drays = r;
  }
  HIP_TRY(counters_clear(c->eng.counters, sm));
  HIP_TRY(hipMemsetAsync(derr, 0, 4, sm));
  if (c->eng.timing) HIP_TRY(hipEventRecord(c->eng.ev[0], sm));
  uint32_t grid = std::min<uint32_t>(TRAVERSE_GRID, (n + 255) / 256);
  HIP_TRY(hipMemsetAsync(c->eng.fetch, 0, FETCH_BYTES, sm));
  hipLaunchKernelGGL(k_wide_closest, dim3(grid), dim3(256), 0, sm, s->wide, s->d.tri, drays, n, t_min, t_max, cull, dtri,
                     dt, du, dv, dfront, c->eng.spill, TRAVERSE_THREADS, c->eng.counters, derr, c->eng.fetch);
  if (c->eng.timing) HIP_TRY(hipEventRecord(c->eng.ev[1], sm));
  HIP_TRY(hipGetLastError());
  if (!device_ptrs) {
    if (t_out) HIP_TRY(hipMemcpyAsync(t_out, dt, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
    if (u_out) HIP_TRY(hipMemcpyAsync(u_out, du, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
    if (v_out) HIP_TRY(hipMemcpyAsync(v_out, dv, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
    if (tri_out) HIP_TRY(hipMemcpyAsync(tri_out, dtri, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
    if (front_out) HIP_TRY(hipMemcpyAsync(front_out, dfront, n, hipMemcpyDeviceToHost, sm));
  }
  uint32_t err = 0;
  HIP_TRY(hipMemcpyAsync(&err, derr, 4, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipStreamSynchronize(sm));
  hipFree(tmp);
  if (err) return fail(MCR_E_INTERNAL, "wide traversal met a corrupt node or overflowed its stack");
  if (c->eng.timing) {
    float ms;
    hipEventElapsedTime(&ms, c->eng.ev[0], c->eng.ev[1]);
    c->eng.last_timing = mcr_timing{ms, 0, 0, 0, ms, 1, 0};
  }
  if (counters) {
    DevCounters hc;
    HIP_TRY(counters_read(c->eng.counters, hc));
    counters->closest_queries += hc.closest_queries;
    counters->node_visits += hc.node_visits;
    counters->leaf_visits += hc.leaf_visits;
    counters->triangle_tests += hc.triangle_tests;
  }
  return MCR_OK;
}

}  // extern "C"
