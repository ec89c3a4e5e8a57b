// bvh_build.hip — binned-SAH BVH builder on gfx950, bit-identical to src/bvh.ts:16-178.
//
// The reference builds depth first from a LIFO stack. Every decision it takes for a node depends
// only on the node's primitive set and on the order its parent's partition left them in, so the
// build here runs level by level over all nodes of a level at once, and reproduces:
//   * per-primitive boxes and centroids in FP64, stored to float (bvh.ts:22-33);
//   * node bounds with JS Math.min/max (-0 < +0): floats are mapped to order-preserving 32-bit
//     keys in which -0 sorts below +0, so integer atomic min/max give exactly Math.min/max;
//   * 16 bins on each axis with extent > 0, per-bin counts and bounds reduced in LDS and then in
//     global memory (min/max are exact, so the reduction order does not matter), and the FP64
//     right and left sweeps over them in the reference's fixed order, first strict minimum over
//     (axis, split) (bvh.ts:88-135);
//   * the SAH leaf test (bvh.ts:138-144) and the all-centroids-equal half split (bvh.ts:158-163);
//   * the in-place two-cursor partition (bvh.ts:146-157) as a closed-form scatter: with nL
//     primitives going left, mid = first + nL and F = nL + [primitive at mid goes right] front
//     positions, left-going front elements keep their slot, right-going back elements move one
//     slot left, the k-th right-going front element lands at last - rank_k (rank_0 = 0,
//     rank_k = 1 + distance from `last` of the (k-1)-th left-going back element), and the j-th
//     left-going back element lands where the j-th right-going front element was;
//   * node numbering: children of the r-th inner node in right-child-first preorder are
//     (1 + 2r, 2 + 2r) (bvh.ts:165-171), computed from subtree inner-node counts after the
//     topology is complete.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mcr.h"
#include "builder.h"
#include "jsmath.h"

namespace {

const int BINS = 16, MAX_LEAF = 4, MAX_DEPTH = 100;  // bvh.ts:3-5
const int NB = 13;                                    // words per bin: count, bmin3, bmax3, cmin3, cmax3
const int BIN_WORDS = 3 * BINS * NB;                  // per active node
const int CHUNK = 256;

enum NodeFlags { NF_LEFT = 1, NF_LEAF = 2, NF_SPLIT = 4, NF_HALF = 8, NF_SAHLEAF = 16, NF_BOUNDS = 32 };

struct BNode {
  int32_t first, n, depth, parent;
  int32_t child, flags, axis, split;
  int32_t nl, F, chunk_begin, chunk_count;
  int32_t ic, rank, id, pad;
  uint32_t bkey[6], ckey[6];
};

// What k_decide chose for the node in an active slot; k_create makes the children from it.
struct Decision {
  int32_t kind;     // 0 SAH leaf, 1 binned split, 2 half split
  int32_t axis, split, nl;
  uint32_t ckey[2][6], bkey[2][6];
};

// Counter words in device memory.
enum { C_NODES = 0, C_NEXT = 1, C_LEAVES = 2, C_BOUNDS = 3, C_ERROR = 4, C_CHUNKS = 5, C_COUNT = 8 };

__device__ __host__ inline uint32_t fkey(float f) {
  uint32_t b = __builtin_bit_cast(uint32_t, f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __host__ inline float fdecode(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __builtin_bit_cast(float, b);
}
#define KEY_POS_INF 0xff800000u  // fkey(+inf)
#define KEY_NEG_INF 0x007fffffu  // fkey(-inf)

struct Prims {  // structure of arrays, `stride` entries per field
  float *f;     // fields 0..2 bmin, 3..5 bmax, 6..8 centroid
  uint32_t *id;
  uint32_t stride;
  __device__ float get(int field, uint32_t p) const { return f[(size_t)field * stride + p]; }
  __device__ void set(int field, uint32_t p, float v) { f[(size_t)field * stride + p] = v; }
};

__device__ inline void copy_prim(const Prims &a, uint32_t p, Prims &b, uint32_t q) {
  for (int k = 0; k < 9; k++) b.set(k, q, a.get(k, p));
  b.id[q] = a.id[p];
}

// bvh.ts:22-33
__global__ void k_prim_init(const float *tris, uint32_t n, Prims P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *t = tris + (size_t)i * 9;
  double x0 = t[0], y0 = t[1], z0 = t[2];
  double x1 = x0 + (double)t[3], y1 = y0 + (double)t[4], z1 = z0 + (double)t[5];
  double x2 = x0 + (double)t[6], y2 = y0 + (double)t[7], z2 = z0 + (double)t[8];
  double ax = js_min3(x0, x1, x2), ay = js_min3(y0, y1, y2), az = js_min3(z0, z1, z2);
  double bx = js_max3(x0, x1, x2), by = js_max3(y0, y1, y2), bz = js_max3(z0, z1, z2);
  P.set(0, i, (float)ax);
  P.set(1, i, (float)ay);
  P.set(2, i, (float)az);
  P.set(3, i, (float)bx);
  P.set(4, i, (float)by);
  P.set(5, i, (float)bz);
  P.set(6, i, (float)((ax + bx) * 0.5));
  P.set(7, i, (float)((ay + by) * 0.5));
  P.set(8, i, (float)((az + bz) * 0.5));
  P.id[i] = i;
}

// Bounds and centroid bounds of the listed nodes, one block per node (bvh.ts:56-75).
__global__ __launch_bounds__(256) void k_node_bounds(BNode *nodes, const int32_t *list, Prims P) {
  __shared__ uint32_t k[12];
  BNode &nd = nodes[list[blockIdx.x]];
  if (threadIdx.x < 12) k[threadIdx.x] = ((threadIdx.x / 3) % 2 == 0) ? KEY_POS_INF : KEY_NEG_INF;
  __syncthreads();
  uint32_t first = nd.first, end = nd.first + nd.n;
  uint32_t m[12];
  for (int j = 0; j < 12; j++) m[j] = ((j / 3) % 2 == 0) ? KEY_POS_INF : KEY_NEG_INF;
  for (uint32_t p = first + threadIdx.x; p < end; p += blockDim.x) {
    for (int a = 0; a < 3; a++) {
      m[a] = min(m[a], fkey(P.get(a, p)));
      m[3 + a] = max(m[3 + a], fkey(P.get(3 + a, p)));
      uint32_t c = fkey(P.get(6 + a, p));
      m[6 + a] = min(m[6 + a], c);
      m[9 + a] = max(m[9 + a], c);
    }
  }
  for (int j = 0; j < 12; j++) {
    if ((j / 3) % 2 == 0) atomicMin(&k[j], m[j]);
    else atomicMax(&k[j], m[j]);
  }
  __syncthreads();
  if (threadIdx.x < 6) nd.bkey[threadIdx.x] = k[threadIdx.x];
  else if (threadIdx.x < 12) nd.ckey[threadIdx.x - 6] = k[threadIdx.x];
}

// Bounds of one large node (the root) over many blocks: per-block LDS reduction, then one global
// atomic per key; the node's keys start at +inf / -inf (k_root).
__global__ __launch_bounds__(256) void k_node_bounds_wide(BNode *nodes, int32_t node, Prims P,
                                                          uint32_t first, uint32_t n) {
  __shared__ uint32_t k[12];
  if (threadIdx.x < 12) k[threadIdx.x] = ((threadIdx.x / 3) % 2 == 0) ? KEY_POS_INF : KEY_NEG_INF;
  __syncthreads();
  uint32_t m[12];
  for (int j = 0; j < 12; j++) m[j] = ((j / 3) % 2 == 0) ? KEY_POS_INF : KEY_NEG_INF;
  for (uint32_t p = first + blockIdx.x * blockDim.x + threadIdx.x; p < first + n; p += gridDim.x * blockDim.x) {
    for (int a = 0; a < 3; a++) {
      m[a] = min(m[a], fkey(P.get(a, p)));
      m[3 + a] = max(m[3 + a], fkey(P.get(3 + a, p)));
      uint32_t c = fkey(P.get(6 + a, p));
      m[6 + a] = min(m[6 + a], c);
      m[9 + a] = max(m[9 + a], c);
    }
  }
  for (int j = 0; j < 12; j++) {
    if ((j / 3) % 2 == 0) atomicMin(&k[j], m[j]);
    else atomicMax(&k[j], m[j]);
  }
  __syncthreads();
  BNode &nd = nodes[node];
  if (threadIdx.x < 12) {
    uint32_t *dst = threadIdx.x < 6 ? &nd.bkey[threadIdx.x] : &nd.ckey[threadIdx.x - 6];
    if ((threadIdx.x / 3) % 2 == 0) atomicMin(dst, k[threadIdx.x]);
    else atomicMax(dst, k[threadIdx.x]);
  }
}

__global__ void k_bins_init(uint32_t *bins, uint32_t words) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= words) return;
  int f = (i % NB);
  bins[i] = f == 0 ? 0u : ((f - 1) / 3 % 2 == 0 ? KEY_POS_INF : KEY_NEG_INF);
}

struct NodeGeom {  // per-node binning parameters (bvh.ts:90-92)
  double lo[3], scale[3];
  bool valid[3];
};
__device__ inline NodeGeom node_geom(const BNode &nd) {
  NodeGeom g;
  for (int a = 0; a < 3; a++) {
    double lo = fdecode(nd.ckey[a]), hi = fdecode(nd.ckey[3 + a]);
    double extent = hi - lo;
    g.lo[a] = lo;
    g.valid[a] = extent > 0;
    g.scale[a] = g.valid[a] ? BINS / extent : 0;
  }
  return g;
}
__device__ inline int bin_of(double c, double lo, double scale) {  // bvh.ts:102,148
  return (int)js_min(BINS - 1, jm_floor((c - lo) * scale));
}

// Binning over chunks of active nodes: LDS bins, then one global atomic per non-empty word. A
// block takes BIN_GROUP consecutive chunks and merges into global memory only when the node changes
// (measured on 1.38M triangles: grouping 8 chunks gained nothing, the LDS atomics dominate).
const int BIN_GROUP = 1;

__device__ inline void bins_lds_init(uint32_t *lb) {
  for (int i = threadIdx.x; i < BIN_WORDS; i += blockDim.x) {
    int f = i % NB;
    lb[i] = f == 0 ? 0u : ((f - 1) / 3 % 2 == 0 ? KEY_POS_INF : KEY_NEG_INF);
  }
}

__device__ inline void bins_flush(const uint32_t *lb, uint32_t *gb) {
  for (int i = threadIdx.x; i < BIN_WORDS; i += blockDim.x) {
    int base = i - i % NB;
    if (lb[base] == 0) continue;
    int f = i % NB;
    if (f == 0) atomicAdd(&gb[i], lb[i]);
    else if ((f - 1) / 3 % 2 == 0) atomicMin(&gb[i], lb[i]);
    else atomicMax(&gb[i], lb[i]);
  }
}

__global__ __launch_bounds__(256) void k_bin(const int4 *chunks, uint32_t nchunks, const int32_t *active,
                                             BNode *nodes, Prims P, uint32_t *bins) {
  __shared__ uint32_t lb[BIN_WORDS];
  uint32_t c0 = blockIdx.x * BIN_GROUP, c1 = min(c0 + BIN_GROUP, nchunks);
  int cur = -1;
  for (uint32_t c = c0; c < c1; c++) {
    int4 ch = chunks[c];  // (slot, start, len, 0); uniform over the block
    if (ch.x != cur) {
      if (cur >= 0) {
        __syncthreads();
        bins_flush(lb, bins + (size_t)cur * BIN_WORDS);
      }
      __syncthreads();
      bins_lds_init(lb);
      __syncthreads();
      cur = ch.x;
    }
    const BNode &nd = nodes[active[ch.x]];
    NodeGeom g = node_geom(nd);
    if ((int)threadIdx.x < ch.z) {
      uint32_t p = (uint32_t)(ch.y + threadIdx.x);
      uint32_t kb[6], kc[3];
      for (int a = 0; a < 3; a++) {
        kb[a] = fkey(P.get(a, p));
        kb[3 + a] = fkey(P.get(3 + a, p));
        kc[a] = fkey(P.get(6 + a, p));
      }
      for (int a = 0; a < 3; a++) {
        if (!g.valid[a]) continue;
        int b = bin_of((double)P.get(6 + a, p), g.lo[a], g.scale[a]);
        uint32_t *w = lb + (a * BINS + b) * NB;
        atomicAdd(&w[0], 1u);
        for (int j = 0; j < 3; j++) {
          atomicMin(&w[1 + j], kb[j]);
          atomicMax(&w[4 + j], kb[3 + j]);
          atomicMin(&w[7 + j], kc[j]);
          atomicMax(&w[10 + j], kc[j]);
        }
      }
    }
  }
  if (cur >= 0) {
    __syncthreads();
    bins_flush(lb, bins + (size_t)cur * BIN_WORDS);
  }
}

__device__ inline double area(double x0, double y0, double z0, double x1, double y1, double z1) {
  double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;  // bvh.ts:180-183
  return dx * dy + dy * dz + dz * dx;
}

__device__ inline void union_bins(const uint32_t *B, int from, int to, uint32_t *bk, uint32_t *ck) {
  for (int j = 0; j < 3; j++) {
    bk[j] = KEY_POS_INF;
    bk[3 + j] = KEY_NEG_INF;
    ck[j] = KEY_POS_INF;
    ck[3 + j] = KEY_NEG_INF;
  }
  for (int i = from; i < to; i++) {
    const uint32_t *w = B + i * NB;
    if (w[0] == 0) continue;
    for (int j = 0; j < 3; j++) {
      bk[j] = min(bk[j], w[1 + j]);
      bk[3 + j] = max(bk[3 + j], w[4 + j]);
      ck[j] = min(ck[j], w[7 + j]);
      ck[3 + j] = max(ck[3 + j], w[10 + j]);
    }
  }
}

// Scans within 16-lane segments (one axis's bins): inclusive prefix (from bin 0) and suffix (to
// bin 15) of min, max and sum. min/max of float keys and integer sums do not depend on the order
// they are taken in, so they equal the reference's sequential sweeps.
__device__ inline uint32_t pre_min(uint32_t v, int l) {
  for (int d = 1; d < 16; d <<= 1) { uint32_t o = __shfl_up(v, d, 16); if (l >= d) v = min(v, o); }
  return v;
}
__device__ inline uint32_t pre_max(uint32_t v, int l) {
  for (int d = 1; d < 16; d <<= 1) { uint32_t o = __shfl_up(v, d, 16); if (l >= d) v = max(v, o); }
  return v;
}
__device__ inline int pre_sum(int v, int l) {
  for (int d = 1; d < 16; d <<= 1) { int o = __shfl_up(v, d, 16); if (l >= d) v += o; }
  return v;
}
__device__ inline uint32_t suf_min(uint32_t v, int l) {
  for (int d = 1; d < 16; d <<= 1) { uint32_t o = __shfl_down(v, d, 16); if (l + d < 16) v = min(v, o); }
  return v;
}
__device__ inline uint32_t suf_max(uint32_t v, int l) {
  for (int d = 1; d < 16; d <<= 1) { uint32_t o = __shfl_down(v, d, 16); if (l + d < 16) v = max(v, o); }
  return v;
}
__device__ inline int suf_sum(int v, int l) {
  for (int d = 1; d < 16; d <<= 1) { int o = __shfl_down(v, d, 16); if (l + d < 16) v += o; }
  return v;
}
__device__ inline uint32_t wave_min(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}
__device__ inline uint32_t wave_max(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}

// Split decision for every active node (bvh.ts:82-171) and children creation: one wave per node,
// lane 16 * axis + bin holds one bin. Lane i of an axis computes the reference's cost of the split
// after bin i (left = bins 0..i, right = bins i+1..15) with the same FP64 expressions; the winner
// is the first strict minimum in (axis, split) order, as the sequential loop finds it.
__global__ __launch_bounds__(256) void k_decide(const int32_t *active, uint32_t nact, BNode *nodes,
                                                const uint32_t *bins, uint32_t *counters,
                                                Decision *decisions) {
  uint32_t slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slot >= nact) return;  // uniform over the wave
  int lane = (int)(threadIdx.x & 63);
  int axis = lane >> 4, bin = lane & 15;
  int32_t ni = active[slot];
  BNode &nd = nodes[ni];
  int n = nd.n;
  const uint32_t *NBINS = bins + (size_t)slot * BIN_WORDS;
  bool in_axis = axis < 3;
  bool axis_ok = false;
  if (in_axis) {
    double lo = fdecode(nd.ckey[axis]), extent = (double)fdecode(nd.ckey[3 + axis]) - lo;
    axis_ok = extent > 0;
  }
  // This lane's bin (empty bins hold +inf / -inf keys and a zero count).
  const uint32_t *w = NBINS + (in_axis ? (axis * BINS + bin) * NB : 0);
  int cnt = in_axis ? (int)w[0] : 0;
  uint32_t kb[6], kc[6];
  for (int j = 0; j < 3; j++) {
    kb[j] = in_axis ? w[1 + j] : KEY_POS_INF;
    kb[3 + j] = in_axis ? w[4 + j] : KEY_NEG_INF;
    kc[j] = in_axis ? w[7 + j] : KEY_POS_INF;
    kc[3 + j] = in_axis ? w[10 + j] : KEY_NEG_INF;
  }
  int lc = pre_sum(cnt, bin), rc = suf_sum(cnt, bin);
  uint32_t L[6], R[6];
  for (int j = 0; j < 3; j++) {
    L[j] = pre_min(kb[j], bin);
    L[3 + j] = pre_max(kb[3 + j], bin);
    R[j] = suf_min(kb[j], bin);
    R[3 + j] = suf_max(kb[3 + j], bin);
  }
  // rightArea[bin] of the reference (bins bin..15), handed to the lane of the split before it.
  double rArea = rc > 0 ? area(fdecode(R[0]), fdecode(R[1]), fdecode(R[2]), fdecode(R[3]), fdecode(R[4]),
                               fdecode(R[5]))
                        : 0;
  double rAreaNext = __shfl_down(rArea, 1, 16);
  int rcNext = __shfl_down(rc, 1, 16);
  double cost = JM_INF;
  bool valid = in_axis && axis_ok && bin < BINS - 1 && lc != 0 && rcNext != 0;
  if (valid) {
    cost = lc * area(fdecode(L[0]), fdecode(L[1]), fdecode(L[2]), fdecode(L[3]), fdecode(L[4]), fdecode(L[5])) +
           rcNext * rAreaNext;
    if (!(cost < JM_INF)) valid = false;  // never below the initial Infinity of the reference
  }
  // First strict minimum: smallest cost, then smallest (axis, split).
  double bc = valid ? cost : JM_INF;
  int bi = valid ? axis * BINS + bin : 1 << 30;
  for (int d = 32; d > 0; d >>= 1) {
    double oc = __shfl_xor(bc, d, 64);
    int oi = __shfl_xor(bi, d, 64);
    if (oc < bc || (oc == bc && oi < bi)) {
      bc = oc;
      bi = oi;
    }
  }
  int bestAxis = bi < (1 << 30) ? bi / BINS : -1;
  int bestSplit = bi < (1 << 30) ? bi % BINS : -1;
  // Children's boxes from the prefix at the split lane and the suffix at the lane after it, and
  // their centroid bounds by masked reductions over the winning axis.
  uint32_t childB[2][6], childC[2][6];
  int nlSplit = 0;
  if (bestAxis >= 0) {
    int src = bestAxis * BINS + bestSplit;
    nlSplit = __shfl(lc, src, 64);
    for (int j = 0; j < 6; j++) {
      childB[0][j] = __shfl(L[j], src, 64);
      childB[1][j] = __shfl(R[j], src + 1, 64);
    }
    bool left = axis == bestAxis && bin <= bestSplit, right = axis == bestAxis && bin > bestSplit;
    for (int j = 0; j < 3; j++) {
      childC[0][j] = wave_min(left ? kc[j] : KEY_POS_INF);
      childC[0][3 + j] = wave_max(left ? kc[3 + j] : KEY_NEG_INF);
      childC[1][j] = wave_min(right ? kc[j] : KEY_POS_INF);
      childC[1][3 + j] = wave_max(right ? kc[3 + j] : KEY_NEG_INF);
    }
  }
  if (lane != 0) return;
  Decision &dec = decisions[slot];
  if (bestAxis >= 0) {
    double nx0 = fdecode(nd.bkey[0]), ny0 = fdecode(nd.bkey[1]), nz0 = fdecode(nd.bkey[2]);
    double nx1 = fdecode(nd.bkey[3]), ny1 = fdecode(nd.bkey[4]), nz1 = fdecode(nd.bkey[5]);
    double parentArea = area(nx0, ny0, nz0, nx1, ny1, nz1);
    double splitCost = 1.0 + (1.0 * bc) / js_max(parentArea, 1e-30);
    if (splitCost >= 1.0 * n && n <= 16) {
      dec.kind = 0;
      return;
    }
    if (nlSplit <= 0 || nlSplit >= n) atomicOr(&counters[C_ERROR], 1u);  // cannot happen: lc > 0, rc2 > 0
    dec.kind = 1;
    dec.axis = bestAxis;
    dec.split = bestSplit;
    dec.nl = nlSplit;
    for (int side = 0; side < 2; side++)
      for (int j = 0; j < 6; j++) {
        dec.bkey[side][j] = childB[side][j];
        dec.ckey[side][j] = childC[side][j];
      }
  } else {
    dec.kind = 2;
    dec.axis = 0;
    dec.split = -1;
    dec.nl = n >> 1;
  }
}

// Exclusive prefix of v over the wave and one atomic per wave on *counter; every lane calls it.
__device__ inline uint32_t wave_reserve(uint32_t *counter, uint32_t v) {
  uint32_t lane = threadIdx.x & 63, x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  uint32_t total = __shfl(x, 63, 64), base = 0;
  if (lane == 63 && total) base = atomicAdd(counter, total);
  base = __shfl(base, 63, 64);
  return base + x - v;
}

// Children of every split node (bvh.ts:165-171): node records in pairs, then the next level's
// active list, the leaf list and the list of nodes whose bounds come from their primitives.
__global__ __launch_bounds__(256) void k_create(const int32_t *active, uint32_t nact, BNode *nodes,
                                                const Decision *decisions, uint32_t *counters,
                                                int32_t *next_active, int32_t *leaf_list,
                                                int32_t *bounds_list) {
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  bool in = slot < nact;
  Decision d;
  d.kind = 0;
  if (in) d = decisions[slot];
  int32_t ni = in ? active[slot] : 0;
  bool split = in && d.kind != 0;
  uint32_t base = wave_reserve(&counters[C_NODES], split ? 2u : 0u);
  bool leaf[2] = {false, false};
  if (in) {
    BNode &nd = nodes[ni];
    if (!split) {
      nd.flags |= NF_LEAF | NF_SAHLEAF;
      nd.child = -1;
    } else {
      int n = nd.n, nl = d.nl;
      nd.child = (int32_t)base;
      nd.axis = d.axis;
      nd.split = d.split;
      nd.nl = nl;
      nd.flags |= d.kind == 2 ? NF_HALF : NF_SPLIT;
      for (int side = 0; side < 2; side++) {
        BNode &c = nodes[base + side];
        c.first = side == 0 ? nd.first : nd.first + nl;
        c.n = side == 0 ? nl : n - nl;
        c.depth = nd.depth + 1;
        c.parent = ni;
        c.child = -1;
        c.flags = (side == 0 ? NF_LEFT : 0) | (d.kind == 2 ? NF_BOUNDS : 0);
        c.axis = 0;
        c.split = -1;
        c.nl = 0;
        c.F = 0;
        c.ic = 0;
        c.rank = 0;
        c.id = 0;
        if (d.kind == 1)
          for (int j = 0; j < 6; j++) {
            c.bkey[j] = d.bkey[side][j];
            c.ckey[j] = d.ckey[side][j];
          }
        leaf[side] = c.n <= MAX_LEAF || c.depth >= MAX_DEPTH;
        if (leaf[side]) c.flags |= NF_LEAF;
      }
    }
  }
  uint32_t nb = split && d.kind == 2 ? 2u : 0u;
  uint32_t nlv = split ? (uint32_t)leaf[0] + (uint32_t)leaf[1] : 0u;
  uint32_t nnx = split ? 2u - nlv : 0u;
  uint32_t bpos = wave_reserve(&counters[C_BOUNDS], nb);
  uint32_t lpos = wave_reserve(&counters[C_LEAVES], nlv);
  uint32_t apos = wave_reserve(&counters[C_NEXT], nnx);
  if (split) {
    for (int side = 0; side < 2; side++) {
      int32_t c = (int32_t)(base + side);
      if (nb) bounds_list[bpos++] = c;
      if (leaf[side]) leaf_list[lpos++] = c;
      else next_active[apos++] = c;
    }
  }
}

__global__ void k_chunk_count(const int32_t *active, uint32_t nact, const BNode *nodes, uint32_t *cnt) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nact) return;
  cnt[s] = (uint32_t)((nodes[active[s]].n + CHUNK - 1) / CHUNK);
}

// Exclusive scan over many blocks: each block of 1024 scans its tile and leaves the tile total;
// one block scans the totals; every tile then adds its offset. total (if given) gets the sum.
__device__ inline uint32_t tile_scan_1024(uint32_t v, uint32_t &tile_total) {
  __shared__ uint32_t wsum[16];
  uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (uint32_t i = 0; i < 16; i++) {
    if (i < wid) pre += wsum[i];
    tot += wsum[i];
  }
  tile_total = tot;
  __syncthreads();
  return pre + x - v;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t *in, uint32_t n, uint32_t *out,
                                                     uint32_t *tile_sums) {
  uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0, tot;
  uint32_t e = tile_scan_1024(v, tot);
  if (i < n) out[i] = e;
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *tile_sums, uint32_t ntiles, uint32_t *total) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < ntiles; base += 1024) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < ntiles ? tile_sums[i] : 0, tot;
    uint32_t e = tile_scan_1024(v, tot);
    if (i < ntiles) tile_sums[i] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_add(uint32_t *out, uint32_t n, const uint32_t *tile_sums) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sums[i / 1024];
}

// Chunk c of the level: its node is the last one whose first chunk is <= c.
__global__ void k_chunk_emit(const int32_t *active, uint32_t nact, BNode *nodes, const uint32_t *cnt,
                             const uint32_t *off, uint32_t nchunks, int4 *chunks) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  int lo = 0, hi = (int)nact - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= c) lo = mid;
    else hi = mid - 1;
  }
  BNode &nd = nodes[active[lo]];
  uint32_t k = c - off[lo];
  if (k == 0) {
    nd.chunk_begin = (int32_t)off[lo];
    nd.chunk_count = (int32_t)cnt[lo];
  }
  int start = nd.first + (int)k * CHUNK;
  chunks[c] = make_int4(lo, start, min(CHUNK, nd.first + nd.n - start), 0);
}

// Block-wide exclusive scan of 0/1 flags (256 threads).
__device__ inline uint32_t block_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t wsum[4];
  uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t i = 0; i < wid; i++) pre += wsum[i];
  total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();
  return pre + x - v;
}

// Partition step 1: class (0 = left, bin <= split) and chunk-local exclusive prefix of lefts.
__global__ __launch_bounds__(256) void k_part_classify(const int4 *chunks, const int32_t *active,
                                                       const BNode *nodes, Prims P, uint8_t *cls,
                                                       uint16_t *lpre, uint32_t *chunk_l) {
  int4 ch = chunks[blockIdx.x];
  const BNode &nd = nodes[active[ch.x]];
  if (!(nd.flags & NF_SPLIT)) {  // uniform over the block
    if (threadIdx.x == 0) chunk_l[blockIdx.x] = 0;
    return;
  }
  NodeGeom g = node_geom(nd);
  uint32_t isL = 0;
  uint32_t p = (uint32_t)(ch.y + threadIdx.x);
  bool in = (int)threadIdx.x < ch.z;
  if (in) {
    int b = bin_of((double)P.get(6 + nd.axis, p), g.lo[nd.axis], g.scale[nd.axis]);
    isL = b <= nd.split ? 1u : 0u;
    cls[p] = isL ? 0 : 1;
  }
  uint32_t total;
  uint32_t pre = block_scan(isL, total);
  if (in) lpre[p] = (uint16_t)pre;
  if (threadIdx.x == 0) chunk_l[blockIdx.x] = total;
}

// Partition step 2: per split node, the left count from the level-wide scan of chunk lefts
// (chunk_off[c] = lefts in chunks before c, chunk_off[nchunks] = all) and the front length F.
__global__ void k_part_offsets(const int32_t *active, uint32_t nact, BNode *nodes, const uint32_t *chunk_off,
                               const uint8_t *cls, uint32_t *counters) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nact) return;
  BNode &nd = nodes[active[s]];
  if (!(nd.flags & NF_SPLIT)) return;
  uint32_t acc = chunk_off[nd.chunk_begin + nd.chunk_count] - chunk_off[nd.chunk_begin];
  if ((int)acc != nd.nl) atomicOr(&counters[C_ERROR], 2u);
  nd.F = nd.nl + (cls[nd.first + nd.nl] ? 1 : 0);
}

// Partition step 3: positions of right-going front elements and left-going back elements.
__global__ __launch_bounds__(256) void k_part_fb(const int4 *chunks, const int32_t *active,
                                                 const BNode *nodes, const uint8_t *cls,
                                                 const uint16_t *lpre, const uint32_t *chunk_off,
                                                 int32_t *fr, int32_t *bl) {
  int4 ch = chunks[blockIdx.x];
  const BNode &nd = nodes[active[ch.x]];
  if (!(nd.flags & NF_SPLIT) || (int)threadIdx.x >= ch.z) return;
  int p = ch.y + (int)threadIdx.x;
  int L = (int)(chunk_off[blockIdx.x] - chunk_off[nd.chunk_begin] + lpre[p]);  // lefts in [first, p)
  int last = nd.first + nd.n - 1;
  if (p < nd.first + nd.F) {
    if (cls[p]) fr[nd.first + ((p - nd.first) - L)] = p - nd.first;
  } else {
    if (!cls[p]) bl[nd.first + (nd.nl - 1 - L)] = last - p;
  }
}

// Partition step 4: move every primitive of the level to its slot in the other buffer; SAH
// leaves write their final order.
__global__ __launch_bounds__(256) void k_scatter(const int4 *chunks, const int32_t *active,
                                                 const BNode *nodes, Prims A, Prims B,
                                                 const uint8_t *cls, const uint16_t *lpre,
                                                 const uint32_t *chunk_off, const int32_t *fr,
                                                 const int32_t *bl, uint32_t *order) {
  int4 ch = chunks[blockIdx.x];
  const BNode &nd = nodes[active[ch.x]];
  if ((int)threadIdx.x >= ch.z) return;
  int p = ch.y + (int)threadIdx.x;
  if (nd.flags & NF_SAHLEAF) {
    order[p] = A.id[p];
    return;
  }
  if (nd.flags & NF_HALF) {
    copy_prim(A, (uint32_t)p, B, (uint32_t)p);
    return;
  }
  int L = (int)(chunk_off[blockIdx.x] - chunk_off[nd.chunk_begin] + lpre[p]);
  int first = nd.first, last = nd.first + nd.n - 1;
  int np;
  if (p < first + nd.F) {
    if (!cls[p]) {
      np = p;
    } else {
      int k = (p - first) - L;
      np = last - (k == 0 ? 0 : bl[first + k - 1] + 1);
    }
  } else {
    if (cls[p]) np = p - 1;
    else np = first + fr[first + (nd.nl - 1 - L)];
  }
  copy_prim(A, (uint32_t)p, B, (uint32_t)np);
}

__global__ void k_flush_leaves(const int32_t *list, uint32_t n, const BNode *nodes, Prims P, uint32_t *order) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BNode &nd = nodes[list[i]];
  for (int p = nd.first; p < nd.first + nd.n; p++) order[p] = P.id[p];
}

// Inner-node counts of subtrees, one level (record range) at a time, deepest first.
__global__ void k_inner_counts(BNode *nodes, uint32_t begin, uint32_t end) {
  uint32_t r = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= end) return;
  BNode &nd = nodes[r];
  nd.ic = (nd.flags & NF_LEAF) ? 0 : 1 + nodes[nd.child].ic + nodes[nd.child + 1].ic;
}

// Right-first preorder ranks and final node ids, one level at a time, root first.
__global__ void k_ranks(BNode *nodes, uint32_t begin, uint32_t end) {
  uint32_t r = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= end) return;
  const BNode &nd = nodes[r];
  if (nd.flags & NF_LEAF) return;
  BNode &L = nodes[nd.child];
  BNode &R = nodes[nd.child + 1];
  R.rank = nd.rank + 1;
  L.rank = nd.rank + 1 + R.ic;
  L.id = 1 + 2 * nd.rank;
  R.id = 2 + 2 * nd.rank;
}

__global__ void k_write_nodes(const BNode *nodes, uint32_t count, float *bounds, int32_t *out) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count) return;
  const BNode &nd = nodes[r];
  int id = nd.id;
  for (int k = 0; k < 6; k++) bounds[(size_t)id * 6 + k] = fdecode(nd.bkey[k]);
  if (nd.flags & NF_LEAF) {
    out[(size_t)id * 3] = nd.first;
    out[(size_t)id * 3 + 1] = nd.n;
    out[(size_t)id * 3 + 2] = 0;
  } else {
    out[(size_t)id * 3] = 1 + 2 * nd.rank;
    out[(size_t)id * 3 + 1] = 0;
    out[(size_t)id * 3 + 2] = nd.axis;
  }
}

__global__ void k_root(BNode *nodes, uint32_t n, int32_t *list, uint32_t *counters) {
  BNode &r = nodes[0];
  r.first = 0;
  r.n = (int32_t)n;
  r.depth = 0;
  r.parent = -1;
  r.child = -1;
  r.flags = NF_BOUNDS | (n <= (uint32_t)MAX_LEAF ? NF_LEAF : 0);
  r.axis = 0;
  r.split = -1;
  r.nl = r.F = r.ic = r.rank = r.id = 0;
  for (int k = 0; k < 6; k++) {
    r.bkey[k] = k < 3 ? KEY_POS_INF : KEY_NEG_INF;
    r.ckey[k] = k < 3 ? KEY_POS_INF : KEY_NEG_INF;
  }
  list[0] = 0;
  for (int k = 0; k < C_COUNT; k++) counters[k] = 0;
  counters[C_NODES] = 1;
}

inline uint32_t nblocks(uint64_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

// Exclusive scan of in[0..n) into out[0..n); the sum goes to *total (device) when given.
void launch_scan(hipStream_t sm, const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *total,
                 uint32_t *tile_sums) {
  uint32_t tiles = nblocks(n, 1024);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(1024), 0, sm, in, n, out, tile_sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, sm, tile_sums, tiles, total);
  if (tiles > 1) hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n)), dim3(256), 0, sm, out, n, tile_sums);
}

struct Layout {
  Prims P[2];
  BNode *nodes;
  int32_t *active[2];
  int32_t *leaf_list, *bounds_list;
  int4 *chunks;
  uint32_t *chunk_cnt, *chunk_off_node, *chunk_l, *chunk_off, *tile_sums;
  uint32_t *bins;
  Decision *decisions;
  uint8_t *cls;
  uint16_t *lpre;
  int32_t *fr, *bl;
  uint32_t *counters;
  size_t max_active, max_chunks;
};

size_t layout(char *base, uint32_t n, Layout &L) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = (off + 255) / 256 * 256;
    off = o + bytes;
    return base ? (void *)(base + o) : nullptr;
  };
  size_t N = n;
  L.max_active = N / 5 + 2;
  L.max_chunks = N / CHUNK + L.max_active + 2;
  for (int b = 0; b < 2; b++) {
    L.P[b].f = (float *)take(sizeof(float) * 9 * N);
    L.P[b].id = (uint32_t *)take(sizeof(uint32_t) * N);
    L.P[b].stride = n;
  }
  L.nodes = (BNode *)take(sizeof(BNode) * (2 * N + 2));
  L.active[0] = (int32_t *)take(sizeof(int32_t) * L.max_active);
  L.active[1] = (int32_t *)take(sizeof(int32_t) * L.max_active);
  L.leaf_list = (int32_t *)take(sizeof(int32_t) * (2 * L.max_active + 2));
  L.bounds_list = (int32_t *)take(sizeof(int32_t) * (2 * L.max_active + 2));
  L.chunks = (int4 *)take(sizeof(int4) * L.max_chunks);
  L.chunk_cnt = (uint32_t *)take(sizeof(uint32_t) * L.max_active);
  L.chunk_off_node = (uint32_t *)take(sizeof(uint32_t) * L.max_active);
  L.chunk_l = (uint32_t *)take(sizeof(uint32_t) * L.max_chunks);
  L.chunk_off = (uint32_t *)take(sizeof(uint32_t) * (L.max_chunks + 1));
  L.tile_sums = (uint32_t *)take(sizeof(uint32_t) * (L.max_chunks / 1024 + 2));
  L.bins = (uint32_t *)take(sizeof(uint32_t) * BIN_WORDS * L.max_active);
  L.decisions = (Decision *)take(sizeof(Decision) * L.max_active);
  L.cls = (uint8_t *)take(N);
  L.lpre = (uint16_t *)take(sizeof(uint16_t) * N);
  L.fr = (int32_t *)take(sizeof(int32_t) * N);
  L.bl = (int32_t *)take(sizeof(int32_t) * N);
  L.counters = (uint32_t *)take(sizeof(uint32_t) * C_COUNT);
  return off + 256;
}

}  // namespace

void builder_free(BuilderWorkspace &ws) {
  hipFree(ws.block);
  if (ws.host) hipHostFree(ws.host);
  ws = BuilderWorkspace();
}

#define BTRY(expr)                                                                    \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
      return MCR_E_DEVICE;                                                            \
    }                                                                                 \
  } while (0)

int builder_run(BuilderWorkspace &ws, hipStream_t sm, const float *tris, uint32_t n, float *bounds,
                int32_t *out_nodes, uint32_t *order, uint32_t *node_count, std::string &err) {
  Layout L;
  size_t need = layout(nullptr, n, L);
  if (ws.capacity < n || !ws.block) {
    hipFree(ws.block);
    ws.block = nullptr;
    ws.capacity = 0;
    BTRY(hipMalloc(&ws.block, need));
    ws.bytes = need;
    ws.capacity = n;
  }
  if (!ws.host) BTRY(hipHostMalloc((void **)&ws.host, sizeof(uint32_t) * C_COUNT, hipHostMallocDefault));
  // Lay out for the workspace's own capacity so buffers keep their places across calls.
  layout((char *)ws.block, ws.capacity, L);
  for (int b = 0; b < 2; b++) L.P[b].stride = ws.capacity;
  hipLaunchKernelGGL(k_prim_init, dim3(nblocks(n)), dim3(256), 0, sm, tris, n, L.P[0]);
  hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, sm, L.nodes, n, L.bounds_list, L.counters);
  hipLaunchKernelGGL(k_node_bounds_wide, dim3(std::min<uint32_t>(1024, nblocks(n))), dim3(256), 0, sm, L.nodes, 0,
                     L.P[0], 0u, n);
  std::vector<uint32_t> level_start = {0, 1};
  uint32_t nact = 0;
  int cur = 0;
  if (n <= (uint32_t)MAX_LEAF) {
    hipLaunchKernelGGL(k_flush_leaves, dim3(1), dim3(1), 0, sm, L.bounds_list, 1u, L.nodes, L.P[0], order);
  } else {
    int32_t zero = 0;
    BTRY(hipMemcpyAsync(L.active[0], &zero, sizeof(int32_t), hipMemcpyHostToDevice, sm));
    nact = 1;
  }
  int act = 0;
  while (nact > 0) {
    if (nact > L.max_active) {
      err = "builder: active node list overflow";
      return MCR_E_OVERFLOW;
    }
    int32_t *active = L.active[act], *next = L.active[act ^ 1];
    // chunks of the level's nodes
    hipLaunchKernelGGL(k_chunk_count, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.chunk_cnt);
    launch_scan(sm, L.chunk_cnt, nact, L.chunk_off_node, L.counters + C_CHUNKS, L.tile_sums);
    BTRY(hipMemsetAsync(L.counters + C_NEXT, 0, sizeof(uint32_t) * 3, sm));
    BTRY(hipMemcpyAsync(ws.host, L.counters, sizeof(uint32_t) * C_COUNT, hipMemcpyDeviceToHost, sm));
    BTRY(hipStreamSynchronize(sm));
    uint32_t nchunks = ws.host[C_CHUNKS];
    if (nchunks > L.max_chunks) {
      err = "builder: chunk list overflow";
      return MCR_E_OVERFLOW;
    }
    hipLaunchKernelGGL(k_chunk_emit, dim3(nblocks(nchunks)), dim3(256), 0, sm, active, nact, L.nodes, L.chunk_cnt,
                       L.chunk_off_node, nchunks, L.chunks);
    uint32_t words = (uint32_t)BIN_WORDS * nact;
    hipLaunchKernelGGL(k_bins_init, dim3(nblocks(words)), dim3(256), 0, sm, L.bins, words);
    hipLaunchKernelGGL(k_bin, dim3(nblocks(nchunks, BIN_GROUP)), dim3(256), 0, sm, L.chunks, nchunks, active, L.nodes,
                       L.P[cur], L.bins);
    hipLaunchKernelGGL(k_decide, dim3(nblocks(nact, 4)), dim3(256), 0, sm, active, nact, L.nodes, L.bins,
                       L.counters, L.decisions);
    hipLaunchKernelGGL(k_create, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.decisions,
                       L.counters, next, L.leaf_list, L.bounds_list);
    hipLaunchKernelGGL(k_part_classify, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.P[cur],
                       L.cls, L.lpre, L.chunk_l);
    launch_scan(sm, L.chunk_l, nchunks, L.chunk_off, L.chunk_off + nchunks, L.tile_sums);
    hipLaunchKernelGGL(k_part_offsets, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.chunk_off,
                       L.cls, L.counters);
    hipLaunchKernelGGL(k_part_fb, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.cls, L.lpre,
                       L.chunk_off, L.fr, L.bl);
    hipLaunchKernelGGL(k_scatter, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.P[cur],
                       L.P[cur ^ 1], L.cls, L.lpre, L.chunk_off, L.fr, L.bl, order);
    BTRY(hipMemcpyAsync(ws.host, L.counters, sizeof(uint32_t) * C_COUNT, hipMemcpyDeviceToHost, sm));
    BTRY(hipStreamSynchronize(sm));
    uint32_t nbounds = ws.host[C_BOUNDS], nleaves = ws.host[C_LEAVES], nnext = ws.host[C_NEXT];
    if (ws.host[C_ERROR]) {
      err = "builder: internal consistency check failed";
      return MCR_E_INVALID;
    }
    if (nbounds)
      hipLaunchKernelGGL(k_node_bounds, dim3(nbounds), dim3(256), 0, sm, L.nodes, L.bounds_list, L.P[cur ^ 1]);
    if (nleaves)
      hipLaunchKernelGGL(k_flush_leaves, dim3(nblocks(nleaves)), dim3(256), 0, sm, L.leaf_list, nleaves,
                         L.nodes, L.P[cur ^ 1], order);
    level_start.push_back(ws.host[C_NODES]);
    cur ^= 1;
    act ^= 1;
    nact = nnext;
  }
  BTRY(hipGetLastError());
  uint32_t total = level_start.back();
  // Inner-node counts bottom-up, then ranks and ids top-down.
  for (int lv = (int)level_start.size() - 2; lv >= 0; lv--) {
    uint32_t b = level_start[lv], e = level_start[lv + 1];
    if (e > b) hipLaunchKernelGGL(k_inner_counts, dim3(nblocks(e - b)), dim3(256), 0, sm, L.nodes, b, e);
  }
  for (size_t lv = 0; lv + 1 < level_start.size(); lv++) {
    uint32_t b = level_start[lv], e = level_start[lv + 1];
    if (e > b) hipLaunchKernelGGL(k_ranks, dim3(nblocks(e - b)), dim3(256), 0, sm, L.nodes, b, e);
  }
  hipLaunchKernelGGL(k_write_nodes, dim3(nblocks(total)), dim3(256), 0, sm, L.nodes, total, bounds, out_nodes);
  BTRY(hipGetLastError());
  BTRY(hipStreamSynchronize(sm));
  *node_count = total;
  return MCR_OK;
}
