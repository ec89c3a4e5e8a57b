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
// Inner nodes of at most SMALL primitives are built whole, subtree and all, by one wave (k_small).
const int SMALL = 64, SMALL_RECS = 2 * SMALL;
const int NB = 13;                                    // words per bin: count, bmin3, bmax3, cmin3, cmax3
const int BIN_WORDS = 3 * BINS * NB;                  // per active node
const int CHUNK = 256;
static_assert(BINS == 16, "k_bin selects a bin by its four bits");

// NF_SUBROOT: the root of a subtree k_small built (its inner count is the subtree's; the levels'
// passes leave it and its children to k_small and k_sub_ranks). NF_UNUSED: a record set aside
// for a small subtree that it did not need.
enum NodeFlags { NF_LEFT = 1, NF_LEAF = 2, NF_SPLIT = 4, NF_HALF = 8, NF_SAHLEAF = 16, NF_BOUNDS = 32, NF_SUBROOT = 64,
                 NF_UNUSED = 128 };

struct BNode {
  int32_t first, n, depth, parent;
  int32_t child, flags, axis, split;
  int32_t nl, F, chunk_begin, chunk_count;
  int32_t ic, rank, id;
  int32_t lrank, sub;  // nodes of a subtree built by k_small: rank below the subtree root, the root
  int32_t pad;
  uint32_t bkey[6], ckey[6];
};

// What k_decide chose for the node in an active slot; k_create makes the children from it.
struct Decision {
  int32_t kind;     // 0 SAH leaf, 1 binned split, 2 half split
  int32_t axis, split, nl;
  uint32_t ckey[2][6], bkey[2][6];
};

// Counter words in device memory.
// C_SMALL_ALLOC: records set aside for the small subtrees; C_SMALL_USED: those they used.
// C_NEXT_CHUNKS: chunks of the nodes the next level bins (so the host learns the next level's
// sizes from the one read per level).
enum { C_NODES = 0, C_NEXT = 1, C_LEAVES = 2, C_BOUNDS = 3, C_NEXT_CHUNKS = 4, C_ERROR = 5, C_CHUNKS = 6,
       C_SMALL = 7, C_SMALL_ALLOC = 8, C_SMALL_USED = 9, C_COUNT = 12 };

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

struct NodeGeom {  // per-node binning parameters (bvh.ts:90-92)
  double lo[3], scale[3];
  bool valid[3];
};
__device__ inline NodeGeom node_geom(const uint32_t *ckey) {
  NodeGeom g;
  for (int a = 0; a < 3; a++) {
    double lo = fdecode(ckey[a]), hi = fdecode(ckey[3 + a]);
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

// Binning over chunks of active nodes, one block per chunk (all its primitives in one node): every
// thread stages its primitive's box and centroid keys in LDS; then in each wave lane 16 * axis +
// bin takes the wave's primitives in that bin (a mask built from five ballots per axis: the bin's
// four bits and validity) and keeps its bin's count, box and centroid keys in registers, reading
// only its own primitives' keys (a sweep of all 64 with every lane reading every key took 3%
// longer over the whole build); the four waves'
// bins are merged in LDS and flushed with one global atomic per word of a non-empty bin. A node
// has `copies` copies of its bins and chunk k of the node adds to copy k % copies, so the few
// large nodes of the top levels spread their blocks' atomics over several addresses; k_decide
// merges the copies. min/max and counts do not depend on the order they are taken in.
__global__ __launch_bounds__(256) void k_bin(const int4 *chunks, uint32_t nchunks, const int32_t *active,
                                             BNode *nodes, Prims P, uint32_t *bins, uint32_t copies) {
  // Per primitive 12 words, read as three 16-byte words: box min, box max and centroid keys, and
  // its bins (byte a: axis a; 0xff: the axis has no extent or the primitive is past the chunk).
  __shared__ uint4 key[CHUNK][3];
  __shared__ uint32_t part[4][3 * BINS][NB];
  const uint32_t c = blockIdx.x;
  const int4 ch = chunks[c];  // (slot, start, len, node); uniform over the block
  const int t = (int)threadIdx.x;
  float pv[9] = {};  // the primitive's loads go out before the node's, which they do not need
  if (t < ch.z)
    for (int k = 0; k < 9; k++) pv[k] = P.get(k, (uint32_t)(ch.y + t));
  const BNode &nd = nodes[ch.w];
  const NodeGeom g = node_geom(nd.ckey);
  {
    uint32_t v[12] = {};
    uint32_t bins3 = 0xffffffu;
    if (t < ch.z) {
      bins3 = 0;
      for (int a = 0; a < 3; a++) {
        float cf = pv[6 + a];
        v[a] = fkey(pv[a]);
        v[3 + a] = fkey(pv[3 + a]);
        v[6 + a] = fkey(cf);
        uint32_t b = g.valid[a] ? (uint32_t)bin_of((double)cf, g.lo[a], g.scale[a]) : 0xffu;
        bins3 |= b << (8 * a);
      }
    }
    v[9] = bins3;
    key[t][0] = make_uint4(v[0], v[1], v[2], v[3]);
    key[t][1] = make_uint4(v[4], v[5], v[6], v[7]);
    key[t][2] = make_uint4(v[8], v[9], v[10], v[11]);
  }
  __syncthreads();
  const int w = t >> 6, lane = t & 63, axis = lane >> 4, bin = lane & 15;
  // Lane 16 * axis + bin: the wave's primitives whose bin on `axis` is `bin` (a primitive past the
  // chunk has no valid bin).
  uint64_t mine = 0;
  {
    const uint32_t bins3 = key[t][2].y;
    for (int a = 0; a < 3; a++) {
      const uint32_t pb = (bins3 >> (8 * a)) & 0xffu;
      const uint64_t V = __ballot(pb < (uint32_t)BINS), B0 = __ballot(pb & 1u), B1 = __ballot(pb & 2u),
                     B2 = __ballot(pb & 4u), B3 = __ballot(pb & 8u);
      if (axis == a)
        mine = V & ((bin & 1) ? B0 : ~B0) & ((bin & 2) ? B1 : ~B1) & ((bin & 4) ? B2 : ~B2) & ((bin & 8) ? B3 : ~B3);
    }
  }
  if (axis < 3) {
    int cnt = 0;
    uint32_t kb[6], kc[6];
    for (int j = 0; j < 3; j++) {
      kb[j] = kc[j] = KEY_POS_INF;
      kb[3 + j] = kc[3 + j] = KEY_NEG_INF;
    }
    while (mine) {
      const int k = 64 * w + (int)__builtin_ctzll(mine);
      mine &= mine - 1;
      uint4 a = key[k][0], b = key[k][1], cc = key[k][2];
      cnt++;
      kb[0] = min(kb[0], a.x);
      kb[1] = min(kb[1], a.y);
      kb[2] = min(kb[2], a.z);
      kb[3] = max(kb[3], a.w);
      kb[4] = max(kb[4], b.x);
      kb[5] = max(kb[5], b.y);
      kc[0] = min(kc[0], b.z);
      kc[1] = min(kc[1], b.w);
      kc[2] = min(kc[2], cc.x);
      kc[3] = max(kc[3], b.z);
      kc[4] = max(kc[4], b.w);
      kc[5] = max(kc[5], cc.x);
    }
    uint32_t *o = part[w][lane];
    o[0] = (uint32_t)cnt;
    for (int j = 0; j < 3; j++) {
      o[1 + j] = kb[j];
      o[4 + j] = kb[3 + j];
      o[7 + j] = kc[j];
      o[10 + j] = kc[3 + j];
    }
  }
  __syncthreads();
  uint32_t copy = (c - (uint32_t)nd.chunk_begin) % copies;
  uint32_t *gb = bins + ((size_t)ch.x * copies + copy) * BIN_WORDS;
  for (int i = t; i < BIN_WORDS; i += 256) {
    int b = i / NB, f = i % NB;
    if (part[0][b][0] + part[1][b][0] + part[2][b][0] + part[3][b][0] == 0) continue;
    uint32_t v0 = part[0][b][f], v1 = part[1][b][f], v2 = part[2][b][f], v3 = part[3][b][f];
    if (f == 0) atomicAdd(&gb[i], v0 + v1 + v2 + v3);
    else if ((f - 1) / 3 % 2 == 0) atomicMin(&gb[i], min(min(v0, v1), min(v2, v3)));
    else atomicMax(&gb[i], max(max(v0, v1), max(v2, v3)));
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

// Scans within the 16-lane rows of a wave (one axis's bins per row): inclusive prefix (from bin 0)
// and suffix (to bin 15) of min, max and sum, with DPP row shifts (a lane whose source lies outside
// its row keeps the identity). min/max of float keys and integer sums do not depend on the order
// they are taken in, so they equal the reference's sequential sweeps.
#define DPP_ROW_SHL(n) (0x100 | (n))
#define DPP_ROW_SHR(n) (0x110 | (n))
#define DPP_ROW_ROR(n) (0x120 | (n))
template <int CTRL>
__device__ inline uint32_t dpp(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline double dpp_f64(double old, double v) {
  uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
  uint32_t lo = dpp<CTRL>((uint32_t)o, (uint32_t)x), hi = dpp<CTRL>((uint32_t)(o >> 32), (uint32_t)(x >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline uint32_t pre_min(uint32_t v) {
  v = min(v, dpp<DPP_ROW_SHR(1)>(~0u, v));
  v = min(v, dpp<DPP_ROW_SHR(2)>(~0u, v));
  v = min(v, dpp<DPP_ROW_SHR(4)>(~0u, v));
  return min(v, dpp<DPP_ROW_SHR(8)>(~0u, v));
}
__device__ inline uint32_t pre_max(uint32_t v) {
  v = max(v, dpp<DPP_ROW_SHR(1)>(0u, v));
  v = max(v, dpp<DPP_ROW_SHR(2)>(0u, v));
  v = max(v, dpp<DPP_ROW_SHR(4)>(0u, v));
  return max(v, dpp<DPP_ROW_SHR(8)>(0u, v));
}
__device__ inline int pre_sum(int v) {
  v += (int)dpp<DPP_ROW_SHR(1)>(0u, (uint32_t)v);
  v += (int)dpp<DPP_ROW_SHR(2)>(0u, (uint32_t)v);
  v += (int)dpp<DPP_ROW_SHR(4)>(0u, (uint32_t)v);
  return v + (int)dpp<DPP_ROW_SHR(8)>(0u, (uint32_t)v);
}
__device__ inline uint32_t suf_min(uint32_t v) {
  v = min(v, dpp<DPP_ROW_SHL(1)>(~0u, v));
  v = min(v, dpp<DPP_ROW_SHL(2)>(~0u, v));
  v = min(v, dpp<DPP_ROW_SHL(4)>(~0u, v));
  return min(v, dpp<DPP_ROW_SHL(8)>(~0u, v));
}
__device__ inline uint32_t suf_max(uint32_t v) {
  v = max(v, dpp<DPP_ROW_SHL(1)>(0u, v));
  v = max(v, dpp<DPP_ROW_SHL(2)>(0u, v));
  v = max(v, dpp<DPP_ROW_SHL(4)>(0u, v));
  return max(v, dpp<DPP_ROW_SHL(8)>(0u, v));
}
__device__ inline int suf_sum(int v) {
  v += (int)dpp<DPP_ROW_SHL(1)>(0u, (uint32_t)v);
  v += (int)dpp<DPP_ROW_SHL(2)>(0u, (uint32_t)v);
  v += (int)dpp<DPP_ROW_SHL(4)>(0u, (uint32_t)v);
  return v + (int)dpp<DPP_ROW_SHL(8)>(0u, (uint32_t)v);
}
__device__ inline uint32_t lane_u32(uint32_t v, int lane) {  // lane: wave-uniform
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ inline double lane_f64(double v, int lane) {
  uint64_t x = __builtin_bit_cast(uint64_t, v);
  uint64_t lo = lane_u32((uint32_t)x, lane), hi = lane_u32((uint32_t)(x >> 32), lane);
  return __builtin_bit_cast(double, (hi << 32) | lo);
}
__device__ inline uint32_t wave_min(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}
__device__ inline uint32_t wave_max(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}

// The split decision of one node of n primitives with box keys nbkey and centroid keys nckey
// (bvh.ts:82-171), from its bins: lane 16 * axis + bin holds one bin's count cnt, box keys kb and
// centroid keys kc (lanes 48..63 and empty bins hold a zero count and +inf / -inf keys). Lane i of
// an axis computes the reference's cost of the split after bin i (left = bins 0..i, right = bins
// i+1..15) with the same FP64 expressions; the winner is the first strict minimum in (axis, split)
// order, as the sequential loop finds it. Every lane of the wave calls it and gets the same
// Decision; bad is set if the split would leave a side empty (it cannot).
__device__ inline void decide_wave(int n, const uint32_t *nbkey, const uint32_t *nckey, int lane, int cnt,
                                   const uint32_t (&kb)[6], const uint32_t (&kc)[6], Decision &dec, bool &bad) {
  int axis = lane >> 4, bin = lane & 15;
  bool in_axis = axis < 3;
  bool axis_ok = false;
  if (in_axis) {
    double lo = fdecode(nckey[axis]), extent = (double)fdecode(nckey[3 + axis]) - lo;
    axis_ok = extent > 0;
  }
  int lc = pre_sum(cnt), rc = suf_sum(cnt);
  uint32_t L[6], R[6];
  for (int j = 0; j < 3; j++) {
    L[j] = pre_min(kb[j]);
    L[3 + j] = pre_max(kb[3 + j]);
    R[j] = suf_min(kb[j]);
    R[3 + j] = suf_max(kb[3 + j]);
  }
  // rightArea[bin] of the reference (bins bin..15), handed to the lane of the split before it
  // (lane 15 of a row keeps its own, unused: no split follows bin 15).
  double rArea = rc > 0 ? area(fdecode(R[0]), fdecode(R[1]), fdecode(R[2]), fdecode(R[3]), fdecode(R[4]),
                               fdecode(R[5]))
                        : 0;
  double rAreaNext = dpp_f64<DPP_ROW_SHL(1)>(rArea, rArea);
  int rcNext = (int)dpp<DPP_ROW_SHL(1)>((uint32_t)rc, (uint32_t)rc);
  double cost = JM_INF;
  bool valid = in_axis && axis_ok && bin < BINS - 1 && lc != 0 && rcNext != 0;
  if (valid) {
    cost = lc * area(fdecode(L[0]), fdecode(L[1]), fdecode(L[2]), fdecode(L[3]), fdecode(L[4]), fdecode(L[5])) +
           rcNext * rAreaNext;
    if (!(cost < JM_INF)) valid = false;  // never below the initial Infinity of the reference
  }
  // First strict minimum: smallest cost, then smallest (axis, split); the order is total, so the
  // minimum of each row (rotations within the row) and then of the rows is the same lane.
  double bc = valid ? cost : JM_INF;
  int bi = valid ? axis * BINS + bin : 1 << 30;
#define MCR_ARGMIN_STEP(n)                                                   \
  {                                                                           \
    double oc = dpp_f64<DPP_ROW_ROR(n)>(bc, bc);                              \
    int oi = (int)dpp<DPP_ROW_ROR(n)>((uint32_t)bi, (uint32_t)bi);            \
    if (oc < bc || (oc == bc && oi < bi)) {                                   \
      bc = oc;                                                                \
      bi = oi;                                                                \
    }                                                                         \
  }
  MCR_ARGMIN_STEP(1)
  MCR_ARGMIN_STEP(2)
  MCR_ARGMIN_STEP(4)
  MCR_ARGMIN_STEP(8)
#undef MCR_ARGMIN_STEP
  {
    double best = lane_f64(bc, 0);
    int besti = (int)lane_u32((uint32_t)bi, 0);
    for (int r = 1; r < 3; r++) {  // row 3 (lanes 48..63) holds no bins
      double oc = lane_f64(bc, 16 * r);
      int oi = (int)lane_u32((uint32_t)bi, 16 * r);
      if (oc < best || (oc == best && oi < besti)) {
        best = oc;
        besti = oi;
      }
    }
    bc = best;
    bi = __builtin_amdgcn_readfirstlane(besti);
  }
  int bestAxis = bi < (1 << 30) ? bi / BINS : -1;
  int bestSplit = bi < (1 << 30) ? bi % BINS : -1;
  // Children's boxes and centroid bounds from the prefix at the split lane and the suffix at the
  // lane after it.
  uint32_t childB[2][6], childC[2][6];
  int nlSplit = 0;
  if (bestAxis >= 0) {
    int src = bi;
    uint32_t CL[6], CR[6];
    for (int j = 0; j < 3; j++) {
      CL[j] = pre_min(kc[j]);
      CL[3 + j] = pre_max(kc[3 + j]);
      CR[j] = suf_min(kc[j]);
      CR[3 + j] = suf_max(kc[3 + j]);
    }
    nlSplit = (int)lane_u32((uint32_t)lc, src);
    for (int j = 0; j < 6; j++) {
      childB[0][j] = lane_u32(L[j], src);
      childB[1][j] = lane_u32(R[j], src + 1);
      childC[0][j] = lane_u32(CL[j], src);
      childC[1][j] = lane_u32(CR[j], src + 1);
    }
  }
  if (bestAxis >= 0) {
    double nx0 = fdecode(nbkey[0]), ny0 = fdecode(nbkey[1]), nz0 = fdecode(nbkey[2]);
    double nx1 = fdecode(nbkey[3]), ny1 = fdecode(nbkey[4]), nz1 = fdecode(nbkey[5]);
    double parentArea = area(nx0, ny0, nz0, nx1, ny1, nz1);
    double splitCost = 1.0 + (1.0 * bc) / js_max(parentArea, 1e-30);
    if (splitCost >= 1.0 * n && n <= 16) {
      dec.kind = 0;
      return;
    }
    if (nlSplit <= 0 || nlSplit >= n) bad = true;  // cannot happen: lc > 0, rc2 > 0
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

// Split decision for every active node: one wave per node over the node's global bins.
__global__ __launch_bounds__(256) void k_decide(const int32_t *active, uint32_t nact, BNode *nodes,
                                                const uint32_t *bins, uint32_t copies, uint32_t *counters,
                                                Decision *decisions) {
  uint32_t slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slot >= nact) return;  // uniform over the wave
  int lane = (int)(threadIdx.x & 63);
  int axis = lane >> 4, bin = lane & 15;
  const BNode &nd = nodes[active[slot]];
  bool in_axis = axis < 3;
  // This lane's bin, merged over the node's copies (empty bins hold +inf / -inf keys and a zero
  // count).
  int cnt = 0;
  uint32_t kb[6], kc[6];
  for (int j = 0; j < 3; j++) {
    kb[j] = kc[j] = KEY_POS_INF;
    kb[3 + j] = kc[3 + j] = KEY_NEG_INF;
  }
  if (in_axis)
    for (uint32_t k = 0; k < copies; k++) {
      const uint32_t *w = bins + ((size_t)slot * copies + k) * BIN_WORDS + (axis * BINS + bin) * NB;
      cnt += (int)w[0];
      for (int j = 0; j < 3; j++) {
        kb[j] = min(kb[j], w[1 + j]);
        kb[3 + j] = max(kb[3 + j], w[4 + j]);
        kc[j] = min(kc[j], w[7 + j]);
        kc[3 + j] = max(kc[3 + j], w[10 + j]);
      }
    }
  Decision dec;
  bool bad = false;
  decide_wave(nd.n, nd.bkey, nd.ckey, lane, cnt, kb, kc, dec, bad);
  if (lane != 0) return;
  if (bad) atomicOr(&counters[C_ERROR], 1u);
  decisions[slot] = dec;
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
                                                int32_t *bounds_list, int32_t *small_list) {
  uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  bool in = slot < nact;
  Decision d;
  d.kind = 0;
  if (in) d = decisions[slot];
  int32_t ni = in ? active[slot] : 0;
  bool split = in && d.kind != 0;
  uint32_t base = wave_reserve(&counters[C_NODES], split ? 2u : 0u);
  bool leaf[2] = {false, false}, small[2] = {false, false};
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
        small[side] = !leaf[side] && c.n <= SMALL;
        c.sub = -1;
        c.lrank = 0;
        if (leaf[side]) c.flags |= NF_LEAF;
      }
    }
  }
  uint32_t nb = split && d.kind == 2 ? 2u : 0u;
  uint32_t nlv = split ? (uint32_t)leaf[0] + (uint32_t)leaf[1] : 0u;
  uint32_t nsm = split ? (uint32_t)small[0] + (uint32_t)small[1] : 0u;
  uint32_t nnx = split ? 2u - nlv - nsm : 0u;
  uint32_t nch = 0;
  if (split)
    for (int side = 0; side < 2; side++)
      if (!leaf[side] && !small[side]) nch += (uint32_t)((nodes[base + side].n + CHUNK - 1) / CHUNK);
  wave_reserve(&counters[C_NEXT_CHUNKS], nch);
  uint32_t bpos = wave_reserve(&counters[C_BOUNDS], nb);
  uint32_t lpos = wave_reserve(&counters[C_LEAVES], nlv);
  uint32_t spos = wave_reserve(&counters[C_SMALL], nsm);
  uint32_t apos = wave_reserve(&counters[C_NEXT], nnx);
  if (split) {
    for (int side = 0; side < 2; side++) {
      int32_t c = (int32_t)(base + side);
      if (nb) bounds_list[bpos++] = c;
      if (leaf[side]) leaf_list[lpos++] = c;
      else if (small[side]) small_list[spos++] = c;
      else next_active[apos++] = c;
    }
  }
}

// ---- small subtrees ------------------------------------------------------------------------------
// One wave builds the whole subtree of a node of at most SMALL primitives, with the decisions the
// level-parallel kernels take: the 16 x 3 bins filled in LDS by the node's primitives (one lane
// each, LDS atomics), read back with lane 16 * axis + bin holding one bin, decide_wave, and the
// closed-form two-cursor partition on the LDS copy. A child's bounds and centroid bounds come from
// its parent's bins (exact: min and max), a half split's children take
// wave reductions. The node's primitives sit in buffer depth % 2 (the level that created it
// scattered them there and no later level touches them).
//
// Nodes come off an LDS stack that takes the left child, then the right, so the right subtree is
// finished first: the wave meets the inner nodes in right-first preorder, which makes a node's
// rank below the root its place in that order (lrank), and the root's inner count the number
// met. k_sub_ranks adds the root's rank once the levels above have theirs. The final primitive
// order of the node's range is the LDS order at the end.
//
// Records: a subtree of n primitives has at most 2n - 2 nodes below its root; the subtrees' counts
// are prefix-summed beforehand (k_small_counts, then a scan), so subtree i takes its records from
// `base` + off[i] on without a shared counter, and marks the ones it did not need NF_UNUSED.
const int SMALL_STACK = SMALL + 8;  // stack entries: at most the subtree depth + 2

// a[i] for a per-lane or runtime i, by selects (an indexed register array would live in scratch).
template <class T>
__device__ inline T pick3(const T (&a)[3], int i) { return i == 0 ? a[0] : i == 1 ? a[1] : a[2]; }

__global__ void k_small_counts(const int32_t *list, uint32_t nsmall, const BNode *nodes, uint32_t *cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsmall) cnt[i] = 2u * (uint32_t)nodes[list[i]].n - 2u;
}

// Four waves per SIMD (128 VGPRs, 7 spilled; the LDS of 15 blocks per CU allows no more): 4.35 ->
// 4.14 ms per 1.38M-triangle build against the compiler's own 136 VGPRs and three waves.
__global__ __launch_bounds__(64, 4) void k_small(const int32_t *list, const uint32_t *off, uint32_t base_rec,
                                              BNode *nodes, Prims P0, Prims P1, uint32_t *order,
                                              uint32_t *counters) {
  __shared__ float pf[9][SMALL];
  __shared__ uint32_t pid[SMALL];
  __shared__ int sfirst[SMALL_STACK], sn[SMALL_STACK], sdepth[SMALL_STACK], sglob[SMALL_STACK];
  __shared__ uint32_t skey[SMALL_STACK][12];  // bkey[6], ckey[6]; skey[.][0] == 0: to be reduced
  __shared__ int fr[SMALL], bl[SMALL];
  __shared__ uint32_t sbin[3 * BINS * NB];  // the current node's bins
  const int lane = (int)threadIdx.x;
  const int32_t root = list[blockIdx.x];
  const int first = nodes[root].first, n = nodes[root].n, depth0 = nodes[root].depth;
  const Prims &P = (depth0 & 1) ? P1 : P0;
  if (lane < n) {
    for (int k = 0; k < 9; k++) pf[k][lane] = P.get(k, (uint32_t)(first + lane));
    pid[lane] = P.id[first + lane];
  }
  if (lane < 12) skey[0][lane] = lane < 6 ? nodes[root].bkey[lane] : nodes[root].ckey[lane - 6];
  if (lane == 0) {
    sfirst[0] = 0;
    sn[0] = n;
    sdepth[0] = depth0;
    sglob[0] = root;
  }
  __syncthreads();
  int sp = 1, inner = 0;  // wave-uniform
  const uint32_t rec0 = base_rec + off[blockIdx.x];
  bool bad = false;
  while (sp > 0) {
    sp--;
    const int f0 = sfirst[sp], nn = sn[sp], dep = sdepth[sp], g = sglob[sp];
    const bool in = lane < nn;
    const int p = f0 + (in ? lane : 0);
    uint32_t nbkey[6], nckey[6];
    if (skey[sp][0] != 0) {
      for (int j = 0; j < 6; j++) {
        nbkey[j] = skey[sp][j];
        nckey[j] = skey[sp][6 + j];
      }
    } else {  // children of a half split
      for (int a = 0; a < 3; a++) {
        nbkey[a] = wave_min(in ? fkey(pf[a][p]) : KEY_POS_INF);
        nbkey[3 + a] = wave_max(in ? fkey(pf[3 + a][p]) : KEY_NEG_INF);
        uint32_t c = fkey(pf[6 + a][p]);
        nckey[a] = wave_min(in ? c : KEY_POS_INF);
        nckey[3 + a] = wave_max(in ? c : KEY_NEG_INF);
      }
    }
    bool leaf = nn <= MAX_LEAF || dep >= MAX_DEPTH;
    Decision dec;
    dec.kind = 0;
    NodeGeom geo = node_geom(nckey);
    if (!leaf) {
      const int axis = lane >> 4, bin = lane & 15;
      // The node's bins in LDS: cleared, then every primitive adds its count and keys to its bin
      // on each axis with LDS atomics (min, max and sums do not depend on their order).
      for (int i = lane; i < 3 * BINS * NB; i += 64) {
        int f = i % NB;
        sbin[i] = f == 0 ? 0u : ((f - 1) / 3 % 2 == 0 ? KEY_POS_INF : KEY_NEG_INF);
      }
      __syncthreads();
      if (in) {
        uint32_t v[9];
        for (int j = 0; j < 9; j++) v[j] = fkey(pf[j][p]);
        for (int a = 0; a < 3; a++) {
          if (!geo.valid[a]) continue;
          uint32_t *wb = sbin + (a * BINS + bin_of((double)pf[6 + a][p], geo.lo[a], geo.scale[a])) * NB;
          atomicAdd(&wb[0], 1u);
          for (int j = 0; j < 3; j++) {
            atomicMin(&wb[1 + j], v[j]);
            atomicMax(&wb[4 + j], v[3 + j]);
            atomicMin(&wb[7 + j], v[6 + j]);
            atomicMax(&wb[10 + j], v[6 + j]);
          }
        }
      }
      __syncthreads();
      int cnt = 0;
      uint32_t kb[6], kc[6];
      for (int j = 0; j < 3; j++) {
        kb[j] = kc[j] = KEY_POS_INF;
        kb[3 + j] = kc[3 + j] = KEY_NEG_INF;
      }
      if (axis < 3 && pick3(geo.valid, axis)) {
        const uint32_t *wb = sbin + (axis * BINS + bin) * NB;
        cnt = (int)wb[0];
        for (int j = 0; j < 3; j++) {
          kb[j] = wb[1 + j];
          kb[3 + j] = wb[4 + j];
          kc[j] = wb[7 + j];
          kc[3 + j] = wb[10 + j];
        }
      }
      decide_wave(nn, nbkey, nckey, lane, cnt, kb, kc, dec, bad);
    }
    BNode &nd = nodes[g];
    if (leaf || dec.kind == 0) {
      if (lane == 0) {
        nd.flags = (nd.flags & (NF_LEFT | NF_SUBROOT)) | NF_LEAF | (leaf ? 0 : NF_SAHLEAF);
        nd.child = -1;
        for (int j = 0; j < 6; j++) {
          nd.bkey[j] = nbkey[j];
          nd.ckey[j] = nckey[j];
        }
      }
      continue;  // the next pop reads entries below this one, written before the last barrier
    }
    int nl = dec.nl;
    if (dec.kind == 1) {  // the two-cursor partition (bvh.ts:146-157) in closed form
      bool isL = in && bin_of((double)pf[6 + dec.axis][p], pick3(geo.lo, dec.axis), pick3(geo.scale, dec.axis)) <=
                           dec.split;
      uint64_t lm = __ballot(isL);
      int cl = (int)__popcll(lm);
      if (cl != nl || cl <= 0 || cl >= nn) bad = true;
      nl = min(max(cl, 1), nn - 1);
      uint64_t below = (1ull << lane) - 1ull;
      int L = (int)__popcll(lm & below);  // lefts before this element
      int F = nl + (((lm >> nl) & 1ull) ? 0 : 1);
      int last = nn - 1;
      // Indices stay inside fr / bl and the node even when bad (which fails the build).
      if (in) {
        if (lane < F) {
          if (!isL) fr[min(lane - L, SMALL - 1)] = lane;
        } else if (isL) {
          bl[max(nl - 1 - L, 0)] = last - lane;
        }
      }
      __syncthreads();
      int np = lane;
      if (in) {
        if (lane < F) {
          if (!isL) {
            int k = lane - L;
            np = last - (k <= 0 ? 0 : bl[min(k - 1, SMALL - 1)] + 1);
          }
        } else {
          np = isL ? fr[max(nl - 1 - L, 0)] : lane - 1;
        }
        np = min(max(np, 0), last);
      }
      float v[9];
      for (int k = 0; k < 9; k++) v[k] = pf[k][p];
      uint32_t id = pid[p];
      __syncthreads();
      if (in) {
        for (int k = 0; k < 9; k++) pf[k][f0 + np] = v[k];
        pid[f0 + np] = id;
      }
    }
    const uint32_t base = rec0 + 2u * (uint32_t)inner;
    if (lane == 0) {
      nd.flags = (nd.flags & (NF_LEFT | NF_SUBROOT)) | (dec.kind == 2 ? NF_HALF : NF_SPLIT);
      nd.child = (int32_t)base;
      nd.axis = dec.kind == 1 ? dec.axis : 0;
      nd.lrank = inner;
      for (int j = 0; j < 6; j++) {
        nd.bkey[j] = nbkey[j];
        nd.ckey[j] = nckey[j];
      }
    }
    if (lane < 2) {  // lane 0 the left child, lane 1 the right (on top of the stack)
      int side = lane, e = sp + side;
      sfirst[e] = side == 0 ? f0 : f0 + nl;
      sn[e] = side == 0 ? nl : nn - nl;
      sdepth[e] = dep + 1;
      sglob[e] = (int)base + side;
      for (int j = 0; j < 6; j++) {
        skey[e][j] = dec.kind == 1 ? (side == 0 ? dec.bkey[0][j] : dec.bkey[1][j]) : 0u;
        skey[e][6 + j] = dec.kind == 1 ? (side == 0 ? dec.ckey[0][j] : dec.ckey[1][j]) : 0u;
      }
      BNode &cn = nodes[base + side];
      cn.first = first + sfirst[e];
      cn.n = sn[e];
      cn.depth = dep + 1;
      cn.parent = g;
      cn.flags = side == 0 ? NF_LEFT : 0;
      cn.child = -1;
      cn.axis = 0;
      cn.split = -1;
      cn.rank = cn.id = cn.ic = 0;
      cn.lrank = 0;
      cn.sub = root;
    }
    inner++;
    sp += 2;
    __syncthreads();
  }
  if (bad && lane == 0) atomicOr(&counters[C_ERROR], 4u);
  if (lane == 0) {
    nodes[root].ic = inner;
    nodes[root].flags |= NF_SUBROOT;
    nodes[root].sub = root;
    atomicAdd(&counters[C_SMALL_USED], 2u * (uint32_t)inner);
  }
  for (int r = 2 * inner + lane; r < 2 * n - 2; r += 64) nodes[rec0 + r].flags = NF_UNUSED;
  if (lane < n) order[first + lane] = pid[lane];
}

// Ranks and child ids inside the small subtrees, once the levels above have ranked the roots:
// records [begin, end) are the ones k_small made; the roots are listed in `roots`.
__global__ void k_sub_ranks(BNode *nodes, uint32_t begin, uint32_t end, const int32_t *roots, uint32_t nroots) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t r;
  if (i < nroots) r = (uint32_t)roots[i];
  else if (i - nroots < end - begin) r = begin + (i - nroots);
  else return;
  BNode &nd = nodes[r];
  if (nd.flags & (NF_LEAF | NF_UNUSED)) return;
  int rank = nodes[nd.sub].rank + nd.lrank;
  if (!(nd.flags & NF_SUBROOT)) nd.rank = rank;
  nodes[nd.child].id = 1 + 2 * rank;
  nodes[nd.child + 1].id = 2 + 2 * rank;
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

// The whole scan in one block for n <= 16 * 1024: each thread sums its run of up to 16 inputs,
// the block scans the run sums, and each thread writes its run's prefixes.
const uint32_t SCAN_ONE_MAX = 16 * 1024;
__global__ __launch_bounds__(1024) void k_scan_one(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *total) {
  const uint32_t per = (n + 1023) / 1024, start = threadIdx.x * per;
  uint32_t run = 0;
  for (uint32_t k = 0; k < per; k++)
    if (start + k < n) run += in[start + k];
  uint32_t tot, acc = tile_scan_1024(run, tot);
  for (uint32_t k = 0; k < per; k++)
    if (start + k < n) {
      uint32_t v = in[start + k];
      out[start + k] = acc;
      acc += v;
    }
  if (threadIdx.x == 0 && total) *total = tot;
}

// k_chunk_count, the scan of its counts and the clearing of `zero_n` counter words, in one block
// for levels of at most SCAN_ONE_MAX active nodes (k_scan_one's arithmetic).
__global__ __launch_bounds__(1024) void k_chunk_scan(const int32_t *active, uint32_t n, const BNode *nodes,
                                                     uint32_t *cnt, uint32_t *out, uint32_t *total,
                                                     uint32_t *zero, uint32_t zero_n) {
  if (threadIdx.x < zero_n) zero[threadIdx.x] = 0;
  const uint32_t per = (n + 1023) / 1024, start = threadIdx.x * per;
  uint32_t run = 0;
  for (uint32_t k = 0; k < per; k++)
    if (start + k < n) {
      uint32_t c = (uint32_t)((nodes[active[start + k]].n + CHUNK - 1) / CHUNK);
      cnt[start + k] = c;
      run += c;
    }
  uint32_t tot, acc = tile_scan_1024(run, tot);
  for (uint32_t k = 0; k < per; k++)
    if (start + k < n) {
      uint32_t v = cnt[start + k];
      out[start + k] = acc;
      acc += v;
    }
  if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_scan_add(uint32_t *out, uint32_t n, const uint32_t *tile_sums) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sums[i / 1024];
}

// Chunk c of the level: its node is the last one whose first chunk is <= c.
__global__ void k_chunk_emit(const int32_t *active, uint32_t nact, BNode *nodes, const uint32_t *cnt,
                             const uint32_t *off, uint32_t nchunks, int4 *chunks, uint32_t *bins, uint32_t words) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < words) {  // the level's bins: zero counts, empty bounds
    int f = (c % NB);
    bins[c] = f == 0 ? 0u : ((f - 1) / 3 % 2 == 0 ? KEY_POS_INF : KEY_NEG_INF);
  }
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
  chunks[c] = make_int4(lo, start, min(CHUNK, nd.first + nd.n - start), active[lo]);
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
  uint32_t p = (uint32_t)(ch.y + threadIdx.x);
  bool in = (int)threadIdx.x < ch.z;
  // The centroid's three coordinates go out with the node's read (the split axis is the node's):
  // one dependent round trip less per block.
  float cen[3] = {0, 0, 0};
  if (in)
    for (int a = 0; a < 3; a++) cen[a] = P.get(6 + a, p);
  const BNode &nd = nodes[ch.w];
  if (!(nd.flags & NF_SPLIT)) {  // uniform over the block
    if (threadIdx.x == 0) chunk_l[blockIdx.x] = 0;
    return;
  }
  NodeGeom g = node_geom(nd.ckey);
  uint32_t isL = 0;
  if (in) {
    const float c = nd.axis == 0 ? cen[0] : nd.axis == 1 ? cen[1] : cen[2];
    int b = bin_of((double)c, g.lo[nd.axis], g.scale[nd.axis]);
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
// Indices outside the node's range (a left count that disagrees with the classification, which
// k_part_offsets also reports) raise the error word instead of being written.
__global__ __launch_bounds__(256) void k_part_fb(const int4 *chunks, const int32_t *active,
                                                 const BNode *nodes, const uint8_t *cls,
                                                 const uint16_t *lpre, const uint32_t *chunk_off,
                                                 int32_t *fr, int32_t *bl, uint32_t *counters) {
  int4 ch = chunks[blockIdx.x];
  if ((int)threadIdx.x >= ch.z) return;
  int p = ch.y + (int)threadIdx.x;
  // The primitive's class and prefix go out with the node's read.
  const uint32_t cp = cls[p], lp = lpre[p], off = chunk_off[blockIdx.x];
  const BNode &nd = nodes[ch.w];
  if (!(nd.flags & NF_SPLIT)) return;
  int L = (int)(off - chunk_off[nd.chunk_begin] + lp);  // lefts in [first, p)
  int last = nd.first + nd.n - 1;
  int j = p < nd.first + nd.F ? (cp ? (p - nd.first) - L : -1) : (cp ? -1 : nd.nl - 1 - L);
  if (j < 0) {
    if (p < nd.first + nd.F ? cp : !cp) atomicOr(&counters[C_ERROR], 8u);
    return;
  }
  if (j >= nd.n) {
    atomicOr(&counters[C_ERROR], 8u);
    return;
  }
  if (p < nd.first + nd.F) fr[nd.first + j] = p - nd.first;
  else bl[nd.first + j] = last - p;
}

// Partition step 4: move every primitive of the level to its slot in the other buffer; SAH
// leaves write their final order.
__global__ __launch_bounds__(256) void k_scatter(const int4 *chunks, const int32_t *active,
                                                 const BNode *nodes, Prims A, Prims B,
                                                 const uint8_t *cls, const uint16_t *lpre,
                                                 const uint32_t *chunk_off, const int32_t *fr,
                                                 const int32_t *bl, uint32_t *order) {
  int4 ch = chunks[blockIdx.x];
  if ((int)threadIdx.x >= ch.z) return;
  int p = ch.y + (int)threadIdx.x;
  // The primitive, its class and prefix go out with the node's read: every case below moves the
  // primitive (or writes its id), so its loads need not wait for the node.
  float pf[9];
  for (int k = 0; k < 9; k++) pf[k] = A.get(k, (uint32_t)p);
  const uint32_t id = A.id[p], cp = cls[p], lp = lpre[p], off = chunk_off[blockIdx.x];
  const BNode &nd = nodes[ch.w];
  auto put = [&](int q) {
    for (int k = 0; k < 9; k++) B.set(k, (uint32_t)q, pf[k]);
    B.id[q] = id;
  };
  if (nd.flags & NF_SAHLEAF) {
    order[p] = id;
    return;
  }
  if (nd.flags & NF_HALF) {
    put(p);
    return;
  }
  int L = (int)(off - chunk_off[nd.chunk_begin] + lp);
  int first = nd.first, last = nd.first + nd.n - 1;
  int np;
  if (p < first + nd.F) {
    if (!cp) {
      np = p;
    } else {
      int k = (p - first) - L;
      np = k <= 0 || k > nd.n ? last : last - (bl[first + k - 1] + 1);
    }
  } else {
    int j = nd.nl - 1 - L;
    if (cp) np = p - 1;
    else np = j < 0 || j >= nd.n ? p : first + fr[first + j];
  }
  // A corrupt partition (reported through the error word) must still write inside the node.
  if (np < first || np > last) np = p;
  put(np);
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
  if (nd.flags & NF_SUBROOT) return;  // k_small counted it
  nd.ic = (nd.flags & NF_LEAF) ? 0 : 1 + nodes[nd.child].ic + nodes[nd.child + 1].ic;
}

// Right-first preorder ranks and final node ids, one level at a time, root first.
__global__ void k_ranks(BNode *nodes, uint32_t begin, uint32_t end) {
  uint32_t r = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= end) return;
  const BNode &nd = nodes[r];
  if (nd.flags & (NF_LEAF | NF_SUBROOT)) return;  // a subtree root's children: k_sub_ranks
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
  if (nd.flags & NF_UNUSED) return;
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
  r.nl = r.F = r.ic = r.rank = r.id = r.lrank = 0;
  r.sub = -1;
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
  if (n <= SCAN_ONE_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, sm, in, n, out, total);
    return;
  }
  uint32_t tiles = nblocks(n, 1024);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(1024), 0, sm, in, n, out, tile_sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, sm, tile_sums, tiles, total);
  if (tiles > 1) hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n)), dim3(256), 0, sm, out, n, tile_sums);
}

struct Layout {
  Prims P[2];
  BNode *nodes;
  int32_t *active[2];
  int32_t *leaf_list, *bounds_list, *small_list;
  uint32_t *small_cnt, *small_off;
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
  L.small_list = (int32_t *)take(sizeof(int32_t) * (2 * L.max_active + 2));
  L.small_cnt = (uint32_t *)take(sizeof(uint32_t) * (2 * L.max_active + 2));
  L.small_off = (uint32_t *)take(sizeof(uint32_t) * (2 * L.max_active + 2));
  L.chunks = (int4 *)take(sizeof(int4) * L.max_chunks);
  L.chunk_cnt = (uint32_t *)take(sizeof(uint32_t) * L.max_active);
  L.chunk_off_node = (uint32_t *)take(sizeof(uint32_t) * L.max_active);
  L.chunk_l = (uint32_t *)take(sizeof(uint32_t) * L.max_chunks);
  L.chunk_off = (uint32_t *)take(sizeof(uint32_t) * (L.max_chunks + 1));
  L.tile_sums = (uint32_t *)take(sizeof(uint32_t) * (std::max(L.max_chunks, 2 * L.max_active + 2) / 1024 + 2));
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
  uint32_t nact = 0, nsmall = 0, nchunks = 0;  // nchunks: chunks of the level's active nodes
  int cur = 0;
  // Host words copied to the device (the copies complete before the synchronisations below).
  static const int32_t zero = 0;
  static const uint32_t one = 1;
  if (n <= (uint32_t)MAX_LEAF) {
    hipLaunchKernelGGL(k_flush_leaves, dim3(1), dim3(1), 0, sm, L.bounds_list, 1u, L.nodes, L.P[0], order);
  } else if (n <= (uint32_t)SMALL) {  // the whole tree is one small subtree
    BTRY(hipMemcpyAsync(L.small_list, &zero, sizeof(int32_t), hipMemcpyHostToDevice, sm));
    BTRY(hipMemcpyAsync(L.counters + C_SMALL, &one, sizeof(uint32_t), hipMemcpyHostToDevice, sm));
    nsmall = 1;
  } else {
    BTRY(hipMemcpyAsync(L.active[0], &zero, sizeof(int32_t), hipMemcpyHostToDevice, sm));
    nact = 1;
    nchunks = (n + CHUNK - 1) / CHUNK;
  }
  int act = 0;
  while (nact > 0) {
    if (nact > L.max_active) {
      err = "builder: active node list overflow";
      return MCR_E_OVERFLOW;
    }
    int32_t *active = L.active[act], *next = L.active[act ^ 1];
    // chunks of the level's nodes; C_NEXT .. C_NEXT_CHUNKS cleared
    if (nact <= SCAN_ONE_MAX) {
      hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, sm, active, nact, L.nodes, L.chunk_cnt,
                         L.chunk_off_node, L.counters + C_CHUNKS, L.counters + C_NEXT, 4u);
    } else {
      hipLaunchKernelGGL(k_chunk_count, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.chunk_cnt);
      launch_scan(sm, L.chunk_cnt, nact, L.chunk_off_node, L.counters + C_CHUNKS, L.tile_sums);
      BTRY(hipMemsetAsync(L.counters + C_NEXT, 0, sizeof(uint32_t) * 4, sm));
    }
    if (nchunks > L.max_chunks) {
      err = "builder: chunk list overflow";
      return MCR_E_OVERFLOW;
    }
    // Copies of each node's bins: about one per 4 of its chunks, at most 16 (the bins buffer holds
    // max_active sets). The chunk records and the bins' initial values in one launch.
    uint32_t copies = (uint32_t)std::min<size_t>(
        {16, std::max<size_t>(1, L.max_active / nact), std::max<size_t>(1, nchunks / (4 * (size_t)nact))});
    uint32_t words = (uint32_t)BIN_WORDS * nact * copies;
    hipLaunchKernelGGL(k_chunk_emit, dim3(nblocks(std::max(nchunks, words))), dim3(256), 0, sm, active, nact, L.nodes,
                       L.chunk_cnt, L.chunk_off_node, nchunks, L.chunks, L.bins, words);
    hipLaunchKernelGGL(k_bin, dim3(nchunks), dim3(256), 0, sm, L.chunks, nchunks, active, L.nodes, L.P[cur], L.bins,
                       copies);
    hipLaunchKernelGGL(k_decide, dim3(nblocks(nact, 4)), dim3(256), 0, sm, active, nact, L.nodes, L.bins, copies,
                       L.counters, L.decisions);
    hipLaunchKernelGGL(k_create, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.decisions,
                       L.counters, next, L.leaf_list, L.bounds_list, L.small_list);
    hipLaunchKernelGGL(k_part_classify, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.P[cur],
                       L.cls, L.lpre, L.chunk_l);
    launch_scan(sm, L.chunk_l, nchunks, L.chunk_off, L.chunk_off + nchunks, L.tile_sums);
    hipLaunchKernelGGL(k_part_offsets, dim3(nblocks(nact)), dim3(256), 0, sm, active, nact, L.nodes, L.chunk_off,
                       L.cls, L.counters);
    hipLaunchKernelGGL(k_part_fb, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.cls, L.lpre,
                       L.chunk_off, L.fr, L.bl, L.counters);
    hipLaunchKernelGGL(k_scatter, dim3(nchunks), dim3(256), 0, sm, L.chunks, active, L.nodes, L.P[cur],
                       L.P[cur ^ 1], L.cls, L.lpre, L.chunk_off, L.fr, L.bl, order);
    BTRY(hipMemcpyAsync(ws.host, L.counters, sizeof(uint32_t) * C_COUNT, hipMemcpyDeviceToHost, sm));
    BTRY(hipStreamSynchronize(sm));
    uint32_t nbounds = ws.host[C_BOUNDS], nleaves = ws.host[C_LEAVES], nnext = ws.host[C_NEXT];
    nchunks = ws.host[C_NEXT_CHUNKS];
    nsmall = ws.host[C_SMALL];
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
  // The small subtrees, all at once; their records follow the levels' records.
  // records [top, span) hold the subtrees' nodes, some of them NF_UNUSED; `total` nodes in all
  uint32_t top = level_start.back(), total = top, span = top;
  if (nsmall) {
    hipLaunchKernelGGL(k_small_counts, dim3(nblocks(nsmall)), dim3(256), 0, sm, L.small_list, nsmall, L.nodes,
                       L.small_cnt);
    launch_scan(sm, L.small_cnt, nsmall, L.small_off, L.counters + C_SMALL_ALLOC, L.tile_sums);
    hipLaunchKernelGGL(k_small, dim3(nsmall), dim3(64), 0, sm, L.small_list, L.small_off, top, L.nodes, L.P[0],
                       L.P[1], order, L.counters);
    BTRY(hipMemcpyAsync(ws.host, L.counters, sizeof(uint32_t) * C_COUNT, hipMemcpyDeviceToHost, sm));
    BTRY(hipStreamSynchronize(sm));
    if (ws.host[C_ERROR]) {
      err = "builder: internal consistency check failed in a small subtree";
      return MCR_E_INVALID;
    }
    span = top + ws.host[C_SMALL_ALLOC];
    total = top + ws.host[C_SMALL_USED];
  }
  // Inner-node counts bottom-up, then ranks and ids top-down.
  for (int lv = (int)level_start.size() - 2; lv >= 0; lv--) {
    uint32_t b = level_start[lv], e = level_start[lv + 1];
    if (e > b) hipLaunchKernelGGL(k_inner_counts, dim3(nblocks(e - b)), dim3(256), 0, sm, L.nodes, b, e);
  }
  for (size_t lv = 0; lv + 1 < level_start.size(); lv++) {
    uint32_t b = level_start[lv], e = level_start[lv + 1];
    if (e > b) hipLaunchKernelGGL(k_ranks, dim3(nblocks(e - b)), dim3(256), 0, sm, L.nodes, b, e);
  }
  if (nsmall)
    hipLaunchKernelGGL(k_sub_ranks, dim3(nblocks((uint64_t)nsmall + (span - top))), dim3(256), 0, sm, L.nodes, top,
                       span, L.small_list, nsmall);
  hipLaunchKernelGGL(k_write_nodes, dim3(nblocks(span)), dim3(256), 0, sm, L.nodes, span, bounds, out_nodes);
  BTRY(hipGetLastError());
  BTRY(hipStreamSynchronize(sm));
  *node_count = total;
  return MCR_OK;
}
