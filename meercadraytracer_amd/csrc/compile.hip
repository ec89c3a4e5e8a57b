// compile.hip — scene flattening on the device: compileGeometry and addMesh (src/compile.ts:
// 180-306, 327-432) over plain mesh arrays (mcr_compile_count, mcr_compile_geometry).
//
// The host lays the meshes' drawn groups out as segments of consecutive output triangles (mesh
// order, then group order, as addMesh pushes them) and copies the vertex arrays to the device. One
// thread per triangle then does addMesh's per-triangle work in FP64 (world transform, normal
// matrix, Math.hypot normalisation or the face normal, uvs and colours) and rounds to the stored
// float arrays; the bit-exact builder orders the triangles; a second pass gathers every array in
// BVH order, writes the tangents (writeTangent) and reduces the vertex bounds for the epsilon.
#include <string.h>

#include <string>
#include <vector>

#include "context.h"
#include "jsmath.h"

namespace {

int fail(int code, const std::string &msg) { return mcr_fail(code, msg); }

const uint32_t INF_COUNT = 0xffffffffu;

struct MeshDev {
  double world[16];
  double nm[9];          // normalMatrix (compile.ts:192-205)
  int32_t mirror;        // det < 0: swap the second and third vertex (compile.ts:241-243)
  uint32_t vertex_count, color_size, pad;
  const float *pos, *nrm, *uv, *col;
  const uint32_t *index;
};

struct SegDev {          // one drawn group: triangles [base, base + count) of the output
  uint32_t base, count;
  uint32_t mesh, first;  // first index position k of the group's first triangle
  uint32_t material;
  uint8_t flags, side, flat, colors;
};

// compile.ts:192-205
void normal_matrix(const double *m, double *n, double &det) {
  double a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  det = a * A + b * B + c * C;
  double s = det != 0 ? 1 / det : 0;
  double v[9] = {A * s, (-(b * i - c * h)) * s, (b * f - c * e) * s,
                 B * s, (a * i - c * g) * s, (-(a * f - c * d)) * s,
                 C * s, (-(a * h - b * g)) * s, (a * e - b * d) * s};
  memcpy(n, v, sizeof(v));
}

// The loop bounds of addMesh for one group (compile.ts:220-225, 250-251): first index position
// and triangle count.
void group_range(const mcr_mesh &m, const mcr_mesh_group &g, uint32_t &first, uint32_t &count) {
  uint64_t total = m.index ? m.index_count : m.vertex_count;
  uint64_t start = m.draw_start;
  uint64_t end = m.draw_count == INF_COUNT ? total : std::min<uint64_t>(total, start + m.draw_count);
  uint64_t gend = g.count == INF_COUNT ? UINT64_MAX : (uint64_t)g.start + g.count;
  uint64_t from = std::max<uint64_t>(start, g.start), to = std::min<uint64_t>(end, gend);
  first = (uint32_t)from;
  count = to > from ? (uint32_t)((to - from) / 3) : 0;
}

int count_meshes(const mcr_mesh *meshes, uint32_t n, uint64_t &tris, bool &colors) {
  tris = 0;
  colors = false;
  for (uint32_t i = 0; i < n; i++) {
    const mcr_mesh &m = meshes[i];
    if (!m.positions && m.vertex_count) return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + " has no positions");
    if (m.group_count && !m.groups) return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + " has no groups array");
    if (m.colors && m.color_size != 3 && m.color_size != 4)
      return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + ": color_size must be 3 or 4");
    for (uint32_t k = 0; k < m.group_count; k++) {
      uint32_t first, count;
      group_range(m, m.groups[k], first, count);
      if (count && m.groups[k].vertex_colors) {
        if (!m.colors) return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + ": vertex colours without colours");
        colors = true;
      }
      tris += count;
    }
  }
  if (tris >= MCR_MAX_TRIANGLES) return fail(MCR_E_OVERFLOW, "too many triangles for this build (2^26 max)");
  return MCR_OK;
}

__device__ inline double fmax_bits(double a, double b) { return a > b ? a : b; }

// addMesh's per-triangle body (compile.ts:238-272) and the sink's push (compile.ts:337-352).
__global__ __launch_bounds__(256) void k_flatten(const SegDev *segs, uint32_t nsegs, const MeshDev *meshes,
                                                 uint32_t n, int any_colors, float *tri, float *nrm, float *uv,
                                                 float *col, uint32_t *mat, uint8_t *flags, uint8_t *sides,
                                                 unsigned long long *max_coord) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double mc = 0;
  if (i < n) {
    int lo = 0, hi = (int)nsegs - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (segs[mid].base <= i) lo = mid;
      else hi = mid - 1;
    }
    const SegDev &sg = segs[lo];
    const MeshDev &m = meshes[sg.mesh];
    uint32_t k = sg.first + 3 * (i - sg.base);
    uint32_t id[3] = {k, k + 1, k + 2};
    if (m.index) {
      id[0] = m.index[k];
      id[1] = m.index[k + 1];
      id[2] = m.index[k + 2];
    }
    if (m.mirror) {
      uint32_t t = id[1];
      id[1] = id[2];
      id[2] = t;
    }
    const double *w = m.world;
    double p[9], nv[9];
    for (int v = 0; v < 3; v++) {
      const float *q = m.pos + (size_t)id[v] * 3;
      double x = q[0], y = q[1], z = q[2];
      p[v * 3] = w[0] * x + w[4] * y + w[8] * z + w[12];
      p[v * 3 + 1] = w[1] * x + w[5] * y + w[9] * z + w[13];
      p[v * 3 + 2] = w[2] * x + w[6] * y + w[10] * z + w[14];
      if (m.nrm) {
        const float *r = m.nrm + (size_t)id[v] * 3;
        double a = r[0], b = r[1], c = r[2];
        double nx = m.nm[0] * a + m.nm[3] * b + m.nm[6] * c;
        double ny = m.nm[1] * a + m.nm[4] * b + m.nm[7] * c;
        double nz = m.nm[2] * a + m.nm[5] * b + m.nm[8] * c;
        double len = js_hypot3(nx, ny, nz);
        if (!(len != 0)) len = 1;  // Math.hypot(...) || 1
        nv[v * 3] = nx / len;
        nv[v * 3 + 1] = ny / len;
        nv[v * 3 + 2] = nz / len;
      }
    }
    if (!m.nrm || sg.flat) {  // the face normal (compile.ts:262-268)
      double ex = p[3] - p[0], ey = p[4] - p[1], ez = p[5] - p[2];
      double fx = p[6] - p[0], fy = p[7] - p[1], fz = p[8] - p[2];
      double cx = ey * fz - ez * fy, cy = ez * fx - ex * fz, cz = ex * fy - ey * fx;
      double len = js_hypot3(cx, cy, cz);
      double f[3] = {0, 0, 1};
      if (len > 0) {
        f[0] = cx / len;
        f[1] = cy / len;
        f[2] = cz / len;
      }
      for (int v = 0; v < 3; v++) {
        nv[v * 3] = f[0];
        nv[v * 3 + 1] = f[1];
        nv[v * 3 + 2] = f[2];
      }
    }
    float *t = tri + (size_t)i * 9;
    for (int c = 0; c < 3; c++) {
      t[c] = (float)p[c];
      t[3 + c] = (float)(p[3 + c] - p[c]);
      t[6 + c] = (float)(p[6 + c] - p[c]);
    }
    for (int c = 0; c < 9; c++) {
      nrm[(size_t)i * 9 + c] = (float)nv[c];
      mc = fmax_bits(mc, jm_fabs(p[c]));
    }
    for (int v = 0; v < 3; v++) {
      uv[(size_t)i * 6 + v * 2] = m.uv ? m.uv[(size_t)id[v] * 2] : 0.0f;
      uv[(size_t)i * 6 + v * 2 + 1] = m.uv ? m.uv[(size_t)id[v] * 2 + 1] : 0.0f;
    }
    if (any_colors) {
      float *o = col + (size_t)i * 12;
      for (int v = 0; v < 3; v++) {
        if (sg.colors) {
          const float *c = m.col + (size_t)id[v] * m.color_size;
          o[v * 4] = c[0];
          o[v * 4 + 1] = c[1];
          o[v * 4 + 2] = c[2];
          o[v * 4 + 3] = m.color_size > 3 ? c[3] : 1.0f;
        } else {
          o[v * 4] = o[v * 4 + 1] = o[v * 4 + 2] = o[v * 4 + 3] = 1.0f;
        }
      }
    }
    mat[i] = sg.material;
    flags[i] = sg.flags;
    sides[i] = sg.side;
  }
  // maxCoordinate (compile.ts:344): non-negative doubles order as their bit patterns.
  for (int off = 32; off > 0; off >>= 1) mc = fmax_bits(mc, __shfl_xor(mc, off, 64));
  if ((threadIdx.x & 63) == 0 && mc > 0) atomicMax(max_coord, (unsigned long long)__double_as_longlong(mc));
}

// writeTangent (compile.ts:406-432) on the stored float arrays.
__device__ void write_tangent(const float *t, const float *u, float *out) {
  double e1x = t[3], e1y = t[4], e1z = t[5], e2x = t[6], e2y = t[7], e2z = t[8];
  double du1 = (double)u[2] - u[0], dv1 = (double)u[3] - u[1];
  double du2 = (double)u[4] - u[0], dv2 = (double)u[5] - u[1];
  double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  double d = du1 * dv2 - du2 * dv1;
  double tx, ty, tz, sign = 1;
  if (jm_fabs(d) > 1e-20) {
    double r = 1 / d;
    tx = (e1x * dv2 - e2x * dv1) * r;
    ty = (e1y * dv2 - e2y * dv1) * r;
    tz = (e1z * dv2 - e2z * dv1) * r;
    double bx = (e2x * du1 - e1x * du2) * r, by = (e2y * du1 - e1y * du2) * r, bz = (e2z * du1 - e1z * du2) * r;
    double cx = ny * tz - nz * ty, cy = nz * tx - nx * tz, cz = nx * ty - ny * tx;
    sign = cx * bx + cy * by + cz * bz < 0 ? -1 : 1;
  } else if (jm_fabs(nx) < 0.9 * js_hypot3(nx, ny, nz)) {
    tx = 0;
    ty = -nz;
    tz = ny;
  } else {
    tx = nz;
    ty = 0;
    tz = -nx;
  }
  double len = js_hypot3(tx, ty, tz);
  if (!(len != 0)) len = 1;
  out[0] = (float)(tx / len);
  out[1] = (float)(ty / len);
  out[2] = (float)(tz / len);
  out[3] = (float)sign;
}

// The BVH-order gather of compile.ts:364-380, tangents, and per-block vertex bounds.
__global__ __launch_bounds__(256) void k_gather(const uint32_t *order, uint32_t n, int any_colors,
                                                const float *tri, const float *nrm, const float *uv,
                                                const float *col, const uint32_t *mat, const uint8_t *flags,
                                                const uint8_t *sides, float *otri, float *onrm, float *ouv,
                                                float *ocol, float *otan, uint32_t *omat, uint8_t *oflags,
                                                uint8_t *osides, double *block_bounds) {
  __shared__ double red[6][256];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double b[6] = {JM_INF, JM_INF, JM_INF, -JM_INF, -JM_INF, -JM_INF};
  if (i < n) {
    uint32_t s = order[i];
    float *t = otri + (size_t)i * 9;
    for (int c = 0; c < 9; c++) {
      t[c] = tri[(size_t)s * 9 + c];
      onrm[(size_t)i * 9 + c] = nrm[(size_t)s * 9 + c];
    }
    for (int c = 0; c < 6; c++) ouv[(size_t)i * 6 + c] = uv[(size_t)s * 6 + c];
    if (any_colors)
      for (int c = 0; c < 12; c++) ocol[(size_t)i * 12 + c] = col[(size_t)s * 12 + c];
    omat[i] = mat[s];
    oflags[i] = flags[s];
    osides[i] = sides[s];
    for (int v = 0; v < 3; v++) {
      double x = (double)t[0] + (v == 1 ? (double)t[3] : v == 2 ? (double)t[6] : 0.0);
      double y = (double)t[1] + (v == 1 ? (double)t[4] : v == 2 ? (double)t[7] : 0.0);
      double z = (double)t[2] + (v == 1 ? (double)t[5] : v == 2 ? (double)t[8] : 0.0);
      b[0] = x < b[0] ? x : b[0];
      b[1] = y < b[1] ? y : b[1];
      b[2] = z < b[2] ? z : b[2];
      b[3] = x > b[3] ? x : b[3];
      b[4] = y > b[4] ? y : b[4];
      b[5] = z > b[5] ? z : b[5];
    }
    write_tangent(t, ouv + (size_t)i * 6, otan + (size_t)i * 4);
  }
  for (int c = 0; c < 6; c++) red[c][threadIdx.x] = b[c];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      for (int c = 0; c < 3; c++) {
        double a = red[c][threadIdx.x + h];
        if (a < red[c][threadIdx.x]) red[c][threadIdx.x] = a;
        double z = red[c + 3][threadIdx.x + h];
        if (z > red[c + 3][threadIdx.x]) red[c + 3][threadIdx.x] = z;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) block_bounds[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// flatten + build + gather on the device (compileGeometry up to its stored arrays). The results
// stay in one device allocation (`dev`, freed with the Pipeline) at the o_* offsets; the caller
// holds the context lock and either downloads them or builds a resident scene from them.
struct Pipeline {
  char *dev = nullptr;
  uint32_t n = 0, node_count = 0;
  bool any_colors = false;
  size_t o_otri = 0, o_onrm = 0, o_ouv = 0, o_ocol = 0, o_otan = 0, o_omat = 0, o_oflg = 0, o_osid = 0;
  size_t o_bounds = 0, o_nodes = 0;
  double epsilon = 1e-5;
  Pipeline() = default;
  Pipeline(const Pipeline &) = delete;
  Pipeline &operator=(const Pipeline &) = delete;
  ~Pipeline() { hipFree(dev); }
};

int run_pipeline(mcr_ctx *c, const mcr_mesh *meshes, uint32_t mesh_count, uint32_t n, bool any_colors,
                 Pipeline &pl) {
  pl.n = n;
  pl.any_colors = any_colors;
  // Segments and per-mesh constants; vertex arrays packed into one upload.
  std::vector<SegDev> segs;
  std::vector<MeshDev> mdev(mesh_count);
  std::vector<size_t> offs(mesh_count * 5, 0);
  std::vector<char> used(mesh_count, 0);
  size_t bytes = 0;
  auto take = [&](size_t b) {
    size_t o = (bytes + 255) / 256 * 256;
    bytes = o + b;
    return o;
  };
  uint32_t base = 0;
  for (uint32_t i = 0; i < mesh_count; i++) {
    const mcr_mesh &m = meshes[i];
    MeshDev &d = mdev[i];
    memset(&d, 0, sizeof(d));
    memcpy(d.world, m.world, sizeof(d.world));
    double det;
    normal_matrix(m.world, d.nm, det);
    d.mirror = det < 0 ? 1 : 0;
    d.vertex_count = m.vertex_count;
    d.color_size = m.color_size;
    for (uint32_t k = 0; k < m.group_count; k++) {
      uint32_t first, count;
      group_range(m, m.groups[k], first, count);
      if (!count) continue;
      used[i] = 1;
      // Every index the group reads must name a vertex.
      if (m.index) {
        if ((uint64_t)first + 3ull * count > m.index_count)
          return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + ": group beyond the index");
        for (uint64_t j = first; j < (uint64_t)first + 3ull * count; j++)
          if (m.index[j] >= m.vertex_count)
            return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + ": index " + std::to_string(j) + " out of range");
      } else if ((uint64_t)first + 3ull * count > m.vertex_count) {
        return fail(MCR_E_INVALID, "mesh " + std::to_string(i) + ": group beyond the vertices");
      }
      const mcr_mesh_group &g = m.groups[k];
      segs.push_back(SegDev{base, count, i, first, g.material, g.flags, g.side, g.flat_shading,
                            (uint8_t)(g.vertex_colors ? 1 : 0)});
      base += count;
    }
    if (!used[i]) continue;
    offs[i * 5 + 0] = take((size_t)m.vertex_count * 12);
    if (m.normals) offs[i * 5 + 1] = take((size_t)m.vertex_count * 12);
    if (m.uvs) offs[i * 5 + 2] = take((size_t)m.vertex_count * 8);
    if (m.colors) offs[i * 5 + 3] = take((size_t)m.vertex_count * m.color_size * 4);
    if (m.index) offs[i * 5 + 4] = take((size_t)m.index_count * 4);
  }
  size_t o_mesh = take(sizeof(MeshDev) * mesh_count), o_seg = take(sizeof(SegDev) * segs.size());
  // Unordered arrays, ordered ones come straight from the builder and the gather.
  size_t o_tri = take((size_t)n * 36), o_nrm = take((size_t)n * 36), o_uv = take((size_t)n * 24);
  size_t o_col = take(any_colors ? (size_t)n * 48 : 4), o_mat = take((size_t)n * 4), o_flg = take(n), o_sid = take(n);
  size_t o_otri = take((size_t)n * 36), o_onrm = take((size_t)n * 36), o_ouv = take((size_t)n * 24);
  size_t o_ocol = take(any_colors ? (size_t)n * 48 : 4), o_otan = take((size_t)n * 16), o_omat = take((size_t)n * 4);
  size_t o_oflg = take(n), o_osid = take(n);
  size_t nodes_cap = 2 * (size_t)n - 1;
  size_t o_bounds = take(nodes_cap * 24), o_nodes = take(nodes_cap * 12), o_order = take((size_t)n * 4);
  uint32_t blocks = (n + 255) / 256;
  size_t o_bb = take((size_t)blocks * 48), o_maxc = take(8);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  char *dev = nullptr;
  HIP_TRY(hipMalloc((void **)&dev, bytes));
  std::vector<double> bb((size_t)blocks * 6);
  unsigned long long maxc_bits = 0;
  uint32_t node_count = 0;
  int rc = [&]() -> int {
    for (uint32_t i = 0; i < mesh_count; i++) {
      const mcr_mesh &m = meshes[i];
      if (!used[i]) continue;
      MeshDev &d = mdev[i];
      HIP_TRY(hipMemcpyAsync(dev + offs[i * 5 + 0], m.positions, (size_t)m.vertex_count * 12, hipMemcpyHostToDevice, sm));
      d.pos = (const float *)(dev + offs[i * 5 + 0]);
      if (m.normals) {
        HIP_TRY(hipMemcpyAsync(dev + offs[i * 5 + 1], m.normals, (size_t)m.vertex_count * 12, hipMemcpyHostToDevice, sm));
        d.nrm = (const float *)(dev + offs[i * 5 + 1]);
      }
      if (m.uvs) {
        HIP_TRY(hipMemcpyAsync(dev + offs[i * 5 + 2], m.uvs, (size_t)m.vertex_count * 8, hipMemcpyHostToDevice, sm));
        d.uv = (const float *)(dev + offs[i * 5 + 2]);
      }
      if (m.colors) {
        HIP_TRY(hipMemcpyAsync(dev + offs[i * 5 + 3], m.colors, (size_t)m.vertex_count * m.color_size * 4,
                               hipMemcpyHostToDevice, sm));
        d.col = (const float *)(dev + offs[i * 5 + 3]);
      }
      if (m.index) {
        HIP_TRY(hipMemcpyAsync(dev + offs[i * 5 + 4], m.index, (size_t)m.index_count * 4, hipMemcpyHostToDevice, sm));
        d.index = (const uint32_t *)(dev + offs[i * 5 + 4]);
      }
    }
    HIP_TRY(hipMemcpyAsync(dev + o_mesh, mdev.data(), sizeof(MeshDev) * mesh_count, hipMemcpyHostToDevice, sm));
    HIP_TRY(hipMemcpyAsync(dev + o_seg, segs.data(), sizeof(SegDev) * segs.size(), hipMemcpyHostToDevice, sm));
    HIP_TRY(hipMemsetAsync(dev + o_maxc, 0, 8, sm));
    hipLaunchKernelGGL(k_flatten, dim3(blocks), dim3(256), 0, sm, (const SegDev *)(dev + o_seg), (uint32_t)segs.size(),
                       (const MeshDev *)(dev + o_mesh), n, any_colors ? 1 : 0, (float *)(dev + o_tri),
                       (float *)(dev + o_nrm), (float *)(dev + o_uv), (float *)(dev + o_col), (uint32_t *)(dev + o_mat),
                       (uint8_t *)(dev + o_flg), (uint8_t *)(dev + o_sid), (unsigned long long *)(dev + o_maxc));
    HIP_TRY(hipGetLastError());
    std::string err;
    int brc = builder_run(c->builder, sm, (const float *)(dev + o_tri), n, (float *)(dev + o_bounds),
                          (int32_t *)(dev + o_nodes), (uint32_t *)(dev + o_order), &node_count, err);
    if (brc) return fail(brc, err);
    hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, sm, (const uint32_t *)(dev + o_order), n,
                       any_colors ? 1 : 0, (const float *)(dev + o_tri), (const float *)(dev + o_nrm),
                       (const float *)(dev + o_uv), (const float *)(dev + o_col), (const uint32_t *)(dev + o_mat),
                       (const uint8_t *)(dev + o_flg), (const uint8_t *)(dev + o_sid), (float *)(dev + o_otri),
                       (float *)(dev + o_onrm), (float *)(dev + o_ouv), (float *)(dev + o_ocol),
                       (float *)(dev + o_otan), (uint32_t *)(dev + o_omat), (uint8_t *)(dev + o_oflg),
                       (uint8_t *)(dev + o_osid), (double *)(dev + o_bb));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(bb.data(), dev + o_bb, bb.size() * 8, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipMemcpyAsync(&maxc_bits, dev + o_maxc, 8, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    return MCR_OK;
  }();
  pl.dev = dev;
  if (rc) return rc;
  double lo[3] = {JM_INF, JM_INF, JM_INF}, hi[3] = {-JM_INF, -JM_INF, -JM_INF};
  for (uint32_t b = 0; b < blocks; b++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], bb[(size_t)b * 6 + k]);
      hi[k] = std::max(hi[k], bb[(size_t)b * 6 + 3 + k]);
    }
  double max_coord;
  memcpy(&max_coord, &maxc_bits, 8);
  double diagonal = js_hypot3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
  pl.node_count = node_count;
  pl.epsilon = std::max(std::max(diagonal * 1e-5, max_coord * 2e-6), 1e-7);
  pl.o_otri = o_otri;
  pl.o_onrm = o_onrm;
  pl.o_ouv = o_ouv;
  pl.o_ocol = o_ocol;
  pl.o_otan = o_otan;
  pl.o_omat = o_omat;
  pl.o_oflg = o_oflg;
  pl.o_osid = o_osid;
  pl.o_bounds = o_bounds;
  pl.o_nodes = o_nodes;
  return MCR_OK;
}

// Triangle records of the resident layout (48 B: v0, e1, e2, side | flags << 8, material).
__global__ __launch_bounds__(256) void k_tri_records(uint32_t n, const float *tri, const uint32_t *mat,
                                                     const uint8_t *flags, const uint8_t *sides, float4 *out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float *t = tri + (size_t)k * 9;
  uint32_t word = (uint32_t)sides[k] | ((uint32_t)flags[k] << 8);
  out[(size_t)k * 3 + 0] = make_float4(t[0], t[1], t[2], t[3]);
  out[(size_t)k * 3 + 1] = make_float4(t[4], t[5], t[6], t[7]);
  out[(size_t)k * 3 + 2] = make_float4(t[8], __uint_as_float(word), __uint_as_float(mat[k]), 0.0f);
}

// Child ref of node c in the resident layout: the inner record of an inner node is its rank r in
// buildBvh's allocation order, read off its children's indices (1 + 2r, 2 + 2r; bvh.ts:165-171);
// a leaf packs into the ref unless it holds over 31 triangles (then *long_leaf is set and the
// caller lays the scene out on the host instead).
__device__ inline int32_t resident_ref(const int32_t *nodes, uint32_t c, uint32_t *long_leaf) {
  int32_t first = nodes[(size_t)c * 3], count = nodes[(size_t)c * 3 + 1];
  if (count == 0) return (first - 1) / 2;
  if (count <= 31 && (uint32_t)first < MCR_MAX_TRIANGLES)
    return (int32_t)(MCR_LEAF_BIT | ((uint32_t)count << MCR_LEAF_FIRST_BITS) | (uint32_t)first);
  atomicOr(long_leaf, 1u);
  return -1;
}

// Inner records of the resident layout, one thread per BVH node: both children's boxes, both refs
// and the split axis, at the node's rank (device.h layout).
__global__ __launch_bounds__(256) void k_inner_records(uint32_t node_count, const int32_t *nodes,
                                                       const float *bounds, float4 *inner, uint32_t *long_leaf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= node_count || nodes[(size_t)i * 3 + 1] != 0) return;
  int32_t left = nodes[(size_t)i * 3];
  const float *bl = bounds + (size_t)left * 6, *br = bounds + (size_t)(left + 1) * 6;
  size_t r = (size_t)((left - 1) / 2) * 4;
  inner[r + 0] = make_float4(bl[0], bl[1], bl[2], bl[3]);
  inner[r + 1] = make_float4(bl[4], bl[5], br[0], br[1]);
  inner[r + 2] = make_float4(br[2], br[3], br[4], br[5]);
  int4 m = make_int4(resident_ref(nodes, (uint32_t)left, long_leaf), resident_ref(nodes, (uint32_t)left + 1, long_leaf),
                     nodes[(size_t)i * 3 + 2], 0);
  *reinterpret_cast<int4 *>(&inner[r + 3]) = m;
}

}  // namespace

extern "C" {

int mcr_compile_count(const mcr_mesh *meshes, uint32_t mesh_count, uint32_t *triangle_count_out,
                      int *any_colors_out) {
  if ((!meshes && mesh_count) || !triangle_count_out) return fail(MCR_E_INVALID, "null argument");
  uint64_t tris;
  bool colors;
  MCR_TRY(count_meshes(meshes, mesh_count, tris, colors));
  *triangle_count_out = (uint32_t)tris;
  if (any_colors_out) *any_colors_out = colors ? 1 : 0;
  return MCR_OK;
}

int mcr_compile_geometry(mcr_ctx *c, const mcr_mesh *meshes, uint32_t mesh_count, mcr_geometry_out *out) {
  if (!c || (!meshes && mesh_count) || !out) return fail(MCR_E_INVALID, "null argument");
  uint64_t tris;
  bool any_colors;
  MCR_TRY(count_meshes(meshes, mesh_count, tris, any_colors));
  uint32_t n = (uint32_t)tris;
  if (out->triangle_count != n)
    return fail(MCR_E_INVALID, "out->triangle_count must be the count mcr_compile_count gives (" +
                                   std::to_string(n) + ")");
  if (n == 0) {
    out->node_count = 0;
    out->epsilon = 1e-5;  // Math.max(1 * 1e-5, 0 * 2e-6, 1e-7): the diagonal of an empty scene is 1
    return MCR_OK;
  }
  if (!out->triangles || !out->normals || !out->uvs || !out->tangents || !out->materials || !out->flags ||
      !out->sides || !out->bvh_bounds || !out->bvh_nodes || (any_colors && !out->colors))
    return fail(MCR_E_INVALID, "output arrays missing");
  std::lock_guard<std::mutex> lock(c->mu);
  Pipeline pl;
  MCR_TRY(run_pipeline(c, meshes, mesh_count, n, any_colors, pl));
  hipStream_t sm = c->stream;
  auto down = [&](void *dst, size_t off, size_t b) { return hipMemcpyAsync(dst, pl.dev + off, b, hipMemcpyDeviceToHost, sm); };
  HIP_TRY(down(out->triangles, pl.o_otri, (size_t)n * 36));
  HIP_TRY(down(out->normals, pl.o_onrm, (size_t)n * 36));
  HIP_TRY(down(out->uvs, pl.o_ouv, (size_t)n * 24));
  HIP_TRY(down(out->tangents, pl.o_otan, (size_t)n * 16));
  if (any_colors) HIP_TRY(down(out->colors, pl.o_ocol, (size_t)n * 48));
  HIP_TRY(down(out->materials, pl.o_omat, (size_t)n * 4));
  HIP_TRY(down(out->flags, pl.o_oflg, n));
  HIP_TRY(down(out->sides, pl.o_osid, n));
  HIP_TRY(down(out->bvh_bounds, pl.o_bounds, (size_t)pl.node_count * 24));
  HIP_TRY(down(out->bvh_nodes, pl.o_nodes, (size_t)pl.node_count * 12));
  HIP_TRY(hipStreamSynchronize(sm));
  out->node_count = pl.node_count;
  out->epsilon = pl.epsilon;
  return MCR_OK;
}

int mcr_scene_compile_geometry(mcr_scene *s, const mcr_mesh *meshes, uint32_t mesh_count, uint32_t version,
                               mcr_compile_info *info) {
  if (!s || (!meshes && mesh_count)) return fail(MCR_E_INVALID, "null argument");
  if (scene_held(s)) return scene_held_error();
  uint64_t tris;
  bool any_colors;
  MCR_TRY(count_meshes(meshes, mesh_count, tris, any_colors));
  uint32_t n = (uint32_t)tris;
  mcr_ctx *c = s->ctx;
  if (n == 0) {  // an empty scene, as mcr_scene_set_geometry lays it out
    mcr_geometry g = {};
    g.epsilon = 1e-5;
    MCR_TRY(mcr_scene_set_geometry(s, &g, version));
    if (info) *info = mcr_compile_info{0, 0, 0, g.epsilon};
    return MCR_OK;
  }
  std::unique_lock<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  HIP_TRY(hipStreamSynchronize(sm));
  Pipeline pl;
  MCR_TRY(run_pipeline(c, meshes, mesh_count, n, any_colors, pl));
  uint32_t nodes = pl.node_count, ninner = (nodes - 1) / 2;
  free_geometry(s);
  uint32_t *long_leaf = nullptr;
  int rc = [&]() -> int {
    HIP_TRY(dalloc(&s->tri, (size_t)n * 3));
    HIP_TRY(dalloc(&s->inner, (size_t)std::max<uint32_t>(ninner, 1) * 4));
    HIP_TRY(dalloc(&s->normals, (size_t)n * 9));
    HIP_TRY(dalloc(&s->uvs, (size_t)n * 6));
    HIP_TRY(dalloc(&s->tangents, (size_t)n * 4));
    if (any_colors) HIP_TRY(dalloc(&s->colors, (size_t)n * 12));
    HIP_TRY(dalloc(&long_leaf, 1));
    HIP_TRY(hipMemsetAsync(long_leaf, 0, 4, sm));
    uint32_t tb = (n + 255) / 256, nb = (nodes + 255) / 256;
    hipLaunchKernelGGL(k_tri_records, dim3(tb), dim3(256), 0, sm, n, (const float *)(pl.dev + pl.o_otri),
                       (const uint32_t *)(pl.dev + pl.o_omat), (const uint8_t *)(pl.dev + pl.o_oflg),
                       (const uint8_t *)(pl.dev + pl.o_osid), (float4 *)s->tri);
    hipLaunchKernelGGL(k_inner_records, dim3(nb), dim3(256), 0, sm, nodes, (const int32_t *)(pl.dev + pl.o_nodes),
                       (const float *)(pl.dev + pl.o_bounds), (float4 *)s->inner, long_leaf);
    HIP_TRY(hipGetLastError());
    auto copy = [&](void *dst, size_t off, size_t b) {
      return hipMemcpyAsync(dst, pl.dev + off, b, hipMemcpyDeviceToDevice, sm);
    };
    HIP_TRY(copy(s->normals, pl.o_onrm, (size_t)n * 36));
    HIP_TRY(copy(s->uvs, pl.o_ouv, (size_t)n * 24));
    HIP_TRY(copy(s->tangents, pl.o_otan, (size_t)n * 16));
    if (any_colors) HIP_TRY(copy(s->colors, pl.o_ocol, (size_t)n * 48));
    return MCR_OK;
  }();
  uint32_t long_host = 0;
  int32_t root[3] = {0, 0, 0};
  float root_box[6] = {0, 0, 0, 0, 0, 0};
  if (!rc) {
    rc = [&]() -> int {
      HIP_TRY(hipMemcpyAsync(&long_host, long_leaf, 4, hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipMemcpyAsync(root, pl.dev + pl.o_nodes, 12, hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipMemcpyAsync(root_box, pl.dev + pl.o_bounds, 24, hipMemcpyDeviceToHost, sm));
      HIP_TRY(hipStreamSynchronize(sm));
      return MCR_OK;
    }();
  }
  hipFree(long_leaf);
  if (rc || long_host || (root[1] > 31)) {
    free_geometry(s);
    if (rc) return rc;
    // Leaves over 31 triangles (only at buildBvh's depth limit) need the host layout's long-leaf
    // table: lay the scene out from the downloaded arrays instead.
    std::vector<float> tri((size_t)n * 9), nrm((size_t)n * 9), uv((size_t)n * 6), tan((size_t)n * 4),
        col(any_colors ? (size_t)n * 12 : 0), bnd((size_t)nodes * 6);
    std::vector<uint32_t> mat(n);
    std::vector<uint8_t> flg(n), sid(n);
    std::vector<int32_t> nod((size_t)nodes * 3);
    auto down = [&](void *dst, size_t off, size_t b) { return hipMemcpyAsync(dst, pl.dev + off, b, hipMemcpyDeviceToHost, sm); };
    HIP_TRY(down(tri.data(), pl.o_otri, tri.size() * 4));
    HIP_TRY(down(nrm.data(), pl.o_onrm, nrm.size() * 4));
    HIP_TRY(down(uv.data(), pl.o_ouv, uv.size() * 4));
    HIP_TRY(down(tan.data(), pl.o_otan, tan.size() * 4));
    if (any_colors) HIP_TRY(down(col.data(), pl.o_ocol, col.size() * 4));
    HIP_TRY(down(mat.data(), pl.o_omat, (size_t)n * 4));
    HIP_TRY(down(flg.data(), pl.o_oflg, n));
    HIP_TRY(down(sid.data(), pl.o_osid, n));
    HIP_TRY(down(bnd.data(), pl.o_bounds, bnd.size() * 4));
    HIP_TRY(down(nod.data(), pl.o_nodes, nod.size() * 4));
    HIP_TRY(hipStreamSynchronize(sm));
    lock.unlock();
    mcr_geometry g = {n, nodes, tri.data(), nrm.data(), uv.data(), tan.data(), any_colors ? col.data() : nullptr,
                      mat.data(), flg.data(), sid.data(), bnd.data(), nod.data(), pl.epsilon};
    MCR_TRY(mcr_scene_set_geometry(s, &g, version));
    if (info) *info = mcr_compile_info{n, nodes, 1, pl.epsilon};
    return MCR_OK;
  }
  DevScene &d = s->d;
  d.triangle_count = n;
  d.inner_count = ninner;
  d.epsilon = pl.epsilon;
  d.root_ref = root[1] == 0 ? 0 : (int32_t)(MCR_LEAF_BIT | ((uint32_t)root[1] << MCR_LEAF_FIRST_BITS) | (uint32_t)root[0]);
  for (int k = 0; k < 6; k++) d.root_box[k] = root_box[k];
  d.tri = s->tri;
  d.inner = s->inner;
  d.longleaf = nullptr;
  d.normals = s->normals;
  d.colors = s->colors;
  d.uvs = s->uvs;
  d.tangents = s->tangents;
  s->bytes_geometry = (size_t)n * 48 + (size_t)ninner * 64 + (size_t)n * (9 + 6 + 4) * 4 + (any_colors ? (size_t)n * 48 : 0);
  s->geometry_version = version;
  s->has_geometry = true;
  if (info) *info = mcr_compile_info{n, nodes, 0, pl.epsilon};
  return MCR_OK;
}

}  // extern "C"
