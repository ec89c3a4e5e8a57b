// api.hip — the C-ABI of include/mcr.h: context, resident scenes, batched tile rendering, ray
// queries, tile encoding and timing. Host orchestration of the wavefront loop lives here; the
// kernels are in trace.hip and bvh_build.hip.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcr.h"
#include "builder.h"
#include "context.h"
#include "device.h"
#include "dfg_lut.h"
#include "kernels.h"
#include "shading.h"
#include "wavefront.h"

namespace {
thread_local std::string g_error;
}  // namespace

int mcr_fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

// Releases a scene's geometry arrays (and its wide BVH).
void free_geometry(mcr_scene *s) {
  wide_free(s);
  hipFree(s->tri);
  hipFree(s->inner);
  hipFree(s->longleaf);
  hipFree(s->normals);
  hipFree(s->colors);
  hipFree(s->tangents);
  hipFree(s->uvs);
  s->tangents = nullptr;
  s->uvs = nullptr;
  s->tri = nullptr;
  s->inner = nullptr;
  s->longleaf = nullptr;
  s->normals = nullptr;
  s->colors = nullptr;
  s->bytes_geometry = 0;
  s->has_geometry = false;
}

namespace {

int fail(int code, const std::string &msg) { return mcr_fail(code, msg); }


int ensure_out_tmp(mcr_ctx *c, size_t floats) {
  if (c->out_tmp_cap >= floats) return MCR_OK;
  hipFree(c->out_tmp);
  c->out_tmp = nullptr;
  c->out_tmp_cap = 0;
  HIP_TRY(dalloc(&c->out_tmp, floats));
  c->out_tmp_cap = floats;
  return MCR_OK;
}

bool shading_supported(const mcr_shading *s, std::string &why) {
  auto tex_ok = [&](int32_t i) { return i == -1 || (i >= 0 && (uint32_t)i < s->texture_count); };
  auto env_ok = [&](int32_t i) { return i == -1 || (i >= 0 && (uint32_t)i < s->environment_count); };
  if ((s->texture_count && !s->textures) || (s->environment_count && !s->environments)) {
    why = "texture or environment array missing";
    return false;
  }
  for (uint32_t i = 0; i < s->texture_count; i++) {
    const mcr_texture &t = s->textures[i];
    if (!t.data || t.width == 0 || t.height == 0 || (uint64_t)t.width * t.height > (1ull << 28)) {
      why = "texture " + std::to_string(i) + " has no data or a bad size";
      return false;
    }
    int modes[2] = {t.wrap_s, t.wrap_t};
    for (int m : modes)
      if (m != MCR_WRAP_REPEAT && m != MCR_WRAP_CLAMP && m != MCR_WRAP_MIRROR) {
        why = "texture " + std::to_string(i) + " has an unknown wrap mode";
        return false;
      }
  }
  for (uint32_t i = 0; i < s->environment_count; i++) {
    const mcr_environment &e = s->environments[i];
    bool ok = e.kind == MCR_ENV_EQUIRECT ? (e.texture >= 0 && tex_ok(e.texture)) : e.kind == MCR_ENV_CUBE;
    if (e.kind == MCR_ENV_CUBE)
      for (int f = 0; f < 6; f++) ok = ok && e.faces[f] >= 0 && tex_ok(e.faces[f]);
    if (!ok) {
      why = "environment " + std::to_string(i) + " names a missing texture";
      return false;
    }
  }
  if (!env_ok(s->environment)) {
    why = "scene environment out of range";
    return false;
  }
  if (s->background_kind == MCR_BACKGROUND_SCREEN && !(s->background_texture >= 0 && tex_ok(s->background_texture))) {
    why = "screen background names a missing texture";
    return false;
  }
  if (s->background_kind == MCR_BACKGROUND_ENVIRONMENT &&
      !(s->background_environment >= 0 && env_ok(s->background_environment))) {
    why = "environment background names a missing environment";
    return false;
  }
  if (s->background_kind < MCR_BACKGROUND_CLEAR || s->background_kind > MCR_BACKGROUND_ENVIRONMENT) {
    why = "unknown background kind";
    return false;
  }
  for (uint32_t i = 0; i < s->material_count; i++) {
    const mcr_material &m = s->materials[i];
    if (!tex_ok(m.map) || !tex_ok(m.alpha_map) || !tex_ok(m.emissive_map) || !tex_ok(m.normal_map) ||
        !tex_ok(m.ao_map) || !tex_ok(m.roughness_map) || !tex_ok(m.metalness_map) || !env_ok(m.env_map)) {
      why = "material " + std::to_string(i) + " names a missing texture or environment";
      return false;
    }
    if (m.model < 0 || m.model > 4) {
      why = "material " + std::to_string(i) + " has an unknown model";
      return false;
    }
  }
  for (uint32_t i = 0; i < s->light_count; i++) {
    int t = s->lights[i].type;
    if (t < 0 || t > 4) {
      why = "light " + std::to_string(i) + " has an unknown type";
      return false;
    }
  }
  // One light slot carries the environment-diffuse ray when the scene has environments.
  uint32_t cap = s->environment_count ? MCR_MAX_LIGHTS - 1 : MCR_MAX_LIGHTS;
  if (s->light_count > cap) {
    why = "more than " + std::to_string(cap) + " lights";
    return false;
  }
  return true;
}

// Child reference of node `node` in the device layout.
int32_t child_ref(const int32_t *nodes, uint32_t node, const std::vector<int32_t> &inner_index,
                  std::vector<int2> &longleaf) {
  int32_t first = nodes[node * 3], count = nodes[node * 3 + 1];
  if (count == 0) return inner_index[node];
  if (count <= 31 && (uint32_t)first < MCR_MAX_TRIANGLES)
    return (int32_t)(MCR_LEAF_BIT | ((uint32_t)count << MCR_LEAF_FIRST_BITS) | (uint32_t)first);
  uint32_t id = (uint32_t)longleaf.size();
  longleaf.push_back(make_int2(first, count));
  return (int32_t)(MCR_LEAF_BIT | id);
}

}  // namespace

// =================================================================================================
// ---- scratch of the entry point that allocates nothing (mcr_intersect_closest_noalloc) ----------
// The closest-hit probe kernel's scratch for a fixed grid of NOALLOC_GRID blocks of 256 threads:
// the spilled part of every thread's traversal stack (MCR_STACK_MAX - MCR_STACK_LDS entries),
// the ray counters (MCR_COUNTER_COPIES lines) and the interleaved work counters. Each GPU the
// library runs on holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress gives on that device.
const uint32_t NOALLOC_GRID = 512;
const size_t NOALLOC_SPILL_BYTES = sizeof(int32_t) * (size_t)(MCR_STACK_MAX - MCR_STACK_LDS) * NOALLOC_GRID * 256;
const size_t NOALLOC_BYTES = NOALLOC_SPILL_BYTES + COUNTER_BYTES + FETCH_BYTES;
static_assert(NOALLOC_BYTES == MCR_NOALLOC_SCRATCH_BYTES, "mcr.h states the scratch size");
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
const int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Clears the counter words of the scratch (a fill kernel, not a memset, so the call can be
// captured into a graph).
__global__ void k_noalloc_clear(uint32_t *words, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0;
}
}  // namespace

// Once per device, with that device current: the device address of its copy of `scratch`.
static int noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(MCR_E_INVALID, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return MCR_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MCR_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_noalloc_ptr[device] = p;
  return MCR_OK;
}

// ---- server holds on scenes ---------------------------------------------------------------------
static std::mutex g_scene_refs_mu;
static void scene_free(mcr_scene *s);

void scene_hold(mcr_scene *s) {
  std::lock_guard<std::mutex> lk(g_scene_refs_mu);
  s->server_refs++;
}

void scene_drop(mcr_scene *s) {
  bool free_now = false;
  {
    std::lock_guard<std::mutex> lk(g_scene_refs_mu);
    s->server_refs--;
    free_now = s->server_refs == 0 && s->release_pending;
  }
  if (free_now) scene_free(s);
}

bool scene_held(const mcr_scene *s) {
  std::lock_guard<std::mutex> lk(g_scene_refs_mu);
  return s->server_refs > 0;
}

int scene_held_error() {
  return fail(MCR_E_INVALID, "the scene is held by a server; mcr_server_remove_scene it first");
}

extern "C" {

const char *mcr_last_error(void) { return g_error.c_str(); }
const char *mcr_version(void) { return "meercadraytracer_amd 0.1 (gfx950)"; }

int mcr_create(int device, mcr_ctx **out) {
  if (!out) return fail(MCR_E_INVALID, "out is null");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(MCR_E_INVALID, "no such device");
  HIP_TRY(hipSetDevice(device));
  mcr_ctx *c = new mcr_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(MCR_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  float lut[512];
  for (int i = 0; i < 512; i++) lut[i] = half_to_float_js(MCR_DFG_LUT_HALF[i]);
  e = mcr_upload_dfg(lut);
  if (e == hipSuccess) {
    float srgb[256];  // color.ts:21-25 with the portable pow
    for (int i = 0; i < 256; i++) {
      double v = i / 255.0;
      srgb[i] = (float)(v < 0.04045 ? v * 0.0773993808 : jm_pow(v * 0.9478672986 + 0.0521327014, 2.4));
    }
    e = mcr_upload_srgb8(srgb);
  }
  if (e != hipSuccess) {
    mcr_destroy(c);
    return fail(MCR_E_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
  }
  int rc = engine_init(c->eng, device, c->stream);
  if (rc == MCR_OK) rc = noalloc_setup(device);
  if (rc) {
    std::string msg = g_error;
    mcr_destroy(c);
    return fail(rc, msg);
  }
  *out = c;
  return MCR_OK;
}

void mcr_destroy(mcr_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  engine_free(c->eng);
  hipFree(c->out_tmp);
  builder_free(c->builder);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int mcr_set_wavefront_capacity(mcr_ctx *c, uint32_t slots) {
  if (!c) return fail(MCR_E_INVALID, "ctx is null");
  if (slots < 256 || slots > (1u << 26)) return fail(MCR_E_INVALID, "capacity must be in [256, 2^26]");
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  Engine &e = c->eng;
  e.capacity_limit = slots;
  if (e.capacity > slots) {
    engine_drop_graphs(e);
    hipFree(e.w_block);
    e.w_block = nullptr;
    e.capacity = 0;
    e.lights = 0;
  }
  return MCR_OK;
}

int mcr_set_timing(mcr_ctx *c, int enabled) {
  if (!c) return fail(MCR_E_INVALID, "ctx is null");
  c->eng.timing = enabled != 0;
  return MCR_OK;
}

int mcr_set_graph(mcr_ctx *c, int enabled) {
  if (!c) return fail(MCR_E_INVALID, "ctx is null");
  std::lock_guard<std::mutex> lock(c->mu);
  c->eng.use_graph = enabled != 0;
  return MCR_OK;
}

int mcr_set_shade_sort(mcr_ctx *c, int enabled) {
  if (!c) return fail(MCR_E_INVALID, "ctx is null");
  std::lock_guard<std::mutex> lock(c->mu);
  c->eng.sort_shade = enabled != 0;
  return MCR_OK;
}

int mcr_last_timing(mcr_ctx *c, mcr_timing *out) {
  if (!c || !out) return fail(MCR_E_INVALID, "null argument");
  *out = c->eng.last_timing;
  return MCR_OK;
}

// ---- scenes -----------------------------------------------------------------------------------
int mcr_scene_create(mcr_ctx *c, mcr_scene **out) {
  if (!c || !out) return fail(MCR_E_INVALID, "null argument");
  mcr_scene *s = new mcr_scene();
  s->ctx = c;
  s->d.root_ref = -1;
  *out = s;
  return MCR_OK;
}

void mcr_scene_release(mcr_scene *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(g_scene_refs_mu);
    if (s->server_refs > 0) {  // freed when the last server holding it lets it go (scene_drop)
      s->release_pending = true;
      return;
    }
  }
  scene_free(s);
}

static void scene_free(mcr_scene *s) {
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  free_geometry(s);
  hipFree(s->materials);
  hipFree(s->lights);
  hipFree(s->texels);
  hipFree(s->textures);
  hipFree(s->environments);
  delete s;
}

uint64_t mcr_scene_bytes(const mcr_scene *s) { return s ? s->bytes_geometry + s->bytes_shading : 0; }

int mcr_scene_set_geometry(mcr_scene *s, const mcr_geometry *g, uint32_t version) {
  if (!s || !g) return fail(MCR_E_INVALID, "null argument");
  if (scene_held(s)) return scene_held_error();
  uint32_t n = g->triangle_count, nodes = g->node_count;
  if (n >= MCR_MAX_TRIANGLES) return fail(MCR_E_OVERFLOW, "too many triangles for this build (2^26 max)");
  if (n > 0 && (!g->triangles || !g->normals || !g->materials || !g->flags || !g->sides || !g->bvh_bounds ||
                !g->bvh_nodes || nodes == 0))
    return fail(MCR_E_INVALID, "geometry arrays missing");
  if (nodes > 0 && n == 0) return fail(MCR_E_INVALID, "nodes without triangles");
  mcr_ctx *c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_geometry(s);
  // Triangle records.
  std::vector<float4> tri((size_t)n * 3);
  for (uint32_t k = 0; k < n; k++) {
    const float *t = g->triangles + (size_t)k * 9;
    uint32_t word = (uint32_t)g->sides[k] | ((uint32_t)g->flags[k] << 8);
    uint32_t mat = g->materials[k];
    float wf, mf;
    memcpy(&wf, &word, 4);
    memcpy(&mf, &mat, 4);
    tri[k * 3 + 0] = make_float4(t[0], t[1], t[2], t[3]);
    tri[k * 3 + 1] = make_float4(t[4], t[5], t[6], t[7]);
    tri[k * 3 + 2] = make_float4(t[8], wf, mf, 0.0f);
  }
  // Inner records: inner nodes numbered in node order.
  std::vector<int32_t> inner_index(nodes, -1);
  uint32_t ninner = 0;
  for (uint32_t i = 0; i < nodes; i++) {
    int32_t cnt = g->bvh_nodes[i * 3 + 1];
    if (cnt < 0) return fail(MCR_E_INVALID, "negative leaf count");
    if (cnt == 0) {
      int32_t left = g->bvh_nodes[i * 3];
      if (left <= (int32_t)i || (uint32_t)left + 1 >= nodes) return fail(MCR_E_INVALID, "bad child index");
      int32_t axis = g->bvh_nodes[i * 3 + 2];
      if (axis < 0 || axis > 2) return fail(MCR_E_INVALID, "bad split axis");
      inner_index[i] = (int32_t)ninner++;
    } else if ((uint32_t)g->bvh_nodes[i * 3] + (uint32_t)cnt > n) {
      return fail(MCR_E_INVALID, "leaf range out of bounds");
    }
  }
  // Every node reached exactly once from the root, and no deeper than the traversal stack holds
  // (a tree of depth d needs at most d stack entries; buildBvh stops at depth 100).
  {
    std::vector<int32_t> depth(nodes, -1);
    if (nodes) depth[0] = 0;
    for (uint32_t i = 0; i < nodes; i++) {
      if (depth[i] < 0) return fail(MCR_E_INVALID, "BVH node " + std::to_string(i) + " is unreachable");
      if (g->bvh_nodes[i * 3 + 1] != 0) continue;
      for (int c = 0; c < 2; c++) {
        uint32_t child = (uint32_t)g->bvh_nodes[i * 3] + c;
        if (depth[child] >= 0) return fail(MCR_E_INVALID, "BVH node " + std::to_string(child) + " has two parents");
        depth[child] = depth[i] + 1;
        if (depth[child] >= MCR_STACK_MAX - 1) return fail(MCR_E_OVERFLOW, "BVH deeper than the traversal stack");
      }
    }
  }
  std::vector<float4> inner((size_t)ninner * 4);
  std::vector<int2> longleaf;
  for (uint32_t i = 0; i < nodes; i++) {
    if (inner_index[i] < 0) continue;
    int32_t left = g->bvh_nodes[i * 3];
    const float *bl = g->bvh_bounds + (size_t)left * 6, *br = g->bvh_bounds + (size_t)(left + 1) * 6;
    size_t r = (size_t)inner_index[i] * 4;
    inner[r + 0] = make_float4(bl[0], bl[1], bl[2], bl[3]);
    inner[r + 1] = make_float4(bl[4], bl[5], br[0], br[1]);
    inner[r + 2] = make_float4(br[2], br[3], br[4], br[5]);
    int32_t refs[4] = {child_ref(g->bvh_nodes, left, inner_index, longleaf),
                       child_ref(g->bvh_nodes, left + 1, inner_index, longleaf), g->bvh_nodes[i * 3 + 2], 0};
    memcpy(&inner[r + 3], refs, 16);
  }
  DevScene &d = s->d;
  d.triangle_count = n;
  d.inner_count = ninner;
  d.epsilon = g->epsilon;
  d.root_ref = nodes ? child_ref(g->bvh_nodes, 0, inner_index, longleaf) : -1;
  for (int k = 0; k < 6; k++) d.root_box[k] = nodes ? g->bvh_bounds[k] : 0.0f;
  HIP_TRY(dalloc(&s->tri, tri.size()));
  HIP_TRY(dalloc(&s->inner, inner.size()));
  HIP_TRY(dalloc(&s->longleaf, longleaf.size()));
  HIP_TRY(dalloc(&s->normals, (size_t)n * 9));
  HIP_TRY(hipMemcpy(s->tri, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice));
  if (!inner.empty()) HIP_TRY(hipMemcpy(s->inner, inner.data(), inner.size() * sizeof(float4), hipMemcpyHostToDevice));
  if (!longleaf.empty())
    HIP_TRY(hipMemcpy(s->longleaf, longleaf.data(), longleaf.size() * sizeof(int2), hipMemcpyHostToDevice));
  if (n) HIP_TRY(hipMemcpy(s->normals, g->normals, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice));
  s->colors = nullptr;
  if (g->colors && n) {
    HIP_TRY(dalloc(&s->colors, (size_t)n * 12));
    HIP_TRY(hipMemcpy(s->colors, g->colors, (size_t)n * 12 * sizeof(float), hipMemcpyHostToDevice));
  }
  // uvs and tangents, read only by textured materials (tracer.ts:151-209)
  HIP_TRY(dalloc(&s->uvs, (size_t)n * 6));
  HIP_TRY(dalloc(&s->tangents, (size_t)n * 4));
  if (n && g->uvs) HIP_TRY(hipMemcpy(s->uvs, g->uvs, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice));
  else if (n) HIP_TRY(hipMemset(s->uvs, 0, (size_t)n * 6 * sizeof(float)));
  if (n && g->tangents)
    HIP_TRY(hipMemcpy(s->tangents, g->tangents, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
  else if (n) HIP_TRY(hipMemset(s->tangents, 0, (size_t)n * 4 * sizeof(float)));
  d.tri = s->tri;
  d.inner = s->inner;
  d.longleaf = s->longleaf;
  d.normals = s->normals;
  d.colors = s->colors;
  d.uvs = s->uvs;
  d.tangents = s->tangents;
  s->bytes_geometry = tri.size() * sizeof(float4) + inner.size() * sizeof(float4) +
                      longleaf.size() * sizeof(int2) + (size_t)n * (9 + 6 + 4) * 4 +
                      (s->colors ? (size_t)n * 48 : 0);
  s->geometry_version = version;
  s->has_geometry = true;
  return MCR_OK;
}

int mcr_scene_set_shading(mcr_scene *s, const mcr_shading *sh, uint32_t version) {
  if (!s || !sh) return fail(MCR_E_INVALID, "null argument");
  if (scene_held(s)) return scene_held_error();
  std::string why;
  if (!shading_supported(sh, why)) return fail(MCR_E_UNSUPPORTED, why);
  mcr_ctx *c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<DevMaterial> mats(std::max<uint32_t>(1, sh->material_count));
  memset(mats.data(), 0, mats.size() * sizeof(DevMaterial));
  for (uint32_t i = 0; i < sh->material_count; i++) {
    const mcr_material &m = sh->materials[i];
    DevMaterial &d = mats[i];
    d.model = m.model;
    d.side = m.side;
    d.transparent = m.transparent;
    d.vertex_colors = m.vertex_colors;
    d.tone_mapped = m.tone_mapped;
    d.physical = m.physical;
    for (int k = 0; k < 3; k++) {
      d.color[k] = m.color[k];
      d.emissive[k] = m.emissive[k];
      d.specular_color[k] = m.specular_color[k];
      d.attenuation[k] = m.attenuation[k];
      d.specular[k] = m.specular[k];
    }
    d.opacity = m.opacity;
    d.alpha_test = m.alpha_test;
    d.roughness = m.roughness;
    d.metalness = m.metalness;
    d.ior = m.ior;
    d.specular_intensity = m.specular_intensity;
    d.transmission = m.transmission;
    d.thickness = m.thickness;
    d.shininess = m.shininess;
    d.map = m.map;
    d.alpha_map = m.alpha_map;
    d.emissive_map = m.emissive_map;
    d.normal_map = m.normal_map;
    d.ao_map = m.ao_map;
    d.roughness_map = m.roughness_map;
    d.metalness_map = m.metalness_map;
    d.env_map = m.env_map;
    d.combine = m.combine;
    d.textured = m.map >= 0 || m.alpha_map >= 0 || m.emissive_map >= 0 || m.normal_map >= 0 || m.ao_map >= 0 ||
                 m.roughness_map >= 0 || m.metalness_map >= 0;
    d.normal_scale[0] = m.normal_scale[0];
    d.normal_scale[1] = m.normal_scale[1];
    d.ao_map_intensity = m.ao_map_intensity;
    d.env_map_intensity = m.env_map_intensity;
    d.reflectivity = m.reflectivity;
  }
  // Textures: every texel array in one allocation, each 256-byte aligned.
  std::vector<DevTexture> texs(std::max<uint32_t>(1, sh->texture_count));
  memset(texs.data(), 0, texs.size() * sizeof(DevTexture));
  std::vector<size_t> toff(sh->texture_count);
  size_t tbytes = 0;
  for (uint32_t i = 0; i < sh->texture_count; i++) {
    const mcr_texture &t = sh->textures[i];
    toff[i] = tbytes;
    tbytes += ((size_t)t.width * t.height * (t.is_float ? 16 : 4) + 255) / 256 * 256;
  }
  std::vector<DevEnvironment> envs(std::max<uint32_t>(1, sh->environment_count));
  memset(envs.data(), 0, envs.size() * sizeof(DevEnvironment));
  for (uint32_t i = 0; i < sh->environment_count; i++) {
    const mcr_environment &e = sh->environments[i];
    DevEnvironment &d = envs[i];
    d.kind = e.kind;
    d.texture = e.texture;
    for (int k = 0; k < 6; k++) d.faces[k] = e.faces[k];
    d.flip_x = e.flip_x;
    for (int k = 0; k < 9; k++) d.rotation[k] = e.rotation[k];
    d.intensity = e.intensity;
  }
  std::vector<DevLight> lights(std::max<uint32_t>(1, sh->light_count));
  memset(lights.data(), 0, lights.size() * sizeof(DevLight));
  for (uint32_t i = 0; i < sh->light_count; i++) {
    const mcr_light &l = sh->lights[i];
    DevLight &d = lights[i];
    d.type = l.type;
    d.cast_shadow = l.cast_shadow;
    for (int k = 0; k < 3; k++) {
      d.color[k] = l.color[k];
      d.ground_color[k] = l.ground_color[k];
      d.direction[k] = l.direction[k];
      d.position[k] = l.position[k];
    }
    d.distance = l.distance;
    d.decay = l.decay;
    d.cone_cos = l.cone_cos;
    d.penumbra_cos = l.penumbra_cos;
  }
  hipFree(s->materials);
  hipFree(s->lights);
  hipFree(s->texels);
  hipFree(s->textures);
  hipFree(s->environments);
  s->materials = nullptr;
  s->lights = nullptr;
  s->texels = nullptr;
  s->textures = nullptr;
  s->environments = nullptr;
  HIP_TRY(dalloc(&s->materials, mats.size()));
  HIP_TRY(dalloc(&s->lights, lights.size()));
  HIP_TRY(hipMalloc(&s->texels, std::max<size_t>(tbytes, 256)));
  for (uint32_t i = 0; i < sh->texture_count; i++) {
    const mcr_texture &t = sh->textures[i];
    DevTexture &d = texs[i];
    size_t b = (size_t)t.width * t.height * (t.is_float ? 16 : 4);
    HIP_TRY(hipMemcpy((char *)s->texels + toff[i], t.data, b, hipMemcpyHostToDevice));
    d.data = (char *)s->texels + toff[i];
    d.width = (int32_t)t.width;
    d.height = (int32_t)t.height;
    d.is_float = t.is_float ? 1 : 0;
    d.srgb = t.srgb && !t.is_float ? 1 : 0;
    d.flip_y = t.flip_y ? 1 : 0;
    d.wrap_s = t.wrap_s;
    d.wrap_t = t.wrap_t;
    for (int k = 0; k < 9; k++) d.matrix[k] = t.matrix[k];
  }
  HIP_TRY(dalloc(&s->textures, texs.size()));
  HIP_TRY(dalloc(&s->environments, envs.size()));
  HIP_TRY(hipMemcpy(s->textures, texs.data(), texs.size() * sizeof(DevTexture), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->environments, envs.data(), envs.size() * sizeof(DevEnvironment), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->materials, mats.data(), mats.size() * sizeof(DevMaterial), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->lights, lights.data(), lights.size() * sizeof(DevLight), hipMemcpyHostToDevice));
  DevScene &d = s->d;
  d.materials = s->materials;
  d.lights = s->lights;
  d.textures = s->textures;
  d.environments = s->environments;
  d.texture_count = sh->texture_count;
  s->shadow_tex = false;
  for (uint32_t i = 0; i < sh->material_count; i++)
    s->shadow_tex |= sh->materials[i].map >= 0 || sh->materials[i].alpha_map >= 0;
  d.environment_count = sh->environment_count;
  d.environment = sh->environment;
  d.background_texture = sh->background_texture;
  d.background_environment = sh->background_environment;
  d.background_intensity = sh->background_intensity;
  d.material_count = sh->material_count;
  d.light_count = sh->light_count;
  d.background_kind = sh->background_kind;
  d.fog_kind = sh->fog_kind;
  for (int k = 0; k < 3; k++) {
    d.background_color[k] = sh->background_color[k];
    d.fog_color[k] = sh->fog_color[k];
  }
  d.fog_near = sh->fog_near;
  d.fog_far = sh->fog_far;
  d.fog_density = sh->fog_density;
  s->bytes_shading = mats.size() * sizeof(DevMaterial) + lights.size() * sizeof(DevLight) + tbytes +
                     texs.size() * sizeof(DevTexture) + envs.size() * sizeof(DevEnvironment);
  s->shading_version = version;
  s->has_shading = true;
  return MCR_OK;
}

}  // extern "C"

// ---- rendering --------------------------------------------------------------------------------
ViewDev make_view(const mcr_scene *s, const mcr_camera &cam, const mcr_settings &st) {
  ViewDev v;
  memset(&v, 0, sizeof(v));
  v.sc = s->d;
  v.wide = s->wide;
  v.cam = cam;
  v.st = st;
  v.light_slots = s->d.light_count;  // at most MCR_MAX_LIGHTS (checked on upload)
  v.shadow_slots = v.light_slots + (s->d.environment_count ? 1 : 0);
  v.shadow_tex = s->shadow_tex ? 1 : 0;
  return v;
}

int64_t layout_job(const mcr_tile_job &j, uint32_t index, uint32_t view, uint64_t start, JobDev &out) {
  uint32_t block = j.block < 1 ? 1 : j.block;
  if (j.tile.width <= 0 || j.tile.height <= 0 || j.tile.x < 0 || j.tile.y < 0 ||
      (uint64_t)j.tile.x + j.tile.width > j.width || (uint64_t)j.tile.y + j.tile.height > j.height) {
    mcr_fail(MCR_E_INVALID, "tile " + std::to_string(index) + " lies outside its image");
    return -1;
  }
  if (j.out_offset % 4) {
    mcr_fail(MCR_E_INVALID, "out_offset must be a multiple of 4 floats");
    return -1;
  }
  uint32_t cols = (j.tile.width + block - 1) / block, rows = (j.tile.height + block - 1) / block;
  out = JobDev{j.tile.x, j.tile.y, j.tile.width, j.tile.height, j.width, j.height, j.pass, block,
               j.out_offset, (uint32_t)start, cols, view, rows};
  return (int64_t)cols * rows;
}

namespace {

// Views and jobs of one call, laid out in sample order. Sample indices are 31-bit, so a call
// with more samples runs as consecutive wavefront runs over consecutive groups of jobs.
int render_views(mcr_ctx *c, const mcr_view *views, uint32_t nviews, const mcr_tile_job *jobs,
                 const uint32_t *job_views, uint32_t njobs, float *dout, mcr_counters *counters) {
  std::vector<ViewDev> vd(nviews);
  for (uint32_t v = 0; v < nviews; v++) vd[v] = make_view(views[v].scene, *views[v].camera, *views[v].settings);
  c->eng.wide = c->intersection == MCR_ISECT_WIDE_FP32;
  if (c->eng.wide)
    for (uint32_t v = 0; v < nviews; v++)
      if (!views[v].scene->has_wide)
        return fail(MCR_E_NO_SCENE, "the scene has no wide BVH (mcr_scene_build_wide)");
  // MCR_RUN_SAMPLES lowers the run size (tests use it to exercise the split on small calls).
  const char *env = getenv("MCR_RUN_SAMPLES");
  const uint64_t RUN_LIMIT = env ? std::max<uint64_t>(1, strtoull(env, nullptr, 10)) : 1ull << 30;
  std::vector<JobDev> jd;
  uint64_t total = 0;
  for (uint32_t i = 0; i < njobs; i++) {
    uint32_t view = job_views ? job_views[i] : 0;
    if (view >= nviews) return fail(MCR_E_INVALID, "job " + std::to_string(i) + " names no view");
    JobDev j;
    int64_t n = layout_job(jobs[i], i, view, total, j);
    if (n < 0) return MCR_E_INVALID;
    if (total + (uint64_t)n > RUN_LIMIT && !jd.empty()) {
      MCR_TRY(engine_run(c->eng, vd, jd, (uint32_t)total, dout, counters));
      jd.clear();
      total = 0;
      layout_job(jobs[i], i, view, 0, j);
    }
    if ((uint64_t)n >= RUN_LIMIT) return fail(MCR_E_OVERFLOW, "job " + std::to_string(i) + " has too many samples");
    jd.push_back(j);
    total += (uint64_t)n;
  }
  return engine_run(c->eng, vd, jd, (uint32_t)total, dout, counters);
}

int check_view(const mcr_view &v, uint32_t i) {
  if (!v.scene || !v.camera || !v.settings) return fail(MCR_E_INVALID, "view " + std::to_string(i) + " has a null field");
  if (!v.scene->has_geometry || !v.scene->has_shading)
    return fail(MCR_E_NO_SCENE, "the tile refers to a scene this worker has not received");
  return MCR_OK;
}

int render_common(mcr_ctx *c, const mcr_view *views, uint32_t nviews, const mcr_tile_job *jobs,
                  const uint32_t *job_views, uint32_t njobs, float *out, int out_device,
                  mcr_counters *counters) {
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  uint64_t floats = 0;
  for (uint32_t i = 0; i < njobs; i++)
    floats = std::max<uint64_t>(floats, jobs[i].out_offset + (uint64_t)jobs[i].tile.width * jobs[i].tile.height * 4);
  float *dout = out;
  if (!out_device) {
    MCR_TRY(ensure_out_tmp(c, floats));
    dout = c->out_tmp;
  }
  MCR_TRY(render_views(c, views, nviews, jobs, job_views, njobs, dout, counters));
  if (!out_device && floats)
    HIP_TRY(hipMemcpyAsync(out, dout, floats * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCR_OK;
}

}  // namespace

extern "C" {

int mcr_set_intersection(mcr_ctx *c, int mode) {
  if (!c) return fail(MCR_E_INVALID, "null argument");
  if (mode != MCR_ISECT_FP64 && mode != MCR_ISECT_WIDE_FP32) return fail(MCR_E_INVALID, "unknown intersection mode");
  std::lock_guard<std::mutex> lock(c->mu);
  c->intersection = mode;
  return MCR_OK;
}

int mcr_render_tiles(mcr_ctx *c, mcr_scene *s, const mcr_camera *camera, const mcr_settings *st,
                     const mcr_tile_job *jobs, uint32_t njobs, float *out, int out_device,
                     mcr_counters *counters) {
  if (!c || !s || !camera || !st || (!jobs && njobs) || !out) return fail(MCR_E_INVALID, "null argument");
  mcr_view v{s, camera, st};
  MCR_TRY(check_view(v, 0));
  return render_common(c, &v, 1, jobs, nullptr, njobs, out, out_device, counters);
}

int mcr_render_batch(mcr_ctx *c, const mcr_view *views, uint32_t nviews, const mcr_tile_job *jobs,
                     const uint32_t *job_views, uint32_t njobs, float *out, int out_device,
                     mcr_counters *counters) {
  if (!c || (!views && nviews) || (!jobs && njobs) || !out) return fail(MCR_E_INVALID, "null argument");
  if (njobs && !nviews) return fail(MCR_E_INVALID, "jobs without views");
  for (uint32_t i = 0; i < nviews; i++) MCR_TRY(check_view(views[i], i));
  return render_common(c, views, nviews, jobs, job_views, njobs, out, out_device, counters);
}

int mcr_render_tile(mcr_ctx *c, mcr_scene *s, uint32_t geometry_version, uint32_t shading_version,
                    const mcr_camera *camera, const mcr_settings *settings, mcr_tile tile,
                    uint32_t width, uint32_t height, uint32_t pass, uint32_t block, float *rgba_out,
                    mcr_counters *counters) {
  if (!c || !s) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_geometry || !s->has_shading || s->geometry_version != geometry_version ||
      s->shading_version != shading_version)
    return fail(MCR_E_NO_SCENE, "the tile refers to a scene this worker has not received");
  mcr_tile_job j{tile, width, height, pass, block, 0};
  return mcr_render_tiles(c, s, camera, settings, &j, 1, rgba_out, 0, counters);
}

int mcr_encode_tile(mcr_ctx *c, uint8_t *image, uint32_t width, mcr_tile tile, const float *data,
                    float *sum, uint32_t count, int srgb, int device_ptrs) {
  if (!c || !image || !data) return fail(MCR_E_INVALID, "null argument");
  if (!device_ptrs) return fail(MCR_E_INVALID, "mcr_encode_tile works on device buffers");
  if (tile.width <= 0 || tile.height <= 0) return MCR_OK;
  if (count < 1) return fail(MCR_E_INVALID, "count must be at least 1");
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  launch_encode(c->stream, image, width, tile, data, sum, nullptr, count, srgb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCR_OK;
}

int mcr_encode_tile64(mcr_ctx *c, uint8_t *image, uint32_t width, mcr_tile tile, const float *data,
                      double *sum, uint32_t count, int srgb) {
  if (!c || !image || !data || !sum) return fail(MCR_E_INVALID, "null argument");
  if (tile.width <= 0 || tile.height <= 0) return MCR_OK;
  if (count < 1) return fail(MCR_E_INVALID, "count must be at least 1");
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  launch_encode(c->stream, image, width, tile, data, nullptr, sum, count, srgb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCR_OK;
}

// ---- ray queries ------------------------------------------------------------------------------
int mcr_intersect_closest(mcr_ctx *c, mcr_scene *s, const double *rays, uint32_t n, double t_min,
                          double t_max, int cull, int32_t *tri_out, double *t_out, double *u_out,
                          double *v_out, uint8_t *front_out, int device_ptrs, mcr_counters *counters) {
  if (!c || !s || (!rays && n)) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_geometry) return fail(MCR_E_NO_SCENE, "the scene has no geometry");
  if (n == 0) return MCR_OK;
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  const double *drays = rays;
  int32_t *dtri = tri_out;
  double *dt = t_out, *du = u_out, *dv = v_out;
  uint8_t *dfront = front_out;
  void *tmp = nullptr;
  if (!device_ptrs) {
    size_t bytes = (size_t)n * (48 + 4 + 24 + 1) + 1024;
    HIP_TRY(hipMalloc(&tmp, bytes));
    char *b = (char *)tmp;
    double *r = (double *)b;
    b += (size_t)n * 48;
    dt = t_out ? (double *)b : nullptr;
    b += (size_t)n * 8;
    du = u_out ? (double *)b : nullptr;
    b += (size_t)n * 8;
    dv = v_out ? (double *)b : nullptr;
    b += (size_t)n * 8;
    dtri = tri_out ? (int32_t *)b : nullptr;
    b += (size_t)n * 4;
    dfront = front_out ? (uint8_t *)b : nullptr;
    HIP_TRY(hipMemcpyAsync(r, rays, (size_t)n * 48, hipMemcpyHostToDevice, sm));
    drays = r;
  }
  HIP_TRY(counters_clear(c->eng.counters, sm));
  if (c->eng.timing) HIP_TRY(hipEventRecord(c->eng.ev[0], sm));
  HIP_TRY(hipMemsetAsync(c->eng.fetch, 0, FETCH_BYTES, sm));
  launch_probe_closest(sm, s->d, drays, n, t_min, t_max, cull, dtri, dt, du, dv, dfront,
                       std::min<uint32_t>(TRAVERSE_GRID, (n + 255) / 256), c->eng.spill, TRAVERSE_THREADS,
                       c->eng.counters, c->eng.fetch);
  if (c->eng.timing) HIP_TRY(hipEventRecord(c->eng.ev[1], sm));
  HIP_TRY(hipGetLastError());
  if (!device_ptrs) {
    if (t_out) HIP_TRY(hipMemcpyAsync(t_out, dt, (size_t)n * 8, hipMemcpyDeviceToHost, sm));
    if (u_out) HIP_TRY(hipMemcpyAsync(u_out, du, (size_t)n * 8, hipMemcpyDeviceToHost, sm));
    if (v_out) HIP_TRY(hipMemcpyAsync(v_out, dv, (size_t)n * 8, hipMemcpyDeviceToHost, sm));
    if (tri_out) HIP_TRY(hipMemcpyAsync(tri_out, dtri, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
    if (front_out) HIP_TRY(hipMemcpyAsync(front_out, dfront, n, hipMemcpyDeviceToHost, sm));
  }
  HIP_TRY(hipStreamSynchronize(sm));
  if (c->eng.timing) {
    float a;
    hipEventElapsedTime(&a, c->eng.ev[0], c->eng.ev[1]);
    c->eng.last_timing = mcr_timing{a, 0, 0, 0, a, 1, 0};
  }
  if (tmp) hipFree(tmp);
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

int mcr_intersect_any_count(mcr_ctx *c, mcr_scene *s, const double *rays, uint32_t n, double t_min,
                            double t_max, uint32_t *count_out, int device_ptrs, mcr_counters *counters) {
  if (!c || !s || (!rays && n) || (!count_out && n)) return fail(MCR_E_INVALID, "null argument");
  if (!s->has_geometry) return fail(MCR_E_NO_SCENE, "the scene has no geometry");
  if (n == 0) return MCR_OK;
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t sm = c->stream;
  const double *drays = rays;
  uint32_t *dcount = count_out;
  void *tmp = nullptr;
  if (!device_ptrs) {
    HIP_TRY(hipMalloc(&tmp, (size_t)n * 52 + 256));
    double *r = (double *)tmp;
    dcount = (uint32_t *)((char *)tmp + (size_t)n * 48);
    HIP_TRY(hipMemcpyAsync(r, rays, (size_t)n * 48, hipMemcpyHostToDevice, sm));
    drays = r;
  }
  HIP_TRY(counters_clear(c->eng.counters, sm));
  HIP_TRY(hipMemsetAsync(c->eng.fetch, 0, FETCH_BYTES, sm));
  launch_probe_any_count(sm, s->d, drays, n, t_min, t_max, dcount,
                         std::min<uint32_t>(TRAVERSE_GRID, (n + 255) / 256), c->eng.spill, TRAVERSE_THREADS,
                         c->eng.counters, c->eng.fetch);
  HIP_TRY(hipGetLastError());
  if (!device_ptrs) HIP_TRY(hipMemcpyAsync(count_out, dcount, (size_t)n * 4, hipMemcpyDeviceToHost, sm));
  HIP_TRY(hipStreamSynchronize(sm));
  if (tmp) hipFree(tmp);
  if (counters) {
    DevCounters hc;
    HIP_TRY(counters_read(c->eng.counters, hc));
    counters->any_queries += hc.any_queries;
    counters->node_visits += hc.node_visits;
    counters->leaf_visits += hc.leaf_visits;
    counters->triangle_tests += hc.triangle_tests;
  }
  return MCR_OK;
}

// As mcr_intersect_closest on device pointers, with no allocation and no wait (see mcr.h).
int mcr_intersect_closest_noalloc(mcr_scene *s, const double *rays, uint32_t n, double t_min, double t_max,
                                  int cull, int32_t *tri_out, double *t_out, double *u_out, double *v_out,
                                  uint8_t *front_out, void *stream) {
  if (!s || !s->ctx || (!rays && n)) return fail(MCR_E_INVALID, "null argument");
  if (n > MCR_NOALLOC_MAX_RAYS) return fail(MCR_E_INVALID, "more than 2^31 rays in one call");
  if (!s->has_geometry) return fail(MCR_E_NO_SCENE, "the scene has no geometry");
  if (n == 0) return MCR_OK;
  int device = s->ctx->device;
  HIP_TRY(hipSetDevice(device));
  void *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    p = g_noalloc_ptr[device];
  }
  if (!p) {
    MCR_TRY(noalloc_setup(device));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    p = g_noalloc_ptr[device];
  }
  hipStream_t sm = (hipStream_t)stream;
  int32_t *spill = (int32_t *)p;
  DevCounters *cnt = (DevCounters *)((char *)p + NOALLOC_SPILL_BYTES);
  uint32_t *fetch = (uint32_t *)((char *)p + NOALLOC_SPILL_BYTES + COUNTER_BYTES);
  // The counters start at zero on every call: the previous call on this GPU left its totals.
  hipLaunchKernelGGL(k_noalloc_clear, dim3(4), dim3(256), 0, sm, (uint32_t *)cnt,
                     (uint32_t)((COUNTER_BYTES + FETCH_BYTES) / sizeof(uint32_t)));
  launch_probe_closest(sm, s->d, rays, n, t_min, t_max, cull, tri_out, t_out, u_out, v_out, front_out,
                       std::min<uint32_t>(NOALLOC_GRID, (n + 255) / 256), spill, NOALLOC_GRID * 256, cnt, fetch);
  HIP_TRY(hipGetLastError());
  return MCR_OK;
}

// ---- device buffers ---------------------------------------------------------------------------
int mcr_device_alloc(mcr_ctx *c, uint64_t bytes, void **out) {
  if (!c || !out) return fail(MCR_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(out, bytes ? bytes : 1));
  return MCR_OK;
}

int mcr_device_free(mcr_ctx *c, void *ptr) {
  if (!c) return fail(MCR_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipFree(ptr));
  return MCR_OK;
}

int mcr_copy(mcr_ctx *c, void *dst, const void *src, uint64_t bytes, int kind) {
  if (!c || (!bytes)) return c ? MCR_OK : fail(MCR_E_INVALID, "null argument");
  if (!dst || !src) return fail(MCR_E_INVALID, "null argument");
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCR_OK;
}

int mcr_memset(mcr_ctx *c, void *dst, int value, uint64_t bytes) {
  if (!c || !dst) return fail(MCR_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemsetAsync(dst, value, bytes, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCR_OK;
}

// ---- builder ----------------------------------------------------------------------------------
int mcr_build_bvh_device(mcr_ctx *c, const float *d_tris, uint32_t n, float *d_bounds,
                         int32_t *d_nodes, uint32_t *d_order, uint32_t *node_count_out) {
  if (!c || !node_count_out) return fail(MCR_E_INVALID, "null argument");
  if (n == 0) {
    *node_count_out = 0;
    return MCR_OK;
  }
  if (!d_tris || !d_bounds || !d_nodes || !d_order) return fail(MCR_E_INVALID, "null argument");
  if (n >= (1u << 30)) return fail(MCR_E_OVERFLOW, "too many triangles");
  std::lock_guard<std::mutex> lock(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  std::string err;
  int rc = builder_run(c->builder, c->stream, d_tris, n, d_bounds, d_nodes, d_order, node_count_out, err);
  if (rc) return fail(rc, err);
  return MCR_OK;
}

int mcr_build_bvh(mcr_ctx *c, const float *tris, uint32_t n, float *bounds_out, int32_t *nodes_out,
                  uint32_t *order_out, uint32_t *node_count_out) {
  if (!c || !node_count_out) return fail(MCR_E_INVALID, "null argument");
  if (n == 0) {
    *node_count_out = 0;
    return MCR_OK;
  }
  if (!tris || !bounds_out || !nodes_out || !order_out) return fail(MCR_E_INVALID, "null argument");
  if (n >= (1u << 30)) return fail(MCR_E_OVERFLOW, "too many triangles");
  HIP_TRY(hipSetDevice(c->device));
  size_t maxn = 2 * (size_t)n - 1;
  float *dt = nullptr, *db = nullptr;
  int32_t *dn = nullptr;
  uint32_t *dord = nullptr;
  HIP_TRY(hipMalloc(&dt, (size_t)n * 9 * sizeof(float)));
  HIP_TRY(hipMalloc(&db, maxn * 6 * sizeof(float)));
  HIP_TRY(hipMalloc(&dn, maxn * 3 * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&dord, (size_t)n * sizeof(uint32_t)));
  HIP_TRY(hipMemcpy(dt, tris, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice));
  uint32_t nc = 0;
  int rc = mcr_build_bvh_device(c, dt, n, db, dn, dord, &nc);
  if (rc == MCR_OK) {
    HIP_TRY(hipMemcpy(bounds_out, db, (size_t)nc * 6 * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nodes_out, dn, (size_t)nc * 3 * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(order_out, dord, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    *node_count_out = nc;
  }
  hipFree(dt);
  hipFree(db);
  hipFree(dn);
  hipFree(dord);
  return rc;
}

}  // extern "C"
