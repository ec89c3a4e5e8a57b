/*
 * mcr_napi.c — the Node-API addon that binds libmcr.so for the reference's TypeScript host.
 *
 * Each export takes the reference's own plain-data objects (src/types.ts, src/protocol.ts) and
 * calls one C-ABI entry point of include/mcr.h. Typed arrays cross without copies. Exports:
 *   version()                                   -> string                  (mcr_version)
 *   compileCount(meshes)                        -> {triangleCount, anyColors}   addMesh's loop
 *                                                  counts (compile.ts:235-251), host only
 *   buildBvh(triangles, count)                  -> {bvh: {bounds, nodes, nodeCount}, order}
 *                                                  (bvh.ts:16, called at compile.ts:361)
 *   sceneCreate() / sceneRelease(scene)         a resident scene (mcr_scene_*)
 *   sceneSetGeometry(scene, geometry, version)  the 'geometry' message (protocol.ts:64-67)
 *   sceneSetShading(scene, shading, pixels, version)
 *                                               the 'shading' message with the pixels of earlier
 *                                                  'textures' messages (protocol.ts:68-74, 81-87)
 *   renderTile(scene, message)                  -> Float32Array            renderTile behind a
 *   renderTileAsync(scene, message)             -> Promise<Float32Array>   'tile' message
 *                                                  (tracer.ts:62, protocol.ts:75-105)
 * Failures throw an Error (or reject) with mcr_last_error()'s text, which carries the protocol's
 * texts where it has them ('the tile refers to a scene this worker has not received',
 * 'texture N was not sent'). The GPU context is created on the first call that needs it, so the
 * host-only exports work on a machine without a GPU.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mcr.h"

static mcr_ctx *g_ctx = NULL;

#define CHECK_NAPI(call)                                            \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, NULL, "Node-API call failed: " #call); \
      return NULL;                                                 \
    }                                                              \
  } while (0)

static napi_value throw_mcr(napi_env env) {
  napi_throw_error(env, NULL, mcr_last_error());
  return NULL;
}

static int ensure_ctx(napi_env env) {
  if (g_ctx) return 1;
  if (mcr_create(0, &g_ctx) != MCR_OK) {
    g_ctx = NULL;
    throw_mcr(env);
    return 0;
  }
  return 1;
}

/* ---- reading the reference's objects ----------------------------------------------------- */
static napi_value prop(napi_env env, napi_value obj, const char *name) {
  napi_value v = NULL;
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return NULL;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

static double num(napi_env env, napi_value obj, const char *name, double dflt) {
  napi_value v = prop(env, obj, name);
  double d;
  if (!v || napi_get_value_double(env, v, &d) != napi_ok) return dflt;
  return d;
}

static int flag(napi_env env, napi_value obj, const char *name) {
  napi_value v = prop(env, obj, name);
  bool b = false;
  if (v) napi_get_value_bool(env, v, &b);
  return b ? 1 : 0;
}

static void vec(napi_env env, napi_value obj, const char *name, double *out, uint32_t n) {
  napi_value v = prop(env, obj, name);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    double d = 0;
    if (v && napi_get_element(env, v, i, &e) == napi_ok) napi_get_value_double(env, e, &d);
    out[i] = d;
  }
}

static int str_is(napi_env env, napi_value obj, const char *name, const char *want) {
  napi_value v = prop(env, obj, name);
  char buf[32];
  size_t len = 0;
  if (!v || napi_get_value_string_utf8(env, v, buf, sizeof buf, &len) != napi_ok) return 0;
  return strcmp(buf, want) == 0;
}

/* A typed array's data (NULL when the value is missing, null or empty). */
static void *typed(napi_env env, napi_value v, size_t *count) {
  bool is = false;
  if (count) *count = 0;
  if (!v || napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  napi_typedarray_type t;
  size_t n, off;
  void *data;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok) return NULL;
  if (count) *count = n;
  return n ? data : NULL;
}

static uint32_t infinite_u32(double d) { return isinf(d) || d >= 4294967295.0 ? 0xffffffffu : (uint32_t)d; }

/* ---- exports ---------------------------------------------------------------------------- */
static napi_value Version(napi_env env, napi_callback_info info) {
  napi_value out;
  CHECK_NAPI(napi_create_string_utf8(env, mcr_version(), NAPI_AUTO_LENGTH, &out));
  return out;
}

/* meshes: [{positions: Float32Array, index: Uint32Array | null, colors: Float32Array | null,
 *           colorSize, drawStart, drawCount,
 *           groups: [{start, count, material, flags, side, flatShading, vertexColors}]}] */
static napi_value CompileCount(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  uint32_t n = 0;
  CHECK_NAPI(napi_get_array_length(env, argv[0], &n));
  mcr_mesh *meshes = calloc(n ? n : 1, sizeof(mcr_mesh));
  mcr_mesh_group **groups = calloc(n ? n : 1, sizeof(mcr_mesh_group *));
  for (uint32_t i = 0; i < n; i++) {
    napi_value m;
    napi_get_element(env, argv[0], i, &m);
    size_t np = 0, ni = 0;
    meshes[i].positions = typed(env, prop(env, m, "positions"), &np);
    meshes[i].index = typed(env, prop(env, m, "index"), &ni);
    meshes[i].vertex_count = (uint32_t)(np / 3);
    meshes[i].index_count = (uint32_t)ni;
    meshes[i].colors = typed(env, prop(env, m, "colors"), NULL);
    meshes[i].color_size = (uint32_t)num(env, m, "colorSize", 0);
    meshes[i].draw_start = (uint32_t)num(env, m, "drawStart", 0);
    meshes[i].draw_count = infinite_u32(num(env, m, "drawCount", INFINITY));
    napi_value gs = prop(env, m, "groups");
    uint32_t ng = 0;
    if (gs) napi_get_array_length(env, gs, &ng);
    groups[i] = calloc(ng ? ng : 1, sizeof(mcr_mesh_group));
    for (uint32_t k = 0; k < ng; k++) {
      napi_value g;
      napi_get_element(env, gs, k, &g);
      mcr_mesh_group *d = &groups[i][k];
      d->start = (uint32_t)num(env, g, "start", 0);
      d->count = infinite_u32(num(env, g, "count", INFINITY));
      d->material = (uint32_t)num(env, g, "material", 0);
      d->flags = (uint8_t)num(env, g, "flags", 0);
      d->side = (uint8_t)num(env, g, "side", 0);
      d->flat_shading = (uint8_t)flag(env, g, "flatShading");
      d->vertex_colors = (uint8_t)flag(env, g, "vertexColors");
    }
    meshes[i].groups = groups[i];
    meshes[i].group_count = ng;
  }
  uint32_t tris = 0;
  int colors = 0;
  int rc = mcr_compile_count(meshes, n, &tris, &colors);
  for (uint32_t i = 0; i < n; i++) free(groups[i]);
  free(groups);
  free(meshes);
  if (rc != MCR_OK) return throw_mcr(env);
  napi_value out, a, b;
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_create_uint32(env, tris, &a));
  CHECK_NAPI(napi_get_boolean(env, colors != 0, &b));
  napi_set_named_property(env, out, "triangleCount", a);
  napi_set_named_property(env, out, "anyColors", b);
  return out;
}

static napi_value new_typed(napi_env env, napi_typedarray_type type, size_t count, size_t elem, void **data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, count * elem, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, type, count, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

/* buildBvh(triangles: Float32Array, count) in the shape bvh.ts:16 returns. */
static napi_value BuildBvh(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  uint32_t n = 0;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &n));
  size_t have = 0;
  const float *tris = typed(env, argv[0], &have);
  if (have < (size_t)n * 9) {
    napi_throw_range_error(env, NULL, "triangles holds fewer than count * 9 floats");
    return NULL;
  }
  if (!ensure_ctx(env)) return NULL;
  size_t maxn = n ? 2 * (size_t)n - 1 : 0;
  float *bounds = malloc((maxn ? maxn : 1) * 6 * sizeof(float));
  int32_t *nodes = malloc((maxn ? maxn : 1) * 3 * sizeof(int32_t));
  void *order_data;
  napi_value order = new_typed(env, napi_uint32_array, n, 4, &order_data);
  uint32_t count = 0;
  int rc = mcr_build_bvh(g_ctx, tris, n, bounds, nodes, order_data, &count);
  if (rc != MCR_OK) {
    free(bounds);
    free(nodes);
    return throw_mcr(env);
  }
  void *bd, *nd;
  napi_value jb = new_typed(env, napi_float32_array, (size_t)count * 6, 4, &bd);
  napi_value jn = new_typed(env, napi_int32_array, (size_t)count * 3, 4, &nd);
  memcpy(bd, bounds, (size_t)count * 6 * 4);
  memcpy(nd, nodes, (size_t)count * 3 * 4);
  free(bounds);
  free(nodes);
  napi_value bvh, out, nc;
  CHECK_NAPI(napi_create_object(env, &bvh));
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_create_uint32(env, count, &nc));
  napi_set_named_property(env, bvh, "bounds", jb);
  napi_set_named_property(env, bvh, "nodes", jn);
  napi_set_named_property(env, bvh, "nodeCount", nc);
  napi_set_named_property(env, out, "bvh", bvh);
  napi_set_named_property(env, out, "order", order);
  return out;
}

static mcr_scene *scene_of(napi_env env, napi_value v) {
  void *p = NULL;
  napi_get_value_external(env, v, &p);
  return (mcr_scene *)p;
}

static napi_value SceneCreate(napi_env env, napi_callback_info info) {
  (void)info;
  if (!ensure_ctx(env)) return NULL;
  mcr_scene *s = NULL;
  if (mcr_scene_create(g_ctx, &s) != MCR_OK) return throw_mcr(env);
  napi_value out;
  CHECK_NAPI(napi_create_external(env, s, NULL, NULL, &out));
  return out;
}

static napi_value SceneRelease(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  mcr_scene_release(scene_of(env, argv[0]));
  return NULL;
}

/* sceneSetGeometry(scene, geometry: CompiledGeometry, version) */
static napi_value SceneSetGeometry(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  napi_value g = argv[1];
  mcr_geometry d;
  memset(&d, 0, sizeof d);
  d.triangle_count = (uint32_t)num(env, g, "triangleCount", 0);
  d.triangles = typed(env, prop(env, g, "triangles"), NULL);
  d.normals = typed(env, prop(env, g, "normals"), NULL);
  d.uvs = typed(env, prop(env, g, "uvs"), NULL);
  d.tangents = typed(env, prop(env, g, "tangents"), NULL);
  d.colors = typed(env, prop(env, g, "colors"), NULL);
  d.materials = typed(env, prop(env, g, "materials"), NULL);
  d.flags = typed(env, prop(env, g, "flags"), NULL);
  d.sides = typed(env, prop(env, g, "sides"), NULL);
  napi_value bvh = prop(env, g, "bvh");
  if (bvh) {
    d.node_count = (uint32_t)num(env, bvh, "nodeCount", 0);
    d.bvh_bounds = typed(env, prop(env, bvh, "bounds"), NULL);
    d.bvh_nodes = typed(env, prop(env, bvh, "nodes"), NULL);
  }
  d.epsilon = num(env, g, "epsilon", 1e-5);
  uint32_t version = 0;
  napi_get_value_uint32(env, argv[2], &version);
  if (mcr_scene_set_geometry(scene_of(env, argv[0]), &d, version) != MCR_OK) return throw_mcr(env);
  return NULL;
}

static int model_of(napi_env env, napi_value m) {
  static const char *names[] = {"physical", "phong", "lambert", "toon", "basic"};
  for (int i = 0; i < 5; i++)
    if (str_is(env, m, "model", names[i])) return i;
  return 0;
}

static void read_material(napi_env env, napi_value m, mcr_material *d) {
  memset(d, 0, sizeof *d);
  d->model = model_of(env, m);
  d->side = (int32_t)num(env, m, "side", 0);
  vec(env, m, "color", d->color, 3);
  d->opacity = num(env, m, "opacity", 1);
  d->transparent = flag(env, m, "transparent");
  d->vertex_colors = flag(env, m, "vertexColors");
  d->alpha_test = num(env, m, "alphaTest", 0);
  d->flat_shading = flag(env, m, "flatShading");
  d->tone_mapped = flag(env, m, "toneMapped");
  d->map = (int32_t)num(env, m, "map", -1);
  d->alpha_map = (int32_t)num(env, m, "alphaMap", -1);
  d->emissive_map = (int32_t)num(env, m, "emissiveMap", -1);
  d->normal_map = (int32_t)num(env, m, "normalMap", -1);
  d->ao_map = (int32_t)num(env, m, "aoMap", -1);
  d->roughness_map = (int32_t)num(env, m, "roughnessMap", -1);
  d->metalness_map = (int32_t)num(env, m, "metalnessMap", -1);
  d->env_map = (int32_t)num(env, m, "envMap", -1);
  vec(env, m, "emissive", d->emissive, 3);
  vec(env, m, "normalScale", d->normal_scale, 2);
  d->ao_map_intensity = num(env, m, "aoMapIntensity", 1);
  d->roughness = num(env, m, "roughness", 1);
  d->metalness = num(env, m, "metalness", 0);
  d->physical = flag(env, m, "physical");
  d->combine = (int32_t)num(env, m, "combine", 0);
  d->ior = num(env, m, "ior", 1.5);
  d->specular_intensity = num(env, m, "specularIntensity", 1);
  vec(env, m, "specularColor", d->specular_color, 3);
  d->transmission = num(env, m, "transmission", 0);
  d->thickness = num(env, m, "thickness", 0);
  vec(env, m, "attenuation", d->attenuation, 3);
  d->shininess = num(env, m, "shininess", 30);
  vec(env, m, "specular", d->specular, 3);
  d->env_map_intensity = num(env, m, "envMapIntensity", 1);
  d->reflectivity = num(env, m, "reflectivity", 1);
}

static void read_light(napi_env env, napi_value l, mcr_light *d) {
  static const char *names[] = {"ambient", "hemisphere", "directional", "point", "spot"};
  memset(d, 0, sizeof *d);
  for (int i = 0; i < 5; i++)
    if (str_is(env, l, "type", names[i])) d->type = i;
  d->cast_shadow = flag(env, l, "castShadow");
  vec(env, l, "color", d->color, 3);
  vec(env, l, "groundColor", d->ground_color, 3);
  vec(env, l, "direction", d->direction, 3);
  vec(env, l, "position", d->position, 3);
  d->distance = num(env, l, "distance", 0);
  d->decay = num(env, l, "decay", 2);
  d->cone_cos = num(env, l, "coneCos", 0);
  d->penumbra_cos = num(env, l, "penumbraCos", 0);
}

/* sceneSetShading(scene, shading: ShadingPayload, pixels: Map<id, Uint8Array | Float32Array>, version) */
static napi_value SceneSetShading(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  napi_value sh = argv[1], pixels = argv[2];
  napi_value mats = prop(env, sh, "materials"), lights = prop(env, sh, "lights");
  napi_value texs = prop(env, sh, "textures"), envs = prop(env, sh, "environments");
  uint32_t nm = 0, nl = 0, nt = 0, ne = 0;
  if (mats) napi_get_array_length(env, mats, &nm);
  if (lights) napi_get_array_length(env, lights, &nl);
  if (texs) napi_get_array_length(env, texs, &nt);
  if (envs) napi_get_array_length(env, envs, &ne);
  mcr_material *dm = calloc(nm ? nm : 1, sizeof(mcr_material));
  mcr_light *dl = calloc(nl ? nl : 1, sizeof(mcr_light));
  mcr_texture *dt = calloc(nt ? nt : 1, sizeof(mcr_texture));
  mcr_environment *de = calloc(ne ? ne : 1, sizeof(mcr_environment));
  napi_value result = NULL;
  for (uint32_t i = 0; i < nm; i++) {
    napi_value m;
    napi_get_element(env, mats, i, &m);
    read_material(env, m, &dm[i]);
  }
  for (uint32_t i = 0; i < nl; i++) {
    napi_value l;
    napi_get_element(env, lights, i, &l);
    read_light(env, l, &dl[i]);
  }
  napi_value get = NULL;
  if (nt) napi_get_named_property(env, pixels, "get", &get);
  for (uint32_t i = 0; i < nt; i++) {
    napi_value t, idv, data;
    napi_get_element(env, texs, i, &t);
    idv = prop(env, t, "id");
    napi_call_function(env, pixels, get, 1, &idv, &data);
    napi_typedarray_type type = napi_uint8_array;
    size_t n = 0, off;
    void *p = NULL;
    napi_value ab;
    bool is = false;
    napi_is_typedarray(env, data, &is);
    if (!is) {  // protocol.ts:84
      char msg[64];
      snprintf(msg, sizeof msg, "texture %.0f was not sent", num(env, t, "id", -1));
      napi_throw_error(env, NULL, msg);
      goto done;
    }
    napi_get_typedarray_info(env, data, &type, &n, &p, &ab, &off);
    dt[i].width = (uint32_t)num(env, t, "width", 0);
    dt[i].height = (uint32_t)num(env, t, "height", 0);
    dt[i].data = p;
    dt[i].is_float = type == napi_float32_array;
    dt[i].srgb = flag(env, t, "srgb");
    dt[i].flip_y = flag(env, t, "flipY");
    dt[i].wrap_s = (int32_t)num(env, t, "wrapS", MCR_WRAP_CLAMP);
    dt[i].wrap_t = (int32_t)num(env, t, "wrapT", MCR_WRAP_CLAMP);
    vec(env, t, "matrix", dt[i].matrix, 9);
  }
  for (uint32_t i = 0; i < ne; i++) {
    napi_value e;
    napi_get_element(env, envs, i, &e);
    de[i].kind = str_is(env, e, "kind", "cube") ? MCR_ENV_CUBE : MCR_ENV_EQUIRECT;
    de[i].texture = (int32_t)num(env, e, "texture", -1);
    double faces[6] = {-1, -1, -1, -1, -1, -1};
    if (prop(env, e, "faces")) vec(env, e, "faces", faces, 6);
    for (int k = 0; k < 6; k++) de[i].faces[k] = (int32_t)faces[k];
    de[i].flip_x = flag(env, e, "flipX");
    vec(env, e, "rotation", de[i].rotation, 9);
    de[i].intensity = num(env, e, "intensity", 1);
  }
  {
    mcr_shading d;
    memset(&d, 0, sizeof d);
    d.material_count = nm;
    d.light_count = nl;
    d.materials = dm;
    d.lights = dl;
    d.texture_count = nt;
    d.environment_count = ne;
    d.textures = dt;
    d.environments = de;
    d.environment = (int32_t)num(env, sh, "environment", -1);
    d.background_texture = -1;
    d.background_environment = -1;
    d.background_intensity = 1;
    napi_value bg = prop(env, sh, "background");
    if (bg) {
      if (str_is(env, bg, "kind", "color")) {
        d.background_kind = MCR_BACKGROUND_COLOR;
        vec(env, bg, "color", d.background_color, 3);
      } else if (str_is(env, bg, "kind", "screen")) {
        d.background_kind = MCR_BACKGROUND_SCREEN;
        d.background_texture = (int32_t)num(env, bg, "texture", -1);
        d.background_intensity = num(env, bg, "intensity", 1);
      } else if (str_is(env, bg, "kind", "environment")) {
        d.background_kind = MCR_BACKGROUND_ENVIRONMENT;
        d.background_environment = (int32_t)num(env, bg, "environment", -1);
        d.background_intensity = num(env, bg, "intensity", 1);
      }
    }
    napi_value fog = prop(env, sh, "fog");
    napi_valuetype ft = napi_null;
    if (fog) napi_typeof(env, fog, &ft);
    if (fog && ft == napi_object) {
      d.fog_kind = str_is(env, fog, "kind", "exp2") ? MCR_FOG_EXP2 : MCR_FOG_LINEAR;
      vec(env, fog, "color", d.fog_color, 3);
      d.fog_near = num(env, fog, "near", 0);
      d.fog_far = num(env, fog, "far", 0);
      d.fog_density = num(env, fog, "density", 0);
    }
    uint32_t version = 0;
    napi_get_value_uint32(env, argv[3], &version);
    if (mcr_scene_set_shading(scene_of(env, argv[0]), &d, version) != MCR_OK) throw_mcr(env);
  }
done:
  free(dm);
  free(dl);
  free(dt);
  free(de);
  return result;
}

/* The tile message's camera, settings and tile (protocol.ts:36-45). */
typedef struct {
  mcr_scene *scene;
  uint32_t gv, sv, width, height, pass, block;
  mcr_camera cam;
  mcr_settings st;
  mcr_tile tile;
  float *out;
  int rc;
  char err[256];
  napi_deferred deferred;
  napi_async_work work;
} TileCall;

static int read_tile(napi_env env, napi_value scene, napi_value msg, TileCall *c) {
  memset(c, 0, sizeof *c);
  c->scene = scene_of(env, scene);
  c->gv = (uint32_t)num(env, msg, "geometryVersion", 0);
  c->sv = (uint32_t)num(env, msg, "shadingVersion", 0);
  c->width = (uint32_t)num(env, msg, "width", 0);
  c->height = (uint32_t)num(env, msg, "height", 0);
  c->pass = (uint32_t)num(env, msg, "pass", 0);
  c->block = (uint32_t)num(env, msg, "block", 1);
  napi_value cam = prop(env, msg, "camera"), st = prop(env, msg, "settings"), tile = prop(env, msg, "tile");
  if (!cam || !st || !tile) return 0;
  vec(env, cam, "projectionInverse", c->cam.projection_inverse, 16);
  vec(env, cam, "world", c->cam.world, 16);
  vec(env, cam, "forward", c->cam.forward, 3);
  vec(env, cam, "position", c->cam.position, 3);
  c->cam.orthographic = flag(env, cam, "orthographic");
  c->st.max_bounces = (int32_t)num(env, st, "maxBounces", 4);
  c->st.shadows = str_is(env, st, "shadows", "castShadow") ? MCR_SHADOWS_CAST
                  : str_is(env, st, "shadows", "none") ? MCR_SHADOWS_NONE : MCR_SHADOWS_ALL;
  c->st.global_illumination = flag(env, st, "globalIllumination");
  c->st.tone_mapping = (int32_t)num(env, st, "toneMapping", 0);
  c->st.tone_mapping_exposure = num(env, st, "toneMappingExposure", 1);
  vec(env, st, "clearColor", c->st.clear_color, 3);
  c->st.clear_alpha = num(env, st, "clearAlpha", 0);
  napi_value seed = prop(env, st, "seed");
  int32_t s32 = 0;
  if (seed) napi_get_value_int32(env, seed, &s32);  /* ToInt32, as random.ts:7 mixes it */
  c->st.seed = (uint32_t)s32;
  c->tile.x = (int32_t)num(env, tile, "x", 0);
  c->tile.y = (int32_t)num(env, tile, "y", 0);
  c->tile.width = (int32_t)num(env, tile, "width", 0);
  c->tile.height = (int32_t)num(env, tile, "height", 0);
  return 1;
}

static void run_tile(TileCall *c) {
  c->rc = mcr_render_tile(g_ctx, c->scene, c->gv, c->sv, &c->cam, &c->st, c->tile, c->width, c->height, c->pass,
                          c->block, c->out, NULL);
  if (c->rc != MCR_OK) snprintf(c->err, sizeof c->err, "%s", mcr_last_error());
}

static napi_value RenderTile(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (!ensure_ctx(env)) return NULL;
  TileCall c;
  if (!read_tile(env, argv[0], argv[1], &c)) {
    napi_throw_type_error(env, NULL, "a tile message needs camera, settings and tile");
    return NULL;
  }
  void *data;
  napi_value out = new_typed(env, napi_float32_array, (size_t)c.tile.width * c.tile.height * 4, 4, &data);
  c.out = data;
  run_tile(&c);
  if (c.rc != MCR_OK) {
    napi_throw_error(env, NULL, c.err);
    return NULL;
  }
  return out;
}

static void tile_execute(napi_env env, void *data) {
  (void)env;
  run_tile((TileCall *)data);
}

static void tile_complete(napi_env env, napi_status status, void *data) {
  TileCall *c = data;
  if (status == napi_ok && c->rc == MCR_OK) {
    void *buf;
    napi_value arr = new_typed(env, napi_float32_array, (size_t)c->tile.width * c->tile.height * 4, 4, &buf);
    memcpy(buf, c->out, (size_t)c->tile.width * c->tile.height * 16);
    napi_resolve_deferred(env, c->deferred, arr);
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, c->rc != MCR_OK ? c->err : "the tile was not rendered", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, c->deferred, err);
  }
  napi_delete_async_work(env, c->work);
  free(c->out);
  free(c);
}

/* renderTileAsync(scene, message) -> Promise<Float32Array>: the render runs on Node's worker pool,
 * so the event loop keeps running; calls on one context are serialised by the library. */
static napi_value RenderTileAsync(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (!ensure_ctx(env)) return NULL;
  TileCall *c = malloc(sizeof(TileCall));
  if (!read_tile(env, argv[0], argv[1], c)) {
    free(c);
    napi_throw_type_error(env, NULL, "a tile message needs camera, settings and tile");
    return NULL;
  }
  c->out = malloc((size_t)(c->tile.width > 0 ? c->tile.width : 1) * (c->tile.height > 0 ? c->tile.height : 1) * 16);
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &c->deferred, &promise));
  CHECK_NAPI(napi_create_string_utf8(env, "mcr.renderTile", NAPI_AUTO_LENGTH, &name));
  CHECK_NAPI(napi_create_async_work(env, NULL, name, tile_execute, tile_complete, c, &c->work));
  CHECK_NAPI(napi_queue_async_work(env, c->work));
  return promise;
}

NAPI_MODULE_INIT() {
  struct {
    const char *name;
    napi_callback fn;
  } fns[] = {{"version", Version},
             {"compileCount", CompileCount},
             {"buildBvh", BuildBvh},
             {"sceneCreate", SceneCreate},
             {"sceneRelease", SceneRelease},
             {"sceneSetGeometry", SceneSetGeometry},
             {"sceneSetShading", SceneSetShading},
             {"renderTile", RenderTile},
             {"renderTileAsync", RenderTileAsync}};
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f) != napi_ok) return NULL;
    if (napi_set_named_property(env, exports, fns[i].name, f) != napi_ok) return NULL;
  }
  return exports;
}
