// shading.h — device restatement of the reference's BRDF, colour and RNG routines, in FP64 with
// the deterministic math of jsmath.h.
#ifndef MCR_SHADING_H
#define MCR_SHADING_H

#include "device.h"
#include "dfg_lut.h"

#define MCR_PI 3.141592653589793
#define MCR_INV_PI (1.0 / 3.141592653589793)  // brdf.ts:6

// ---- random.ts ------------------------------------------------------------------------------
__device__ inline uint32_t rng_hash(uint32_t x) {  // random.ts:24-31 (lowbias32)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ inline uint32_t rng_seed(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {  // random.ts:6-10
  uint32_t h = rng_hash(a ^ rng_hash(b ^ rng_hash(c ^ rng_hash(d))));
  return h == 0 ? 0x9e3779b9u : h;
}
__device__ inline double rng_next(uint32_t &state) {  // random.ts:13-20
  uint32_t x = state;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  state = x;
  return (double)x * (1.0 / 4294967296.0);  // exact: x < 2^32, a power-of-two scale
}

// ---- color.ts -------------------------------------------------------------------------------
__device__ inline double linear_to_srgb(double c) {  // color.ts:16-18
  return c < 0.0031308 ? c * 12.92 : 1.055 * jm_pow(c, 0.41666) - 0.055;
}
__device__ inline double saturate(double x) { return x < 0 ? 0 : x > 1 ? 1 : x; }
__device__ inline void mul3(const double *m, double *o) {  // color.ts:54-59
  double x = o[0], y = o[1], z = o[2];
  o[0] = m[0] * x + m[3] * y + m[6] * z;
  o[1] = m[1] * x + m[4] * y + m[7] * z;
  o[2] = m[2] * x + m[5] * y + m[8] * z;
}

// color.ts:81-136
__device__ inline void tone_map(double *c, int mapping, double exposure) {
  // three.js GLSL mat3 constructors, column by column (color.ts:62-73).
  const double ACES_IN[9] = {0.59719, 0.076, 0.0284, 0.35458, 0.90834, 0.13383, 0.04823, 0.01566, 0.83777};
  const double ACES_OUT[9] = {1.60475, -0.10208, -0.00327, -0.53108, 1.10813, -0.07276, -0.07367, -0.00605, 1.07602};
  const double TO2020[9] = {0.6274, 0.0691, 0.0164, 0.3293, 0.9195, 0.088, 0.0433, 0.0113, 0.8956};
  const double FROM2020[9] = {1.6605, -0.1246, -0.0182, -0.5876, 1.1329, -0.1006, -0.0728, -0.0083, 1.1187};
  const double INSET[9] = {0.856627153315983, 0.137318972929847, 0.11189821299995, 0.0951212405381588,
                           0.761241990602591, 0.0767994186031903, 0.0482516061458583, 0.101439036467562,
                           0.811302368396859};
  const double OUTSET[9] = {1.1271005818144368, -0.1413297634984383, -0.14132976349843826,
                            -0.11060664309660323, 1.157823702216272, -0.11060664309660294,
                            -0.016493938717834573, -0.016493938717834257, 1.2519364065950405};
  switch (mapping) {
    case 1:  // Linear
      for (int i = 0; i < 3; i++) c[i] = saturate(exposure * c[i]);
      return;
    case 2:  // Reinhard
      for (int i = 0; i < 3; i++) {
        double v = exposure * c[i];
        c[i] = saturate(v / (1 + v));
      }
      return;
    case 3:  // Cineon
      for (int i = 0; i < 3; i++) {
        double v = js_max(0, exposure * c[i] - 0.004);
        c[i] = jm_pow((v * (6.2 * v + 0.5)) / (v * (6.2 * v + 1.7) + 0.06), 2.2);
      }
      return;
    case 4:  // ACES filmic
      for (int i = 0; i < 3; i++) c[i] = (c[i] * exposure) / 0.6;
      mul3(ACES_IN, c);
      for (int i = 0; i < 3; i++) {
        double v = c[i];
        double a = v * (v + 0.0245786) - 0.000090537;
        double b = v * (0.983729 * v + 0.432951) + 0.238081;
        c[i] = a / b;
      }
      mul3(ACES_OUT, c);
      for (int i = 0; i < 3; i++) c[i] = saturate(c[i]);
      return;
    case 6:  // AgX
      for (int i = 0; i < 3; i++) c[i] = c[i] * exposure;
      mul3(TO2020, c);
      mul3(INSET, c);
      for (int i = 0; i < 3; i++) {
        double v = (jm_log2(js_max(c[i], 1e-10)) - (-12.47393)) / (4.026069 - (-12.47393));
        double x = saturate(v);
        double x2 = x * x, x4 = x2 * x2;
        c[i] = 15.5 * x4 * x2 - 40.14 * x4 * x + 31.96 * x4 - 6.868 * x2 * x + 0.4298 * x2 + 0.1191 * x - 0.00232;
      }
      mul3(OUTSET, c);
      for (int i = 0; i < 3; i++) c[i] = jm_pow(js_max(0, c[i]), 2.2);
      mul3(FROM2020, c);
      for (int i = 0; i < 3; i++) c[i] = saturate(c[i]);
      return;
    case 7: {  // Neutral
      double start = 0.8 - 0.04, desat = 0.15;
      for (int i = 0; i < 3; i++) c[i] = c[i] * exposure;
      double x = js_min3(c[0], c[1], c[2]);
      double offset = x < 0.08 ? x - 6.25 * x * x : 0.04;
      for (int i = 0; i < 3; i++) c[i] = c[i] - offset;
      double peak = js_max3(c[0], c[1], c[2]);
      if (peak < start) return;
      double d = 1 - start;
      double np = 1 - (d * d) / (peak + d - start);
      double g = 1 - 1 / (desat * (peak - np) + 1);
      for (int i = 0; i < 3; i++) c[i] = (c[i] * np) / peak * (1 - g) + np * g;
      return;
    }
    default:
      return;
  }
}

// ---- brdf.ts --------------------------------------------------------------------------------
// color.ts:28-35 halfToFloat, for the DFG table (brdf.ts:8-12 stores it as float32).
__host__ __device__ inline float half_to_float_js(uint32_t h) {
  double sign = (h & 0x8000) ? -1 : 1;
  int e = (h >> 10) & 0x1f;
  int f = h & 0x3ff;
  double v;
  if (e == 0) v = sign * jm_pow2i(-14) * (f / 1024.0);
  else if (e == 31) v = f ? JM_NAN : sign * JM_INF;
  else v = sign * jm_pow2i(e - 15) * (1 + f / 1024.0);
  return (float)v;
}

// brdf.ts:18-30, reading the table from `lut` (512 floats in constant memory).
__device__ inline void dfg(const float *lut, double roughness, double dotNV, double &s0, double &s1) {
  const int n = MCR_DFG_LUT_SIZE;
  double x = js_min(js_max(roughness * n - 0.5, 0), n - 1);
  double y = js_min(js_max(dotNV * n - 0.5, 0), n - 1);
  double x0 = jm_floor(x), y0 = jm_floor(y);
  double x1 = js_min(x0 + 1, n - 1), y1 = js_min(y0 + 1, n - 1);
  double fx = x - x0, fy = y - y0;
  int ix0 = (int)x0, ix1 = (int)x1, iy0 = (int)y0, iy1 = (int)y1;
  double out[2];
  for (int c = 0; c < 2; c++) {
    double a = lut[(iy0 * n + ix0) * 2 + c], b = lut[(iy0 * n + ix1) * 2 + c];
    double d = lut[(iy1 * n + ix0) * 2 + c], e = lut[(iy1 * n + ix1) * 2 + c];
    out[c] = (a * (1 - fx) + b * fx) * (1 - fy) + (d * (1 - fx) + e * fx) * fy;
  }
  s0 = out[0];
  s1 = out[1];
}
__device__ inline double fresnel_weight(double dotVH) {  // brdf.ts:33-35
  return jm_pow(2, (-5.55473 * dotVH - 6.98316) * dotVH);
}
__device__ inline double visibility_ggx(double alpha, double dotNL, double dotNV) {  // brdf.ts:38-43
  double a2 = alpha * alpha;
  double gv = dotNL * jm_sqrt(a2 + (1 - a2) * dotNV * dotNV);
  double gl = dotNV * jm_sqrt(a2 + (1 - a2) * dotNL * dotNL);
  return 0.5 / js_max(gv + gl, 1e-6);
}
__device__ inline double distribution_ggx(double alpha, double dotNH) {  // brdf.ts:46-50
  double a2 = alpha * alpha;
  double denom = dotNH * dotNH * (a2 - 1) + 1;
  return (MCR_INV_PI * a2) / (denom * denom);
}
__device__ inline double blinn_phong(double shininess, double dotNH) {  // brdf.ts:53-55
  return 0.25 * MCR_INV_PI * (shininess * 0.5 + 1) * jm_pow(dotNH, shininess);
}
__device__ inline double smoothstep_js(double e0, double e1, double x) {  // brdf.ts:58-62
  if (e0 == e1) return x < e0 ? 0 : 1;
  double t = js_min(js_max((x - e0) / (e1 - e0), 0), 1);
  return t * t * (3 - 2 * t);
}
__device__ inline double distance_attenuation(double distance, double cutoff, double decay) {  // brdf.ts:65-72
  double falloff = 1 / js_max(jm_pow(distance, decay), 0.01);
  if (cutoff > 0) {
    double r = js_min(js_max(1 - jm_pow(distance / cutoff, 4), 0), 1);
    falloff *= r * r;
  }
  return falloff;
}
__device__ inline double luminance(double r, double g, double b) {  // tracer.ts:694-696
  return 0.2126 * r + 0.7152 * g + 0.0722 * b;
}
__device__ inline double multi_scatter(double fss, double f0, double ems) {  // tracer.ts:699-702
  double fAvg = f0 + (1 - f0) * 0.047619;
  return ((fss * fAvg) / (1 - ems * fAvg)) * ems;
}
__device__ inline double or_one(double v) { return (v == 0 || v != v) ? 1.0 : v; }  // `x || 1`

// tracer.ts:636-646 (Duff et al. 2017 orthonormal basis)
__device__ inline void to_world(double nx, double ny, double nz, double a, double b, double c,
                                double &sx, double &sy, double &sz) {
  double sign = nz >= 0 ? 1 : -1;
  double e = -1 / (sign + nz);
  double f = nx * ny * e;
  double tx = 1 + sign * nx * nx * e, ty = sign * f, tz = -sign * nx;
  double bx = f, by = sign + ny * ny * e, bz = -ny;
  sx = tx * a + bx * b + nx * c;
  sy = ty * a + by * b + ny * c;
  sz = tz * a + bz * b + nz * c;
}

#endif  // MCR_SHADING_H
