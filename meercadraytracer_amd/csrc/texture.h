// texture.h — texture and environment lookups on the device (textures.ts:229-317), the same
// arithmetic as the oracle's restatement with the portable math of jsmath.h.
#ifndef MCR_TEXTURE_H
#define MCR_TEXTURE_H

#include "device.h"
#include "jsmath.h"

// color.ts:21-25: linear values of the 256 sRGB-encoded levels, as a Float32Array holds them.
extern __constant__ float g_srgb8[256];

// textures.ts:229-236
MCR_INLINE double tex_wrap(double c, int mode) {
  if (mode == MCR_WRAP_CLAMP) return c < 0 ? 0 : c > 1 ? 1 : c;
  if (mode == MCR_WRAP_MIRROR) {
    double period = c - 2 * jm_floor(c / 2);
    return period > 1 ? 2 - period : period;
  }
  return c - jm_floor(c);
}

// textures.ts:238-241 (i is an integer-valued double)
MCR_INLINE int64_t tex_index(double i, int64_t size, int mode) {
  int64_t k = (int64_t)i;
  if (mode == MCR_WRAP_CLAMP || mode == MCR_WRAP_MIRROR) return k < 0 ? 0 : k >= size ? size - 1 : k;
  return ((k % size) + size) % size;
}

// textures.ts:255-285: bilinear, sRGB decode of 8-bit colour channels.
MCR_INLINE void tex_texel(const DevTexture &tx, double s, double t, double *out) {
  int64_t w = tx.width, h = tx.height;
  double row = tx.flip_y ? 1 - t : t;
  double x = s * w - 0.5, y = row * h - 0.5;
  double x0 = jm_floor(x), y0 = jm_floor(y);
  double fx = x - x0, fy = y - y0;
  int64_t xa = tex_index(x0, w, tx.wrap_s), xb = tex_index(x0 + 1, w, tx.wrap_s);
  int64_t ya = tex_index(y0, h, tx.wrap_t), yb = tex_index(y0 + 1, h, tx.wrap_t);
  int64_t i00 = (ya * w + xa) * 4, i10 = (ya * w + xb) * 4, i01 = (yb * w + xa) * 4, i11 = (yb * w + xb) * 4;
  double w00 = (1 - fx) * (1 - fy), w10 = fx * (1 - fy), w01 = (1 - fx) * fy, w11 = fx * fy;
  if (tx.is_float) {
    const float *d = (const float *)tx.data;
    for (int c = 0; c < 4; c++)
      out[c] = (double)d[i00 + c] * w00 + (double)d[i10 + c] * w10 + (double)d[i01 + c] * w01 +
               (double)d[i11 + c] * w11;
    return;
  }
  const uint8_t *d = (const uint8_t *)tx.data;
  for (int c = 0; c < 4; c++) {
    if (tx.srgb && c < 3) {
      out[c] = (double)g_srgb8[d[i00 + c]] * w00 + (double)g_srgb8[d[i10 + c]] * w10 +
               (double)g_srgb8[d[i01 + c]] * w01 + (double)g_srgb8[d[i11 + c]] * w11;
    } else {
      out[c] = ((double)d[i00 + c] * w00 + (double)d[i10 + c] * w10 + (double)d[i01 + c] * w01 +
                (double)d[i11 + c] * w11) / 255;
    }
  }
}

// textures.ts:247-253: uv transform, wrap, sample.
MCR_INLINE void tex_sample(const DevTexture &tx, double u, double v, double *out) {
  const double *m = tx.matrix;
  double tu = m[0] * u + m[3] * v + m[6];
  double tv = m[1] * u + m[4] * v + m[7];
  tex_texel(tx, tex_wrap(tu, tx.wrap_s), tex_wrap(tv, tx.wrap_t), out);
}

// textures.ts:287-317: radiance of an environment in a world direction (RGB).
MCR_INLINE void env_sample(const DevEnvironment &env, const DevTexture *textures, double dx, double dy,
                           double dz, double *out) {
  const double *r = env.rotation;
  double x = r[0] * dx + r[3] * dy + r[6] * dz;
  double y = r[1] * dx + r[4] * dy + r[7] * dz;
  double z = r[2] * dx + r[5] * dy + r[8] * dz;
  double t4[4];
  if (env.kind == MCR_ENV_EQUIRECT) {
    double length = js_hypot3(x, y, z);
    if (!(length != 0)) length = 1;
    double u = jm_atan2(z, x) / (2 * JM_PI) + 0.5;
    double v = jm_asin(js_max(-1, js_min(1, y / length))) / JM_PI + 0.5;
    const DevTexture &tx = textures[env.texture];
    tex_texel(tx, tex_wrap(u, tx.wrap_s), js_max(0, js_min(1, v)), t4);
  } else {
    if (env.flip_x) x = -x;
    double ax = jm_fabs(x), ay = jm_fabs(y), az = jm_fabs(z);
    int face;
    double sc, tc, ma;
    if (ax >= ay && ax >= az) {
      face = x > 0 ? 0 : 1;
      ma = ax;
      sc = x > 0 ? -z : z;
      tc = -y;
    } else if (ay >= az) {
      face = y > 0 ? 2 : 3;
      ma = ay;
      sc = x;
      tc = y > 0 ? z : -z;
    } else {
      face = z > 0 ? 4 : 5;
      ma = az;
      sc = z > 0 ? x : -x;
      tc = -y;
    }
    double s2 = (sc / ma + 1) * 0.5, t2 = (tc / ma + 1) * 0.5;
    tex_texel(textures[env.faces[face]], js_max(0, js_min(1, s2)), js_max(0, js_min(1, t2)), t4);
  }
  out[0] = t4[0] * env.intensity;
  out[1] = t4[1] * env.intensity;
  out[2] = t4[2] * env.intensity;
}

#endif  // MCR_TEXTURE_H
