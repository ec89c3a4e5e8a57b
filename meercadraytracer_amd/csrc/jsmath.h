// jsmath.h — deterministic double-precision math with JavaScript semantics, for host and device.
//
// The trace path evaluates every formula in IEEE double, as the reference's JS `number`
// arithmetic does. The functions here use only correctly rounded IEEE operations (+, -, *, /,
// sqrt), exact ones (floor, fabs, bit casts) and fixed polynomials, so a host (gcc) build and a
// gfx950 (hipcc) build give bit-identical results. Both sides must compile with
// -ffp-contract=off: JS never fuses a multiply into an add.
//
//   js_min / js_max  ECMAScript Math.min / Math.max: NaN propagates, -0 < +0.
//   js_hypot2/3      Math.hypot as V8 evaluates it: scale by the largest magnitude, then a
//                    Kahan-compensated sum of squares, then sqrt(sum) * max.
//   jm_exp           Cody & Waite argument reduction (Software Manual for the Elementary
//                    Functions, 1980) with a two-part ln 2, then the Taylor polynomial of degree
//                    13 on |r| <= ln2/2; error about 1 ulp.
//   jm_log           Reduction to m in [sqrt(1/2), sqrt(2)), then log(1+f) = f - s*(f - R(s^2))
//                    with s = f/(2+f) and R the atanh series; error about 1 ulp.
//   jm_exp2          2^z as 2^k * e^(f ln 2), f = z - k.
//   jm_pow           ECMAScript Math.pow special cases, exact x*x for y = 2 and (x*x)^2 for y = 4,
//                    2^y through jm_exp2, otherwise exp(y * log x) with log x and the product
//                    y * log x carried in double-double (Dekker's exact products, no FMA), so
//                    the error stays near 1 ulp for every |y log x| < 709.
//   jm_sin / jm_cos  Cody & Waite reduction by a two-part pi/2 (valid for |x| < 2^19), Taylor
//                    polynomials of degree 17 / 18 on |r| <= pi/4.
//   jm_log2          jm_log(x) / ln 2.
#ifndef MCR_JSMATH_H
#define MCR_JSMATH_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define JM_HD __host__ __device__ inline
#else
#define JM_HD static inline
#endif

#define JM_INF (__builtin_inf())
#define JM_NAN (__builtin_nan(""))
#define JM_PI 3.141592653589793

JM_HD bool jm_isnan(double x) { return x != x; }
JM_HD double jm_fabs(double x) { return __builtin_fabs(x); }
JM_HD double jm_floor(double x) { return __builtin_floor(x); }
JM_HD double jm_sqrt(double x) { return __builtin_sqrt(x); }
JM_HD bool jm_signbit(double x) { return (__builtin_bit_cast(uint64_t, x) >> 63) != 0; }

// 2^k as a double, for -1022 <= k <= 1023 (exact).
JM_HD double jm_pow2i(int k) { return __builtin_bit_cast(double, (uint64_t)(k + 1023) << 52); }

// x * 2^k for any k, by multiplications with exact powers of two (rounds only when the result is
// subnormal, identically on every IEEE machine).
JM_HD double jm_scale(double x, int k) {
  while (k > 1023) { x *= jm_pow2i(1023); k -= 1023; }
  while (k < -1022) { x *= jm_pow2i(-1022); k += 1022; }
  return x * jm_pow2i(k);
}

// ECMAScript Math.min / Math.max of two numbers.
JM_HD double js_min(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a < b) return a;
  if (b < a) return b;
  return jm_signbit(a) ? a : b;
}
JM_HD double js_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a > b) return a;
  if (b > a) return b;
  return jm_signbit(a) ? b : a;
}
JM_HD double js_min3(double a, double b, double c) { return js_min(js_min(a, b), c); }
JM_HD double js_max3(double a, double b, double c) { return js_max(js_max(a, b), c); }

// Math.hypot over n <= 3 arguments, V8's evaluation order.
JM_HD double js_hypot_n(double x, double y, double z, int count) {
  double v[3] = {x, y, z};
  bool nan = false;
  double m = 0;
  for (int i = 0; i < count; i++) {
    if (v[i] != v[i]) {
      nan = true;
    } else {
      v[i] = jm_fabs(v[i]);
      if (v[i] > m) m = v[i];
    }
  }
  if (m == JM_INF) return JM_INF;
  if (nan) return JM_NAN;
  if (m == 0) return 0;
  double sum = 0, comp = 0;
  for (int i = 0; i < count; i++) {
    double n = v[i] / m;
    double summand = n * n - comp;
    double pre = sum + summand;
    comp = (pre - sum) - summand;
    sum = pre;
  }
  return jm_sqrt(sum) * m;
}
JM_HD double js_hypot3(double x, double y, double z) { return js_hypot_n(x, y, z, 3); }
JM_HD double js_hypot2(double x, double y) { return js_hypot_n(x, y, 0, 2); }

// e^r for |r| <= ~0.35: Taylor polynomial, Horner form.
JM_HD double jm_exp_poly(double r) {
  double p = 1.6059043836821613e-10;          // 1/13!
  p = 2.08767569878681e-09 + r * p;           // 1/12!
  p = 2.505210838544172e-08 + r * p;          // 1/11!
  p = 2.755731922398589e-07 + r * p;          // 1/10!
  p = 2.7557319223985893e-06 + r * p;         // 1/9!
  p = 2.48015873015873e-05 + r * p;           // 1/8!
  p = 0.0001984126984126984 + r * p;          // 1/7!
  p = 0.001388888888888889 + r * p;           // 1/6!
  p = 0.008333333333333333 + r * p;           // 1/5!
  p = 0.041666666666666664 + r * p;           // 1/4!
  p = 0.16666666666666666 + r * p;            // 1/3!
  p = 0.5 + r * p;                            // 1/2!
  // 1 + r + r^2 * (...): keep the leading terms exact as long as possible.
  return 1.0 + (r + (r * r) * p);
}

JM_HD double jm_exp(double x) {
  const double LN2_HI = 0.6931471803691238;    // ln 2 to 32 bits: k * LN2_HI is exact
  const double LN2_LO = 1.9082149292705877e-10;
  const double INV_LN2 = 1.4426950408889634;
  if (x != x) return x;
  if (x > 709.782712893384) return JM_INF;
  if (x < -745.1332191019412) return 0.0;
  double kd = jm_floor(x * INV_LN2 + 0.5);
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  return jm_scale(jm_exp_poly(r), (int)kd);
}

// 2^z.
JM_HD double jm_exp2(double z) {
  const double LN2 = 0.6931471805599453;
  if (z != z) return z;
  if (z >= 1024.0) return JM_INF;
  if (z < -1075.0) return 0.0;
  double kd = jm_floor(z + 0.5);
  double f = z - kd;  // exact
  return jm_scale(jm_exp_poly(f * LN2), (int)kd);
}

JM_HD double jm_log(double x) {
  const double LN2_HI = 0.6931471803691238;
  const double LN2_LO = 1.9082149292705877e-10;
  const double SQRT_HALF = 0.7071067811865476;
  if (x != x || x < 0) return JM_NAN;
  if (x == 0) return -JM_INF;
  if (x == JM_INF) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) {  // subnormal: scale into the normal range first
    x *= 18014398509481984.0;         // 2^54
    e = -54;
  }
  uint64_t bits = __builtin_bit_cast(uint64_t, x);
  e += (int)((bits >> 52) & 0x7ff) - 1022;
  double m = __builtin_bit_cast(double, (bits & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL);  // [0.5,1)
  if (m < SQRT_HALF) {
    m = m * 2.0;
    e -= 1;
  }
  double f = m - 1.0;  // exact
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 0.09523809523809523;   // 2/21
  R = 0.10526315789473684 + z * R;  // 2/19
  R = 0.11764705882352941 + z * R;  // 2/17
  R = 0.13333333333333333 + z * R;  // 2/15
  R = 0.15384615384615385 + z * R;  // 2/13
  R = 0.18181818181818182 + z * R;  // 2/11
  R = 0.2222222222222222 + z * R;   // 2/9
  R = 0.2857142857142857 + z * R;   // 2/7
  R = 0.4 + z * R;                  // 2/5
  R = 0.6666666666666666 + z * R;   // 2/3
  R = z * R;
  double lp = f - s * (f - R);
  double ed = (double)e;
  return ed * LN2_HI + (lp + ed * LN2_LO);
}

JM_HD double jm_log2(double x) { return jm_log(x) / 0.6931471805599453; }

// Error-free transformations (Dekker 1971, "A floating-point technique for extending the
// available precision"): a*b = p + e exactly, by Veltkamp splitting, without fused multiply-add.
JM_HD void jm_split(double a, double &hi, double &lo) {
  double c = 134217729.0 * a;  // 2^27 + 1
  hi = c - (c - a);
  lo = a - hi;
}
JM_HD void jm_two_prod(double a, double b, double &p, double &e) {
  p = a * b;
  double ah, al, bh, bl;
  jm_split(a, ah, al);
  jm_split(b, bh, bl);
  e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

// log(x) as hi + lo (about 2^-100 relative), for x > 0 finite and normal.
JM_HD void jm_log_dd(double x, double &hi, double &lo) {
  const double LN2_HI = 0.6931471803691238;
  const double LN2_LO = 1.9082149292705877e-10;
  const double SQRT_HALF = 0.7071067811865476;
  uint64_t bits = __builtin_bit_cast(uint64_t, x);
  int e = (int)((bits >> 52) & 0x7ff) - 1022;
  double m = __builtin_bit_cast(double, (bits & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL);
  if (m < SQRT_HALF) {
    m = m * 2.0;
    e -= 1;
  }
  double f = m - 1.0;                 // exact
  double uh = 2.0 + f;                // 2 + f as uh + ul exactly (|f| < 2)
  double ul = f - (uh - 2.0);
  double sh = f / uh;                 // s = f / (2 + f) as sh + sl
  double p, pe;
  jm_two_prod(sh, uh, p, pe);
  double sl = (((f - p) - pe) - sh * ul) / uh;
  double z = sh * sh;
  double R = 0.09523809523809523 * 0.5;  // atanh series: s + s^3/3 + s^5/5 + ...: here 1/(2k+1)
  R = 0.05263157894736842 + z * R;       // 1/19
  R = 0.058823529411764705 + z * R;      // 1/17
  R = 0.06666666666666667 + z * R;       // 1/15
  R = 0.07692307692307693 + z * R;       // 1/13
  R = 0.09090909090909091 + z * R;       // 1/11
  R = 0.1111111111111111 + z * R;        // 1/9
  R = 0.14285714285714285 + z * R;       // 1/7
  R = 0.2 + z * R;                       // 1/5
  R = 0.3333333333333333 + z * R;        // 1/3
  double tail = 2.0 * (sh * z) * R;      // 2 s^3 (1/3 + s^2/5 + ...)
  double lh = 2.0 * sh;                  // exact
  double ll = 2.0 * sl + tail;
  // + e ln 2
  double ed = (double)e;
  double a = ed * LN2_HI;                // exact
  double s1 = a + lh;
  double bb = s1 - a;
  double err = (a - (s1 - bb)) + (lh - bb);  // two-sum
  hi = s1;
  lo = err + ll + ed * LN2_LO;
  double h2 = hi + lo;                   // renormalise
  lo = lo - (h2 - hi);
  hi = h2;
}

// e^(hi + lo) for |hi| < 709.
JM_HD double jm_exp_dd(double hi, double lo) {
  const double LN2_HI = 0.6931471803691238;
  const double LN2_LO = 1.9082149292705877e-10;
  const double INV_LN2 = 1.4426950408889634;
  if (hi != hi) return hi;
  if (hi > 709.782712893384) return JM_INF;
  if (hi < -745.1332191019412) return 0.0;
  double kd = jm_floor(hi * INV_LN2 + 0.5);
  double r = (hi - kd * LN2_HI) + (lo - kd * LN2_LO);
  return jm_scale(jm_exp_poly(r), (int)kd);
}

JM_HD bool jm_is_integer(double y) { return jm_floor(y) == y && y - y == 0; }

// ECMAScript Math.pow(x, y) (Number::exponentiate).
JM_HD double jm_pow(double x, double y) {
  if (y != y) return JM_NAN;
  if (y == 0) return 1.0;
  if (x != x) return JM_NAN;
  double ay = jm_fabs(y);
  if (ay == JM_INF) {
    double ax = jm_fabs(x);
    if (ax == 1.0) return JM_NAN;
    if (ax > 1.0) return y > 0 ? JM_INF : 0.0;
    return y > 0 ? 0.0 : JM_INF;
  }
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  if (y == 4.0) {
    double x2 = x * x;
    return x2 * x2;
  }
  bool yint = jm_is_integer(y);
  bool yodd = yint && jm_fabs(y) < 9007199254740992.0 && jm_fabs(jm_floor(y * 0.5) * 2.0 - y) == 1.0;
  if (x == 0) {
    bool neg = jm_signbit(x) && yodd;
    if (y > 0) return neg ? -0.0 : 0.0;
    return neg ? -JM_INF : JM_INF;
  }
  if (x == JM_INF) return y > 0 ? JM_INF : 0.0;
  if (x == -JM_INF) {
    if (y > 0) return yodd ? -JM_INF : JM_INF;
    return yodd ? -0.0 : 0.0;
  }
  double sign = 1.0;
  if (x < 0) {
    if (!yint) return JM_NAN;
    if (yodd) sign = -1.0;
    x = -x;
  }
  if (x == 2.0) return sign * jm_exp2(y);
  if (x < 2.2250738585072014e-308) return sign * jm_exp(y * jm_log(x));  // subnormal base
  double lh, ll;
  jm_log_dd(x, lh, ll);
  double wp, we;
  jm_two_prod(y, lh, wp, we);
  return sign * jm_exp_dd(wp, we + y * ll);
}

// sin and cos of r, |r| <= pi/4.
JM_HD double jm_sin_poly(double r) {
  double z = r * r;
  double p = 2.8114572543455206e-15;       // 1/17!
  p = -7.647163731819816e-13 + z * p;      // -1/15!
  p = 1.6059043836821613e-10 + z * p;      // 1/13!
  p = -2.505210838544172e-08 + z * p;      // -1/11!
  p = 2.7557319223985893e-06 + z * p;      // 1/9!
  p = -0.0001984126984126984 + z * p;      // -1/7!
  p = 0.008333333333333333 + z * p;        // 1/5!
  p = -0.16666666666666666 + z * p;        // -1/3!
  return r + (r * z) * p;
}
JM_HD double jm_cos_poly(double r) {
  double z = r * r;
  double p = -1.5619206968586225e-16;      // -1/18!
  p = 4.779477332387385e-14 + z * p;       // 1/16!
  p = -1.1470745597729725e-11 + z * p;     // -1/14!
  p = 2.08767569878681e-09 + z * p;        // 1/12!
  p = -2.755731922398589e-07 + z * p;      // -1/10!
  p = 2.48015873015873e-05 + z * p;        // 1/8!
  p = -0.001388888888888889 + z * p;       // -1/6!
  p = 0.041666666666666664 + z * p;        // 1/4!
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  return w + (((1.0 - w) - hz) + (z * z) * p);
}

// Reduces x to r in [-pi/4, pi/4] and returns the quadrant n mod 4.
JM_HD int jm_reduce_pio2(double x, double *r) {
  const double PIO2_1 = 1.5707963267341256;    // pi/2 to 33 bits: n * PIO2_1 is exact
  const double PIO2_1T = 6.077100506506192e-11;
  const double TWO_OVER_PI = 0.6366197723675814;
  double n = jm_floor(x * TWO_OVER_PI + 0.5);
  *r = (x - n * PIO2_1) - n * PIO2_1T;
  long long q = (long long)n;
  return (int)(q & 3);
}

JM_HD double jm_sin(double x) {
  if (x != x || jm_fabs(x) == JM_INF) return JM_NAN;
  double r;
  int q = jm_reduce_pio2(x, &r);
  switch (q) {
    case 0: return jm_sin_poly(r);
    case 1: return jm_cos_poly(r);
    case 2: return -jm_sin_poly(r);
    default: return -jm_cos_poly(r);
  }
}

JM_HD double jm_cos(double x) {
  if (x != x || jm_fabs(x) == JM_INF) return JM_NAN;
  double r;
  int q = jm_reduce_pio2(x, &r);
  switch (q) {
    case 0: return jm_cos_poly(r);
    case 1: return -jm_sin_poly(r);
    case 2: return -jm_cos_poly(r);
    default: return jm_sin_poly(r);
  }
}

// atan on |x| <= 1: atan(x) = atan(c) + atan((x - c) / (1 + x c)) with c the nearest multiple of
// 1/16, whose arctangents are tabulated (rounded from a 60-digit evaluation), and the Taylor series
// of atan on the remainder (|r| <= 1/32, nine terms leave < 1e-26).
JM_HD double jm_atan_unit(double x) {
  const double T[17] = {0.0, 0.06241880999595735, 0.12435499454676144, 0.18534794999569476,
                        0.24497866312686414, 0.3028848683749714, 0.35877067027057225, 0.4124104415973873,
                        0.4636476090008061, 0.5123894603107377, 0.5585993153435624, 0.6022873461349642,
                        0.6435011087932844, 0.6823165548747481, 0.7188299996216245, 0.7531512809621944,
                        0.7853981633974483};
  double a = jm_fabs(x);
  int i = (int)(a * 16 + 0.5);
  double c = i * 0.0625;
  double r = (a - c) / (1 + a * c);
  double r2 = r * r;
  double p = 1.0 / 17;
  p = 1.0 / 15 - r2 * p;
  p = 1.0 / 13 - r2 * p;
  p = 1.0 / 11 - r2 * p;
  p = 1.0 / 9 - r2 * p;
  p = 1.0 / 7 - r2 * p;
  p = 1.0 / 5 - r2 * p;
  p = 1.0 / 3 - r2 * p;
  double v = T[i] + (r - r * r2 * p);
  return x < 0 ? -v : v;
}

// Math.atan2 with the IEEE special cases (zeros, infinities, NaN).
JM_HD double jm_atan2(double y, double x) {
  const double PIO2 = 1.5707963267948966, PIO4 = 0.7853981633974483;
  if (y != y || x != x) return JM_NAN;
  bool yneg = jm_signbit(y);
  double r;
  if (y == 0) {
    r = jm_signbit(x) ? JM_PI : 0.0;
  } else if (x == 0) {
    r = PIO2;
  } else if (jm_fabs(x) == JM_INF) {
    if (jm_fabs(y) == JM_INF) r = x > 0 ? PIO4 : 3 * PIO4;
    else r = x > 0 ? 0.0 : JM_PI;
  } else if (jm_fabs(y) == JM_INF) {
    r = PIO2;
  } else {
    double ay = jm_fabs(y), ax = jm_fabs(x);
    double a = ay <= ax ? jm_atan_unit(ay / ax) : PIO2 - jm_atan_unit(ax / ay);
    r = x > 0 ? a : JM_PI - a;
  }
  return yneg ? -r : r;
}

// Math.asin: atan2(x, sqrt((1 - x)(1 + x))), NaN outside [-1, 1].
JM_HD double jm_asin(double x) {
  if (!(jm_fabs(x) <= 1)) return JM_NAN;
  return jm_atan2(x, jm_sqrt((1 - x) * (1 + x)));
}

#endif  // MCR_JSMATH_H
