// Checks, on the CPU, the argument behind the FP32 slab pre-test of the any-hit traversals
// (meercadraytracer_amd/csrc/device.h: slab_setup, slab_threshold, slab_class): whenever it calls
// a box a hit or a miss, the reference's FP64 slab test (bvh.ts:340-356) agrees, and when both of
// two boxes are hit and their FP32 entry estimates differ by more than the margin, their FP64
// entries are ordered the same way. The functions below restate device.h's arithmetic with
// host fmaf and directed rounding; random rays and boxes cover scene scales from 1e-3 to 1e3,
// nearly axis-parallel rays and boxes placed so a face or an edge passes through a point of the
// ray (the cases the margin exists for). Exit status 1 on any disagreement.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include <algorithm>
static inline float rd(double x) { if (std::isnan(x)) return NAN; float f = (float)x; if ((double)f > x) f = std::nextafterf(f, -INFINITY); return f; }
static inline float ru(double x) { if (std::isnan(x)) return NAN; float f = (float)x; if ((double)f < x) f = std::nextafterf(f, INFINITY); return f; }
struct SlabF32 { float ix, iy, iz, oix, oiy, oiz, M; bool ok; };
static void slab_threshold(double t, float M, float &below, float &above) {
  if (!(std::fabs(t) < INFINITY)) { below = above = (float)t; return; }
  double slack = std::fabs(t) * 0x1p-20 + (double)M;
  below = rd(t - slack); above = ru(t + slack);
}
static SlabF32 slab_setup(const float *rb, double ox, double oy, double oz, double ix, double iy, double iz) {
  SlabF32 r; float B = 0; for (int k = 0; k < 6; k++) B = std::fmax(B, std::fabs(rb[k]));
  double ax = std::fabs(ix), ay = std::fabs(iy), az = std::fabs(iz);
  double mx = ((double)B + std::fabs(ox)) * ax, my = ((double)B + std::fabs(oy)) * ay, mz = ((double)B + std::fabs(oz)) * az;
  double m = std::fmax(std::fmax(mx, my), mz) * 0x1p-21; m = std::fmax(m, 0x1p-100);
  r.ok = ax <= 0x1p64 && ay <= 0x1p64 && az <= 0x1p64 && m <= 0x1p80;
  r.ix = (float)ix; r.iy = (float)iy; r.iz = (float)iz;
  r.oix = (float)(ox * ix); r.oiy = (float)(oy * iy); r.oiz = (float)(oz * iz);
  r.M = ru(m); return r;
}
static int slab_class(const float *bb, const SlabF32 &r, float th, float tm, float &lo_out) {
  float t0 = std::fmaf(bb[0], r.ix, -r.oix), t1 = std::fmaf(bb[3], r.ix, -r.oix);
  float lo = std::fmin(t0, t1), hi = std::fmax(t0, t1);
  t0 = std::fmaf(bb[1], r.iy, -r.oiy); t1 = std::fmaf(bb[4], r.iy, -r.oiy);
  lo = std::fmax(lo, std::fmin(t0, t1)); hi = std::fmin(hi, std::fmax(t0, t1));
  t0 = std::fmaf(bb[2], r.iz, -r.oiz); t1 = std::fmaf(bb[5], r.iz, -r.oiz);
  lo = std::fmax(lo, std::fmin(t0, t1)); hi = std::fmin(hi, std::fmax(t0, t1));
  lo_out = lo; float d = hi - lo;
  if (d >= r.M && hi >= r.M && lo <= th) return 1;
  if (d < -r.M || hi < -r.M || lo > tm) return 0;
  return 2;
}
static double box_entry(const float *b, double ox, double oy, double oz, double ix, double iy, double iz, double tmin, double tmax) {
  double t0 = ((double)b[0] - ox) * ix, t1 = ((double)b[3] - ox) * ix;
  double lo = std::fmin(t0, t1), hi = std::fmax(t0, t1);
  t0 = ((double)b[1] - oy) * iy; t1 = ((double)b[4] - oy) * iy;
  lo = std::fmax(lo, std::fmin(t0, t1)); hi = std::fmin(hi, std::fmax(t0, t1));
  t0 = ((double)b[2] - oz) * iz; t1 = ((double)b[5] - oz) * iz;
  lo = std::fmax(lo, std::fmin(t0, t1)); hi = std::fmin(hi, std::fmax(t0, t1));
  if (hi >= lo && hi >= tmin && lo <= tmax) return std::fmax(lo, tmin);
  return INFINITY;
}
int main(int argc, char **argv) {
  long N = argc > 1 ? atol(argv[1]) : 10000000;
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> U(-1, 1);
  long bad = 0, amb = 0, total = 0, bad_order = 0, amb_order = 0, orders = 0;
  for (long it = 0; it < N; it++) {
    double scale = std::pow(10.0, (int)(g() % 7) - 3);   // scene scales 1e-3 .. 1e3
    float rb[6] = {(float)(-scale), (float)(-scale), (float)(-scale), (float)scale, (float)scale, (float)scale};
    double ox = U(g) * scale * 2, oy = U(g) * scale * 2, oz = U(g) * scale * 2;
    double dx = U(g), dy = U(g), dz = U(g);
    int mode = g() % 4;
    if (mode == 1) dx *= 1e-6;                 // nearly axis-parallel rays
    if (mode == 2) { dy = 0; }                 // exactly parallel: inverse 1e30 -> FP64 path
    double len = std::sqrt(dx * dx + dy * dy + dz * dz); dx /= len; dy /= len; dz /= len;
    double ix = dx != 0 ? 1 / dx : 1e30, iy = dy != 0 ? 1 / dy : 1e30, iz = dz != 0 ? 1 / dz : 1e30;
    double tmin = 0;
    double tmax = (g() % 3) ? INFINITY : std::fabs(U(g)) * scale * 3;
    SlabF32 r = slab_setup(rb, ox, oy, oz, ix, iy, iz);
    if (!r.ok) continue;
    float th, tm; slab_threshold(tmax, r.M, th, tm);
    // two boxes; half the time placed so a face passes through a point of the ray (grazing)
    float bx[2][6];
    double e[2];
    for (int c = 0; c < 2; c++) {
      for (int k = 0; k < 3; k++) {
        double a = U(g) * scale, b = U(g) * scale;
        bx[c][k] = (float)std::min(a, b); bx[c][k + 3] = (float)std::max(a, b);
      }
      if (g() % 2) {  // put a face (or edge) exactly where the ray is at some t
        double t = U(g) * scale * 2; int ax = g() % 3; int side = g() % 2;
        double p[3] = {ox + dx * t, oy + dy * t, oz + dz * t};
        bx[c][ax + 3 * side] = (float)p[ax];
        if (g() % 2) { int ax2 = (ax + 1) % 3; bx[c][ax2 + 3 * (g() % 2)] = (float)p[ax2]; }
        if (g() % 2) {  // then move that face by a few times the FP32 error bound, either way
          double org[3] = {ox, oy, oz};
          bx[c][ax + 3 * side] = (float)(p[ax] + U(g) * 8 * 0x1p-23 * (scale + std::fabs(org[ax])));
        }
        for (int k = 0; k < 3; k++) if (bx[c][k] > bx[c][k + 3]) std::swap(bx[c][k], bx[c][k + 3]);
      }
      if (g() % 8 == 0) { int ax = g() % 3; bx[c][ax + 3] = bx[c][ax]; }  // flat box
      double ex = box_entry(bx[c], ox, oy, oz, ix, iy, iz, tmin, tmax);
      e[c] = ex;
      float lo;
      int cl = slab_class(bx[c], r, th, tm, lo);
      total++;
      if (cl == 2) amb++;
      else if ((cl == 1) != (ex < INFINITY)) {
        bad++;
        if (bad < 10) printf("BAD cls=%d exact=%g lo=%g\n", cl, ex, lo);
      }
    }
    // order decision as in closest_impl
    float l0, l1;
    int c0 = slab_class(bx[0], r, th, tm, l0), c1 = slab_class(bx[1], r, th, tm, l1);
    if (c0 == 1 && c1 == 1) {
      orders++;
      float e0 = std::fmax(l0, 0.0f), e1 = std::fmax(l1, 0.0f);
      if (std::fabs(e0 - e1) > r.M) {
        if ((e0 < e1) != (e[0] < e[1])) { bad_order++; if (bad_order < 10) printf("BAD ORDER %g %g vs %g %g\n", e0, e1, e[0], e[1]); }
      } else amb_order++;
    }
  }
  printf("tests %ld bad %ld ambiguous %ld (%.5f)  orders %ld bad %ld ambiguous %ld\n", total, bad, amb, (double)amb / total, orders, bad_order, amb_order);
  return bad || bad_order;
}
