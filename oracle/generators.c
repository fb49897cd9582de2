/*
 * generators.c -- test-infrastructure restatement of paintop's field.fbm@1
 * value-noise generator (crates/paintop-cpu/src/noise.rs), the pixel-centre
 * polygon mask (mask.polygon@1's inside test, mask_polygon.rs:155-195) and the
 * seeded Gaussian noise of the screened workload. Used by the tests as the
 * checker for the device-side generators, by the fixture scripts to rebuild the
 * full-size workloads on the CPU, and by the benchmark's cpu_baseline leg.
 * Build: -O2 -ffp-contract=off, no fast-math.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

/* noise.rs:72-76: the SplitMix64 finalizer. */
uint64_t or_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* noise.rs:86-109: fold (gx, gy) into the seed, top 24 bits -> [-1, 1].
 * Parity unpinned for the two fold multipliers: the reference holds no fbm
 * golden values, and its fbm tests (range, position purity, value noise equal
 * to the lattice value at integer points) hold for any odd multipliers. The
 * finalizer (or_mix64) is pinned by SplitMix64's published vectors. */
double or_lattice_value(int64_t gx, int64_t gy, uint64_t seed) {
  uint64_t ux = (uint64_t)gx;
  uint64_t uy = (uint64_t)gy;
  uint64_t hx = or_mix64(seed ^ (ux * 0x9e3779b97f4a7c15ULL));
  uint64_t h = or_mix64(hx ^ (uy * 0xc2b2ae3d27d4eb4fULL));
  uint64_t top24 = h >> 40;
  double unit = (double)(uint32_t)top24 / (double)(1u << 24);
  return fma(unit, 2.0, -1.0);
}

/* noise.rs:113-115: ((t*t)*t) * fma(t, fma(t, 6, -15), 10). */
static double fade(double t) { return t * t * t * fma(t, fma(t, 6.0, -15.0), 10.0); }

/* noise.rs:118-120 */
static double lerp(double a, double b, double t) { return fma(b - a, t, a); }

/* noise.rs:127-160 */
double or_value_noise(double x, double y, uint64_t seed) {
  double x0 = floor(x);
  double y0 = floor(y);
  int64_t gx0 = (int64_t)x0;
  int64_t gy0 = (int64_t)y0;
  double tx = fade(x - x0);
  double ty = fade(y - y0);
  double v00 = or_lattice_value(gx0, gy0, seed);
  double v10 = or_lattice_value(gx0 + 1, gy0, seed);
  double v01 = or_lattice_value(gx0, gy0 + 1, seed);
  double v11 = or_lattice_value(gx0 + 1, gy0 + 1, seed);
  double top = lerp(v00, v10, tx);
  double bottom = lerp(v01, v11, tx);
  return lerp(top, bottom, ty);
}

/* noise.rs:218-240 */
double or_fbm_eval(double x, double y, uint64_t seed, uint32_t octaves, double lacunarity,
                   double gain) {
  double frequency = 1.0, amplitude = 1.0, sum = 0.0, total = 0.0;
  for (uint32_t o = 0; o < octaves; ++o) {
    uint64_t octave_seed = or_mix64(seed ^ ((uint64_t)o + 1u));
    sum = fma(amplitude, or_value_noise(x * frequency, y * frequency, octave_seed), sum);
    total += amplitude;
    frequency *= lacunarity;
    amplitude *= gain;
  }
  return total > 0.0 ? sum / total : 0.0;
}

/* noise.rs:660-694 with the pixel-centre convention of
 * paintop-ir/src/resource.rs:394-397 and FREQUENCY_PIXELS = 64 (noise.rs:646). */
void or_fbm(size_t width, size_t height, double frequency, uint64_t seed, uint32_t octaves,
            double lacunarity, double gain, float *out) {
  double scale = frequency / 64.0;
  /* pixels are independent: rows in parallel give the same bits */
#pragma omp parallel for schedule(dynamic, 16)
  for (long yy = 0; yy < (long)height; ++yy)
    for (size_t x = 0; x < width; ++x) {
      size_t y = (size_t)yy;
      double px = (double)(uint32_t)x + 0.5;
      double py = (double)(uint32_t)y + 0.5;
      out[y * width + x] =
          (float)or_fbm_eval(px * scale, py * scale, seed, octaves, lacunarity, gain);
    }
}

/* mask_polygon.rs:161-195: half-open ray cast to +x at each pixel centre. */
void or_polygon_mask(size_t width, size_t height, const double *points_xy, size_t n_points,
                     int even_odd, float *out) {
#pragma omp parallel for schedule(static)
  for (long yy = 0; yy < (long)height; ++yy)
    for (size_t xx = 0; xx < width; ++xx) {
      double x = (double)xx + 0.5, y = (double)yy + 0.5;
      int winding = 0, parity = 0;
      if (n_points >= 3)
        for (size_t i = 0; i < n_points; ++i) {
          size_t k = i + 1 == n_points ? 0 : i + 1;
          double x0 = points_xy[2 * i], y0 = points_xy[2 * i + 1];
          double x1 = points_xy[2 * k], y1 = points_xy[2 * k + 1];
          int below0 = y0 <= y, below1 = y1 <= y;
          if (below0 == below1) continue;
          double t = (y - y0) / (y1 - y0);
          double cross_x = fma(t, x1 - x0, x0);
          if (cross_x > x) {
            parity = !parity;
            winding += below0 ? 1 : -1;
          }
        }
      int inside = n_points >= 3 && (even_odd ? parity : winding != 0);
      out[(size_t)yy * width + xx] = inside ? 1.0f : 0.0f;
    }
}

/* Box-Muller on two SplitMix64 draws keyed by (seed, i). */
void or_gauss_noise(float *out, size_t n, uint64_t seed, double sigma) {
  const uint64_t G = 0x9e3779b97f4a7c15ULL;
#pragma omp parallel for schedule(static)
  for (long ii = 0; ii < (long)n; ++ii) {
    uint64_t i = (uint64_t)ii;
    uint64_t a = or_mix64(seed + G * (2 * i + 1));
    uint64_t b = or_mix64(seed + G * (2 * i + 2));
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    out[i] = out[i] + (float)(sigma * z);
  }
}
