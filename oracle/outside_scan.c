/* outside_scan.c -- CPU restatement of assert.no_change_outside_mask@1's scan
 * (test infrastructure only; the tests use it as the checker of the device
 * scan, never as the product).
 *
 * Restates crates/paintop-cpu/src/assert.rs:
 *   scan_outside      :405-446  (outside = coverage <= epsilon, per-pixel max
 *                               channel delta, strict > for every running max)
 *   srgb_decode       :74-85    (f32; powf as Rust's f32::powf calls it)
 *   to_space          :178-182
 *   pixel_coord       :196-201  (x = i % width, y = i / width)
 *   MAX_LOCATIONS     :72       (256)
 */
#include <math.h>

#include "oracle.h"

static float srgb_decode(float c) {
  if (c <= 0.04045f) return c / 12.92f;
  return powf((c + 0.055f) / 1.055f, 2.4f);
}

static float to_space(float s, int decode) { return decode ? srgb_decode(s) : s; }

void or_outside_scan(const float *before, const float *after, const float *mask, size_t width,
                     size_t height, size_t channels, int decode, double outside_threshold,
                     double coverage_epsilon, or_outside_result *out) {
  out->changed_outside = 0;
  out->max_outside = 0.0;
  out->worst_index = -1;
  out->worst_delta = 0.0;
  out->n_locations = 0;
  if (channels == 0) return;
  const size_t n = width * height;
  for (size_t i = 0; i < n; ++i) {
    const double coverage = (double)mask[i];
    if (coverage > coverage_epsilon) continue;
    double pixel_delta = 0.0;
    for (size_t c = 0; c < channels; ++c) {
      const float d = fabsf(to_space(after[i * channels + c], decode) -
                            to_space(before[i * channels + c], decode));
      if ((double)d > pixel_delta) pixel_delta = (double)d;
    }
    if (pixel_delta > out->max_outside) out->max_outside = pixel_delta;
    if (pixel_delta > outside_threshold) {
      out->changed_outside += 1;
      if (out->n_locations < OR_MAX_LOCATIONS) out->locations[out->n_locations++] = (int64_t)i;
      if (out->worst_index < 0 || pixel_delta > out->worst_delta) {
        out->worst_index = (int64_t)i;
        out->worst_delta = pixel_delta;
      }
    }
  }
}
