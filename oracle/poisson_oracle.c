/*
 * poisson_oracle.c -- test-infrastructure restatement of paintop's cpu.reference
 * Poisson solver (crates/paintop-cpu/src/poisson.rs) and its per-channel op
 * driver (crates/paintop-cpu/src/poisson_blend.rs:153-246). Only the tests, the
 * smoke entry point and the benchmark's cpu_baseline leg use it, as the checker.
 * Build: -O2 -ffp-contract=off, no fast-math (see oracle/Makefile).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* poisson.rs:73 */
static const double CENTRE_WEIGHT = 4.0;
/* poisson.rs:81, :85 */
static const uint32_t STALL_WINDOW = 16;
static const double STALL_IMPROVEMENT = 1e-12;
/* poisson.rs:329 */
static const float MASK_THRESHOLD = 0.5f;

/* poisson.rs:340-355: Interior iff mask > 0.5 and not on the image border. */
void or_classify_cells(const float *mask, size_t width, size_t height, uint8_t *cells) {
  memset(cells, 0, width * height);
  if (width < 3 || height < 3) return;
  for (size_t row = 1; row < height - 1; ++row)
    for (size_t col = 1; col < width - 1; ++col) {
      size_t idx = row * width + col;
      if (mask[idx] > MASK_THRESHOLD) cells[idx] = 1;
    }
}

/* poisson.rs:275-298: ((left + right) + up) + down, out-of-grid = centre. */
static double neighbour_sum(const double *field, size_t col, size_t row, size_t width,
                            size_t height) {
  double centre = field[row * width + col];
  double left = col == 0 ? centre : field[row * width + col - 1];
  double right = col + 1 == width ? centre : field[row * width + col + 1];
  double up = row == 0 ? centre : field[(row - 1) * width + col];
  double down = row + 1 == height ? centre : field[(row + 1) * width + col];
  return left + right + up + down;
}

/* poisson.rs:364-373: fma(4, -centre, neighbour_sum). */
double or_guidance_laplacian(const double *field, size_t col, size_t row, size_t width,
                             size_t height) {
  double centre = field[row * width + col];
  return fma(CENTRE_WEIGHT, -centre, neighbour_sum(field, col, row, width, height));
}

/* poisson.rs:305-325: running fma sum of squares in row-major order, then sqrt. */
double or_residual_l2(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                      const double *anchor, double lambda, const double *field) {
  double sum_sq = 0.0;
  for (size_t row = 0; row < height; ++row)
    for (size_t col = 0; col < width; ++col) {
      size_t idx = row * width + col;
      if (!cells[idx]) continue;
      double nsum = neighbour_sum(field, col, row, width, height);
      double lhs = fma(lambda + CENTRE_WEIGHT, -field[idx], nsum);
      double r_rhs = fma(lambda, -anchor[idx], rhs[idx]);
      double r = lhs - r_rhs;
      sum_sq = fma(r, r, sum_sq);
    }
  return sqrt(sum_sq);
}

/* poisson.rs:244-265: row-major SOR over Interior cells, in place. */
static void sweep_once(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                       const double *anchor, double lambda, double *field, double denom,
                       double omega) {
  for (size_t row = 0; row < height; ++row)
    for (size_t col = 0; col < width; ++col) {
      size_t idx = row * width + col;
      if (!cells[idx]) continue;
      double nsum = neighbour_sum(field, col, row, width, height);
      double target = fma(lambda, anchor[idx], nsum - rhs[idx]) / denom;
      field[idx] = fma(omega, target - field[idx], field[idx]);
    }
}

/* poisson.rs:168-240. */
void or_solve(size_t width, size_t height, const uint8_t *cells, const double *rhs,
              const double *anchor, double lambda, uint32_t max_iterations, double tolerance,
              double omega, double *field, double *history, or_report *report) {
  size_t n = width * height;
  memcpy(field, anchor, n * sizeof(double));
  double initial = or_residual_l2(width, height, cells, rhs, anchor, lambda, field);
  memset(report, 0, sizeof(*report));
  report->tolerance = tolerance;
  if (initial == 0.0) {
    report->iterations = 0;
    report->converged = 1;
    report->stop_reason = OR_CONVERGED;
    report->final_residual = 0.0;
    report->history_len = 0;
    return;
  }
  double denom = CENTRE_WEIGHT + lambda;
  int converged = 0;
  int stop = OR_MAX_ITERATIONS;
  double last = 1.0;
  uint32_t iterations = 0;
  double best = INFINITY;
  uint32_t since = 0;
  for (uint32_t sweep = 0; sweep < max_iterations; ++sweep) {
    sweep_once(width, height, cells, rhs, anchor, lambda, field, denom, omega);
    iterations = sweep + 1;
    double rel = or_residual_l2(width, height, cells, rhs, anchor, lambda, field) / initial;
    history[sweep] = rel;
    last = rel;
    if (rel <= tolerance) {
      converged = 1;
      stop = OR_CONVERGED;
      break;
    }
    if (rel < best * (1.0 - STALL_IMPROVEMENT)) {
      best = rel;
      since = 0;
    } else {
      since += 1;
      if (since >= STALL_WINDOW) {
        stop = OR_STALLED;
        break;
      }
    }
  }
  report->iterations = iterations;
  report->converged = converged;
  report->stop_reason = stop;
  report->final_residual = last;
  report->history_len = iterations;
}

/* poisson_blend.rs:240-246: (iterations, final_residual) lexicographic max;
 * ties keep the earlier channel. */
static int worse(const or_report *a, const or_report *b) {
  if (b->iterations != a->iterations) return b->iterations > a->iterations;
  return b->final_residual > a->final_residual;
}

/* poisson_blend.rs:168-236. */
void or_blend_channels(const float *source, const float *target, const float *mask,
                       size_t channels, size_t width, size_t height, uint32_t max_iterations,
                       double tolerance, double lambda, double omega, float *out,
                       double *history, or_report *report) {
  size_t count = width * height;
  uint8_t *cells = (uint8_t *)malloc(count ? count : 1);
  double *src = (double *)malloc((count ? count : 1) * sizeof(double));
  double *anchor = (double *)malloc((count ? count : 1) * sizeof(double));
  double *rhs = (double *)malloc((count ? count : 1) * sizeof(double));
  double *field = (double *)malloc((count ? count : 1) * sizeof(double));
  double *hist = (double *)malloc((max_iterations ? max_iterations : 1) * sizeof(double));
  or_classify_cells(mask, width, height, cells);
  int have_worst = 0;
  or_report worst;
  memset(&worst, 0, sizeof(worst));
  for (size_t c = 0; c < channels; ++c) {
    for (size_t p = 0; p < count; ++p) {
      src[p] = (double)source[p * channels + c];
      anchor[p] = (double)target[p * channels + c];
    }
    memset(rhs, 0, count * sizeof(double));
    for (size_t row = 0; row < height; ++row)
      for (size_t col = 0; col < width; ++col) {
        size_t idx = row * width + col;
        if (cells[idx]) rhs[idx] = or_guidance_laplacian(src, col, row, width, height);
      }
    or_report rep;
    or_solve(width, height, cells, rhs, anchor, lambda, max_iterations, tolerance, omega, field,
             hist, &rep);
    for (size_t p = 0; p < count; ++p) out[p * channels + c] = (float)field[p];
    if (!have_worst || worse(&worst, &rep)) {
      worst = rep;
      memcpy(history, hist, rep.history_len * sizeof(double));
      have_worst = 1;
    }
  }
  if (!have_worst) {
    memset(&worst, 0, sizeof(worst));
    worst.converged = 1;
    worst.stop_reason = OR_CONVERGED;
    worst.tolerance = tolerance;
  }
  *report = worst;
  free(cells);
  free(src);
  free(anchor);
  free(rhs);
  free(field);
  free(hist);
}
