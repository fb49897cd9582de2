/*
 * rbsor_omp.c -- the all-cores CPU baseline. paintop has no cpu.optimized (or
 * Rayon) Poisson kernel (crates/paintop-cpu/src/optimized/bench.rs:36-52 lists
 * only pointwise ops), so this OpenMP red-black SOR in f64 stands in for it. It
 * keeps the reference's stopping rules (poisson.rs:168-240: relative residual
 * against the anchor-start residual, per-sweep history, 16-sweep stall window,
 * iteration cap) and its update formula (poisson.rs:256-262), only the visiting
 * order is red-black so the rows can be split across threads, and the SOR
 * factor may follow a per-half-sweep table (the GPU path's schedule). The residual sum
 * is taken per row and the row partials are added in row order, so a rerun is
 * bit-identical for any thread count.
 * Test infrastructure: used only by the tests and the cpu_baseline leg.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline double nsum_at(const double *f, size_t idx, size_t w) {
  return f[idx - 1] + f[idx + 1] + f[idx - w] + f[idx + w];
}

static double residual_norm(size_t width, size_t height, const uint8_t *cells,
                            const double *bprime, double denom, const double *field,
                            double *row_part, int nthreads) {
  long h1 = (long)height - 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long row = 1; row < h1; ++row) {
    double s = 0.0;
    for (size_t col = 1; col + 1 < width; ++col) {
      size_t idx = (size_t)row * width + col;
      if (!cells[idx]) continue;
      double r = fma(denom, -field[idx], nsum_at(field, idx, width)) - bprime[idx];
      s = fma(r, r, s);
    }
    row_part[row] = s;
  }
  double acc = 0.0;
  for (long row = 1; row < h1; ++row) acc += row_part[row];
  return sqrt(acc);
}

int or_rbsor_omp(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                 const double *anchor, double lambda, uint32_t max_iterations, double tolerance,
                 double omega, const double *omega_tab, int threads, uint32_t stall_window,
                 double *field, double *history, or_report *report) {
  size_t n = width * height;
  int nthreads = threads > 0 ? threads : omp_get_max_threads();
  memcpy(field, anchor, n * sizeof(double));
  memset(report, 0, sizeof(*report));
  report->tolerance = tolerance;
  /* Interior cells are never on the border (classify_cells), so every
   * interior cell has four in-grid neighbours. Rows 1..h-2 only. */
  double *row_part = (double *)calloc(height ? height : 1, sizeof(double));
  double *bprime = (double *)malloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; ++i) bprime[i] = fma(lambda, -anchor[i], rhs[i]);
  double denom = 4.0 + lambda;

  double initial;
  if (width < 3 || height < 3) {
    initial = 0.0;
  } else {
    initial = residual_norm(width, height, cells, bprime, denom, field, row_part, nthreads);
  }
  if (initial == 0.0) {
    report->converged = 1;
    report->stop_reason = OR_CONVERGED;
    free(row_part);
    free(bprime);
    return nthreads;
  }
  int converged = 0, stop = OR_MAX_ITERATIONS;
  double last = 1.0, best = INFINITY;
  uint32_t iterations = 0, since = 0;
  for (uint32_t sweep = 0; sweep < max_iterations; ++sweep) {
    for (int colour = 0; colour < 2; ++colour) {
      const double w = omega_tab ? omega_tab[2 * (size_t)sweep + (size_t)colour] : omega;
      long h1 = (long)height - 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (long row = 1; row < h1; ++row) {
        size_t start = 1 + (((size_t)row + 1 + (size_t)colour) & 1);
        for (size_t col = start; col + 1 < width; col += 2) {
          size_t idx = (size_t)row * width + col;
          if (!cells[idx]) continue;
          double target = (nsum_at(field, idx, width) - bprime[idx]) / denom;
          field[idx] = fma(w, target - field[idx], field[idx]);
        }
      }
    }
    iterations = sweep + 1;
    double r = residual_norm(width, height, cells, bprime, denom, field, row_part, nthreads);
    double rel = r / initial;
    history[sweep] = rel;
    last = rel;
    if (rel <= tolerance) {
      converged = 1;
      stop = OR_CONVERGED;
      break;
    }
    if (rel < best * (1.0 - 1e-12)) {
      best = rel;
      since = 0;
    } else if (++since >= stall_window && stall_window > 0) {
      stop = OR_STALLED;
      break;
    }
  }
  report->iterations = iterations;
  report->converged = converged;
  report->stop_reason = stop;
  report->final_residual = last;
  report->history_len = iterations;
  free(row_part);
  free(bprime);
  return nthreads;
}
