/*
 * oracle.h -- CPU restatement of paintop's Poisson / screened-Poisson solver
 * path, used ONLY as test infrastructure: the checker that the tests, the smoke
 * entry point and the benchmark's cpu_baseline leg compare the HIP path against.
 * Nothing in the product (paintop_amd/) includes, links or calls this code.
 *
 * What each part restates (file:line in bobisme/paintop):
 *   or_classify_cells        crates/paintop-cpu/src/poisson.rs:329-355
 *   or_guidance_laplacian    crates/paintop-cpu/src/poisson.rs:275-298, 364-373
 *   or_residual_l2           crates/paintop-cpu/src/poisson.rs:305-325
 *   or_solve                 crates/paintop-cpu/src/poisson.rs:168-265 (+ STALL_* :75-85)
 *   or_blend_channels        crates/paintop-cpu/src/poisson_blend.rs:153-246
 *   or_fbm                   crates/paintop-cpu/src/noise.rs:72-160, 218-240, 660-694
 *   or_outside_scan          crates/paintop-cpu/src/assert.rs:72-85, 178-201, 405-446
 *   or_tight_solve           the system of crates/paintop-cpu/src/poisson.rs:168-240,
 *                            305-325 solved to a tight FP64 residual by
 *                            multigrid-preconditioned conjugate gradients
 *                            (the large-size parity reference)
 *   or_polygon_mask          crates/paintop-cpu/src/mask_polygon.rs:155-195 inside
 *                            test at pixel centres (the workload generator)
 *   or_rbsor_omp             no reference counterpart: the reference has no
 *                            cpu.optimized Poisson kernel, so this OpenMP
 *                            red-black SOR is the stand-in timed as the
 *                            all-cores CPU baseline.
 *
 * Bit-faithfulness recipe: every Rust `a.mul_add(b, c)` is `fma(a, b, c)`, the
 * left-to-right evaluation order of every sum is kept, and the library is built
 * with -ffp-contract=off and without fast-math, so the f64 trajectory is the
 * one IEEE-754 fixes for the Rust source.
 */
#ifndef PAINTOP_ORACLE_H
#define PAINTOP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SolverStopReason (paintop-ir/src/resource.rs:1186-1196). */
enum { OR_CONVERGED = 0, OR_MAX_ITERATIONS = 1, OR_STALLED = 2 };

typedef struct {
  uint32_t iterations;
  int32_t stop_reason;
  int32_t converged;
  int32_t _pad;
  double tolerance;
  double final_residual;
  uint32_t history_len;
  uint32_t _pad2;
} or_report;

/* cells: 1 = Interior, 0 = Boundary. */
void or_classify_cells(const float *mask, size_t width, size_t height, uint8_t *cells);

double or_guidance_laplacian(const double *field, size_t col, size_t row, size_t width,
                             size_t height);

double or_residual_l2(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                      const double *anchor, double lambda, const double *field);

/* field: out, width*height. history: out, capacity max_iterations. */
void or_solve(size_t width, size_t height, const uint8_t *cells, const double *rhs,
              const double *anchor, double lambda, uint32_t max_iterations, double tolerance,
              double omega, double *field, double *history, or_report *report);

/* Full op semantics of blend_channels: interleaved f32 in, interleaved f32 out,
 * the worst channel's report (and its history, capacity max_iterations). */
void or_blend_channels(const float *source, const float *target, const float *mask,
                       size_t channels, size_t width, size_t height, uint32_t max_iterations,
                       double tolerance, double lambda, double omega, float *out,
                       double *history, or_report *report);

/* field.fbm@1 sample grid: value at pixel centre (x+0.5, y+0.5) * frequency/64. */
uint64_t or_mix64(uint64_t z);
double or_lattice_value(int64_t gx, int64_t gy, uint64_t seed);
double or_value_noise(double x, double y, uint64_t seed);
double or_fbm_eval(double x, double y, uint64_t seed, uint32_t octaves, double lacunarity,
                   double gain);
void or_fbm(size_t width, size_t height, double frequency, uint64_t seed, uint32_t octaves,
            double lacunarity, double gain, float *out);

/* The all-cores CPU stand-in: red-black SOR in f64 on interior cells, with the
 * reference's per-sweep relative residual and stop rules (stall_window 16 is
 * the reference's; 0 turns the stall guard off, for timing fixed sweep counts).
 * omega_tab (NULL: the fixed omega) gives the factor of half-sweep h, red
 * first. threads <= 0 means the OpenMP default. Returns the thread count used. */
int or_rbsor_omp(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                 const double *anchor, double lambda, uint32_t max_iterations, double tolerance,
                 double omega, const double *omega_tab, int threads, uint32_t stall_window,
                 double *field, double *history, or_report *report);

/* Tight FP64 solve of the system poisson.rs:168-240 defines (Interior cells:
 * nsum(u) - (4 + lambda) u = rhs - lambda anchor; u = anchor elsewhere), by
 * conjugate gradients preconditioned with a Galerkin-aggregation multigrid
 * V-cycle, restarted from the true residual until the relative residual
 * ||r(u)|| / ||r(anchor)|| (or_residual_l2, the reference's formula and order)
 * is <= tolerance or stops halving (the FP64 floor). alpha: coarse-correction
 * factor in (0, 2) (<= 0: 1). threads <= 0: the OpenMP default. Returns the
 * thread count used; the result is bit-identical for any thread count. */
typedef struct {
  double r0;             /* ||r(anchor)|| */
  double final_residual; /* ||r(field)|| / r0 */
  uint32_t iterations;   /* CG iterations, all restarts */
  uint32_t restarts;
  uint32_t levels;
  int32_t converged;
  int32_t threads;
  int32_t _pad;
} or_tight_report;

int or_tight_solve(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                   const double *anchor, double lambda, double tolerance,
                   uint32_t max_iterations, double alpha, int threads, double *field,
                   or_tight_report *report);

/* mask.polygon@1's inside test (mask_polygon.rs:161-195: half-open ray cast to
 * +x, cross_x = fma(t, x1 - x0, x0), nonzero winding or even-odd) at each pixel
 * centre (x + 0.5, y + 0.5): 1 inside, 0 outside. points_xy = x0,y0,x1,y1,... */
void or_polygon_mask(size_t width, size_t height, const double *points_xy, size_t n_points,
                     int even_odd, float *out);

/* out[i] += (float)(sigma * z_i), z_i = sqrt(-2 ln u1) cos(2 pi u2) with
 * u1 = ((mix64(seed + G (2i+1)) >> 11) + 1) 2^-53, u2 = (mix64(seed + G (2i+2)) >> 11) 2^-53,
 * G = 0x9e3779b97f4a7c15 (SplitMix64's increment, mix64 = noise.rs:72-76). */
void or_gauss_noise(float *out, size_t n, uint64_t seed, double sigma);

/* assert.no_change_outside_mask@1's scan (crates/paintop-cpu/src/assert.rs:405-446).
 * Pixel indices are row-major; (x, y) = (i % width, i / width) (assert.rs:196-201). */
#define OR_MAX_LOCATIONS 256 /* assert.rs:72 */
typedef struct or_outside_result {
  uint64_t changed_outside;
  double max_outside;
  int64_t worst_index; /* -1: no leak */
  double worst_delta;
  uint32_t n_locations;
  uint32_t _pad;
  int64_t locations[OR_MAX_LOCATIONS];
} or_outside_result;

void or_outside_scan(const float *before, const float *after, const float *mask, size_t width,
                     size_t height, size_t channels, int decode, double outside_threshold,
                     double coverage_epsilon, or_outside_result *out);

#ifdef __cplusplus
}
#endif
#endif
