/*
 * tight_solve.c -- test infrastructure: an FP64 solve of the Poisson /
 * screened-Poisson system that crates/paintop-cpu/src/poisson.rs:168-240
 * defines, driven to a tight relative residual (1e-12) at full image sizes,
 * where the lexicographic restatement (or_solve) would need 10^5..10^6 sweeps.
 * The tests and the benchmark compare the HIP path's results against it; the
 * product never links it.
 *
 * The system (poisson.rs:305-325): at Interior cells
 *     nsum(u) - (4 + lambda) u = rhs - lambda * anchor,
 * u = anchor at every other cell (Dirichlet). With u = anchor + e and e = 0
 * off the interior this is the SPD problem
 *     L e = f,  L = (4 + lambda) I - N,  f = r(anchor),
 * N the adjacency of interior cells. It is solved by conjugate gradients
 * preconditioned with one multigrid V-cycle:
 *   - levels by 2x2 aggregation: a coarse cell is the sum of its active fine
 *     cells; prolongation P copies a coarse value to the active cells of its
 *     block, restriction is P^T, and the coarse operator is the Galerkin
 *     product P^T L P, which for piecewise-constant P is again a 5-point
 *     operator: diag = sum of the block's fine diagonals - 2 x (couplings
 *     inside the block), east/south coupling = sum of the fine couplings that
 *     cross the block face;
 *   - red/black Gauss-Seidel smoothing, red then black before the coarse
 *     correction and black then red after it, so the V-cycle is a symmetric
 *     operator (the adjoint of the pre-smoother is the post-smoother);
 *   - the coarse correction scaled by a fixed factor (an over-correction that
 *     compensates the energy piecewise-constant interpolation loses; any
 *     factor in (0, 2) keeps the preconditioner SPD);
 *   - a dense Cholesky factorisation on the coarsest level.
 * The preconditioner only changes the speed: the stopping test is the true
 * FP64 residual of u, recomputed from scratch, in the reference's own formula
 * and running-sum order (or_residual_l2). Every reduction sums per-row
 * partials in row order, so a rerun is bit-identical for any thread count.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  long w, h;
  uint8_t *act;
  double *D, *wE, *wS; /* L = D - couplings; wE[i]: i <-> i+1, wS[i]: i <-> i+w */
  double *x, *b, *r;
  double *row; /* per-row partials */
} Level;

#define MAX_LEVELS 24
#define DENSE_MAX 1600

typedef struct {
  int nlev;
  Level lv[MAX_LEVELS];
  /* dense coarsest level */
  long nd;
  long *didx; /* cell -> dense index, -1 if inactive */
  double *chol;
  double *dv;
  double alpha; /* coarse-correction factor */
  int nu;       /* smoothing sweeps before and after */
  int threads;
} MG;

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

/* One red/black Gauss-Seidel half-sweep of colour c on level L: x_a =
 * (b_a + sum of couplings x neighbours) / D_a at active cells with (x+y)&1 == c. */
static void gs_colour(Level *L, int c, int nth) {
  const long w = L->w, h = L->h;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < h; ++y) {
    for (long x = (y + c) & 1; x < w; x += 2) {
      const long i = y * w + x;
      if (!L->act[i]) continue;
      double s = L->b[i];
      if (x > 0) s += L->wE[i - 1] * L->x[i - 1];
      if (x + 1 < w) s += L->wE[i] * L->x[i + 1];
      if (y > 0) s += L->wS[i - w] * L->x[i - w];
      if (y + 1 < h) s += L->wS[i] * L->x[i + w];
      L->x[i] = s / L->D[i];
    }
  }
}

/* r = b - L x on level L. */
static void level_residual(Level *L, int nth) {
  const long w = L->w, h = L->h;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < h; ++y) {
    for (long x = 0; x < w; ++x) {
      const long i = y * w + x;
      if (!L->act[i]) {
        L->r[i] = 0.0;
        continue;
      }
      double s = L->D[i] * L->x[i];
      if (x > 0) s -= L->wE[i - 1] * L->x[i - 1];
      if (x + 1 < w) s -= L->wE[i] * L->x[i + 1];
      if (y > 0) s -= L->wS[i - w] * L->x[i - w];
      if (y + 1 < h) s -= L->wS[i] * L->x[i + w];
      L->r[i] = L->b[i] - s;
    }
  }
}

/* Galerkin coarsening of F into C (2x2 aggregation). */
static void coarsen(const Level *F, Level *C, int nth) {
  const long W = (F->w + 1) / 2, H = (F->h + 1) / 2, fw = F->w, fh = F->h;
  C->w = W;
  C->h = H;
  const size_t n = (size_t)W * H;
  C->act = (uint8_t *)xcalloc(n, 1);
  C->D = (double *)xcalloc(n, sizeof(double));
  C->wE = (double *)xcalloc(n, sizeof(double));
  C->wS = (double *)xcalloc(n, sizeof(double));
  C->x = (double *)xcalloc(n, sizeof(double));
  C->b = (double *)xcalloc(n, sizeof(double));
  C->r = (double *)xcalloc(n, sizeof(double));
  C->row = (double *)xcalloc((size_t)H, sizeof(double));
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long Y = 0; Y < H; ++Y)
    for (long X = 0; X < W; ++X) {
      const long a = Y * W + X;
      double d = 0.0, inner = 0.0, ce = 0.0, cs = 0.0;
      int any = 0;
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const long fx = 2 * X + dx, fy = 2 * Y + dy;
          if (fx >= fw || fy >= fh) continue;
          const long i = fy * fw + fx;
          if (!F->act[i]) continue;
          any = 1;
          d += F->D[i];
          /* couplings inside the block: east of column 2X, south of row 2Y */
          if (dx == 0 && fx + 1 < fw) inner += F->wE[i];
          if (dy == 0 && fy + 1 < fh) inner += F->wS[i];
          /* couplings across the east / south face */
          if (dx == 1 && fx + 1 < fw) ce += F->wE[i];
          if (dy == 1 && fy + 1 < fh) cs += F->wS[i];
        }
      C->act[a] = (uint8_t)any;
      C->D[a] = any ? d - 2.0 * inner : 1.0;
      C->wE[a] = ce;
      C->wS[a] = cs;
    }
}

static void restrict_to(const Level *F, Level *C, int nth) {
  const long W = C->w, H = C->h, fw = F->w, fh = F->h;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long Y = 0; Y < H; ++Y)
    for (long X = 0; X < W; ++X) {
      double s = 0.0;
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const long fx = 2 * X + dx, fy = 2 * Y + dy;
          if (fx < fw && fy < fh) s += F->r[fy * fw + fx];
        }
      C->b[Y * W + X] = s;
      C->x[Y * W + X] = 0.0;
    }
}

static void prolong_add(Level *F, const Level *C, double alpha, int nth) {
  const long fw = F->w, fh = F->h, W = C->w;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < fh; ++y)
    for (long x = 0; x < fw; ++x) {
      const long i = y * fw + x;
      if (F->act[i]) F->x[i] += alpha * C->x[(y / 2) * W + x / 2];
    }
}

static void dense_setup(MG *M) {
  Level *L = &M->lv[M->nlev - 1];
  const long n = L->w * L->h;
  M->didx = (long *)xcalloc((size_t)n, sizeof(long));
  long m = 0;
  for (long i = 0; i < n; ++i) M->didx[i] = L->act[i] ? m++ : -1;
  M->nd = m;
  M->chol = (double *)xcalloc((size_t)(m * m), sizeof(double));
  M->dv = (double *)xcalloc((size_t)m, sizeof(double));
  double *A = M->chol;
  for (long y = 0; y < L->h; ++y)
    for (long x = 0; x < L->w; ++x) {
      const long i = y * L->w + x, k = M->didx[i];
      if (k < 0) continue;
      A[k * m + k] = L->D[i];
      if (x + 1 < L->w && M->didx[i + 1] >= 0) {
        A[k * m + M->didx[i + 1]] = -L->wE[i];
        A[M->didx[i + 1] * m + k] = -L->wE[i];
      }
      if (y + 1 < L->h && M->didx[i + L->w] >= 0) {
        A[k * m + M->didx[i + L->w]] = -L->wS[i];
        A[M->didx[i + L->w] * m + k] = -L->wS[i];
      }
    }
  /* in-place Cholesky, lower triangle */
  for (long j = 0; j < m; ++j) {
    double s = A[j * m + j];
    for (long k = 0; k < j; ++k) s -= A[j * m + k] * A[j * m + k];
    if (!(s > 0.0)) abort(); /* the coarse operator is SPD */
    const double ljj = sqrt(s);
    A[j * m + j] = ljj;
    for (long i = j + 1; i < m; ++i) {
      double t = A[i * m + j];
      for (long k = 0; k < j; ++k) t -= A[i * m + k] * A[j * m + k];
      A[i * m + j] = t / ljj;
    }
  }
}

static void dense_solve(MG *M) {
  Level *L = &M->lv[M->nlev - 1];
  const long n = L->w * L->h, m = M->nd;
  const double *A = M->chol;
  double *v = M->dv;
  for (long i = 0; i < n; ++i)
    if (M->didx[i] >= 0) v[M->didx[i]] = L->b[i];
  for (long i = 0; i < m; ++i) {
    double s = v[i];
    for (long k = 0; k < i; ++k) s -= A[i * m + k] * v[k];
    v[i] = s / A[i * m + i];
  }
  for (long i = m - 1; i >= 0; --i) {
    double s = v[i];
    for (long k = i + 1; k < m; ++k) s -= A[k * m + i] * v[k];
    v[i] = s / A[i * m + i];
  }
  for (long i = 0; i < n; ++i) L->x[i] = M->didx[i] >= 0 ? v[M->didx[i]] : 0.0;
}

static void vcycle(MG *M, int l) {
  Level *L = &M->lv[l];
  const int nth = M->threads;
  if (l == M->nlev - 1) {
    dense_solve(M);
    return;
  }
  const size_t n = (size_t)L->w * L->h;
  memset(L->x, 0, n * sizeof(double));
  for (int s = 0; s < M->nu; ++s) {
    gs_colour(L, 0, nth);
    gs_colour(L, 1, nth);
  }
  level_residual(L, nth);
  restrict_to(L, &M->lv[l + 1], nth);
  vcycle(M, l + 1);
  prolong_add(L, &M->lv[l + 1], M->alpha, nth);
  for (int s = 0; s < M->nu; ++s) {
    gs_colour(L, 1, nth);
    gs_colour(L, 0, nth);
  }
}

/* Dot product over active cells with per-row partials summed in row order. */
static double dot_rows(const Level *L, const double *a, const double *b, int nth) {
  const long w = L->w, h = L->h;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < h; ++y) {
    double s = 0.0;
    for (long x = 0; x < w; ++x) {
      const long i = y * w + x;
      if (L->act[i]) s = fma(a[i], b[i], s);
    }
    L->row[y] = s;
  }
  double t = 0.0;
  for (long y = 0; y < h; ++y) t += L->row[y];
  return t;
}

/* q = L p on the fine level (p = 0 off the interior). */
static void apply_fine(const Level *L, const double *p, double *q, int nth) {
  const long w = L->w, h = L->h;
  const double dg = L->D[0];
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < h; ++y)
    for (long x = 0; x < w; ++x) {
      const long i = y * w + x;
      if (!L->act[i]) {
        q[i] = 0.0;
        continue;
      }
      /* interior cells are never on the image border (poisson.rs:340-355) */
      q[i] = dg * p[i] - (((p[i - 1] + p[i + 1]) + p[i - w]) + p[i + w]);
    }
}

/* r(u) = nsum(u) - (4 + lambda) u - (rhs - lambda anchor) at interior cells
 * (the formula of poisson.rs:305-325), 0 elsewhere. */
static void true_residual(size_t w, size_t h, const uint8_t *cells, const double *rhs,
                          const double *anchor, double lambda, const double *u, double *r,
                          int nth) {
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long y = 0; y < (long)h; ++y)
    for (size_t x = 0; x < w; ++x) {
      const size_t i = (size_t)y * w + x;
      if (!cells[i]) {
        r[i] = 0.0;
        continue;
      }
      const double nsum = ((u[i - 1] + u[i + 1]) + u[i - w]) + u[i + w];
      const double lhs = fma(lambda + 4.0, -u[i], nsum);
      r[i] = lhs - fma(lambda, -anchor[i], rhs[i]);
    }
}

int or_tight_solve(size_t width, size_t height, const uint8_t *cells, const double *rhs,
                   const double *anchor, double lambda, double tolerance,
                   uint32_t max_iterations, double alpha, int threads, double *field,
                   or_tight_report *report) {
  const int nth = threads > 0 ? threads : omp_get_max_threads();
  const size_t n = width * height;
  memset(report, 0, sizeof(*report));
  memcpy(field, anchor, n * sizeof(double));
  report->r0 = or_residual_l2(width, height, cells, rhs, anchor, lambda, field);
  report->final_residual = 0.0;
  report->threads = nth;
  if (report->r0 == 0.0 || !isfinite(report->r0)) {
    report->final_residual = report->r0 == 0.0 ? 0.0 : NAN;
    report->converged = report->r0 == 0.0;
    return nth;
  }

  MG M;
  memset(&M, 0, sizeof(M));
  M.alpha = alpha > 0.0 ? alpha : 1.0;
  M.nu = 2;
  M.threads = nth;
  Level *F = &M.lv[0];
  F->w = (long)width;
  F->h = (long)height;
  F->act = (uint8_t *)xcalloc(n, 1);
  memcpy(F->act, cells, n);
  F->D = (double *)xcalloc(n, sizeof(double));
  F->wE = (double *)xcalloc(n, sizeof(double));
  F->wS = (double *)xcalloc(n, sizeof(double));
  F->x = (double *)xcalloc(n, sizeof(double));
  F->b = (double *)xcalloc(n, sizeof(double));
  F->r = (double *)xcalloc(n, sizeof(double));
  F->row = (double *)xcalloc(height, sizeof(double));
  long nact = 0;
#pragma omp parallel for num_threads(nth) schedule(static) reduction(+ : nact)
  for (long y = 0; y < (long)height; ++y)
    for (size_t x = 0; x < width; ++x) {
      const size_t i = (size_t)y * width + x;
      F->D[i] = 4.0 + lambda;
      if (!cells[i]) continue;
      ++nact;
      if (x + 1 < width && cells[i + 1]) F->wE[i] = 1.0;
      if ((size_t)y + 1 < height && cells[i + width]) F->wS[i] = 1.0;
    }
  M.nlev = 1;
  long act_l = nact;
  while (act_l > DENSE_MAX && M.nlev < MAX_LEVELS) {
    Level *C = &M.lv[M.nlev];
    coarsen(&M.lv[M.nlev - 1], C, nth);
    ++M.nlev;
    act_l = 0;
    for (long i = 0; i < C->w * C->h; ++i) act_l += C->act[i];
  }
  dense_setup(&M);
  report->levels = M.nlev;

  double *e = (double *)xcalloc(n, sizeof(double));
  double *r = (double *)xcalloc(n, sizeof(double));
  double *p = (double *)xcalloc(n, sizeof(double));
  double *q = (double *)xcalloc(n, sizeof(double));
  const double r0 = report->r0;
  uint32_t it = 0;
  double prev = INFINITY;
  /* outer restarts: each solves L e = r(u) from the true residual of the
   * current field, then u += e */
  for (uint32_t restart = 0; restart < 64 && it < max_iterations; ++restart) {
    const double rel = or_residual_l2(width, height, cells, rhs, anchor, lambda, field) / r0;
    report->final_residual = rel;
    if (rel <= tolerance) break;
    /* the FP64 floor: a restart no longer halves the true residual */
    if (!(rel < 0.5 * prev)) break;
    prev = rel;
    report->restarts = restart + 1;
    true_residual(width, height, cells, rhs, anchor, lambda, field, r, nth);
    memset(e, 0, n * sizeof(double));
    memcpy(F->b, r, n * sizeof(double));
    vcycle(&M, 0);
    memcpy(p, F->x, n * sizeof(double));
    double rz = dot_rows(F, r, F->x, nth);
    const double stop = 0.05 * tolerance * r0; /* inner target below the outer one */
    for (; it < max_iterations; ++it) {
      apply_fine(F, p, q, nth);
      const double pq = dot_rows(F, p, q, nth);
      if (!(pq > 0.0)) break;
      const double a = rz / pq;
#pragma omp parallel for num_threads(nth) schedule(static)
      for (long i = 0; i < (long)n; ++i) {
        e[i] = fma(a, p[i], e[i]);
        r[i] = fma(-a, q[i], r[i]);
      }
      const double rn = sqrt(dot_rows(F, r, r, nth));
      if (rn <= stop) {
        ++it;
        break;
      }
      memcpy(F->b, r, n * sizeof(double));
      vcycle(&M, 0);
      const double rz2 = dot_rows(F, r, F->x, nth);
      const double beta = rz2 / rz;
      rz = rz2;
#pragma omp parallel for num_threads(nth) schedule(static)
      for (long i = 0; i < (long)n; ++i) p[i] = fma(beta, p[i], F->x[i]);
    }
#pragma omp parallel for num_threads(nth) schedule(static)
    for (long i = 0; i < (long)n; ++i)
      if (cells[i]) field[i] += e[i];
  }
  report->iterations = it;
  report->final_residual = or_residual_l2(width, height, cells, rhs, anchor, lambda, field) / r0;
  report->converged = report->final_residual <= tolerance;

  free(e);
  free(r);
  free(p);
  free(q);
  for (int l = 0; l < M.nlev; ++l) {
    Level *L = &M.lv[l];
    free(L->act);
    free(L->D);
    free(L->wE);
    free(L->wS);
    free(L->x);
    free(L->b);
    free(L->r);
    free(L->row);
  }
  free(M.didx);
  free(M.chol);
  free(M.dv);
  return nth;
}
