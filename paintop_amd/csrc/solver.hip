// solver.hip -- host orchestration of the HIP Poisson solve and the C ABI
// (include/paintop_hip.h).
//
// Per op call (PoissonBlend::compute, poisson_blend.rs:535-600, and
// ScreenedPoisson::compute, screened_poisson.rs:362-430, after parameter
// resolution):
//   1. classify the mask once (poisson.rs:340-355) and find the bounding box
//      of the interior cells;
//   2. per channel (blend_channels, poisson_blend.rs:168-236), assemble the
//      FP64 system in a padded plane around that box; x starts at the anchor
//      (poisson.rs:169-170);
//   3. r0 = ||r(anchor)|| in FP64 (poisson.rs:172); r0 == 0 ends the channel
//      at iteration 0 (poisson.rs:177-189);
//   4. FP32 red-black SOR sweeps on the correction equation, K half-sweeps per
//      launch, with the per-sweep relative residual and the stop rules run on
//      the device after every launch; an FP64 refinement step (x += e,
//      g = -r(x)) whenever the inner residual has fallen 1e3-fold, and to
//      confirm every stop with the FP64 residual;
//   5. write interior cells back as f32; every other sample is the target's.
// All channels of an image run in the same launches. The report is the worst
// channel's (poisson_blend.rs:240-246).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/paintop_hip.h"
#include "host_common.h"
#include "ph_internal.h"
#include "sor_tuning.h"

namespace ph {

template <int K, int PF>
__global__ void k_sor(const PhTile *, int, const PhInst *, const PhCtl *, double *, int);
template <int K>
__global__ void k_sor_control(const PhInst *, PhCtl *, const double *, int, int);
__global__ void k_classify_spans(const PhSpan *, int, const PhItemIO *, int2 *);
__global__ void k_bbox_spans(const PhSpan *, const int2 *, int *);
__global__ void k_tile_occupancy_all(const PhItemGeo *, const int *, const int2 *, int, int, int,
                                     int *);
__global__ void k_coarse_count_all(const PhItemGeo *, const PhSpan *, int, uint8_t *,
                                   uint64_t *);
__global__ void k_coarse_hash_items(const PhSpan *, const uint64_t *, uint64_t *);
template <int CC>
__global__ void k_assemble_spans(const PhInst *, const PhSpan *, int, const PhItemIO *, double *);
__global__ void k_apply_correction_spans(const PhInst *, const PhSpan *, int);
__global__ void k_residual_spans(const PhInst *, const PhSpan *, int, double *);
__global__ void k_residual_norm_spans(const PhSpan *, const double *, double *);
__global__ void k_probe_residual_spans(const PhInst *, const PhSpan *, int, double *);
__global__ void k_scatter(PhInst, float *);
template <int CC>
__global__ void k_scatter_spans(const PhInst *, const PhSpan *, int, const PhItemIO *);
__global__ void k_gather_hist(const PhHistCopy *, double *);
__global__ void k_coarsen(const float *, const float *, const float *, const float *, int, int, int,
                          int, double, int, int, int, int, int, float *, float *, float *);
__global__ void k_coarsen_groups(const PhInst *, const PhCorrArgs *, const PhSpan *, int,
                                 const PhItemIO *, double, int, int, float *);
__global__ void k_warm_apply_spans(const PhInst *, const PhSpan *, int, const PhCorrArgs *,
                                   const PhItemIO *, const float *, const float *, int);
__global__ void k_coarsen_plane(PhInst, int, int, int, int, int, float *);
__global__ void k_corr_input_list(const PhInst *, const PhCtl *, const PhCorrArgs *, int, float *,
                                  double *);
__global__ void k_corr_apply_spans(const PhInst *, const PhCtl *, const PhSpan *, int,
                                   const PhCorrArgs *, const float *, const float *,
                                   const double *, int, double *);
__global__ void k_put_values(const PhValue *, int);
__global__ void k_copy_anchor_spans(const PhInst *, const PhInst *, const PhSpan *, int);
int coarsen_plane_cpb(int f);
int coarsen_blocks(int f, int res);
__global__ void k_in_norm_part(const float *, int, double *);
__global__ void k_in_norm_apply(float *, int, const double *, float *);
__global__ void k_corr_alpha_part(const float *, const float *, const double *, int, double *);
__global__ void k_corr_alpha(const double *, int, double, double *, double *);
__global__ void k_start_apply(PhInst, const float *);
bool warm_net_ready();
int warm_net_res();
int warm_net_batch();
void warm_net_forward(const float *d_in, float *d_out, int batch, hipStream_t st);

// ---------------------------------------------------------------------------
// errors

Fail fail(int cls, const char *code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return Fail{cls, code, buf};
}

void put_error(ph_error *err, const Fail &f) {
  if (!err) return;
  err->error_class = f.cls;
  snprintf(err->code, sizeof(err->code), "%s", f.code.c_str());
  snprintf(err->message, sizeof(err->message), "%s", f.msg.c_str());
}

void clear_error(ph_error *err) {
  if (!err) return;
  err->error_class = -1;
  err->code[0] = 0;
  err->message[0] = 0;
}

// ---------------------------------------------------------------------------
// device context (one per process, behind a mutex)

std::mutex g_mu;
int g_forced = 0;
Ctx g_ctx;

static bool forced_unavailable() {
  if (g_forced) return true;
  const char *e = getenv("PAINTOP_HIP_FORCE_UNAVAILABLE");
  return e && e[0] == '1';
}

void ensure_device() {
  if (forced_unavailable())
    throw fail(PH_CLASS_POLICY, "E_GPU_UNAVAILABLE",
               "no compatible GPU is available: forced unavailable");
  if (!g_ctx.probed) {
    g_ctx.probed = true;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
      g_ctx.ok = false;
      g_ctx.why = "no HIP device";
    } else {
      g_ctx.dev = 0;
      if (hipGetDeviceProperties(&g_ctx.prop, 0) != hipSuccess) {
        g_ctx.ok = false;
        g_ctx.why = "hipGetDeviceProperties failed";
      } else if (strncmp(g_ctx.prop.gcnArchName, "gfx950", 6) != 0) {
        g_ctx.ok = false;
        g_ctx.why = std::string("device 0 is ") + g_ctx.prop.gcnArchName + ", not gfx950";
      } else {
        g_ctx.ok = true;
      }
    }
  }
  if (!g_ctx.ok)
    throw fail(PH_CLASS_POLICY, "E_GPU_UNAVAILABLE", "no compatible GPU is available: %s",
               g_ctx.why.c_str());
  HIP_OK(hipSetDevice(g_ctx.dev));
}

// Device workspace of a solve. Normally a grow-only allocation reused across
// calls; during a call of the allocation-free entry point it is the fixed
// `scratch` block instead, and a solve that would need more is rejected before
// any launch.
struct Pool {
  char *own = nullptr;  // hipMalloc'd, grow-only
  size_t own_cap = 0;
  char *ext = nullptr;  // fixed block in use instead (allocation-free calls), or null
  size_t ext_cap = 0;
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void reset(size_t need) {
    if (ext) {
      if (need > ext_cap)
        throw fail(PH_CLASS_POLICY, "E_HIP_SCRATCH_LIMIT",
                   "the solve needs %zu bytes of device workspace; the allocation-free entry "
                   "point has %zu (PH_NOALLOC_SCRATCH_BYTES)",
                   need, ext_cap);
      base = ext;
      cap = ext_cap;
    } else {
      if (need > own_cap) {
        if (own) HIP_OK(hipFree(own));
        own = nullptr;
        own_cap = 0;
        HIP_OK(hipMalloc(&own, need));
        own_cap = need;
      }
      base = own;
      cap = own_cap;
    }
    used = 0;
  }
  template <class T>
  T *take(size_t n) {
    size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
    if (used + bytes > cap) throw fail(PH_CLASS_EXECUTION, "E_HIP_WORKSPACE", "workspace overrun");
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
};
static Pool g_pool;

// Pinned host memory for the readbacks of the sweep loop (control blocks,
// residual norms): a small copy to pageable memory goes through a staging
// buffer and waits on it; to pinned memory it is one DMA. Host code must not
// change a pinned buffer while a copy from it is queued. Allocation-free calls
// use a fixed pageable block instead (ext).
struct PinnedPool {
  char *own = nullptr;
  size_t own_cap = 0;
  char *ext = nullptr;
  size_t ext_cap = 0;
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void reset(size_t need) {
    if (ext) {
      if (need > ext_cap)
        throw fail(PH_CLASS_POLICY, "E_HIP_SCRATCH_LIMIT",
                   "the solve needs %zu bytes of host staging; the allocation-free entry "
                   "point has %zu",
                   need, ext_cap);
      base = ext;
      cap = ext_cap;
    } else {
      if (need > own_cap) {
        if (own) HIP_OK(hipHostFree(own));
        own = nullptr;
        own_cap = 0;
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&own), need, hipHostMallocDefault));
        own_cap = need;
      }
      base = own;
      cap = own_cap;
    }
    used = 0;
  }
  template <class T>
  T *take(size_t n) {
    size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
    if (used + bytes > cap)
      throw fail(PH_CLASS_EXECUTION, "E_HIP_WORKSPACE", "pinned staging overrun");
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
};
static PinnedPool g_pinned;

// True while an allocation-free call runs: sweep batches are launched
// directly (capturing and instantiating graphs may allocate).
static bool g_noalloc = false;

// ---------------------------------------------------------------------------
// parameter validation (the same ranges and codes as resolve_controls,
// poisson_blend.rs:102-139 and screened_poisson.rs:66-112)

static const uint32_t MAX_ITERATIONS_LIMIT = 100000;  // poisson_blend.rs:71

static void validate_shape(const ph_solve_desc *d, uint32_t width, uint32_t height,
                           uint32_t channels) {
  const bool screened = d->op == PH_OP_SCREENED_POISSON;
  const char *opn = screened ? "repair.screened_poisson" : "repair.poisson_blend";
  if (channels == 0 || channels > 64)
    throw fail(PH_CLASS_SEMANTIC, "E_POISSON_BLEND_SHAPE", "%s: channel count %u out of range",
               opn, channels);
  if ((size_t)width * height > ((size_t)1 << 31))
    throw fail(PH_CLASS_POLICY, "E_HIP_LIMIT", "image of %ux%u exceeds the 2^31-pixel limit",
               width, height);
}

static void validate(const ph_solve_desc *d, bool check_shape = true) {
  if (!d) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null descriptor");
  const bool screened = d->op == PH_OP_SCREENED_POISSON;
  if (d->op != PH_OP_POISSON_BLEND && !screened)
    throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "unknown op %d", d->op);
  const char *pcode = screened ? "E_SCREENED_POISSON_PARAM" : "E_POISSON_BLEND_PARAM";
  const char *opn = screened ? "repair.screened_poisson" : "repair.poisson_blend";
  if (d->max_iterations == 0 || d->max_iterations > MAX_ITERATIONS_LIMIT)
    throw fail(PH_CLASS_SCHEMA, pcode, "%s: max_iterations must be in 1..=%u", opn,
               MAX_ITERATIONS_LIMIT);
  if (!std::isfinite(d->tolerance) || d->tolerance <= 0.0 || d->tolerance >= 1.0)
    throw fail(PH_CLASS_SCHEMA, pcode, "%s: tolerance must be in (0, 1)", opn);
  if (!std::isfinite(d->lambda) || d->lambda < 0.0)
    throw fail(PH_CLASS_SCHEMA, pcode, "%s: lambda must be a finite value >= 0", opn);
  if (!screened && d->lambda != 0.0)
    throw fail(PH_CLASS_SCHEMA, pcode, "%s: lambda must be 0 for the pure Poisson blend", opn);
  if (d->omega != 0.0 && !(d->omega > 0.0 && d->omega < 2.0))
    throw fail(PH_CLASS_SCHEMA, pcode, "%s: omega must be in (0, 2)", opn);

  const uint32_t k = d->half_sweeps;
  if (k != 0 && k != 2 && k != 4 && k != 8 && k != 12 && k != 16)
    throw fail(PH_CLASS_SCHEMA, "E_HIP_TUNING", "half_sweeps must be one of 2, 4, 8, 12, 16");
  if (d->tile_rows != 0 && (d->tile_rows < 16 || d->tile_rows > 1024 || (d->tile_rows & 1)))
    throw fail(PH_CLASS_SCHEMA, "E_HIP_TUNING", "tile_rows must be 0 or even in 16..1024");
  if (check_shape) validate_shape(d, d->width, d->height, d->channels);
}

// ---------------------------------------------------------------------------
// tuning

// SOR factor tuner. omega_opt = 2 / (1 + sqrt(1 - rho^2)) with the Jacobi
// spectral radius rho = (4 - mu) / (4 + lambda), where mu is the smallest
// eigenvalue of the 5-point negative Laplacian on the interior cells. mu is
// measured on the mask coarsened by f (a coarse cell is interior when at least
// half of its f x f cells are), by inverse iteration with conjugate gradients,
// and scaled by 1/f^2. On a 4096^2 star domain a <= 128^2 coarse grid lands
// within a few percent of the fine eigenvalue, on the low side (a slightly
// large omega, where SOR degrades gently). Results are cached per coarse mask,
// keyed by its size, factor and a 64-bit hash the device computes with the
// mask (the sum of a mix of every inside cell's index), so a cached mask is
// never copied to the host.
struct TuneKey {
  int cw, ch, f;
  uint64_t hash;
  bool operator==(const TuneKey &o) const {
    return cw == o.cw && ch == o.ch && f == o.f && hash == o.hash;
  }
};
static std::vector<std::pair<TuneKey, double>> g_tune_cache;

static double coarse_min_eigenvalue(const std::vector<uint8_t> &in, int cw, int ch) {
  const int n = cw * ch;
  std::vector<int> idx(n, -1);
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (in[i]) idx[i] = m++;
  if (m == 0) return 4.0;
  std::vector<int> nb((size_t)m * 4, -1);
  for (int y = 0; y < ch; ++y)
    for (int x = 0; x < cw; ++x) {
      const int k = idx[y * cw + x];
      if (k < 0) continue;
      nb[4 * k + 0] = x > 0 ? idx[y * cw + x - 1] : -1;
      nb[4 * k + 1] = x + 1 < cw ? idx[y * cw + x + 1] : -1;
      nb[4 * k + 2] = y > 0 ? idx[(y - 1) * cw + x] : -1;
      nb[4 * k + 3] = y + 1 < ch ? idx[(y + 1) * cw + x] : -1;
    }
  auto apply = [&](const std::vector<double> &v, std::vector<double> &out) {
    for (int k = 0; k < m; ++k) {
      double s = 4.0 * v[k];
      for (int j = 0; j < 4; ++j)
        if (nb[4 * k + j] >= 0) s -= v[nb[4 * k + j]];
      out[k] = s;
    }
  };
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0;
    for (int k = 0; k < m; ++k) s += a[k] * b[k];
    return s;
  };
  std::vector<double> v(m, 1.0), x(m, 0.0), r(m), p(m), q(m);
  double mu = 4.0;
  for (int it = 0; it < 8; ++it) {
    // solve A x = v by CG (warm start from the previous iterate's direction)
    apply(x, q);
    for (int k = 0; k < m; ++k) r[k] = v[k] - q[k];
    p = r;
    double rr = dot(r, r);
    // relative residual 1e-6: inexact inverse iteration still converges to the
    // smallest eigenpair (each solve starts at v / mu), and omega needs mu to
    // about 1e-3 only
    const double stop = 1e-12 * dot(v, v);
    for (int cg = 0; cg < 4 * m + 100 && rr > stop; ++cg) {
      apply(p, q);
      const double alpha = rr / dot(p, q);
      for (int k = 0; k < m; ++k) {
        x[k] += alpha * p[k];
        r[k] -= alpha * q[k];
      }
      const double rr2 = dot(r, r);
      for (int k = 0; k < m; ++k) p[k] = r[k] + (rr2 / rr) * p[k];
      rr = rr2;
    }
    const double nx = sqrt(dot(x, x));
    for (int k = 0; k < m; ++k) v[k] = x[k] / nx;
    apply(v, q);
    mu = dot(v, q);  // Rayleigh quotient of the normalised iterate
    x = v;
    for (int k = 0; k < m; ++k) x[k] /= mu;  // next solve starts at A^-1 v ~ v / mu
  }
  return mu;
}

static double omega_for(double mu, double lambda) {
  const double rho = (4.0 - std::min(mu, 4.0)) / (4.0 + lambda);
  const double w = 2.0 / (1.0 + sqrt(std::max(1.0 - rho * rho, 0.0)));
  return std::min(std::max(w, 1.0), 1.9999);
}

// Window of the stall guard (poisson.rs:215-228). The reference's 16 sweeps
// fit its fixed omega = 1.8; the over-relaxation transient, during which the
// residual of red/black SOR may rise before it falls, lasts ~1/(2 - omega)
// sweeps, so a tuned omega gets a window of 3/(2 - omega) (16 for any
// omega <= 1.81). A fixed omega from the caller keeps the reference's 16.
// Sweeps of the omega ramp after a kept warm start: a multiple of 1/(2 - omega),
// which grows with the domain's size like the SOR transient itself.
// Measured best for the shipped network on 2048^2 and 4096^2 blends: 1 for a
// warm start alone, 1/4 with coarse-correction cycles (each cycle starts over
// from a fresh correction, so a short ramp that reaches omega early pays).
static uint32_t warm_ramp_sweeps(double omega, bool cycles) {
  double fct = cycles ? 0.25 : 1.0;
  if (const char *e = getenv("PAINTOP_HIP_RAMP_FACTOR")) fct = atof(e);
  if (!(fct >= 0.0)) fct = 0.0;
  return (uint32_t)std::min(1e5, std::ceil(fct / (2.0 - omega)));
}

// Sweeps of the warm-start race before the FP64 judgement: a quarter of the
// ramp, within [8, 96] (always shorter than the warm member's stall window,
// which is the tuned window plus the ramp).
static uint32_t warm_probe_sweeps(uint32_t ramp) {
  return std::min<uint32_t>(96, std::max<uint32_t>(8, ramp / 4));
}
// The race is judged on the FP64 residual of each member's state at the
// probe sweep. A warm start keeps the lead only when its residual is at most
// RACE_MARGIN of the plain start's; a smaller lead extends the race to twice
// the probe, at most RACE_MAX_EXTENSIONS times, after which the smaller
// residual wins (the shipped network leads ~20-fold at the first probe).
static const double RACE_MARGIN = 0.25;
static const int RACE_MAX_EXTENSIONS = 4;
// A warm start whose initial residual exceeds the anchor start's this many
// times is dropped without a race. A coarse prediction leaves the fine
// boundary texture unresolved, so good starts sit at 5-8x (the shipped
// network, and even the exact coarse solution); far above that the start
// carries large error everywhere, which a short probe can fail to expose.
static const double WARM_MAX_START_RATIO = 32.0;

// Learned coarse corrections. After a kept warm start, every cycle of sweeps
// ends in an FP64 refinement whose residual g = -r(x) defines the correction
// problem Laplacian(e) = g with e = 0 on the ring; the network solves it on
// the warm start's coarse grid and x += bilinear(e), after which the omega
// ramp restarts. The cycle is a multiple of the ramp (PAINTOP_HIP_CYCLE_FACTOR,
// default 1.5), on a launch boundary; PAINTOP_HIP_CORRECTIONS caps the count per
// channel (0 = none).
// The network's start and its corrections pay only where sweeps are dear: each
// costs a race twin or a network round (a forward, a few plane passes and a
// host round trip) whatever the instance's size, while a small instance's
// sweeps cost little, and nothing at all in a batch, where they ride in the
// launches of the larger instances. So a channel takes the learned path when
// its interior cells reach max(2^14, the call's largest instance / 8), or
// PAINTOP_HIP_WARM_MIN_CELLS / PAINTOP_HIP_CORR_MIN_CELLS (start / corrections)
// when set. Measured on MI355X (ms, learned against plain):
//   4-channel star blends: 256^2 (20k cells per channel) 3.0 vs 4.7 with the
//   gate at 2^14; 512^2 3.4 vs 4.9; 1024^2 4.9 vs 15.0 (8.6 with the start
//   alone);
//   the batch of 512 blends of 32..1024 px (largest 0.6M cells): fixed gates
//   of 2^12 / 2^14 / 2^15 / 2^16 gave 49.5 / 43.4 / 40.0 / 38.5 against 55.0
//   plain (2^18 for the start and 2^20 for the corrections: 59.7).
static const uint64_t WARM_GATE_FLOOR = 1ull << 14;
static const uint64_t WARM_GATE_DIVISOR = 8;
static uint64_t warm_gate(const char *env, uint64_t largest) {
  if (const char *e = getenv(env)) return (uint64_t)std::max(0LL, atoll(e));
  return std::max(WARM_GATE_FLOOR, largest / WARM_GATE_DIVISOR);
}
static uint32_t warm_max_corrections() {
  long n = 16;
  if (const char *e = getenv("PAINTOP_HIP_CORRECTIONS")) n = atol(e);
  return (uint32_t)std::max(0L, std::min(n, 1000L));
}
static uint32_t warm_cycle_sweeps(uint32_t ramp, uint32_t half) {
  double fct = 1.5;  // measured best at 2048^2 and 4096^2 (1..2 all within 20%)
  if (const char *e = getenv("PAINTOP_HIP_CYCLE_FACTOR")) fct = atof(e);
  if (!(fct > 0.0)) fct = 1.0;
  const uint32_t n = (uint32_t)std::min(1e5, std::ceil(fct * ramp));
  return std::max<uint32_t>(2 * half, (n + half - 1) / half * half);
}
// Corrections go on while each cycle ends well below the end of the previous
// one. Plain SOR gains about exp(c (2 - omega)) over c sweeps, ~1.45x over a
// cycle of the default length; a cycle that gains less than CORR_MIN_GAIN
// ends the corrections, and so do two in a row below CORR_GOOD_GAIN (a single
// weak cycle, after a mediocre correction, still beats the plain rate that
// ending would fall back to).
static const double CORR_MIN_GAIN = 2.0;
static const double CORR_GOOD_GAIN = 4.0;
// Largest step of a correction (its Galerkin step on the coarse grid is
// clamped to this).
static const double CORR_MAX_STEP = 1.5;
// A good correction leaves the residual within about 1.6x of the uncorrected
// one (bilinear interpolation error at the fine scale, which the next sweeps
// remove). One that raises it more than this (PAINTOP_HIP_CORR_MAX_JUMP,
// default 4) is taken back; the cycle then runs on uncorrected, and a second
// one taken back in a row ends the cycles.
static double corr_max_jump() {
  double j = 4.0;
  if (const char *e = getenv("PAINTOP_HIP_CORR_MAX_JUMP")) j = atof(e);
  return j > 0.0 ? j : 4.0;
}

static uint32_t stall_window_for(double omega, bool tuned) {
  if (!tuned) return PH_STALL_WINDOW;
  const double w = ceil(3.0 / std::max(2.0 - omega, 1e-6));
  return (uint32_t)std::max<double>(PH_STALL_WINDOW, std::min(w, 1e6));
}

// One sweep launch: the tiles (one wave each), then the per-instance control.
static void launch_sor(int K, int PF, int ntiles, int ninst, hipStream_t st, const PhTile *tiles,
                       const PhInst *insts, PhCtl *ctl, double *partials, int hb) {
  const dim3 grid(ntiles), block(64), cgrid(ninst), cblock(64 * K);
#define PH_SOR(KK, PP)                                                                      \
  hipLaunchKernelGGL((k_sor<KK, PP>), grid, block, 0, st, tiles, ntiles, insts, ctl, partials, hb)
  switch (K * 8 + PF) {
    case 2 * 8 + 1: PH_SOR(2, 1); break;
    case 4 * 8 + 1: PH_SOR(4, 1); break;
    case 8 * 8 + 1: PH_SOR(8, 1); break;
    case 12 * 8 + 1: PH_SOR(12, 1); break;
    case 16 * 8 + 1: PH_SOR(16, 1); break;
    case 2 * 8 + 3: PH_SOR(2, 3); break;
    case 4 * 8 + 3: PH_SOR(4, 3); break;
    case 8 * 8 + 3: PH_SOR(8, 3); break;
    case 12 * 8 + 3: PH_SOR(12, 3); break;
    case 16 * 8 + 3: PH_SOR(16, 3); break;
    case 8 * 8 + 5: PH_SOR(8, 5); break;
    case 16 * 8 + 5: PH_SOR(16, 5); break;
    default: throw fail(PH_CLASS_SCHEMA, "E_HIP_TUNING", "no sweep kernel for K=%d prefetch=%d", K, PF);
  }
#undef PH_SOR
  HIP_OK(hipGetLastError());
  switch (K) {
    case 2: hipLaunchKernelGGL(k_sor_control<2>, cgrid, cblock, 0, st, insts, ctl, partials, ninst, ntiles); break;
    case 4: hipLaunchKernelGGL(k_sor_control<4>, cgrid, cblock, 0, st, insts, ctl, partials, ninst, ntiles); break;
    case 8: hipLaunchKernelGGL(k_sor_control<8>, cgrid, cblock, 0, st, insts, ctl, partials, ninst, ntiles); break;
    case 12: hipLaunchKernelGGL(k_sor_control<12>, cgrid, cblock, 0, st, insts, ctl, partials, ninst, ntiles); break;
    default: hipLaunchKernelGGL(k_sor_control<16>, cgrid, cblock, 0, st, insts, ctl, partials, ninst, ntiles); break;
  }
  HIP_OK(hipGetLastError());
}

// Sweep batches as hipGraphs: a batch of n launch pairs (k_sor +
// k_sor_control) is replayed from a graph of 2^j pairs per set bit of n,
// captured once per kernel-argument set on a private stream (default
// priority) and launched on the solve's stream, so a host iteration costs a
// few graph launches instead of up to 128 kernel launches. The graphs hold
// only kernel nodes. PAINTOP_HIP_GRAPH=0 launches the kernels directly.
// The cache holds at most 64 graphs; launch_sor_batch evicts, never get().
struct SweepGraphs {
  struct Entry {
    int K, PF, ntiles, ninst, hb, n;
    const void *tiles, *insts, *ctl, *partials;
    hipGraphExec_t exec;
  };
  hipStream_t cs = nullptr;
  std::vector<Entry> cache;
  bool broken = false;
  void clear() {
    for (Entry &e : cache) (void)hipGraphExecDestroy(e.exec);
    cache.clear();
  }
  hipGraphExec_t get(int K, int PF, int ntiles, int ninst, const PhTile *tiles,
                     const PhInst *insts, PhCtl *ctl, double *partials, int hb, int n) {
    for (const Entry &e : cache)
      if (e.K == K && e.PF == PF && e.ntiles == ntiles && e.ninst == ninst && e.hb == hb &&
          e.n == n && e.tiles == tiles && e.insts == insts && e.ctl == ctl &&
          e.partials == partials)
        return e.exec;
    if (!cs) HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    try {
      for (int i = 0; i < n; ++i) launch_sor(K, PF, ntiles, ninst, cs, tiles, insts, ctl, partials, hb);
    } catch (...) {
      (void)hipStreamEndCapture(cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_OK(hipStreamEndCapture(cs, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    cache.push_back(Entry{K, PF, ntiles, ninst, hb, n, tiles, insts, ctl, partials, ex});
    return ex;
  }
};
static SweepGraphs g_graphs;

// A second stream of default priority (the learned start's network pass runs
// on it beside the solve's stream) and the two events that order them.
static hipStream_t side_stream() {
  static hipStream_t s = nullptr;
  if (!s) HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}
static std::pair<hipEvent_t, hipEvent_t> side_events() {
  static hipEvent_t a = nullptr, b = nullptr;
  if (!a) {
    HIP_OK(hipEventCreateWithFlags(&a, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&b, hipEventDisableTiming));
  }
  return {a, b};
}

// PAINTOP_HIP_TRACE=1: the warm-start race's events on stderr.
static bool trace_on() {
  const char *e = getenv("PAINTOP_HIP_TRACE");
  return e && e[0] == '1';
}

static bool use_graphs() {
  const char *e = getenv("PAINTOP_HIP_GRAPH");
  return !(e && e[0] == '0');
}

// n launch pairs on st: from cached graphs when enabled, else directly.
static void launch_sor_batch(int K, int PF, int ntiles, int ninst, hipStream_t st,
                             const PhTile *tiles, const PhInst *insts, PhCtl *ctl,
                             double *partials, int hb, int n) {
  if (n <= 0 || ntiles <= 0) return;
  if (n == 1 || !use_graphs() || g_graphs.broken || g_noalloc) {
    for (int i = 0; i < n; ++i) launch_sor(K, PF, ntiles, ninst, st, tiles, insts, ctl, partials, hb);
    return;
  }
  // a batch adds at most 7 graphs; make room before its first launch, once
  // the stream (which may still run graphs of earlier batches) has drained
  if (g_graphs.cache.size() + 7 > 64) {
    HIP_OK(hipStreamSynchronize(st));
    g_graphs.clear();
  }
  for (int b = 6; b >= 0; --b)
    if (n & (1 << b)) {
      hipGraphExec_t ex = g_graphs.get(K, PF, ntiles, ninst, tiles, insts, ctl, partials, hb, 1 << b);
      HIP_OK(hipGraphLaunch(ex, st));
    }
  if (n >= 128)
    for (int i = 0; i < (n >> 7) * 128; ++i)
      launch_sor(K, PF, ntiles, ninst, st, tiles, insts, ctl, partials, hb);
}

// Rows a sweep tile loads ahead of use (1, 3 or 5; PAINTOP_HIP_PREFETCH overrides).
static int prefetch_rows(int K) {
  const char *e = getenv("PAINTOP_HIP_PREFETCH");
  int pf = e ? atoi(e) : 3;
  if (pf != 1 && pf != 3 && pf != 5) pf = 3;
  if (pf == 5 && K != 8 && K != 16) pf = 3;
  return pf;
}

// ---------------------------------------------------------------------------
// the solve

struct ChanResult {
  uint32_t iterations = 0;
  int stop = PH_STOP_CONVERGED;
  bool converged = true;
  double final_residual = 0.0;
  bool active = false;
  bool done = true;
  bool nonfinite = false;
};

// Refinement triggers: the inner residual fell 1e3-fold, or REFINE_EVERY
// sweeps passed. The second bounds |e| by the error at the start of the window,
// so each sweep's change to e (~ (2 - omega) x error) stays far above
// ulp(|e|): an FP32 correction that accumulated the whole solve would stop
// absorbing the late, small updates.
static const double REFINE_RATIO = 1e-3;
static const uint32_t REFINE_EVERY = 1024;
// A cycle end refines (and its correction follows on the refined state) once
// the residual has fallen 1/REFINE_RATIO/CYCLE_REFINE_AHEAD-fold since the
// last refinement, about 30x: a cycle gains 10-50x, so the refinement
// trigger would otherwise fire inside the next cycle, halting the instance
// in the middle of a batch of launches planned to the cycle's end.
static const double CYCLE_REFINE_AHEAD = 0.03;
// Right after a correction the residual dips within a quarter of its omega
// ramp; near the tolerance a batch then plans one launch at a time.
static const double NEAR_TOLERANCE = 30.0;
// Below this many interior cells the SOR factor comes from the geometric
// bounds instead of the coarse-mask eigenvalue (small domains converge in a
// few hundred sweeps; the host-side eigen solve would cost more than it saves).
static const double PRECISE_TUNE_MIN_CELLS = 32768.0;

// One image of a solve (one op call).
struct ItemIn {
  int W, H, C;
  const float *src, *tgt, *mask;
  float *out;
  ph_report *rep;
  const float *start = nullptr;  // caller's warm start (same layout as out), or null
};

struct ItemWork {
  ItemIn in;
  uint8_t *cls_img = nullptr;
  bool any = false;
  int rx0 = 0, ry0 = 0, rw = 0, rh = 0, ox = 0, oy = 0, pw = 0, phh = 0, pitch = 0;
  std::vector<int2> origins;
  size_t origin_off = 0;
  std::vector<int2> live;  // tile origins whose core holds interior cells
  std::vector<char> live_interior;  // 1: every cell the live tile loads is interior
  double n_interior = 0;
  int cf = 1, cw = 0, ch = 0;
  size_t coarse_off = 0;
  int coarse_entry = -1;  // entry in the coarse-mask launch, or -1: no mask
  bool precise = false;
  double omega = 1.0;
  uint32_t stall_window = PH_STALL_WINDOW;
  uint8_t *cls_plane = nullptr;
  int inst0 = -1;
};

struct InstWork {
  int item = 0, channel = 0;
  PhInst I{};
  double r0 = 0.0;
  ChanResult R;
  double *hist = nullptr;
  dim3 pgrd;
};

// Geometric bounds on mu (the smallest eigenvalue of the 5-point operator on
// the interior): the equal-area disc (Faber-Krahn, mu >= (j01/R)^2) and the
// bounding rectangle (domain monotonicity); the larger bound is used.
static double mu_bound(double n_interior, int bw, int bh) {
  const double j01 = 2.404825557695773;
  const double R = sqrt(std::max(n_interior, 1.0) / M_PI);
  const double mu_disc = (j01 / R) * (j01 / R);
  const double mu_rect = (M_PI / std::max(bw, 1)) * (M_PI / std::max(bw, 1)) +
                         (M_PI / std::max(bh, 1)) * (M_PI / std::max(bh, 1));
  return std::min(std::max(mu_disc, mu_rect), 4.0);
}


// PAINTOP_HIP_TUNE_CACHE=0: every solve pays the tuner (the cost a new mask
// sees), e.g. to time it.
static bool tune_lookup(const TuneKey &key, double &mu) {
  const char *ev = getenv("PAINTOP_HIP_TUNE_CACHE");
  if (ev && ev[0] == '0') return false;
  for (const auto &e : g_tune_cache)
    if (e.first == key) {
      mu = e.second;
      return true;
    }
  return false;
}

static void tune_store(const TuneKey &key, double mu) {
  if (g_tune_cache.size() > 4096) g_tune_cache.erase(g_tune_cache.begin());
  g_tune_cache.push_back({key, mu});
}

// Upper bound of one image's workspace, so the pool never regrows mid-call.
// Host side of a row-batched launch: entries with the running row count.
struct SpanTable {
  std::vector<PhSpan> h;
  int rows = 0;
  void clear() {
    h.clear();
    rows = 0;
  }
  void add(int slot, int nrows, int item, int buf = 0, int out = 0, double *hist = nullptr,
           double r0 = 0.0) {
    h.push_back(PhSpan{slot, rows, nrows, buf, out, item, hist, r0});
    rows += nrows;
  }
  // the host vector is pageable staging: keep it unchanged until the stream
  // has been synchronised
  void upload(PhSpan *d, hipStream_t st) const {
    if (!h.empty())
      HIP_OK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(PhSpan), hipMemcpyHostToDevice, st));
  }
  // through pinned staging (one DMA, no staging wait): the sweep loop's
  // tables; the stage must hold the entries and stay unchanged until the
  // stream has been synchronised
  void upload(PhSpan *d, hipStream_t st, PhSpan *stage) const {
    if (h.empty()) return;
    memcpy(stage, h.data(), h.size() * sizeof(PhSpan));
    HIP_OK(hipMemcpyAsync(d, stage, h.size() * sizeof(PhSpan), hipMemcpyHostToDevice, st));
  }
};

// n values of T to the device through pinned staging (see SpanTable::upload).
template <class T>
static void upload_pinned(T *d, const T *h, size_t n, T *stage, hipStream_t st) {
  if (n == 0) return;
  memcpy(stage, h, n * sizeof(T));
  HIP_OK(hipMemcpyAsync(d, stage, n * sizeof(T), hipMemcpyHostToDevice, st));
}

static size_t item_need(const ph_solve_desc *d, int W, int H, int C, int K, int hb) {
  const int WV = 256 - 2 * K;
  const size_t npx = (size_t)W * H;
  const size_t pw_up = (size_t)W + 2 * PH_PAD + 4 + WV + 2 * K + 64;
  const size_t ph_up = (size_t)H + 2 * PH_PAD + 4 + hb + 2 * K;
  const size_t plane_up = pw_up * ph_up;
  const size_t ntile_up = (pw_up / WV + 2) * (ph_up / hb + 2);
  return npx + 8 * 1024 + 2 * (size_t)H * sizeof(int) +
         ntile_up * (sizeof(int2) + 3 * sizeof(int)) + sizeof(PhItemGeo) +
         (size_t)C * 5 * ph_up * sizeof(double) + (size_t)(12 * C + 4) * sizeof(PhSpan) +
         (size_t)C * (3 * sizeof(PhInst) + 4 * sizeof(PhCorrArgs) + sizeof(PhValue)) +
         sizeof(PhItemIO) + sizeof(PhHistCopy) + (size_t)(d->max_iterations + 2) * sizeof(double) +
         plane_up * (1 + (size_t)C * (3 * sizeof(float) + 2 * sizeof(double))) +
         (size_t)(C + 1) * (d->max_iterations + 2) * sizeof(double) +
         (size_t)C * ph_up * sizeof(double) +
         ntile_up * (size_t)C * (sizeof(PhTile) + K * sizeof(double)) +
         (size_t)C * (sizeof(PhInst) + sizeof(PhCtl) + 256) + 40 * 256 +
         130 * 130 * sizeof(int);
}

// Per-image launches (assembly, write-back) specialise on the channel count
// when every listed image has the same one (1..4); 0 = any.
static int common_channels(const SpanTable &t, const std::vector<ItemWork> &iw) {
  int c = -1;
  for (const PhSpan &e : t.h) {
    const int ci = iw[e.item].in.C;
    if (c < 0) c = ci;
    else if (c != ci) return 0;
  }
  return c >= 1 && c <= 4 ? c : 0;
}

template <template <int> class L, class... A>
static void launch_by_channels(int cc, A... args) {
  switch (cc) {
    case 1: L<1>::go(args...); break;
    case 2: L<2>::go(args...); break;
    case 3: L<3>::go(args...); break;
    case 4: L<4>::go(args...); break;
    default: L<0>::go(args...); break;
  }
}
template <int CC>
struct AssembleL {
  static void go(dim3 g, hipStream_t st, const PhInst *i, const PhSpan *sp, int n,
                 const PhItemIO *io, double *row_sums) {
    hipLaunchKernelGGL(k_assemble_spans<CC>, g, dim3(256), 0, st, i, sp, n, io, row_sums);
  }
};
template <int CC>
struct ScatterL {
  static void go(dim3 g, hipStream_t st, const PhInst *i, const PhSpan *sp, int n,
                 const PhItemIO *io) {
    hipLaunchKernelGGL(k_scatter_spans<CC>, g, dim3(256), 0, st, i, sp, n, io);
  }
};

// Half-sweeps per launch and tile rows of a call: the caller's, or the tuning
// table's row for the call's work (sum over images of channels x interior
// bounding-box area). One wave per tile streams its hb core rows and 2K halo
// rows per launch in periods of U = K + PF + 1 row steps; two waves fit a
// SIMD (2048 on the chip), so a launch takes about ceil(live tiles / 2048)
// rounds of one tile pass: large calls want tall tiles and K = 16 (fewer HBM
// round trips and halo rows per sweep), small ones short tiles and K = 8 (a
// lone tile's pass is serial, and shorter with fewer half-sweeps).
// (One table serves plain calls and calls with the learned start: a learned
// start that loses its race leaves the plain call's result bit for bit only
// under the same launch boundaries.)
static void sweep_config(const ph_solve_desc *d, double work, int &K, int &hb) {
  const PhSorTune *row = &PH_SOR_TUNING[0];
  while (work > row->w_max) ++row;
  K = d->half_sweeps ? (int)d->half_sweeps : row->k;
  if (d->tile_rows) hb = (int)d->tile_rows;
  else if (K == row->k) hb = row->hb;
  else {
    // a caller's K without tile rows: the ring period U, twice it for large calls
    const int U = K + prefetch_rows(K) + 1;
    hb = (work >= 3.0e7 ? 2 : 1) * ((16 + U - 1) / U) * U;
  }
}

static void solve_items(const ph_solve_desc *d, std::vector<ItemIn> &items, hipStream_t st) {
  const int nitem = (int)items.size();
  const bool timing = (d->flags & PH_FLAG_TIMING) != 0;
  const bool tuned = d->omega <= 0.0;
  const bool want_warm =
      (d->flags & PH_FLAG_WARM_START_NET) && !(d->flags & PH_FLAG_FORCE_PLAIN);
  bool race_any = want_warm;
  for (const ItemIn &it : items) race_any = race_any || it.start != nullptr;
  if (want_warm && !warm_net_ready())
    throw fail(PH_CLASS_POLICY, "E_HIP_WARM_START_UNSET",
               "warm start requested but no network weights are set (ph_warm_start_set_weights)");
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (timing) {
    HIP_OK(hipEventCreate(&ev0));
    HIP_OK(hipEventCreate(&ev1));
    HIP_OK(hipEventRecord(ev0, st));
  }
  {
    // (the sweep configuration is chosen after the classification: the
    // workspace covers every configuration the call may choose)
    auto item_need_any = [&](const ItemIn &it) {
      size_t m = 0;
      for (const PhSorTune &r : PH_SOR_TUNING) {
        int k = 0, h = 0;
        sweep_config(d, r.w_max, k, h);
        m = std::max(m, item_need(d, it.W, it.H, it.C, k, h));
      }
      return m;
    };
    size_t need = 128 * 1024 + (size_t)nitem * 64;
    for (const ItemIn &it : items) need += item_need_any(it);
    // the plain twins of the warm-start race
    if (race_any)
      for (const ItemIn &it : items) need += item_need_any(it);
    if (want_warm) {
      const size_t R2 = (size_t)warm_net_res() * warm_net_res();
      size_t nch = 0;
      for (const ItemIn &it : items) nch += it.C;
      need += (R2 * 4 * warm_net_batch() + (R2 + 1) * nch) * sizeof(float) +
              (nch + 2 * PH_NET_SLICES * (size_t)warm_net_batch()) * sizeof(double) + 1536;
      // the correction rounds' squared-residual cells and norms
      need += (R2 + 1) * nch * sizeof(double) + 512;
    }
    g_pool.reset(need);
    size_t nch = 0;
    for (const ItemIn &it : items) nch += it.C;
    // control blocks, norms and the sweep loop's staging (instances and twins)
    g_pinned.reset(4096 + 2 * std::max<size_t>(nch, 1) *
                              (4 * sizeof(double) + sizeof(PhCtl) + 2 * sizeof(PhSpan) +
                               sizeof(PhCorrArgs) + sizeof(PhValue) + sizeof(PhInst)) +
                   10 * 256);
  }
  // host time per phase (printed with PAINTOP_HIP_TRACE=1)
  const bool phase_trace = trace_on();
  auto phase_t0 = std::chrono::steady_clock::now();
  auto phase = [&](const char *name) {
    if (!phase_trace) return;
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "[phase] %-12s %8.3f ms\n", name,
            std::chrono::duration<double, std::milli>(t - phase_t0).count());
    phase_t0 = t;
  };
  std::vector<ItemWork> iw(nitem);
  for (int i = 0; i < nitem; ++i) {
    iw[i].in = items[i];
    ph_report *rep = items[i].rep;
    rep->launches = 0;
    rep->refinements = 0;
    rep->unknowns = 0;
    rep->unknown_sweeps = 0;
    rep->warm_start_used = 0;
    rep->warm_start_rel = 0.0;
    rep->coarse_corrections = 0;
    rep->reserved_ = 0;
    rep->sweep_ms = 0.0;
    rep->device_ms = 0.0;
    rep->omega_used = d->omega;
  }

  // --- 1. classification and bounding boxes of every image: one launch per
  // step over all image rows, one readback
  int *bbox_d = g_pool.take<int>(4 * (size_t)nitem);
  std::vector<int> bbox_h(4 * (size_t)nitem, -1);
  std::vector<PhItemIO> io(nitem);
  SpanTable img_rows;
  for (int i = 0; i < nitem; ++i) {
    const ItemIn &it = iw[i].in;
    PhItemIO &G = io[i];
    memset(&G, 0, sizeof(G));
    G.src = it.src;
    G.tgt = it.tgt;
    G.mask = it.mask;
    G.out = it.out;
    G.W = it.W;
    G.H = it.H;
    G.C = it.C;
    if ((size_t)it.W * it.H == 0) continue;
    iw[i].cls_img = G.cls_img = g_pool.take<uint8_t>((size_t)it.W * it.H);
    img_rows.add(i, it.H, i);
  }
  PhItemIO *io_d = g_pool.take<PhItemIO>(nitem);
  PhSpan *img_rows_d = g_pool.take<PhSpan>(nitem);
  if (img_rows.rows > 0) {
    int2 *rowx_d = g_pool.take<int2>((size_t)img_rows.rows);
    HIP_OK(hipMemcpyAsync(io_d, io.data(), nitem * sizeof(PhItemIO), hipMemcpyHostToDevice, st));
    img_rows.upload(img_rows_d, st);
    hipLaunchKernelGGL(k_classify_spans, dim3(img_rows.rows), dim3(256), 0, st, img_rows_d,
                       (int)img_rows.h.size(), io_d, rowx_d);
    hipLaunchKernelGGL(k_bbox_spans, dim3(img_rows.h.size()), dim3(256), 0, st, img_rows_d, rowx_d,
                       bbox_d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(bbox_h.data(), bbox_d, bbox_h.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
  }
  HIP_OK(hipStreamSynchronize(st));
  for (int i = 0; i < nitem; ++i)
    if ((size_t)iw[i].in.W * iw[i].in.H == 0) bbox_h[4 * i + 2] = -1;
  // the sweep configuration for the call's work
  double work = 0.0;
  for (int i = 0; i < nitem; ++i)
    if (bbox_h[4 * i + 2] >= 0)
      work += (double)iw[i].in.C * (bbox_h[4 * i + 2] - bbox_h[4 * i] + 1) *
              (bbox_h[4 * i + 3] - bbox_h[4 * i + 1] + 1);
  int K = 16, hb = 40;
  sweep_config(d, work, K, hb);
  const int WV = 256 - 2 * K;
  const int PF = prefetch_rows(K);
  for (int i = 0; i < nitem; ++i) iw[i].in.rep->half_sweeps_used = (uint32_t)K;

  phase("classify");
  // --- 2. planes, tile occupancy, coarse masks for the tuner (one readback)
  std::vector<int2> all_origins;
  size_t coarse_total = 0;
  for (int i = 0; i < nitem; ++i) {
    ItemWork &w = iw[i];
    if (bbox_h[4 * i + 2] < 0) continue;
    w.any = true;
    w.rx0 = bbox_h[4 * i] - 1;
    w.ry0 = bbox_h[4 * i + 1] - 1;
    const int rx1 = bbox_h[4 * i + 2] + 2, ry1 = bbox_h[4 * i + 3] + 2;
    if ((w.rx0 + w.ry0) & 1) w.rx0 -= 1;  // red/black colour of plane cells = image colour
    w.rw = rx1 - w.rx0;
    w.rh = ry1 - w.ry0;
    w.ox = w.rx0 - PH_PAD;
    w.oy = w.ry0 - PH_PAD;
    const int nstrips = (PH_PAD + w.rw - K + WV - 1) / WV;
    const int nblocks = (PH_PAD + w.rh - K + hb - 1) / hb;
    w.pw = nstrips * WV + 2 * K;
    w.phh = nblocks * hb + 2 * K;
    w.pitch = (w.pw + 63) & ~63;
    w.origin_off = all_origins.size();
    for (int m = 0; m < nblocks; ++m)
      for (int n = 0; n < nstrips; ++n) all_origins.push_back(make_int2(n * WV, m * hb));
    w.origins.assign(all_origins.begin() + w.origin_off, all_origins.end());
    // (a coarse mask only where the interior's bounding box could hold the
    // precise tuner's minimum of interior cells)
    const double bbox_area = (double)(bbox_h[4 * i + 2] - bbox_h[4 * i] + 1) *
                             (bbox_h[4 * i + 3] - bbox_h[4 * i + 1] + 1);
    if (tuned && bbox_area >= PRECISE_TUNE_MIN_CELLS) {
      w.cf = std::max(1, (std::max(w.rw, w.rh) + 127) / 128);
      w.cw = (w.rw + w.cf - 1) / w.cf;
      w.ch = (w.rh + w.cf - 1) / w.cf;
      w.coarse_off = coarse_total;
      coarse_total += (size_t)w.cw * w.ch;
    }
  }
  std::vector<int> counts(2 * all_origins.size());
  std::vector<uint64_t> coarse_hash;  // per coarse-mask image (ItemWork::coarse_entry)
  uint8_t *coarse_d = coarse_total ? g_pool.take<uint8_t>(coarse_total) : nullptr;
  if (!all_origins.empty()) {
    int2 *origins_d = g_pool.take<int2>(all_origins.size());
    int *counts_d = g_pool.take<int>(2 * all_origins.size());
    // one launch each for every image: per-image geometry and the image of
    // every tile origin
    std::vector<PhItemGeo> geo(nitem);
    std::vector<int> origin_item(all_origins.size());
    SpanTable coarse_rows;  // ph_coarse_parts blocks per coarse row of every tuned image
    for (int i = 0; i < nitem; ++i) {
      const ItemWork &w = iw[i];
      PhItemGeo &G = geo[i];
      memset(&G, 0, sizeof(G));
      G.cls = w.cls_img;
      G.w = w.in.W;
      G.h = w.in.H;
      G.ox = w.ox;
      G.oy = w.oy;
      G.rx0 = w.rx0;
      G.ry0 = w.ry0;
      if (w.any && w.ch > 0) {
        G.cf = w.cf;
        G.cw = w.cw;
        G.ch = w.ch;
        G.coarse_off = (int64_t)w.coarse_off;
        iw[i].coarse_entry = (int)coarse_rows.h.size();
        coarse_rows.add(i, w.ch * ph_coarse_parts(w.cw, w.cf), i);
      }
      if (w.any)
        for (size_t k = 0; k < w.origins.size(); ++k) origin_item[w.origin_off + k] = i;
    }
    PhItemGeo *geo_d = g_pool.take<PhItemGeo>(nitem);
    int *oitem_d = g_pool.take<int>(all_origins.size());
    HIP_OK(hipMemcpyAsync(origins_d, all_origins.data(), all_origins.size() * sizeof(int2),
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(geo_d, geo.data(), nitem * sizeof(PhItemGeo), hipMemcpyHostToDevice,
                          st));
    HIP_OK(hipMemcpyAsync(oitem_d, origin_item.data(), origin_item.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tile_occupancy_all, dim3(all_origins.size()), dim3(256), 0, st, geo_d,
                       oitem_d, origins_d, K, WV, hb, counts_d);
    uint64_t *hash_d = nullptr;
    if (coarse_rows.rows > 0) {
      const int ne = (int)coarse_rows.h.size();
      PhSpan *coarse_rows_d = g_pool.take<PhSpan>(ne);
      uint64_t *row_hash_d = g_pool.take<uint64_t>(coarse_rows.rows);
      hash_d = g_pool.take<uint64_t>(ne);
      coarse_rows.upload(coarse_rows_d, st);
      hipLaunchKernelGGL(k_coarse_count_all, dim3(coarse_rows.rows), dim3(256), 0, st, geo_d,
                         coarse_rows_d, ne, coarse_d, row_hash_d);
      hipLaunchKernelGGL(k_coarse_hash_items, dim3(ne), dim3(256), 0, st, coarse_rows_d,
                         row_hash_d, hash_d);
      coarse_hash.resize(ne);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(counts.data(), counts_d, counts.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    if (hash_d)
      HIP_OK(hipMemcpyAsync(coarse_hash.data(), hash_d, coarse_hash.size() * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  }

  phase("occupancy");
  // --- 3. live tiles, SOR factor and stall window per image
  for (int i = 0; i < nitem; ++i) {
    ItemWork &w = iw[i];
    if (!w.any) continue;
    const int full = 256 * (hb + 2 * K);
    for (size_t k = 0; k < w.origins.size(); ++k)
      if (counts[2 * (w.origin_off + k)] > 0) {
        w.live.push_back(w.origins[k]);
        w.live_interior.push_back(counts[2 * (w.origin_off + k) + 1] == full ? 1 : 0);
        w.n_interior += counts[2 * (w.origin_off + k)];
      }
  }
  // SOR factor per image: the coarse-mask eigenvalue for large domains (cached
  // per coarse mask; the masks not in the cache are solved on host threads in
  // parallel), the geometric bound for small ones
  std::vector<double> mu_of(nitem, 0.0);
  if (tuned) {
    struct Job { int item; TuneKey key; std::vector<uint8_t> in; double mu; };
    std::vector<Job> jobs;
    for (int i = 0; i < nitem; ++i) {
      ItemWork &w = iw[i];
      if (!w.any) continue;
      if (w.n_interior < PRECISE_TUNE_MIN_CELLS || w.coarse_entry < 0) {
        mu_of[i] = mu_bound(w.n_interior, w.rw, w.rh);
        continue;
      }
      Job j;
      j.item = i;
      j.key = TuneKey{w.cw, w.ch, w.cf, coarse_hash[w.coarse_entry]};
      j.mu = -1.0;
      if (tune_lookup(j.key, mu_of[i])) continue;
      jobs.push_back(std::move(j));
    }
    if (!jobs.empty()) {
      // the masks not in the cache, from the device (one readback)
      std::vector<uint8_t> masks(coarse_total);
      HIP_OK(hipMemcpyAsync(masks.data(), coarse_d, coarse_total, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      for (Job &j : jobs) {
        const ItemWork &w = iw[j.item];
        j.in.assign(masks.begin() + w.coarse_off,
                    masks.begin() + w.coarse_off + (size_t)w.cw * w.ch);
      }
      const int nth = (int)std::min<size_t>(jobs.size(),
                                            std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
      std::atomic<size_t> next{0};
      auto worker = [&]() {
        for (size_t q; (q = next.fetch_add(1)) < jobs.size();) {
          Job &j = jobs[q];
          const ItemWork &w = iw[j.item];
          j.mu = coarse_min_eigenvalue(j.in, w.cw, w.ch) / ((double)w.cf * w.cf);
        }
      };
      std::vector<std::thread> pool;
      for (int t = 1; t < nth; ++t) pool.emplace_back(worker);
      worker();
      for (auto &th : pool) th.join();
      for (const Job &j : jobs) {
        mu_of[j.item] = j.mu;
        tune_store(j.key, j.mu);
      }
    }
  }
  for (int i = 0; i < nitem; ++i) {
    ItemWork &w = iw[i];
    if (!w.any) continue;
    w.omega = tuned ? omega_for(mu_of[i], d->lambda) : d->omega;
    w.stall_window = stall_window_for(w.omega, tuned);
    w.in.rep->omega_used = w.omega;
  }

  phase("tuner");
  // --- 4. per-channel planes, assembly and r0 (one readback)
  std::vector<InstWork> inst;
  for (int i = 0; i < nitem; ++i) {
    ItemWork &w = iw[i];
    if (!w.any) continue;
    const size_t plane = (size_t)w.pitch * w.phh;
    w.cls_plane = g_pool.take<uint8_t>(plane);
    w.inst0 = (int)inst.size();
    for (int c = 0; c < w.in.C; ++c) {
      InstWork X;
      X.item = i;
      X.channel = c;
      PhInst &I = X.I;
      I.e[0] = g_pool.take<float>(plane);
      I.e[1] = g_pool.take<float>(plane);
      I.g = g_pool.take<float>(plane);
      I.x = g_pool.take<double>(plane);
      I.bp = g_pool.take<double>(plane);
      I.cls = w.cls_plane;
      I.pitch = w.pitch;
      I.pw = w.pw;
      I.ph = w.phh;
      I.ox = w.ox;
      I.oy = w.oy;
      I.img_w = w.in.W;
      I.img_h = w.in.H;
      I.channels = w.in.C;
      I.channel = c;
      I.dg64 = 4.0 + d->lambda;
      I.dg = (float)I.dg64;
      I.omega = w.omega;
      I.ramp_log = 0.0;
      I.ramp_sweeps = 0;
      I.lambda = d->lambda;
      X.pgrd = dim3((w.pitch + 255) / 256, w.phh);
      X.hist = g_pool.take<double>(std::max(d->max_iterations, 1u));
      inst.push_back(X);
    }
  }
  const int ninst = (int)inst.size();  // the images' instances (twins come after)
  const int ncap = std::max(race_any ? 2 * ninst : ninst, 1);
  double *resv_d = g_pool.take<double>(2 * (size_t)ncap);
  double *resv = g_pinned.take<double>(2 * (size_t)ncap);
  // assembly and the anchor-start residual of every instance: one launch per
  // step over all plane rows (inst_all_d is the instances' descriptors)
  PhInst *inst_all_d = g_pool.take<PhInst>(std::max(ninst, 1));
  PhSpan *inst_rows_d = g_pool.take<PhSpan>(std::max(ninst, 1));
  PhSpan *img_planes_d = g_pool.take<PhSpan>(std::max(nitem, 1));
  SpanTable inst_rows, img_planes;
  std::vector<PhInst> inst_all(ninst);
  for (int k = 0; k < ninst; ++k) {
    inst_all[k] = inst[k].I;
    inst_rows.add(k, inst[k].I.ph, inst[k].item, 0, k);
  }
  // (an image's entry carries its first instance's first row in the
  // instances' row sums: the assembly takes the anchor residual too)
  for (int i = 0; i < nitem; ++i)
    if (iw[i].any) img_planes.add(iw[i].inst0, iw[i].phh, i, 0, inst_rows.h[iw[i].inst0].row0);
  if (ninst) {
    for (int i = 0; i < nitem; ++i) io[i].cls_plane = iw[i].cls_plane;
    double *inst_sums_d = g_pool.take<double>((size_t)inst_rows.rows);
    HIP_OK(hipMemcpyAsync(io_d, io.data(), nitem * sizeof(PhItemIO), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(inst_all_d, inst_all.data(), ninst * sizeof(PhInst),
                          hipMemcpyHostToDevice, st));
    inst_rows.upload(inst_rows_d, st);
    img_planes.upload(img_planes_d, st);
    launch_by_channels<AssembleL>(common_channels(img_planes, iw), dim3(img_planes.rows), st,
                                  (const PhInst *)inst_all_d, (const PhSpan *)img_planes_d,
                                  (int)img_planes.h.size(), (const PhItemIO *)io_d, inst_sums_d);
    hipLaunchKernelGGL(k_residual_norm_spans, dim3(ninst), dim3(256), 0, st, inst_rows_d,
                       inst_sums_d, resv_d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(resv, resv_d, 2 * ninst * sizeof(double), hipMemcpyDeviceToHost,
                          st));
    HIP_OK(hipStreamSynchronize(st));
  }

  phase("assembly");
  // --- 5. instances that need sweeps
  std::vector<int> act;
  for (int k = 0; k < ninst; ++k) {
    InstWork &X = inst[k];
    X.r0 = resv[2 * k];
    X.R = ChanResult{};
    if (X.r0 == 0.0) continue;  // poisson.rs:177-189
    if (!std::isfinite(X.r0)) {
      // Non-finite inputs: the reference's residuals are NaN, no sweep ever
      // improves on the best, and the stall guard ends it after 16 sweeps.
      X.R.nonfinite = true;
      X.R.iterations = std::min(PH_STALL_WINDOW, d->max_iterations);
      X.R.stop = X.R.iterations == PH_STALL_WINDOW ? PH_STOP_STALLED : PH_STOP_MAX_ITERATIONS;
      X.R.converged = false;
      X.R.final_residual = NAN;
      continue;
    }
    X.R.active = true;
    X.R.done = false;
    act.push_back(k);
    iw[X.item].in.rep->unknowns += (uint64_t)iw[X.item].n_interior;
  }

  // --- 5b. learned warm start: x = P + bilinear(net(coarse inputs)) at the
  // interior cells of instance k, raced against a plain twin (x = anchor) that
  // runs the same first sweeps in the same launches. At the probe sweep the
  // solver's own FP64 residuals decide: the smaller one continues, the other
  // is dropped; when the plain twin wins, its state replaces the warm one, so
  // the result is the plain solve's, bit for bit. Both keep the anchor-start
  // denominator r0 (poisson.rs:172).
  std::vector<double> rstart;
  std::vector<int> twin, probe;
  std::vector<char> is_twin, silent;
  // coarse-correction cycles of network instances: next cycle end (0 = none),
  // cycle length, corrections kept, residual at the previous cycle end
  std::vector<uint32_t> cyc_next, cyc_len, cyc_count, cyc_undone, cyc_weak;
  std::vector<double> cyc_prev;
  auto grow = [&]() {
    const size_t n = inst.size();
    rstart.resize(n, 0.0);
    twin.resize(n, -1);
    probe.resize(n, 0);
    is_twin.resize(n, 0);
    silent.resize(n, 0);
    cyc_next.resize(n, 0);
    cyc_len.resize(n, 0);
    cyc_count.resize(n, 0);
    cyc_undone.resize(n, 0);
    cyc_weak.resize(n, 0);
    cyc_prev.resize(n, 0.0);
  };
  grow();
  for (int k = 0; k < ninst; ++k) rstart[k] = inst[k].r0;
  // candidates: images with a caller's start field, and (with the network
  // flag) every other Poisson-blend instance
  std::vector<int> cand, net_list;
  uint64_t largest = 0;
  for (int k : act) largest = std::max(largest, (uint64_t)iw[inst[k].item].n_interior);
  const uint64_t warm_min_cells = warm_gate("PAINTOP_HIP_WARM_MIN_CELLS", largest);
  for (int k : act) {
    if (iw[inst[k].item].in.start) cand.push_back(k);
    else if (want_warm && d->lambda == 0.0 &&
             (uint64_t)iw[inst[k].item].n_interior >= warm_min_cells) {
      cand.push_back(k);
      net_list.push_back(k);
    }
  }
  // coarse-grid framing of each network instance (kept for the corrections)
  struct Geo { int x0, y0, f; };
  std::vector<Geo> wgeo(ninst, Geo{0, 0, 1});
  const int NR = warm_net_res(), NNB = warm_net_batch();
  if (NR > PH_CORR_RES_MAX)
    throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "network grid %d exceeds %d", NR,
               PH_CORR_RES_MAX);
  const size_t NR2 = (size_t)NR * NR;
  float *net_in = nullptr, *net_h = nullptr, *net_scale = nullptr;
  double *net_part = nullptr;
  if (!net_list.empty()) {
    net_in = g_pool.take<float>(NR2 * 4 * NNB);
    // one prediction and one input scale per network instance (the warm
    // starts, then the corrections of any subset of them)
    net_h = g_pool.take<float>(NR2 * net_list.size());
    net_scale = g_pool.take<float>(net_list.size());
    net_part = g_pool.take<double>((size_t)3 * PH_NET_SLICES * NNB);
  }
  const uint32_t max_corr = warm_max_corrections();
  const uint64_t corr_min_cells = warm_gate("PAINTOP_HIP_CORR_MIN_CELLS", largest);
  // host staging of the warm start (kept until the stream is synchronised)
  std::vector<PhCorrArgs> warm_args;
  std::vector<PhInst> anchor_from, anchor_to;
  SpanTable warm_rows, warm_groups, cand_rows, anchor_rows;
  if (!cand.empty()) {
    const int nw = (int)cand.size();
    hipStream_t side = nullptr;
    PhCorrArgs *warm_args_d = nullptr;
    PhSpan *warm_spans_d = nullptr;
    std::function<void()> launch_net;  // the side stream's work, enqueued after the copy
    if (!net_list.empty()) {
      const int nn = (int)net_list.size();
      for (int k : net_list) {
        const ItemWork &w = iw[inst[k].item];
        const int f = std::max(1, (std::max(w.rw, w.rh) + NR - 1) / NR);
        wgeo[k] = Geo{w.rx0 + w.rw / 2 - (NR * f) / 2, w.ry0 + w.rh / 2 - (NR * f) / 2, f};
      }
      // inputs, network and application for every network instance: one
      // coarsening launch per network batch, one application launch
      warm_args.resize(nn);
      warm_rows.clear();
      for (int j = 0; j < nn; ++j) {
        const int k = net_list[j];
        const Geo &g = wgeo[k];
        warm_args[j] = PhCorrArgs{k, g.x0, g.y0, g.f, 0, j, inst[k].item, 0};
        warm_rows.add(k, inst[k].I.ph, inst[k].item);
      }
      // the coarsening takes an image's channels together: groups of up to 4
      // consecutive entries of one image inside one network batch, each with
      // its blocks (row0: the group's first block within its batch's launch)
      warm_groups.clear();
      std::vector<int> group0, blocks0;  // per batch: first group, blocks of its launch
      for (int j0 = 0; j0 < nn; j0 += NNB) {
        group0.push_back((int)warm_groups.h.size());
        const int j1 = std::min(nn, j0 + NNB);
        int nblk = 0;
        for (int j = j0; j < j1;) {
          int e = j + 1;
          while (e < j1 && e - j < 4 && warm_args[e].item == warm_args[j].item) ++e;
          warm_groups.h.push_back(PhSpan{j, nblk, e - j, 0, 0, warm_args[j].item, nullptr, 0.0});
          nblk += coarsen_blocks(warm_args[j].f, NR);
          j = e;
        }
        blocks0.push_back(nblk);
      }
      group0.push_back((int)warm_groups.h.size());
      warm_args_d = g_pool.take<PhCorrArgs>(nn);
      warm_spans_d = g_pool.take<PhSpan>(nn);
      PhSpan *warm_groups_d = g_pool.take<PhSpan>(nn);
      HIP_OK(hipMemcpyAsync(warm_args_d, warm_args.data(), nn * sizeof(PhCorrArgs),
                            hipMemcpyHostToDevice, st));
      warm_rows.upload(warm_spans_d, st);
      warm_groups.upload(warm_groups_d, st);
      // the inputs (HBM-bound) here; the network passes (latency-bound: small
      // launches) on a second stream beside the copy of the twins' anchor
      // state (HBM-bound) that follows. One network batch's input buffer
      // holds NNB inputs, so batches past the first wait for the previous pass.
      side = side_stream();
      const int nbatch = (nn + NNB - 1) / NNB;
      if (nbatch == 1) {
        hipLaunchKernelGGL(k_coarsen_groups, dim3(blocks0[0]), dim3(256), 0, st, inst_all_d,
                           warm_args_d, warm_groups_d, group0[1] - group0[0], io_d, d->lambda, NR,
                           0, net_in);
        hipLaunchKernelGGL(k_in_norm_part, dim3(PH_NET_SLICES, nn), dim3(256), 0, st, net_in, NR,
                           net_part);
        hipLaunchKernelGGL(k_in_norm_apply, dim3(PH_NET_SLICES, nn), dim3(256), 0, st, net_in, NR,
                           net_part, net_scale);
        HIP_OK(hipGetLastError());
      }
      launch_net = [=, &side]() {
        HIP_OK(hipStreamWaitEvent(side, side_events().first, 0));
        for (int j0 = 0, bi = 0; j0 < nn; j0 += NNB, ++bi) {
          const int nb = std::min(NNB, nn - j0);
          if (nbatch > 1) {
            hipLaunchKernelGGL(k_coarsen_groups, dim3(blocks0[bi]), dim3(256), 0, side,
                               inst_all_d, warm_args_d, warm_groups_d + group0[bi],
                               group0[bi + 1] - group0[bi], io_d, d->lambda, NR, j0, net_in);
            hipLaunchKernelGGL(k_in_norm_part, dim3(PH_NET_SLICES, nb), dim3(256), 0, side,
                               net_in, NR, net_part);
            hipLaunchKernelGGL(k_in_norm_apply, dim3(PH_NET_SLICES, nb), dim3(256), 0, side,
                               net_in, NR, net_part, net_scale + j0);
            HIP_OK(hipGetLastError());
          }
          warm_net_forward(net_in, net_h + NR2 * j0, nb, side);
        }
        HIP_OK(hipEventRecord(side_events().second, side));
      };
    }
    // the anchor-start state of every candidate, copied before its start
    // overwrites it: the plain twin of the race, or the way back to the anchor
    // start for a start rejected outright (the source term b' is shared)
    for (int j = 0; j < nw; ++j) {
      const InstWork &X = inst[cand[j]];
      const size_t plane = (size_t)X.I.pitch * X.I.ph;
      PhInst A = X.I;
      A.e[0] = g_pool.take<float>(plane);
      A.e[1] = g_pool.take<float>(plane);
      A.g = g_pool.take<float>(plane);
      A.x = g_pool.take<double>(plane);
      anchor_from.push_back(X.I);
      anchor_to.push_back(A);
      anchor_rows.add(j, X.I.ph, X.item);
    }
    {
      PhInst *from_d = g_pool.take<PhInst>(nw);
      PhInst *to_d = g_pool.take<PhInst>(nw);
      PhSpan *rows_d = g_pool.take<PhSpan>(nw);
      HIP_OK(hipMemcpyAsync(from_d, anchor_from.data(), nw * sizeof(PhInst),
                            hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(to_d, anchor_to.data(), nw * sizeof(PhInst), hipMemcpyHostToDevice,
                            st));
      anchor_rows.upload(rows_d, st);
      // the side stream starts behind every upload, beside the copy
      if (launch_net) HIP_OK(hipEventRecord(side_events().first, st));
      hipLaunchKernelGGL(k_copy_anchor_spans, dim3(anchor_rows.rows), dim3(256), 0, st, from_d,
                         to_d, rows_d, nw);
      HIP_OK(hipGetLastError());
    }
    if (launch_net) launch_net();
    if (!net_list.empty()) {
      // the predictions join the solve's stream, then every start is applied
      HIP_OK(hipStreamWaitEvent(st, side_events().second, 0));
      hipLaunchKernelGGL(k_warm_apply_spans, dim3(warm_rows.rows), dim3(256), 0, st, inst_all_d,
                         warm_spans_d, (int)net_list.size(), warm_args_d, io_d, net_h, net_scale,
                         NR);
      HIP_OK(hipGetLastError());
    }
    // the caller's start fields, then the starts' FP64 residuals in three launches
    cand_rows.clear();
    for (int k : cand) {
      const InstWork &X = inst[k];
      if (iw[X.item].in.start) {
        hipLaunchKernelGGL(k_start_apply, X.pgrd, dim3(256), 0, st, X.I, iw[X.item].in.start);
        HIP_OK(hipGetLastError());
      }
      cand_rows.add(k, X.I.ph, X.item, 0, k);
    }
    PhSpan *cand_spans_d = g_pool.take<PhSpan>(nw);
    double *cand_sums_d = g_pool.take<double>((size_t)cand_rows.rows);
    cand_rows.upload(cand_spans_d, st);
    hipLaunchKernelGGL(k_residual_spans, dim3(cand_rows.rows), dim3(256), 0, st, inst_all_d,
                       cand_spans_d, nw, cand_sums_d);
    hipLaunchKernelGGL(k_residual_norm_spans, dim3(nw), dim3(256), 0, st, cand_spans_d,
                       cand_sums_d, resv_d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(resv, resv_d, 2 * ninst * sizeof(double), hipMemcpyDeviceToHost,
                          st));
    HIP_OK(hipStreamSynchronize(st));
    for (int j = 0; j < nw; ++j) {
      const int k = cand[j];
      const ItemWork &w = iw[inst[k].item];
      const double rw = resv[2 * k];
      ph_report *rep = w.in.rep;
      const double r0 = inst[k].r0;
      rep->warm_start_rel = std::max(rep->warm_start_rel, std::isfinite(rw) ? rw / r0 : INFINITY);
      if (!(std::isfinite(rw) && rw <= WARM_MAX_START_RATIO * r0)) {
        // no race for a start this far off: back to the anchor start at once
        inst[k].I = anchor_to[j];
        continue;
      }
      // the plain twin: the anchor-start copy
      InstWork T = inst[k];
      T.I = anchor_to[j];
      T.hist = g_pool.take<double>(std::max(d->max_iterations, 1u));
      inst.push_back(T);
      const int kt = (int)inst.size() - 1;
      grow();
      InstWork &Xw = inst[k];
      const bool cycles = !w.in.start && max_corr > 0 && (uint64_t)w.n_interior >= corr_min_cells;
      const uint32_t n = warm_ramp_sweeps(w.omega, cycles);
      Xw.I.ramp_sweeps = n;
      Xw.I.ramp_log = n ? log(2.0 - w.omega) / n : 0.0;
      twin[k] = kt;
      twin[kt] = k;
      is_twin[kt] = 1;
      rstart[k] = rw;
      rstart[kt] = r0;
      // the probe ends on a launch boundary (a multiple of K/2 sweeps) that a
      // plain solve also passes through, so the plain member's trajectory is
      // the plain solve's, bit for bit
      const uint32_t half = (uint32_t)(K / 2);
      const uint32_t pr = (warm_probe_sweeps(n) + half - 1) / half * half;
      probe[k] = probe[kt] = (int)std::min<uint32_t>(pr, d->max_iterations);
      act.push_back(kt);
      if (cycles) {
        // the first cycle ends past the probe, so the race is settled first
        cyc_len[k] = warm_cycle_sweeps(n, half);
        cyc_next[k] = std::max<uint32_t>(cyc_len[k], (uint32_t)probe[k] + half);
      }
    }
  }

  phase("warm");
  const int nact = (int)act.size();
  if (nact > 0) {
    // the tile list of the running instances (rebuilt as instances finish)
    std::vector<PhInst> ai(nact);
    for (int a = 0; a < nact; ++a) ai[a] = inst[act[a]].I;
    size_t ntile_cap = 0;
    for (int a = 0; a < nact; ++a) ntile_cap += iw[inst[act[a]].item].live.size();
    PhTile *tiles_d = g_pool.take<PhTile>(ntile_cap);
    PhInst *inst_d = g_pool.take<PhInst>(nact);
    PhCtl *ctl_d = g_pool.take<PhCtl>(nact);
    double *partials_d = g_pool.take<double>(ntile_cap * K);
    // refinement passes: one span table for all pending instances (three
    // launches a pass)
    size_t row_cap = 0;
    for (int a = 0; a < nact; ++a) row_cap += (size_t)inst[act[a]].I.ph;
    double *ref_sums_d = g_pool.take<double>(row_cap);
    PhSpan *ref_spans_d = g_pool.take<PhSpan>(nact);
    // correction rounds: arguments and span tables of the corrections and
    // of the take-backs, the history entries of the kept ones
    PhCorrArgs *corr_args_d = nullptr;
    PhSpan *corr_spans_d = nullptr;
    PhValue *corr_vals_d = nullptr;
    // a round's results in one array, read back in one copy: per entry j the
    // corrected state's norm and its square (2j, 2j + 1), from 2 nn on the
    // steps, from 3 nn on the norms of the states before the corrections
    const size_t nn = net_list.size();
    double *corr_rsq_d = nullptr, *round_d = nullptr, *corr_alpha = nullptr,
           *corr_rnorm_d = nullptr;
    if (nn) {
      corr_args_d = g_pool.take<PhCorrArgs>(nact);
      corr_spans_d = g_pool.take<PhSpan>(nact);
      corr_vals_d = g_pool.take<PhValue>(nact);
      // per coarse cell of each correction: the sum of r(x + e)^2 (a round
      // corrects only network instances)
      corr_rsq_d = g_pool.take<double>(NR2 * nn);
      round_d = g_pool.take<double>(4 * nn);
      corr_alpha = round_d + 2 * nn;
      corr_rnorm_d = round_d + 3 * nn;
    }
    SpanTable ref;
    std::vector<PhTile> tl;
    int ntiles = 0;
    auto build_tiles = [&](const std::vector<char> &running) {
      tl.clear();
      for (int a = 0; a < nact; ++a) {
        ai[a].tile_begin = (int)tl.size();
        if (running[a]) {
          const ItemWork &w = iw[inst[act[a]].item];
          for (size_t q = 0; q < w.live.size(); ++q)
            tl.push_back(PhTile{a, w.live[q].x, w.live[q].y,
                                (int)(tl.size() << 1) | (w.live_interior[q] ? PH_TILE_INTERIOR : 0)});
        }
        ai[a].tile_end = (int)tl.size();
      }
      ntiles = (int)tl.size();
      // Dispatch order: within each XCD's range, the masked tiles first. A
      // launch of about one tile per wave slot then puts at most one masked
      // tile (the slower kind) on a SIMD where the ranges allow, instead of
      // two on some; the partial sums keep instance order (the slot in flags).
      for (int x = 0; x < 8; ++x)
        std::stable_partition(tl.begin() + xcd_range_begin(x, ntiles),
                              tl.begin() + xcd_range_begin(x + 1, ntiles),
                              [](const PhTile &t) { return !(t.flags & PH_TILE_INTERIOR); });
      if (trace_on()) {
        int nin = 0;
        for (const PhTile &t : tl) nin += t.flags & PH_TILE_INTERIOR;
        fprintf(stderr, "[tiles] %d tiles of %d rows, %d all interior\n", ntiles, hb, nin);
      }
      if (ntiles)
        HIP_OK(hipMemcpyAsync(tiles_d, tl.data(), ntiles * sizeof(PhTile), hipMemcpyHostToDevice,
                              st));
      HIP_OK(hipMemcpyAsync(inst_d, ai.data(), nact * sizeof(PhInst), hipMemcpyHostToDevice, st));
      HIP_OK(hipStreamSynchronize(st));  // tl and ai are host staging
    };
    std::vector<char> running(nact, 1);
    build_tiles(running);

    PhCtl *ctl = g_pinned.take<PhCtl>(nact);
    // pinned staging of the loop's small transfers (each region is written
    // once between two synchronisations of the stream)
    PhSpan *ref_stage = g_pinned.take<PhSpan>(nact);
    PhSpan *corr_stage = g_pinned.take<PhSpan>(nact);
    PhCorrArgs *args_stage = g_pinned.take<PhCorrArgs>(nact);
    PhValue *vals_stage = g_pinned.take<PhValue>(nact);
    double *round_h = g_pinned.take<double>(4 * (size_t)nact);
    for (int a = 0; a < nact; ++a) {
      const InstWork &X = inst[act[a]];
      const ItemWork &w = iw[X.item];
      PhCtl &Cc = ctl[a];
      memset(&Cc, 0, sizeof(Cc));
      Cc.r0 = X.r0;
      Cc.tol = d->tolerance;
      const double rs = rstart[act[a]];
      Cc.refine_abs2 = (REFINE_RATIO * rs) * (REFINE_RATIO * rs);
      Cc.best = INFINITY;
      Cc.history = X.hist;
      Cc.max_it = d->max_iterations;
      Cc.sweep_limit = probe[act[a]] ? (uint32_t)probe[act[a]] : d->max_iterations;
      Cc.halt = PH_RUN;
      Cc.mode = PH_MODE_NORMAL;
      // the stall guard's window starts after a warm start's omega ramp
      Cc.stall_window = w.stall_window + X.I.ramp_sweeps;
      Cc.refine_every = REFINE_EVERY;
      Cc.last_refine = 0;
    }
    HIP_OK(hipMemcpyAsync(ctl_d, ctl, nact * sizeof(PhCtl), hipMemcpyHostToDevice, st));

    auto finish = [&](int a, int stop, uint32_t m, double rel) {
      ChanResult &R = inst[act[a]].R;
      R.done = true;
      R.iterations = m;
      R.stop = stop;
      R.converged = stop == PH_STOP_CONVERGED;
      R.final_residual = rel;
      ctl[a].halt = PH_HALT_DONE;
    };
    // stall rule of the reference for a host-evaluated sweep (poisson.rs:220-228)
    auto stall_step = [&](int a, PhCtl &Cc, uint32_t m, double rel) -> bool {
      if (ph_in_ramp(ai[a], Cc.ramp_from, m)) return false;  // on hold inside an omega ramp
      if (rel < Cc.best * (1.0 - PH_STALL_IMPROVEMENT)) {
        Cc.best = rel;
        Cc.since = 0;
        return false;
      }
      Cc.since += 1;
      return Cc.since >= Cc.stall_window;
    };
    // warm-start race state: members held at the probe, races decided, the
    // probe residuals, the extensions taken and the members awaiting a probe
    std::vector<char> held(nact, 0), hold_refine(nact, 0), resolved(inst.size(), 0);
    std::vector<double> pval(inst.size(), -1.0);
    std::vector<int> race_ext(inst.size(), 0), probe_pend;
    // sweep limit of a running instance: the end of its correction cycle, if any
    auto run_limit = [&](int a) -> uint32_t {
      const PhCtl &Cc = ctl[a];
      const int k = act[a];
      uint32_t lim = Cc.max_it;
      const uint32_t c = cyc_next[k];
      if (c > Cc.sweeps_done) lim = std::min(lim, c);
      // an open warm-start race holds both members at its probe
      if (twin[k] >= 0 && !resolved[k] && (uint32_t)probe[k] > Cc.sweeps_done)
        lim = std::min<uint32_t>(lim, (uint32_t)probe[k]);
      return lim;
    };
    auto resume = [&](int a, double gnorm) {
      PhCtl &Cc = ctl[a];
      Cc.last_refine = Cc.sweeps_done;
      Cc.halt = PH_RUN;
      Cc.mode = PH_MODE_NORMAL;
      Cc.sweep_limit = run_limit(a);
      Cc.refine_pending = 0;
      Cc.refine_abs2 = (REFINE_RATIO * gnorm) * (REFINE_RATIO * gnorm);
    };

    const bool tracing = trace_on();
    std::vector<int> slot(inst.size(), -1);
    for (int a = 0; a < nact; ++a) slot[act[a]] = a;
    int batch = 4;
    int remaining = nact;
    double guard = 0;
    uint64_t last_progress = ~0ull;
    int stuck = 0;
    hipEvent_t eb0 = nullptr, eb1 = nullptr;
    if (timing) {
      HIP_OK(hipEventCreate(&eb0));
      HIP_OK(hipEventCreate(&eb1));
    }
    struct Pending {
      int a;
      uint32_t m;
      int reason;
    };
    std::vector<Pending> pend;
    struct Trend {
      uint32_t m = 0;
      double rel = 0.0;
    };
    std::vector<Trend> trend(nact);
    uint32_t total_launches = 0, total_refine = 0;
    double sweep_ms = 0.0;
    // Learned coarse corrections at the cycle ends. An entry comes either from
    // a refinement at its cycle end (gnorm = its FP64 residual, e = 0) or from
    // a sweep halt at its cycle end with no refinement due (gnorm < 0: the
    // residual of sweep m is not known yet). All entries of a pass share the
    // network batches and one readback.
    struct Corr {
      int a;
      uint32_t m;
      double gnorm;
    };
    std::vector<Corr> corr, cyc_end;
    // host staging of the correction rounds (kept until the stream is synchronised)
    std::vector<PhCorrArgs> corr_args;
    std::vector<PhValue> corr_vals;
    SpanTable corr_rows;
    // a cycle that ends without a usable correction: the cycles end
    auto end_cycles = [&](int a, uint32_t m, double rel, const char *why) {
      const int k = act[a];
      cyc_next[k] = 0;
      ctl[a].sweep_limit = run_limit(a);
      if (trace_on())
        fprintf(stderr, "[corr] inst %d: cycles end at sweep %u (rel %.3e, %u kept, %s)\n", k, m,
                rel, cyc_count[k], why);
    };
    // the gain rules on the residual at a cycle end: false ends the cycles
    auto cycle_gains = [&](int k, double rel) -> bool {
      // (after a correction taken back, the cycle ran uncorrected: no test)
      bool gain = true;
      if (cyc_count[k] > 0 && !cyc_undone[k]) {
        const bool weak = rel * CORR_GOOD_GAIN > cyc_prev[k];
        gain = rel * CORR_MIN_GAIN <= cyc_prev[k] && !(weak && cyc_weak[k]);
        cyc_weak[k] = weak ? 1 : 0;
      }
      cyc_prev[k] = rel;
      return gain && cyc_count[k] < max_corr;
    };
    auto correct = [&]() {
      // refined entries: cycles that stopped paying for themselves end here,
      // without a correction
      std::vector<Corr> go;
      for (const Corr &c : corr) {
        const double rel = c.gnorm / inst[act[c.a]].r0;
        if (!(c.gnorm > 0.0) || !cycle_gains(act[c.a], rel)) {
          end_cycles(c.a, c.m, rel, "gain");
          continue;
        }
        go.push_back(c);
      }
      // unrefined entries run the round speculatively; their sweep-m residual
      // comes back with it and decides afterwards
      for (const Corr &c : cyc_end) go.push_back(c);
      corr.clear();
      cyc_end.clear();
      if (go.empty()) return;
      const int nc = (int)go.size();
      // the round's tables depend on its instances only (the kernels take
      // each state's buffer from its control block): uploaded when they change
      bool same = (int)corr_args.size() == nc;
      for (int j = 0; same && j < nc; ++j) same = corr_args[j].slot == go[j].a;
      if (!same) {
        corr_args.resize(nc);
        corr_rows.clear();
        for (int j = 0; j < nc; ++j) {
          const int a = go[j].a, k = act[a];
          const Geo &g = wgeo[k];
          corr_args[j] = PhCorrArgs{a, g.x0, g.y0, g.f, coarsen_plane_cpb(g.f), j, inst[k].item,
                                    0};
          corr_rows.add(a, ai[a].ph, 0, 0, j);
        }
        upload_pinned(corr_args_d, corr_args.data(), nc, args_stage, st);
        corr_rows.upload(corr_spans_d, st, corr_stage);
      }
      for (int c0 = 0; c0 < nc; c0 += NNB) {
        const int nb = std::min(NNB, nc - c0);
        int gx = 1;
        for (int j = c0; j < c0 + nb; ++j)
          gx = std::max(gx, (NR + corr_args[j].cpb - 1) / corr_args[j].cpb);
        hipLaunchKernelGGL(k_corr_input_list, dim3(gx, NR, nb), dim3(256), 0, st, inst_d,
                           (const PhCtl *)ctl_d, corr_args_d + c0, NR, net_in,
                           corr_rsq_d + NR2 * c0);
        hipLaunchKernelGGL(k_in_norm_part, dim3(PH_NET_SLICES, nb), dim3(256), 0, st, net_in, NR,
                           net_part);
        hipLaunchKernelGGL(k_in_norm_apply, dim3(PH_NET_SLICES, nb), dim3(256), 0, st, net_in, NR,
                           net_part, net_scale + c0);
        HIP_OK(hipGetLastError());
        warm_net_forward(net_in, net_h + NR2 * c0, nb, st);
        hipLaunchKernelGGL(k_corr_alpha_part, dim3(PH_NET_SLICES, nb), dim3(256), 0, st, net_in,
                           net_h + NR2 * c0, (const double *)(corr_rsq_d + NR2 * c0), NR,
                           net_part);
        hipLaunchKernelGGL(k_corr_alpha, dim3((nb + 63) / 64), dim3(64), 0, st, net_part, nb,
                           CORR_MAX_STEP, corr_alpha + c0, corr_rnorm_d + c0);
        HIP_OK(hipGetLastError());
      }
      // every correction into the other correction buffer and the residual
      // of the corrected state (the norm of the state each entry's sweeps
      // left came with its step)
      hipLaunchKernelGGL(k_corr_apply_spans, dim3(corr_rows.rows), dim3(256), 0, st, inst_d,
                         (const PhCtl *)ctl_d, corr_spans_d, nc, corr_args_d, net_h, net_scale,
                         corr_alpha, NR, ref_sums_d);
      hipLaunchKernelGGL(k_residual_norm_spans, dim3(nc), dim3(256), 0, st, corr_spans_d,
                         ref_sums_d, round_d);
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpyAsync(round_h, round_d, 4 * nn * sizeof(double), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      const double *alpha = round_h + 2 * nn, *rnorm = round_h + 3 * nn;
      const double max_jump = corr_max_jump();
      corr_vals.clear();
      for (int j = 0; j < nc; ++j) {
        const int a = go[j].a, k = act[a];
        const uint32_t m = go[j].m;
        PhCtl &Cc = ctl[a];
        double gnorm = go[j].gnorm;
        int hist_at = -1;  // this entry's history value of sweep m in corr_vals
        if (gnorm < 0.0) {
          // sweep m of an unrefined entry: its residual, ||L e - g|| in FP64
          // from the stored state (the grid covers every interior cell)
          gnorm = rnorm[j];
          const double rel = gnorm / Cc.r0;
          // a stop, a stall or a refinement due at sweep m: the FP64 rules of
          // a refinement decide at m instead (the correction is not taken;
          // the refinement's pass queues the cycle's correction again)
          const bool hold = ph_in_ramp(ai[a], Cc.ramp_from, m);
          const bool better = rel < Cc.best * (1.0 - PH_STALL_IMPROVEMENT);
          const uint32_t since = (hold || better) ? Cc.since : Cc.since + 1;
          if (!(rel > Cc.tol) || !std::isfinite(rel) || since >= Cc.stall_window ||
              (!hold && !better && since == (Cc.stall_window >> 1)) ||
              gnorm * gnorm <= Cc.refine_abs2) {
            Cc.halt = PH_HALT_REFINE;
            if (trace_on())
              fprintf(stderr, "[corr] inst %d sweep %u: rel %.3e, refinement first\n", k, m, rel);
            continue;
          }
          // the reference's rules for sweep m (no rule fires)
          if (!hold) {
            if (better) Cc.best = rel;
            Cc.since = since;
          }
          Cc.sweeps_final = m;
          Cc.last_rel = rel;
          hist_at = (int)corr_vals.size();
          corr_vals.push_back(PhValue{Cc.history + (m - 1), rel});
          if (!cycle_gains(k, rel)) {
            end_cycles(a, m, rel, "gain");
            Cc.halt = PH_RUN;
            continue;
          }
        }
        const double jump = round_h[2 * j] / gnorm;
        const bool usable = alpha[j] != 0.0;
        const bool undo = usable && !(std::isfinite(jump) && jump <= max_jump);
        const double rel_new = round_h[2 * j] / Cc.r0;
        if (trace_on())
          fprintf(stderr, "[corr] inst %d sweep %u: rel %.3e -> %.3e (step %.3f%s)\n", k, m,
                  gnorm / Cc.r0, rel_new, alpha[j], undo ? ", taken back" : "");
        if (!usable || (undo && cyc_undone[k])) {
          end_cycles(a, m, gnorm / Cc.r0, usable ? "taken back twice" : "no step");
        } else if (undo) {
          // the correction buffer is not switched: the state stays the sweeps'
          cyc_undone[k] = 1;
          cyc_next[k] = m + cyc_len[k];
        } else {
          cyc_undone[k] = 0;
          cyc_count[k] += 1;
          iw[inst[k].item].in.rep->coarse_corrections += 1;
          cyc_next[k] = m + cyc_len[k];
          Cc.cur_buf ^= 1u;  // the corrected state
          // the omega ramp restarts at the corrected state
          Cc.ramp_from = m;
          if (hist_at >= 0) corr_vals[hist_at].val = rel_new;
          else corr_vals.push_back(PhValue{Cc.history + (m - 1), rel_new});
          // the residual dips and then rises through the restarted ramp: the
          // stall guard starts again after it, the decay estimate from here
          Cc.best = INFINITY;
          Cc.since = 0;
          Cc.sweeps_final = m;
          Cc.last_rel = rel_new;
          trend[a] = Trend{m, rel_new};
          if (!(rel_new > Cc.tol)) {
            // converged on the correction: a refinement (x += e) confirms it
            // with the FP64 rules
            Cc.halt = PH_HALT_REFINE;
            continue;
          }
        }
        Cc.halt = PH_RUN;
        Cc.mode = PH_MODE_NORMAL;
        Cc.sweep_limit = run_limit(a);
      }
      // the restarted ramps and the history entries of sweep m
      if (!corr_vals.empty()) {
        upload_pinned(corr_vals_d, corr_vals.data(), corr_vals.size(), vals_stage, st);
        hipLaunchKernelGGL(k_put_values, dim3((corr_vals.size() + 255) / 256), dim3(256), 0, st,
                           corr_vals_d, (int)corr_vals.size());
        HIP_OK(hipGetLastError());
      }
    };

    while (remaining > 0) {
      // no instance running and one waiting for its refinement: no launch
      // (the host copy of the control blocks is the current one)
      bool any_run = false, any_refine = false;
      for (int a = 0; a < nact; ++a) {
        if (inst[act[a]].R.done || held[a]) continue;
        any_run = any_run || ctl[a].halt == PH_RUN;
        any_refine = any_refine || ctl[a].halt == PH_HALT_REFINE;
      }
      if (any_run || !any_refine) {
        if (timing) HIP_OK(hipEventRecord(eb0, st));
        launch_sor_batch(K, PF, ntiles, nact, st, tiles_d, inst_d, ctl_d, partials_d, hb, batch);
        total_launches += (uint32_t)batch;
        if (timing) HIP_OK(hipEventRecord(eb1, st));
        HIP_OK(hipMemcpyAsync(ctl, ctl_d, nact * sizeof(PhCtl), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (timing) {
          float bms = 0.f;
          HIP_OK(hipEventElapsedTime(&bms, eb0, eb1));
          sweep_ms += bms;
        }
      }
      // pass 1: every instance that needs an FP64 evaluation gets its
      // refinement kernels; one readback serves all of them
      pend.clear();
      for (int a = 0; a < nact; ++a) {
        PhCtl &Cc = ctl[a];
        if (inst[act[a]].R.done || held[a]) continue;
        switch (Cc.halt) {
          case PH_RUN:
            break;
          case PH_HALT_STOP:
            if (Cc.halt_sweep > Cc.halt_base) {  // re-run up to the stop sweep
              Cc.mode = PH_MODE_REDO;
              Cc.sweeps_done = Cc.halt_base;
              Cc.sweep_limit = Cc.halt_sweep;
              Cc.cur_buf = Cc.halt_buf;
              Cc.halt = PH_RUN;
            } else {
              Cc.cur_buf = Cc.halt_buf;
              pend.push_back({a, Cc.halt_sweep, (int)Cc.reason});
            }
            break;
          case PH_HALT_REDO_DONE:
            pend.push_back({a, Cc.halt_sweep, (int)Cc.reason});
            break;
          case PH_HALT_LIMIT:
            if (twin[act[a]] >= 0 && !resolved[act[a]] && Cc.sweeps_done < Cc.max_it) {
              // warm-start race probe: judged on the FP64 residual of the
              // member's state (x + e), computed without a refinement so
              // neither member's trajectory changes
              held[a] = 1;
              hold_refine[a] = Cc.refine_pending ? 1 : 0;
              probe_pend.push_back(a);
              if (tracing)
                fprintf(stderr, "[race] hold inst %d (%s) at sweep %u (final %u) rel %.3e\n",
                        act[a], is_twin[act[a]] ? "plain" : "warm", Cc.sweeps_done,
                        Cc.sweeps_final, Cc.last_rel);
              break;
            }
            if (cyc_next[act[a]] && Cc.sweeps_done >= cyc_next[act[a]] && !Cc.refine_pending &&
                Cc.mode == PH_MODE_NORMAL && Cc.sweeps_done < Cc.max_it &&
                !(Cc.last_rel * Cc.r0 * CYCLE_REFINE_AHEAD * Cc.last_rel * Cc.r0 *
                      CYCLE_REFINE_AHEAD <=
                  Cc.refine_abs2)) {
              // a cycle end with no refinement due: the correction round reads
              // the sweeps' state as it is
              cyc_end.push_back(Corr{a, Cc.sweeps_done, -1.0});
              break;
            }
            pend.push_back({a, Cc.sweeps_done, -1});
            break;
          case PH_HALT_REFINE:
            pend.push_back({a, Cc.sweeps_done, -1});
            break;
          default:
            throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "unexpected halt state %u", Cc.halt);
        }
      }
      if (!pend.empty()) {
        // x += e, e = 0 and the FP64 residual of every pending instance; the
        // norm kernel also writes each one's history entry ||r|| / r0 of sweep m
        ref.clear();
        for (const Pending &P : pend) {
          const PhCtl &Cc = ctl[P.a];
          ref.add(P.a, ai[P.a].ph, 0, (int)Cc.cur_buf, act[P.a], Cc.history + (P.m - 1), Cc.r0);
        }
        ref.upload(ref_spans_d, st, ref_stage);
        const int nref = (int)ref.h.size();
        hipLaunchKernelGGL(k_apply_correction_spans, dim3(ref.rows), dim3(256), 0, st, inst_d,
                           ref_spans_d, nref);
        hipLaunchKernelGGL(k_residual_spans, dim3(ref.rows), dim3(256), 0, st, inst_d, ref_spans_d,
                           nref, ref_sums_d);
        hipLaunchKernelGGL(k_residual_norm_spans, dim3(nref), dim3(256), 0, st, ref_spans_d,
                           ref_sums_d, resv_d);
        HIP_OK(hipGetLastError());
        // every instance's slot, the warm-start race's plain twins included
        HIP_OK(hipMemcpyAsync(resv, resv_d, 2 * inst.size() * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        total_refine += (uint32_t)pend.size();
        // pass 2: the reference's rules on the exact FP64 residual of sweep m
        for (size_t q = 0; q < pend.size(); ++q) {
          const Pending &P = pend[q];
          PhCtl &Cc = ctl[P.a];
          const double gnorm = resv[2 * act[P.a]];
          const double rel = gnorm / Cc.r0;  // the value k_residual_norm_spans wrote
          if (tracing)
            fprintf(stderr, "[refine] inst %d sweep %u: rel %.3e (halt %u, est %.3e)\n", act[P.a],
                    P.m, rel, Cc.halt, Cc.last_rel);
          Cc.sweeps_final = P.m;
          Cc.sweeps_done = P.m;
          const int k = act[P.a];
          const bool paired = twin[k] >= 0 && !resolved[k];
          if (rel <= Cc.tol) finish(P.a, PH_STOP_CONVERGED, P.m, rel);
          else if (P.reason == PH_REASON_STALLED) finish(P.a, PH_STOP_STALLED, P.m, rel);
          else if (stall_step(P.a, Cc, P.m, rel)) finish(P.a, PH_STOP_STALLED, P.m, rel);
          else if (P.m >= Cc.max_it) finish(P.a, PH_STOP_MAX_ITERATIONS, P.m, rel);
          else resume(P.a, gnorm);
          if (inst[k].R.done) {
            --remaining;
            if (paired) pval[k] = gnorm;
            if (paired && tracing)
              fprintf(stderr, "[race] inst %d (%s) finished at sweep %u rel %.3e\n", k,
                      is_twin[k] ? "plain" : "warm", P.m, rel);
          } else if (paired && tracing && P.m < (uint32_t)probe[k]) {
            fprintf(stderr, "[race] inst %d (%s) evaluated at sweep %u before the probe\n", k,
                    is_twin[k] ? "plain" : "warm", P.m);
          } else if (!paired && cyc_next[k] && P.m >= cyc_next[k]) {
            corr.push_back(Corr{P.a, P.m, gnorm});
          }
        }
      }
      if (!corr.empty() || !cyc_end.empty()) correct();
      if (!probe_pend.empty()) {
        // the FP64 residuals of the race members held at their probe
        ref.clear();
        for (int a : probe_pend)
          ref.add(a, ai[a].ph, 0, (int)ctl[a].cur_buf, act[a], nullptr, ctl[a].r0);
        ref.upload(ref_spans_d, st, ref_stage);
        const int nref = (int)ref.h.size();
        hipLaunchKernelGGL(k_probe_residual_spans, dim3(ref.rows), dim3(256), 0, st, inst_d,
                           ref_spans_d, nref, ref_sums_d);
        hipLaunchKernelGGL(k_residual_norm_spans, dim3(nref), dim3(256), 0, st, ref_spans_d,
                           ref_sums_d, resv_d);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(resv, resv_d, 2 * inst.size() * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (int a : probe_pend) {
          pval[act[a]] = resv[2 * act[a]];
          if (tracing)
            fprintf(stderr, "[race] inst %d (%s) at sweep %u: fp64 %.3e\n", act[a],
                    is_twin[act[a]] ? "plain" : "warm", ctl[a].sweeps_done,
                    resv[2 * act[a]] / ctl[a].r0);
        }
        probe_pend.clear();
      }
      {
        // warm-start races whose two members both have a value: the smaller
        // residual continues (a tie goes to the plain start); a warm lead
        // smaller than RACE_MARGIN extends the race while both members run
        for (int k = 0; k < (int)inst.size(); ++k) {
          if (is_twin[k] || twin[k] < 0 || resolved[k]) continue;
          const int kt = twin[k];
          if (pval[k] < 0.0 || pval[kt] < 0.0) continue;
          const bool warm_wins = pval[k] < pval[kt];
          if (warm_wins && pval[k] > RACE_MARGIN * pval[kt] && race_ext[k] < RACE_MAX_EXTENSIONS &&
              !inst[k].R.done && !inst[kt].R.done && held[slot[k]] && held[slot[kt]] &&
              (uint32_t)probe[k] < d->max_iterations) {
            race_ext[k] += 1;
            probe[k] = probe[kt] = (int)std::min<uint32_t>(2u * (uint32_t)probe[k],
                                                           d->max_iterations);
            if (tracing)
              fprintf(stderr, "[race] extend inst %d: warm %.3e plain %.3e -> probe %d\n", k,
                      pval[k], pval[kt], probe[k]);
            pval[k] = pval[kt] = -1.0;
            for (const int a : {slot[k], slot[kt]}) {
              held[a] = 0;
              PhCtl &Cc = ctl[a];
              Cc.sweep_limit = run_limit(a);
              Cc.halt = hold_refine[a] ? PH_HALT_REFINE : PH_RUN;
            }
            continue;
          }
          const int win = warm_wins ? k : kt, lose = warm_wins ? kt : k;
          if (tracing)
            fprintf(stderr, "[race] judge inst %d: warm %.3e (sweep %u) plain %.3e (sweep %u) -> %s\n",
                    k, pval[k], ctl[slot[k]].sweeps_done, pval[kt], ctl[slot[kt]].sweeps_done,
                    warm_wins ? "warm" : "plain");
          resolved[k] = resolved[kt] = 1;
          silent[lose] = 1;
          if (!inst[lose].R.done) {
            inst[lose].R.done = true;
            inst[lose].R.iterations = ctl[slot[lose]].sweeps_done;
            ctl[slot[lose]].halt = PH_HALT_DONE;
            --remaining;
          }
          if (held[slot[win]]) {
            const int a = slot[win];
            held[a] = 0;
            PhCtl &Cc = ctl[a];
            if (win != k) cyc_next[k] = 0;
            Cc.sweep_limit = run_limit(a);
            // a refinement that was due at the probe launch happens now, as it
            // would have without the probe
            Cc.halt = hold_refine[a] ? PH_HALT_REFINE : PH_RUN;
          }
          if (warm_wins) iw[inst[k].item].in.rep->warm_start_used += 1;
        }
      }
      HIP_OK(hipMemcpyAsync(ctl_d, ctl, nact * sizeof(PhCtl), hipMemcpyHostToDevice, st));
      if (remaining == 0) break;
      // drop the tiles of finished instances once they are most of the list
      {
        size_t dead = 0;
        for (int a = 0; a < nact; ++a)
          if (running[a] && inst[act[a]].R.done) dead += ai[a].tile_end - ai[a].tile_begin;
        if (dead > 0 && 2 * dead >= (size_t)ntiles) {
          for (int a = 0; a < nact; ++a) running[a] = !inst[act[a]].R.done;
          build_tiles(running);
        }
      }
      // next batch: up to the next scheduled halt (refinement or sweep limit) of
      // the running instances, so no launch finds every instance halted
      // and, once the residual decays steadily, not far past the sweep at
      // which the observed decay rate reaches the tolerance
      {
        uint32_t need = 0;
        for (int a = 0; a < nact; ++a) {
          const PhCtl &Cc = ctl[a];
          if (inst[act[a]].R.done || Cc.halt != PH_RUN || held[a]) continue;
          uint32_t next = std::min(Cc.sweep_limit, Cc.last_refine + Cc.refine_every);
          if (Cc.mode == PH_MODE_REDO) next = Cc.sweep_limit;
          uint32_t todo = next > Cc.sweeps_done ? next - Cc.sweeps_done : 1;
          if (Cc.mode == PH_MODE_NORMAL && ai[a].ramp_sweeps != 0 &&
              Cc.sweeps_done >= Cc.ramp_from &&
              Cc.sweeps_done - Cc.ramp_from < ai[a].ramp_sweeps / 4 && Cc.last_rel > 0.0 &&
              Cc.last_rel < NEAR_TOLERANCE * Cc.tol)
            todo = std::min<uint32_t>(todo, (uint32_t)(K / 2));
          Trend &T = trend[a];
          if (Cc.mode == PH_MODE_NORMAL && Cc.sweeps_final >= T.m + 16 && T.rel > 0.0 &&
              Cc.last_rel > 0.0 && Cc.last_rel < T.rel) {
            const double rate = log(Cc.last_rel / T.rel) / (double)(Cc.sweeps_final - T.m);
            const double left = log(Cc.tol / Cc.last_rel) / rate;  // sweeps to tolerance
            const double lead = (double)Cc.sweeps_done - (double)Cc.sweeps_final;
            const double est = std::max(1.0, 1.15 * left - lead + 8.0);
            if (est < (double)todo) todo = (uint32_t)est;
          }
          if (Cc.sweeps_final >= T.m + 16) {
            T.m = Cc.sweeps_final;
            T.rel = Cc.last_rel;
          }
          need = std::max(need, (todo + (uint32_t)(K / 2) - 1) / (uint32_t)(K / 2));
        }
        batch = (int)std::max<uint32_t>(1, std::min<uint32_t>(need, 64));
      }
      // every batch must move some instance forward (sweeps or halt state)
      uint64_t progress = 0;
      for (int a = 0; a < nact; ++a)
        progress += (uint64_t)ctl[a].sweeps_done * 8u + ctl[a].halt + ctl[a].launches * 131u +
                    (inst[act[a]].R.done ? 7919u : 0u);
      if (progress == last_progress && ++stuck >= 3)
        throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "sweep control made no progress");
      if (progress != last_progress) stuck = 0;
      last_progress = progress;
      guard += 1;
      if (guard > 4.0 * MAX_ITERATIONS_LIMIT)
        throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "solve loop did not terminate");
    }
    HIP_OK(hipStreamSynchronize(st));  // the correction rounds' host staging
    if (timing) {
      hipEventDestroy(eb0);
      hipEventDestroy(eb1);
    }
    for (int a = 0; a < nact; ++a) {
      const int k = act[a];
      InstWork &X = inst[k];
      // every sweep run counts as work, the dropped member's included
      iw[X.item].in.rep->unknown_sweeps += (uint64_t)iw[X.item].n_interior * X.R.iterations;
      if (is_twin[k] || silent[k]) continue;
      if (twin[k] >= 0 && silent[k] == 0 && resolved[k] == 0)
        throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "warm-start race left unresolved");
    }
    for (int k = 0; k < ninst; ++k) {
      if (twin[k] >= 0 && silent[k]) {  // the plain twin won: its result is the image's
        InstWork &X = inst[k];
        const InstWork &T = inst[twin[k]];
        X.I = T.I;
        X.R = T.R;
        X.hist = T.hist;
      }
    }
    // write-back of every solved instance in one launch (a plain twin's
    // planes stand in for its image's instance where the twin won)
    // (each image's channels together: C consecutive descriptors per image,
    // with no planes for the channels that were not solved)
    SpanTable out_rows;
    std::vector<PhInst> out_inst;
    std::vector<int> group(nitem, -1);
    for (int a = 0; a < nact; ++a) {
      const int k = act[a];
      if (k >= ninst || !inst[k].R.active) continue;
      const int i = inst[k].item;
      if (group[i] < 0) {
        group[i] = (int)out_inst.size();
        PhInst blank = inst[k].I;
        blank.x = nullptr;
        out_inst.insert(out_inst.end(), iw[i].in.C, blank);
        out_rows.add(group[i], inst[k].I.ph, i);
      }
      out_inst[group[i] + inst[k].channel] = inst[k].I;
    }
    if (!out_inst.empty()) {
      HIP_OK(hipMemcpyAsync(inst_all_d, out_inst.data(), out_inst.size() * sizeof(PhInst),
                            hipMemcpyHostToDevice, st));
      out_rows.upload(inst_rows_d, st);
      launch_by_channels<ScatterL>(common_channels(out_rows, iw), dim3(out_rows.rows), st,
                                   (const PhInst *)inst_all_d, (const PhSpan *)inst_rows_d,
                                   (int)out_rows.h.size(), (const PhItemIO *)io_d);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(st));  // out_rows and out_inst are host staging
    }
    for (int i = 0; i < nitem; ++i) {
      iw[i].in.rep->launches = total_launches;
      iw[i].in.rep->refinements = total_refine;
      iw[i].in.rep->sweep_ms = sweep_ms;
    }
  }
  phase("sweeps");
  // non-finite channels: interior samples become NaN (the reference's
  // residuals and field are NaN there)
  for (int k = 0; k < ninst; ++k)
    if (inst[k].R.nonfinite) {
      PhInst I = inst[k].I;
      I.x = I.bp;  // reuse: fill with NaN then scatter
      std::vector<double> nanv((size_t)I.pitch * I.ph, NAN);
      HIP_OK(hipMemcpyAsync(I.x, nanv.data(), nanv.size() * sizeof(double), hipMemcpyHostToDevice,
                            st));
      hipLaunchKernelGGL(k_scatter, inst[k].pgrd, dim3(256), 0, st, I, iw[inst[k].item].in.out);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(st));
    }

  // --- worst channel of each image (poisson_blend.rs:240-246: more
  // iterations, then larger final residual; ties keep the earlier channel);
  // the reported histories are gathered for one readback
  std::vector<PhHistCopy> hcopy;
  std::vector<double *> hdst;
  int64_t htotal = 0;
  for (int i = 0; i < nitem; ++i) {
    ItemWork &w = iw[i];
    ph_report *rep = w.in.rep;
    ChanResult zero;
    int worst = 0;
    const ChanResult *Wr = &zero;
    double *whist = nullptr;
    if (w.any) {
      for (int c = 0; c < w.in.C; ++c) {
        const InstWork &X = inst[w.inst0 + c];
        const bool worse = c == 0 ? true
                           : X.R.iterations != Wr->iterations
                               ? X.R.iterations > Wr->iterations
                               : X.R.final_residual > Wr->final_residual;
        if (worse) {
          worst = c;
          Wr = &X.R;
          whist = X.hist;
        }
      }
    }
    rep->worst_channel = (uint32_t)worst;
    rep->iterations = Wr->iterations;
    rep->stop_reason = Wr->stop;
    rep->converged = Wr->converged ? 1 : 0;
    rep->tolerance = d->tolerance;
    rep->final_residual = Wr->final_residual;
    rep->history_len = Wr->iterations;
    if (rep->residual_history && rep->history_capacity >= Wr->iterations && Wr->iterations > 0) {
      if (Wr->nonfinite) {
        for (uint32_t q = 0; q < Wr->iterations; ++q) rep->residual_history[q] = NAN;
      } else {
        hcopy.push_back(PhHistCopy{whist, htotal, (int32_t)Wr->iterations, 0});
        hdst.push_back(rep->residual_history);
        htotal += Wr->iterations;
      }
    } else if (Wr->iterations > 0 && rep->residual_history) {
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "history capacity %u < %u iterations",
                 rep->history_capacity, Wr->iterations);
    }
  }
  std::vector<double> hbuf((size_t)htotal);
  if (!hcopy.empty()) {
    PhHistCopy *hcopy_d = g_pool.take<PhHistCopy>(hcopy.size());
    double *hbuf_d = g_pool.take<double>((size_t)htotal);
    HIP_OK(hipMemcpyAsync(hcopy_d, hcopy.data(), hcopy.size() * sizeof(PhHistCopy),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gather_hist, dim3(hcopy.size()), dim3(256), 0, st, hcopy_d, hbuf_d);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(hbuf.data(), hbuf_d, (size_t)htotal * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  }
  if (timing) {
    HIP_OK(hipEventRecord(ev1, st));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    for (int i = 0; i < nitem; ++i) iw[i].in.rep->device_ms = ms;
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
  }
  HIP_OK(hipStreamSynchronize(st));
  for (size_t q = 0; q < hcopy.size(); ++q)
    memcpy(hdst[q], hbuf.data() + hcopy[q].off, (size_t)hcopy[q].n * sizeof(double));
  phase("reports");
}

static void solve_device(const ph_solve_desc *d, const float *src, const float *tgt,
                         const float *mask, float *out, hipStream_t st, ph_report *rep,
                         const float *start = nullptr) {
  std::vector<ItemIn> items{
      ItemIn{(int)d->width, (int)d->height, (int)d->channels, src, tgt, mask, out, rep, start}};
  solve_items(d, items, st);
}

}  // namespace ph

// ---------------------------------------------------------------------------
// C ABI

using namespace ph;

// ---------------------------------------------------------------------------
// allocation-free entry point

// The workspace of ph_poisson_solve_device_noalloc: a fixed device array of
// PH_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends. Host code never passes `scratch` itself (in host code the name
// stands for a host-side shadow, not a device address): the device address
// comes from hipGetSymbolAddress, once per device, and is kept per device.
static __device__ __align__(256) char scratch[PH_NOALLOC_SCRATCH_BYTES];
static void *g_scratch_dev[64];
// Host staging of the allocation-free calls' readbacks (pageable).
static char g_noalloc_host[1 << 20] __attribute__((aligned(256)));

static void *noalloc_scratch() {
  const int dev = g_ctx.dev;  // current device (ensure_device made it current)
  if (dev < 0 || dev >= 64)
    throw fail(PH_CLASS_POLICY, "E_HIP_LIMIT", "device index %d out of range", dev);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw fail(PH_CLASS_EXECUTION, "E_HIP_RUNTIME", "hipGetSymbolAddress(scratch) failed: %s",
                 hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

// Routes the pools to the fixed blocks for the duration of one call.
struct NoAllocScope {
  explicit NoAllocScope(void *dev_block) {
    g_pool.ext = static_cast<char *>(dev_block);
    g_pool.ext_cap = PH_NOALLOC_SCRATCH_BYTES;
    g_pinned.ext = g_noalloc_host;
    g_pinned.ext_cap = sizeof(g_noalloc_host);
    g_noalloc = true;
  }
  ~NoAllocScope() {
    g_pool.ext = nullptr;
    g_pool.ext_cap = 0;
    g_pinned.ext = nullptr;
    g_pinned.ext_cap = 0;
    g_noalloc = false;
  }
};

extern "C" int32_t ph_abi_version(void) { return PH_ABI_VERSION; }

extern "C" void ph_force_unavailable(int32_t forced) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_forced = forced ? 1 : 0;
}

extern "C" void ph_sweep_config(double work, uint32_t *half_sweeps, uint32_t *tile_rows) {
  ph_solve_desc d{};
  int K = 0, hb = 0;
  sweep_config(&d, work, K, hb);
  if (half_sweeps) *half_sweeps = (uint32_t)K;
  if (tile_rows) *tile_rows = (uint32_t)hb;
}

extern "C" ph_status ph_probe(ph_device_info *out, ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    ensure_device();
    if (out) {
      memset(out, 0, sizeof(*out));
      snprintf(out->name, sizeof(out->name), "%s",
               g_ctx.prop.name[0] ? g_ctx.prop.name : "AMD Instinct (gfx950)");
      snprintf(out->arch, sizeof(out->arch), "%s", g_ctx.prop.gcnArchName);
      out->compute_units = g_ctx.prop.multiProcessorCount;
      out->device_index = g_ctx.dev;
      out->hbm_bytes = g_ctx.prop.totalGlobalMem;
      int rv = 0, dv = 0;
      hipRuntimeGetVersion(&rv);
      hipDriverGetVersion(&dv);
      out->runtime_version = rv;
      out->driver_version = dv;
    }
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return f.code == "E_GPU_UNAVAILABLE" ? PH_UNAVAILABLE : PH_ERROR;
  }
}

static ph_status run_guarded(const ph_solve_desc *desc, ph_report *rep, ph_error *err,
                             const std::function<void()> &body, bool check_shape = true) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    if (!rep) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null report");
    validate(desc, check_shape);
    ensure_device();
    body();
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return f.code == "E_GPU_UNAVAILABLE" ? PH_UNAVAILABLE : PH_ERROR;
  } catch (const std::exception &e) {
    put_error(err, fail(PH_CLASS_EXECUTION, "E_HIP_INTERNAL", "%s", e.what()));
    return PH_ERROR;
  }
}

extern "C" ph_status ph_poisson_solve_device(const ph_solve_desc *desc, const float *d_source,
                                             const float *d_target, const float *d_mask,
                                             float *d_out, void *stream, ph_report *rep,
                                             ph_error *err) {
  return run_guarded(desc, rep, err, [&]() {
    if ((size_t)desc->width * desc->height > 0 && (!d_source || !d_target || !d_mask || !d_out))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    solve_device(desc, d_source, d_target, d_mask, d_out, (hipStream_t)stream, rep);
  });
}

extern "C" ph_status ph_poisson_solve_device_noalloc(const ph_solve_desc *desc,
                                                     const float *d_source, const float *d_target,
                                                     const float *d_mask, float *d_out,
                                                     void *stream, ph_report *rep,
                                                     ph_error *err) {
  return run_guarded(desc, rep, err, [&]() {
    if (desc->flags & (PH_FLAG_WARM_START_NET | PH_FLAG_TIMING))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_NOALLOC_FLAG",
                 "the allocation-free solve takes neither PH_FLAG_WARM_START_NET nor "
                 "PH_FLAG_TIMING");
    if ((size_t)desc->width * desc->height > 0 && (!d_source || !d_target || !d_mask || !d_out))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    void *blk = noalloc_scratch();
    NoAllocScope scope(blk);
    solve_device(desc, d_source, d_target, d_mask, d_out, (hipStream_t)stream, rep);
  });
}

extern "C" ph_status ph_poisson_solve_start_device(const ph_solve_desc *desc,
                                                   const float *d_source, const float *d_target,
                                                   const float *d_mask, const float *d_start,
                                                   float *d_out, void *stream, ph_report *rep,
                                                   ph_error *err) {
  return run_guarded(desc, rep, err, [&]() {
    if ((size_t)desc->width * desc->height > 0 &&
        (!d_source || !d_target || !d_mask || !d_out || !d_start))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    solve_device(desc, d_source, d_target, d_mask, d_out, (hipStream_t)stream, rep, d_start);
  });
}

extern "C" ph_status ph_poisson_solve(const ph_solve_desc *desc, const float *source,
                                      const float *target, const float *mask, float *out,
                                      ph_report *rep, ph_error *err) {
  return run_guarded(desc, rep, err, [&]() {
    const size_t npx = (size_t)desc->width * desc->height;
    const size_t n = npx * desc->channels;
    if (npx > 0 && (!source || !target || !mask || !out))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    float *buf = nullptr;
    const size_t bytes = (2 * n + npx + n) * sizeof(float) + 1024;
    HIP_OK(hipMalloc(&buf, bytes));
    struct Free {
      float *p;
      ~Free() { hipFree(p); }
    } guard{buf};
    float *ds = buf, *dt = ds + n, *dm = dt + n, *dout = dm + npx;
    hipStream_t st = nullptr;
    if (npx > 0) {
      HIP_OK(hipMemcpyAsync(ds, source, n * sizeof(float), hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(dt, target, n * sizeof(float), hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(dm, mask, npx * sizeof(float), hipMemcpyHostToDevice, st));
    }
    solve_device(desc, ds, dt, dm, dout, st, rep);
    if (npx > 0) {
      HIP_OK(hipMemcpyAsync(out, dout, n * sizeof(float), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
    }
  });
}

extern "C" ph_status ph_poisson_solve_batch_device(const ph_solve_desc *desc,
                                                   const ph_batch_item *items, uint32_t n_items,
                                                   void *stream, ph_report *reports,
                                                   ph_error *err) {
  return run_guarded(
      desc, reports, err,
      [&]() {
        if (n_items > 0 && !items) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null items");
        std::vector<ItemIn> v;
        v.reserve(n_items);
        for (uint32_t i = 0; i < n_items; ++i) {
          const ph_batch_item &b = items[i];
          validate_shape(desc, b.width, b.height, b.channels);
          if ((size_t)b.width * b.height > 0 &&
              (!b.source || !b.target || !b.mask || !b.out_candidate))
            throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "item %u: null buffer", i);
          v.push_back(ItemIn{(int)b.width, (int)b.height, (int)b.channels, b.source, b.target,
                             b.mask, b.out_candidate, &reports[i]});
        }
        if (!v.empty()) solve_items(desc, v, (hipStream_t)stream);
      },
      false);
}
