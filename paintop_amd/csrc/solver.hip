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
#include <mutex>
#include <string>
#include <vector>

#include "../../include/paintop_hip.h"
#include "ph_internal.h"

namespace ph {

template <int K, int PF>
__global__ void k_sor(const PhTile *, int, const PhInst *, const PhCtl *, double *, int);
template <int K>
__global__ void k_sor_control(const PhInst *, PhCtl *, const double *, int);
__global__ void k_classify(const float *, int, int, uint8_t *, int *);
__global__ void k_tile_occupancy(const uint8_t *, int, int, int, int, int, int, int, const int2 *,
                                 int *);
__global__ void k_coarse_count(const uint8_t *, int, int, int, int, int, int, int, int *);
__global__ void k_assemble(PhInst, const float *, const float *, const uint8_t *, uint8_t *);
__global__ void k_apply_correction(PhInst, int);
__global__ void k_residual(PhInst, double *, uint32_t *, double *);
__global__ void k_scatter(PhInst, float *);

// ---------------------------------------------------------------------------
// errors

struct Fail {
  int cls;
  std::string code;
  std::string msg;
};

static Fail fail(int cls, const char *code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return Fail{cls, code, buf};
}

#define HIP_OK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      throw fail(PH_CLASS_EXECUTION, "E_HIP_RUNTIME", "%s failed: %s (%s:%d)", #expr, \
                 hipGetErrorString(_e), __FILE__, __LINE__);                           \
  } while (0)

static void put_error(ph_error *err, const Fail &f) {
  if (!err) return;
  err->error_class = f.cls;
  snprintf(err->code, sizeof(err->code), "%s", f.code.c_str());
  snprintf(err->message, sizeof(err->message), "%s", f.msg.c_str());
}

static void clear_error(ph_error *err) {
  if (!err) return;
  err->error_class = -1;
  err->code[0] = 0;
  err->message[0] = 0;
}

// ---------------------------------------------------------------------------
// device context (one per process, behind a mutex)

static std::mutex g_mu;
static int g_forced = 0;

struct Ctx {
  bool probed = false;
  bool ok = false;
  std::string why;
  int dev = 0;
  hipDeviceProp_t prop{};
};
static Ctx g_ctx;

static bool forced_unavailable() {
  if (g_forced) return true;
  const char *e = getenv("PAINTOP_HIP_FORCE_UNAVAILABLE");
  return e && e[0] == '1';
}

static void ensure_device() {
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

// Grow-only device workspace: one allocation reused across calls.
struct Pool {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void reset(size_t need) {
    if (need > cap) {
      if (base) HIP_OK(hipFree(base));
      base = nullptr;
      cap = 0;
      HIP_OK(hipMalloc(&base, need));
      cap = need;
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

// ---------------------------------------------------------------------------
// parameter validation (the same ranges and codes as resolve_controls,
// poisson_blend.rs:102-139 and screened_poisson.rs:66-112)

static const uint32_t MAX_ITERATIONS_LIMIT = 100000;  // poisson_blend.rs:71

static void validate(const ph_solve_desc *d) {
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
  if (d->channels == 0 || d->channels > 64)
    throw fail(PH_CLASS_SEMANTIC, screened ? "E_SCREENED_POISSON_SHAPE" : "E_POISSON_BLEND_SHAPE",
               "%s: channel count %u out of range", opn, d->channels);
  const uint32_t k = d->half_sweeps;
  if (k != 0 && k != 2 && k != 4 && k != 8 && k != 12 && k != 16)
    throw fail(PH_CLASS_SCHEMA, "E_HIP_TUNING", "half_sweeps must be one of 2, 4, 8, 12, 16");
  if (d->tile_rows != 0 && (d->tile_rows < 16 || d->tile_rows > 1024 || (d->tile_rows & 1)))
    throw fail(PH_CLASS_SCHEMA, "E_HIP_TUNING", "tile_rows must be 0 or even in 16..1024");
  if ((size_t)d->width * d->height > ((size_t)1 << 31))
    throw fail(PH_CLASS_POLICY, "E_HIP_LIMIT", "image of %ux%u exceeds the 2^31-pixel limit",
               d->width, d->height);
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
// large omega, where SOR degrades gently). Results are cached per coarse mask.
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
    const double stop = 1e-16 * dot(v, v);
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

// mu (fine grid) of the interior cells in region [rx0, rx0+rw) x [ry0, ry0+rh).
static double tuned_mu(const uint8_t *cls_img, int W, int H, int rx0, int ry0, int rw, int rh,
                       hipStream_t st) {
  const int f = std::max(1, (std::max(rw, rh) + 127) / 128);
  const int cw = (rw + f - 1) / f, ch = (rh + f - 1) / f;
  int *counts_d = g_pool.take<int>((size_t)cw * ch);
  hipLaunchKernelGGL(k_coarse_count, dim3((cw * ch + 255) / 256), dim3(256), 0, st, cls_img, W, H,
                     rx0, ry0, f, cw, ch, counts_d);
  HIP_OK(hipGetLastError());
  std::vector<int> counts((size_t)cw * ch);
  HIP_OK(hipMemcpyAsync(counts.data(), counts_d, counts.size() * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  std::vector<uint8_t> in(counts.size());
  uint64_t hsh = 1469598103934665603ull;  // FNV-1a of the coarse mask
  for (size_t i = 0; i < counts.size(); ++i) {
    in[i] = 2 * counts[i] >= f * f ? 1 : 0;
    hsh = (hsh ^ in[i]) * 1099511628211ull;
  }
  const TuneKey key{cw, ch, f, hsh};
  for (const auto &e : g_tune_cache)
    if (e.first == key) return e.second;
  const double mu = coarse_min_eigenvalue(in, cw, ch) / ((double)f * f);
  if (g_tune_cache.size() > 256) g_tune_cache.erase(g_tune_cache.begin());
  g_tune_cache.push_back({key, mu});
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
static uint32_t stall_window_for(double omega, bool tuned) {
  if (!tuned) return PH_STALL_WINDOW;
  const double w = ceil(3.0 / std::max(2.0 - omega, 1e-6));
  return (uint32_t)std::max<double>(PH_STALL_WINDOW, std::min(w, 1e6));
}

// One sweep launch: the tiles, then the per-instance control.
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
    case 2: hipLaunchKernelGGL(k_sor_control<2>, cgrid, cblock, 0, st, insts, ctl, partials, ninst); break;
    case 4: hipLaunchKernelGGL(k_sor_control<4>, cgrid, cblock, 0, st, insts, ctl, partials, ninst); break;
    case 8: hipLaunchKernelGGL(k_sor_control<8>, cgrid, cblock, 0, st, insts, ctl, partials, ninst); break;
    case 12: hipLaunchKernelGGL(k_sor_control<12>, cgrid, cblock, 0, st, insts, ctl, partials, ninst); break;
    default: hipLaunchKernelGGL(k_sor_control<16>, cgrid, cblock, 0, st, insts, ctl, partials, ninst); break;
  }
  HIP_OK(hipGetLastError());
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
  uint64_t interior = 0;
};

// Refinement triggers: the inner residual fell 1e3-fold, or REFINE_EVERY
// sweeps passed. The second bounds |e| by the error at the start of the window,
// so each sweep's change to e (~ (2 - omega) x error) stays far above
// ulp(|e|): an FP32 correction that accumulated the whole solve would stop
// absorbing the late, small updates.
static const double REFINE_RATIO = 1e-3;
static const uint32_t REFINE_EVERY = 1024;

static void solve_device(const ph_solve_desc *d, const float *src, const float *tgt,
                         const float *mask, float *out, hipStream_t st, ph_report *rep) {
  const int W = (int)d->width, H = (int)d->height, C = (int)d->channels;
  const size_t npx = (size_t)W * H;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const bool timing = (d->flags & PH_FLAG_TIMING) != 0;
  if (timing) {
    HIP_OK(hipEventCreate(&ev0));
    HIP_OK(hipEventCreate(&ev1));
    HIP_OK(hipEventRecord(ev0, st));
  }
  std::vector<ChanResult> res(C);
  rep->launches = 0;
  rep->refinements = 0;
  rep->unknowns = 0;
  rep->unknown_sweeps = 0;
  rep->warm_start_used = 0;
  rep->warm_start_rel = 0.0;
  rep->sweep_ms = 0.0;

  if (npx == 0) {
    rep->iterations = 0;
    rep->stop_reason = PH_STOP_CONVERGED;
    rep->converged = 1;
    rep->history_len = 0;
    rep->tolerance = d->tolerance;
    rep->final_residual = 0.0;
    rep->worst_channel = 0;
    rep->device_ms = 0.0;
    return;
  }
  // the candidate starts as the target: every non-interior sample keeps its bits
  HIP_OK(hipMemcpyAsync(out, tgt, npx * C * sizeof(float), hipMemcpyDeviceToDevice, st));

  // --- classification and bounding box
  const int K = d->half_sweeps ? (int)d->half_sweeps : 16;
  const int CPL = 4, WS = 64 * CPL, WV = WS - 2 * K;
  int hb = d->tile_rows ? (int)d->tile_rows : 48;
  {
    // upper bound of the workspace for this call, so the pool never regrows mid-call
    const size_t pw_up = (size_t)W + 2 * PH_PAD + 4 + WV + 2 * K + 64;
    const size_t ph_up = (size_t)H + 2 * PH_PAD + 4 + hb + 2 * K;
    const size_t plane_up = pw_up * ph_up;
    const size_t ntile_up = (pw_up / WV + 2) * (ph_up / hb + 2);
    const size_t need = npx + 16 * 1024 + ntile_up * (sizeof(int2) + sizeof(int)) +
                        plane_up * (1 + (size_t)C * (3 * sizeof(float) + 2 * sizeof(double))) +
                        (size_t)(C + 2) * (d->max_iterations + 1) * sizeof(double) +
                        (size_t)C * ph_up * sizeof(double) +
                        ntile_up * (size_t)C * (sizeof(PhTile) + K * sizeof(double)) +
                        (size_t)C * (sizeof(PhInst) + sizeof(PhCtl) + 64) + 64 * 256 +
                        130 * 130 * sizeof(int);
    g_pool.reset(need);
  }
  uint8_t *cls_img = g_pool.take<uint8_t>(npx);
  int *bbox = g_pool.take<int>(4);
  int bbox_h[4] = {0x7fffffff, 0x7fffffff, -1, -1};
  HIP_OK(hipMemcpyAsync(bbox, bbox_h, sizeof(bbox_h), hipMemcpyHostToDevice, st));
  {
    dim3 blk(256), grd((W + 255) / 256, H);
    hipLaunchKernelGGL(k_classify, grd, blk, 0, st, mask, W, H, cls_img, bbox);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipMemcpyAsync(bbox_h, bbox, sizeof(bbox_h), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));

  const bool any_interior = bbox_h[2] >= 0;
  std::vector<PhInst> inst_h;
  std::vector<int> inst_chan;
  std::vector<PhTile> tiles_h;
  std::vector<int2> chan_origins;
  double n_interior = 0;
  int region_w = 0, region_h = 0;
  int ox = 0, oy = 0, pw = 0, phh = 0, pitch = 0;

  if (any_interior) {
    int rx0 = bbox_h[0] - 1, ry0 = bbox_h[1] - 1;
    const int rx1 = bbox_h[2] + 2, ry1 = bbox_h[3] + 2;
    if ((rx0 + ry0) & 1) rx0 -= 1;  // red/black colour of plane cells = image colour
    region_w = rx1 - rx0;
    region_h = ry1 - ry0;
    ox = rx0 - PH_PAD;
    oy = ry0 - PH_PAD;
    const int nstrips = (PH_PAD + region_w - K + WV - 1) / WV;
    const int nblocks = (PH_PAD + region_h - K + hb - 1) / hb;
    pw = nstrips * WV + 2 * K;
    phh = nblocks * hb + 2 * K;
    pitch = (pw + 63) & ~63;
    // tile occupancy
    std::vector<int2> origins;
    for (int m = 0; m < nblocks; ++m)
      for (int n = 0; n < nstrips; ++n) origins.push_back(make_int2(n * WV, m * hb));
    int2 *origins_d = g_pool.take<int2>(origins.size());
    int *counts_d = g_pool.take<int>(origins.size());
    HIP_OK(hipMemcpyAsync(origins_d, origins.data(), origins.size() * sizeof(int2),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tile_occupancy, dim3(origins.size()), dim3(256), 0, st, cls_img, W, H,
                       ox, oy, K, WV, hb, origins_d, counts_d);
    HIP_OK(hipGetLastError());
    std::vector<int> counts(origins.size());
    HIP_OK(hipMemcpyAsync(counts.data(), counts_d, counts.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (size_t i = 0; i < origins.size(); ++i)
      if (counts[i] > 0) {
        chan_origins.push_back(origins[i]);
        n_interior += counts[i];
      }
  }

  const bool tuned = d->omega <= 0.0;
  double omega = d->omega;
  if (tuned) {
    const double mu = any_interior ? tuned_mu(cls_img, W, H, ox + PH_PAD, oy + PH_PAD, region_w,
                                              region_h, st)
                                   : 4.0;
    omega = omega_for(mu, d->lambda);
  }
  rep->omega_used = omega;
  rep->half_sweeps_used = (uint32_t)K;
  const uint32_t stall_window = stall_window_for(omega, tuned);
  // SOR factor of every half-sweep (one value for now; the table lets a
  // schedule vary it per half-sweep)
  const size_t ntab = 2 * (size_t)d->max_iterations + 2;
  std::vector<double> wtab(ntab, omega);
  double *omega_d = g_pool.take<double>(ntab);
  HIP_OK(hipMemcpyAsync(omega_d, wtab.data(), ntab * sizeof(double), hipMemcpyHostToDevice, st));

  // --- per-channel planes and r0
  std::vector<double> r0(C, 0.0);
  uint8_t *cls_plane = nullptr;
  double *res_d = nullptr;
  uint32_t *arrival_d = nullptr;
  double *rowsum_d = nullptr;
  std::vector<double *> hist_d(C, nullptr);
  dim3 pblk(256), pgrd(1, 1);
  if (any_interior) {
    const size_t plane = (size_t)pitch * phh;
    cls_plane = g_pool.take<uint8_t>(plane);
    res_d = g_pool.take<double>(2 * (size_t)C);
    arrival_d = g_pool.take<uint32_t>(C + 1);
    rowsum_d = g_pool.take<double>((size_t)phh * C);
    HIP_OK(hipMemsetAsync(arrival_d, 0, (C + 1) * sizeof(uint32_t), st));
    pgrd = dim3((pitch + 255) / 256, phh);
    for (int c = 0; c < C; ++c) {
      PhInst I{};
      I.e[0] = g_pool.take<float>(plane);
      I.e[1] = g_pool.take<float>(plane);
      I.g = g_pool.take<float>(plane);
      I.x = g_pool.take<double>(plane);
      I.bp = g_pool.take<double>(plane);
      I.cls = cls_plane;
      I.pitch = pitch;
      I.pw = pw;
      I.ph = phh;
      I.ox = ox;
      I.oy = oy;
      I.img_w = W;
      I.img_h = H;
      I.channels = C;
      I.channel = c;
      I.dg64 = 4.0 + d->lambda;
      I.dg = (float)I.dg64;
      I.omega_tab = omega_d;
      I.lambda = d->lambda;
      hipLaunchKernelGGL(k_assemble, pgrd, pblk, 0, st, I, src, tgt, cls_img, cls_plane);
      HIP_OK(hipGetLastError());
      hist_d[c] = g_pool.take<double>(std::max(d->max_iterations, 1u));
      inst_h.push_back(I);
      inst_chan.push_back(c);
    }
    for (int c = 0; c < C; ++c) {
      hipLaunchKernelGGL(k_residual, dim3(phh), dim3(256), 0, st, inst_h[c],
                         rowsum_d + (size_t)c * phh, arrival_d + c, res_d + 2 * c);
      HIP_OK(hipGetLastError());
    }
    std::vector<double> rr(2 * C);
    HIP_OK(hipMemcpyAsync(rr.data(), res_d, rr.size() * sizeof(double), hipMemcpyDeviceToHost,
                          st));
    HIP_OK(hipStreamSynchronize(st));
    for (int c = 0; c < C; ++c) r0[c] = rr[2 * c];
  }

  // --- instances that need sweeps
  std::vector<int> act;  // channel ids
  for (int c = 0; c < C; ++c) {
    res[c] = ChanResult{};
    res[c].interior = (uint64_t)n_interior;
    if (!any_interior || r0[c] == 0.0) continue;  // poisson.rs:177-189
    if (!std::isfinite(r0[c])) {
      // Non-finite inputs: the reference's residuals are NaN, no sweep ever
      // improves on the best, and the stall guard ends it after 16 sweeps.
      res[c].nonfinite = true;
      res[c].iterations = std::min(PH_STALL_WINDOW, d->max_iterations);
      res[c].stop = res[c].iterations == PH_STALL_WINDOW ? PH_STOP_STALLED : PH_STOP_MAX_ITERATIONS;
      res[c].converged = false;
      res[c].final_residual = NAN;
      continue;
    }
    res[c].active = true;
    res[c].done = false;
    act.push_back(c);
  }
  rep->unknowns = (uint64_t)n_interior * act.size();

  if (!act.empty()) {
    const int nact = (int)act.size();
    std::vector<PhInst> ai(nact);
    std::vector<PhTile> tl;
    for (int a = 0; a < nact; ++a) {
      ai[a] = inst_h[act[a]];
      ai[a].tile_begin = (int)tl.size();
      for (const int2 &o : chan_origins) tl.push_back(PhTile{a, o.x, o.y, 0});
      ai[a].tile_end = (int)tl.size();
    }
    const int ntiles = (int)tl.size();
    PhTile *tiles_d = g_pool.take<PhTile>(ntiles);
    PhInst *inst_d = g_pool.take<PhInst>(nact);
    PhCtl *ctl_d = g_pool.take<PhCtl>(nact);
    double *partials_d = g_pool.take<double>((size_t)ntiles * K);
    HIP_OK(hipMemcpyAsync(tiles_d, tl.data(), ntiles * sizeof(PhTile), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(inst_d, ai.data(), nact * sizeof(PhInst), hipMemcpyHostToDevice, st));
    std::vector<PhCtl> ctl(nact);
    for (int a = 0; a < nact; ++a) {
      PhCtl &Cc = ctl[a];
      memset(&Cc, 0, sizeof(Cc));
      const int c = act[a];
      Cc.r0 = r0[c];
      Cc.tol = d->tolerance;
      Cc.refine_abs2 = (REFINE_RATIO * r0[c]) * (REFINE_RATIO * r0[c]);
      Cc.best = INFINITY;
      Cc.history = hist_d[c];
      Cc.max_it = d->max_iterations;
      Cc.sweep_limit = d->max_iterations;
      Cc.halt = PH_RUN;
      Cc.mode = PH_MODE_NORMAL;
      Cc.stall_window = stall_window;
      Cc.refine_every = REFINE_EVERY;
      Cc.last_refine = 0;
    }
    HIP_OK(hipMemcpyAsync(ctl_d, ctl.data(), nact * sizeof(PhCtl), hipMemcpyHostToDevice, st));

    std::vector<double> rr(2);
    auto refine = [&](int a) -> double {
      PhInst &I = ai[a];
      hipLaunchKernelGGL(k_apply_correction, pgrd, pblk, 0, st, I, (int)ctl[a].cur_buf);
      HIP_OK(hipGetLastError());
      const int c = act[a];
      hipLaunchKernelGGL(k_residual, dim3(phh), dim3(256), 0, st, I, rowsum_d + (size_t)c * phh,
                         arrival_d + c, res_d + 2 * c);
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpyAsync(rr.data(), res_d + 2 * c, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            st));
      HIP_OK(hipStreamSynchronize(st));
      rep->refinements += 1;
      return rr[0];
    };
    auto put_hist = [&](int a, uint32_t m, double rel) {
      HIP_OK(hipMemcpyAsync(ctl[a].history + (m - 1), &rel, sizeof(double),
                            hipMemcpyHostToDevice, st));
      HIP_OK(hipStreamSynchronize(st));
    };
    auto finish = [&](int a, int stop, uint32_t m, double rel) {
      ChanResult &R = res[act[a]];
      R.done = true;
      R.iterations = m;
      R.stop = stop;
      R.converged = stop == PH_STOP_CONVERGED;
      R.final_residual = rel;
      ctl[a].halt = PH_HALT_DONE;
    };
    // stall rule of the reference for a host-evaluated sweep (poisson.rs:220-228)
    auto stall_step = [&](PhCtl &Cc, double rel) -> bool {
      if (rel < Cc.best * (1.0 - PH_STALL_IMPROVEMENT)) {
        Cc.best = rel;
        Cc.since = 0;
        return false;
      }
      Cc.since += 1;
      return Cc.since >= Cc.stall_window;
    };
    auto resume = [&](int a, double gnorm) {
      PhCtl &Cc = ctl[a];
      Cc.last_refine = Cc.sweeps_done;
      Cc.halt = PH_RUN;
      Cc.mode = PH_MODE_NORMAL;
      Cc.sweep_limit = Cc.max_it;
      Cc.refine_pending = 0;
      Cc.refine_abs2 = (REFINE_RATIO * gnorm) * (REFINE_RATIO * gnorm);
    };
    // an exact FP64 evaluation of sweep m (state in e[cur_buf]) and the reference's rules
    auto settle = [&](int a, uint32_t m, int device_reason) {
      PhCtl &Cc = ctl[a];
      const double gnorm = refine(a);
      const double rel = gnorm / Cc.r0;
      put_hist(a, m, rel);
      Cc.sweeps_final = m;
      Cc.sweeps_done = m;
      if (rel <= Cc.tol) return finish(a, PH_STOP_CONVERGED, m, rel);
      if (device_reason == PH_REASON_STALLED) return finish(a, PH_STOP_STALLED, m, rel);
      if (stall_step(Cc, rel)) return finish(a, PH_STOP_STALLED, m, rel);
      if (m >= Cc.max_it) return finish(a, PH_STOP_MAX_ITERATIONS, m, rel);
      resume(a, gnorm);
    };

    const int PF = prefetch_rows(K);
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
    while (remaining > 0) {
      if (timing) HIP_OK(hipEventRecord(eb0, st));
      for (int i = 0; i < batch; ++i) {
        launch_sor(K, PF, ntiles, nact, st, tiles_d, inst_d, ctl_d, partials_d, hb);
        rep->launches += 1;
      }
      if (timing) HIP_OK(hipEventRecord(eb1, st));
      HIP_OK(hipMemcpyAsync(ctl.data(), ctl_d, nact * sizeof(PhCtl), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      if (timing) {
        float bms = 0.f;
        HIP_OK(hipEventElapsedTime(&bms, eb0, eb1));
        rep->sweep_ms += bms;
      }
      bool any_halt = false;
      for (int a = 0; a < nact; ++a) {
        PhCtl &Cc = ctl[a];
        if (res[act[a]].done) continue;
        switch (Cc.halt) {
          case PH_RUN:
            break;
          case PH_HALT_STOP:
            any_halt = true;
            if (Cc.halt_sweep > Cc.halt_base) {
              Cc.mode = PH_MODE_REDO;
              Cc.sweeps_done = Cc.halt_base;
              Cc.sweep_limit = Cc.halt_sweep;
              Cc.cur_buf = Cc.halt_buf;
              Cc.halt = PH_RUN;
            } else {
              Cc.cur_buf = Cc.halt_buf;
              settle(a, Cc.halt_sweep, (int)Cc.reason);
            }
            break;
          case PH_HALT_REDO_DONE:
            any_halt = true;
            settle(a, Cc.halt_sweep, (int)Cc.reason);
            break;
          case PH_HALT_LIMIT:
            any_halt = true;
            settle(a, Cc.sweeps_done, -1);
            break;
          case PH_HALT_REFINE:
            any_halt = true;
            settle(a, Cc.sweeps_done, -1);
            break;
          default:
            throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "unexpected halt state %u", Cc.halt);
        }
        if (res[act[a]].done) --remaining;
      }
      HIP_OK(hipMemcpyAsync(ctl_d, ctl.data(), nact * sizeof(PhCtl), hipMemcpyHostToDevice, st));
      // next batch: up to the next scheduled halt (refinement or sweep limit) of
      // the running instances, so no launch finds every instance halted
      {
        uint32_t need = 0;
        for (int a = 0; a < nact; ++a) {
          const PhCtl &Cc = ctl[a];
          if (res[act[a]].done || Cc.halt != PH_RUN) continue;
          uint32_t next = std::min(Cc.sweep_limit, Cc.last_refine + Cc.refine_every);
          if (Cc.mode == PH_MODE_REDO) next = Cc.sweep_limit;
          const uint32_t todo = next > Cc.sweeps_done ? next - Cc.sweeps_done : 1;
          need = std::max(need, (todo + (uint32_t)(K / 2) - 1) / (uint32_t)(K / 2));
        }
        batch = (int)std::max<uint32_t>(1, std::min<uint32_t>(need, 64));
      }
      // every batch must move some instance forward (sweeps or halt state)
      uint64_t progress = 0;
      for (int a = 0; a < nact; ++a)
        progress += (uint64_t)ctl[a].sweeps_done * 8u + ctl[a].halt + ctl[a].launches * 131u;
      if (progress == last_progress && ++stuck >= 3)
        throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "sweep control made no progress");
      if (progress != last_progress) stuck = 0;
      last_progress = progress;
      guard += 1;
      if (guard > 4.0 * MAX_ITERATIONS_LIMIT)
        throw fail(PH_CLASS_EXECUTION, "E_HIP_CONTROL", "solve loop did not terminate");
    }
    if (timing) {
      hipEventDestroy(eb0);
      hipEventDestroy(eb1);
    }
    for (int a = 0; a < nact; ++a) {
      hipLaunchKernelGGL(k_scatter, pgrd, pblk, 0, st, ai[a], out);
      HIP_OK(hipGetLastError());
      rep->unknown_sweeps += (uint64_t)n_interior * res[act[a]].iterations;
    }
  }
  // non-finite channels: interior samples become NaN (the reference's
  // residuals and field are NaN there)
  for (int c = 0; c < C; ++c)
    if (res[c].nonfinite) {
      PhInst I = inst_h[c];
      I.x = I.bp;  // reuse: fill with NaN then scatter
      std::vector<double> nanv((size_t)pitch * phh, NAN);
      HIP_OK(hipMemcpyAsync(I.x, nanv.data(), nanv.size() * sizeof(double), hipMemcpyHostToDevice,
                            st));
      hipLaunchKernelGGL(k_scatter, pgrd, pblk, 0, st, I, out);
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(st));
    }

  // --- worst channel (poisson_blend.rs:240-246: more iterations, then larger
  // final residual; ties keep the earlier channel)
  int worst = 0;
  for (int c = 1; c < C; ++c) {
    const ChanResult &a = res[worst], &b = res[c];
    const bool worse = b.iterations != a.iterations ? b.iterations > a.iterations
                                                    : b.final_residual > a.final_residual;
    if (worse) worst = c;
  }
  const ChanResult &Wr = res[worst];
  rep->worst_channel = (uint32_t)worst;
  rep->iterations = Wr.iterations;
  rep->stop_reason = Wr.stop;
  rep->converged = Wr.converged ? 1 : 0;
  rep->tolerance = d->tolerance;
  rep->final_residual = Wr.final_residual;
  rep->history_len = Wr.iterations;
  if (rep->residual_history && rep->history_capacity >= Wr.iterations && Wr.iterations > 0) {
    if (Wr.nonfinite) {
      for (uint32_t i = 0; i < Wr.iterations; ++i) rep->residual_history[i] = NAN;
    } else {
      HIP_OK(hipMemcpyAsync(rep->residual_history, hist_d[worst], Wr.iterations * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    }
  } else if (Wr.iterations > 0 && rep->residual_history) {
    throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "history capacity %u < %u iterations",
               rep->history_capacity, Wr.iterations);
  }
  if (timing) {
    HIP_OK(hipEventRecord(ev1, st));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    rep->device_ms = ms;
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
  } else {
    rep->device_ms = 0.0;
  }
  HIP_OK(hipStreamSynchronize(st));
}

}  // namespace ph

// ---------------------------------------------------------------------------
// C ABI

using namespace ph;

extern "C" int32_t ph_abi_version(void) { return PH_ABI_VERSION; }

extern "C" void ph_force_unavailable(int32_t forced) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_forced = forced ? 1 : 0;
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
                             const std::function<void()> &body) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    if (!rep) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null report");
    validate(desc);
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
