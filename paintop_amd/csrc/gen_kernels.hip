// gen_kernels.hip -- device-side input generators, so large synthetic inputs
// never cross PCIe:
//   field.fbm@1     crates/paintop-cpu/src/noise.rs:72-160, 218-240, 660-694
//   mask.polygon@1  crates/paintop-cpu/src/mask_polygon.rs:155-195 (inside test),
//                   sampled once at each pixel centre (binary coverage)
//   ellipse mask    binary coverage of an axis-aligned ellipse at pixel centres
// The fbm value is bit-identical to the reference formula: integer hashing is
// exact, and every f64 step keeps the reference's operation order with
// explicit fma where the reference uses mul_add (built with -ffp-contract=off).
#include <math.h>
#include <stdio.h>

#include <cmath>

#include "../../include/paintop_hip.h"
#include "ph_internal.h"

namespace ph {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // noise.rs:72-76
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double lattice_value(int64_t gx, int64_t gy, uint64_t seed) {
  const uint64_t hx = mix64(seed ^ ((uint64_t)gx * 0x9e3779b97f4a7c15ULL));
  const uint64_t h = mix64(hx ^ ((uint64_t)gy * 0xc2b2ae3d27d4eb4fULL));
  const double unit = (double)(uint32_t)(h >> 40) / 16777216.0;
  return fma(unit, 2.0, -1.0);
}

__device__ __forceinline__ double fade(double t) {
  return t * t * t * fma(t, fma(t, 6.0, -15.0), 10.0);
}

__device__ __forceinline__ double lerp(double a, double b, double t) { return fma(b - a, t, a); }

__device__ double value_noise(double x, double y, uint64_t seed) {
  const double x0 = floor(x), y0 = floor(y);
  const int64_t gx0 = (int64_t)x0, gy0 = (int64_t)y0;
  const double tx = fade(x - x0), ty = fade(y - y0);
  const double v00 = lattice_value(gx0, gy0, seed);
  const double v10 = lattice_value(gx0 + 1, gy0, seed);
  const double v01 = lattice_value(gx0, gy0 + 1, seed);
  const double v11 = lattice_value(gx0 + 1, gy0 + 1, seed);
  return lerp(lerp(v00, v10, tx), lerp(v01, v11, tx), ty);
}

__global__ void k_fbm(float *out, int w, int h, int stride, int offset, double scale,
                      uint64_t seed, int octaves, double lacunarity, double gain, float a,
                      float b) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w) return;
  const double px = (double)x + 0.5, py = (double)y + 0.5;
  const double fx = px * scale, fy = py * scale;
  double frequency = 1.0, amplitude = 1.0, sum = 0.0, total = 0.0;
  for (int o = 0; o < octaves; ++o) {
    const uint64_t os = mix64(seed ^ ((uint64_t)o + 1u));
    sum = fma(amplitude, value_noise(fx * frequency, fy * frequency, os), sum);
    total += amplitude;
    frequency *= lacunarity;
    amplitude *= gain;
  }
  const double v = total > 0.0 ? sum / total : 0.0;
  out[((size_t)y * w + x) * stride + offset] = fmaf((float)v, a, b);
}

#define PH_MAX_POLY 64
struct Poly {
  double x[PH_MAX_POLY], y[PH_MAX_POLY];
  int n;
  int even_odd;
};

// mask_polygon.rs:161-195: half-open ray cast to +x.
__device__ bool poly_contains(const Poly &P, double x, double y) {
  if (P.n < 3) return false;
  int winding = 0;
  bool parity = false;
  for (int i = 0; i < P.n; ++i) {
    const double x0 = P.x[i], y0 = P.y[i];
    const int k = i + 1 == P.n ? 0 : i + 1;
    const double x1 = P.x[k], y1 = P.y[k];
    const bool below0 = y0 <= y, below1 = y1 <= y;
    if (below0 == below1) continue;
    const double t = (y - y0) / (y1 - y0);
    const double cross_x = fma(t, x1 - x0, x0);
    if (cross_x > x) {
      parity = !parity;
      winding += below0 ? 1 : -1;
    }
  }
  return P.even_odd ? parity : winding != 0;
}

__global__ void k_polygon(float *out, int w, int h, Poly P) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w) return;
  out[(size_t)y * w + x] = poly_contains(P, (double)x + 0.5, (double)y + 0.5) ? 1.0f : 0.0f;
}

__global__ void k_ellipse(float *out, int w, int h, double cx, double cy, double rx, double ry) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w) return;
  const double dx = ((double)x + 0.5 - cx) / rx, dy = ((double)y + 0.5 - cy) / ry;
  out[(size_t)y * w + x] = (dx * dx + dy * dy <= 1.0) ? 1.0f : 0.0f;
}

// out[i] += sigma * N(0, 1): Box-Muller on two SplitMix64 draws keyed by (seed, i).
__global__ void k_gauss_noise(float *out, size_t n, uint64_t seed, double sigma) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * (uint64_t)i + 1));
  const uint64_t b = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * (uint64_t)i + 2));
  const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  out[i] = out[i] + (float)(sigma * z);
}

}  // namespace ph

using namespace ph;

static void gen_err(ph_error *err, int cls, const char *code, const char *msg) {
  if (!err) return;
  err->error_class = cls;
  snprintf(err->code, sizeof(err->code), "%s", code);
  snprintf(err->message, sizeof(err->message), "%s", msg);
}

static ph_status gen_finish(hipError_t e, ph_error *err) {
  if (e != hipSuccess) {
    gen_err(err, PH_CLASS_EXECUTION, "E_HIP_RUNTIME", hipGetErrorString(e));
    return PH_ERROR;
  }
  return PH_OK;
}

extern "C" ph_status ph_fbm_device(float *d_out, uint32_t width, uint32_t height,
                                   uint32_t channel_stride, uint32_t channel_offset,
                                   double frequency, uint64_t seed, uint32_t octaves,
                                   double lacunarity, double gain, float scale, float bias,
                                   void *stream, ph_error *err) {
  if (!d_out || channel_stride == 0 || channel_offset >= channel_stride || octaves == 0 ||
      octaves > 32 || !(frequency > 0.0) || !(lacunarity > 0.0) || !(gain > 0.0 && gain <= 1.0)) {
    gen_err(err, PH_CLASS_SCHEMA, "E_NOISE_PARAM", "field.fbm: invalid parameters");
    return PH_ERROR;
  }
  if (width == 0 || height == 0) return PH_OK;
  hipLaunchKernelGGL(k_fbm, dim3((width + 255) / 256, height), dim3(256), 0, (hipStream_t)stream,
                     d_out, (int)width, (int)height, (int)channel_stride, (int)channel_offset,
                     frequency / 64.0, seed, (int)octaves, lacunarity, gain, scale, bias);
  return gen_finish(hipGetLastError(), err);
}

extern "C" ph_status ph_gauss_noise_device(float *d_out, uint64_t n, uint64_t seed, double sigma,
                                           void *stream, ph_error *err) {
  if (!d_out || !std::isfinite(sigma)) {
    gen_err(err, PH_CLASS_SCHEMA, "E_NOISE_PARAM", "gauss noise: invalid parameters");
    return PH_ERROR;
  }
  if (n == 0) return PH_OK;
  hipLaunchKernelGGL(k_gauss_noise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_out, (size_t)n, seed, sigma);
  return gen_finish(hipGetLastError(), err);
}

extern "C" ph_status ph_polygon_mask_device(float *d_out, uint32_t width, uint32_t height,
                                            const double *points_xy, uint32_t n_points,
                                            int32_t even_odd, void *stream, ph_error *err) {
  if (!d_out || n_points > PH_MAX_POLY || (n_points > 0 && !points_xy)) {
    gen_err(err, PH_CLASS_SCHEMA, "E_POLYGON_PARAM", "mask.polygon: at most 64 vertices");
    return PH_ERROR;
  }
  Poly P{};
  P.n = (int)n_points;
  P.even_odd = even_odd ? 1 : 0;
  for (uint32_t i = 0; i < n_points; ++i) {
    P.x[i] = points_xy[2 * i];
    P.y[i] = points_xy[2 * i + 1];
    if (!std::isfinite(P.x[i]) || !std::isfinite(P.y[i])) {
      gen_err(err, PH_CLASS_SCHEMA, "E_POLYGON_PARAM", "mask.polygon: non-finite vertex");
      return PH_ERROR;
    }
  }
  if (width == 0 || height == 0) return PH_OK;
  hipLaunchKernelGGL(k_polygon, dim3((width + 255) / 256, height), dim3(256), 0,
                     (hipStream_t)stream, d_out, (int)width, (int)height, P);
  return gen_finish(hipGetLastError(), err);
}

extern "C" ph_status ph_ellipse_mask_device(float *d_out, uint32_t width, uint32_t height,
                                            double cx, double cy, double rx, double ry,
                                            void *stream, ph_error *err) {
  if (!d_out || !(rx > 0.0) || !(ry > 0.0)) {
    gen_err(err, PH_CLASS_SCHEMA, "E_ELLIPSE_PARAM", "ellipse: radii must be positive");
    return PH_ERROR;
  }
  if (width == 0 || height == 0) return PH_OK;
  hipLaunchKernelGGL(k_ellipse, dim3((width + 255) / 256, height), dim3(256), 0,
                     (hipStream_t)stream, d_out, (int)width, (int)height, cx, cy, rx, ry);
  return gen_finish(hipGetLastError(), err);
}
