// assert_kernels.hip -- assert.no_change_outside_mask@1's scan on the device,
// so the outside-mask check after a solve runs without copying the images to
// the host (crates/paintop-cpu/src/assert.rs:405-446).
//
// A pixel is outside when its mask coverage (as f64) is at most epsilon; its
// delta is the largest channel |to_space(after) - to_space(before)| (f32
// arithmetic, NaN never counts); it leaks when the delta strictly exceeds the
// threshold. The scan reports the leak count, the largest outside delta, the
// first pixel that reaches it (the reference's running strict maximum keeps
// the earliest) and the first 256 leaking pixels in row-major order
// (assert.rs:72). Everything is integer or a maximum, so the result does not
// depend on the order of the reductions.
//
// Three kernels, all HBM-bound reads of before, after and mask:
//   k_os_scan   per block of OS_PIX pixels: leak count, max delta, first argmax;
//   k_os_total  one block: exclusive prefix of the block counts (block
//               offsets into the location list), totals, global max / argmax;
//   k_os_locs   the blocks whose offset is below 256 re-scan their pixels and
//               write their leaks in row-major order (a block-wide prefix over
//               each 256-pixel step).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "../../include/paintop_hip.h"
#include "host_common.h"

namespace ph {

constexpr int OS_THREADS = 256;
constexpr int OS_STEPS = 16;
constexpr int OS_PIX = OS_THREADS * OS_STEPS;  // pixels per scan block
constexpr int OS_MAX_LOC = 256;                // assert.rs:72

// sRGB decode (assert.rs:74-85): the power is taken in double and rounded
// once, which is within one f32 ulp of the host's powf.
__device__ __forceinline__ float os_decode(float c) {
  if (c <= 0.04045f) return c / 12.92f;
  return (float)pow((double)((c + 0.055f) / 1.055f), 2.4);
}

struct OsArgs {
  const float *before, *after, *mask;
  uint64_t n;    // pixels
  int channels;
  int decode;
  double threshold, epsilon;
};

// delta of pixel i, or -1 when it is inside
__device__ __forceinline__ double os_delta(const OsArgs &A, uint64_t i) {
  const double coverage = (double)A.mask[i];
  if (coverage > A.epsilon) return -1.0;
  double pd = 0.0;
  const size_t base = (size_t)i * A.channels;
  for (int c = 0; c < A.channels; ++c) {
    float a = A.after[base + c], b = A.before[base + c];
    if (A.decode) {
      a = os_decode(a);
      b = os_decode(b);
    }
    const double d = (double)fabsf(a - b);
    if (d > pd) pd = d;
  }
  return pd;
}

// (max, first index) of two candidates
__device__ __forceinline__ void os_better(double &m, int64_t &arg, double m2, int64_t arg2) {
  if (arg2 < 0) return;
  if (arg < 0 || m2 > m || (m2 == m && arg2 < arg)) {
    m = m2;
    arg = arg2;
  }
}

__global__ __launch_bounds__(OS_THREADS) void k_os_scan(OsArgs A, uint64_t *__restrict__ cnt,
                                                        double *__restrict__ mx,
                                                        int64_t *__restrict__ ag) {
  const uint64_t b = blockIdx.x;
  uint64_t count = 0;
  double m = 0.0;
  int64_t arg = -1;  // first outside pixel with the largest delta
  for (int k = 0; k < OS_STEPS; ++k) {
    const uint64_t i = b * OS_PIX + (uint64_t)k * OS_THREADS + threadIdx.x;
    if (i >= A.n) break;
    const double d = os_delta(A, i);
    if (d < 0.0) continue;
    if (d > A.threshold) count += 1;
    if (arg < 0 || d > m) {
      m = d;
      arg = (int64_t)i;
    }
  }
  __shared__ uint64_t s_cnt[OS_THREADS];
  __shared__ double s_m[OS_THREADS];
  __shared__ int64_t s_arg[OS_THREADS];
  s_cnt[threadIdx.x] = count;
  s_m[threadIdx.x] = m;
  s_arg[threadIdx.x] = arg;
  __syncthreads();
  for (int h = OS_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
      double mm = s_m[threadIdx.x];
      int64_t aa = s_arg[threadIdx.x];
      os_better(mm, aa, s_m[threadIdx.x + h], s_arg[threadIdx.x + h]);
      s_m[threadIdx.x] = mm;
      s_arg[threadIdx.x] = aa;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[b] = s_cnt[0];
    mx[b] = s_m[0];
    ag[b] = s_arg[0];
  }
}

// The same scan over groups of 4 consecutive pixels (C <= 4 channels, 16-byte
// aligned buffers, n a multiple of 4): one 16-byte load of the mask, then the
// group's 4*C samples of before and after as C 16-byte loads each, skipped
// when all four pixels are inside. Thread t of block b takes groups
// b*OS_PIX/4 + s*OS_THREADS + t, s < OS_STEPS/4 (the block's pixel range is
// the generic kernel's).
template <int C>
__global__ __launch_bounds__(OS_THREADS) void k_os_scan_v(OsArgs A, uint64_t *__restrict__ cnt,
                                                          double *__restrict__ mx,
                                                          int64_t *__restrict__ ag) {
  const uint64_t b = blockIdx.x;
  uint64_t count = 0;
  double m = 0.0;
  int64_t arg = -1;
  const float4 *mk4 = reinterpret_cast<const float4 *>(A.mask);
  const float4 *bf4 = reinterpret_cast<const float4 *>(A.before);
  const float4 *af4 = reinterpret_cast<const float4 *>(A.after);
  for (int s = 0; s < OS_STEPS / 4; ++s) {
    const uint64_t g = b * (OS_PIX / 4) + (uint64_t)s * OS_THREADS + threadIdx.x;
    if (4 * g >= A.n) break;
    const float4 mv = mk4[g];
    const float cov[4] = {mv.x, mv.y, mv.z, mv.w};
    bool out[4], any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      out[j] = !((double)cov[j] > A.epsilon);
      any |= out[j];
    }
    if (!any) continue;
    float bv[4 * C], av[4 * C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const float4 x = bf4[g * C + q], y = af4[g * C + q];
      bv[4 * q] = x.x; bv[4 * q + 1] = x.y; bv[4 * q + 2] = x.z; bv[4 * q + 3] = x.w;
      av[4 * q] = y.x; av[4 * q + 1] = y.y; av[4 * q + 2] = y.z; av[4 * q + 3] = y.w;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!out[j]) continue;
      double pd = 0.0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        float x = av[j * C + c], y = bv[j * C + c];
        if (A.decode) {
          x = os_decode(x);
          y = os_decode(y);
        }
        const double d = (double)fabsf(x - y);
        if (d > pd) pd = d;
      }
      if (pd > A.threshold) count += 1;
      if (arg < 0 || pd > m) {
        m = pd;
        arg = (int64_t)(4 * g + j);
      }
    }
  }
  __shared__ uint64_t s_cnt[OS_THREADS];
  __shared__ double s_m[OS_THREADS];
  __shared__ int64_t s_arg[OS_THREADS];
  s_cnt[threadIdx.x] = count;
  s_m[threadIdx.x] = m;
  s_arg[threadIdx.x] = arg;
  __syncthreads();
  for (int h = OS_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + h];
      double mm = s_m[threadIdx.x];
      int64_t aa = s_arg[threadIdx.x];
      os_better(mm, aa, s_m[threadIdx.x + h], s_arg[threadIdx.x + h]);
      s_m[threadIdx.x] = mm;
      s_arg[threadIdx.x] = aa;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[b] = s_cnt[0];
    mx[b] = s_m[0];
    ag[b] = s_arg[0];
  }
}

// res: [0] total leaks (u64), [1] max (f64 bits), [2] first argmax (i64)
__global__ __launch_bounds__(OS_THREADS) void k_os_total(const uint64_t *__restrict__ cnt,
                                                         const double *__restrict__ mx,
                                                         const int64_t *__restrict__ ag,
                                                         int nblocks, uint64_t *__restrict__ off,
                                                         uint64_t *__restrict__ res) {
  const int per = (nblocks + OS_THREADS - 1) / OS_THREADS;
  const int b0 = threadIdx.x * per, b1 = min(nblocks, b0 + per);
  uint64_t s = 0;
  double m = 0.0;
  int64_t arg = -1;
  for (int b = b0; b < b1; ++b) {
    s += cnt[b];
    os_better(m, arg, mx[b], ag[b]);
  }
  __shared__ uint64_t s_sum[OS_THREADS];
  __shared__ double s_m[OS_THREADS];
  __shared__ int64_t s_arg[OS_THREADS];
  s_sum[threadIdx.x] = s;
  s_m[threadIdx.x] = m;
  s_arg[threadIdx.x] = arg;
  __syncthreads();
  // inclusive scan of the thread sums (Hillis-Steele, integer: exact)
  for (int h = 1; h < OS_THREADS; h <<= 1) {
    const uint64_t v = (int)threadIdx.x >= h ? s_sum[threadIdx.x - h] : 0;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t run = s_sum[threadIdx.x] - s;  // exclusive prefix of this thread's run
  for (int b = b0; b < b1; ++b) {
    off[b] = run;
    run += cnt[b];
  }
  for (int h = OS_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double mm = s_m[threadIdx.x];
      int64_t aa = s_arg[threadIdx.x];
      os_better(mm, aa, s_m[threadIdx.x + h], s_arg[threadIdx.x + h]);
      s_m[threadIdx.x] = mm;
      s_arg[threadIdx.x] = aa;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[0] = s_sum[OS_THREADS - 1];
    res[1] = __double_as_longlong(s_m[0]);
    res[2] = (uint64_t)s_arg[0];
  }
}

__global__ __launch_bounds__(OS_THREADS) void k_os_locs(OsArgs A,
                                                        const uint64_t *__restrict__ cnt,
                                                        const uint64_t *__restrict__ off,
                                                        int64_t *__restrict__ loc) {
  const uint64_t b = blockIdx.x;
  if (cnt[b] == 0 || off[b] >= (uint64_t)OS_MAX_LOC) return;  // uniform over the block
  __shared__ uint32_t s_f[OS_THREADS];
  uint64_t run = off[b];
  for (int k = 0; k < OS_STEPS && run < (uint64_t)OS_MAX_LOC; ++k) {
    const uint64_t i = b * OS_PIX + (uint64_t)k * OS_THREADS + threadIdx.x;
    uint32_t leak = 0;
    if (i < A.n) {
      const double d = os_delta(A, i);
      leak = d >= 0.0 && d > A.threshold ? 1u : 0u;
    }
    s_f[threadIdx.x] = leak;
    __syncthreads();
    for (int h = 1; h < OS_THREADS; h <<= 1) {
      const uint32_t v = (int)threadIdx.x >= h ? s_f[threadIdx.x - h] : 0u;
      __syncthreads();
      s_f[threadIdx.x] += v;
      __syncthreads();
    }
    const uint64_t pos = run + s_f[threadIdx.x] - leak;  // exclusive rank in row-major order
    if (leak && pos < (uint64_t)OS_MAX_LOC) loc[pos] = (int64_t)i;
    run += s_f[OS_THREADS - 1];
    __syncthreads();
  }
}

// Grow-only device workspace of the scan.
static char *g_os_buf = nullptr;
static size_t g_os_cap = 0;

}  // namespace ph

using namespace ph;

extern "C" ph_status ph_assert_no_change_outside_mask_device(
    const float *d_before, const float *d_after, const float *d_mask, uint32_t width,
    uint32_t height, uint32_t channels, int32_t decode_srgb, double outside_threshold,
    double coverage_epsilon, void *stream, ph_outside_scan *out, ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    if (!out) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null result");
    if (!std::isfinite(outside_threshold) || outside_threshold < 0.0)
      throw fail(PH_CLASS_SCHEMA, "E_ASSERT_PARAM",
                 "assert.no_change_outside_mask `outside_threshold` must be a finite "
                 "non-negative number");
    if (!std::isfinite(coverage_epsilon) || coverage_epsilon < 0.0)
      throw fail(PH_CLASS_SCHEMA, "E_ASSERT_PARAM",
                 "assert.no_change_outside_mask `coverage_epsilon` must be a finite "
                 "non-negative number");
    if ((size_t)width * height > ((size_t)1 << 31) || channels > 64)
      throw fail(PH_CLASS_POLICY, "E_HIP_LIMIT", "scan of %ux%ux%u exceeds the limits", width,
                 height, channels);
    memset(out, 0, sizeof(*out));
    out->worst_x = out->worst_y = -1;
    const uint64_t n = (uint64_t)width * height;
    if (channels == 0 || n == 0) return PH_OK;  // nothing to compare (assert.rs:407-409)
    if (!d_before || !d_after || !d_mask)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    ensure_device();
    hipStream_t st = (hipStream_t)stream;
    const int nblocks = (int)((n + OS_PIX - 1) / OS_PIX);
    auto r256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t need = 4 * r256((size_t)nblocks * sizeof(uint64_t)) +
                        r256(4 * sizeof(uint64_t)) + r256(OS_MAX_LOC * sizeof(int64_t));
    if (need > g_os_cap) {
      if (g_os_buf) HIP_OK(hipFree(g_os_buf));
      g_os_buf = nullptr;
      g_os_cap = 0;
      HIP_OK(hipMalloc(&g_os_buf, need));
      g_os_cap = need;
    }
    char *p = g_os_buf;
    auto take = [&](size_t bytes) {
      char *q = p;
      p += (bytes + 255) & ~(size_t)255;
      return q;
    };
    uint64_t *cnt = (uint64_t *)take(nblocks * sizeof(uint64_t));
    double *mx = (double *)take(nblocks * sizeof(double));
    int64_t *ag = (int64_t *)take(nblocks * sizeof(int64_t));
    uint64_t *off = (uint64_t *)take(nblocks * sizeof(uint64_t));
    uint64_t *res = (uint64_t *)take(4 * sizeof(uint64_t));
    int64_t *loc = (int64_t *)take(OS_MAX_LOC * sizeof(int64_t));
    if ((size_t)(p - g_os_buf) > g_os_cap)
      throw fail(PH_CLASS_EXECUTION, "E_HIP_WORKSPACE", "scan workspace overrun");
    OsArgs A{d_before, d_after, d_mask, n, (int)channels, decode_srgb ? 1 : 0,
             outside_threshold, coverage_epsilon};
    const bool aligned = ((uintptr_t)d_before % 16 == 0) && ((uintptr_t)d_after % 16 == 0) &&
                         ((uintptr_t)d_mask % 16 == 0) && (n % 4 == 0);
    if (aligned && channels >= 1 && channels <= 4) {
      switch (channels) {
        case 1: hipLaunchKernelGGL(k_os_scan_v<1>, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, mx, ag); break;
        case 2: hipLaunchKernelGGL(k_os_scan_v<2>, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, mx, ag); break;
        case 3: hipLaunchKernelGGL(k_os_scan_v<3>, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, mx, ag); break;
        default: hipLaunchKernelGGL(k_os_scan_v<4>, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, mx, ag); break;
      }
    } else {
      hipLaunchKernelGGL(k_os_scan, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, mx, ag);
    }
    hipLaunchKernelGGL(k_os_total, dim3(1), dim3(OS_THREADS), 0, st, cnt, mx, ag, nblocks, off,
                       res);
    hipLaunchKernelGGL(k_os_locs, dim3(nblocks), dim3(OS_THREADS), 0, st, A, cnt, off, loc);
    HIP_OK(hipGetLastError());
    uint64_t hres[3];
    int64_t hloc[OS_MAX_LOC];
    HIP_OK(hipMemcpyAsync(hres, res, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hloc, loc, OS_MAX_LOC * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double maxv;
    memcpy(&maxv, &hres[1], sizeof(double));
    const int64_t arg = (int64_t)hres[2];
    out->changed_outside = hres[0];
    out->max_outside = arg >= 0 ? maxv : 0.0;
    if (hres[0] > 0 && arg >= 0) {  // the first pixel with the largest delta leaks
      out->worst_x = arg % width;
      out->worst_y = arg / width;
      out->worst_delta = maxv;
    }
    out->n_locations = (uint32_t)std::min<uint64_t>(hres[0], OS_MAX_LOC);
    for (uint32_t k = 0; k < out->n_locations; ++k) {
      out->locations[k][0] = hloc[k] % width;
      out->locations[k][1] = hloc[k] / width;
    }
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return PH_ERROR;
  }
}
