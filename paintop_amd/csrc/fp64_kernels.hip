// fp64_kernels.hip -- classification, system assembly, the FP64 residual of the
// refinement loop, and the write-back of the candidate.
//
// All arithmetic that fixes what the solve converges to runs in FP64 with the
// reference's operation order:
//   classify      poisson.rs:340-355  (mask > 0.5 and not on the image border)
//   b             poisson.rs:364-373  (fma(4, -c, ((l + r) + u) + d) of the source)
//   b'            poisson.rs:319      (fma(lambda, -anchor, b))
//   r             poisson.rs:318-320  (fma(lambda + 4, -u, nsum) - b')
// The squared-residual sums are reduced in a fixed order (per-row sums, then a
// fixed-order sum of rows), so reruns are bit-identical.
#include <math.h>

#include "ph_internal.h"

namespace ph {

constexpr float MASK_THRESHOLD = 0.5f;  // poisson.rs:329

// Classify every image pixel (one block per 256 pixels of a row). Each block
// folds its interior columns into the row's [min, max] (few atomics per
// row); k_bbox then reduces the rows.
__global__ __launch_bounds__(256) void k_classify(const float *__restrict__ mask, int w, int h,
                                                  uint8_t *__restrict__ cls,
                                                  int *__restrict__ rowx) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  uint8_t in = 0;
  if (x < w) {
    const size_t idx = (size_t)y * w + x;
    if (w >= 3 && h >= 3 && x >= 1 && x < w - 1 && y >= 1 && y < h - 1)
      in = mask[idx] > MASK_THRESHOLD ? 1 : 0;
    cls[idx] = in;
  }
  int xmin = in ? x : 0x7fffffff, xmax = in ? x : -1;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    xmin = min(xmin, __shfl_xor(xmin, off, 64));
    xmax = max(xmax, __shfl_xor(xmax, off, 64));
  }
  __shared__ int smin[4], smax[4];
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = xmin;
    smax[threadIdx.x >> 6] = xmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    xmin = min(min(smin[0], smin[1]), min(smin[2], smin[3]));
    xmax = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    if (xmax >= 0) {
      atomicMin(&rowx[2 * y], xmin);
      atomicMax(&rowx[2 * y + 1], xmax);
    }
  }
}

// Bounding box of the interior cells from the per-row [min, max] (one block).
__global__ __launch_bounds__(256) void k_bbox(const int *__restrict__ rowx, int h,
                                              int *__restrict__ bbox) {
  int xmin = 0x7fffffff, xmax = -1, ymin = 0x7fffffff, ymax = -1;
  for (int y = threadIdx.x; y < h; y += blockDim.x)
    if (rowx[2 * y + 1] >= 0) {
      xmin = min(xmin, rowx[2 * y]);
      xmax = max(xmax, rowx[2 * y + 1]);
      ymin = min(ymin, y);
      ymax = max(ymax, y);
    }
  __shared__ int s[4][256];
  s[0][threadIdx.x] = xmin;
  s[1][threadIdx.x] = ymin;
  s[2][threadIdx.x] = xmax;
  s[3][threadIdx.x] = ymax;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      s[0][threadIdx.x] = min(s[0][threadIdx.x], s[0][threadIdx.x + k]);
      s[1][threadIdx.x] = min(s[1][threadIdx.x], s[1][threadIdx.x + k]);
      s[2][threadIdx.x] = max(s[2][threadIdx.x], s[2][threadIdx.x + k]);
      s[3][threadIdx.x] = max(s[3][threadIdx.x], s[3][threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) bbox[threadIdx.x] = s[threadIdx.x][0];
}

// Per-row [min, max] start empty.
__global__ void k_rowx_init(int *__restrict__ rowx, int h) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y < h) {
    rowx[2 * y] = 0x7fffffff;
    rowx[2 * y + 1] = -1;
  }
}

// Interior-cell count of each tile's valid core, for every tile origin of
// every image of a call in one launch (one block per origin; origin_item
// gives its image). Images of a batch have few tiles each, so one launch per
// image left the GPU idle between launches.
__global__ void k_tile_occupancy_all(const PhItemGeo *__restrict__ items,
                                     const int *__restrict__ origin_item,
                                     const int2 *__restrict__ origins, int K, int wv, int hb,
                                     int *__restrict__ counts) {
  const PhItemGeo G = items[origin_item[blockIdx.x]];
  const int2 o = origins[blockIdx.x];
  int n = 0;
  for (int r = 0; r < hb; ++r) {
    const int iy = G.oy + o.y + K + r;
    if (iy < 0 || iy >= G.h) continue;
    for (int cidx = threadIdx.x; cidx < wv; cidx += blockDim.x) {
      const int ix = G.ox + o.x + K + cidx;
      if (ix >= 0 && ix < G.w) n += G.cls[(size_t)iy * G.w + ix];
    }
  }
  __shared__ int red[256];
  red[threadIdx.x] = n;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// Interior-cell count of each f x f block of an image's region (the coarse
// mask the SOR-factor tuner works on), for every image of a call in one
// launch (blockIdx.y = image; images without a coarse grid have cw * ch = 0).
__global__ void k_coarse_count_all(const PhItemGeo *__restrict__ items, int *__restrict__ counts) {
  const PhItemGeo G = items[blockIdx.y];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.cw * G.ch) return;
  const int cx = c % G.cw, cy = c / G.cw;
  int n = 0;
  for (int dy = 0; dy < G.cf; ++dy) {
    const int iy = G.ry0 + cy * G.cf + dy;
    if (iy < 0 || iy >= G.h) continue;
    for (int dx = 0; dx < G.cf; ++dx) {
      const int ix = G.rx0 + cx * G.cf + dx;
      if (ix >= 0 && ix < G.w) n += G.cls[(size_t)iy * G.w + ix];
    }
  }
  counts[G.coarse_off + c] = n;
}

// Fill an instance's planes from the interleaved image channel.
__global__ void k_assemble(PhInst I, const float *__restrict__ src, const float *__restrict__ tgt,
                           const uint8_t *__restrict__ cls_img, uint8_t *__restrict__ cls) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y;
  if (px >= I.pitch) return;
  const size_t p = (size_t)py * I.pitch + px;
  const int ix = I.ox + px, iy = I.oy + py;
  const int W = I.img_w, H = I.img_h, C = I.channels, ch = I.channel;
  double a = 0.0, bp = 0.0;
  uint8_t in = 0;
  if (px < I.pw && ix >= 0 && ix < W && iy >= 0 && iy < H) {
    const size_t q = (size_t)iy * W + ix;
    a = (double)tgt[q * C + ch];
    in = cls_img[q];
    if (in) {
      // interior cells are never on the image border: four in-image neighbours
      const double cc = (double)src[q * C + ch];
      const double l = (double)src[(q - 1) * C + ch];
      const double r = (double)src[(q + 1) * C + ch];
      const double u = (double)src[(q - W) * C + ch];
      const double d = (double)src[(q + W) * C + ch];
      const double b = fma(4.0, -cc, ((l + r) + u) + d);
      bp = fma(I.lambda, -a, b);
    }
  }
  I.x[p] = a;
  I.bp[p] = bp;
  cls[p] = in;
  I.e[0][p] = 0.0f;
  I.e[1][p] = 0.0f;
  I.g[p] = __builtin_nanf("");
}

// g = -r(x) in FP64, stored FP32 (NaN at non-interior cells), and the per-row
// sums of r^2 in a fixed order; k_residual_norm adds the rows in order and
// stores ||r|| in out[0] and ||r||^2 in out[1]. (A last-block reduction inside
// this kernel would cost one device-scope atomic per row block, which
// serialise across the XCDs and took longer than the rows themselves.)
// One block per plane row. Each thread takes cell pairs (px, px + 1), px even,
// so the loads of a pair are independent 16-byte vectors; rows 0 and ph - 1
// and columns 0 and pw - 1 are padding (no interior cells), so every
// neighbour read stays inside the plane. Non-interior cells of a pair get
// g = NaN, the value they hold already.
__global__ __launch_bounds__(256) void k_residual(PhInst I, double *__restrict__ row_sums) {
  const int py = blockIdx.x;
  double s = 0.0;
  if (py >= 1 && py < I.ph - 1) {
    const size_t row = (size_t)py * I.pitch;
    for (int px = 2 * threadIdx.x; px < I.pw; px += 2 * blockDim.x) {
      const size_t p = row + px;
      const uchar2 c = *reinterpret_cast<const uchar2 *>(I.cls + p);
      if (!(c.x | c.y)) continue;
      const double2 xc = *reinterpret_cast<const double2 *>(I.x + p);
      const double2 xu = *reinterpret_cast<const double2 *>(I.x + p - I.pitch);
      const double2 xd = *reinterpret_cast<const double2 *>(I.x + p + I.pitch);
      const double2 b = *reinterpret_cast<const double2 *>(I.bp + p);
      const double xl = I.x[p - 1], xr = I.x[p + 2];
      const double n0 = ((xl + xc.y) + xu.x) + xd.x;
      const double n1 = ((xc.x + xr) + xu.y) + xd.y;
      const double r0 = fma(I.dg64, -xc.x, n0) - b.x;
      const double r1 = fma(I.dg64, -xc.y, n1) - b.y;
      float2 gv;
      gv.x = c.x ? (float)(-r0) : __builtin_nanf("");
      gv.y = c.y ? (float)(-r1) : __builtin_nanf("");
      *reinterpret_cast<float2 *>(I.g + p) = gv;
      if (c.x) s = fma(r0, r0, s);
      if (c.y) s = fma(r1, r1, s);
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_sums[py] = red[0];
}

// ||r|| from the row sums of k_residual, added in a fixed order: thread t
// takes a contiguous run of rows, then a tree over the threads. One block.
__global__ __launch_bounds__(256) void k_residual_norm(const double *__restrict__ row_sums,
                                                       int nrows, double *__restrict__ out) {
  __shared__ double red[256];
  const int per = (nrows + 255) / 256;
  double v = 0.0;
  for (int r = threadIdx.x * per; r < min(nrows, (int)(threadIdx.x + 1) * per); ++r)
    v += row_sums[r];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sqrt(red[0]);
    out[1] = red[0];
  }
}

// The refinement of a list of instances in three launches instead of three
// per instance (a batch of many small images refines hundreds of instances a
// pass). rows[j] = (instance slot in insts, plane row, correction buffer).
// x += e, e = 0 at the interior cells of each listed row.
__global__ __launch_bounds__(256) void k_apply_correction_rows(const PhInst *__restrict__ insts,
                                                               const int4 *__restrict__ rows) {
  const int4 r = rows[blockIdx.x];
  const PhInst &I = insts[r.x];
  const size_t row = (size_t)r.y * I.pitch;
  float *e = I.e[r.z];
  for (int px = threadIdx.x; px < I.pw; px += 256) {
    const size_t p = row + px;
    if (I.cls[p]) {
      I.x[p] += (double)e[p];
      e[p] = 0.0f;
    }
  }
}

// k_residual for every listed row (same arithmetic, same pairs, same row sum
// order); row_sums[j] is the sum of row j of the list.
__global__ __launch_bounds__(256) void k_residual_rows(const PhInst *__restrict__ insts,
                                                       const int4 *__restrict__ rows,
                                                       double *__restrict__ row_sums) {
  const int4 rr = rows[blockIdx.x];
  const PhInst &I = insts[rr.x];
  const int py = rr.y;
  double s = 0.0;
  if (py >= 1 && py < I.ph - 1) {
    const size_t row = (size_t)py * I.pitch;
    for (int px = 2 * threadIdx.x; px < I.pw; px += 2 * blockDim.x) {
      const size_t p = row + px;
      const uchar2 c = *reinterpret_cast<const uchar2 *>(I.cls + p);
      if (!(c.x | c.y)) continue;
      const double2 xc = *reinterpret_cast<const double2 *>(I.x + p);
      const double2 xu = *reinterpret_cast<const double2 *>(I.x + p - I.pitch);
      const double2 xd = *reinterpret_cast<const double2 *>(I.x + p + I.pitch);
      const double2 b = *reinterpret_cast<const double2 *>(I.bp + p);
      const double xl = I.x[p - 1], xr = I.x[p + 2];
      const double n0 = ((xl + xc.y) + xu.x) + xd.x;
      const double n1 = ((xc.x + xr) + xu.y) + xd.y;
      const double r0 = fma(I.dg64, -xc.x, n0) - b.x;
      const double r1 = fma(I.dg64, -xc.y, n1) - b.y;
      float2 gv;
      gv.x = c.x ? (float)(-r0) : __builtin_nanf("");
      gv.y = c.y ? (float)(-r1) : __builtin_nanf("");
      *reinterpret_cast<float2 *>(I.g + p) = gv;
      if (c.x) s = fma(r0, r0, s);
      if (c.y) s = fma(r1, r1, s);
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_sums[blockIdx.x] = red[0];
}

// ||r|| of each listed instance from its rows (list[i] = (first row in
// row_sums, row count, output slot)), added exactly as k_residual_norm adds
// an instance's rows, so an instance's norm does not depend on the list.
__global__ __launch_bounds__(256) void k_residual_norm_list(const int4 *__restrict__ list,
                                                            const double *__restrict__ row_sums,
                                                            double *__restrict__ out) {
  const int4 L = list[blockIdx.x];
  const double *rs = row_sums + L.x;
  const int nrows = L.y;
  __shared__ double red[256];
  const int per = (nrows + 255) / 256;
  double v = 0.0;
  for (int r = threadIdx.x * per; r < min(nrows, (int)(threadIdx.x + 1) * per); ++r) v += rs[r];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * (size_t)L.z] = sqrt(red[0]);
    out[2 * (size_t)L.z + 1] = red[0];
  }
}

// Write the solved interior cells of an instance into the interleaved
// candidate; every other sample already holds the target's bits.
__global__ void k_scatter(PhInst I, float *__restrict__ out) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y;
  if (px >= I.pw) return;
  const size_t p = (size_t)py * I.pitch + px;
  if (I.cls[p]) {
    const size_t q = (size_t)(I.oy + py) * I.img_w + (I.ox + px);
    out[q * I.channels + I.channel] = (float)I.x[p];
  }
}

}  // namespace ph
