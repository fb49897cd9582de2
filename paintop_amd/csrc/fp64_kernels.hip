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

// Block-wide min and max of one int pair per thread (256 threads); the
// result is valid in thread 0.
__device__ __forceinline__ int2 block_minmax(int lo, int hi) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  __shared__ int slo[4], shi[4];
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  return make_int2(min(min(slo[0], slo[1]), min(slo[2], slo[3])),
                   max(max(shi[0], shi[1]), max(shi[2], shi[3])));
}

// Set-up of every image of a call, one block per image row (spans: slot =
// image, nrows = H). The candidate row starts as the target row, so every
// non-interior sample keeps its bits; the row's pixels are classified; and
// the row's interior column range goes to rowx[b] ((INT_MAX, -1) when the
// row has none).
__global__ __launch_bounds__(256) void k_classify_spans(const PhSpan *__restrict__ spans, int n,
                                                        const PhItemIO *__restrict__ io,
                                                        int2 *__restrict__ rowx) {
  const int b = blockIdx.x;
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const PhItemIO &G = io[S.slot];
  const int y = b - S.row0;
  const int w = G.W, h = G.H;
  const size_t rl = (size_t)w * G.C;
  const float *t = G.tgt + (size_t)y * rl;
  float *o = G.out + (size_t)y * rl;
  if (((reinterpret_cast<uintptr_t>(G.tgt) | reinterpret_cast<uintptr_t>(G.out)) & 15) == 0 &&
      (rl & 3) == 0) {
    for (size_t q = threadIdx.x; q < rl / 4; q += 256)
      reinterpret_cast<float4 *>(o)[q] = reinterpret_cast<const float4 *>(t)[q];
  } else {
    for (size_t q = threadIdx.x; q < rl; q += 256) o[q] = t[q];
  }
  const bool row_in = w >= 3 && h >= 3 && y >= 1 && y < h - 1;
  int xmin = 0x7fffffff, xmax = -1;
  for (int x = threadIdx.x; x < w; x += 256) {
    const size_t idx = (size_t)y * w + x;
    uint8_t in = 0;
    if (row_in && x >= 1 && x < w - 1) in = G.mask[idx] > MASK_THRESHOLD ? 1 : 0;
    G.cls_img[idx] = in;
    if (in) {
      xmin = min(xmin, x);
      xmax = max(xmax, x);
    }
  }
  const int2 r = block_minmax(xmin, xmax);
  if (threadIdx.x == 0) rowx[b] = r;
}

// Bounding box of each image's interior cells from its rows' column ranges
// (one block per image): bbox[4 * image + {0, 1, 2, 3}] = xmin, ymin, xmax,
// ymax (xmax = -1: no interior cell).
__global__ __launch_bounds__(256) void k_bbox_spans(const PhSpan *__restrict__ spans,
                                                    const int2 *__restrict__ rowx,
                                                    int *__restrict__ bbox) {
  const PhSpan S = spans[blockIdx.x];
  int xmin = 0x7fffffff, xmax = -1, ymin = 0x7fffffff, ymax = -1;
  for (int y = threadIdx.x; y < S.nrows; y += 256) {
    const int2 r = rowx[S.row0 + y];
    if (r.y >= 0) {
      xmin = min(xmin, r.x);
      xmax = max(xmax, r.y);
      ymin = min(ymin, y);
      ymax = max(ymax, y);
    }
  }
  const int2 bx = block_minmax(xmin, xmax);
  __syncthreads();  // block_minmax's LDS is reused
  const int2 by = block_minmax(ymin, ymax);
  if (threadIdx.x == 0) {
    int *o = bbox + 4 * (size_t)S.slot;
    o[0] = bx.x;
    o[1] = by.x;
    o[2] = bx.y;
    o[3] = by.y;
  }
}

// Interior cells of each tile origin, for every image of a call in one launch
// (one block per origin): counts[2i] over the tile's valid core (hb rows x
// wv columns from (K, K)), counts[2i + 1] over every cell the tile loads
// (hb + 2K rows x 256 columns); a tile whose loaded cells are all interior
// needs no masking of non-interior cells.
__global__ void k_tile_occupancy_all(const PhItemGeo *__restrict__ items,
                                     const int *__restrict__ origin_item,
                                     const int2 *__restrict__ origins, int K, int wv, int hb,
                                     int *__restrict__ counts) {
  const PhItemGeo G = items[origin_item[blockIdx.x]];
  const int2 o = origins[blockIdx.x];
  int n = 0, nall = 0;
  for (int r = 0; r < hb + 2 * K; ++r) {
    const int iy = G.oy + o.y + r;
    if (iy < 0 || iy >= G.h) continue;
    const bool core_row = r >= K && r < K + hb;
    for (int cidx = threadIdx.x; cidx < wv + 2 * K; cidx += blockDim.x) {
      const int ix = G.ox + o.x + cidx;
      if (ix >= 0 && ix < G.w) {
        const int v = G.cls[(size_t)iy * G.w + ix];
        nall += v;
        if (core_row && cidx >= K && cidx < K + wv) n += v;
      }
    }
  }
  __shared__ int red[2][256];
  red[0][threadIdx.x] = n;
  red[1][threadIdx.x] = nall;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[2 * blockIdx.x] = red[0][0];
    counts[2 * blockIdx.x + 1] = red[1][0];
  }
}

// splitmix64's output function (Steele, Lea and Flood 2014): a bijective mix
// of a 64-bit integer.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The coarse mask the SOR-factor tuner works on, for every image of a call in
// one launch: a block of f x f cells of an image's region is inside when at
// least half of its cells are interior. ph_coarse_parts blocks per coarse row
// of every listed image, each for a contiguous part of the row's cells
// (spans: slot = image, nrows = the parts x its coarse rows). The 4 waves
// share the coarse row's f image rows and read along them (coalesced), each
// lane adds its columns over its rows, and the sums meet in LDS counters
// (integer adds: the counts do not depend on the order). Writes the row's
// mask bytes and row_hash[block] = the sum mod 2^64 of mix64(index) over the
// part's inside cells (index = cy * cw + c within the image's mask), so an
// image's key (k_coarse_hash_items) is independent of the summation order.
// cw <= PH_COARSE_MAX.
constexpr int PH_COARSE_MAX = 256;
__global__ __launch_bounds__(256) void k_coarse_count_all(const PhItemGeo *__restrict__ items,
                                                          const PhSpan *__restrict__ spans,
                                                          int n, uint8_t *__restrict__ mask,
                                                          uint64_t *__restrict__ row_hash) {
  const PhSpan S = spans[ph_span_find(spans, n, blockIdx.x)];
  const PhItemGeo G = items[S.slot];
  const int local = blockIdx.x - S.row0;
  const int np = ph_coarse_parts(G.cw, G.cf);
  const int cy = local / np, part = local - cy * np;
  const int c_lo = part * G.cw / np, c_hi = (part + 1) * G.cw / np;
  __shared__ int cnt[PH_COARSE_MAX];
  for (int c = threadIdx.x; c < c_hi - c_lo; c += 256) cnt[c] = 0;
  __syncthreads();
  const int xs = c_lo * G.cf, ncols = (c_hi - c_lo) * G.cf;
  // wave w takes the image rows w, w + 4, ...; its lanes take columns
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int J = 16;  // columns per lane per chunk
  for (int x0 = 0; x0 < ncols; x0 += 64 * J) {
    int n[J];
#pragma unroll
    for (int j = 0; j < J; ++j) n[j] = 0;
    for (int dy = wave; dy < G.cf; dy += 4) {
      const int iy = G.ry0 + cy * G.cf + dy;
      if (iy < 0 || iy >= G.h) continue;
      const uint8_t *row = G.cls + (size_t)iy * G.w;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int x = x0 + j * 64 + lane;
        const int ix = G.rx0 + xs + x;
        if (x < ncols && ix >= 0 && ix < G.w) n[j] += row[ix];
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = x0 + j * 64 + lane;
      if (n[j]) atomicAdd(&cnt[x / G.cf], n[j]);
    }
  }
  __syncthreads();
  uint64_t h = 0;
  const int full = G.cf * G.cf;
  for (int c = threadIdx.x; c < c_hi - c_lo; c += 256) {
    const bool in = 2 * cnt[c] >= full;
    const int64_t i = (int64_t)cy * G.cw + c_lo + c;
    mask[G.coarse_off + i] = in ? 1 : 0;
    if (in) h += mix64((uint64_t)i);
  }
  __shared__ uint64_t hs[256];
  hs[threadIdx.x] = h;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) hs[threadIdx.x] += hs[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_hash[blockIdx.x] = hs[0];
}

// Per listed image (one block per span entry of k_coarse_count_all's table):
// hash[e] = the sum mod 2^64 of its rows' row_hash.
__global__ __launch_bounds__(256) void k_coarse_hash_items(const PhSpan *__restrict__ spans,
                                                           const uint64_t *__restrict__ row_hash,
                                                           uint64_t *__restrict__ hash) {
  const PhSpan S = spans[blockIdx.x];
  uint64_t h = 0;
  for (int r = threadIdx.x; r < S.nrows; r += 256) h += row_hash[S.row0 + r];
  __shared__ uint64_t hs[256];
  hs[threadIdx.x] = h;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) hs[threadIdx.x] += hs[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) hash[blockIdx.x] = hs[0];
}

// Tree sum over the 256 threads of a block, in a fixed order (thread 0).
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  return red[0];
}

// Fill the planes of the listed images' channel instances, one block per plane
// row (spans: slot = the first of the image's C consecutive descriptors, one
// per channel, a descriptor with no planes (x == nullptr) is skipped; item =
// the image; out = the first descriptor's first row in row_sums, the
// channels' rows following each other, ph apart). Each pixel's channels are
// read together, so the interleaved source and target are read once per
// image, not once per channel.
// b = fma(4, -c, ((l + r) + u) + d) of the source (poisson.rs:364-373),
// b' = fma(lambda, -anchor, b) (poisson.rs:319); x starts at the anchor.
// The anchor start's residual comes in the same pass: at an interior cell
// r = fma(4 + lambda, -x, ((x_l + x_r) + x_u) + x_d) - b' with x the anchor
// at the cell and its four neighbours (all in the image: interior cells are
// never on its border), the FP64 residual's own operation order
// (poisson.rs:305-325); g = -r (FP32) there and NaN elsewhere, and the row's
// sum of r^2 goes to row_sums (each thread's cells in column order, then the
// block's fixed tree).
// CC = C for 1..4 (the descriptors' pointers then sit in registers), 0 for
// any C (read per cell; the residual then takes a pass per channel).
template <int CC>
__global__ __launch_bounds__(256) void k_assemble_spans(const PhInst *__restrict__ insts,
                                                       const PhSpan *__restrict__ spans, int n,
                                                       const PhItemIO *__restrict__ io,
                                                       double *__restrict__ row_sums) {
  const int b = xcd_row_group(blockIdx.x, gridDim.x);  // reads the source rows around b
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const PhItemIO &G = io[S.item];
  const PhInst I0 = insts[S.slot];  // the plane geometry, shared by the channels
  const int py = b - S.row0;
  const int W = G.W, H = G.H, C = CC ? CC : G.C;
  const float *__restrict__ src = G.src;
  const float *__restrict__ tgt = G.tgt;
  const uint8_t *__restrict__ cls_img = G.cls_img;
  uint8_t *__restrict__ cls_plane = G.cls_plane;
  constexpr int R = CC ? CC : 1;
  double *xs[R], *bps[R];
  float *e0s[R], *e1s[R], *gs[R];
  double rs[R];
#pragma unroll
  for (int c = 0; c < R; ++c) rs[c] = 0.0;
  if (CC) {
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const PhInst &I = insts[S.slot + c];
      xs[c] = I.x;
      bps[c] = I.bp;
      e0s[c] = I.e[0];
      e1s[c] = I.e[1];
      gs[c] = I.g;
    }
  }
  const double lambda = I0.lambda;
  const double dg = I0.dg64;
  // b' of channel c at interior image cell q, and its anchor residual
  auto bprime = [&](size_t q, int c, double a) {
    const double cc = (double)src[q * C + c];
    const double l = (double)src[(q - 1) * C + c];
    const double r = (double)src[(q + 1) * C + c];
    const double u = (double)src[(q - W) * C + c];
    const double d = (double)src[(q + W) * C + c];
    const double bb = fma(4.0, -cc, ((l + r) + u) + d);
    return fma(lambda, -a, bb);
  };
  auto anchor_r = [&](size_t q, int c, double a, double bp) {
    const double xl = (double)tgt[(q - 1) * C + c], xr = (double)tgt[(q + 1) * C + c];
    const double xu = (double)tgt[(q - W) * C + c], xd = (double)tgt[(q + W) * C + c];
    return fma(dg, -a, ((xl + xr) + xu) + xd) - bp;
  };
  if constexpr (CC != 0) {
    // a known channel count: each pixel of the stencil is loaded once for all
    // of its channels (one 16-byte load for four channels when both images are
    // 16-byte aligned), and every channel is formed from registers with the
    // operation order of bprime / anchor_r
    const bool vec = CC == 4 && ((((uintptr_t)src) | ((uintptr_t)tgt)) & 15) == 0;  // uniform
    auto ldpx = [&](const float *__restrict__ base, size_t q, float (&v)[R]) {
      if (CC == 4 && vec) {
        const float4 t = *reinterpret_cast<const float4 *>(base + q * 4);
        v[0] = t.x;
        v[1 % R] = t.y;
        v[2 % R] = t.z;
        v[3 % R] = t.w;
      } else {
#pragma unroll
        for (int c = 0; c < R; ++c) v[c] = base[q * CC + c];
      }
    };
    for (int px = threadIdx.x; px < I0.pitch; px += 256) {
      const size_t p = (size_t)py * I0.pitch + px;
      const int ix = I0.ox + px, iy = I0.oy + py;
      const bool inside = px < I0.pw && ix >= 0 && ix < W && iy >= 0 && iy < H;
      const size_t q = inside ? (size_t)iy * W + ix : 0;
      const uint8_t in = inside ? cls_img[q] : 0;
      cls_plane[p] = in;
      float ta[R], sc[R], sl[R], sr[R], su[R], sd[R], tl[R], tr[R], tu[R], td[R];
#pragma unroll
      for (int c = 0; c < R; ++c) ta[c] = 0.0f;
      if (inside) ldpx(tgt, q, ta);
      if (in) {
        // interior cells are never on the image border: four in-image neighbours
        ldpx(src, q, sc);
        ldpx(src, q - 1, sl);
        ldpx(src, q + 1, sr);
        ldpx(src, q - W, su);
        ldpx(src, q + W, sd);
        ldpx(tgt, q - 1, tl);
        ldpx(tgt, q + 1, tr);
        ldpx(tgt, q - W, tu);
        ldpx(tgt, q + W, td);
      }
#pragma unroll
      for (int c = 0; c < R; ++c) {
        if (!xs[c]) continue;
        const double a = (double)ta[c];
        double bp = 0.0;
        float gv = __builtin_nanf("");
        if (in) {
          const double bb = fma(4.0, -(double)sc[c],
                                (((double)sl[c] + (double)sr[c]) + (double)su[c]) + (double)sd[c]);
          bp = fma(lambda, -a, bb);
          const double r =
              fma(dg, -a, (((double)tl[c] + (double)tr[c]) + (double)tu[c]) + (double)td[c]) - bp;
          gv = (float)(-r);
          rs[c] = fma(r, r, rs[c]);
        }
        xs[c][p] = a;
        bps[c][p] = bp;
        e0s[c][p] = 0.0f;
        e1s[c][p] = 0.0f;
        gs[c][p] = gv;
      }
    }
  } else {
    // any C: one channel at a time (the residual follows in a second pass)
    for (int px = threadIdx.x; px < I0.pitch; px += 256) {
      const size_t p = (size_t)py * I0.pitch + px;
      const int ix = I0.ox + px, iy = I0.oy + py;
      const bool inside = px < I0.pw && ix >= 0 && ix < W && iy >= 0 && iy < H;
      const size_t q = inside ? (size_t)iy * W + ix : 0;
      const uint8_t in = inside ? cls_img[q] : 0;
      cls_plane[p] = in;
      for (int c = 0; c < C; ++c) {
        const PhInst &I = insts[S.slot + c];
        if (!I.x) continue;
        double a = 0.0, bp = 0.0;
        if (inside) {
          a = (double)tgt[q * C + c];
          // interior cells are never on the image border: four in-image neighbours
          if (in) bp = bprime(q, c, a);
        }
        I.x[p] = a;
        I.bp[p] = bp;
        I.e[0][p] = 0.0f;
        I.e[1][p] = 0.0f;
        I.g[p] = __builtin_nanf("");
      }
    }
  }
  if (CC) {
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const double t = block_sum(rs[c]);
      if (threadIdx.x == 0 && xs[c]) row_sums[S.out + (size_t)c * I0.ph + py] = t;
      __syncthreads();
    }
  } else {
    // any C: the residual in a pass per channel over the row (its g written
    // by this thread in the pass above, so no other thread's writes are read)
    for (int c = 0; c < C; ++c) {
      const PhInst &I = insts[S.slot + c];
      if (!I.x) continue;
      double sr = 0.0;
      for (int px = threadIdx.x; px < I0.pitch; px += 256) {
        const size_t p = (size_t)py * I0.pitch + px;
        const int ix = I0.ox + px, iy = I0.oy + py;
        const bool inside = px < I0.pw && ix >= 0 && ix < W && iy >= 0 && iy < H;
        if (!inside) continue;
        const size_t q = (size_t)iy * W + ix;
        if (!cls_img[q]) continue;
        const double a = (double)tgt[q * C + c];
        const double r = anchor_r(q, c, a, bprime(q, c, a));
        I.g[p] = (float)(-r);
        sr = fma(r, r, sr);
      }
      const double t = block_sum(sr);
      if (threadIdx.x == 0) row_sums[S.out + (size_t)c * I0.ph + py] = t;
      __syncthreads();
    }
  }
}
template __global__ void k_assemble_spans<0>(const PhInst *, const PhSpan *, int,
                                             const PhItemIO *, double *);
template __global__ void k_assemble_spans<1>(const PhInst *, const PhSpan *, int,
                                             const PhItemIO *, double *);
template __global__ void k_assemble_spans<2>(const PhInst *, const PhSpan *, int,
                                             const PhItemIO *, double *);
template __global__ void k_assemble_spans<3>(const PhInst *, const PhSpan *, int,
                                             const PhItemIO *, double *);
template __global__ void k_assemble_spans<4>(const PhInst *, const PhSpan *, int,
                                             const PhItemIO *, double *);

// g = -r(x) in FP64, stored FP32 (NaN at non-interior cells), and this
// thread's part of the row's sum of r^2 (the block adds the parts in a fixed
// order; a separate launch adds the rows, since a last-block reduction in the
// same kernel costs one device-scope atomic per row block, which serialise
// across the XCDs and took longer than the rows themselves).
// One block per plane row. Each thread takes cell pairs (px, px + 1), px even,
// so the loads of a pair are independent 16-byte vectors; rows 0 and ph - 1
// and columns 0 and pw - 1 are padding (no interior cells), so every
// neighbour read stays inside the plane. Non-interior cells of a pair get
// g = NaN, the value they hold already.
__device__ __forceinline__ double residual_row(const PhInst &I, int py) {
  double s = 0.0;
  if (py >= 1 && py < I.ph - 1) {
    // the instance's fields in registers: the store to g could alias them
    const double *__restrict__ x = I.x;
    const double *__restrict__ bp = I.bp;
    const uint8_t *__restrict__ cls = I.cls;
    float *__restrict__ g = I.g;
    const double dg = I.dg64;
    const int pw = I.pw;
    const size_t pitch = (size_t)I.pitch;
    const size_t row = (size_t)py * pitch;
    for (int px = 2 * threadIdx.x; px < pw; px += 2 * blockDim.x) {
      const size_t p = row + px;
      const uchar2 c = *reinterpret_cast<const uchar2 *>(cls + p);
      if (!(c.x | c.y)) continue;
      const double2 xc = *reinterpret_cast<const double2 *>(x + p);
      const double2 xu = *reinterpret_cast<const double2 *>(x + p - pitch);
      const double2 xd = *reinterpret_cast<const double2 *>(x + p + pitch);
      const double2 b = *reinterpret_cast<const double2 *>(bp + p);
      const double xl = x[p - 1], xr = x[p + 2];
      const double n0 = ((xl + xc.y) + xu.x) + xd.x;
      const double n1 = ((xc.x + xr) + xu.y) + xd.y;
      const double r0 = fma(dg, -xc.x, n0) - b.x;
      const double r1 = fma(dg, -xc.y, n1) - b.y;
      float2 gv;
      gv.x = c.x ? (float)(-r0) : __builtin_nanf("");
      gv.y = c.y ? (float)(-r1) : __builtin_nanf("");
      *reinterpret_cast<float2 *>(g + p) = gv;
      if (c.x) s = fma(r0, r0, s);
      if (c.y) s = fma(r1, r1, s);
    }
  }
  return s;
}

// Fixed-order sum of nrows row sums (thread 0): thread t takes a contiguous
// run of rows, then the tree over the threads.
__device__ __forceinline__ double rows_sum(const double *__restrict__ rs, int nrows) {
  const int per = (nrows + 255) / 256;
  double v = 0.0;
  for (int r = threadIdx.x * per; r < min(nrows, (int)(threadIdx.x + 1) * per); ++r) v += rs[r];
  return block_sum(v);
}

// The refinement of many instances in three launches, not three per instance
// (a batch of small images refines hundreds of instances a pass).
// x += e[buf], e[buf] = 0 at the interior cells of every listed row.
// ||r(x + e)||^2 of a plane row in FP64, written nowhere: the residual of the
// state a member of the warm-start race has reached (x and the correction e
// it carries), the value a refinement would compute, judged without one so
// neither member's trajectory changes. e is 0 off the interior.
__device__ __forceinline__ double probe_residual_row(const PhInst &I, const float *e, int py) {
  double s = 0.0;
  if (py >= 1 && py < I.ph - 1) {
    const size_t row = (size_t)py * I.pitch;
    const size_t pitch = (size_t)I.pitch;
    for (int px = 2 * threadIdx.x; px < I.pw; px += 2 * blockDim.x) {
      const size_t p = row + px;
      const uchar2 c = *reinterpret_cast<const uchar2 *>(I.cls + p);
      if (!(c.x | c.y)) continue;
      const double xc0 = I.x[p] + (double)e[p], xc1 = I.x[p + 1] + (double)e[p + 1];
      const double xl = I.x[p - 1] + (double)e[p - 1], xr = I.x[p + 2] + (double)e[p + 2];
      const double xu0 = I.x[p - pitch] + (double)e[p - pitch];
      const double xu1 = I.x[p + 1 - pitch] + (double)e[p + 1 - pitch];
      const double xd0 = I.x[p + pitch] + (double)e[p + pitch];
      const double xd1 = I.x[p + 1 + pitch] + (double)e[p + 1 + pitch];
      const double n0 = ((xl + xc1) + xu0) + xd0;
      const double n1 = ((xc0 + xr) + xu1) + xd1;
      const double r0 = fma(I.dg64, -xc0, n0) - I.bp[p];
      const double r1 = fma(I.dg64, -xc1, n1) - I.bp[p + 1];
      if (c.x) s = fma(r0, r0, s);
      if (c.y) s = fma(r1, r1, s);
    }
  }
  return s;
}
__global__ __launch_bounds__(256) void k_probe_residual_spans(const PhInst *__restrict__ insts,
                                                              const PhSpan *__restrict__ spans,
                                                              int n,
                                                              double *__restrict__ row_sums) {
  const int b = xcd_row_group(blockIdx.x, gridDim.x);
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const PhInst &I = insts[S.slot];
  const double s = block_sum(probe_residual_row(I, I.e[S.buf], b - S.row0));
  if (threadIdx.x == 0) row_sums[b] = s;
}
// A learned coarse correction added to the FP32 state, and the residual of
// the corrected state, in one pass: one block per plane row of every listed
// correction (span entry i takes args[i]). The state is x + e with e in the
// buffer the entry's control block names; the corrected correction e' = e + alpha * scale * E goes to
// the other buffer (so taking the correction back costs nothing: the solve
// keeps S.buf), where E interpolates the prediction h bilinearly in FP32
// (image cell ix sits at coarse coordinate (ix - x0 + 0.5) / f - 0.5, clamped
// to the grid): vertically once per row into LDS, then along the row.
// row_sums[b] = the row's sum of r(x + e')^2 = (L e' - g)^2 in FP64 over
// each cell's FP32 residual (the sweeps' own formula). A block computes e' at its row's neighbour
// cells with the same arithmetic as the blocks that own them, so it reads no
// other block's writes.
__global__ __launch_bounds__(256) void k_corr_apply_spans(const PhInst *__restrict__ insts,
                                                          const PhCtl *__restrict__ ctl,
                                                          const PhSpan *__restrict__ spans, int n,
                                                          const PhCorrArgs *__restrict__ args,
                                                          const float *__restrict__ hc,
                                                          const float *__restrict__ scale,
                                                          const double *__restrict__ alpha,
                                                          int res, double *__restrict__ row_sums) {
  const int b = xcd_row_group(blockIdx.x, gridDim.x);
  const int i = ph_span_find(spans, n, b);
  const PhSpan S = spans[i];
  const PhCorrArgs A = args[i];
  const PhInst &I = insts[S.slot];
  const float ca = (float)(alpha[A.net] * (double)scale[A.net]);
  const float *h = hc + (size_t)res * res * A.net;
  const int py = b - S.row0;
  const float invf = 1.0f / (float)A.f;
  __shared__ float vrow[3][PH_CORR_RES_MAX];  // rows py - 1, py, py + 1
  for (int t = threadIdx.x; t < 3 * res; t += 256) {
    const int r = t / res, c = t - r * res;
    const int iy = I.oy + py - 1 + r;
    const float v = fminf(fmaxf(((float)(iy - A.y0) + 0.5f) * invf - 0.5f, 0.0f), (float)(res - 1));
    const int j0 = min((int)v, res - 1), j1 = min(j0 + 1, res - 1);
    const float h0 = h[(size_t)j0 * res + c], h1 = h[(size_t)j1 * res + c];
    vrow[r][c] = fmaf(v - (float)j0, h1 - h0, h0);
  }
  __syncthreads();
  // The instance's fields in registers (a store to e could alias them).
  const uint32_t buf = ctl[S.slot].cur_buf;
  const float *__restrict__ eo = I.e[buf];
  float *__restrict__ en = I.e[buf ^ 1u];
  const float *__restrict__ g = I.g;
  const uint8_t *__restrict__ cls = I.cls;
  const float dg = I.dg;
  const int pw = I.pw, ox = I.ox;
  const size_t pitch = (size_t)I.pitch;
  // the horizontal interpolation position of plane column px
  auto col = [&](int px, int &i0, int &i1, float &w) {
    const float u =
        fminf(fmaxf(((float)(ox + px - A.x0) + 0.5f) * invf - 0.5f, 0.0f), (float)(res - 1));
    i0 = min((int)u, res - 1);
    i1 = min(i0 + 1, res - 1);
    w = u - (float)i0;
  };
  // e' at a cell of row r (0: py - 1, 1: py, 2: py + 1) from its e; e = 0 and
  // no correction off the interior
  auto upd = [&](float ev, int r, bool in, int i0, int i1, float w) -> float {
    return in ? ev + ca * fmaf(w, vrow[r][i1] - vrow[r][i0], vrow[r][i0]) : ev;
  };
  double s = 0.0;
  if (py >= 1 && py < I.ph - 1) {
    const size_t row = (size_t)py * pitch;
    // two cells per thread (pw is even; interior cells lie in [1, pw - 1))
    for (int px = 2 * threadIdx.x; px < pw; px += 512) {
      const size_t p = row + px;
      const uchar2 c = *reinterpret_cast<const uchar2 *>(cls + p);
      if (!(c.x | c.y)) continue;
      const uchar2 cu = *reinterpret_cast<const uchar2 *>(cls + p - pitch);
      const uchar2 cd = *reinterpret_cast<const uchar2 *>(cls + p + pitch);
      const bool cl = cls[p - 1] != 0, cr = cls[p + 2] != 0;
      int a0, a1, b0, b1, l0, l1, r0_, r1_;
      float wa, wb, wl, wr;
      col(px, a0, a1, wa);
      col(px + 1, b0, b1, wb);
      col(px - 1, l0, l1, wl);
      col(px + 2, r0_, r1_, wr);
      const float2 ec = *reinterpret_cast<const float2 *>(eo + p);
      const float2 eu = *reinterpret_cast<const float2 *>(eo + p - pitch);
      const float2 ed = *reinterpret_cast<const float2 *>(eo + p + pitch);
      const float2 gv = *reinterpret_cast<const float2 *>(g + p);
      const float e0 = upd(ec.x, 1, c.x, a0, a1, wa), e1 = upd(ec.y, 1, c.y, b0, b1, wb);
      if (c.x) en[p] = e0;
      if (c.y) en[p + 1] = e1;
      const float el = upd(eo[p - 1], 1, cl, l0, l1, wl);
      const float er = upd(eo[p + 2], 1, cr, r0_, r1_, wr);
      const float eu0 = upd(eu.x, 0, cu.x, a0, a1, wa), eu1 = upd(eu.y, 0, cu.y, b0, b1, wb);
      const float ed0 = upd(ed.x, 2, cd.x, a0, a1, wa), ed1 = upd(ed.y, 2, cd.y, b0, b1, wb);
      // the sweeps' own FP32 residual of each cell, squares summed in FP64
      const double q0 = fmaf(-dg, e0, (((el + e1) + eu0) + ed0) - gv.x);
      const double q1 = fmaf(-dg, e1, (((e0 + er) + eu1) + ed1) - gv.y);
      if (c.x) s = fma(q0, q0, s);
      if (c.y) s = fma(q1, q1, s);
    }
  }
  const double tot = block_sum(s);
  if (threadIdx.x == 0) row_sums[b] = tot;
}

__global__ __launch_bounds__(256) void k_apply_correction_spans(const PhInst *__restrict__ insts,
                                                                const PhSpan *__restrict__ spans,
                                                                int n) {
  const int b = blockIdx.x;
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const PhInst &I = insts[S.slot];
  // the instance's fields in registers (the stores could alias them); cell
  // pairs (pw and the pitch are even) with vector loads and stores
  double *__restrict__ x = I.x;
  const uint8_t *__restrict__ cls = I.cls;
  float *__restrict__ e = I.e[S.buf];
  const int pw = I.pw;
  const size_t row = (size_t)(b - S.row0) * I.pitch;
  for (int px = 2 * threadIdx.x; px < pw; px += 512) {
    const size_t p = row + px;
    const uchar2 c = *reinterpret_cast<const uchar2 *>(cls + p);
    if (!(c.x | c.y)) continue;
    double2 xv = *reinterpret_cast<const double2 *>(x + p);
    float2 ev = *reinterpret_cast<const float2 *>(e + p);
    if (c.x) {
      xv.x += (double)ev.x;
      ev.x = 0.0f;
    }
    if (c.y) {
      xv.y += (double)ev.y;
      ev.y = 0.0f;
    }
    *reinterpret_cast<double2 *>(x + p) = xv;
    *reinterpret_cast<float2 *>(e + p) = ev;
  }
}

// g = -r(x) and the row's sum of r^2 (residual_row) for every listed row;
// row_sums[b] is the sum of block b's row.
__global__ __launch_bounds__(256) void k_residual_spans(const PhInst *__restrict__ insts,
                                                        const PhSpan *__restrict__ spans, int n,
                                                        double *__restrict__ row_sums) {
  const int b = xcd_row_group(blockIdx.x, gridDim.x);
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const double s = block_sum(residual_row(insts[S.slot], b - S.row0));
  if (threadIdx.x == 0) row_sums[b] = s;
}

// ||r|| of each listed instance (one block per entry) from its row sums in a
// fixed order (rows_sum), so the norm does not depend on the table or the
// other entries: out[2 * out_slot] = ||r||, out[2 * out_slot + 1] = ||r||^2,
// and the history entry, when given, receives ||r|| / r0.
__global__ __launch_bounds__(256) void k_residual_norm_spans(const PhSpan *__restrict__ spans,
                                                             const double *__restrict__ row_sums,
                                                             double *__restrict__ out) {
  const PhSpan S = spans[blockIdx.x];
  const double s = rows_sum(row_sums + S.row0, S.nrows);
  if (threadIdx.x == 0) {
    const double r = sqrt(s);
    out[2 * (size_t)S.out] = r;
    out[2 * (size_t)S.out + 1] = s;
    if (S.hist) *S.hist = r / S.r0;
  }
}

// Write the solved interior cells of the listed images into their candidates,
// one block per plane row (spans as for k_assemble_spans: C consecutive
// descriptors per image, one per channel, those with no planes skipped), all
// channels of a pixel by one thread, so each touched line is written once.
// CC as for k_assemble_spans.
template <int CC>
__global__ __launch_bounds__(256) void k_scatter_spans(const PhInst *__restrict__ insts,
                                                       const PhSpan *__restrict__ spans, int n,
                                                       const PhItemIO *__restrict__ io) {
  const int b = blockIdx.x;
  const PhSpan S = spans[ph_span_find(spans, n, b)];
  const PhItemIO &G = io[S.item];
  const PhInst I0 = insts[S.slot];
  const int py = b - S.row0;
  const int C = CC ? CC : G.C;
  float *__restrict__ out = G.out;
  constexpr int R = CC ? CC : 1;
  const double *xs[R];
  if (CC) {
#pragma unroll
    for (int c = 0; c < R; ++c) xs[c] = insts[S.slot + c].x;
  }
  for (int px = threadIdx.x; px < I0.pw; px += 256) {
    const size_t p = (size_t)py * I0.pitch + px;
    if (!I0.cls[p]) continue;
    const size_t q = (size_t)(I0.oy + py) * I0.img_w + (I0.ox + px);
#pragma unroll
    for (int c = 0; c < (CC ? CC : 64); ++c) {
      if (!CC && c >= C) break;
      const double *x = CC ? xs[c] : insts[S.slot + c].x;
      if (x) out[q * C + c] = (float)x[p];
    }
  }
}
template __global__ void k_scatter_spans<0>(const PhInst *, const PhSpan *, int, const PhItemIO *);
template __global__ void k_scatter_spans<1>(const PhInst *, const PhSpan *, int, const PhItemIO *);
template __global__ void k_scatter_spans<2>(const PhInst *, const PhSpan *, int, const PhItemIO *);
template __global__ void k_scatter_spans<3>(const PhInst *, const PhSpan *, int, const PhItemIO *);
template __global__ void k_scatter_spans<4>(const PhInst *, const PhSpan *, int, const PhItemIO *);

// Copy the anchor-start state of every listed instance (x, g) into the planes
// of its copy and clear the copy's corrections, one block per plane row (span
// entry i: from[i] -> to[i]). Taken before a warm start overwrites x and g, the
// copy is the plain start of the warm-start race, equal to a fresh assembly
// and residual bit for bit.
__global__ __launch_bounds__(256) void k_copy_anchor_spans(const PhInst *__restrict__ from,
                                                           const PhInst *__restrict__ to,
                                                           const PhSpan *__restrict__ spans,
                                                           int n) {
  const int b = blockIdx.x;
  const int i = ph_span_find(spans, n, b);
  const PhSpan S = spans[i];
  const PhInst &A = from[i];
  const PhInst &B = to[i];
  const size_t row = (size_t)(b - S.row0) * A.pitch;
  for (int px = threadIdx.x; px < A.pitch; px += 256) {
    const size_t p = row + px;
    B.x[p] = A.x[p];
    B.g[p] = A.g[p];
    B.e[0][p] = 0.0f;
    B.e[1][p] = 0.0f;
  }
}

// Write n values to their device addresses (one launch for many small updates).
__global__ void k_put_values(const PhValue *__restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) *v[i].dst = v[i].val;
}

// Gather residual histories into one array for a single readback (one block
// per history).
__global__ __launch_bounds__(256) void k_gather_hist(const PhHistCopy *__restrict__ c,
                                                     double *__restrict__ dst) {
  const PhHistCopy C = c[blockIdx.x];
  for (int q = threadIdx.x; q < C.n; q += 256) dst[C.off + q] = C.src[q];
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
