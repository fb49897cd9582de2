// warm_kernels.hip -- the learned warm start's data path.
//
// The solution is written as u = P + h with P = the source (guidance) image.
// For the blend operator the source's own 5-point Laplacian is the right-hand
// side (poisson_blend.rs:186-236, poisson.rs:364-373), so h satisfies
//   (Laplacian - lambda) h = lambda * (P - anchor)   at interior cells
//   h = anchor - P                                   on the Dirichlet ring
// which for lambda = 0 is a Laplace problem: h is smooth inside the mask and
// carries no fine texture. The network predicts h on a coarse grid (f x f
// block means of a region around the mask); the solver's start is
// x = P + bilinear(h) at interior cells.
#include <math.h>

#include <algorithm>

#include "../../include/paintop_hip.h"
#include "host_common.h"
#include "ph_internal.h"

namespace ph {

constexpr float WARM_MASK_THRESHOLD = 0.5f;  // poisson.rs:329

__device__ __forceinline__ bool warm_interior(const float *mask, int W, int H, int ix, int iy) {
  return ix >= 1 && ix < W - 1 && iy >= 1 && iy < H - 1 &&
         mask[(size_t)iy * W + ix] > WARM_MASK_THRESHOLD;
}

// Cells of the coarsening kernels' blocks (256 threads). Below
// COARSEN_WAVE_F a block takes whole cells of one coarse row, f columns
// each, at most 256 columns (one thread per column); from COARSEN_WAVE_F on
// a cell has f^2 >= 256 pixels and takes a wave of its own (4 cells a block),
// whose lanes stride its pixels: a thread's chain of dependent loads stays
// short either way (a column of 30 rows per thread ran 25% slower than a wave
// per 30 x 30 cell).
constexpr int COARSEN_WAVE_F = 16;
__host__ __device__ __forceinline__ int coarsen_cells_per_block(int f) {
  return f >= COARSEN_WAVE_F ? 4 : 256 / f;
}
__host__ __device__ __forceinline__ int coarsen_blocks_per_row(int f, int res) {
  const int c = coarsen_cells_per_block(f);
  return (res + c - 1) / c;
}

// Sums of one coarse cell: interior and ring counts, and per channel the rhs
// over interior cells, (anchor - P) over ring cells and (solution - P) over
// both (the training target).
template <int NE>
struct CoarseChannels {
  int c[NE];
};

template <int NE>
struct CoarseSums {
  int n_int = 0, n_ring = 0;
  double rhs[NE] = {}, ring[NE] = {}, t[NE] = {};
};

// Adds pixel (ix, iy) (inside the image) to s, for channels ch[0 .. NE).
// Ring cells are the non-interior cells with an interior 4-neighbour: the
// Dirichlet data of the solve (poisson.rs:374-392 reads them as neighbours).
// mode 0 (warm start): P = source, the rhs is lambda (P - anchor).
// mode 1 (correction problem): P = 0 and the rhs is the 5-point Laplacian of
// the source (poisson.rs:364-373): a blend with a zero anchor is the problem
// Laplacian(u) = rhs, u = 0 on the ring, that a coarse correction solves.
template <int NE, bool TGT, int MODE>
__device__ __forceinline__ void coarsen_pixel(CoarseSums<NE> &cs, const float *__restrict__ src,
                                              const float *__restrict__ anc,
                                              const float *__restrict__ mask,
                                              const float *__restrict__ sol, int W, int H, int C,
                                              CoarseChannels<NE> ch, double lambda, int ix,
                                              int iy) {
  const size_t q = (size_t)iy * W + ix;
  const bool in = warm_interior(mask, W, H, ix, iy);
  const bool rg = !in && (warm_interior(mask, W, H, ix - 1, iy) ||
                          warm_interior(mask, W, H, ix + 1, iy) ||
                          warm_interior(mask, W, H, ix, iy - 1) ||
                          warm_interior(mask, W, H, ix, iy + 1));
  cs.n_int += in ? 1 : 0;
  cs.n_ring += rg ? 1 : 0;
  if (!(in || rg)) return;
  // every sum takes a term (0 where the pixel does not count: the sums start
  // at +0, so an added +0 changes no bit); adds into one array or the other
  // under a branch would be merged into one add through a selected address,
  // which puts the sums in scratch memory
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const double S = (double)src[q * C + ch.c[e]], A = (double)anc[q * C + ch.c[e]];
    const double P = MODE ? 0.0 : S;
    double vr = 0.0;
    if constexpr (MODE) {
      if (in) {
        // interior cells are never on the image border: four in-image neighbours
        const double l = src[(q - 1) * C + ch.c[e]], r = src[(q + 1) * C + ch.c[e]];
        const double u = src[(q - W) * C + ch.c[e]], d = src[(q + W) * C + ch.c[e]];
        vr = fma(4.0, -S, ((l + r) + u) + d);
      }
    } else {
      vr = lambda * (P - A);
    }
    cs.rhs[e] += in ? vr : 0.0;
    cs.ring[e] += rg ? A - P : 0.0;
    if (TGT) cs.t[e] += in ? (double)sol[q * C + ch.c[e]] - P : A - P;
  }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// LDS of a coarsening block: the column sums of its 256 threads (ints: the
// two counts; doubles: up to 3 sums for each of up to 4 channels).
struct CoarseLds {
  int n[2][256];
  double d[12][256];
};

// The f x f blocks (cells) cell0 .. cell0 + ncell - 1 of coarse row cy of the
// region [x0, x0 + res*f) x [y0, y0 + res*f), for channels ch[0 .. NE) of an
// image; put(c, sums) receives cell c's sums, once per cell, from one thread.
// f < COARSEN_WAVE_F: thread t sums plane column t of the block down the
// cells' f rows (in row order), the column sums meet in LDS, and thread c
// adds its cell's f columns in column order. Otherwise wave w takes cell w:
// lane l adds the cell's pixels l, l + 64, ... (row-major), and a fixed
// butterfly adds the lanes. Either order is fixed for a given f, the same for
// every launch of either kernel, so results are deterministic and a channel's
// sums do not depend on which other channels share the block.
template <int NE, bool TGT, int MODE, class Put>
__device__ __forceinline__ void coarsen_block(
    CoarseLds &lds, const float *__restrict__ src, const float *__restrict__ anc,
    const float *__restrict__ mask, const float *__restrict__ sol, int W, int H, int C,
    CoarseChannels<NE> ch, double lambda, int x0, int y0, int f, int cy, int cell0, int ncell,
    const Put &put) {
  const int t = threadIdx.x;
  const int ix0 = x0 + cell0 * f, iy0 = y0 + cy * f;
  if (f >= COARSEN_WAVE_F) {
    const int w = t >> 6, lane = t & 63;
    CoarseSums<NE> cs;
    if (w < ncell) {
      for (int i = lane; i < f * f; i += 64) {
        const int dy = i / f, dx = i - dy * f;
        const int iy = iy0 + dy, ix = ix0 + w * f + dx;
        if (iy < 0 || iy >= H || ix < 0 || ix >= W) continue;
        coarsen_pixel<NE, TGT, MODE>(cs, src, anc, mask, sol, W, H, C, ch, lambda, ix, iy);
      }
    }
    cs.n_int = wave_sum_i(cs.n_int);
    cs.n_ring = wave_sum_i(cs.n_ring);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      cs.rhs[e] = wave_sum_d(cs.rhs[e]);
      cs.ring[e] = wave_sum_d(cs.ring[e]);
      if (TGT) cs.t[e] = wave_sum_d(cs.t[e]);
    }
    if (lane == 0 && w < ncell) put(w, cs);
    return;
  }
  const int ncol = ncell * f;  // <= 256
  const int dy0 = max(0, -iy0), dy1 = min(f, H - iy0);
  {
    const int ix = ix0 + t;
    CoarseSums<NE> cs;
    if (t < ncol && ix >= 0 && ix < W)
      for (int dy = dy0; dy < dy1; ++dy)
        coarsen_pixel<NE, TGT, MODE>(cs, src, anc, mask, sol, W, H, C, ch, lambda, ix, iy0 + dy);
    lds.n[0][t] = cs.n_int;
    lds.n[1][t] = cs.n_ring;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      lds.d[3 * e][t] = cs.rhs[e];
      lds.d[3 * e + 1][t] = cs.ring[e];
      if (TGT) lds.d[3 * e + 2][t] = cs.t[e];
    }
  }
  __syncthreads();
  if (t < ncell) {
    CoarseSums<NE> cell;
    for (int j = t * f; j < t * f + f; ++j) {
      cell.n_int += lds.n[0][j];
      cell.n_ring += lds.n[1][j];
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        cell.rhs[e] += lds.d[3 * e][j];
        cell.ring[e] += lds.d[3 * e + 1][j];
        if (TGT) cell.t[e] += lds.d[3 * e + 2][j];
      }
    }
    put(t, cell);
  }
}

// The four input channels of one cell (laid out [res][res][4]).
template <int NE>
__device__ __forceinline__ void coarsen_put(float *__restrict__ o, const CoarseSums<NE> &s, int e,
                                            int f) {
  const double n_int = s.n_int, n_ring = s.n_ring;
  o[0] = (float)s.rhs[e];  // f^2 * mean
  o[1] = n_ring > 0.0 ? (float)(s.ring[e] / n_ring) : 0.0f;
  o[2] = (float)(n_int * (1.0 / ((double)f * f)));
  o[3] = (float)(n_ring / f);
}

template <bool TGT, int MODE>
__device__ __forceinline__ void coarsen_one(CoarseLds &lds, const float *__restrict__ src,
                                            const float *__restrict__ anc,
                                            const float *__restrict__ mask,
                                            const float *__restrict__ sol, int W, int H, int C,
                                            int ch, double lambda, int x0, int y0, int f, int res,
                                            float *__restrict__ in4, float *__restrict__ tgt,
                                            float *__restrict__ wt) {
  const int nbx = coarsen_blocks_per_row(f, res), cpb = coarsen_cells_per_block(f);
  const int cy = blockIdx.x / nbx, cell0 = (blockIdx.x - cy * nbx) * cpb;
  const int ncell = min(cpb, res - cell0);
  const CoarseChannels<1> chs = {{ch}};
  coarsen_block<1, TGT, MODE>(lds, src, anc, mask, sol, W, H, C, chs, lambda, x0, y0, f, cy,
                              cell0, ncell, [&](int k, const CoarseSums<1> &s) {
    const size_t c = (size_t)cy * res + cell0 + k;
    coarsen_put(in4 + c * 4, s, 0, f);
    if (TGT) {
      const double nt = (double)s.n_int + s.n_ring;
      if (tgt) tgt[c] = nt > 0.0 ? (float)(s.t[0] / nt) : 0.0f;
      if (wt) wt[c] = s.n_int > 0 ? 1.0f : (s.n_ring > 0 ? 0.5f : 0.0f);
    }
  });
}

// One channel: grid res * coarsen_blocks_per_row(f, res) blocks of 256.
__global__ __launch_bounds__(256) void k_coarsen(const float *__restrict__ src,
                                                 const float *__restrict__ anc,
                                                 const float *__restrict__ mask,
                                                 const float *__restrict__ sol, int W, int H,
                                                 int C, int ch, double lambda, int x0, int y0,
                                                 int f, int res, int mode, float *__restrict__ in4,
                                                 float *__restrict__ tgt, float *__restrict__ wt) {
  __shared__ CoarseLds lds;
  if (sol && mode)
    coarsen_one<true, 1>(lds, src, anc, mask, sol, W, H, C, ch, lambda, x0, y0, f, res, in4, tgt, wt);
  else if (sol)
    coarsen_one<true, 0>(lds, src, anc, mask, sol, W, H, C, ch, lambda, x0, y0, f, res, in4, tgt, wt);
  else if (mode)
    coarsen_one<false, 1>(lds, src, anc, mask, nullptr, W, H, C, ch, lambda, x0, y0, f, res, in4,
                          nullptr, nullptr);
  else
    coarsen_one<false, 0>(lds, src, anc, mask, nullptr, W, H, C, ch, lambda, x0, y0, f, res, in4,
                          nullptr, nullptr);
}

// The warm-start inputs (mode 0) of every network instance of a batch in one
// launch, the channels of an image together: a group (span entry) is a run of
// up to 4 consecutive entries args[slot .. slot + nrows) of one image (item),
// so each pixel's channels and its interior and ring tests are read once for
// the group; its blocks are row0 .. row0 + res * coarsen_blocks_per_row(f, res) - 1.
// A channel's inputs equal k_coarsen's bit for bit (the same sums in the same
// order). Entry j's input goes to in4 + res*res*4*(net - j0).
template <int NE>
__device__ __forceinline__ void coarsen_group(CoarseLds &lds, const PhInst *__restrict__ insts,
                                              const PhCorrArgs *__restrict__ args, const PhSpan &Gs,
                                              const PhItemIO &G, double lambda, int res, int j0,
                                              float *__restrict__ in4) {
  const PhCorrArgs A0 = args[Gs.slot];
  const int f = A0.f;
  CoarseChannels<NE> ch;
  int net[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const PhCorrArgs A = args[Gs.slot + e];
    ch.c[e] = insts[A.slot].channel;
    net[e] = A.net;
  }
  const int nbx = coarsen_blocks_per_row(f, res), cpb = coarsen_cells_per_block(f);
  const int local = blockIdx.x - Gs.row0;
  const int cy = local / nbx, cell0 = (local - cy * nbx) * cpb;
  const int ncell = min(cpb, res - cell0);
  coarsen_block<NE, false, 0>(lds, G.src, G.tgt, G.mask, nullptr, G.W, G.H, G.C, ch, lambda,
                              A0.x0, A0.y0, f, cy, cell0, ncell,
                              [&](int k, const CoarseSums<NE> &s) {
    const size_t c = (size_t)cy * res + cell0 + k;
#pragma unroll
    for (int e = 0; e < NE; ++e)
      coarsen_put(in4 + ((size_t)(net[e] - j0) * res * res + c) * 4, s, e, f);
  });
}

__global__ __launch_bounds__(256) void k_coarsen_groups(const PhInst *__restrict__ insts,
                                                        const PhCorrArgs *__restrict__ args,
                                                        const PhSpan *__restrict__ groups,
                                                        int ngroups,
                                                        const PhItemIO *__restrict__ io,
                                                        double lambda, int res, int j0,
                                                        float *__restrict__ in4) {
  __shared__ CoarseLds lds;
  const PhSpan Gs = groups[ph_span_find(groups, ngroups, blockIdx.x)];
  const PhItemIO &G = io[Gs.item];
  switch (Gs.nrows) {  // the group's channels, uniform across the block
    case 1: coarsen_group<1>(lds, insts, args, Gs, G, lambda, res, j0, in4); break;
    case 2: coarsen_group<2>(lds, insts, args, Gs, G, lambda, res, j0, in4); break;
    case 3: coarsen_group<3>(lds, insts, args, Gs, G, lambda, res, j0, in4); break;
    default: coarsen_group<4>(lds, insts, args, Gs, G, lambda, res, j0, in4); break;
  }
}

// Net input of a coarse correction, from an instance's plane: per block the
// sum of the correction problem's rhs over interior cells, zero boundary
// values, interior fraction and ring density, exactly as k_coarsen lays them
// out. The state is x + e (x FP64 with g = -r(x) stored FP32, e the FP32
// correction in the buffer given); its residual is r = L e - g with
// L e = nsum(e) - (4 + lambda) e, and the correction problem for the error
// of x + e is L d = -r. Without e (WE false, or e = 0 after a refinement)
// the rhs is g itself, bit for bit. With WE every cell's r is the sweeps'
// FP32 residual (the per-sweep residual estimates use the same formula), and
// rsq receives the sum of r^2 of each coarse cell (their total is
// ||r(x + e)||^2: the grid covers every interior cell). Sums run in FP32 down
// a column's f rows and in FP64 from there on.
// Block (bx, cy) takes coarse cells cx0 = bx*cpb .. cx0+cpb-1 of coarse row cy.
// Each thread sums whole plane columns over the row's f plane rows (in row
// order; adjacent threads read adjacent columns), the column sums go to LDS,
// and one thread per coarse cell adds its f columns in column order, so the
// result is deterministic. Columns go through LDS in chunks of
// CP_MAX_COLS; a coarse cell wider than that (cpb = 1) adds chunk by chunk.
// (a block's columns, ncc * f, stay within 1024 for f <= 1024: see
// coarsen_plane_cpb; 16 KB of LDS a block, so six blocks fit a CU)
constexpr int CP_MAX_COLS = 1024;
// plane rows a thread takes per batch of loads
constexpr int CP_ROWS = 6;
template <bool WE>
__device__ __forceinline__ void coarsen_plane_block(const PhInst &I, const float *__restrict__ e,
                                                    int x0, int y0, int f, int res, int cpb,
                                                    float *__restrict__ in4,
                                                    double *__restrict__ rsq, int bx, int cy) {
  __shared__ float s_rhs[CP_MAX_COLS];
  __shared__ float s_sq[WE ? CP_MAX_COLS : 1];
  __shared__ int s_int[CP_MAX_COLS], s_ring[CP_MAX_COLS];
  const int cx0 = bx * cpb;
  const int ncc = min(cpb, res - cx0);
  const int ncols = ncc * f;
  const int pxb = x0 + cx0 * f - I.ox;  // plane column of the block's column 0
  const int pyb = y0 + cy * f - I.oy;
  const int ya = max(pyb, 1), yb = min(pyb + f, I.ph - 1);
  const float dgf = I.dg;
  double wr = 0.0, wi = 0.0, wn = 0.0, wq = 0.0;  // running sums of a wide cell (thread 0)
  for (int j0 = 0; j0 < ncols; j0 += CP_MAX_COLS) {
    const int nj = min(CP_MAX_COLS, ncols - j0);
    // plane columns [c0, c1) of this chunk, inside the plane's non-padding
    // columns, in groups of 4 aligned plane columns (one 4-byte load of cls
    // and one 16-byte load of g per row); the rows above and below give the
    // ring test, the bytes left and right of the group its edges
    const int c0 = max(pxb + j0, 1), c1 = min(pxb + j0 + nj, I.pw - 1);
    for (int j = threadIdx.x; j < nj; j += 256) {
      s_rhs[j] = 0.0f;
      if (WE) s_sq[j] = 0.0f;
      s_int[j] = 0;
      s_ring[j] = 0;
    }
    __syncthreads();
    const int ga = c0 & ~3;
    const size_t pitch = (size_t)I.pitch;
    const int lane = threadIdx.x & 63;
    // A wave's 64 groups walk the rows together (lanes past c1 idle along,
    // reading a fixed cell of row 1), so the values left and right of a group
    // come from the lanes beside it; the wave's edge lanes and the group
    // before c1 load them. Rows go in batches of CP_ROWS: the batch's g and e
    // loads (a row of the group without interior cells is not read: e is 0
    // there) are issued together with the next batch's classification, so a
    // batch waits for one round of loads, not one per row.
    for (int gb = ga + 4 * (int)(threadIdx.x & ~63); gb < c1; gb += 4 * 256) {
      const int gx0 = gb + 4 * lane;
      const bool act = gx0 < c1;
      const int gx = act ? gx0 : ga;
      const bool ld_l = lane == 0, ld_r = lane == 63 || gx0 + 4 >= c1;
      const size_t dummy = pitch + gx;  // row 1: in the plane, in cache
      // column sums over the coarse row's f plane rows, in row order (FP32:
      // at most f terms of one cell's size)
      float sr[4] = {0.f, 0.f, 0.f, 0.f}, sq[4] = {0.f, 0.f, 0.f, 0.f};
      int ni[4] = {0, 0, 0, 0}, nr[4] = {0, 0, 0, 0};
      // classification of rows py0 - 1 .. py0 + CP_ROWS of the batch at py0
      // (rows clamped into the plane; a clamped row is never used)
      unsigned cb[CP_ROWS + 2];
#pragma unroll
      for (int r = 0; r < CP_ROWS + 2; ++r)
        cb[r] = *reinterpret_cast<const unsigned *>(
            I.cls + (size_t)min(max(ya - 1 + r, 0), I.ph - 1) * pitch + gx);
      for (int py0 = ya; py0 < yb; py0 += CP_ROWS) {
        // the batch's g (rows py0 ..), e (rows py0 - 1 ..) and edges
        float4 gb4[CP_ROWS], eb4[CP_ROWS + 2];
        float elb[CP_ROWS], erb[CP_ROWS];
        unsigned char clb[CP_ROWS], crb[CP_ROWS];
#pragma unroll
        for (int r = 0; r < CP_ROWS + 2; ++r) eb4[r] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int r = 0; r < CP_ROWS; ++r) elb[r] = erb[r] = 0.f;
#pragma unroll
        for (int r = 0; r < CP_ROWS; ++r) {
          const int py = py0 + r;
          const size_t p = (size_t)min(py, I.ph - 2) * pitch + gx;
          const bool any = act && py < yb && cb[r + 1] != 0;
          gb4[r] = *reinterpret_cast<const float4 *>(I.g + (any ? p : dummy));
          clb[r] = ld_l ? I.cls[act ? p - 1 : dummy] : 0;
          crb[r] = ld_r ? I.cls[act ? p + 4 : dummy] : 0;
          if (WE) {
            elb[r] = ld_l ? e[any ? p - 1 : dummy] : 0.f;
            erb[r] = ld_r ? e[any ? p + 4 : dummy] : 0.f;
          }
        }
        if (WE) {
#pragma unroll
          for (int r = 0; r < CP_ROWS + 2; ++r) {
            // e is 0 off the interior: a row of the group without interior
            // cells reads as 0
            const size_t pr = (size_t)min(max(py0 - 1 + r, 0), I.ph - 1) * pitch + gx;
            const float4 v = *reinterpret_cast<const float4 *>(e + (cb[r] ? pr : dummy));
            const float z = cb[r] ? 1.0f : 0.0f;
            eb4[r] = make_float4(v.x * z, v.y * z, v.z * z, v.w * z);
          }
        }
        // the next batch's classification (its first two rows are this
        // batch's last two)
        unsigned cn[CP_ROWS + 2];
        cn[0] = cb[CP_ROWS];
        cn[1] = cb[CP_ROWS + 1];
#pragma unroll
        for (int r = 2; r < CP_ROWS + 2; ++r)
          cn[r] = *reinterpret_cast<const unsigned *>(
              I.cls + (size_t)min(py0 + CP_ROWS - 1 + r, I.ph - 1) * pitch + gx);
#pragma unroll
        for (int r = 0; r < CP_ROWS; ++r) {
          if (py0 + r >= yb) continue;  // uniform across the block
          const unsigned cu = cb[r], cc = cb[r + 1], cd = cb[r + 2];
          const float4 eu4 = eb4[r], ec4 = eb4[r + 1], ed4 = eb4[r + 2];
          const unsigned cl_n =
              (unsigned)__builtin_amdgcn_update_dpp(0, (int)cc, 0x138, 0xf, 0xf, false);
          const unsigned cr_n =
              (unsigned)__builtin_amdgcn_update_dpp(0, (int)cc, 0x130, 0xf, 0xf, false);
          const unsigned char cl = ld_l ? clb[r] : (unsigned char)(cl_n >> 24);
          const unsigned char cr = ld_r ? crb[r] : (unsigned char)(cr_n & 0xff);
          float el = 0.f, erx = 0.f;
          if (WE) {
            const float el_n = __builtin_bit_cast(
                float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, ec4.w), 0x138, 0xf,
                                                   0xf, false));
            const float er_n = __builtin_bit_cast(
                float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, ec4.x), 0x130, 0xf,
                                                   0xf, false));
            el = ld_l ? elb[r] : el_n;
            erx = ld_r ? erb[r] : er_n;
          }
          const uchar4 c4 = __builtin_bit_cast(uchar4, cc), u4 = __builtin_bit_cast(uchar4, cu),
                       d4 = __builtin_bit_cast(uchar4, cd);
          const unsigned char c[6] = {cl, c4.x, c4.y, c4.z, c4.w, cr};
          const unsigned char u[4] = {u4.x, u4.y, u4.z, u4.w}, d[4] = {d4.x, d4.y, d4.z, d4.w};
          const float gg[4] = {gb4[r].x, gb4[r].y, gb4[r].z, gb4[r].w};
          const float er[6] = {el, ec4.x, ec4.y, ec4.z, ec4.w, erx};
          const float eu[4] = {eu4.x, eu4.y, eu4.z, eu4.w}, ed[4] = {ed4.x, ed4.y, ed4.z, ed4.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (c[k + 1]) {
              ni[k] += 1;
              if (WE) {
                // the sweeps' own FP32 residual of the cell (r = -g exactly for e = 0)
                const float rr =
                    fmaf(-dgf, er[k + 1], (((er[k] + er[k + 2]) + eu[k]) + ed[k]) - gg[k]);
                sr[k] -= rr;
                sq[k] = fmaf(rr, rr, sq[k]);
              } else {
                sr[k] += gg[k];
              }
            } else if (c[k] | c[k + 2] | u[k] | d[k]) {
              nr[k] += 1;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < CP_ROWS + 2; ++r) cb[r] = cn[r];
      }
      if (!act) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int px = gx + k;
        if (px < c0 || px >= c1) continue;
        const int j = px - pxb - j0;
        s_rhs[j] = sr[k];
        if (WE) s_sq[j] = sq[k];
        s_int[j] = ni[k];
        s_ring[j] = nr[k];
      }
    }
    __syncthreads();
    if (f > CP_MAX_COLS) {
      if (threadIdx.x == 0)
        for (int j = 0; j < nj; ++j) {
          wr += s_rhs[j];
          if (WE) wq += s_sq[j];
          wi += s_int[j];
          wn += s_ring[j];
        }
    } else if ((int)threadIdx.x < ncc) {
      const int c = threadIdx.x;
      double sr = 0.0, sq = 0.0, ni = 0.0, nr = 0.0;
      for (int j = c * f; j < c * f + f; ++j) {
        sr += s_rhs[j];
        if (WE) sq += s_sq[j];
        ni += s_int[j];
        nr += s_ring[j];
      }
      float *o = in4 + ((size_t)cy * res + cx0 + c) * 4;
      o[0] = (float)sr;
      o[1] = 0.0f;
      o[2] = (float)(ni / ((double)f * f));
      o[3] = (float)(nr / f);
      if (WE) rsq[(size_t)cy * res + cx0 + c] = sq;
    }
    __syncthreads();
  }
  if (f > CP_MAX_COLS && threadIdx.x == 0) {
    float *o = in4 + ((size_t)cy * res + cx0) * 4;
    o[0] = (float)wr;
    o[1] = 0.0f;
    o[2] = (float)(wi / ((double)f * f));
    o[3] = (float)(wn / f);
    if (WE) rsq[(size_t)cy * res + cx0] = wq;
  }
}

// The input from g alone (the plane right after a refinement, e = 0).
__global__ __launch_bounds__(256) void k_coarsen_plane(PhInst I, int x0, int y0, int f, int res,
                                                       int cpb, float *__restrict__ in4) {
  coarsen_plane_block<false>(I, nullptr, x0, y0, f, res, cpb, in4, nullptr, blockIdx.x,
                             blockIdx.y);
}

// The input from the state x + e of a plane (e in the given buffer), with the
// per-coarse-cell sums of r(x + e)^2.
__global__ __launch_bounds__(256) void k_corr_input(PhInst I, const float *__restrict__ e, int x0,
                                                    int y0, int f, int res, int cpb,
                                                    float *__restrict__ in4,
                                                    double *__restrict__ rsq) {
  coarsen_plane_block<true>(I, e, x0, y0, f, res, cpb, in4, rsq, blockIdx.x, blockIdx.y);
}

// k_corr_input for every correction of a round in one launch: blockIdx.z =
// entry j, whose input goes to in4 + res*res*4*j and its squared-residual
// cells to rsq + res*res*j; its state is in the correction buffer its control
// block names. gridDim.x covers the entry with the most blocks per coarse row;
// blocks past an entry's row end exit.
__global__ __launch_bounds__(256) void k_corr_input_list(const PhInst *__restrict__ insts,
                                                         const PhCtl *__restrict__ ctl,
                                                         const PhCorrArgs *__restrict__ args,
                                                         int res, float *__restrict__ in4,
                                                         double *__restrict__ rsq) {
  const PhCorrArgs A = args[blockIdx.z];
  if ((int)blockIdx.x * A.cpb >= res) return;
  const PhInst &I = insts[A.slot];
  coarsen_plane_block<true>(I, I.e[ctl[A.slot].cur_buf], A.x0, A.y0, A.f, res, A.cpb,
                            in4 + (size_t)res * res * 4 * blockIdx.z,
                            rsq + (size_t)res * res * blockIdx.z, blockIdx.x, blockIdx.y);
}

// Coarse cells per block of k_coarsen_plane: about one 4-column group per
// thread (each thread's serial chain is then the f rows of its group), at most
// 256 cells (one finishing thread per cell). The result does not depend on it:
// a cell's columns are always added in column order by one thread.
int coarsen_plane_cpb(int f) { return std::max(1, std::min(256, 1024 / f)); }
int coarsen_blocks(int f, int res) { return res * coarsen_blocks_per_row(f, res); }

// The problems are linear: each net input is scaled so that its rhs and
// boundary channels have unit RMS over the grid (scale[b] = that RMS, 1 when
// both are zero), and the prediction is scaled back. Two kernels over
// (NORM_SLICES slices, inputs): per-slice sums of squares, then every slice
// adds the slices' sums in order (deterministic) and divides its cells.
constexpr int NORM_SLICES = PH_NET_SLICES;
__global__ __launch_bounds__(256) void k_in_norm_part(const float *__restrict__ in4, int res,
                                                      double *__restrict__ part) {
  const int sl = blockIdx.x, b = blockIdx.y;
  const int n = res * res, per = (n + NORM_SLICES - 1) / NORM_SLICES;
  const int i0 = sl * per, i1 = min(n, i0 + per);
  const float *v = in4 + (size_t)b * n * 4;
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const double x = v[4 * i], y = v[4 * i + 1];
    s += x * x + y * y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)b * NORM_SLICES + sl] = red[0];
}

__global__ __launch_bounds__(256) void k_in_norm_apply(float *__restrict__ in4, int res,
                                                       const double *__restrict__ part,
                                                       float *__restrict__ scale) {
  const int sl = blockIdx.x, b = blockIdx.y;
  const int n = res * res, per = (n + NORM_SLICES - 1) / NORM_SLICES;
  const int i0 = sl * per, i1 = min(n, i0 + per);
  double t = 0.0;
  for (int k = 0; k < NORM_SLICES; ++k) t += part[(size_t)b * NORM_SLICES + k];
  const double rms = sqrt(t / ((double)n));
  const float sc = rms > 0.0 ? (float)rms : 1.0f;
  float *v = in4 + (size_t)b * n * 4;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    v[4 * i] /= sc;
    v[4 * i + 1] /= sc;
  }
  if (sl == 0 && threadIdx.x == 0) scale[b] = sc;
}

__device__ __forceinline__ double warm_bilinear(const float *hc, int res, int x0, int y0, int f,
                                                int ix, int iy) {
  const double inv = 1.0 / f;
  double u = ((ix - x0) + 0.5) * inv - 0.5, v = ((iy - y0) + 0.5) * inv - 0.5;
  u = fmin(fmax(u, 0.0), (double)(res - 1));
  v = fmin(fmax(v, 0.0), (double)(res - 1));
  const int i0 = min((int)u, res - 1), j0 = min((int)v, res - 1);
  const int i1 = min(i0 + 1, res - 1), j1 = min(j0 + 1, res - 1);
  const double a = u - i0, b = v - j0;
  const double h00 = hc[(size_t)j0 * res + i0], h01 = hc[(size_t)j0 * res + i1];
  const double h10 = hc[(size_t)j1 * res + i0], h11 = hc[(size_t)j1 * res + i1];
  return (1.0 - b) * ((1.0 - a) * h00 + a * h01) + b * ((1.0 - a) * h10 + a * h11);
}

// Step length of a coarse correction, on the coarse grid. For block means E of
// a smooth error e, the block sums of Laplacian(e) are the net fluxes through
// the block faces, sum over the 4 neighbours of (E_n - E_b) (A_c E): the
// correction problem's coarse form is A_c E = G with G = the block sums of g
// (channel 0 of the net input). The Galerkin step alpha = <G, H> / <A_c H, H>
// for the prediction H is 1 for a prediction that solves it and scales a
// wrong-sized one back; it is taken over bulk blocks only (every cell
// interior, 4 neighbour blocks with interior cells), where A_c holds, and
// clamped to [0, max_step] (0: no usable correction). Input and prediction
// carry the same normalisation, so alpha is scale-free.
// Two kernels over (NORM_SLICES slices, corrections): per-slice sums of
// G H and (A_c H) H (and, given rsq, of the input's squared-residual cells),
// then one lane per correction adds the slices in order: alpha, and the norm
// ||r|| of the state the input was taken from.
__global__ __launch_bounds__(256) void k_corr_alpha_part(const float *__restrict__ in4_all,
                                                         const float *__restrict__ h_all,
                                                         const double *__restrict__ rsq_all,
                                                         int res, double *__restrict__ part) {
  const int sl = blockIdx.x, b = blockIdx.y;
  const int n = res * res, per = (n + NORM_SLICES - 1) / NORM_SLICES;
  const int i0 = sl * per, i1 = min(n, i0 + per);
  const float *in4 = in4_all + (size_t)b * n * 4;
  const float *hc = h_all + (size_t)b * n;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int c = i0 + threadIdx.x; c < i1; c += 256) {
    if (rsq_all) s3 += rsq_all[(size_t)b * n + c];
    const int cx = c % res, cy = c / res;
    if (cx < 1 || cy < 1 || cx >= res - 1 || cy >= res - 1) continue;
    if (!(in4[(size_t)c * 4 + 2] >= 1.0f)) continue;
    if (!(in4[(size_t)(c - 1) * 4 + 2] > 0.0f && in4[(size_t)(c + 1) * 4 + 2] > 0.0f &&
          in4[(size_t)(c - res) * 4 + 2] > 0.0f && in4[(size_t)(c + res) * 4 + 2] > 0.0f))
      continue;
    const double h = hc[c];
    const double ah = (((double)hc[c - 1] + hc[c + 1]) + hc[c - res]) + hc[c + res] - 4.0 * h;
    s1 = fma((double)in4[(size_t)c * 4], h, s1);
    s2 = fma(ah, h, s2);
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = s3;
  __syncthreads();
  for (int k = 128; k >= 1; k >>= 1) {
    if ((int)threadIdx.x < k) {
      red[0][threadIdx.x] += red[0][threadIdx.x + k];
      red[1][threadIdx.x] += red[1][threadIdx.x + k];
      red[2][threadIdx.x] += red[2][threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *o = part + ((size_t)b * NORM_SLICES + sl) * 3;
    o[0] = red[0][0];
    o[1] = red[1][0];
    o[2] = red[2][0];
  }
}

__global__ void k_corr_alpha(const double *__restrict__ part, int n, double max_step,
                             double *__restrict__ alpha, double *__restrict__ rnorm) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int k = 0; k < NORM_SLICES; ++k) {
    const double *o = part + ((size_t)b * NORM_SLICES + k) * 3;
    s1 += o[0];
    s2 += o[1];
    s3 += o[2];
  }
  const double a = s1 / s2;
  alpha[b] = (s2 < 0.0 && isfinite(a)) ? fmin(fmax(a, 0.0), max_step) : 0.0;
  if (rnorm) rnorm[b] = sqrt(s3);
}

// x = P + scale * bilinear(h) at the interior cells of every listed network
// instance, one block per plane row (span entry i takes args[i]; cell centres:
// image cell ix sits at coarse coordinate (ix - x0 + 0.5) / f - 0.5, clamped
// to the grid).
__global__ __launch_bounds__(256) void k_warm_apply_spans(const PhInst *__restrict__ insts,
                                                          const PhSpan *__restrict__ spans, int n,
                                                          const PhCorrArgs *__restrict__ args,
                                                          const PhItemIO *__restrict__ io,
                                                          const float *__restrict__ hc,
                                                          const float *__restrict__ scale,
                                                          int res) {
  const int b = blockIdx.x;
  const int i = ph_span_find(spans, n, b);
  const PhSpan S = spans[i];
  const PhCorrArgs A = args[i];
  const PhInst &I = insts[S.slot];
  const float *src = io[A.item].src;
  const float *h = hc + (size_t)res * res * A.net;
  const double sc = (double)scale[A.net];
  const int py = b - S.row0;
  const int iy = I.oy + py;
  for (int px = threadIdx.x; px < I.pw; px += 256) {
    const size_t p = (size_t)py * I.pitch + px;
    if (!I.cls[p]) continue;
    const int ix = I.ox + px;
    const double hv = sc * warm_bilinear(h, res, A.x0, A.y0, A.f, ix, iy);
    const size_t q = (size_t)iy * I.img_w + ix;
    I.x[p] = (double)src[q * I.channels + I.channel] + hv;
  }
}

// x = the caller's start field at the interior cells of an instance's plane.
__global__ void k_start_apply(PhInst I, const float *__restrict__ start) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y;
  if (px >= I.pw) return;
  const size_t p = (size_t)py * I.pitch + px;
  if (!I.cls[p]) return;
  const size_t q = (size_t)(I.oy + py) * I.img_w + (I.ox + px);
  I.x[p] = (double)start[q * I.channels + I.channel];
}

}  // namespace ph

using namespace ph;

extern "C" ph_status ph_warm_coarsen_device(const float *d_source, const float *d_anchor,
                                            const float *d_mask, const float *d_solution,
                                            uint32_t width, uint32_t height, uint32_t channels,
                                            uint32_t channel, double lambda, int32_t x0,
                                            int32_t y0, int32_t factor, int32_t res, int32_t mode,
                                            float *d_in, float *d_target, float *d_weight,
                                            void *stream, ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    ensure_device();
    if (!d_source || !d_anchor || !d_mask || !d_in)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    if (channels == 0 || channel >= channels || factor < 1 || res < 1 || res > 4096 ||
        (mode != 0 && mode != 1))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                 "coarsen: channel %u of %u, factor %d, resolution %d, mode %d", channel, channels,
                 factor, res, (int)mode);
    if (d_solution && (!d_target || !d_weight))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "coarsen: target/weight buffers required");
    hipLaunchKernelGGL(k_coarsen, dim3(res * coarsen_blocks_per_row(factor, res)), dim3(256), 0,
                       (hipStream_t)stream, d_source,
                       d_anchor, d_mask, d_solution, (int)width, (int)height, (int)channels,
                       (int)channel, lambda, x0, y0, factor, res, (int)mode, d_in,
                       d_solution ? d_target : nullptr, d_solution ? d_weight : nullptr);
    HIP_OK(hipGetLastError());
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return f.code == "E_GPU_UNAVAILABLE" ? PH_UNAVAILABLE : PH_ERROR;
  }
}

extern "C" ph_status ph_correction_input_device(const float *d_e, const float *d_g,
                                                const uint8_t *d_cls, int32_t pitch, int32_t pw,
                                                int32_t ph, int32_t ox, int32_t oy, int32_t x0,
                                                int32_t y0, int32_t factor, int32_t res,
                                                double lambda, float *d_in, double *d_rsq,
                                                void *stream, ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    ensure_device();
    if (!d_e || !d_g || !d_cls || !d_in || !d_rsq)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    if (pw < 3 || ph < 3 || pitch < pw || pitch % 4 != 0 || factor < 1 || res < 1 ||
        res > 4096 || (uintptr_t)d_g % 16 != 0 || (uintptr_t)d_e % 16 != 0 ||
        (uintptr_t)d_cls % 4 != 0 || !(lambda >= 0.0))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                 "correction_input: plane %dx%d pitch %d, factor %d, resolution %d (pitch a "
                 "multiple of 4, e and g 16-byte and cls 4-byte aligned, lambda >= 0)",
                 pw, ph, pitch, factor, res);
    PhInst I{};
    I.g = const_cast<float *>(d_g);
    I.cls = d_cls;
    I.pitch = pitch;
    I.pw = pw;
    I.ph = ph;
    I.ox = ox;
    I.oy = oy;
    I.dg64 = 4.0 + lambda;
    I.dg = (float)I.dg64;
    const int cpb = coarsen_plane_cpb(factor);
    hipLaunchKernelGGL(k_corr_input, dim3((res + cpb - 1) / cpb, res), dim3(256), 0,
                       (hipStream_t)stream, I, d_e, x0, y0, factor, res, cpb, d_in, d_rsq);
    HIP_OK(hipGetLastError());
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return f.code == "E_GPU_UNAVAILABLE" ? PH_UNAVAILABLE : PH_ERROR;
  }
}

extern "C" ph_status ph_coarsen_plane_device(const float *d_g, const uint8_t *d_cls, int32_t pitch,
                                             int32_t pw, int32_t ph, int32_t ox, int32_t oy,
                                             int32_t x0, int32_t y0, int32_t factor,
                                             int32_t res, float *d_in, void *stream,
                                             ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    ensure_device();
    if (!d_g || !d_cls || !d_in) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null buffer");
    if (pw < 3 || ph < 3 || pitch < pw || pitch % 4 != 0 || factor < 1 || res < 1 ||
        res > 4096 || (uintptr_t)d_g % 16 != 0 || (uintptr_t)d_cls % 4 != 0)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                 "coarsen_plane: plane %dx%d pitch %d, factor %d, resolution %d (pitch a "
                 "multiple of 4, g 16-byte and cls 4-byte aligned)",
                 pw, ph, pitch, factor, res);
    PhInst I{};
    I.g = const_cast<float *>(d_g);
    I.cls = d_cls;
    I.pitch = pitch;
    I.pw = pw;
    I.ph = ph;
    I.ox = ox;
    I.oy = oy;
    const int cpb = coarsen_plane_cpb(factor);
    hipLaunchKernelGGL(k_coarsen_plane, dim3((res + cpb - 1) / cpb, res), dim3(256), 0,
                       (hipStream_t)stream, I, x0, y0, factor, res, cpb, d_in);
    HIP_OK(hipGetLastError());
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return f.code == "E_GPU_UNAVAILABLE" ? PH_UNAVAILABLE : PH_ERROR;
  }
}
