// ph_internal.h -- device-side data layout shared by the solver kernels and the
// host orchestration of libpaintop_hip.
//
// One solve instance = one channel of one image region. Its unknowns live in
// a padded rectangle ("plane") that covers the bounding box of the interior
// cells plus the one-cell Dirichlet ring, extended by PH_PAD cells of inert
// padding on every side so that the halo reads of the sweep tiles never leave
// the allocation. Plane cell (px, py) is image cell (ox + px, oy + py); the
// plane origin is chosen with (ox + oy) even, so the red/black colour of a
// plane cell is (px + py) & 1 and equals the colour of its image cell
// (red = even).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define PH_PAD 16          // >= the largest half-sweep count per launch
#define PH_MAX_K 16
#define PH_STALL_WINDOW 16u       // poisson.rs:81 (the window for a fixed omega)
#define PH_STALL_IMPROVEMENT 1e-12 // poisson.rs:85

// Per-instance descriptor (read-only for the sweep kernels).
struct PhInst {
  float *e[2];      // FP32 correction, ping-pong (plane layout, pitch floats per row)
  float *g;         // FP32 right-hand side of the correction equation; NaN = not interior
  double *x;        // FP64 solution (plane layout)
  double *bp;       // FP64 b' = b - lambda*anchor at interior cells, 0 elsewhere
  const uint8_t *cls; // 1 = interior (plane layout, pitch bytes per row)
  int32_t pitch;    // elements per plane row
  int32_t pw, ph;   // plane width / height in cells
  int32_t ox, oy;   // image coordinates of plane cell (0, 0)
  int32_t img_w, img_h, channels, channel; // source image geometry for gather/scatter
  double omega;        // SOR factor (after the ramp)
  double ramp_log;     // ln(2 - omega) / ramp_sweeps
  uint32_t ramp_sweeps; // sweeps of the warm-start ramp (0: none; it starts at PhCtl::ramp_from)
  float dg;         // 4 + lambda (FP32)
  double dg64;      // 4 + lambda (FP64)
  double lambda;
  int32_t tile_begin, tile_end; // this instance's tiles in the tile list
};

// Halt states of the per-instance control block.
enum : uint32_t {
  PH_RUN = 0,
  PH_HALT_STOP = 1,     // a stop rule fired at sweep halt_sweep (estimate)
  PH_HALT_REFINE = 2,   // the inner residual fell below the refinement threshold
  PH_HALT_LIMIT = 3,    // sweeps_done reached sweep_limit
  PH_HALT_REDO_DONE = 4,
  PH_HALT_DONE = 5,     // solve finished (host side)
};
enum : uint32_t { PH_MODE_NORMAL = 0, PH_MODE_REDO = 1 };
enum : uint32_t { PH_REASON_CONVERGED = 0, PH_REASON_MAXIT = 1, PH_REASON_STALLED = 2 };

// Per-instance control block: the reference's stop rules (poisson.rs:202-229)
// run here, on the device, once per sweep launch.
struct PhCtl {
  double r0;          // ||r(anchor start)|| in FP64: the relative-residual denominator
  double tol;
  double refine_abs2; // refine when the squared inner residual falls to this
  double best;        // best relative residual so far (stall guard)
  double last_rel;    // relative residual of sweep sweeps_final
  double prev_black;  // (1 - omega)^2 x black pre-update residual sum, last half-sweep of the previous launch
  double *history;    // relative residual per sweep, capacity max_it
  uint32_t max_it;
  uint32_t sweep_limit;
  uint32_t sweeps_done;  // sweeps applied to the state in e[cur_buf]
  uint32_t sweeps_final; // sweeps whose residual is in history
  uint32_t since;
  uint32_t halt;
  uint32_t reason;
  uint32_t halt_sweep;
  uint32_t halt_base;  // sweeps_done at the start of the launch that halted
  uint32_t halt_buf;   // its input buffer
  uint32_t cur_buf;
  uint32_t mode;
  uint32_t refine_pending;
  uint32_t launches;
  uint32_t stall_window; // sweeps without a new best that end the solve as stalled
  uint32_t refine_every; // refine at least this often (sweeps): bounds |e| and the FP32 floor
  uint32_t last_refine;  // sweep count of the last refinement
  uint32_t ramp_from;    // first sweep of the current omega ramp (it restarts after a coarse correction)
};

// SOR factor of half-sweep h (0-based). Both half-sweeps of sweep k = h/2
// use one factor. After a warm start or a coarse correction (at sweep
// ramp_from), sweep k in [ramp_from, ramp_from + ramp_sweeps) uses
// 2 - (2 - omega)^((k - ramp_from) / ramp_sweeps): Gauss-Seidel first (the
// high-frequency error a coarse prediction leaves decays fastest there), then
// omega.
__host__ __device__ inline double ph_omega_at(const PhInst &I, uint32_t ramp_from, uint64_t h) {
  const uint64_t k = h >> 1;
  if (k < ramp_from || k - ramp_from >= I.ramp_sweeps) return I.omega;
  return 2.0 - exp((double)(k - ramp_from) * I.ramp_log);
}

// Sweep m (1-based) runs inside an omega ramp. The residual of a ramp is not
// monotone (it dips while omega is near 1, then rises as omega grows and the
// smooth error shows), so the stall guard holds still during one.
__host__ __device__ inline bool ph_in_ramp(const PhInst &I, uint32_t ramp_from, uint32_t m) {
  const uint32_t k = m - 1;
  return I.ramp_sweeps != 0 && k >= ramp_from && k - ramp_from < I.ramp_sweeps;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of work items (tiles, rows) so neighbours, which share
// halo rows, hit the same L2. A bijection of [0, n) for b < n; XCD x gets the
// items [xcd_range_begin(x, n), xcd_range_begin(x + 1, n)), in block order.
__host__ __device__ inline int xcd_range_begin(int x, int n) {
  const int per = (n + 7) >> 3;
  const int full = n - (per - 1) * 8;
  return x <= full ? x * per : full * per + (x - full) * (per - 1);
}
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int per = (n + 7) >> 3;
  const int xcd = b & 7, slot = b >> 3;
  const int full = n - (per - 1) * 8;  // XCDs that get `per` items
  int t;
  if (xcd < full) t = xcd * per + slot;
  else t = full * per + (xcd - full) * (per - 1) + slot;
  return t;
}

// Row of block b in a launch of one block per plane row (n blocks) whose rows
// read the rows above and below: groups of 8 consecutive rows go to one XCD,
// the groups dealt over the XCDs in turn, so neighbouring rows are mostly in
// the L2 that read them while every XCD gets rows from all over the planes
// (contiguous bands per XCD leave some XCDs most of a mask). Rows past the
// last whole 64 keep their own block. A bijection of [0, n).
__device__ __forceinline__ int xcd_row_group(int b, int n) {
  if (b >= (n & ~63)) return b;
  const int slot = b >> 3;
  return (((slot >> 3) << 3) + (b & 7)) * 8 + (slot & 7);
}

// Slices of the two-stage reductions over a warm-start net input (input
// normalisation, correction step).
constexpr int PH_NET_SLICES = 16;

// Per-image geometry for the set-up kernels that serve every image of a call
// in one launch (tile occupancy, tuner coarse masks).
// Blocks per coarse row of the SOR-factor tuner's coarse-mask launch: one per
// started 1024 image columns of the row, at most 4 (a narrow row is one
// block's work).
constexpr int PH_COARSE_PARTS = 4;
__host__ __device__ inline int ph_coarse_parts(int cw, int cf) {
  const int p = (cw * cf + 1023) / 1024;
  return p < 1 ? 1 : (p > PH_COARSE_PARTS ? PH_COARSE_PARTS : p);
}

struct PhItemGeo {
  const uint8_t *cls;   // image classification, w x h
  int32_t w, h;
  int32_t ox, oy;       // plane origin in image coordinates
  int32_t rx0, ry0;     // region origin (interior box plus ring)
  int32_t cf, cw, ch;   // tuner coarse grid: factor and size (0 x 0: none)
  int64_t coarse_off;   // offset of the image's coarse mask (bytes)
};

// One entry of a row-batched launch (many instances or images in one grid,
// one block per row): block b with row0 <= b < row0 + nrows works on row
// b - row0 of entry `slot`. row0 is the running sum of nrows over the table.
struct PhSpan {
  int32_t slot;   // instance (or image) index in the launch's descriptor array
  int32_t row0;   // first block of the entry; also its first row-sum index
  int32_t nrows;
  int32_t buf;    // correction buffer (x += e[buf])
  int32_t out;    // slot of the entry's norm in the output array
  int32_t item;   // image index (assembly, write-back)
  double *hist;   // when set: receives ||r|| / r0, the residual-history entry
  double r0;
};

// Per-image buffers of the row-batched set-up and write-back kernels.
struct PhItemIO {
  const float *src, *tgt, *mask;
  float *out;
  uint8_t *cls_img;   // image classification, W x H
  uint8_t *cls_plane; // plane classification (shared by the image's channels)
  int32_t W, H, C, pad;
};

// One residual history to gather for the readback.
struct PhHistCopy {
  const double *src;
  int64_t off;  // first element in the gathered array
  int32_t n, pad;
};

// One network instance of a batched warm start or correction round: entry j
// of the round's list (the same order as its span table).
struct PhCorrArgs {
  int32_t slot;       // instance index in the descriptor array
  int32_t x0, y0, f;  // coarse-grid framing (image cell of coarse cell (0, 0), cell size)
  int32_t cpb;        // coarse cells per block of the input coarsening
  int32_t net;        // network batch index: prediction h + res*res*net, scale[net], alpha[net]
  int32_t item, pad;  // image index (the warm start reads the image)
};

// Largest coarse-grid side a correction kernel stages in LDS.
constexpr int PH_CORR_RES_MAX = 256;

// A value for one device address (batched small updates).
struct PhValue {
  double *dst;
  double val;
};

// Entry of a span table that holds block b: the last entry with row0 <= b.
__device__ __forceinline__ int ph_span_find(const PhSpan *__restrict__ s, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s[mid].row0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct PhTile {
  int32_t inst;
  int32_t px0; // first loaded plane column (valid columns start at px0 + K)
  int32_t py0; // first loaded plane row (valid rows start at py0 + K)
  // bit 0 PH_TILE_INTERIOR: every cell the tile loads is interior; bits 1..:
  // the tile's slot in the residual partial sums (its index in instance order)
  int32_t flags;
};
#define PH_TILE_INTERIOR 1
