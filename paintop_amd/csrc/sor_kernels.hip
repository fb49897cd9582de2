// sor_kernels.hip -- FP32 red-black SOR sweeps with temporal blocking in
// registers, and the per-instance stop-rule control that follows each launch.
//
// The equation relaxed is the correction equation of the FP64 refinement loop:
//     nsum(e) - (4 + lambda) e = g      at interior cells, e = 0 elsewhere,
// where g = -(residual of the FP64 iterate x). Its residual
//     r = nsum(e) - (4 + lambda) e - g
// is the residual of x + e, i.e. of the reference's equation
// (crates/paintop-cpu/src/poisson.rs:305-325), and one SOR update is
//     e <- e + omega/(4+lambda) * r          (poisson.rs:256-262 rearranged).
//
// Schedule. A wave owns a tile: 64 lanes x CPL adjacent columns, and streams
// down its rows. At step t it applies half-sweep s (s = 1..K, red for odd s)
// to row t - s + 1, in increasing s, so every update reads exactly the values
// a global red/black sweep would read; the K+2 live rows stay in registers
// (a ring indexed at compile time), left/right neighbours across lanes come
// from DPP wavefront shifts. Each row is read once and written once per K
// half-sweeps: 12 bytes per unknown per launch (e in, g in, e out). The tile
// loads a K-cell halo on every side; its valid core is written to the other
// ping-pong buffer, so tiles never see each other's writes within a launch and
// the result is the same for every tile shape.
//
// Memory access. Rows are loaded PF steps ahead of use and stored once done,
// through buffer resources rooted at the tile: a row past the tile reads the
// last row again, and a store outside the valid core (halo rows and lanes)
// takes an offset past the resource's extent, which the hardware drops. So
// every step issues the same memory instructions with no branch around them.
// (Behind a branch, a load or store leaves the compiler unsure how many are in
// flight at the join, and it then waits for all of them, the prefetched rows
// included: the 4096^2 RGBA launch went from 92.0 to 88.8 us.)
//
// Residuals. With r_pre the residual of a cell just before its update, the
// residual after the update is (1 - omega) r_pre exactly, and black cells do
// not move after their own update within a sweep. So, with omega_h the factor
// of half-sweep h,
//     ||r(after sweep j)||^2 = (1-omega_2j)^2 sum_black r_pre(2j)^2 + sum_red r_pre(2j+1)^2,
// available one half-sweep late, with no extra pass. Per-lane sums are FP32 in
// fixed row order, lanes are combined by a fixed butterfly in FP64, tiles by a
// fixed-order sum: reruns are bit-identical.
#include <math.h>

#include "ph_internal.h"

namespace ph {

// DPP wavefront shifts: lane l receives lane l-1 (shr) / lane l+1 (shl); the
// lanes shifted in from outside the wave receive 0.
__device__ __forceinline__ float from_left(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

// Fixed-order butterfly sum across the 64 lanes of a wave; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}


// Stop rules for one finished sweep m (poisson.rs:202-229), given its squared
// residual res2 and relative residual rel = sqrt(res2) / r0. Returns true when
// the solve must halt at m.
__device__ bool finalize_sweep(PhCtl &C, const PhInst &I, uint32_t m, double res2, double rel) {
  C.history[m - 1] = rel;
  C.sweeps_final = m;
  C.last_rel = rel;
  if (rel <= C.tol) {
    C.reason = PH_REASON_CONVERGED;
    C.halt_sweep = m;
    return true;
  }
  if (ph_in_ramp(I, C.ramp_from, m)) {
    // stall guard on hold
  } else if (rel < C.best * (1.0 - PH_STALL_IMPROVEMENT)) {
    C.best = rel;
    C.since = 0;
  } else {
    C.since += 1;
    if (C.since >= C.stall_window) {
      C.reason = PH_REASON_STALLED;
      C.halt_sweep = m;
      return true;
    }
    // Half a window without progress: the FP32 correction may have reached
    // its rounding floor (~ 8 ulp(|e|) / (2 - omega) in residual terms) before
    // the 1e3-fold trigger; an FP64 refinement restarts it from e = 0 so a
    // real stall, not the FP32 floor, is what the stall rule sees.
    if (C.since == (C.stall_window >> 1)) C.refine_pending = 1;
  }
  if (res2 <= C.refine_abs2) C.refine_pending = 1;
  return false;
}

// Squared residual of sweep S + j of a launch that starts at sweep S (A: the
// launch's K per-half-sweep sums; j = 0 is the previous launch's last sweep,
// whose red half came in this launch).
template <int K>
__device__ __forceinline__ double sweep_res2(const PhCtl &C, const PhInst &I, const double *A,
                                             int j) {
  if (j == 0) return C.prev_black + A[0];
  const double wb = ph_omega_at(I, C.ramp_from, 2 * (uint64_t)C.sweeps_done + 2 * j - 1);  // black half of S + j
  return (1.0 - wb) * (1.0 - wb) * A[2 * j - 1] + A[2 * j];
}

// Control for one instance after all of its tiles finished a launch. A holds
// the instance's residual sums of the launch's K half-sweeps, R2 / RL the
// squared and relative residual of each sweep (computed by parallel lanes);
// the rules run on one lane.
template <int K>
__device__ void control(PhCtl &C, const PhInst &I, const double *A, const double *R2,
                        const double *RL) {
  const uint32_t S = C.sweeps_done;
  const int k_run = min(K, (int)(2 * (C.sweep_limit - S)));
  const int nsw = k_run >> 1;
  C.launches += 1;
  if (C.mode == PH_MODE_REDO) {
    C.sweeps_done = S + nsw;
    C.cur_buf ^= 1u;
    C.halt = PH_HALT_REDO_DONE;
    return;
  }
  const uint32_t in_buf = C.cur_buf;
  const uint64_t h0 = 2 * (uint64_t)S;  // first half-sweep of this launch
  bool stop = false;
  if (S >= 1 && C.sweeps_final < S) stop = finalize_sweep(C, I, S, R2[0], RL[0]);
  for (int j = 1; j < nsw && !stop; ++j) stop = finalize_sweep(C, I, S + j, R2[j], RL[j]);
  const double wl = ph_omega_at(I, C.ramp_from, h0 + 2 * nsw - 1);
  C.prev_black = (1.0 - wl) * (1.0 - wl) * A[2 * nsw - 1];
  if (stop) {
    C.halt_base = S;
    C.halt_buf = in_buf;
    C.halt = PH_HALT_STOP;
    return;
  }
  C.sweeps_done = S + nsw;
  C.cur_buf = in_buf ^ 1u;
  if (C.sweeps_done - C.last_refine >= C.refine_every) C.refine_pending = 1;
  if (C.sweeps_done >= C.sweep_limit) C.halt = PH_HALT_LIMIT;
  else if (C.refine_pending) C.halt = PH_HALT_REFINE;
}

// One SOR update of cell j of a lane's row, given its four neighbours.
template <bool MASK>
__device__ __forceinline__ float relax(float &e, float g, float l, float r, float u, float d,
                                       float c, float dg) {
  const float rp = fmaf(-dg, e, (((l + r) + u) + d) - g);
  if constexpr (!MASK) {  // a tile whose cells are all interior
    e = fmaf(c, rp, e);
    return rp;
  }
  // g is NaN at non-interior cells, so rp is NaN exactly there (e and its
  // neighbours are finite). Testing rp, which is new at every stage, keeps the
  // compare next to its use: a test of g would be hoisted out of the row's
  // stages, and the lane masks of a ring of rows overflow the scalar registers.
  const float rr = (rp == rp) ? rp : 0.0f;
  e = fmaf(c, rr, e);
  return rr;
}

typedef float v4f __attribute__((ext_vector_type(4)));

// Buffer-resource access to a tile: loads and stores at an offset past the
// resource's extent return 0 / are dropped, so every step of a period issues
// the same memory instructions with no branch around them (a conditional load
// or store makes the compiler wait for every outstanding load at the join,
// which would defeat the row prefetch).
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
// An offset at or past 2^30 is outside every tile's resource (a tile spans
// at most 1056 rows of at most 2^17 bytes); the sum of two stays below 2^31.
constexpr unsigned PH_OOB_ROW = 1u << 30;
// Fourth word of a raw buffer resource (32-bit data format, no swizzle).
constexpr int PH_BUF_FLAGS = 0x00020000;
__device__ __forceinline__ v4f buf_load4(__amdgpu_buffer_rsrc_t rs, int off) {
  return __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}
__device__ __forceinline__ void buf_store4(__amdgpu_buffer_rsrc_t rs, int off, v4f v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, v), rs, off, 0, 0);
}
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t rs, int off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), rs, off, 0, 0);
}

// The ring's lead: the first period starts at step -M, so that the periods
// holding the steps t < 2K - 1 end exactly at t = 2K - 2 and every later
// period is a whole ring of steps (ring slot of row r: (r + M) mod U).
template <int K, int PF>
constexpr int ring_lead() {
  constexpr int U = K + PF + 1;
  return (U - (2 * K - 1) % U) % U;
}

// One ring period of U row steps of k_sor from step t0 (t0 = -M mod U),
// every step unconditional. At step t the wave applies half-sweep s (s =
// 1..K, red for odd s) to row t - s + 1, in increasing s. HEAD: the period
// holds steps t < 2K - 1 (t0 < 0 for the first), the only ones with updates
// outside the dependence cone (rows i < s), which are skipped (the test is
// uniform across the wave, and the skipped updates are closed under
// dependence, so no core value or residual sum changes). Rows before the
// tile read as 0 (their offsets fall outside the buffer); rows past it read
// its last row again; rows outside the valid core are not stored; when hb is
// not a multiple of U, steps past the last useful one touch only rows
// outside the core's cone.
template <int K, int PF, bool HEAD, bool MASK>
__device__ __forceinline__ void sor_period(int t0, int nrows, int hb, int rowb, int loff,
                                             int soff, float (&E)[K + PF + 1][4],
                                             float (&G)[K + PF + 1][4], float (&acc)[K],
                                             const float (&cs)[K], const float (&cw)[4], float dg,
                                             __amdgpu_buffer_rsrc_t rse,
                                             __amdgpu_buffer_rsrc_t rsg,
                                             __amdgpu_buffer_rsrc_t rso) {
  constexpr int U = K + PF + 1;
  constexpr int M = ring_lead<K, PF>();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = t0 + u;
    {
      const int sl = (u + PF) % U;  // row t+PF replaces row t+PF-U (< t-K, dead)
      // the row part of the offset is uniform (scalar); a row past the tile
      // reads the last row again, whose values are never used; a row before
      // it has a negative offset, outside the buffer as an unsigned offset
      const int off = loff + min(t + PF, nrows - 1) * rowb;
      const v4f ev = buf_load4(rse, off);
      const v4f gv = buf_load4(rsg, off);
      E[sl][0] = ev.x; E[sl][1] = ev.y; E[sl][2] = ev.z; E[sl][3] = ev.w;
      G[sl][0] = gv.x; G[sl][1] = gv.y; G[sl][2] = gv.z; G[sl][3] = gv.w;
    }
#pragma unroll
    for (int s = 1; s <= K; ++s) {
      if (HEAD && t - s + 1 < s) continue;
      const int si = ((u - s + 1) % U + U) % U;
      const int su = ((u - s) % U + U) % U;
      const int sd = ((u - s + 2) % U + U) % U;
      const int colour = (s - 1) & 1;
      const int j0 = (colour + (u - s + 1 + M + 2 * U)) & 1;  // parity of row t - s + 1
      const int i = t - s + 1;
      float *Er = E[si];
      const float *Eu = E[su];
      const float *Ed = E[sd];
      const float *Gr = G[si];
      const float c = cs[s - 1];
      float ra, rb;
      if (j0 == 0) {
        const float l0 = from_left(Er[3]);
        ra = relax<MASK>(Er[0], Gr[0], l0, Er[1], Eu[0], Ed[0], c, dg);
        rb = relax<MASK>(Er[2], Gr[2], Er[1], Er[3], Eu[2], Ed[2], c, dg);
      } else {
        const float r3n = from_right(Er[0]);
        ra = relax<MASK>(Er[1], Gr[1], Er[0], Er[2], Eu[1], Ed[1], c, dg);
        rb = relax<MASK>(Er[3], Gr[3], Er[2], r3n, Eu[3], Ed[3], c, dg);
      }
      if (i >= K && i < K + hb) {  // uniform across the wave (becomes a select)
        if constexpr (K % 4 == 0) {
          acc[s - 1] = fmaf(ra, ra, acc[s - 1]);
          acc[s - 1] = fmaf(rb, rb, acc[s - 1]);
        } else {
          acc[s - 1] = fmaf(ra * cw[j0], ra, acc[s - 1]);
          acc[s - 1] = fmaf(rb * cw[j0 + 2], rb, acc[s - 1]);
        }
      }
    }
    // row t-K+1 has had all of its half-sweeps: store its valid core. soff is
    // the lane's offset, or past the extent for a lane outside the core; rows
    // outside the core add a scalar offset past the extent as well.
    const int w = t - K + 1;
    const int sw = ((u - K + 1) % U + U) % U;
    // (unsigned: two out-of-extent parts sum to 2^31, still past the extent)
    const unsigned wrow = (w >= K && w < K + hb) ? (unsigned)(w * rowb) : PH_OOB_ROW;
    if constexpr (K % 4 == 0) {
      buf_store4(rso, (int)((unsigned)soff + wrow), v4f{E[sw][0], E[sw][1], E[sw][2], E[sw][3]});
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        buf_store1(rso, (int)((cw[j] != 0.0f ? (unsigned)(loff + 4 * j) : PH_OOB_ROW) + wrow),
                   E[sw][j]);
    }
  }
}

// PF: rows loaded ahead of use (odd, so the ring size stays even and the
// red/black colour of a ring slot is fixed at compile time).
// One tile's pass of K half-sweeps (MASK: the tile holds non-interior cells,
// whose updates are masked out; a tile of interior cells only needs no mask,
// and its result is the same).
template <int K, int PF, bool MASK>
__device__ __forceinline__ void sor_tile(const PhTile &T, const PhInst &I, const PhCtl &C,
                                         int lane, int slot, int ntiles,
                                         double *__restrict__ partials, int hb) {
  constexpr int CPL = 4;
  constexpr int U = K + PF + 1;
  const int k_run = min(K, (int)(2 * (C.sweep_limit - C.sweeps_done)));
  const uint32_t cur = C.cur_buf;
  const size_t pitch = (size_t)I.pitch;
  const size_t tile0 = (size_t)T.py0 * pitch + T.px0;
  const float dg = I.dg;
  // omega/(4+lambda) of each half-sweep of this launch; 0 past k_run, which
  // leaves e unchanged (the control reads only the first k_run residual sums)
  float cs[K];
  {
    const uint64_t h0 = 2 * (uint64_t)C.sweeps_done;
#pragma unroll
    for (int s = 0; s < K; ++s) cs[s] = s < k_run ? (float)(ph_omega_at(I, C.ramp_from, h0 + s) / I.dg64) : 0.0f;
  }
  // columns of the valid core, as 0/1 weights of the residual sums
  float cw[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j)
    cw[j] = (lane * CPL + j >= K) && (lane * CPL + j < 64 * CPL - K) ? 1.0f : 0.0f;
  const bool lane_full = cw[0] != 0.0f && cw[CPL - 1] != 0.0f;

  float E[U][CPL], G[U][CPL];
  float acc[K];
#pragma unroll
  for (int s = 0; s < K; ++s) acc[s] = 0.0f;
#pragma unroll
  for (int r = 0; r < U; ++r)
#pragma unroll
    for (int j = 0; j < CPL; ++j) E[r][j] = G[r][j] = 0.0f;

  // The tile's rows through buffer resources rooted at its origin: every step
  // of a period issues the same loads and stores, none behind a branch.
  const int nrows = hb + 2 * K;
  const int rowb = (int)(pitch * 4);
  const int loff = lane * CPL * 4;
  const int soff = lane_full ? loff : (int)PH_OOB_ROW;  // the lane's store offset (K % 4 == 0)
  const int nbytes = nrows * rowb;
  const __amdgpu_buffer_rsrc_t rse =
      __builtin_amdgcn_make_buffer_rsrc(I.e[cur] + tile0, 0, nbytes, PH_BUF_FLAGS);
  const __amdgpu_buffer_rsrc_t rsg =
      __builtin_amdgcn_make_buffer_rsrc(I.g + tile0, 0, nbytes, PH_BUF_FLAGS);
  const __amdgpu_buffer_rsrc_t rso =
      __builtin_amdgcn_make_buffer_rsrc(I.e[cur ^ 1u] + tile0, 0, nbytes, PH_BUF_FLAGS);
  // Steps t = 0 .. nrows - 2 carry every update the core needs. The head
  // periods (from step -M) end at t = 2K - 2; whole periods follow, exactly
  // hb / U of them when U divides hb. Rows the first period's steps do not
  // load (r < PF - M) are loaded here.
  constexpr int M = ring_lead<K, PF>();
#pragma unroll
  for (int r = 0; r < PF - M; ++r) {
    const v4f ev = buf_load4(rse, loff + r * rowb);
    const v4f gv = buf_load4(rsg, loff + r * rowb);
    const int sl = (r + M) % U;
    E[sl][0] = ev.x; E[sl][1] = ev.y; E[sl][2] = ev.z; E[sl][3] = ev.w;
    G[sl][0] = gv.x; G[sl][1] = gv.y; G[sl][2] = gv.z; G[sl][3] = gv.w;
  }
  const int nsteps = nrows - 1;
  int t0 = -M;
  for (; t0 < 2 * K - 1; t0 += U)
    sor_period<K, PF, true, MASK>(t0, nrows, hb, rowb, loff, soff, E, G, acc, cs, cw, dg, rse, rsg, rso);
  for (; t0 < nsteps; t0 += U)
    sor_period<K, PF, false, MASK>(t0, nrows, hb, rowb, loff, soff, E, G, acc, cs, cw, dg, rse, rsg, rso);
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const float a = (K % 4 == 0) ? (lane_full ? acc[s] : 0.0f) : acc[s];
    const double tot = wave_sum((double)a);
    if (lane == 0) partials[(size_t)s * ntiles + slot] = tot;  // stage-major: coalesced sums
  }
}

template <int K, int PF>
__global__ __launch_bounds__(64) void k_sor(const PhTile *__restrict__ tiles, int ntiles,
                                            const PhInst *__restrict__ insts,
                                            const PhCtl *__restrict__ ctl,
                                            double *__restrict__ partials, int hb) {
  static_assert(((K + PF + 1) & 1) == 0, "ring size must be even");
  const int lane = threadIdx.x;
  const int b = xcd_tile(blockIdx.x, ntiles);
  const PhTile T = tiles[b];
  const PhInst &I = insts[T.inst];
  const PhCtl &C = ctl[T.inst];
  if (C.halt != PH_RUN) return;
  if (T.flags & PH_TILE_INTERIOR)
    sor_tile<K, PF, false>(T, I, C, lane, T.flags >> 1, ntiles, partials, hb);
  else
    sor_tile<K, PF, true>(T, I, C, lane, T.flags >> 1, ntiles, partials, hb);
}

// After the sweep launch (the kernel boundary orders it after every tile):
// one block of K waves per instance; wave s adds the tiles' sums of
// half-sweep s in a fixed order, then one lane runs the stop rules.
template <int K>
__global__ __launch_bounds__(64 * K) void k_sor_control(const PhInst *__restrict__ insts,
                                                        PhCtl *ctl,
                                                        const double *__restrict__ partials,
                                                        int ninst, int ntiles) {
  __shared__ double A[K];
  __shared__ double R2[K / 2], RL[K / 2];
  // (the grid has one block per instance: blockIdx.x < ninst)
  const int a = blockIdx.x;
  PhCtl &C = ctl[a];
  const PhInst &I = insts[a];
  // the halt state and the instance's tile range in one memory round trip:
  // one test on all of them keeps both loads ahead of the branch
  const uint32_t halt = C.halt;
  const int tb = I.tile_begin, te = I.tile_end;
  if (a >= ninst || halt != PH_RUN || tb > te) return;  // uniform across the block
  const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // The rules run on copies of the control block and the instance (a chain of
  // dependent global loads and stores would dominate this kernel otherwise).
  // Wave 0 loads them, one dword per lane, together with its tiles' sums, so
  // the whole block waits for one memory round trip; they meet in LDS.
  constexpr int NC = (int)(sizeof(PhCtl) / 4), NI = (int)(sizeof(PhInst) / 4);
  static_assert(NC <= 64 && NI <= 64 && sizeof(PhCtl) % 4 == 0 && sizeof(PhInst) % 4 == 0,
                "one dword per lane");
  __shared__ __align__(16) uint32_t cpy_c[NC];
  __shared__ __align__(16) uint32_t cpy_i[NI];
  uint32_t wc = 0, wi = 0;
  if (s == 0) {
    if (lane < NC) wc = reinterpret_cast<const uint32_t *>(&C)[lane];
    if (lane < NI) wi = reinterpret_cast<const uint32_t *>(&I)[lane];
  }
  // wave s adds the tiles' sums of half-sweep s in tile order: each lane loads
  // its tiles in batches of 16 (one memory latency for up to 1024 tiles), then
  // adds them in order (a past-the-end slot adds 0, which leaves a sum of
  // squares unchanged)
  double v = 0.0;
  const double *ps = partials + (size_t)s * ntiles;
  constexpr int B = 16;
  for (int t = tb + lane; t < te; t += 64 * B) {
    double q[B];
#pragma unroll
    for (int k = 0; k < B; ++k) q[k] = t + 64 * k < te ? ps[t + 64 * k] : 0.0;
#pragma unroll
    for (int k = 0; k < B; ++k) v += q[k];
  }
  v = wave_sum(v);
  if (lane == 0) A[s] = v;
  if (s == 0) {
    if (lane < NC) cpy_c[lane] = wc;
    if (lane < NI) cpy_i[lane] = wi;
  }
  __syncthreads();
  const PhCtl &Cs = *reinterpret_cast<const PhCtl *>(cpy_c);
  const PhInst &Is = *reinterpret_cast<const PhInst *>(cpy_i);
  // the sweeps' residuals in parallel (lane j: sweep S + j)
  if (threadIdx.x < K / 2) {
    const int j = threadIdx.x;
    const double r2 = sweep_res2<K>(Cs, Is, A, j);
    R2[j] = r2;
    RL[j] = sqrt(r2) / Cs.r0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    PhCtl c = Cs;
    PhInst in = Is;
    control<K>(c, in, A, R2, RL);
    C = c;
  }
}

#define PH_INSTANTIATE(K, PF)                                                                \
  template __global__ void k_sor<K, PF>(const PhTile *, int, const PhInst *, const PhCtl *,  \
                                        double *, int);
PH_INSTANTIATE(2, 1)
PH_INSTANTIATE(4, 1)
PH_INSTANTIATE(8, 1)
PH_INSTANTIATE(12, 1)
PH_INSTANTIATE(16, 1)
PH_INSTANTIATE(2, 3)
PH_INSTANTIATE(4, 3)
PH_INSTANTIATE(8, 3)
PH_INSTANTIATE(12, 3)
PH_INSTANTIATE(16, 3)
PH_INSTANTIATE(8, 5)
PH_INSTANTIATE(16, 5)
template __global__ void k_sor_control<2>(const PhInst *, PhCtl *, const double *, int, int);
template __global__ void k_sor_control<4>(const PhInst *, PhCtl *, const double *, int, int);
template __global__ void k_sor_control<8>(const PhInst *, PhCtl *, const double *, int, int);
template __global__ void k_sor_control<12>(const PhInst *, PhCtl *, const double *, int, int);
template __global__ void k_sor_control<16>(const PhInst *, PhCtl *, const double *, int, int);

}  // namespace ph
