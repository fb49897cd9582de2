// unet_kernels.hip -- the warm-start network's kernels: BF16 3x3 convolutions
// on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation) for
// the forward pass and both gradients, plus the small elementwise kernels
// around them (average pooling, gradient routing through skip/upsample
// connections, the FP32 1x1 output head, the loss, Adam).
//
// Activations are NHWC bf16: pixel-major, channels innermost, so the four
// consecutive channels one MFMA lane needs are one 8-byte load.
//
// Convolution as a GEMM with the output transposed (out^T = W * patches^T):
//   A = weights   [cout][kp]   k = tap*cin + ci, zero padded to kp (multiple of 16)
//   B = patches   [kp][pixels] gathered on the fly (SAME zero padding)
//   C = out^T     [cout][pixels]
// v_mfma_f32_16x16x16_bf16 operand layout (lane l, i = 0..3):
//   A[m = l%16][k = 4*(l/16)+i], B[k = 4*(l/16)+i][n = l%16], C[m = 4*(l/16)+i][n = l%16]
// so a lane holds four consecutive output channels of one pixel: one 8-byte store.
// The forward / input-gradient convolutions take K in chunks of 32 with gfx950's
// v_mfma_f32_16x16x32_bf16 (lane l holds k = 8*(l/16)+i, i = 0..7, of A and B:
// one 16-byte load each; C as above), and a 16-wide chunk where cin % 32 == 16.
//
// The input of a decoder convolution is concat(skip, up2(coarser)) with a
// nearest-neighbour x2 upsample; both are read in place (channel ci < ca from
// the skip tensor, the rest from the coarser tensor at (y/2, x/2)), so neither
// the upsampled tensor nor the concatenation is ever stored.
//
// Every reduction is deterministic: weight gradients are per-chunk partial sums
// (fixed pixel chunks) added in chunk order by a second kernel.
#include <stdint.h>

#include "ph_internal.h"
#include "unet_internal.h"

namespace ph {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const bf16x4 gcbf4;
typedef __attribute__((address_space(1))) bf16x4 gbf4;

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ float bf2f(short v) { return bf2f((uint16_t)v); }

// round to nearest even (finite inputs; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) { return *(gcbf4 *)p; }
__device__ __forceinline__ void st4(uint16_t *p, bf16x4 v) { *(gbf4 *)p = v; }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8m __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(1))) const bf16x8 gcbf8;
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) { return *(gcbf8 *)p; }
__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
  return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}
// gfx950's double-K form: the products of a 32-wide K chunk in one instruction
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8m, a),
                                                 __builtin_bit_cast(bf16x8m, b), c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// 3x3 convolution (taps dilated by A.dil), forward and input gradient (the
// adjoint of a dilated correlation is the same dilation with flipped taps).
// One wave computes NB*16 pixels x MB*16 output channels; a block has 4 waves.
// ---------------------------------------------------------------------------
template <int MB, int NB>
__global__ __launch_bounds__(256) void k_conv3(UnConvArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hw = A.h * A.w;
  const int npix = A.n * hw;
  const int p0 = (blockIdx.x * 4 + wave) * (NB * 16);
  if (p0 >= npix) return;
  const int col = lane & 15, grp = lane >> 4;
  const int cin = A.ca + A.cb;
  const int h2 = A.h >> 1, w2 = A.w >> 1;

  int pimg[NB], py[NB], px[NB];
  bool pv[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int p = p0 + nb * 16 + col;
    pv[nb] = p < npix;
    const int pc = pv[nb] ? p : 0;
    pimg[nb] = pc / hw;
    const int rem = pc - pimg[nb] * hw;
    py[nb] = rem / A.w;
    px[nb] = rem - py[nb] * A.w;
  }

  f32x4 acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  const uint16_t *wrow[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) wrow[mb] = A.wgt + (size_t)(mb * 16 + col) * A.kp;

  if (cin >= 16) {
    // cin is a multiple of 16: each tap's K in 32-wide chunks, then a 16-wide
    // one when cin % 32 == 16 (the skip/coarser boundary ca is a multiple of 8)
    for (int t = 0; t < 9; ++t) {
      const int dy = (t / 3 - 1) * A.dil, dx = (t % 3 - 1) * A.dil;
      const uint16_t *sa[NB];
      const uint16_t *sb[NB];
      bool ok[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int yy = py[nb] + dy, xx = px[nb] + dx;
        ok[nb] = pv[nb] && yy >= 0 && yy < A.h && xx >= 0 && xx < A.w;
        const int ys = ok[nb] ? yy : 0, xs = ok[nb] ? xx : 0;
        sa[nb] = A.a + ((size_t)(pimg[nb] * A.h + ys) * A.w + xs) * A.ca;
        sb[nb] = A.cb ? A.b + ((size_t)(pimg[nb] * h2 + (ys >> 1)) * w2 + (xs >> 1)) * A.cb - A.ca
                      : nullptr;
      }
      const int kt = t * cin;
      int c0 = 0;
      for (; c0 + 32 <= cin; c0 += 32) {
        const int c = c0 + grp * 8;
        bf16x8 bv[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          bv[nb] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
          if (ok[nb]) bv[nb] = (c < A.ca) ? ld8(sa[nb] + c) : ld8(sb[nb] + c);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const bf16x8 av = ld8(wrow[mb] + kt + c);
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma32(av, bv[nb], acc[mb][nb]);
        }
      }
      if (c0 < cin) {
        const int c = c0 + grp * 4;
        bf16x4 bv[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          bv[nb] = bf16x4{0, 0, 0, 0};
          if (ok[nb]) bv[nb] = (c < A.ca) ? ld4(sa[nb] + c) : ld4(sb[nb] + c);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const bf16x4 av = ld4(wrow[mb] + kt + c);
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma16(av, bv[nb], acc[mb][nb]);
        }
      }
    }
  } else {
    // cin == 4 (the network input): k = 4 * tap + channel; a 32-wide K chunk
    // spans eight taps (two per lane group), a 16-wide one four (kp is a
    // multiple of 16: 48 = 32 + 16)
    auto tap4 = [&](int t, int nb) -> bf16x4 {
      const int dy = (t / 3 - 1) * A.dil, dx = (t % 3 - 1) * A.dil;
      const int yy = py[nb] + dy, xx = px[nb] + dx;
      if (t < 9 && pv[nb] && yy >= 0 && yy < A.h && xx >= 0 && xx < A.w)
        return ld4(A.a + ((size_t)(pimg[nb] * A.h + yy) * A.w + xx) * 4);
      return bf16x4{0, 0, 0, 0};
    };
    int k0 = 0;
    for (; k0 + 32 <= A.kp; k0 += 32) {
      const int t = (k0 >> 2) + 2 * grp;
      bf16x8 bv[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) bv[nb] = cat8(tap4(t, nb), tap4(t + 1, nb));
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const bf16x8 av = ld8(wrow[mb] + k0 + grp * 8);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma32(av, bv[nb], acc[mb][nb]);
      }
    }
    if (k0 < A.kp) {
      const int t = (k0 >> 2) + grp;
      bf16x4 bv[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) bv[nb] = tap4(t, nb);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const bf16x4 av = ld4(wrow[mb] + k0 + grp * 4);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma16(av, bv[nb], acc[mb][nb]);
      }
    }
  }

  // epilogue: bias, ReLU, gate (x (gate > 0)), bf16 store of 4 channels
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    if (!pv[nb]) continue;
    const size_t p = (size_t)(p0 + nb * 16 + col);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int co = mb * 16 + grp * 4;
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = acc[mb][nb][i];
        if (A.bias) v[i] += A.bias[co + i];
        if (A.relu) v[i] = fmaxf(v[i], 0.0f);
      }
      if (A.gate) {
        const bf16x4 g = ld4(A.gate + p * A.cout + co);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (!(bf2f(g[i]) > 0.0f)) v[i] = 0.0f;
      }
      st4(A.out + p * A.cout + co,
          bf16x4{(short)f2bf(v[0]), (short)f2bf(v[1]), (short)f2bf(v[2]), (short)f2bf(v[3])});
    }
  }
}

// Split-K variant for convolutions over few pixels. In k_conv3 a wave's serial
// chain is 9 taps x cin/16 chunks of load -> MFMA steps whatever the pixel
// count, so a layer of a few thousand pixels runs on a few dozen waves, each
// bound by load latency (~40 us for 64 channels). Here the 4 waves of a block
// share one 16-pixel tile and split the taps (wave w: taps w, w+4, w+8); their
// accumulators are added in wave order through LDS (deterministic), and wave w
// finishes output blocks mb = w, w + 4. Requires cin % 16 == 0.
template <int MB>
__global__ __launch_bounds__(256) void k_conv3_sk(UnConvArgs A) {
  __shared__ f32x4 red[4][MB][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hw = A.h * A.w;
  const int npix = A.n * hw;
  const int col = lane & 15, grp = lane >> 4;
  const int cin = A.ca + A.cb;
  const int h2 = A.h >> 1, w2 = A.w >> 1;
  const int p = blockIdx.x * 16 + col;
  const bool pv = p < npix;
  const int pc = pv ? p : 0;
  const int pimg = pc / hw;
  const int rem = pc - pimg * hw;
  const int py = rem / A.w, px = rem - py * A.w;

  f32x4 acc[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint16_t *wrow[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) wrow[mb] = A.wgt + (size_t)(mb * 16 + col) * A.kp;

  if (MB >= 4 && cin % 32 == 0) {
    // wave w takes taps w, w + 4, w + 8. With 64 output channels or more its
    // (tap, 32-wide chunk) steps run as one loop whose operands are loaded a
    // step ahead (the products still enter each accumulator in (tap, chunk)
    // order): 12.3 -> 9.9 us for the 64-channel layers at a batch of three
    // 128^2 grids; narrower layers keep the plain loop, which the extra
    // registers slowed.
    const int nchunk = cin >> 5;
    const int ntap = (9 - wave + 3) / 4, nstep = ntap * nchunk;
    auto load_b = [&](int st) -> bf16x8 {
      const int i = st / nchunk, c = (st - i * nchunk) * 32 + grp * 8, t = wave + 4 * i;
      const int dy = (t / 3 - 1) * A.dil, dx = (t % 3 - 1) * A.dil;
      const int yy = py + dy, xx = px + dx;
      const bool ok = pv && yy >= 0 && yy < A.h && xx >= 0 && xx < A.w;
      const int ys = ok ? yy : 0, xs = ok ? xx : 0;
      if (!ok) return bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
      return (c < A.ca) ? ld8(A.a + ((size_t)(pimg * A.h + ys) * A.w + xs) * A.ca + c)
                        : ld8(A.b + ((size_t)(pimg * h2 + (ys >> 1)) * w2 + (xs >> 1)) * A.cb -
                              A.ca + c);
    };
    auto load_a = [&](int st, bf16x8 (&av)[MB]) {
      const int i = st / nchunk, c = (st - i * nchunk) * 32 + grp * 8, t = wave + 4 * i;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) av[mb] = ld8(wrow[mb] + t * cin + c);
    };
    if (nstep > 0) {
      bf16x8 bv = load_b(0), av[MB];
      load_a(0, av);
      for (int st = 0; st < nstep; ++st) {
        bf16x8 bn = bv, an[MB];
        if (st + 1 < nstep) {
          bn = load_b(st + 1);
          load_a(st + 1, an);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[mb] = mfma32(av[mb], bv, acc[mb]);
        if (st + 1 < nstep) {
          bv = bn;
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) av[mb] = an[mb];
        }
      }
    }
  } else {
    for (int t = wave; t < 9; t += 4) {
      const int dy = (t / 3 - 1) * A.dil, dx = (t % 3 - 1) * A.dil;
      const int yy = py + dy, xx = px + dx;
      const bool ok = pv && yy >= 0 && yy < A.h && xx >= 0 && xx < A.w;
      const int ys = ok ? yy : 0, xs = ok ? xx : 0;
      const uint16_t *sa = A.a + ((size_t)(pimg * A.h + ys) * A.w + xs) * A.ca;
      const uint16_t *sb =
          A.cb ? A.b + ((size_t)(pimg * h2 + (ys >> 1)) * w2 + (xs >> 1)) * A.cb - A.ca : nullptr;
      const int kt = t * cin;
      int c0 = 0;
      for (; c0 + 32 <= cin; c0 += 32) {
        const int c = c0 + grp * 8;
        bf16x8 bv = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        if (ok) bv = (c < A.ca) ? ld8(sa + c) : ld8(sb + c);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[mb] = mfma32(ld8(wrow[mb] + kt + c), bv, acc[mb]);
      }
      if (c0 < cin) {  // cin % 32 == 16
        const int c = c0 + grp * 4;
        bf16x4 bv = bf16x4{0, 0, 0, 0};
        if (ok) bv = (c < A.ca) ? ld4(sa + c) : ld4(sb + c);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[mb] = mfma16(ld4(wrow[mb] + kt + c), bv, acc[mb]);
      }
    }
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) red[wave][mb][lane] = acc[mb];
  __syncthreads();
  for (int mb = wave; mb < MB; mb += 4) {
    const f32x4 v4 = ((red[0][mb][lane] + red[1][mb][lane]) + red[2][mb][lane]) + red[3][mb][lane];
    if (!pv) continue;
    const int co = mb * 16 + grp * 4;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = v4[i];
      if (A.bias) v[i] += A.bias[co + i];
      if (A.relu) v[i] = fmaxf(v[i], 0.0f);
    }
    if (A.gate) {
      const bf16x4 g = ld4(A.gate + (size_t)p * A.cout + co);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (!(bf2f(g[i]) > 0.0f)) v[i] = 0.0f;
    }
    st4(A.out + (size_t)p * A.cout + co,
        bf16x4{(short)f2bf(v[0]), (short)f2bf(v[1]), (short)f2bf(v[2]), (short)f2bf(v[3])});
  }
}

template __global__ void k_conv3_sk<1>(UnConvArgs);
template __global__ void k_conv3_sk<2>(UnConvArgs);
template __global__ void k_conv3_sk<3>(UnConvArgs);
template __global__ void k_conv3_sk<4>(UnConvArgs);
template __global__ void k_conv3_sk<6>(UnConvArgs);
template __global__ void k_conv3_sk<8>(UnConvArgs);

template __global__ void k_conv3<1, 4>(UnConvArgs);
template __global__ void k_conv3<2, 4>(UnConvArgs);
template __global__ void k_conv3<3, 4>(UnConvArgs);
template __global__ void k_conv3<4, 4>(UnConvArgs);
template __global__ void k_conv3<6, 2>(UnConvArgs);
template __global__ void k_conv3<8, 2>(UnConvArgs);

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][k] = sum_p dY[p][co] * patch_k(p), db[co] = sum_p dY[p][co].
// Block = (pixel chunk, K range). Each 64-pixel sub-tile of dY and of the
// patch columns is staged transposed in LDS (pixels contiguous), so an MFMA
// operand (eight consecutive pixels of one row: the pixels are this GEMM's K,
// taken 32 at a time by v_mfma_f32_16x16x32_bf16) is one 16-byte LDS read.
// ---------------------------------------------------------------------------
constexpr int DW_PX = 64;
constexpr int DW_LD = DW_PX + 8;  // LDS row pitch (bf16), keeps 16-byte alignment

__global__ __launch_bounds__(256) void k_conv3_dw(UnDwArgs A) {
  __shared__ __align__(16) uint16_t dyT[64 * DW_LD];
  __shared__ __align__(16) uint16_t xT[128 * DW_LD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int hw = A.h * A.w, npix = A.n * hw;
  const int cin = A.ca + A.cb;
  const int h2 = A.h >> 1, w2 = A.w >> 1;
  const int chunk = blockIdx.x, kr0 = blockIdx.y * A.kr;
  const int mbt = A.cout >> 4, nbt = A.kr >> 4, nblk = mbt * nbt;
  const int cg_dy = A.cout >> 2, cg_x = A.kr >> 2;

  f32x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float dbs = 0.0f;

  const int pbeg = chunk * A.chunk;
  const int pend = min(npix, pbeg + A.chunk);
  for (int sp = pbeg; sp < pend; sp += DW_PX) {
    // stage dY^T
    for (int i = threadIdx.x; i < DW_PX * cg_dy; i += 256) {
      const int pxl = i / cg_dy, cg = i - pxl * cg_dy;
      const int p = sp + pxl;
      bf16x4 v = bf16x4{0, 0, 0, 0};
      if (p < pend) v = ld4(A.dy + (size_t)p * A.cout + cg * 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) dyT[(cg * 4 + j) * DW_LD + pxl] = (uint16_t)v[j];
    }
    // stage the patch columns kr0 .. kr0 + kr of the sub-tile
    for (int i = threadIdx.x; i < DW_PX * cg_x; i += 256) {
      const int pxl = i / cg_x, g = i - pxl * cg_x;
      const int p = sp + pxl;
      const int k = kr0 + g * 4;
      const int t = k / cin, ci = k - t * cin;
      bf16x4 v = bf16x4{0, 0, 0, 0};
      if (p < pend && t < 9) {
        const int img = p / hw, rem = p - img * hw;
        const int y = rem / A.w, x = rem - y * A.w;
        const int yy = y + (t / 3 - 1) * A.dil, xx = x + (t % 3 - 1) * A.dil;
        if (yy >= 0 && yy < A.h && xx >= 0 && xx < A.w) {
          if (ci < A.ca)
            v = ld4(A.a + ((size_t)(img * A.h + yy) * A.w + xx) * A.ca + ci);
          else
            v = ld4(A.b + ((size_t)(img * h2 + (yy >> 1)) * w2 + (xx >> 1)) * A.cb + (ci - A.ca));
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) xT[(g * 4 + j) * DW_LD + pxl] = (uint16_t)v[j];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < DW_PX / 32; ++ks) {
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = wave + 4 * jj;
        if (j < nblk) {
          const int mb = j % mbt, nb = j / mbt;
          const bf16x8 av = *(const bf16x8 *)&dyT[(mb * 16 + col) * DW_LD + ks * 32 + grp * 8];
          const bf16x8 bv = *(const bf16x8 *)&xT[(nb * 16 + col) * DW_LD + ks * 32 + grp * 8];
          acc[jj] = mfma32(av, bv, acc[jj]);
        }
      }
    }
    if (blockIdx.y == 0 && (int)threadIdx.x < A.cout) {
      float s = 0.0f;
      for (int q = 0; q < DW_PX; ++q) s += bf2f(dyT[threadIdx.x * DW_LD + q]);
      dbs += s;
    }
    __syncthreads();
  }

  float *part = A.part + (size_t)chunk * A.cout * A.kp;
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const int j = wave + 4 * jj;
    if (j < nblk) {
      const int mb = j % mbt, nb = j / mbt;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        part[(size_t)(mb * 16 + grp * 4 + i) * A.kp + kr0 + nb * 16 + col] = acc[jj][i];
    }
  }
  if (blockIdx.y == 0 && (int)threadIdx.x < A.cout)
    A.part_db[(size_t)chunk * A.cout + threadIdx.x] = dbs;
}

// Sum the chunk partials in chunk order (deterministic) into the FP32 gradient.
__global__ void k_dw_reduce(const float *__restrict__ part, const float *__restrict__ part_db,
                            int nchunk, int cout, int kp, float *__restrict__ gw,
                            float *__restrict__ gb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = cout * kp;
  // (unrolled so that several chunks' loads are in flight; the adds keep
  // chunk order)
  if (i < nw) {
    float s = 0.0f;
#pragma unroll 8
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * nw + i];
    gw[i] = s;
  } else if (i < nw + cout) {
    const int co = i - nw;
    float s = 0.0f;
#pragma unroll 8
    for (int c = 0; c < nchunk; ++c) s += part_db[(size_t)c * cout + co];
    gb[co] = s;
  }
}

// ---------------------------------------------------------------------------
// Pooling and gradient routing
// ---------------------------------------------------------------------------
// 2x2 average pool, NHWC bf16 (c multiple of 4)
__global__ void k_pool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int n, int h,
                        int w, int c) {
  const int h2 = h >> 1, w2 = w >> 1, cg = c >> 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * h2 * w2 * cg) return;
  const int g = i % cg, q = i / cg;
  const int img = q / (h2 * w2), rem = q - img * h2 * w2;
  const int y = rem / w2, x = rem - y * w2;
  const size_t b00 = ((size_t)(img * h + 2 * y) * w + 2 * x) * c + g * 4;
  const bf16x4 a = ld4(in + b00), b = ld4(in + b00 + c), d = ld4(in + b00 + (size_t)w * c),
               e = ld4(in + b00 + (size_t)w * c + c);
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    o[j] = (short)f2bf(((bf2f(a[j]) + bf2f(b[j])) + (bf2f(d[j]) + bf2f(e[j]))) * 0.25f);
  st4(out + (size_t)q * c + g * 4, o);
}

// Gradient at the pre-activation of a skip tensor e (res h x w, c1 channels):
// (dcat[:, 0:c1] + dpool/4 at the pooled pixel) x (e > 0).
__global__ void k_skip_grad(const uint16_t *__restrict__ dcat, int ccat,
                            const uint16_t *__restrict__ dpool, const uint16_t *__restrict__ e,
                            int c1, int n, int h, int w, uint16_t *__restrict__ dz) {
  const int cg = c1 >> 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * h * w * cg) return;
  const int g = i % cg, p = i / cg;
  const int img = p / (h * w), rem = p - img * h * w;
  const int y = rem / w, x = rem - y * w;
  const size_t q = (size_t)(img * (h >> 1) + (y >> 1)) * (w >> 1) + (x >> 1);
  const bf16x4 a = ld4(dcat + (size_t)p * ccat + g * 4);
  const bf16x4 b = ld4(dpool + q * c1 + g * 4);
  const bf16x4 s = ld4(e + (size_t)p * c1 + g * 4);
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float v = fmaf(bf2f(b[j]), 0.25f, bf2f(a[j]));
    o[j] = (short)f2bf(bf2f(s[j]) > 0.0f ? v : 0.0f);
  }
  st4(dz + (size_t)p * c1 + g * 4, o);
}

// Gradient at the pre-activation of the coarser decoder input (res h/2 x w/2,
// c2 channels): the nearest-upsample's 2x2 sum of dcat[:, c1:c1+c2], x (prev > 0).
__global__ void k_up_grad(const uint16_t *__restrict__ dcat, int ccat, int c1,
                          const uint16_t *__restrict__ prev, int c2, int n, int h, int w,
                          uint16_t *__restrict__ dz) {
  const int h2 = h >> 1, w2 = w >> 1, cg = c2 >> 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * h2 * w2 * cg) return;
  const int g = i % cg, q = i / cg;
  const int img = q / (h2 * w2), rem = q - img * h2 * w2;
  const int y = rem / w2, x = rem - y * w2;
  const size_t p00 = (size_t)(img * h + 2 * y) * w + 2 * x;
  const size_t off = c1 + g * 4;
  const bf16x4 a = ld4(dcat + p00 * ccat + off), b = ld4(dcat + (p00 + 1) * ccat + off),
               d = ld4(dcat + (p00 + w) * ccat + off), e = ld4(dcat + (p00 + w + 1) * ccat + off);
  const bf16x4 s = ld4(prev + (size_t)q * c2 + g * 4);
  bf16x4 o;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float v = (bf2f(a[j]) + bf2f(b[j])) + (bf2f(d[j]) + bf2f(e[j]));
    o[j] = (short)f2bf(bf2f(s[j]) > 0.0f ? v : 0.0f);
  }
  st4(dz + (size_t)q * c2 + g * 4, o);
}

// ---------------------------------------------------------------------------
// FP32 1x1 output head, loss and its gradient
// ---------------------------------------------------------------------------
__global__ void k_head(const uint16_t *__restrict__ d0, const float *__restrict__ wh, int npix,
                       float *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float s = wh[UN_HEAD_C];
#pragma unroll
  for (int g = 0; g < UN_HEAD_C / 4; ++g) {
    const bf16x4 v = ld4(d0 + (size_t)p * UN_HEAD_C + g * 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) s = fmaf(wh[g * 4 + j], bf2f(v[j]), s);
  }
  out[p] = s;
}

// Block means of the weighted error e = wt * (out - y) over 2x2, 4x4, 8x8
// and 16x16 blocks, one 256-thread block per 16x16 tile (res a multiple of
// 16), reduced in a fixed order. pyr holds, per image, the four levels
// [res/2]^2, [res/4]^2, [res/8]^2, [res/16]^2 back to back.
__global__ __launch_bounds__(256) void k_err_pyramid(const float *__restrict__ out,
                                                     const float *__restrict__ y,
                                                     const float *__restrict__ wt, int res,
                                                     float *__restrict__ pyr) {
  const int tiles = res >> 4;
  const int img = blockIdx.x / (tiles * tiles), t = blockIdx.x % (tiles * tiles);
  const int ty = t / tiles, tx = t % tiles;
  const int ly = threadIdx.x >> 4, lx = threadIdx.x & 15;
  const size_t p = ((size_t)img * res + ty * 16 + ly) * res + tx * 16 + lx;
  __shared__ float e[256];
  e[threadIdx.x] = wt[p] * (out[p] - y[p]);
  __syncthreads();
  const size_t per_img = (size_t)(res / 2) * (res / 2) + (res / 4) * (res / 4) +
                         (res / 8) * (res / 8) + (res / 16) * (res / 16);
  float *base = pyr + img * per_img;
  int off = 0;
  for (int lvl = 1; lvl <= 4; ++lvl) {
    const int bs = 1 << lvl, nb = 16 >> lvl, r = res >> lvl;
    if ((int)threadIdx.x < nb * nb) {
      const int by = threadIdx.x / nb, bx = threadIdx.x % nb;
      float sum = 0.0f;
      for (int yy = 0; yy < bs; ++yy)
        for (int xx = 0; xx < bs; ++xx) sum += e[(by * bs + yy) * 16 + bx * bs + xx];
      base[off + (size_t)(ty * nb + by) * r + tx * nb + bx] = sum / (float)(bs * bs);
    }
    off += r * r;
  }
}

// loss = (1/n) sum_p [wt (out - y)^2 + alpha sum_s m_s(p)^2] with m_s the
// pyramid block means (alpha = 0: plain weighted MSE);
// dout = (2/n) wt (out - y + alpha sum_s m_s(p)).
// Writes dz_d0 (gradient at the pre-activation of the last decoder conv) and
// per-block partials of [dwh (16), dbh, loss] for a fixed-order reduction.
constexpr int UN_HEAD_PART = UN_HEAD_C + 2;

__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ out,
                                                  const float *__restrict__ y,
                                                  const float *__restrict__ wt,
                                                  const uint16_t *__restrict__ d0,
                                                  const float *__restrict__ wh, int npix,
                                                  float inv_n, const float *__restrict__ pyr,
                                                  int res, float alpha,
                                                  uint16_t *__restrict__ dz,
                                                  float *__restrict__ part) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float acc[UN_HEAD_PART];
#pragma unroll
  for (int j = 0; j < UN_HEAD_PART; ++j) acc[j] = 0.0f;
  if (p < npix) {
    const float diff = out[p] - y[p];
    float msum = 0.0f, m2 = 0.0f;
    if (alpha != 0.0f) {
      const int hw = res * res, img = p / hw, rem = p - img * hw, py = rem / res, px = rem % res;
      const size_t per_img = (size_t)(res / 2) * (res / 2) + (res / 4) * (res / 4) +
                             (res / 8) * (res / 8) + (res / 16) * (res / 16);
      const float *base = pyr + img * per_img;
      int off = 0;
      for (int lvl = 1; lvl <= 4; ++lvl) {
        const int r = res >> lvl;
        const float m = base[off + (size_t)(py >> lvl) * r + (px >> lvl)];
        msum += m;
        m2 = fmaf(m, m, m2);
        off += r * r;
      }
    }
    const float dout = 2.0f * wt[p] * fmaf(alpha, msum, diff) * inv_n;
    acc[UN_HEAD_C] = dout;
    acc[UN_HEAD_C + 1] = fmaf(alpha, m2, wt[p] * diff * diff) * inv_n;
#pragma unroll
    for (int g = 0; g < UN_HEAD_C / 4; ++g) {
      const bf16x4 v = ld4(d0 + (size_t)p * UN_HEAD_C + g * 4);
      bf16x4 o;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float a = bf2f(v[j]);
        acc[g * 4 + j] = dout * a;
        o[j] = (short)f2bf(a > 0.0f ? dout * wh[g * 4 + j] : 0.0f);
      }
      st4(dz + (size_t)p * UN_HEAD_C + g * 4, o);
    }
  }
  __shared__ float red[UN_HEAD_PART][256];
#pragma unroll
  for (int j = 0; j < UN_HEAD_PART; ++j) red[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k)
#pragma unroll
      for (int j = 0; j < UN_HEAD_PART; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + k];
    __syncthreads();
  }
  if ((int)threadIdx.x < UN_HEAD_PART) part[(size_t)blockIdx.x * UN_HEAD_PART + threadIdx.x] =
      red[threadIdx.x][0];
}

// Fixed-order sum of the head partials: gh[0..16] = [dwh, dbh], loss[0] (FP64).
// One block per sum j (grid UN_HEAD_PART, 256 threads): thread t adds the
// partials t, t + 256, ... in order, then the block adds its threads' sums by
// a fixed tree (one thread walking all partials took ~0.3 ms at 2048 blocks:
// a chain of dependent loads).
__global__ __launch_bounds__(256) void k_head_reduce(const float *__restrict__ part, int nblocks,
                                                     float *__restrict__ gh,
                                                     double *__restrict__ loss) {
  __shared__ double red[256];
  const int j = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < nblocks; b += 256) s += (double)part[(size_t)b * UN_HEAD_PART + j];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) {
    if (j < UN_HEAD_C + 1)
      gh[j] = (float)red[0];
    else
      loss[0] = red[0];
  }
}

// ---------------------------------------------------------------------------
// Parameters
// ---------------------------------------------------------------------------
// Adam with decoupled weight decay on the FP32 master copy.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, float wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
  const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / bc1, vh = vi / bc2;
  p[i] = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

// bf16 copies of one layer: w16[co][k] and (when wt16 is set) the flipped,
// transposed weights of the input-gradient convolution,
// wt16[ci][t'*cout + co] = w[co][(8 - t')*cp + ci].
__global__ void k_cast_w(const float *__restrict__ w, int cout, int cp, int kp,
                         uint16_t *__restrict__ w16, uint16_t *__restrict__ wt16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cout * kp) return;
  const int co = i / kp, k = i - co * kp;
  const uint16_t v = f2bf(w[i]);
  w16[i] = v;
  if (wt16 && k < 9 * cp) {
    const int t = k / cp, ci = k - t * cp;
    wt16[(size_t)ci * (9 * cout) + (8 - t) * cout + co] = v;
  }
}

// Training batch gather: sample j of the batch is sample idx[j] of the
// dataset under the dihedral map mode[j] (bit 2: transpose, bit 0: mirror x,
// bit 1: mirror y) -- the Poisson problem is invariant under the 8 maps, and
// every input and the target are scalar fields on the grid.
__global__ void k_gather_dihedral(const float *__restrict__ xin, const float *__restrict__ yin,
                                  const float *__restrict__ win, UnGather G, int res,
                                  float *__restrict__ xo, float *__restrict__ yo,
                                  float *__restrict__ wo) {
  const int j = blockIdx.y, y = blockIdx.x;
  const int src = G.idx[j], mode = G.mode[j];
  for (int x = threadIdx.x; x < res; x += blockDim.x) {
    int sy = (mode & 4) ? x : y, sx = (mode & 4) ? y : x;
    if (mode & 1) sx = res - 1 - sx;
    if (mode & 2) sy = res - 1 - sy;
    const size_t q = ((size_t)src * res + sy) * res + sx;
    const size_t p = ((size_t)j * res + y) * res + x;
    *(float4 *)(xo + 4 * p) = *(const float4 *)(xin + 4 * q);
    yo[p] = yin[q];
    wo[p] = win[q];
  }
}

// FP32 [n][res][res][4] network input -> bf16
__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

}  // namespace ph
