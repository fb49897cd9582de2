// unet_internal.h -- the warm-start network's layer table and kernel arguments.
//
// A 4-level U-Net on a res x res grid (res a multiple of 8):
//   level 0 (res)     e0a 4->16, e0b 16->16             pool
//   level 1 (res/2)   e1a 16->32, e1b 32->32            pool
//   level 2 (res/4)   e2a 32->64, e2b 64->64            pool
//   level 3 (res/8)   b3a 64->64 (dilation 2), b3b 64->64 (dilation 4)
//   level 2           d2a cat(e2b, up(b3b)) 128->64, d2b 64->64
//   level 1           d1a cat(e1b, up(d2b)) 96->32,  d1b 32->32
//   level 0           d0a cat(e0b, up(d1b)) 48->16,  d0b 16->16
//   head              1x1, 16->1, FP32
// The dilated bottleneck widens the receptive field to the whole grid (the
// harmonic correction at a cell depends on the entire boundary). Every 3x3
// convolution has a bias and a ReLU. The input has 4 channels: the
// right-hand side of the correction problem, the boundary values (zero off
// the Dirichlet ring), the interior fraction and the ring density.
//
// Parameter storage (FP32, also the on-disk layout): per convolution
// W[cout][kp] (k = tap*cp + ci, cp = input channels, kp = 9*cp rounded up
// to 16, padding zero) followed by bias[cout]; then the
// head's 16 weights and its bias.
#pragma once
#include <stdint.h>

namespace ph {

constexpr int UN_LAYERS = 14;
constexpr int UN_HEAD_C = 16;
constexpr int UN_IN_C = 4;

// activation buffers
enum : int {
  UB_X = 0, UB_A0, UB_E0, UB_P0, UB_A1, UB_E1, UB_P1, UB_A2, UB_E2, UB_P2, UB_A3, UB_B3,
  UB_D2A, UB_D2, UB_D1A, UB_D1, UB_D0A, UB_D0, UB_COUNT
};

struct UnLayerDef {
  int level;  // resolution res >> level
  int in_a, in_b;  // activation buffers (in_b = -1: none; else nearest-upsampled x2)
  int ca, cb;      // channels of in_a / in_b
  int cout;
  int out;         // activation buffer written
  int dil;         // dilation of the 3x3 taps
};

// clang-format off
constexpr UnLayerDef UN_DEF[UN_LAYERS] = {
    {0, UB_X,   -1,     4,  0, 16, UB_A0, 1},
    {0, UB_A0,  -1,    16,  0, 16, UB_E0, 1},
    {1, UB_P0,  -1,    16,  0, 32, UB_A1, 1},
    {1, UB_A1,  -1,    32,  0, 32, UB_E1, 1},
    {2, UB_P1,  -1,    32,  0, 64, UB_A2, 1},
    {2, UB_A2,  -1,    64,  0, 64, UB_E2, 1},
    {3, UB_P2,  -1,    64,  0, 64, UB_A3, 2},
    {3, UB_A3,  -1,    64,  0, 64, UB_B3, 4},
    {2, UB_E2,  UB_B3, 64, 64, 64, UB_D2A, 1},
    {2, UB_D2A, -1,    64,  0, 64, UB_D2, 1},
    {1, UB_E1,  UB_D2, 32, 64, 32, UB_D1A, 1},
    {1, UB_D1A, -1,    32,  0, 32, UB_D1, 1},
    {0, UB_E0,  UB_D1, 16, 32, 16, UB_D0A, 1},
    {0, UB_D0A, -1,    16,  0, 16, UB_D0, 1},
};
// clang-format on

constexpr int un_kp(int cin) { return ((9 * cin + 15) / 16) * 16; }

struct UnConvArgs {
  const uint16_t *a;     // [n][h][w][ca]
  const uint16_t *b;     // [n][h/2][w/2][cb] read through a nearest x2 upsample, or null
  const uint16_t *wgt;   // [cout][kp]
  const float *bias;     // [cout] or null
  const uint16_t *gate;  // [n][h][w][cout] or null: out = 0 where gate <= 0
  uint16_t *out;         // [n][h][w][cout]
  int ca, cb, kp, cout;
  int n, h, w;
  int relu;
  int dil;               // dilation of the 3x3 taps
};

struct UnDwArgs {
  const uint16_t *dy;  // [n][h][w][cout]
  const uint16_t *a, *b;
  int ca, cb;
  float *part;         // [nchunk][cout][kp]
  float *part_db;      // [nchunk][cout]
  int cout, kp, kr;    // kr: K columns per block
  int n, h, w;
  int chunk;           // pixels per chunk (multiple of 64)
  int dil;             // dilation of the 3x3 taps
};

constexpr int UN_GATHER_MAX = 256;
struct UnGather {
  int32_t idx[UN_GATHER_MAX];
  int32_t mode[UN_GATHER_MAX];
};

template <int MB, int NB>
__global__ void k_conv3(UnConvArgs A);
template <int MB>
__global__ void k_conv3_sk(UnConvArgs A);
// Convolutions over at most this many pixels take the split-K kernel: the
// solver's batch of three 128x128 grids (one RGB image) has too few waves for
// the one-wave-per-tile kernel to hide its load chain, while a 64-grid batch's
// 32x32 level (65536 pixels) is faster without the split.
constexpr int UN_SPLITK_MAX_PIX = 49152;
__global__ void k_gather_dihedral(const float *xin, const float *yin, const float *win, UnGather G,
                                  int res, float *xo, float *yo, float *wo);
__global__ void k_conv3_dw(UnDwArgs A);
__global__ void k_dw_reduce(const float *part, const float *part_db, int nchunk, int cout, int kp,
                            float *gw, float *gb);
__global__ void k_pool2(const uint16_t *in, uint16_t *out, int n, int h, int w, int c);
__global__ void k_skip_grad(const uint16_t *dcat, int ccat, const uint16_t *dpool,
                            const uint16_t *e, int c1, int n, int h, int w, uint16_t *dz);
__global__ void k_up_grad(const uint16_t *dcat, int ccat, int c1, const uint16_t *prev, int c2,
                          int n, int h, int w, uint16_t *dz);
__global__ void k_head(const uint16_t *d0, const float *wh, int npix, float *out);
__global__ void k_err_pyramid(const float *out, const float *y, const float *wt, int res,
                              float *pyr);
__global__ void k_head_bwd(const float *out, const float *y, const float *wt, const uint16_t *d0,
                           const float *wh, int npix, float inv_n, const float *pyr, int res,
                           float alpha, uint16_t *dz, float *part);
__global__ void k_head_reduce(const float *part, int nblocks, float *gh, double *loss);
__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float wd);
__global__ void k_cast_w(const float *w, int cout, int cp, int kp, uint16_t *w16, uint16_t *wt16);
__global__ void k_to_bf16(const float *in, uint16_t *out, int n);

}  // namespace ph
