// unet.hip -- host side of the warm-start network: parameter layout, device
// buffers, forward pass, backward pass, Adam, and the C ABI (ph_unet_*).
//
// The network is trained and run entirely by the kernels in unet_kernels.hip;
// no library GEMM or convolution is called. Training state (FP32 master
// weights, Adam moments, step count) round-trips through the C ABI so a
// caller can checkpoint and resume.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "host_common.h"
#include "unet_internal.h"

namespace ph {

struct UnLayout {
  size_t off_w[UN_LAYERS], off_b[UN_LAYERS], off_head, storage, count;
  int kp[UN_LAYERS], cp[UN_LAYERS];
  UnLayout() {
    size_t o = 0;
    count = 0;
    for (int l = 0; l < UN_LAYERS; ++l) {
      const UnLayerDef &d = UN_DEF[l];
      cp[l] = d.ca + d.cb;
      kp[l] = un_kp(cp[l]);
      off_w[l] = o;
      o += (size_t)d.cout * kp[l];
      off_b[l] = o;
      o += d.cout;
      const int cin_real = cp[l];
      count += (size_t)9 * cin_real * d.cout + d.cout;
    }
    off_head = o;
    o += UN_HEAD_C + 1;
    count += UN_HEAD_C + 1;
    storage = o;
  }
};
static const UnLayout g_lay;

static int buf_channels(int b) {
  switch (b) {
    case UB_X: return UN_IN_C;
    case UB_A0: case UB_E0: case UB_P0: case UB_D0A: case UB_D0: return 16;
    case UB_A1: case UB_E1: case UB_P1: case UB_D1A: case UB_D1: return 32;
    default: return 64;
  }
}
static int buf_level(int b) {
  switch (b) {
    case UB_X: case UB_A0: case UB_E0: case UB_D0A: case UB_D0: return 0;
    case UB_P0: case UB_A1: case UB_E1: case UB_D1A: case UB_D1: return 1;
    case UB_P1: case UB_A2: case UB_E2: case UB_D2A: case UB_D2: return 2;
    default: return 3;
  }
}

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Seeded He-normal initialisation (SplitMix64 + Box-Muller), zero biases and
// zero padding.
static void init_params(std::vector<float> &p, uint64_t seed) {
  p.assign(g_lay.storage, 0.0f);
  uint64_t s = seed;
  auto next = [&]() {
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() { return sqrt(-2.0 * log(next())) * cos(6.283185307179586 * next()); };
  for (int l = 0; l < UN_LAYERS; ++l) {
    const UnLayerDef &d = UN_DEF[l];
    const int cin_real = g_lay.cp[l];
    const double sd = sqrt(2.0 / (9.0 * cin_real));
    for (int co = 0; co < d.cout; ++co)
      for (int t = 0; t < 9; ++t)
        for (int ci = 0; ci < cin_real; ++ci)
          p[g_lay.off_w[l] + (size_t)co * g_lay.kp[l] + t * g_lay.cp[l] + ci] = (float)(sd * normal());
  }
  const double sd = sqrt(1.0 / UN_HEAD_C);
  for (int c = 0; c < UN_HEAD_C; ++c) p[g_lay.off_head + c] = (float)(sd * normal());
}

// Every K-padding slot must be zero.
static void check_padding(const float *p) {
  for (int l = 0; l < UN_LAYERS; ++l) {
    const UnLayerDef &d = UN_DEF[l];
    const int cin_real = g_lay.cp[l];
    for (int co = 0; co < d.cout; ++co)
      for (int k = 0; k < g_lay.kp[l]; ++k) {
        const int t = k / g_lay.cp[l], ci = k % g_lay.cp[l];
        if ((t >= 9 || ci >= cin_real) && p[g_lay.off_w[l] + (size_t)co * g_lay.kp[l] + k] != 0.0f)
          throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                     "network parameters: padding slot (layer %d, row %d, column %d) is not zero",
                     l, co, k);
      }
  }
}

struct UnNet {
  int max_batch = 0, res = 0;
  int step = 0;
  float *master = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *w16[UN_LAYERS] = {}, *wt16[UN_LAYERS] = {};
  uint16_t *act[UB_COUNT] = {};
  uint16_t *dz[UN_LAYERS] = {};
  uint16_t *dcat[3] = {}, *dpool[3] = {};
  float *out = nullptr;
  float *part = nullptr, *part_db = nullptr, *head_part = nullptr, *pyr = nullptr;
  double *loss = nullptr;
  size_t part_cap = 0, part_db_cap = 0;
  std::vector<void *> allocs;

  template <class T>
  T *alloc(size_t n) {
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)));
    allocs.push_back(p);
    return (T *)p;
  }
  ~UnNet() {
    for (void *p : allocs) (void)hipFree(p);
  }

  size_t npix(int level, int batch) const {
    const size_t r = (size_t)(res >> level);
    return (size_t)batch * r * r;
  }

  // chunk size of the weight-gradient reduction for layer l at this batch
  int dw_chunk(int l, int batch) const {
    const size_t np = npix(UN_DEF[l].level, batch);
    const int kr = g_lay.cp[l] >= 16 ? g_lay.cp[l] : g_lay.kp[l];
    const int nkr = g_lay.kp[l] / kr;
    size_t c = np * nkr / 1024;
    c = std::max<size_t>(256, std::min<size_t>(8192, c));
    return (int)((c + 63) / 64 * 64);
  }

  void create(int mb, int r, const float *host_params, uint64_t seed) {
    if (mb < 1 || r < 16 || r % 8 != 0 || r > 1024)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                 "network buffers: batch %d / resolution %d (need batch >= 1 and resolution a "
                 "multiple of 8 in [16, 1024])",
                 mb, r);
    max_batch = mb;
    res = r;
    const size_t S = g_lay.storage;
    master = alloc<float>(S);
    grad = alloc<float>(S);
    m = alloc<float>(S);
    v = alloc<float>(S);
    for (int l = 0; l < UN_LAYERS; ++l) {
      const UnLayerDef &d = UN_DEF[l];
      w16[l] = alloc<uint16_t>((size_t)d.cout * g_lay.kp[l]);
      if (l > 0) wt16[l] = alloc<uint16_t>((size_t)g_lay.cp[l] * 9 * d.cout);
    }
    for (int b = 0; b < UB_COUNT; ++b) act[b] = alloc<uint16_t>(npix(buf_level(b), mb) * buf_channels(b));
    for (int l = 0; l < UN_LAYERS; ++l)
      dz[l] = alloc<uint16_t>(npix(UN_DEF[l].level, mb) * UN_DEF[l].cout);
    // concat gradients (layers 12, 10, 8 -> levels 0, 1, 2) and pool gradients
    dcat[0] = alloc<uint16_t>(npix(0, mb) * 48);
    dcat[1] = alloc<uint16_t>(npix(1, mb) * 96);
    dcat[2] = alloc<uint16_t>(npix(2, mb) * 128);
    dpool[0] = alloc<uint16_t>(npix(1, mb) * 16);
    dpool[1] = alloc<uint16_t>(npix(2, mb) * 32);
    dpool[2] = alloc<uint16_t>(npix(3, mb) * 64);
    out = alloc<float>(npix(0, mb));
    // the chunking depends on the batch: size the partials for every batch
    for (int b = 1; b <= mb; ++b)
      for (int l = 0; l < UN_LAYERS; ++l) {
        const size_t nchunk = cdiv(npix(UN_DEF[l].level, b), dw_chunk(l, b));
        part_cap = std::max(part_cap, nchunk * UN_DEF[l].cout * g_lay.kp[l]);
        part_db_cap = std::max(part_db_cap, nchunk * UN_DEF[l].cout);
      }
    part = alloc<float>(part_cap);
    part_db = alloc<float>(part_db_cap);
    head_part = alloc<float>((size_t)cdiv(npix(0, mb), 256) * (UN_HEAD_C + 2));
    pyr = alloc<float>(npix(0, mb) / 3 + 64);  // sum of the 2x2..16x16 levels < npix/3
    loss = alloc<double>(1);
    std::vector<float> p;
    if (host_params) {
      check_padding(host_params);
      p.assign(host_params, host_params + S);
    } else {
      init_params(p, seed);
    }
    HIP_OK(hipMemcpy(master, p.data(), S * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(m, 0, S * sizeof(float)));
    HIP_OK(hipMemset(v, 0, S * sizeof(float)));
    HIP_OK(hipMemset(grad, 0, S * sizeof(float)));
    cast_weights(nullptr);
    HIP_OK(hipDeviceSynchronize());
  }

  void cast_weights(hipStream_t st) {
    for (int l = 0; l < UN_LAYERS; ++l) {
      const UnLayerDef &d = UN_DEF[l];
      const int n = d.cout * g_lay.kp[l];
      hipLaunchKernelGGL(k_cast_w, dim3(cdiv(n, 256)), dim3(256), 0, st, master + g_lay.off_w[l],
                         d.cout, g_lay.cp[l], g_lay.kp[l], w16[l], wt16[l]);
    }
    HIP_OK(hipGetLastError());
  }

  void conv(const UnConvArgs &A, hipStream_t st) {
    const size_t np = (size_t)A.n * A.h * A.w;
    const int MB = A.cout / 16;
    // (a 32-wide K chunk's lane group reads 8 channels of one tensor: from 16
    // input channels on, the skip/coarser boundary ca is a multiple of 8)
    if (A.cout % 16 != 0 || (A.ca % 4) || (A.cb % 4) || (A.ca + A.cb != 4 && (A.ca + A.cb) % 16) ||
        (A.ca + A.cb >= 16 && A.ca % 8))
      throw fail(PH_CLASS_EXECUTION, "E_HIP_INTERNAL", "conv shape %d+%d -> %d", A.ca, A.cb,
                 A.cout);
    // few pixels (the solver's batches of a handful of 128x128 grids below the
    // top levels): split the taps over the block's waves instead
    if (np <= UN_SPLITK_MAX_PIX && A.ca + A.cb >= 16) {
      switch (MB) {
#define UN_CONV_SK(M)                                                                    \
  case M:                                                                                \
    hipLaunchKernelGGL((k_conv3_sk<M>), dim3(cdiv(np, 16)), dim3(256), 0, st, A);        \
    break;
        UN_CONV_SK(1)
        UN_CONV_SK(2)
        UN_CONV_SK(3)
        UN_CONV_SK(4)
        UN_CONV_SK(6)
        UN_CONV_SK(8)
#undef UN_CONV_SK
        default:
          throw fail(PH_CLASS_EXECUTION, "E_HIP_INTERNAL", "conv: %d output channels", A.cout);
      }
      HIP_OK(hipGetLastError());
      return;
    }
    switch (MB) {
#define UN_CONV(M, N)                                                                    \
  case M:                                                                                \
    hipLaunchKernelGGL((k_conv3<M, N>), dim3(cdiv(np, 4 * N * 16)), dim3(256), 0, st, A); \
    break;
      UN_CONV(1, 4)
      UN_CONV(2, 4)
      UN_CONV(3, 4)
      UN_CONV(4, 4)
      UN_CONV(6, 2)
      UN_CONV(8, 2)
#undef UN_CONV
      default:
        throw fail(PH_CLASS_EXECUTION, "E_HIP_INTERNAL", "conv: %d output channels", A.cout);
    }
    HIP_OK(hipGetLastError());
  }

  // Forward pass; the FP32 output goes to `o` ([batch][res][res]).
  void forward(const float *d_in, float *o, int batch, hipStream_t st) {
    if (batch < 1 || batch > max_batch)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "network batch %d outside [1, %d]", batch,
                 max_batch);
    const int nin = (int)(npix(0, batch) * UN_IN_C);
    hipLaunchKernelGGL(k_to_bf16, dim3(cdiv(nin, 256)), dim3(256), 0, st, d_in, act[UB_X], nin);
    for (int l = 0; l < UN_LAYERS; ++l) {
      const UnLayerDef &d = UN_DEF[l];
      const int r = res >> d.level;
      UnConvArgs A{};
      A.a = act[d.in_a];
      A.b = d.in_b >= 0 ? act[d.in_b] : nullptr;
      A.wgt = w16[l];
      A.bias = master + g_lay.off_b[l];
      A.gate = nullptr;
      A.out = act[d.out];
      A.ca = d.ca;
      A.cb = d.cb;
      A.kp = g_lay.kp[l];
      A.cout = d.cout;
      A.n = batch;
      A.h = r;
      A.w = r;
      A.relu = 1;
      A.dil = d.dil;
      conv(A, st);
      if (d.out == UB_E0 || d.out == UB_E1 || d.out == UB_E2) {
        const int pb = d.out == UB_E0 ? UB_P0 : d.out == UB_E1 ? UB_P1 : UB_P2;
        const int c = d.cout;
        const size_t n = npix(d.level + 1, batch) * (c / 4);
        hipLaunchKernelGGL(k_pool2, dim3(cdiv(n, 256)), dim3(256), 0, st, act[d.out], act[pb],
                           batch, r, r, c);
      }
    }
    const int np0 = (int)npix(0, batch);
    hipLaunchKernelGGL(k_head, dim3(cdiv(np0, 256)), dim3(256), 0, st, act[UB_D0],
                       master + g_lay.off_head, np0, o);
    HIP_OK(hipGetLastError());
  }

  void weight_grad(int l, const uint16_t *dy, int batch, hipStream_t st) {
    const UnLayerDef &d = UN_DEF[l];
    const int r = res >> d.level;
    UnDwArgs A{};
    A.dy = dy;
    A.a = act[d.in_a];
    A.b = d.in_b >= 0 ? act[d.in_b] : nullptr;
    A.ca = d.ca;
    A.cb = d.cb;
    A.part = part;
    A.part_db = part_db;
    A.cout = d.cout;
    A.kp = g_lay.kp[l];
    A.kr = g_lay.cp[l] >= 16 ? g_lay.cp[l] : g_lay.kp[l];
    A.n = batch;
    A.h = r;
    A.w = r;
    A.chunk = dw_chunk(l, batch);
    A.dil = d.dil;
    const int nchunk = (int)cdiv(npix(d.level, batch), A.chunk);
    hipLaunchKernelGGL(k_conv3_dw, dim3(nchunk, A.kp / A.kr), dim3(256), 0, st, A);
    const int nw = d.cout * A.kp + d.cout;
    hipLaunchKernelGGL(k_dw_reduce, dim3(cdiv(nw, 256)), dim3(256), 0, st, part, part_db, nchunk,
                       d.cout, A.kp, grad + g_lay.off_w[l], grad + g_lay.off_b[l]);
    HIP_OK(hipGetLastError());
  }

  // input gradient of layer l into `dst` (channels cp[l]), gated by `gate`
  void input_grad(int l, const uint16_t *dy, const uint16_t *gate, uint16_t *dst, int batch,
                  hipStream_t st) {
    const UnLayerDef &d = UN_DEF[l];
    const int r = res >> d.level;
    UnConvArgs A{};
    A.a = dy;
    A.b = nullptr;
    A.wgt = wt16[l];
    A.bias = nullptr;
    A.gate = gate;
    A.out = dst;
    A.ca = d.cout;
    A.cb = 0;
    A.kp = 9 * d.cout;
    A.cout = g_lay.cp[l];
    A.n = batch;
    A.h = r;
    A.w = r;
    A.relu = 0;
    A.dil = d.dil;
    conv(A, st);
  }

  // Loss and gradients for one batch after forward(); returns nothing (the
  // loss stays on the device in `loss`).
  void backward(const float *d_target, const float *d_weight, int batch, float alpha,
                hipStream_t st) {
    const int np0 = (int)npix(0, batch);
    const int nhb = (int)cdiv(np0, 256);
    if (alpha != 0.0f) {
      if (res % 16) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "pyramid loss needs res % 16 == 0");
      hipLaunchKernelGGL(k_err_pyramid, dim3(batch * (res / 16) * (res / 16)), dim3(256), 0, st,
                         out, d_target, d_weight, res, pyr);
    }
    hipLaunchKernelGGL(k_head_bwd, dim3(nhb), dim3(256), 0, st, out, d_target, d_weight,
                       act[UB_D0], master + g_lay.off_head, np0, 1.0f / (float)np0, pyr, res,
                       alpha, dz[13], head_part);
    hipLaunchKernelGGL(k_head_reduce, dim3(UN_HEAD_C + 2), dim3(256), 0, st, head_part, nhb,
                       grad + g_lay.off_head, loss);
    // decoder level 0
    weight_grad(13, dz[13], batch, st);
    input_grad(13, dz[13], act[UB_D0A], dz[12], batch, st);
    weight_grad(12, dz[12], batch, st);
    input_grad(12, dz[12], nullptr, dcat[0], batch, st);
    up_grad(dcat[0], 48, 16, UB_D1, 32, 0, dz[11], batch, st);
    // decoder level 1
    weight_grad(11, dz[11], batch, st);
    input_grad(11, dz[11], act[UB_D1A], dz[10], batch, st);
    weight_grad(10, dz[10], batch, st);
    input_grad(10, dz[10], nullptr, dcat[1], batch, st);
    up_grad(dcat[1], 96, 32, UB_D2, 64, 1, dz[9], batch, st);
    // decoder level 2
    weight_grad(9, dz[9], batch, st);
    input_grad(9, dz[9], act[UB_D2A], dz[8], batch, st);
    weight_grad(8, dz[8], batch, st);
    input_grad(8, dz[8], nullptr, dcat[2], batch, st);
    up_grad(dcat[2], 128, 64, UB_B3, 64, 2, dz[7], batch, st);
    // bottleneck
    weight_grad(7, dz[7], batch, st);
    input_grad(7, dz[7], act[UB_A3], dz[6], batch, st);
    weight_grad(6, dz[6], batch, st);
    input_grad(6, dz[6], nullptr, dpool[2], batch, st);
    skip_grad(dcat[2], 128, dpool[2], UB_E2, 64, 2, dz[5], batch, st);
    // encoder level 2
    weight_grad(5, dz[5], batch, st);
    input_grad(5, dz[5], act[UB_A2], dz[4], batch, st);
    weight_grad(4, dz[4], batch, st);
    input_grad(4, dz[4], nullptr, dpool[1], batch, st);
    skip_grad(dcat[1], 96, dpool[1], UB_E1, 32, 1, dz[3], batch, st);
    // encoder level 1
    weight_grad(3, dz[3], batch, st);
    input_grad(3, dz[3], act[UB_A1], dz[2], batch, st);
    weight_grad(2, dz[2], batch, st);
    input_grad(2, dz[2], nullptr, dpool[0], batch, st);
    skip_grad(dcat[0], 48, dpool[0], UB_E0, 16, 0, dz[1], batch, st);
    // encoder level 0
    weight_grad(1, dz[1], batch, st);
    input_grad(1, dz[1], act[UB_A0], dz[0], batch, st);
    weight_grad(0, dz[0], batch, st);
    HIP_OK(hipGetLastError());
  }

  void up_grad(const uint16_t *dc, int ccat, int c1, int prev, int c2, int level, uint16_t *dst,
               int batch, hipStream_t st) {
    const int r = res >> level;
    const size_t n = npix(level + 1, batch) * (c2 / 4);
    hipLaunchKernelGGL(k_up_grad, dim3(cdiv(n, 256)), dim3(256), 0, st, dc, ccat, c1, act[prev],
                       c2, batch, r, r, dst);
  }
  void skip_grad(const uint16_t *dc, int ccat, const uint16_t *dp, int e, int c1, int level,
                 uint16_t *dst, int batch, hipStream_t st) {
    const int r = res >> level;
    const size_t n = npix(level, batch) * (c1 / 4);
    hipLaunchKernelGGL(k_skip_grad, dim3(cdiv(n, 256)), dim3(256), 0, st, dc, ccat, dp, act[e],
                       c1, batch, r, r, dst);
  }

  void adam(const ph_train_opts &o, hipStream_t st) {
    ++step;
    const float bc1 = 1.0f - powf(o.beta1, (float)step);
    const float bc2 = 1.0f - powf(o.beta2, (float)step);
    const int n = (int)g_lay.storage;
    hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, st, master, grad, m, v, n,
                       o.learning_rate, o.beta1, o.beta2, o.epsilon, bc1, bc2, o.weight_decay);
    cast_weights(st);
    HIP_OK(hipGetLastError());
  }
};

// The network the solver's warm start runs (set by ph_warm_start_set_weights).
static std::unique_ptr<UnNet> g_warm;
static std::vector<float> g_warm_params;
static const int WARM_RES = 128;
static const int WARM_BATCH = 64;

bool warm_net_ready() { return !g_warm_params.empty(); }
int warm_net_res() { return WARM_RES; }
int warm_net_batch() { return WARM_BATCH; }

// The solver's forward passes as hipGraphs: its ~20 small launches (about
// 10 us each at the solver's batches) replayed from one graph launch. A key
// (input, output, batch) runs directly the first time and is captured the
// second (a batch of many images sees many keys once); the graphs hold kernel
// nodes only, captured on a private default-priority stream. They refer to
// the network's buffers, so they go when the network does.
// PAINTOP_HIP_GRAPH=0 runs every pass directly.
struct NetGraph {
  const float *in;
  float *out;
  int batch;
  hipGraphExec_t exec;
};
static std::vector<NetGraph> g_net_graphs;
static hipStream_t g_net_cs = nullptr;
static void net_graphs_clear() {
  if (!g_net_graphs.empty()) (void)hipDeviceSynchronize();
  for (NetGraph &g : g_net_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g_net_graphs.clear();
}

// Inference for the solver: d_in [batch][res][res][4] -> d_out [batch][res][res].
void warm_net_forward(const float *d_in, float *d_out, int batch, hipStream_t st) {
  if (g_warm_params.empty())
    throw fail(PH_CLASS_POLICY, "E_HIP_WARM_START_UNSET",
               "warm start requested but no network weights are set (ph_warm_start_set_weights)");
  if (!g_warm) {
    net_graphs_clear();
    auto n = std::make_unique<UnNet>();
    n->create(WARM_BATCH, WARM_RES, g_warm_params.data(), 0);
    g_warm = std::move(n);
  }
  const char *ge = getenv("PAINTOP_HIP_GRAPH");
  if (ge && ge[0] == '0') {
    g_warm->forward(d_in, d_out, batch, st);
    return;
  }
  NetGraph *e = nullptr;
  for (NetGraph &g : g_net_graphs)
    if (g.in == d_in && g.out == d_out && g.batch == batch) e = &g;
  if (!e) {
    if (g_net_graphs.size() >= 32) {
      HIP_OK(hipStreamSynchronize(st));
      net_graphs_clear();
    }
    g_net_graphs.push_back(NetGraph{d_in, d_out, batch, nullptr});
    g_warm->forward(d_in, d_out, batch, st);
    return;
  }
  if (!e->exec) {
    if (!g_net_cs) HIP_OK(hipStreamCreateWithFlags(&g_net_cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(g_net_cs, hipStreamCaptureModeRelaxed));
    try {
      g_warm->forward(d_in, d_out, batch, g_net_cs);
    } catch (...) {
      (void)hipStreamEndCapture(g_net_cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_OK(hipStreamEndCapture(g_net_cs, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t r = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(r);
    e->exec = ex;
  }
  HIP_OK(hipGraphLaunch(e->exec, st));
}

}  // namespace ph

using namespace ph;

struct ph_unet {
  UnNet net;
};

template <class F>
static ph_status un_guarded(ph_error *err, F &&body) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
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

extern "C" uint64_t ph_unet_param_storage(void) { return g_lay.storage; }
extern "C" uint64_t ph_unet_param_count(void) { return g_lay.count; }

extern "C" ph_status ph_unet_create(int32_t max_batch, int32_t res, const float *params,
                                    uint64_t seed, ph_unet **net, ph_error *err) {
  return un_guarded(err, [&]() {
    if (!net) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null network handle");
    *net = nullptr;
    auto h = std::make_unique<ph_unet>();
    h->net.create(max_batch, res, params, seed);
    *net = h.release();
  });
}

extern "C" void ph_unet_destroy(ph_unet *net) {
  std::lock_guard<std::mutex> lk(g_mu);
  delete net;
}

extern "C" ph_status ph_unet_forward(ph_unet *net, const float *d_in, float *d_out, int32_t batch,
                                     void *stream, ph_error *err) {
  return un_guarded(err, [&]() {
    if (!net || !d_in || !d_out) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null argument");
    hipStream_t st = (hipStream_t)stream;
    net->net.forward(d_in, net->net.out, batch, st);
    HIP_OK(hipMemcpyAsync(d_out, net->net.out,
                          net->net.npix(0, batch) * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
  });
}

extern "C" ph_status ph_unet_train_step(ph_unet *net, const float *d_in, const float *d_target,
                                        const float *d_weight, int32_t batch,
                                        const ph_train_opts *opts, double *loss, void *stream,
                                        ph_error *err) {
  return un_guarded(err, [&]() {
    if (!net || !d_in || !d_target || !d_weight || !opts)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null argument");
    hipStream_t st = (hipStream_t)stream;
    UnNet &n = net->net;
    n.forward(d_in, n.out, batch, st);
    if (!(opts->pyramid_weight >= 0.0f))
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "pyramid_weight must be >= 0");
    n.backward(d_target, d_weight, batch, opts->pyramid_weight, st);
    if (opts->apply_update) n.adam(*opts, st);
    if (loss) HIP_OK(hipMemcpyAsync(loss, n.loss, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  });
}

extern "C" ph_status ph_unet_get_state(ph_unet *net, float *params, float *grads, float *adam_m,
                                       float *adam_v, int32_t *step, ph_error *err) {
  return un_guarded(err, [&]() {
    if (!net) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null network handle");
    UnNet &n = net->net;
    const size_t bytes = g_lay.storage * sizeof(float);
    HIP_OK(hipDeviceSynchronize());
    if (params) HIP_OK(hipMemcpy(params, n.master, bytes, hipMemcpyDeviceToHost));
    if (grads) HIP_OK(hipMemcpy(grads, n.grad, bytes, hipMemcpyDeviceToHost));
    if (adam_m) HIP_OK(hipMemcpy(adam_m, n.m, bytes, hipMemcpyDeviceToHost));
    if (adam_v) HIP_OK(hipMemcpy(adam_v, n.v, bytes, hipMemcpyDeviceToHost));
    if (step) *step = n.step;
  });
}

extern "C" ph_status ph_unet_set_state(ph_unet *net, const float *params, const float *adam_m,
                                       const float *adam_v, int32_t step, ph_error *err) {
  return un_guarded(err, [&]() {
    if (!net || !params) throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null argument");
    check_padding(params);
    UnNet &n = net->net;
    const size_t bytes = g_lay.storage * sizeof(float);
    HIP_OK(hipMemcpy(n.master, params, bytes, hipMemcpyHostToDevice));
    if (adam_m) HIP_OK(hipMemcpy(n.m, adam_m, bytes, hipMemcpyHostToDevice));
    if (adam_v) HIP_OK(hipMemcpy(n.v, adam_v, bytes, hipMemcpyHostToDevice));
    n.step = step;
    n.cast_weights(nullptr);
    HIP_OK(hipDeviceSynchronize());
  });
}

extern "C" ph_status ph_unet_gather_device(const float *d_in_all, const float *d_target_all,
                                           const float *d_weight_all, uint64_t n_samples,
                                           const int32_t *index, const int32_t *mode,
                                           int32_t batch, int32_t res, float *d_in,
                                           float *d_target, float *d_weight, void *stream,
                                           ph_error *err) {
  return un_guarded(err, [&]() {
    if (!d_in_all || !d_target_all || !d_weight_all || !index || !d_in || !d_target || !d_weight)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "null argument");
    if (batch < 1 || batch > UN_GATHER_MAX || res < 1 || res > 4096)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "gather: batch %d (at most %d), res %d", batch,
                 UN_GATHER_MAX, res);
    UnGather G{};
    for (int j = 0; j < batch; ++j) {
      if (index[j] < 0 || (uint64_t)index[j] >= n_samples)
        throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "gather: index %d outside [0, %llu)",
                   index[j], (unsigned long long)n_samples);
      G.idx[j] = index[j];
      G.mode[j] = mode ? (mode[j] & 7) : 0;
    }
    hipLaunchKernelGGL(k_gather_dihedral, dim3(res, batch), dim3(std::min(res, 256)), 0,
                       (hipStream_t)stream, d_in_all, d_target_all, d_weight_all, G, res, d_in,
                       d_target, d_weight);
    HIP_OK(hipGetLastError());
  });
}

extern "C" ph_status ph_warm_start_set_weights(const float *params, uint64_t n_params,
                                               ph_error *err) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_error(err);
  try {
    if (params == nullptr || n_params == 0) {  // clear
      net_graphs_clear();
      g_warm.reset();
      g_warm_params.clear();
      return PH_OK;
    }
    if (n_params != g_lay.storage)
      throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT",
                 "warm-start weights: %llu values, the network stores %llu",
                 (unsigned long long)n_params, (unsigned long long)g_lay.storage);
    check_padding(params);
    for (uint64_t i = 0; i < n_params; ++i)
      if (!std::isfinite(params[i]))
        throw fail(PH_CLASS_SCHEMA, "E_HIP_ARGUMENT", "warm-start weights: value %llu not finite",
                   (unsigned long long)i);
    g_warm_params.assign(params, params + n_params);
    net_graphs_clear();
    g_warm.reset();
    return PH_OK;
  } catch (const Fail &f) {
    put_error(err, f);
    return PH_ERROR;
  }
}
