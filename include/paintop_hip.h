/*
 * paintop_hip.h -- C ABI of libpaintop_hip, the MI355X (gfx950) backend for
 * paintop's Poisson / screened-Poisson solver ops.
 *
 * The library replaces, behind paintop's existing plugin boundary, the compute
 * of the two ops that share the solver core:
 *   repair.poisson_blend@1     PoissonBlend::compute
 *                              crates/paintop-cpu/src/poisson_blend.rs:535-600
 *   repair.screened_poisson@1  ScreenedPoisson::compute
 *                              crates/paintop-cpu/src/screened_poisson.rs:362-430
 * both of which reach blend_channels (poisson_blend.rs:168-236) and
 * poisson::solve (crates/paintop-cpu/src/poisson.rs:168-240). A host-side
 * OpImplementation (paintop-core/src/executor/op_impl.rs:51-70) resolves the
 * plan params exactly as resolve_controls does (poisson_blend.rs:102-139,
 * screened_poisson.rs:66-112) and calls ph_poisson_solve with plain values.
 *
 * Conventions: plain pointers and sizes only; no allocation ownership crosses
 * the boundary; no exception crosses it. Every entry point returns a
 * ph_status and, on failure, fills the caller's ph_error with the paintop
 * ErrorClass ordinal (paintop-ir/src/error.rs:23-37), a stable code and a
 * message. Every entry point is reentrant: one lazily created device context
 * sits behind an internal mutex, so concurrent callers serialise.
 */
#ifndef PAINTOP_HIP_H
#define PAINTOP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PH_ABI_VERSION 2

typedef enum ph_status {
  PH_OK = 0,
  PH_ERROR = 1,       /* see ph_error */
  PH_UNAVAILABLE = 2, /* no usable gfx950 device (class Policy, E_GPU_UNAVAILABLE) */
} ph_status;

/* paintop_ir::ErrorClass, in declaration order (error.rs:23-37). */
typedef enum ph_error_class {
  PH_CLASS_PARSE = 0,
  PH_CLASS_SCHEMA = 1,
  PH_CLASS_REFERENCE = 2,
  PH_CLASS_TYPE = 3,
  PH_CLASS_SEMANTIC = 4,
  PH_CLASS_POLICY = 5,
  PH_CLASS_EXECUTION = 6,
} ph_error_class;

/* paintop_ir::Error (error.rs:210-219) flattened. */
typedef struct ph_error {
  int32_t error_class; /* ph_error_class */
  char code[64];       /* e.g. E_POISSON_BLEND_PARAM, E_GPU_UNAVAILABLE, E_HIP_RUNTIME */
  char message[512];
} ph_error;

/* The op whose contract a solve follows: it fixes the report kind
 * ("poisson" / "screened-poisson") and the error-code family. */
typedef enum ph_op {
  PH_OP_POISSON_BLEND = 0,    /* repair.poisson_blend@1: lambda must be 0 */
  PH_OP_SCREENED_POISSON = 1, /* repair.screened_poisson@1 */
} ph_op;

/* SolverStopReason (paintop-ir/src/resource.rs:1186-1196). */
typedef enum ph_stop_reason {
  PH_STOP_CONVERGED = 0,
  PH_STOP_MAX_ITERATIONS = 1,
  PH_STOP_STALLED = 2,
} ph_stop_reason;

/* Solve flags. */
#define PH_FLAG_TIMING 0x1u         /* fill ph_report.device_ms with hipEvent timing */
#define PH_FLAG_WARM_START_NET 0x2u /* try the learned warm start (falls back to the plain start) */
#define PH_FLAG_FORCE_PLAIN 0x4u    /* never use a warm start */

/* Descriptor of one solve (one image, all of its channels). */
typedef struct ph_solve_desc {
  uint32_t width;
  uint32_t height;
  uint32_t channels;       /* interleaved channel count (Gray 1 .. RGBA 4) */
  uint32_t max_iterations; /* 1..=100000 (poisson_blend.rs:71) */
  double tolerance;        /* (0, 1) */
  double lambda;           /* >= 0, finite; 0 for PH_OP_POISSON_BLEND */
  double omega;            /* fixed SOR factor in (0, 2); 0 = tuned per instance from the mask's coarse spectral radius */
  uint32_t flags;          /* PH_FLAG_* */
  uint32_t half_sweeps;    /* red/black half-sweeps per HBM round trip; 0 = tuned */
  int32_t op;              /* ph_op */
  uint32_t tile_rows;      /* rows per sweep tile (even, 16..1024); 0 = tuned */
} ph_solve_desc;

/* SolveReport (poisson.rs:126-141) of the worst channel (poisson_blend.rs:240-246),
 * plus measurement side data that stays outside the op's Report schema. */
typedef struct ph_report {
  uint32_t iterations;
  int32_t stop_reason; /* ph_stop_reason */
  int32_t converged;
  uint32_t history_len;
  double tolerance;
  double final_residual;
  double *residual_history;  /* caller-owned, capacity history_capacity (>= max_iterations) */
  uint32_t history_capacity; /* 0 with residual_history NULL: history not returned */
  uint32_t worst_channel;
  /* side data (not part of the paintop Report) */
  uint32_t refinements;      /* FP64 outer refinement steps, all channels */
  uint32_t launches;         /* sweep-kernel launches */
  uint64_t unknowns;         /* interior cells summed over channels */
  uint64_t unknown_sweeps;   /* sum over channels of interior cells x sweeps */
  double device_ms;          /* solve time on the device stream (PH_FLAG_TIMING) */
  double omega_used;
  uint32_t half_sweeps_used;
  uint32_t warm_start_used;  /* channels that kept the learned start (0: plain start) */
  double warm_start_rel;     /* worst channel's ||r(warm start)|| / ||r(anchor start)||, 0 if not tried */
  double sweep_ms;           /* hipEvent time of the sweep-kernel batches (PH_FLAG_TIMING) */
  uint32_t coarse_corrections; /* learned coarse corrections kept, all channels */
  uint32_t reserved_;
} ph_report;

typedef struct ph_device_info {
  char name[128];
  char arch[64]; /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int32_t compute_units;
  int32_t device_index;
  int32_t runtime_version;
  int32_t driver_version;
  uint64_t hbm_bytes;
} ph_device_info;

/* Result of assert.no_change_outside_mask@1's scan (the assertion's metrics,
 * crates/paintop-cpu/src/assert.rs:405-446 and :690-710). */
#define PH_MAX_LOCATIONS 256 /* assert.rs:72 */
typedef struct ph_outside_scan {
  uint64_t changed_outside; /* outside pixels whose delta strictly exceeds the threshold */
  double max_outside;       /* largest delta over outside pixels, 0 if none */
  int64_t worst_x, worst_y; /* first pixel with the largest leaking delta; -1 if no leak */
  double worst_delta;       /* its delta, 0 if no leak */
  uint32_t n_locations;     /* min(changed_outside, PH_MAX_LOCATIONS) */
  uint32_t reserved_;
  int64_t locations[PH_MAX_LOCATIONS][2]; /* first leaking pixels, row-major order, (x, y) */
} ph_outside_scan;

/* Library ABI version (PH_ABI_VERSION). */
int32_t ph_abi_version(void);

/* Device probe. Replaces the adapter probe of the GPU backend template
 * (paintop-wgpu/src/gpu/probe.rs:95-101, 209-230): PH_UNAVAILABLE with class
 * Policy / E_GPU_UNAVAILABLE when no gfx950 device is usable, or when the
 * probe is forced unavailable (ph_force_unavailable, or the environment
 * variable PAINTOP_HIP_FORCE_UNAVAILABLE=1), mirroring probe_forced(true)
 * (probe.rs:223). */
ph_status ph_probe(ph_device_info *out, ph_error *err);

/* Force the probe (and every solve) to report PH_UNAVAILABLE. For tests. */
void ph_force_unavailable(int32_t forced);

/* The sweep configuration (half-sweeps per launch, tile rows) a call with
 * half_sweeps = tile_rows = 0 takes for its work: the sum over its images of
 * channels x the area of the interior's bounding box. No counterpart in the
 * reference (its solver has no launch configuration); an executor that
 * gathers several nodes into one batched call passes each group's own
 * configuration, so every image is solved as it would be alone. Host-only:
 * no device needed. */
void ph_sweep_config(double work, uint32_t *half_sweeps, uint32_t *tile_rows);

/* One op invocation on host buffers in the ResourceValue layout
 * (paintop-core/src/executor/value.rs:35-85): row-major, channel-interleaved
 * f32, source/target (guidance/anchor) width*height*channels, mask
 * width*height. Writes out_candidate (width*height*channels) and report.
 * Replaces PoissonBlend::compute / ScreenedPoisson::compute after param
 * resolution (poisson_blend.rs:535-600, screened_poisson.rs:362-430). */
ph_status ph_poisson_solve(const ph_solve_desc *desc, const float *source, const float *target,
                           const float *mask, float *out_candidate, ph_report *report,
                           ph_error *err);

/* The same on device-resident buffers (hipMalloc'd, same layout), enqueued on
 * `stream` (a hipStream_t, NULL for the default stream); returns when the
 * result is in out_candidate. */
ph_status ph_poisson_solve_device(const ph_solve_desc *desc, const float *d_source,
                                  const float *d_target, const float *d_mask,
                                  float *d_out_candidate, void *stream, ph_report *report,
                                  ph_error *err);

/* Device workspace of ph_poisson_solve_device_noalloc (128 MiB): enough for a
 * 512x512 RGBA or a 1024x1024 single-channel solve. */
#define PH_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* ph_poisson_solve_device without any allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory during a solve
 * (serving loops, harnesses that forbid it): no hipMalloc, hipFree,
 * hipHostMalloc or graph instantiation happens in the call. The solve's
 * workspace is a fixed device array of PH_NOALLOC_SCRATCH_BYTES that each GPU
 * the library runs on holds until the process ends; a solve that needs more
 * is rejected before any launch (class Policy, E_HIP_SCRATCH_LIMIT). The
 * plain solve only: PH_FLAG_WARM_START_NET and PH_FLAG_TIMING are rejected
 * (class Schema, E_HIP_NOALLOC_FLAG). Results are bit-identical to
 * ph_poisson_solve_device's. Every call on a GPU shares that GPU's array,
 * which keeps the last call's contents (nothing in it is assumed zero); calls
 * that overlapped on one GPU would race on it, and the library's lock
 * serialises them (each call finishes its work on `stream` before it
 * returns). */
ph_status ph_poisson_solve_device_noalloc(const ph_solve_desc *desc, const float *d_source,
                                          const float *d_target, const float *d_mask,
                                          float *d_out_candidate, void *stream,
                                          ph_report *report, ph_error *err);

/* ph_poisson_solve_device from a caller's initial field d_start (same layout
 * as d_out_candidate; only its interior samples are read), e.g. the previous
 * frame's solution. No reference counterpart (the reference starts from the
 * anchor, poisson.rs:169-170). The start races a plain (anchor-start) twin
 * for a short probe and the smaller residual continues, exactly as the
 * learned warm start does; the relative residual keeps the anchor-start
 * denominator (poisson.rs:172). */
ph_status ph_poisson_solve_start_device(const ph_solve_desc *desc, const float *d_source,
                                        const float *d_target, const float *d_mask,
                                        const float *d_start, float *d_out_candidate,
                                        void *stream, ph_report *report, ph_error *err);

/* One image of a batched solve: device buffers in the ResourceValue layout. */
typedef struct ph_batch_item {
  uint32_t width;
  uint32_t height;
  uint32_t channels;
  uint32_t reserved;
  const float *source; /* or guidance */
  const float *target; /* or anchor */
  const float *mask;
  float *out_candidate;
} ph_batch_item;

/* Many independent op invocations solved together: every channel of every
 * image is one instance, all instances share the sweep launches, and each
 * instance stops on its own (so each report equals the one its image would
 * get alone). desc supplies the controls (op, max_iterations, tolerance,
 * lambda, omega, tuning); its width/height/channels are ignored. reports[i]
 * (with its own history buffer) receives image i's worst-channel report.
 * No reference counterpart: paintop dispatches nodes one at a time
 * (paintop-core/src/executor/mod.rs:175-193); this is the batching entry
 * point an executor-level batcher, or a caller with many masked regions,
 * would use. */
ph_status ph_poisson_solve_batch_device(const ph_solve_desc *desc, const ph_batch_item *items,
                                        uint32_t n_items, void *stream, ph_report *reports,
                                        ph_error *err);

/* ---- device-side input generators (inputs stay in HBM) ------------------ */

/* field.fbm@1 (crates/paintop-cpu/src/noise.rs:660-694, value formula
 * noise.rs:72-160, 218-240) evaluated at pixel centres with scale
 * frequency/64; writes fmaf((float)fbm, scale, bias) to
 * d_out[(y*width + x)*channel_stride + channel_offset]. */
ph_status ph_fbm_device(float *d_out, uint32_t width, uint32_t height, uint32_t channel_stride,
                        uint32_t channel_offset, double frequency, uint64_t seed,
                        uint32_t octaves, double lacunarity, double gain, float scale,
                        float bias, void *stream, ph_error *err);

/* mask.polygon@1's inside test (crates/paintop-cpu/src/mask_polygon.rs:155-195,
 * nonzero winding or even-odd) sampled once at each pixel centre: a binary
 * coverage mask. At most 64 vertices, points_xy = x0,y0,x1,y1,... */
ph_status ph_polygon_mask_device(float *d_out, uint32_t width, uint32_t height,
                                 const double *points_xy, uint32_t n_points, int32_t even_odd,
                                 void *stream, ph_error *err);

/* d_out[i] += sigma * N(0, 1) for i < n (seeded SplitMix64 + Box-Muller). */
ph_status ph_gauss_noise_device(float *d_out, uint64_t n, uint64_t seed, double sigma,
                                void *stream, ph_error *err);

/* Binary axis-aligned ellipse mask: 1 where ((x+.5-cx)/rx)^2 + ((y+.5-cy)/ry)^2 <= 1. */
ph_status ph_ellipse_mask_device(float *d_out, uint32_t width, uint32_t height, double cx,
                                 double cy, double rx, double ry, void *stream, ph_error *err);

/* ---- learned warm start ------------------------------------------------- *
 * No reference counterpart: paintop's solvers always start from the anchor
 * (poisson.rs:169-170). The warm start keeps the reference's relative-residual
 * denominator ||r(anchor start)|| (poisson.rs:172) and its stop rules.
 *
 * The network: a 4-level U-Net (3x3 convolutions, channels 16/32/64/64, BF16
 * weights and activations, FP32 accumulation on the matrix cores, FP32 1x1
 * head) on a res x res grid, input [batch][res][res][4] FP32 (channel 0: the
 * right-hand side of the correction problem, 1: boundary values (0 off the
 * Dirichlet ring), 2: interior fraction, 3: ring density), output
 * [batch][res][res] FP32.
 * Parameters are ph_unet_param_storage() FP32 values in the layout
 * documented in paintop_amd/csrc/unet_internal.h (padding slots must be 0). */

/* FP32 values of the parameter storage (padding included). */
uint64_t ph_unet_param_storage(void);
/* Learnable values (padding excluded). */
uint64_t ph_unet_param_count(void);

typedef struct ph_unet ph_unet; /* a network with its training state and buffers */

typedef struct ph_train_opts {
  float learning_rate;
  float beta1, beta2, epsilon; /* Adam */
  float weight_decay;          /* decoupled (AdamW) */
  int32_t apply_update;        /* 0: compute loss and gradients only */
  float pyramid_weight;        /* alpha of the multi-scale term (0: weighted MSE) */
} ph_train_opts;

/* Buffers for batches up to max_batch at resolution res (a multiple of 8 in
 * [16, 1024]); params = NULL draws a seeded He-normal initialisation. */
ph_status ph_unet_create(int32_t max_batch, int32_t res, const float *params, uint64_t seed,
                         ph_unet **net, ph_error *err);
void ph_unet_destroy(ph_unet *net);

/* d_out = net(d_in) for `batch` inputs (device buffers). */
ph_status ph_unet_forward(ph_unet *net, const float *d_in, float *d_out, int32_t batch,
                          void *stream, ph_error *err);

/* One training step: forward, loss, backward, and (if opts->apply_update) an
 * AdamW update. loss = (1/n) sum_p [w[p] d[p]^2 + alpha sum_s m_s(p)^2] with
 * d = out - target and m_s(p) the mean of w d over the 2x2, 4x4, 8x8 and
 * 16x16 block holding p (alpha = opts->pyramid_weight; res % 16 == 0 when
 * alpha > 0); the multi-scale term weighs the smooth error that costs SOR the
 * most sweeps. *loss (host) gets the loss before the update. */
ph_status ph_unet_train_step(ph_unet *net, const float *d_in, const float *d_target,
                             const float *d_weight, int32_t batch, const ph_train_opts *opts,
                             double *loss, void *stream, ph_error *err);

/* Training state to / from host arrays of ph_unet_param_storage() floats
 * (any pointer may be NULL); the gradients are those of the last step. */
ph_status ph_unet_get_state(ph_unet *net, float *params, float *grads, float *adam_m,
                            float *adam_v, int32_t *step, ph_error *err);
ph_status ph_unet_set_state(ph_unet *net, const float *params, const float *adam_m,
                            const float *adam_v, int32_t step, ph_error *err);

/* Training batch from a device-resident dataset of n_samples examples
 * (inputs [n][res][res][4], targets and weights [n][res][res]): example
 * index[j] (host array) under the dihedral map mode[j] (host array or NULL;
 * bit 2 transpose, bit 0 mirror x, bit 1 mirror y) becomes batch entry j,
 * for j < batch <= 256. */
ph_status ph_unet_gather_device(const float *d_in_all, const float *d_target_all,
                                const float *d_weight_all, uint64_t n_samples,
                                const int32_t *index, const int32_t *mode, int32_t batch,
                                int32_t res, float *d_in, float *d_target, float *d_weight,
                                void *stream, ph_error *err);

/* The weights the solver's warm start (PH_FLAG_WARM_START_NET) runs with;
 * params = NULL clears them. */
ph_status ph_warm_start_set_weights(const float *params, uint64_t n_params, ph_error *err);

/* assert.no_change_outside_mask@1 on device buffers: the scan of
 * NoChangeOutsideMask::compute (crates/paintop-cpu/src/assert.rs:630-710,
 * scan_outside :405-446) over `before`/`after` (width*height*channels f32,
 * channel-interleaved) and the `allowed` coverage mask (width*height f32).
 * A pixel is outside when its coverage is <= coverage_epsilon; it leaks when
 * its largest channel |after - before| (sRGB-decoded first when decode_srgb,
 * i.e. comparison_space "decoded-linear" on an sRGB image, assert.rs:677-679)
 * strictly exceeds outside_threshold. Both parameters must be finite and
 * >= 0 (E_ASSERT_PARAM, class Schema, as resolve_non_negative :157-176);
 * channels == 0 gives the empty scan. passed = (changed_outside == 0); the
 * severity and the report wrapping stay on the host. With decode_srgb the
 * power is taken in double and rounded once (within one f32 ulp of powf). */
ph_status ph_assert_no_change_outside_mask_device(const float *d_before, const float *d_after,
                                                  const float *d_mask, uint32_t width,
                                                  uint32_t height, uint32_t channels,
                                                  int32_t decode_srgb, double outside_threshold,
                                                  double coverage_epsilon, void *stream,
                                                  ph_outside_scan *out, ph_error *err);

/* The coarse-correction input the solver builds from an instance's plane
 * (no counterpart in the reference; exposed so the kernel can be checked on
 * its own): per f x f block of the region [x0, x0 + res*f) x [y0, y0 + res*f)
 * of image coordinates, in[..][0] = sum of g over interior cells,
 * in[..][1] = 0, in[..][2] = interior cells / f^2, in[..][3] = ring cells / f,
 * where interior = cls != 0 and ring = non-interior cells with an interior
 * 4-neighbour. The plane is pitch x ph (g f32, cls u8 row-major; plane cell
 * (px, py) is image cell (ox + px, oy + py)); rows 0 and ph - 1 and columns 0
 * and pw - 1 are padding (never read as cells), pitch is a multiple of 4, g is
 * 16-byte and cls 4-byte aligned. d_in is [res][res][4]. */
ph_status ph_coarsen_plane_device(const float *d_g, const uint8_t *d_cls, int32_t pitch,
                                  int32_t pw, int32_t ph, int32_t ox, int32_t oy, int32_t x0,
                                  int32_t y0, int32_t factor, int32_t res, float *d_in,
                                  void *stream, ph_error *err);

/* The same input from the state x + e of a plane between refinements (no
 * counterpart in the reference; the solver's correction rounds use it): g is
 * the FP32 copy of -r(x) taken at the last refinement and d_e the FP32
 * correction since (0 off the interior), so the state's residual is
 * r = nsum(e) - (4 + lambda) e - g, formed per cell in FP32 as the sweeps form
 * it, with nsum the sum of the four neighbours' e. in[..][0] = the sum of -r
 * over interior cells (= the sum of g for e = 0, bit for bit); channels 1..3
 * as above; d_rsq ([res][res] f64) receives each block's sum of r^2 over its
 * interior cells. Sums run in FP32 down a column's rows of a block, in FP64
 * across its columns. d_e has g's layout and alignment. */
ph_status ph_correction_input_device(const float *d_e, const float *d_g, const uint8_t *d_cls,
                                     int32_t pitch, int32_t pw, int32_t ph, int32_t ox,
                                     int32_t oy, int32_t x0, int32_t y0, int32_t factor,
                                     int32_t res, double lambda, float *d_in, double *d_rsq,
                                     void *stream, ph_error *err);

/* Net input for one channel of an image (and, when d_solution is given, the
 * training target): the region [x0, x0 + res*f) x [y0, y0 + res*f) of the
 * image in f x f blocks. With interior = (mask > 0.5 and not on the image
 * border) (poisson.rs:340-355), ring = non-interior cells with an interior
 * 4-neighbour (the Dirichlet data), per block:
 *   in[..][0] = sum over interior cells of rhs
 *   in[..][1] = mean over ring cells of (anchor - P), 0 without ring cells
 *   in[..][2] = interior cells / f^2
 *   in[..][3] = ring cells / f
 *   target    = mean over interior and ring cells of (solution - P)
 *   weight    = 1 with interior cells, 0.5 with ring cells only, else 0.
 * mode 0 (warm start): P = source, rhs = lambda*(P - anchor).
 * mode 1 (coarse correction): P = 0, rhs = the source's 5-point Laplacian
 * (poisson.rs:364-373); with a zero anchor this is the correction problem
 * Laplacian(e) = g, e = 0 on the ring, that the solver's coarse corrections
 * hand to the network. The solver divides channels 0 and 1 of each input by
 * their RMS over the grid and scales the prediction back; this entry point
 * returns them unscaled.
 * d_in is [res][res][4], d_target / d_weight [res][res]. */
ph_status ph_warm_coarsen_device(const float *d_source, const float *d_anchor,
                                 const float *d_mask, const float *d_solution, uint32_t width,
                                 uint32_t height, uint32_t channels, uint32_t channel,
                                 double lambda, int32_t x0, int32_t y0, int32_t factor,
                                 int32_t res, int32_t mode, float *d_in, float *d_target,
                                 float *d_weight, void *stream, ph_error *err);

#ifdef __cplusplus
}
#endif
#endif
