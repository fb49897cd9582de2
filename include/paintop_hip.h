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

#define PH_ABI_VERSION 1

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
  double omega;            /* fixed SOR factor in (0, 2); 0 = Chebyshev schedule to the tuned value */
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
  uint32_t warm_start_used;  /* 1 when the learned start was kept */
  double warm_start_rel;     /* ||r(warm start)|| / ||r(anchor start)||, 0 if unused */
  double sweep_ms;           /* hipEvent time of the sweep-kernel batches (PH_FLAG_TIMING) */
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

#ifdef __cplusplus
}
#endif
#endif
