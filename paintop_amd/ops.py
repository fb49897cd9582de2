"""The hip backend's implementations of the two Poisson ops.

``PoissonBlend`` and ``ScreenedPoisson`` have the reference ops' contract
(crates/paintop-cpu/src/poisson_blend.rs, crates/paintop-cpu/src/screened_poisson.rs):
the same input/output ports, the same plan params with the same defaults,
ranges, error codes and classes, the same candidate layout (the target's
descriptor) and the same report (``SolverData`` of the worst channel). The
compute runs on the GPU through the C ABI (``ph_poisson_solve``); there is no
CPU path.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import hip
from .errors import E_GPU_UNAVAILABLE, ErrorClass, PaintopError
from .values import (AssertionOutcome, ImageDescriptor, MaskDescriptor, Report, ResourceValue,
                     SolverData)

POISSON_BLEND_OP_ID = "repair.poisson_blend@1"
SCREENED_POISSON_OP_ID = "repair.screened_poisson@1"
HIP_IMPL_ID = "hip.poisson@1"

# poisson_blend.rs:53-78
E_POISSON_BLEND_INPUT = "E_POISSON_BLEND_INPUT"
E_POISSON_BLEND_SHAPE = "E_POISSON_BLEND_SHAPE"
E_POISSON_BLEND_PARAM = "E_POISSON_BLEND_PARAM"
E_POISSON_BLEND_BUFFER = "E_POISSON_BLEND_BUFFER"
DEFAULT_MAX_ITERATIONS = 2000
MAX_ITERATIONS_LIMIT = 100_000
DEFAULT_TOLERANCE = 1e-6
# screened_poisson.rs:50-63
E_SCREENED_POISSON_INPUT = "E_SCREENED_POISSON_INPUT"
E_SCREENED_POISSON_PARAM = "E_SCREENED_POISSON_PARAM"
E_SCREENED_POISSON_BUFFER = "E_SCREENED_POISSON_BUFFER"
DEFAULT_LAMBDA = 0.0

_U32_MAX = 2**32 - 1
_MANIFEST_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "manifests")


def load_manifest(op_id: str) -> dict:
    """The op manifest with ``hip.poisson@1`` declared beside ``cpu.reference@1``."""
    with open(os.path.join(_MANIFEST_DIR, op_id + ".json"), encoding="utf-8") as f:
        return json.load(f)


# --- serde_json accessors -------------------------------------------------------


def _as_u64(v):
    """serde_json ``Value::as_u64``: only a non-negative JSON integer."""
    if isinstance(v, bool) or not isinstance(v, int) or v < 0 or v > 2**64 - 1:
        return None
    return v


def _as_f64(v):
    """serde_json ``Value::as_f64``: any JSON number."""
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        return None
    return float(v)


def _param_error(code: str, message: str, value) -> PaintopError:
    return PaintopError(ErrorClass.SCHEMA, code, message, actual=json.dumps(value))


def _resolve_iterations(params: dict, code: str, prefix: str) -> int:
    if "max_iterations" not in params:
        return DEFAULT_MAX_ITERATIONS
    value = params["max_iterations"]
    n = _as_u64(value)
    if n is None or n > _U32_MAX:
        raise _param_error(code, f"{prefix}max_iterations must be a positive integer", value)
    if n == 0 or n > MAX_ITERATIONS_LIMIT:
        raise _param_error(code, f"{prefix}max_iterations must be in 1..={MAX_ITERATIONS_LIMIT}",
                           value)
    return n


def _resolve_tolerance(params: dict, code: str, prefix: str) -> float:
    if "tolerance" not in params:
        return DEFAULT_TOLERANCE
    value = params["tolerance"]
    t = _as_f64(value)
    if t is None:
        raise _param_error(code, f"{prefix}tolerance must be a number", value)
    if not math.isfinite(t) or t <= 0.0 or t >= 1.0:
        raise _param_error(code, f"{prefix}tolerance must be in (0, 1)", value)
    return t


def resolve_blend_controls(params: dict) -> tuple:
    """poisson_blend.rs:102-139 -> (max_iterations, tolerance, lambda=0)."""
    params = params or {}
    pre = "repair.poisson_blend: "
    n = _resolve_iterations(params, E_POISSON_BLEND_PARAM, pre)
    t = _resolve_tolerance(params, E_POISSON_BLEND_PARAM, pre)
    return n, t, 0.0


def resolve_screened_controls(params: dict) -> tuple:
    """screened_poisson.rs:66-112 -> (max_iterations, tolerance, lambda)."""
    params = params or {}
    pre = "repair.screened_poisson: "
    lam = DEFAULT_LAMBDA
    if "lambda" in params:
        value = params["lambda"]
        lv = _as_f64(value)
        if lv is None:
            raise _param_error(E_SCREENED_POISSON_PARAM, pre + "lambda must be a number", value)
        if not math.isfinite(lv) or lv < 0.0:
            raise _param_error(E_SCREENED_POISSON_PARAM,
                               pre + "lambda must be a finite value >= 0", value)
        lam = lv
    n = _resolve_iterations(params, E_SCREENED_POISSON_PARAM, pre)
    t = _resolve_tolerance(params, E_SCREENED_POISSON_PARAM, pre)
    return n, t, lam


def check_shapes(source: ImageDescriptor, target: ImageDescriptor, mask_extent, op: str):
    """poisson_blend.rs:417-445 (both ops use E_POISSON_BLEND_SHAPE)."""
    if source.extent != target.extent:
        raise PaintopError(ErrorClass.SEMANTIC, E_POISSON_BLEND_SHAPE,
                           f"{op}: the source and target images must share an extent",
                           actual=f"source {source.extent} vs target {target.extent}")
    if source.layout != target.layout:
        raise PaintopError(ErrorClass.SEMANTIC, E_POISSON_BLEND_SHAPE,
                           f"{op}: the source and target images must share a channel layout",
                           actual=f"source {source.layout} vs target {target.layout}")
    if mask_extent != target.extent:
        raise PaintopError(ErrorClass.SEMANTIC, E_POISSON_BLEND_SHAPE,
                           f"{op}: the mask must share the images' extent",
                           actual=f"mask {mask_extent} vs image {target.extent}")
    return target


def interior_bbox_area(mask: np.ndarray) -> int:
    """Area of the bounding box of the interior cells (mask > 0.5 and not on
    the image border, poisson.rs:329-355): the work measure of the solver's
    sweep configuration, per channel."""
    h, w = mask.shape
    inside = np.asarray(mask[1:h - 1, 1:w - 1]) > 0.5
    ys = np.flatnonzero(inside.any(axis=1))
    xs = np.flatnonzero(inside.any(axis=0))
    if len(ys) == 0:
        return 0
    return int((ys[-1] - ys[0] + 1) * (xs[-1] - xs[0] + 1))


def solver_data(kind: str, r: hip.SolveResult) -> SolverData:
    """poisson_blend.rs:249-266."""
    return SolverData(kind=kind, steps=r.iterations, stability_number=0.0, stability_limit=0.0,
                      stable=True, residual_history=list(r.residual_history), total_energy=0.0,
                      iterations=r.iterations, stop_reason=r.stop_reason,
                      converged=r.converged, tolerance=r.tolerance,
                      final_residual=r.final_residual)


class _HipPoissonOp:
    """Shared compute of the two ops (the blend_channels role)."""

    op_id = ""
    op_name = ""
    ports = ("source", "target", "mask")
    input_code = ""
    buffer_code = ""
    kind = ""
    ph_op = hip.PH_OP_POISSON_BLEND

    def __init__(self, omega: float = 0.0, half_sweeps: int = 0, warm_start: bool = False,
                 split_components: bool = False):
        # SOR factor and half-sweeps per launch: tuned when 0. These are backend
        # tuning knobs, not plan params (the reference fixes omega = 1.8 in its
        # own backend, poisson_blend.rs:78). warm_start: the learned warm start
        # (the shipped network weights, raced against the anchor start; the
        # report keeps the reference's fields and denominator).
        self.omega = omega
        self.half_sweeps = half_sweeps
        self.warm_start = warm_start
        # split_components: solve each 4-connected component of the interior
        # on its own bounding box, all in one batched call (components.py);
        # the report is the worst component's
        self.split_components = split_components
        self.last_components: list = []
        if warm_start:
            from . import warmstart

            warmstart.enable()
        self.last_result: hip.SolveResult | None = None

    @classmethod
    def manifest(cls) -> dict:
        return load_manifest(cls.op_id)

    def _resolve(self, params):
        raise NotImplementedError

    def _input(self, inputs, port):
        if port not in inputs:
            raise PaintopError(ErrorClass.REFERENCE, self.input_code,
                               f"{self.op_name} requires a `{port}` input value")
        return inputs[port]

    def _prepare(self, inputs: dict, params: dict):
        """Resolved controls, the three input values and the output descriptor."""
        max_it, tol, lam = self._resolve(params)
        a, b, m = (self._input(inputs, p) for p in self.ports)
        if not isinstance(a.descriptor, ImageDescriptor):
            raise PaintopError(ErrorClass.TYPE, self.input_code,
                               f"{self.op_name} `{self.ports[0]}` input must be a image resource")
        if not isinstance(b.descriptor, ImageDescriptor):
            raise PaintopError(ErrorClass.TYPE, self.input_code,
                               f"{self.op_name} `{self.ports[1]}` input must be a image resource")
        if not isinstance(m.descriptor, MaskDescriptor):
            raise PaintopError(ErrorClass.TYPE, self.input_code,
                               f"{self.op_name} `mask` input must be a mask resource")
        out_desc = check_shapes(a.descriptor, b.descriptor, m.descriptor.extent, self.op_name)
        return (max_it, tol, lam), (a, b, m), out_desc

    def _outputs(self, out: np.ndarray, res: "hip.SolveResult", out_desc) -> dict:
        self.last_result = res
        samples = out.reshape(-1)
        all_finite = bool(np.isfinite(samples).all())
        try:
            candidate = ResourceValue(out_desc, out_desc.channels, samples)
        except PaintopError:
            raise PaintopError(ErrorClass.EXECUTION, self.buffer_code,
                               f"{self.op_name} produced a candidate buffer of unexpected length "
                               f"{samples.size}") from None
        report = Report(extent=out_desc.extent, channels=out_desc.channels, all_finite=all_finite,
                        solver=solver_data(self.kind, res))
        return {"candidate": candidate, "report": ResourceValue.report_value(report)}

    def batch_key(self, inputs: dict, params: dict):
        """What a node must share with others to be solved in one batched call
        with each image's result the one it gets alone: the resolved controls,
        and the sweep configuration its own work takes (an image's result
        depends on its launch boundaries). None: this node is solved alone (the
        learned start gates on the call's largest image; split components)."""
        if self.warm_start or self.split_components:
            return None
        controls, (_, _, m), out_desc = self._prepare(inputs, params)
        k, hb = (self.half_sweeps, 0) if self.half_sweeps else hip.sweep_config(
            out_desc.channels * interior_bbox_area(m.array()[..., 0]))
        return (self.op_id, controls, self.omega, k, hb)

    def compute_batch(self, calls: list) -> list:
        """The nodes of one batch key in one batched call
        (ph_poisson_solve_batch_device), each with the outputs its own compute
        gives."""
        from .device import DeviceArray

        keys = {self.batch_key(i, p) for i, p in calls}
        if len(keys) != 1 or None in keys:
            raise PaintopError(ErrorClass.EXECUTION, "E_HIP_BATCH",
                               f"{self.op_name}: nodes of different batch keys in one call")
        (_, (max_it, tol, lam), omega, k, hb), = keys
        prepared = [self._prepare(i, p) for i, p in calls]
        arrays, items = [], []
        try:
            for _, (a, b, m), od in prepared:
                ext = od.extent
                s_ = DeviceArray.from_numpy(np.ascontiguousarray(a.array(), dtype=np.float32))
                t_ = DeviceArray.from_numpy(np.ascontiguousarray(b.array(), dtype=np.float32))
                m_ = DeviceArray.from_numpy(np.ascontiguousarray(m.array()[..., 0],
                                                                 dtype=np.float32))
                o_ = DeviceArray((ext.height, ext.width, od.channels), np.float32)
                arrays += [s_, t_, m_, o_]
                items.append((ext.width, ext.height, od.channels, s_.ptr, t_.ptr, m_.ptr, o_.ptr))
            desc = hip.make_desc(0, 0, 1, max_it, tol, lam, op=self.ph_op, omega=omega,
                                 half_sweeps=k, tile_rows=hb)
            try:
                results = hip.poisson_solve_batch_device(desc, items)
            except hip.HipCallError as e:
                raise PaintopError(ErrorClass(e.error_class), e.code, e.message) from None
            outs = [arrays[4 * j + 3].numpy() for j in range(len(calls))]
        finally:
            for arr in arrays:
                arr.free()
        return [self._outputs(o, r, od) for o, r, (_, _, od) in zip(outs, results, prepared)]

    def compute(self, inputs: dict, params: dict) -> dict:
        """OpImplementation::compute (op_impl.rs:51-70) for this op."""
        (max_it, tol, lam), (a, b, m), out_desc = self._prepare(inputs, params)
        channels = out_desc.channels
        ext = out_desc.extent
        desc = hip.make_desc(ext.width, ext.height, channels, max_it, tol, lam, op=self.ph_op,
                             omega=self.omega, half_sweeps=self.half_sweeps,
                             flags=hip.PH_FLAG_WARM_START_NET if self.warm_start else 0)
        try:
            res = None
            if self.split_components:
                from . import components

                out, res, self.last_components = components.solve_components(
                    desc, a.array(), b.array(), m.array()[..., 0])
            if res is None:  # not split, or no interior cell at all
                out, res = hip.poisson_solve(desc, a.samples, b.samples, m.samples)
        except hip.HipCallError as e:
            cls = ErrorClass(e.error_class)
            raise PaintopError(cls, e.code, e.message) from None
        return self._outputs(out, res, out_desc)


class PoissonBlend(_HipPoissonOp):
    """repair.poisson_blend@1 on the GPU (poisson_blend.rs:535-600)."""

    op_id = POISSON_BLEND_OP_ID
    op_name = "repair.poisson_blend"
    ports = ("source", "target", "mask")
    input_code = E_POISSON_BLEND_INPUT
    buffer_code = E_POISSON_BLEND_BUFFER
    kind = "poisson"
    ph_op = hip.PH_OP_POISSON_BLEND

    def _resolve(self, params):
        return resolve_blend_controls(params)


class ScreenedPoisson(_HipPoissonOp):
    """repair.screened_poisson@1 on the GPU (screened_poisson.rs:362-430)."""

    op_id = SCREENED_POISSON_OP_ID
    op_name = "repair.screened_poisson"
    ports = ("guidance", "anchor", "mask")
    input_code = E_SCREENED_POISSON_INPUT
    buffer_code = E_SCREENED_POISSON_BUFFER
    kind = "screened-poisson"
    ph_op = hip.PH_OP_SCREENED_POISSON

    def _resolve(self, params):
        return resolve_screened_controls(params)


def gpu_unavailable(reason: str) -> PaintopError:
    return PaintopError(ErrorClass.POLICY, E_GPU_UNAVAILABLE,
                        f"no compatible GPU adapter is available: {reason}")


# ---------------------------------------------------------------------------
# assert.no_change_outside_mask@1 (crates/paintop-cpu/src/assert.rs), the check
# that runs after a blend: the scan on the GPU, the params, errors and verdict
# as the reference op has them.

NO_CHANGE_OUTSIDE_MASK_OP_ID = "assert.no_change_outside_mask@1"
E_ASSERT_INPUT = "E_ASSERT_INPUT"  # assert.rs:59
E_ASSERT_SHAPE = "E_ASSERT_SHAPE"  # assert.rs:63
E_ASSERT_PARAM = "E_ASSERT_PARAM"  # assert.rs:66


def _resolve_comparison_space(params: dict) -> str:
    """ComparisonSpace::resolve (assert.rs:104-128): decoded-linear by default."""
    if "comparison_space" not in params:
        return "decoded-linear"
    v = params["comparison_space"]
    if v in ("decoded-linear", "encoded"):
        return v
    raise PaintopError(ErrorClass.SCHEMA, E_ASSERT_PARAM,
                       "assert.no_change_outside_mask `comparison_space` must be "
                       "`decoded-linear` or `encoded`")


def _resolve_severity(params: dict) -> str:
    """resolve_severity (assert.rs:132-153): error by default."""
    if "severity" not in params:
        return "error"
    v = params["severity"]
    if v in ("error", "warning", "metric"):
        return v
    raise PaintopError(ErrorClass.SCHEMA, E_ASSERT_PARAM,
                       "assertion `severity` must be `error`, `warning`, or `metric`")


def _resolve_non_negative(params: dict, name: str) -> float:
    """resolve_non_negative (assert.rs:157-176): finite and >= 0, default 0."""
    if name not in params:
        return 0.0
    v = params[name]
    ok = isinstance(v, (int, float)) and not isinstance(v, bool)
    if not ok or not math.isfinite(float(v)) or float(v) < 0.0:
        raise PaintopError(ErrorClass.SCHEMA, E_ASSERT_PARAM,
                           f"assert.no_change_outside_mask `{name}` must be a finite "
                           f"non-negative number")
    return float(v)


class NoChangeOutsideMask:
    """assert.no_change_outside_mask@1 with the scan on the GPU
    (NoChangeOutsideMask::compute, assert.rs:630-710;
    ph_assert_no_change_outside_mask_device). The inputs are host values, as
    the executor hands them; ``scan_device`` takes device buffers directly."""

    op_id = NO_CHANGE_OUTSIDE_MASK_OP_ID
    op_name = NO_CHANGE_OUTSIDE_MASK_OP_ID

    def _input(self, inputs, port):
        if port not in inputs:
            raise PaintopError(ErrorClass.REFERENCE, E_ASSERT_INPUT,
                               f"{self.op_name} requires a `{port}` input value")
        return inputs[port]

    def compute(self, inputs: dict, params: dict) -> dict:
        params = params or {}
        space = _resolve_comparison_space(params)
        severity = _resolve_severity(params)
        threshold = _resolve_non_negative(params, "outside_threshold")
        epsilon = _resolve_non_negative(params, "coverage_epsilon")
        before, after, allowed = (self._input(inputs, p) for p in ("before", "after", "allowed"))
        for port, v in (("before", before), ("after", after)):
            if not isinstance(v.descriptor, ImageDescriptor):
                raise PaintopError(ErrorClass.TYPE, E_ASSERT_INPUT,
                                   f"{self.op_name} `{port}` input must be a image resource")
        bd, ad = before.descriptor, after.descriptor
        if bd.extent != ad.extent or before.channels != after.channels:
            raise PaintopError(ErrorClass.SEMANTIC, E_ASSERT_SHAPE,
                               f"{self.op_name}: the `before` and `after` images must share an "
                               f"extent and channel count")
        if not isinstance(allowed.descriptor, MaskDescriptor):
            raise PaintopError(ErrorClass.TYPE, E_ASSERT_INPUT,
                               f"{self.op_name} `allowed` input must be a mask resource")
        if allowed.descriptor.extent != bd.extent:
            raise PaintopError(ErrorClass.SEMANTIC, E_ASSERT_SHAPE,
                               f"{self.op_name}: the `allowed` mask must share the images' extent")
        decode = space == "decoded-linear" and bd.color == "Srgb"
        from .device import DeviceArray

        arrs = [DeviceArray.from_numpy(x.samples) for x in (before, after, allowed)]
        try:
            scan = self.scan_device(arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, bd.extent.width,
                                    bd.extent.height, before.channels, decode, threshold, epsilon)
        finally:
            for a in arrs:
                a.free()
        outcome = AssertionOutcome(
            assertion=self.op_id, passed=scan["changed_outside"] == 0, severity=severity,
            max_abs_delta_outside=scan["max_outside"],
            changed_pixels_outside=scan["changed_outside"], worst_pixel=scan["worst_pixel"],
            locations=scan["locations"])
        report = Report(extent=bd.extent, channels=0, all_finite=True, solver=None,
                        assertion=outcome)
        return {"report": ResourceValue.report_value(report)}

    @staticmethod
    def scan_device(d_before, d_after, d_mask, width, height, channels, decode=False,
                    outside_threshold=0.0, coverage_epsilon=0.0, stream=0) -> dict:
        try:
            return hip.assert_no_change_outside_mask_device(
                d_before, d_after, d_mask, width, height, channels, decode, outside_threshold,
                coverage_epsilon, stream)
        except hip.HipCallError as e:
            raise PaintopError(ErrorClass(e.error_class), e.code, e.message) from None
