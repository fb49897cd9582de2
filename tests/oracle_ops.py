"""Reference kernels for the host-side harness tests: the two Poisson ops'
compute (poisson_blend.rs:535-600, screened_poisson.rs:362-430) on the CPU
oracle (oracle/poisson_oracle.c: blend_channels with the reference's fixed
omega = 1.8, poisson_blend.rs:78). Test infrastructure: registered as the
``cpu.reference`` slot of an ImplRegistry so the differential harness, the
plan runner and the cache can be exercised against the oracle."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from paintop_amd import ops
from paintop_amd.values import Report, ResourceValue, SolverData

OMEGA = 1.8  # poisson_blend.rs:78


class _OracleOp:
    ports = ("source", "target", "mask")
    kind = "poisson"
    op_name = "repair.poisson_blend"

    def _resolve(self, params):
        return ops.resolve_blend_controls(params)

    def compute(self, inputs: dict, params: dict) -> dict:
        max_it, tol, lam = self._resolve(params)
        a, b, m = (inputs[p] for p in self.ports)
        out_desc = ops.check_shapes(a.descriptor, b.descriptor, m.descriptor.extent, self.op_name)
        out, rep = orc.blend_channels(a.array(), b.array(), m.array()[..., 0], lam, max_it, tol,
                                      omega=OMEGA)
        samples = out.reshape(-1)
        sd = SolverData(kind=self.kind, steps=rep.iterations, stability_number=0.0,
                        stability_limit=0.0, stable=True,
                        residual_history=list(rep.residual_history), total_energy=0.0,
                        iterations=rep.iterations, stop_reason=rep.stop_reason,
                        converged=rep.converged, tolerance=rep.tolerance,
                        final_residual=rep.final_residual)
        report = Report(extent=out_desc.extent, channels=out_desc.channels,
                        all_finite=bool(np.isfinite(samples).all()), solver=sd)
        return {"candidate": ResourceValue(out_desc, out_desc.channels, samples),
                "report": ResourceValue.report_value(report)}


class PoissonBlendOracle(_OracleOp):
    pass


class ScreenedPoissonOracle(_OracleOp):
    ports = ("guidance", "anchor", "mask")
    kind = "screened-poisson"
    op_name = "repair.screened_poisson"

    def _resolve(self, params):
        return ops.resolve_screened_controls(params)


def reference_registry():
    """An ImplRegistry with the oracle kernels in the reference slots, and the
    op manifests by id."""
    from paintop_amd.executor import ImplRegistry

    impls = ImplRegistry()
    impls.register(ops.POISSON_BLEND_OP_ID, PoissonBlendOracle())
    impls.register(ops.SCREENED_POISSON_OP_ID, ScreenedPoissonOracle())
    manifests = {op: ops.load_manifest(op)
                 for op in (ops.POISSON_BLEND_OP_ID, ops.SCREENED_POISSON_OP_ID)}
    return impls, manifests
