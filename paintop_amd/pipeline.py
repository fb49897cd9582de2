"""Plan-level backend selection and run provenance for the hip backend.

In the reference, ``run_parsed_plan`` always dispatches with the reference
policy (crates/paintop-cpu/src/pipeline.rs:115-125) and the plan's
``policy.execution.allowed_backends`` (IR_SPEC §15; opaque JSON in
paintop-ir/src/plan.rs:79) is carried but not read, so a plan can never be
served by a GPU backend. This module is the adjacent change a maintainer adds
with the hip backend:

* ``node_policy`` maps ``allowed_backends`` to a ``BackendPolicy`` per node:
  the non-reference backends the plan allows, in its order, restricted to
  those the op's manifest declares, are preferred; when ``cpu-reference`` is
  allowed too the policy falls back to it (``BackendPolicy::prefer``,
  dispatch.rs:138-175), otherwise a missing backend is an error
  (``BackendPolicy::require``). A plan without the key keeps the reference
  policy, exactly as today.
* ``run_parsed_plan`` resolves ``input:`` / ``node:`` references
  (paintop-ir/src/resolve.rs:149-170), dispatches every node through
  ``executor.dispatch_node`` (which traces ``implementation_selected`` with
  the fall-back flag) and resolves the exports.
* ``Platform`` mirrors the evidence bundle's platform record
  (paintop-core/src/evidence/manifest.rs:80-110); when a hip implementation
  served a node, ``gpu`` and ``driver`` are filled from ``ph_probe``, the way
  the wgpu adapter's ``to_platform`` fills them
  (paintop-wgpu/src/gpu/probe.rs:60-101).
"""
from __future__ import annotations

import platform as _platform
from dataclasses import dataclass, field
from typing import Optional

from .errors import E_BACKEND_UNSUPPORTED, ErrorClass, PaintopError
from .executor import (BackendId, BackendPolicy, ImplId, ImplRegistry, dispatch_batch,
                       dispatch_node, select_backend)

# Names a plan may list in policy.execution.allowed_backends (IR_SPEC §15) and
# the backends each admits: an exact backend id, or a whole backend family.
ALLOWED_BACKENDS = {
    "cpu-reference": ("id", BackendId("cpu", "reference")),
    "cpu-optimized": ("id", BackendId("cpu", "optimized")),
    "wgpu": ("family", "wgpu"),
    "hip": ("family", "hip"),
}
E_POLICY_BACKEND = "E_POLICY_BACKEND"  # class Schema: an unknown allowed_backends name


def _admits(name: str, b: BackendId) -> bool:
    if name not in ALLOWED_BACKENDS:
        raise PaintopError(ErrorClass.SCHEMA, E_POLICY_BACKEND,
                           f"policy.execution.allowed_backends names unknown backend `{name}` "
                           f"(known: {', '.join(sorted(ALLOWED_BACKENDS))})")
    kind, what = ALLOWED_BACKENDS[name]
    return b == what if kind == "id" else b.backend == what


def allowed_backends(plan: dict) -> Optional[list]:
    """The plan's policy.execution.allowed_backends list, or None when absent."""
    ex = ((plan or {}).get("policy") or {}).get("execution") or {}
    allowed = ex.get("allowed_backends")
    if allowed is None:
        return None
    if not isinstance(allowed, list) or not all(isinstance(a, str) for a in allowed):
        raise PaintopError(ErrorClass.SCHEMA, E_POLICY_BACKEND,
                           "policy.execution.allowed_backends must be a list of backend names")
    return allowed


def node_policy(plan: dict, op_id: str, manifest: dict) -> BackendPolicy:
    """The BackendPolicy a node of ``op_id`` runs under in ``plan``."""
    allowed = allowed_backends(plan)
    if allowed is None:
        return BackendPolicy.reference()
    declared = [ImplId.parse(t).backend_id() for t in manifest.get("implementations", [])]
    preferred = []
    for name in allowed:
        for b in declared:
            if not b.is_reference() and _admits(name, b) and b not in preferred:
                preferred.append(b)
    reference_ok = any(_admits(n, BackendId.reference()) for n in allowed)
    if not preferred:
        if reference_ok:
            return BackendPolicy.reference()
        raise PaintopError(ErrorClass.POLICY, E_BACKEND_UNSUPPORTED,
                           f"operation {op_id} declares no implementation the plan's "
                           f"allowed_backends {allowed} admit")
    if reference_ok:
        return BackendPolicy.prefer(preferred)
    return BackendPolicy(preferred, True)


@dataclass
class Platform:
    """evidence/manifest.rs:80-110; the device fields are absent (None) when
    unknown."""

    os: str
    arch: str
    cpu: Optional[str] = None
    gpu: Optional[str] = None
    driver: Optional[str] = None

    @staticmethod
    def current() -> "Platform":
        return Platform(os=_platform.system().lower(), arch=_platform.machine())

    def to_json(self) -> dict:
        return {k: v for k, v in (("os", self.os), ("arch", self.arch), ("cpu", self.cpu),
                                  ("gpu", self.gpu), ("driver", self.driver)) if v is not None}


def hip_platform(info: dict) -> Platform:
    """The host platform with the hip device's provenance (probe.rs:60-101
    pattern: gpu = "<name> (<backend>)", driver when known)."""
    p = Platform.current()
    p.gpu = f"{info['name']} (hip {info['arch']})"
    rv, dv = info.get("runtime_version"), info.get("driver_version")
    if rv or dv:
        p.driver = f"HIP runtime {rv}, driver {dv}"
    return p


@dataclass
class RunOutcome:
    node_outputs: dict
    exports: dict
    trace: list
    platform: Platform
    selected: dict = field(default_factory=dict)  # node id -> implementation id


def resolve_reference(ref: str, inputs: dict, node_outputs: dict):
    """``input:<id>`` or ``node:<id>/<port>`` (resolve.rs:149-170)."""
    if ref.startswith("input:"):
        key = ref[len("input:"):]
        if key in inputs:
            return inputs[key]
    elif ref.startswith("node:"):
        nid, _, port = ref[len("node:"):].partition("/")
        if nid in node_outputs and port in node_outputs[nid]:
            return node_outputs[nid][port]
    raise PaintopError(ErrorClass.REFERENCE, "E_INPUT_NOT_AVAILABLE",
                       f"resource reference `{ref}` is not available")


def _batch_key(impl, ins: dict, params: dict):
    """The node's batch key when its implementation batches, else None."""
    key = getattr(impl, "batch_key", None)
    if key is None or not hasattr(impl, "compute_batch"):
        return None
    try:
        return key(ins, params)
    except PaintopError:
        return None  # dispatched alone, where compute reports the error


def run_parsed_plan(plan: dict, impls: ImplRegistry, manifests: dict,
                    inputs: Optional[dict] = None, batch: bool = True) -> RunOutcome:
    """Run every node of ``plan`` in its listed (topological-compatible) order
    under the node's policy, then resolve the exports.

    Executor-level batching (``batch``): a run of consecutive nodes of one op
    whose implementation batches (the hip Poisson ops), selected to the same
    implementation, with equal batch keys and no input of one taken from
    another, goes to the device as one batched call
    (``executor.dispatch_batch``). The batch key makes each node's result the
    one it gets alone, bit for bit, so batching changes only the launches.
    The reference dispatches node by node (executor/mod.rs:175-193)."""
    inputs = inputs or {}
    node_outputs, trace, selected = {}, [], {}
    nodes = plan.get("nodes", [])

    def prepare(node):
        nid, op_id = node["id"], node["op"]
        manifest = manifests.get(op_id)
        if manifest is None:
            raise PaintopError(ErrorClass.REFERENCE, "E_OP_UNKNOWN",
                               f"node {nid}: operation {op_id} is not registered")
        ins = {port: resolve_reference(ref, inputs, node_outputs)
               for port, ref in (node.get("in") or {}).items()}
        return nid, op_id, manifest, ins, node.get("params") or {}

    i = 0
    while i < len(nodes):
        nid, op_id, manifest, ins, params = prepare(nodes[i])
        policy = node_policy(plan, op_id, manifest)
        group = [(nid, ins, params)]
        if batch:
            sel = select_backend(op_id, manifest, impls, policy)
            impl = impls.get_backend(op_id, sel.impl_id.backend_id())
            key = _batch_key(impl, ins, params)
            j = i + 1
            while key is not None and j < len(nodes) and nodes[j]["op"] == op_id:
                refs = (nodes[j].get("in") or {}).values()
                if any(r.startswith("node:") and r[5:].partition("/")[0] in
                       {g[0] for g in group} for r in refs):
                    break
                try:
                    nj, _, _, ins_j, params_j = prepare(nodes[j])
                except PaintopError:
                    break
                if select_backend(op_id, manifest, impls,
                                  node_policy(plan, op_id, manifest)) != sel or \
                        _batch_key(impl, ins_j, params_j) != key:
                    break
                group.append((nj, ins_j, params_j))
                j += 1
        if len(group) > 1:
            outs = dispatch_batch([g[0] for g in group], op_id, manifest, impl, sel,
                                  [(g[1], g[2]) for g in group], trace)
            for (gid, _, _), out in zip(group, outs):
                node_outputs[gid] = out
                selected[gid] = str(sel.impl_id)
        else:
            node_outputs[nid] = dispatch_node(nid, op_id, manifest, impls, policy, ins, params,
                                              trace)
            selected[nid] = next(e["implementation"] for e in reversed(trace)
                                 if e["event"] == "implementation_selected" and e["node"] == nid)
        i += len(group)
    exports = {}
    for eid, spec in (plan.get("exports") or {}).items():
        ref = spec["resource"] if isinstance(spec, dict) else spec
        exports[eid] = resolve_reference(ref, inputs, node_outputs)
    platform = Platform.current()
    if any(ImplId.parse(s).backend == "hip" for s in selected.values()):
        from . import hip

        platform = hip_platform(hip.probe())
    return RunOutcome(node_outputs, exports, trace, platform, selected)
