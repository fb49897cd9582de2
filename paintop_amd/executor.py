"""The dispatch boundary the hip backend plugs into, mirrored from
paintop-core/src/executor (op_impl.rs, dispatch.rs, mod.rs).

* ``ImplId`` -- ``<backend>.<name>@<version>`` (paintop-ir/src/manifest.rs:192-330).
* ``ImplRegistry`` -- one reference slot per op plus a per-backend table;
  ``register_backend`` rejects reference ids and duplicates (op_impl.rs:92-200).
* ``BackendPolicy`` / ``select_backend`` -- a backend is selectable only when the
  manifest declares it and a kernel is registered; the first preferred backend
  is the intended one, anything later is a recorded fall-back; ``require``
  forbids the fall-back (dispatch.rs:121-318).
* ``dispatch_node`` -- selection, trace events around ``compute``, and the
  declared-output check (executor/mod.rs:232-349).
"""
from __future__ import annotations

import re
import time
from dataclasses import dataclass, field

from .errors import E_BACKEND_UNSUPPORTED, E_OUTPUT_NOT_PRODUCED, ErrorClass, PaintopError

_IMPL_RE = re.compile(r"^([a-z][a-z0-9_]*)\.([a-z][a-z0-9_]*)@([1-9][0-9]*)$")


@dataclass(frozen=True)
class BackendId:
    backend: str
    name: str

    @staticmethod
    def reference() -> "BackendId":
        return BackendId("cpu", "reference")

    def is_reference(self) -> bool:
        return self == BackendId.reference()

    def matches(self, impl_id: "ImplId") -> bool:
        return impl_id.backend == self.backend and impl_id.name == self.name

    def __str__(self) -> str:
        return f"{self.backend}.{self.name}"


@dataclass(frozen=True)
class ImplId:
    backend: str
    name: str
    version: int

    @staticmethod
    def parse(text: str) -> "ImplId":
        m = _IMPL_RE.match(text)
        if not m:
            raise PaintopError(ErrorClass.SCHEMA, "E_IMPL_ID", f"invalid implementation id `{text}`")
        return ImplId(m.group(1), m.group(2), int(m.group(3)))

    def backend_id(self) -> BackendId:
        return BackendId(self.backend, self.name)

    def __str__(self) -> str:
        return f"{self.backend}.{self.name}@{self.version}"


class ImplRegistry:
    def __init__(self):
        self._reference = {}
        self._backends = {}

    def register(self, op_id: str, impl) -> None:
        if op_id in self._reference:
            raise PaintopError(ErrorClass.SCHEMA, "E_IMPL_DUPLICATE",
                               f"operation {op_id} already has a reference implementation")
        self._reference[op_id] = impl

    def register_backend(self, op_id: str, impl_id: ImplId, impl) -> None:
        b = impl_id.backend_id()
        if b.is_reference():
            raise PaintopError(ErrorClass.SCHEMA, "E_IMPL_REFERENCE",
                               f"{impl_id} is the reference slot; use register()")
        key = (op_id, b)
        if key in self._backends:
            raise PaintopError(ErrorClass.SCHEMA, "E_IMPL_DUPLICATE",
                               f"operation {op_id} already has a `{b}` implementation")
        self._backends[key] = impl

    def contains(self, op_id: str) -> bool:
        return op_id in self._reference

    def has_backend(self, op_id: str, backend: BackendId) -> bool:
        return (op_id, backend) in self._backends

    def get_backend(self, op_id: str, backend: BackendId):
        if backend.is_reference():
            return self._reference.get(op_id)
        return self._backends.get((op_id, backend))


@dataclass
class BackendPolicy:
    preferred: list = field(default_factory=lambda: [BackendId.reference()])
    required: bool = False

    @staticmethod
    def reference() -> "BackendPolicy":
        return BackendPolicy([BackendId.reference()], False)

    @staticmethod
    def prefer(backends) -> "BackendPolicy":
        pref = list(backends)
        if not any(b.is_reference() for b in pref):
            pref.append(BackendId.reference())
        return BackendPolicy(pref, False)

    @staticmethod
    def require(backend: BackendId) -> "BackendPolicy":
        return BackendPolicy([backend], True)


@dataclass(frozen=True)
class BackendSelection:
    impl_id: ImplId
    fell_back: bool


def _resolve(op_id, backend: BackendId, declared, impls: ImplRegistry):
    found = next((d for d in declared if backend.matches(d)), None)
    if found is None:
        return None
    if impls.has_backend(op_id, backend):
        return found
    if backend.is_reference() and impls.contains(op_id):
        return found
    return None


def select_backend(op_id: str, manifest: dict, impls: ImplRegistry,
                   policy: BackendPolicy) -> BackendSelection:
    """dispatch.rs:248-297."""
    declared = [ImplId.parse(s) for s in (manifest or {}).get("implementations", [])]
    intended = policy.preferred[0] if policy.preferred else None
    for b in policy.preferred:
        impl_id = _resolve(op_id, b, declared, impls)
        if impl_id is not None:
            return BackendSelection(impl_id, intended != b)
    if policy.required:
        wanted = str(policy.preferred[0]) if policy.preferred else str(BackendId.reference())
        raise PaintopError(
            ErrorClass.POLICY, E_BACKEND_UNSUPPORTED,
            f"operation {op_id} has no registered `{wanted}` implementation and the policy "
            "forbids falling back to the reference oracle")
    impl_id = _resolve(op_id, BackendId.reference(), declared, impls)
    if impl_id is not None:
        return BackendSelection(impl_id, True)
    raise PaintopError(ErrorClass.POLICY, E_BACKEND_UNSUPPORTED,
                       f"operation {op_id} has no registered `cpu.reference` oracle to dispatch")


def dispatch_node(node_id: str, op_id: str, manifest: dict, impls: ImplRegistry,
                  policy: BackendPolicy, inputs: dict, params: dict, trace: list) -> dict:
    """Select, trace, compute and check the declared outputs (executor/mod.rs:232-349)."""
    sel = select_backend(op_id, manifest, impls, policy)
    impl = impls.get_backend(op_id, sel.impl_id.backend_id())
    trace.append({"event": "implementation_selected", "node": node_id,
                  "implementation": str(sel.impl_id), "fell_back": sel.fell_back})
    trace.append({"event": "dispatch_started", "node": node_id})
    trace.append({"event": "cache_lookup", "node": node_id, "outcome": "bypass"})
    t0 = time.perf_counter()
    outputs = impl.compute(inputs, params)
    elapsed = (time.perf_counter() - t0) * 1e3
    for port in [o["name"] for o in manifest.get("outputs", [])]:
        if port not in outputs:
            raise PaintopError(ErrorClass.EXECUTION, E_OUTPUT_NOT_PRODUCED,
                               f"node {node_id} did not produce declared output `{port}`")
    trace.append({"event": "dispatch_completed", "node": node_id, "elapsed_ms": elapsed})
    return outputs


def dispatch_batch(node_ids: list, op_id: str, manifest: dict, impl, sel: BackendSelection,
                   calls: list, trace: list) -> list:
    """Several nodes of one op, selected to the same implementation, in one
    batched call (the hip backend's ``compute_batch``): the per-node trace
    events of ``dispatch_node`` plus one ``dispatch_batched`` event naming the
    nodes; every node's outputs are checked against the manifest."""
    for nid in node_ids:
        trace.append({"event": "implementation_selected", "node": nid,
                      "implementation": str(sel.impl_id), "fell_back": sel.fell_back})
        trace.append({"event": "dispatch_started", "node": nid})
        trace.append({"event": "cache_lookup", "node": nid, "outcome": "bypass"})
    trace.append({"event": "dispatch_batched", "nodes": list(node_ids),
                  "implementation": str(sel.impl_id)})
    t0 = time.perf_counter()
    outputs = impl.compute_batch(calls)
    elapsed = (time.perf_counter() - t0) * 1e3
    for nid, out in zip(node_ids, outputs):
        for port in [o["name"] for o in manifest.get("outputs", [])]:
            if port not in out:
                raise PaintopError(ErrorClass.EXECUTION, E_OUTPUT_NOT_PRODUCED,
                                   f"node {nid} did not produce declared output `{port}`")
        trace.append({"event": "dispatch_completed", "node": nid, "elapsed_ms": elapsed,
                      "batched": len(node_ids)})
    return outputs


def register_hip_backend(impls: ImplRegistry, **tuning) -> None:
    """Register the hip implementations of both Poisson ops under ``hip.poisson@1``
    (the line domains/repair.rs:23-34 would add next to the reference ones)."""
    from .ops import HIP_IMPL_ID, POISSON_BLEND_OP_ID, SCREENED_POISSON_OP_ID, PoissonBlend, \
        ScreenedPoisson

    hid = ImplId.parse(HIP_IMPL_ID)
    impls.register_backend(POISSON_BLEND_OP_ID, hid, PoissonBlend(**tuning))
    impls.register_backend(SCREENED_POISSON_OP_ID, hid, ScreenedPoisson(**tuning))
