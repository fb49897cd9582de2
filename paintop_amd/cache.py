"""Content-addressed node cache with backend identity in the key.

Mirrors paintop-core/src/cache (key.rs, replay.rs). The reference key holds the
op id, its semantic version, the canonical params, every input's content hash
and semantics (sorted by port), the seed and ``BACKEND_SEMANTICS_VERSION``
(key.rs:48-57, 100-190), and ``execute_cached`` always dispatches the
reference kernel (replay.rs:109-290). A key without the implementation that
served the node would let a bounded-tier hip result (equal to the oracle only
within max-abs 1e-4) be served for a reference request. So here the key also
carries the implementation id whenever it is not the reference: reference
keys keep their exact composition, and a hip entry can never answer a
reference lookup, nor the reverse.

Hashing: the reference uses BLAKE3 over canonical JSON, length-framed with the
domain label "paintop/cache-entry/v1" (paintop-ir/src/hash.rs:225-266). No
BLAKE3 module is installed here, so BLAKE2b-256 stands in for it with the same
framing; canonical JSON is sorted keys without whitespace.
"""
from __future__ import annotations

import hashlib
import json
import struct
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .errors import E_OUTPUT_NOT_PRODUCED, ErrorClass, PaintopError
from .executor import BackendPolicy, ImplId, ImplRegistry, select_backend
from .values import ImageDescriptor, MaskDescriptor, ResourceValue

BACKEND_SEMANTICS_VERSION = 1  # key.rs:45-57
CACHE_ENTRY_DOMAIN = "paintop/cache-entry/v1"  # hash.rs:128
CONTENT_DOMAIN = "paintop/content/v1"


def _hash(domain: str, payload: bytes) -> str:
    h = hashlib.blake2b(digest_size=32)
    label = domain.encode()
    h.update(struct.pack("<Q", len(label)))
    h.update(label)
    h.update(payload)
    return "blake2b:" + h.hexdigest()


def canonical_bytes(value) -> bytes:
    return json.dumps(value, sort_keys=True, separators=(",", ":"), allow_nan=False).encode()


def semantics(value: ResourceValue) -> dict:
    d = value.descriptor
    if isinstance(d, ImageDescriptor):
        return {"kind": "image", "width": d.extent.width, "height": d.extent.height,
                "layout": d.layout, "color": d.color}
    if isinstance(d, MaskDescriptor):
        return {"kind": "mask", "width": d.extent.width, "height": d.extent.height}
    return {"kind": "report"}


def content_hash(value: ResourceValue) -> str:
    """The identity of a value's content: its semantics and its f32 sample bytes."""
    s = np.ascontiguousarray(value.samples, dtype="<f4")
    return _hash(CONTENT_DOMAIN, canonical_bytes(semantics(value)) + s.tobytes())


@dataclass
class InputContribution:
    port: str
    content_hash: str
    semantics: dict


@dataclass
class CacheKeyInputs:
    op_id: str
    op_semantic_version: int
    params: dict
    inputs: list
    seed: int = 0
    backend_semantics_version: int = BACKEND_SEMANTICS_VERSION
    implementation: Optional[str] = None  # the implementation id that serves the node

    def to_canonical_value(self) -> dict:
        v = {
            "op_id": self.op_id,
            "op_semantic_version": self.op_semantic_version,
            "params": self.params,
            "inputs": [{"port": i.port, "content_hash": i.content_hash, "semantics": i.semantics}
                       for i in sorted(self.inputs, key=lambda i: i.port)],
            "seed": self.seed,
            "backend_semantics_version": self.backend_semantics_version,
        }
        if self.implementation is not None and not ImplId.parse(
                self.implementation).backend_id().is_reference():
            v["implementation"] = self.implementation
        return v


@dataclass(frozen=True)
class CacheKey:
    digest: str

    @staticmethod
    def compute(inputs: CacheKeyInputs) -> "CacheKey":
        return CacheKey(_hash(CACHE_ENTRY_DOMAIN, canonical_bytes(inputs.to_canonical_value())))

    def for_output(self, port: str) -> "CacheKey":
        """key.rs:256-266: the per-output key (the port length-framed after the key)."""
        p = port.encode()
        return CacheKey(_hash(CACHE_ENTRY_DOMAIN, self.digest.encode() + struct.pack("<Q", len(p))
                              + p))

    def __str__(self) -> str:
        return self.digest


@dataclass
class CacheStore:
    """An in-memory store of values by per-output key."""

    entries: dict = field(default_factory=dict)

    def get(self, key: CacheKey) -> Optional[ResourceValue]:
        return self.entries.get(key.digest)

    def put(self, key: CacheKey, value: ResourceValue) -> None:
        self.entries[key.digest] = value


def node_cache_key(op_id: str, manifest: dict, params: dict, inputs: dict,
                   implementation: str) -> CacheKey:
    contribs = [InputContribution(p, content_hash(v), semantics(v)) for p, v in inputs.items()]
    return CacheKey.compute(CacheKeyInputs(op_id, int(manifest.get("impl_version", 1)), params,
                                           contribs, implementation=implementation))


def dispatch_cached(node_id: str, op_id: str, manifest: dict, impls: ImplRegistry,
                    policy: BackendPolicy, inputs: dict, params: dict, cache: CacheStore,
                    trace: list) -> dict:
    """replay.rs:198-290 with backend selection: the implementation is chosen
    first, the key carries it, and a full hit (every declared output present)
    skips the compute."""
    sel = select_backend(op_id, manifest, impls, policy)
    impl_id = str(sel.impl_id)
    key = node_cache_key(op_id, manifest, params, inputs, impl_id)
    ports = [o["name"] for o in manifest.get("outputs", [])]
    hits = {p: cache.get(key.for_output(p)) for p in ports}
    trace.append({"event": "implementation_selected", "node": node_id,
                  "implementation": impl_id, "fell_back": sel.fell_back})
    if ports and all(v is not None for v in hits.values()):
        trace.append({"event": "cache_lookup", "node": node_id, "outcome": "hit",
                      "key": str(key)})
        return hits
    trace.append({"event": "cache_lookup", "node": node_id, "outcome": "miss", "key": str(key)})
    trace.append({"event": "dispatch_started", "node": node_id, "implementation": impl_id})
    produced = impls.get_backend(op_id, sel.impl_id.backend_id()).compute(inputs, params)
    out = {}
    for p in ports:
        if p not in produced:
            raise PaintopError(ErrorClass.EXECUTION, E_OUTPUT_NOT_PRODUCED,
                               f"node {node_id} did not produce declared output `{p}`")
        cache.put(key.for_output(p), produced[p])
        out[p] = produced[p]
    trace.append({"event": "dispatch_completed", "node": node_id, "implementation": impl_id})
    return out


def execute_cached(plan: dict, impls: ImplRegistry, manifests: dict, inputs: dict,
                   cache: CacheStore) -> dict:
    """Run the plan's nodes in order with the cache interposed (replay.rs:109-175);
    each node's policy comes from the plan (pipeline.node_policy). Returns
    {"node_outputs", "trace", "cache_hits", "recomputed"}."""
    from .pipeline import node_policy, resolve_reference

    node_outputs, trace, hits, recomputed = {}, [], [], []
    for node in plan.get("nodes", []):
        nid, op_id = node["id"], node["op"]
        manifest = manifests[op_id]
        ins = {port: resolve_reference(ref, inputs, node_outputs)
               for port, ref in (node.get("in") or {}).items()}
        n0 = len(trace)
        node_outputs[nid] = dispatch_cached(nid, op_id, manifest, impls,
                                            node_policy(plan, op_id, manifest), ins,
                                            node.get("params") or {}, cache, trace)
        hit = any(e["event"] == "cache_lookup" and e["outcome"] == "hit" for e in trace[n0:])
        (hits if hit else recomputed).append(nid)
    return {"node_outputs": node_outputs, "trace": trace, "cache_hits": hits,
            "recomputed": recomputed}
