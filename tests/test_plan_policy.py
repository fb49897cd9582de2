"""Host-side tests of the layers around the path: plan-level backend selection
(policy.execution.allowed_backends -> BackendPolicy; run_parsed_plan with
evidence provenance), backend identity in the cache key, and the differential
harness (Tolerance / ErrorMap / availability), mirroring
crates/paintop-cpu/src/pipeline.rs:115-125, paintop-core/src/cache/key.rs,
replay.rs and paintop-testkit/src/differential.rs. No GPU: the hip kernels are
either not registered (as on a host whose probe fails) or a test double."""
import numpy as np
import pytest

from paintop_amd import cache as C
from paintop_amd import differential as D
from paintop_amd import ops, pipeline
from paintop_amd.errors import ErrorClass, PaintopError
from paintop_amd.executor import BackendId, BackendPolicy, ImplId
from paintop_amd.inputs import gaussian_bump_case
from paintop_amd.values import image, mask
from tests.oracle_ops import reference_registry

HIP = BackendId("hip", "poisson")


def plan_with(allowed, params=None, n=48):
    p = {
        "paintop": "1.0",
        "name": "poisson-blend-hip",
        "inputs": {"source": {"kind": "image"}, "target": {"kind": "image"},
                   "mask": {"kind": "mask"}},
        "nodes": [{"id": "blend", "op": ops.POISSON_BLEND_OP_ID,
                   "in": {"source": "input:source", "target": "input:target",
                          "mask": "input:mask"},
                   "params": params or {"tolerance": 1e-6, "max_iterations": 4000}}],
        "exports": {"final": {"resource": "node:blend/candidate", "kind": "image"}},
    }
    if allowed is not None:
        p["policy"] = {"execution": {"deadline_ms": 10000, "allowed_backends": allowed}}
    return p


def small_inputs(n=48):
    src, tgt, m = gaussian_bump_case(n)
    return {"source": image(n, n, src), "target": image(n, n, tgt), "mask": mask(n, n, m)}


# ---- allowed_backends -> BackendPolicy ----------------------------------------------


def test_node_policy_mapping():
    man = ops.load_manifest(ops.POISSON_BLEND_OP_ID)
    op = ops.POISSON_BLEND_OP_ID
    assert pipeline.node_policy(plan_with(None), op, man) == BackendPolicy.reference()
    assert pipeline.node_policy(plan_with(["cpu-reference"]), op, man) == BackendPolicy.reference()
    assert pipeline.node_policy(plan_with(["hip", "cpu-reference"]), op, man) == \
        BackendPolicy.prefer([HIP])
    # order of the list does not matter for the reference fall-back
    assert pipeline.node_policy(plan_with(["cpu-reference", "hip"]), op, man) == \
        BackendPolicy.prefer([HIP])
    assert pipeline.node_policy(plan_with(["hip"]), op, man) == BackendPolicy([HIP], True)
    with pytest.raises(PaintopError) as e:
        pipeline.node_policy(plan_with(["wgpu"]), op, man)  # nothing admitted, no fall-back
    assert e.value.code == "E_BACKEND_UNSUPPORTED" and e.value.exit_code == 4
    with pytest.raises(PaintopError) as e:
        pipeline.node_policy(plan_with(["cuda"]), op, man)
    assert e.value.code == "E_POLICY_BACKEND" and e.value.error_class == ErrorClass.SCHEMA


def test_reference_conformance_plans_keep_the_reference_policy():
    """Every plan in the reference's conformance suite allows only
    cpu-reference (conformance/plans/*.json); such a plan never reaches hip."""
    man = ops.load_manifest(ops.POISSON_BLEND_OP_ID)
    plan = {"policy": {"execution": {"deadline_ms": 10000, "allowed_backends": ["cpu-reference"]}}}
    assert pipeline.node_policy(plan, ops.POISSON_BLEND_OP_ID, man) == BackendPolicy.reference()


def test_plan_prefers_hip_and_falls_back_when_it_is_not_registered(oracle):
    impls, manifests = reference_registry()  # hip not registered (no device here)
    out = pipeline.run_parsed_plan(plan_with(["hip", "cpu-reference"]), impls, manifests,
                                   small_inputs())
    sel = [e for e in out.trace if e["event"] == "implementation_selected"]
    assert sel == [{"event": "implementation_selected", "node": "blend",
                    "implementation": "cpu.reference@1", "fell_back": True}]
    assert out.selected == {"blend": "cpu.reference@1"}
    assert out.platform.gpu is None and out.platform.driver is None  # no device served it
    assert set(out.platform.to_json()) == {"os", "arch"}
    cand = out.exports["final"]
    assert cand.array().shape == (48, 48, 1)
    rep = out.node_outputs["blend"]["report"].as_report()
    assert rep.solver.converged


def test_plan_requiring_hip_fails_without_it(oracle):
    impls, manifests = reference_registry()
    with pytest.raises(PaintopError) as e:
        pipeline.run_parsed_plan(plan_with(["hip"]), impls, manifests, small_inputs())
    assert e.value.code == "E_BACKEND_UNSUPPORTED" and e.value.exit_code == 4


def test_plan_reference_only_never_selects_hip(oracle):
    impls, manifests = reference_registry()
    impls.register_backend(ops.POISSON_BLEND_OP_ID, ImplId.parse("hip.poisson@1"), _FakeHip())
    out = pipeline.run_parsed_plan(plan_with(["cpu-reference"]), impls, manifests, small_inputs())
    assert out.selected == {"blend": "cpu.reference@1"}


def test_unresolved_reference_is_an_error():
    impls, manifests = reference_registry()
    with pytest.raises(PaintopError) as e:
        pipeline.run_parsed_plan(plan_with(None), impls, manifests, {})
    assert e.value.code == "E_INPUT_NOT_AVAILABLE"


# ---- backend identity in the cache key --------------------------------------------


class _FakeHip:
    """A bounded-tier stand-in for the hip kernel: the oracle's result moved
    by 5e-5 inside the mask, so a served entry is recognisable."""

    calls = 0

    def compute(self, inputs, params):
        from tests.oracle_ops import PoissonBlendOracle

        _FakeHip.calls += 1
        out = PoissonBlendOracle().compute(inputs, params)
        c = out["candidate"]
        m = inputs["mask"].array()[..., 0] > 0.5
        a = c.array().copy()
        a[m] += np.float32(5e-5)
        out["candidate"] = image(a.shape[1], a.shape[0], a)
        return out


def test_cache_key_carries_the_serving_implementation():
    ins = small_inputs(16)
    man = ops.load_manifest(ops.POISSON_BLEND_OP_ID)
    k_ref = C.node_cache_key(ops.POISSON_BLEND_OP_ID, man, {"tolerance": 1e-6}, ins,
                             "cpu.reference@1")
    k_hip = C.node_cache_key(ops.POISSON_BLEND_OP_ID, man, {"tolerance": 1e-6}, ins,
                             "hip.poisson@1")
    assert k_ref != k_hip
    # the reference key keeps the reference's composition (no implementation field)
    contribs = [C.InputContribution(p, C.content_hash(v), C.semantics(v)) for p, v in ins.items()]
    v = C.CacheKeyInputs(ops.POISSON_BLEND_OP_ID, 1, {"tolerance": 1e-6}, contribs,
                         implementation="cpu.reference@1").to_canonical_value()
    assert set(v) == {"op_id", "op_semantic_version", "params", "inputs", "seed",
                      "backend_semantics_version"}
    assert [i["port"] for i in v["inputs"]] == ["mask", "source", "target"]
    # order of inputs and param keys does not change the key; content does
    rev = dict(reversed(list(ins.items())))
    assert C.node_cache_key(ops.POISSON_BLEND_OP_ID, man, {"tolerance": 1e-6}, rev,
                            "hip.poisson@1") == k_hip
    ins2 = dict(ins)
    a = ins["source"].array().copy()
    a[3, 3, 0] += np.float32(1e-3)
    ins2["source"] = image(16, 16, a)
    assert C.node_cache_key(ops.POISSON_BLEND_OP_ID, man, {"tolerance": 1e-6}, ins2,
                            "hip.poisson@1") != k_hip
    assert k_hip.for_output("candidate") != k_hip.for_output("report")


def test_hip_entry_is_never_served_for_a_reference_request(oracle):
    impls, manifests = reference_registry()
    impls.register_backend(ops.POISSON_BLEND_OP_ID, ImplId.parse("hip.poisson@1"), _FakeHip())
    store = C.CacheStore()
    ins = small_inputs()
    _FakeHip.calls = 0
    r1 = C.execute_cached(plan_with(["hip", "cpu-reference"]), impls, manifests, ins, store)
    assert r1["recomputed"] == ["blend"] and _FakeHip.calls == 1
    r2 = C.execute_cached(plan_with(["hip", "cpu-reference"]), impls, manifests, ins, store)
    assert r2["cache_hits"] == ["blend"] and _FakeHip.calls == 1
    hip_cand = r2["node_outputs"]["blend"]["candidate"].samples
    # a reference request misses the hip entries and computes on the reference
    r3 = C.execute_cached(plan_with(["cpu-reference"]), impls, manifests, ins, store)
    assert r3["recomputed"] == ["blend"] and _FakeHip.calls == 1
    ref_cand = r3["node_outputs"]["blend"]["candidate"].samples
    assert ref_cand.tobytes() != hip_cand.tobytes()
    assert float(np.abs(ref_cand - hip_cand).max()) == pytest.approx(5e-5, rel=0.05)
    r4 = C.execute_cached(plan_with(["cpu-reference"]), impls, manifests, ins, store)
    assert r4["cache_hits"] == ["blend"]
    assert r4["node_outputs"]["blend"]["candidate"].samples.tobytes() == ref_cand.tobytes()
    assert len(store.entries) == 4  # candidate + report, per implementation


# ---- differential harness ---------------------------------------------------------


def _val(fill, n=4):
    return image(n, n, np.full((n, n), fill, np.float32))


def test_tolerance_for_tier():
    assert D.Tolerance.for_tier("exact") == D.EXACT and D.EXACT.exact
    for t in ("bounded", "reproducible", "stochastic"):
        assert D.Tolerance.for_tier(t) == D.BOUNDED
    assert D.BOUNDED.with_bounds(1e-3, 1e-4).max_abs == 1e-3 and not D.BOUNDED.exact


def test_error_map_statistics():
    o = image(3, 1, np.array([0.0, 1.0, 2.0], np.float32))
    c = image(3, 1, np.array([0.0, 1.5, 1.0], np.float32))
    em = D.ErrorMap.compute(o, c)
    assert em.argmax == 2 and em.max_abs == 1.0
    assert em.rms == pytest.approx(np.sqrt((0.25 + 1.0) / 3), rel=1e-6)
    assert list(em.abs_diff) == [0.0, 0.5, 1.0]
    assert not em.within(D.BOUNDED) and D.ErrorMap.compute(o, o).within(D.EXACT)
    # ties keep the first index; a shape mismatch never passes
    t = D.ErrorMap.compute(image(2, 1, [0.0, 0.0]), image(2, 1, [1.0, 1.0]))
    assert t.argmax == 0
    sm = D.ErrorMap.compute(_val(0.0, 4), _val(0.0, 5))
    assert sm.shape_mismatch and not sm.within(D.BOUNDED.with_bounds(1e9, 1e9))
    tiny = D.ErrorMap.compute(_val(0.5), _val(0.5 + 5e-6))
    assert tiny.within(D.BOUNDED) and not tiny.within(D.EXACT)


class _Const:
    def __init__(self, fill):
        self.fill = fill

    def compute(self, inputs, params):
        return {"candidate": _val(self.fill)}


def _registry(tier, impl_ids, kernels):
    from paintop_amd.executor import ImplRegistry

    impls = ImplRegistry()
    impls.register("t.op@1", _Const(0.25))
    for text, k in kernels.items():
        impls.register_backend("t.op@1", ImplId.parse(text), k)
    return impls, {"t.op@1": {"determinism": tier, "implementations": impl_ids}}


def test_differential_outcomes():
    inv = D.OpInvocation({}, {}, "candidate")
    impls, man = _registry("exact", ["cpu.reference@1", "cpu.optimized@1"],
                           {"cpu.optimized@1": _Const(0.25)})
    rep = D.differential_check("t.op@1", man, impls, inv, D.AllAvailable())
    assert rep.tolerance.exact and rep.all_pass() and rep.backends[0][1].is_pass()
    impls, man = _registry("exact", ["cpu.reference@1", "cpu.optimized@1"],
                           {"cpu.optimized@1": _Const(0.75)})
    rep = D.differential_check("t.op@1", man, impls, inv, D.AllAvailable())
    assert not rep.all_pass() and rep.failures()[0][1].max_abs == 0.5
    impls, man = _registry("bounded", ["cpu.reference@1", "cpu.optimized@1"],
                           {"cpu.optimized@1": _Const(0.25 + 5e-6)})
    assert D.differential_check("t.op@1", man, impls, inv, D.AllAvailable()).all_pass()
    # unavailable and unregistered backends are skipped, not failed
    impls, man = _registry("bounded", ["cpu.reference@1", "wgpu.x@1", "hip.poisson@1",
                                       "cpu.optimized@1"], {})
    rep = D.differential_check("t.op@1", man, impls, inv, D.GpuAdapter(False))
    assert rep.all_pass() and all(o.is_skipped() for _, o in rep.backends)
    assert "no GPU adapter" in rep.backends[0][1].reason
    rep = D.differential_check("t.op@1", man, impls, inv, D.HipDevice("forced"))
    assert "hip backend skipped" in rep.backends[1][1].reason
    # an override cannot relax an exact op
    impls, man = _registry("exact", ["cpu.reference@1", "cpu.optimized@1"],
                           {"cpu.optimized@1": _Const(0.25 + 5e-6)})
    rep = D.differential_check("t.op@1", man, impls, inv, D.AllAvailable(),
                               D.BOUNDED.with_bounds(1.0, 1.0))
    assert rep.tolerance.exact and not rep.all_pass()


def test_hip_availability_from_the_probe():
    from paintop_amd import hip

    hip.force_unavailable(True)
    try:
        gate = D.HipDevice()
    finally:
        hip.force_unavailable(False)
    assert gate.reason is not None and "E_GPU_UNAVAILABLE" in gate.reason
    assert gate.unavailable(HIP) is not None
    assert gate.unavailable(BackendId("cpu", "optimized")) is None


# ---- executor-level batching -----------------------------------------------------


class _BatchingHip:
    """A hip stand-in that batches: its batch key is the resolved controls
    (None for a node carrying the `solo` marker param), compute_batch records
    each call's node count and returns what compute returns per node."""

    def __init__(self):
        self.batches = []
        self.singles = 0

    def batch_key(self, inputs, params):
        if params.get("_solo"):
            return None
        p = {k: v for k, v in params.items() if k != "_solo"}
        return tuple(sorted(ops.resolve_blend_controls(p)))

    def compute(self, inputs, params):
        from tests.oracle_ops import PoissonBlendOracle

        self.singles += 1
        p = {k: v for k, v in params.items() if k != "_solo"}
        return PoissonBlendOracle().compute(inputs, p)

    def compute_batch(self, calls):
        from tests.oracle_ops import PoissonBlendOracle

        self.batches.append(len(calls))
        return [PoissonBlendOracle().compute(i, {k: v for k, v in p.items() if k != "_solo"})
                for i, p in calls]


def _multi_plan(nodes, allowed=("hip", "cpu-reference")):
    return {"paintop": "1.0", "name": "many-blends",
            "inputs": {"source": {"kind": "image"}, "target": {"kind": "image"},
                       "mask": {"kind": "mask"}, "source2": {"kind": "image"}},
            "nodes": nodes,
            "exports": {n["id"]: {"resource": f"node:{n['id']}/candidate", "kind": "image"}
                        for n in nodes},
            "policy": {"execution": {"deadline_ms": 10000, "allowed_backends": list(allowed)}}}


def _blend(nid, src="input:source", params=None):
    return {"id": nid, "op": ops.POISSON_BLEND_OP_ID,
            "in": {"source": src, "target": "input:target", "mask": "input:mask"},
            "params": params or {"tolerance": 1e-6, "max_iterations": 3000}}


FAKE_PROBE = {"name": "test double", "arch": "gfx950", "compute_units": 256,
              "runtime_version": 0, "driver_version": 0, "hbm_bytes": 0}


def test_consecutive_independent_nodes_go_in_one_batch(oracle, monkeypatch):
    """run_parsed_plan gathers a run of consecutive nodes of one op with equal
    batch keys and no input from each other into one compute_batch call; a
    node reading another's output, a node with other controls, a node that
    batches alone and a node whose policy selects another implementation each
    end the run. Outputs equal the unbatched run's bit for bit."""
    from paintop_amd import hip

    monkeypatch.setattr(hip, "probe", lambda: dict(FAKE_PROBE))  # no device here
    impls, manifests = reference_registry()
    fake = _BatchingHip()
    impls.register_backend(ops.POISSON_BLEND_OP_ID, ImplId.parse(ops.HIP_IMPL_ID), fake)
    ins = small_inputs(24)
    ins["source2"] = image(24, 24, ins["source"].array()[..., 0] * np.float32(0.5))
    nodes = [
        _blend("a"), _blend("b", "input:source2"), _blend("c"),           # one batch of 3
        _blend("d", "node:c/candidate"),                                  # reads c: new run
        _blend("e"),                                                      # batch with d
        _blend("f", params={"tolerance": 1e-5, "max_iterations": 3000}),  # other controls
        _blend("g", params={"tolerance": 1e-6, "max_iterations": 3000, "_solo": True}),
        _blend("h"), _blend("i"),                                         # batch of 2
    ]
    out = pipeline.run_parsed_plan(_multi_plan(nodes), impls, manifests, ins)
    assert fake.batches == [3, 2, 2] and fake.singles == 2
    batched = [e["nodes"] for e in out.trace if e["event"] == "dispatch_batched"]
    assert batched == [["a", "b", "c"], ["d", "e"], ["h", "i"]]
    assert set(out.selected.values()) == {ops.HIP_IMPL_ID}
    sel = [e["node"] for e in out.trace if e["event"] == "implementation_selected"]
    assert sel == list("abcdefghi")
    # the same plan node by node
    impls2, _ = reference_registry()
    fake2 = _BatchingHip()
    impls2.register_backend(ops.POISSON_BLEND_OP_ID, ImplId.parse(ops.HIP_IMPL_ID), fake2)
    out1 = pipeline.run_parsed_plan(_multi_plan(nodes), impls2, manifests, ins, batch=False)
    assert fake2.batches == [] and fake2.singles == 9
    for nid in "abcdefghi":
        assert out.exports[nid].array().tobytes() == out1.exports[nid].array().tobytes()


def test_batching_respects_the_selected_implementation(oracle):
    """A plan that allows only the reference never batches (the oracle does
    not batch), and a hip node is not batched with one the policy sends to
    another implementation."""
    impls, manifests = reference_registry()
    fake = _BatchingHip()
    impls.register_backend(ops.POISSON_BLEND_OP_ID, ImplId.parse(ops.HIP_IMPL_ID), fake)
    ins = small_inputs(16)
    ins["source2"] = ins["source"]
    nodes = [_blend("a"), _blend("b")]
    out = pipeline.run_parsed_plan(_multi_plan(nodes, allowed=("cpu-reference",)), impls,
                                   manifests, ins)
    assert fake.batches == [] and fake.singles == 0
    assert not [e for e in out.trace if e["event"] == "dispatch_batched"]
    assert set(out.selected.values()) == {"cpu.reference.poisson_blend@1"} or \
        all(s.startswith("cpu.reference") for s in out.selected.values())
