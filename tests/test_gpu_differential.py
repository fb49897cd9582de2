"""The hip backend through the reference's own judging path on the GPU: the
differential harness (paintop-testkit/src/differential.rs:434-490, mirrored in
paintop_amd/differential.py) runs the cpu.reference oracle kernel (the C
restatement, omega 1.8) and the registered hip kernel on the same inputs and
holds the hip candidate to the op's tier (bounded: max-abs 1e-4, RMS 1e-5),
for small versions of each workload: the Gaussian-bump case (C1), a star-mask
RGBA fbm blend (C2), a screened case with a border mask (C3) and uneven
instances of the batch generator (C4). Both run to 1e-9, where both solvers
sit on the fixed point. Then a plan runs with hip preferred (provenance from
the probe) and the cache keeps hip and reference entries apart."""
import math

import numpy as np
import pytest

from paintop_amd import cache as C
from paintop_amd import differential as D
from paintop_amd import ops, pipeline
from paintop_amd.executor import register_hip_backend
from paintop_amd.inputs import SplitMix64, gaussian_bump_case, random_polygon, star_polygon
from paintop_amd.values import image, mask
from tests.oracle_ops import reference_registry

pytestmark = pytest.mark.gpu

TIGHT = {"tolerance": 1e-9, "max_iterations": 100000}


def _registry():
    impls, manifests = reference_registry()
    register_hip_backend(impls)
    return impls, manifests


def _check(op, inputs, params=TIGHT, label=""):
    impls, manifests = _registry()
    rep = D.differential_check(op, manifests, impls, D.OpInvocation(inputs, params, "candidate"),
                               D.HipDevice())
    (b, o), = rep.backends
    assert str(b) == "hip.poisson" and not o.is_skipped(), o.reason
    print(f"[diff] {label}: {o.kind}, max_abs {o.error_map.max_abs:.3e}, "
          f"rms {o.error_map.rms:.3e}", flush=True)
    assert rep.all_pass() and not rep.tolerance.exact
    return o.error_map


def test_c1_gaussian_bump(gpu, oracle):
    s, t, m = gaussian_bump_case(256)
    _check(ops.POISSON_BLEND_OP_ID, {"source": image(256, 256, s), "target": image(256, 256, t),
                                     "mask": mask(256, 256, m)}, label="C1 256^2 bump")


def test_c2_star_rgba(gpu, oracle):
    w, h = 192, 160
    src = np.stack([oracle.fbm(w, h, 1.5, 101 + c, 6) * np.float32(0.5) + np.float32(0.5)
                    for c in range(3)] + [np.ones((h, w), np.float32)], axis=-1)
    tgt = np.stack([oracle.fbm(w, h, 1.5, 201 + c, 6) * np.float32(0.5) + np.float32(0.5)
                    for c in range(3)] + [np.ones((h, w), np.float32)], axis=-1)
    m = oracle.polygon_mask(w, h, star_polygon(w / 2, h / 2, 70.0, 0.5, 6, 0.1, 301))
    em = _check(ops.POISSON_BLEND_OP_ID,
                {"source": image(w, h, src, "Rgba"), "target": image(w, h, tgt, "Rgba"),
                 "mask": mask(w, h, m)}, label="C2 192x160 RGBA star")
    assert em.abs_diff.reshape(h, w, 4)[..., 3].max() == 0.0  # alpha passes through


def test_c3_screened_border(gpu, oracle):
    n = 128
    a = oracle.fbm(n, n, 4.0, 401, 6) * np.float32(0.5) + np.float32(0.5)
    g = oracle.fbm(n, n, 4.0, 402, 6) * np.float32(0.5) + np.float32(0.5)
    oracle.gauss_noise(g, 403, 0.01)
    m = oracle.polygon_mask(n, n, [(8, 8), (n - 8, 8), (n - 8, n - 8), (8, n - 8)])
    _check(ops.SCREENED_POISSON_OP_ID,
           {"guidance": image(n, n, g), "anchor": image(n, n, a), "mask": mask(n, n, m)},
           dict(TIGHT, **{"lambda": 0.01}), label="C3 128^2 screened lambda 0.01")


@pytest.mark.parametrize("i", range(4))
def test_c4_uneven_instances(gpu, oracle, i):
    rng = SplitMix64(500 + i)
    s = math.exp(rng.uniform(math.log(32.0), math.log(160.0)))
    asp = rng.uniform(0.5, 2.0)
    w, h = max(8, round(s * math.sqrt(asp))), max(8, round(s / math.sqrt(asp)))
    verts = random_polygon(w / 2, h / 2, 0.45 * w, 0.45 * h, 6 + i, rng)
    m = oracle.polygon_mask(w, h, verts)
    src = oracle.fbm(w, h, 2.0, 10_000 + 2 * i, 6) * np.float32(0.5) + np.float32(0.5)
    tgt = oracle.fbm(w, h, 2.0, 10_001 + 2 * i, 6) * np.float32(0.5) + np.float32(0.5)
    _check(ops.POISSON_BLEND_OP_ID, {"source": image(w, h, src), "target": image(w, h, tgt),
                                     "mask": mask(w, h, m)}, label=f"C4 instance {i} {w}x{h}")


def _plan(allowed):
    return {"paintop": "1.0",
            "policy": {"execution": {"allowed_backends": allowed}},
            "inputs": {"source": {}, "target": {}, "mask": {}},
            "nodes": [{"id": "blend", "op": ops.POISSON_BLEND_OP_ID,
                       "in": {"source": "input:source", "target": "input:target",
                              "mask": "input:mask"}, "params": TIGHT}],
            "exports": {"final": {"resource": "node:blend/candidate"}}}


def test_plan_runs_on_hip_with_device_provenance(gpu, oracle):
    impls, manifests = _registry()
    s, t, m = gaussian_bump_case(128)
    ins = {"source": image(128, 128, s), "target": image(128, 128, t), "mask": mask(128, 128, m)}
    out = pipeline.run_parsed_plan(_plan(["hip", "cpu-reference"]), impls, manifests, ins)
    sel = [e for e in out.trace if e["event"] == "implementation_selected"]
    assert sel[0]["implementation"] == "hip.poisson@1" and sel[0]["fell_back"] is False
    assert out.platform.gpu and "gfx950" in out.platform.gpu and out.platform.driver
    print(f"[plan] platform {out.platform.to_json()}", flush=True)
    ref = pipeline.run_parsed_plan(_plan(["cpu-reference"]), impls, manifests, ins)
    assert ref.selected == {"blend": "cpu.reference@1"} and ref.platform.gpu is None
    em = D.ErrorMap.compute(ref.exports["final"], out.exports["final"])
    assert em.within(D.BOUNDED)


def test_cache_keeps_hip_and_reference_entries_apart(gpu, oracle):
    impls, manifests = _registry()
    s, t, m = gaussian_bump_case(96)
    ins = {"source": image(96, 96, s), "target": image(96, 96, t), "mask": mask(96, 96, m)}
    store = C.CacheStore()
    h1 = C.execute_cached(_plan(["hip"]), impls, manifests, ins, store)
    r1 = C.execute_cached(_plan(["cpu-reference"]), impls, manifests, ins, store)
    h2 = C.execute_cached(_plan(["hip"]), impls, manifests, ins, store)
    assert h1["recomputed"] == ["blend"] and r1["recomputed"] == ["blend"]
    assert h2["cache_hits"] == ["blend"]
    hip_c = h2["node_outputs"]["blend"]["candidate"].samples
    ref_c = r1["node_outputs"]["blend"]["candidate"].samples
    assert hip_c.tobytes() == h1["node_outputs"]["blend"]["candidate"].samples.tobytes()
    assert hip_c.tobytes() != ref_c.tobytes()  # bounded, not exact: distinct entries


def test_plan_of_many_blends_runs_as_one_batched_call(gpu, oracle):
    """Executor-level batching on the device: a plan of five independent
    blend nodes (uneven C4-style images, two of them reading the same inputs,
    one RGBA star) plus a screened node runs its blends as one batched call
    (ph_poisson_solve_batch_device, each group passing its own sweep
    configuration) and gives every node's candidate and report bit for bit as
    the same plan run node by node; each is within the bounded tier of the
    reference run."""
    impls, manifests = _registry()
    ins, nodes = {}, []
    for i in range(4):
        rng = SplitMix64(700 + i)
        s = math.exp(rng.uniform(math.log(32.0), math.log(160.0)))
        asp = rng.uniform(0.5, 2.0)
        w, h = max(8, round(s * math.sqrt(asp))), max(8, round(s / math.sqrt(asp)))
        verts = random_polygon(w / 2, h / 2, 0.45 * w, 0.45 * h, 6 + i, rng)
        m = oracle.polygon_mask(w, h, verts)
        src = oracle.fbm(w, h, 2.0, 20_000 + 2 * i, 6) * np.float32(0.5) + np.float32(0.5)
        tgt = oracle.fbm(w, h, 2.0, 20_001 + 2 * i, 6) * np.float32(0.5) + np.float32(0.5)
        ins.update({f"s{i}": image(w, h, src), f"t{i}": image(w, h, tgt), f"m{i}": mask(w, h, m)})
    n = 96
    src, tgt = (np.stack([oracle.fbm(n, n, 1.5, k + c, 6) * np.float32(0.5) + np.float32(0.5)
                          for c in range(3)] + [np.ones((n, n), np.float32)], axis=-1)
                for k in (30_100, 30_200))
    m4 = oracle.polygon_mask(n, n, star_polygon(n / 2, n / 2, 36.0, 0.5, 6, 0.1, 301))
    ins.update({"s4": image(n, n, src, "Rgba"), "t4": image(n, n, tgt, "Rgba"),
                "m4": mask(n, n, m4)})
    for i, srcs in enumerate([0, 1, 2, 3, 4, 0]):
        nodes.append({"id": f"b{i}", "op": ops.POISSON_BLEND_OP_ID,
                      "in": {"source": f"input:s{srcs}", "target": f"input:t{srcs}",
                             "mask": f"input:m{srcs}"}, "params": TIGHT})
    nodes.append({"id": "scr", "op": ops.SCREENED_POISSON_OP_ID,
                  "in": {"guidance": "input:s0", "anchor": "input:t0", "mask": "input:m0"},
                  "params": {"tolerance": 1e-9, "max_iterations": 100000, "lambda": 0.25}})
    plan = {"paintop": "1.0", "policy": {"execution": {"allowed_backends": ["hip"]}},
            "inputs": {k: {} for k in ins}, "nodes": nodes,
            "exports": {nd["id"]: {"resource": f"node:{nd['id']}/candidate"} for nd in nodes}}
    out = pipeline.run_parsed_plan(plan, impls, manifests, ins)
    batched = [e["nodes"] for e in out.trace if e["event"] == "dispatch_batched"]
    print(f"[plan] batched dispatches {batched}", flush=True)
    assert batched and sum(len(b) for b in batched) >= 2
    assert all(s == "hip.poisson@1" for s in out.selected.values())
    one = pipeline.run_parsed_plan(plan, impls, manifests, ins, batch=False)
    assert not [e for e in one.trace if e["event"] == "dispatch_batched"]
    for nd in nodes:
        nid = nd["id"]
        assert out.exports[nid].array().tobytes() == one.exports[nid].array().tobytes(), nid
        ra = out.node_outputs[nid]["report"].as_report().solver
        rb = one.node_outputs[nid]["report"].as_report().solver
        assert (ra.iterations, ra.final_residual, ra.residual_history) == \
            (rb.iterations, rb.final_residual, rb.residual_history), nid
    ref_plan = dict(plan, policy={"execution": {"allowed_backends": ["cpu-reference"]}})
    ref = pipeline.run_parsed_plan(ref_plan, impls, manifests, ins)
    for nd in nodes:
        em = D.ErrorMap.compute(ref.exports[nd["id"]], out.exports[nd["id"]])
        assert em.within(D.BOUNDED), (nd["id"], em)
