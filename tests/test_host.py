"""CPU-side tests of the host layer: the C-ABI library loads and exports every
symbol the header declares, parameter resolution matches the reference's
codes and classes, the backend-selection boundary behaves like dispatch.rs,
and without a GPU the ops fail loudly (no CPU fallback)."""
import os
import re

import numpy as np
import pytest

from paintop_amd import executor, hip, ops
from paintop_amd.errors import ErrorClass, PaintopError
from paintop_amd.values import image, mask

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "paintop_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    names = declared_symbols()
    assert names, "header declares no entry points"
    assert set(names) == set(hip.EXPORTED_SYMBOLS)
    for n in names:
        assert hasattr(L, n), n
    assert hip.abi_version() == 2


def test_probe_forced_unavailable_is_policy_error():
    hip.force_unavailable(True)
    try:
        with pytest.raises(hip.HipCallError) as ei:
            hip.probe()
        assert ei.value.status == hip.PH_UNAVAILABLE
        assert ei.value.error_class == "policy" and ei.value.code == "E_GPU_UNAVAILABLE"
    finally:
        hip.force_unavailable(False)


def test_op_without_device_fails_loudly_with_policy_error():
    hip.force_unavailable(True)
    try:
        w = h = 8
        inputs = {"source": image(w, h, np.zeros(64)), "target": image(w, h, np.zeros(64)),
                  "mask": mask(w, h, np.ones(64))}
        with pytest.raises(PaintopError) as ei:
            ops.PoissonBlend().compute(inputs, {})
        assert ei.value.error_class == ErrorClass.POLICY
        assert ei.value.code == "E_GPU_UNAVAILABLE" and ei.value.exit_code == 4
    finally:
        hip.force_unavailable(False)


@pytest.mark.parametrize("params,msg", [
    ({"max_iterations": 0}, "1..=100000"),
    ({"max_iterations": 100001}, "1..=100000"),
    ({"max_iterations": -1}, "positive integer"),
    ({"max_iterations": 3.0}, "positive integer"),
    ({"max_iterations": "5"}, "positive integer"),
    ({"tolerance": 2.0}, "(0, 1)"),  # poisson_blend/tests.rs:268
    ({"tolerance": 0}, "(0, 1)"),
    ({"tolerance": "x"}, "must be a number"),
    ({"tolerance": True}, "must be a number"),
])
def test_blend_param_errors(params, msg):
    with pytest.raises(PaintopError) as ei:
        ops.resolve_blend_controls(params)
    e = ei.value
    assert e.error_class == ErrorClass.SCHEMA and e.code == "E_POISSON_BLEND_PARAM"
    assert msg in e.message and e.exit_code == 2


def test_blend_defaults_and_unknown_params_ignored():
    assert ops.resolve_blend_controls({}) == (2000, 1e-6, 0.0)
    assert ops.resolve_blend_controls({"omega": 1.2, "lambda": 3}) == (2000, 1e-6, 0.0)
    assert ops.resolve_blend_controls({"max_iterations": 7, "tolerance": 0.5}) == (7, 0.5, 0.0)


def test_screened_params():
    assert ops.resolve_screened_controls({}) == (2000, 1e-6, 0.0)
    assert ops.resolve_screened_controls({"lambda": 2}) == (2000, 1e-6, 2.0)
    for bad in (-1.0, float("inf"), "a"):
        with pytest.raises(PaintopError) as ei:  # screened tests :284
            ops.resolve_screened_controls({"lambda": bad})
        assert ei.value.code == "E_SCREENED_POISSON_PARAM"


def test_shape_errors_before_any_device_work():
    # mismatched extents are Semantic errors (poisson_blend/tests.rs:258)
    inputs = {"source": image(8, 8, np.zeros(64)), "target": image(8, 8, np.zeros(64)),
              "mask": mask(4, 4, np.zeros(16))}
    with pytest.raises(PaintopError) as ei:
        ops.PoissonBlend().compute(inputs, {})
    assert ei.value.error_class == ErrorClass.SEMANTIC
    assert ei.value.code == "E_POISSON_BLEND_SHAPE" and ei.value.exit_code == 3
    inputs = {"source": image(8, 8, np.zeros(64)), "target": image(8, 8, np.zeros(64 * 3), "Rgb"),
              "mask": mask(8, 8, np.zeros(64))}
    with pytest.raises(PaintopError) as ei:
        ops.PoissonBlend().compute(inputs, {})
    assert "channel layout" in ei.value.message
    with pytest.raises(PaintopError) as ei:
        ops.ScreenedPoisson().compute({"guidance": image(8, 8, np.zeros(64))}, {})
    assert ei.value.error_class == ErrorClass.REFERENCE
    assert ei.value.code == "E_SCREENED_POISSON_INPUT"


def test_manifests_declare_hip_beside_reference():
    for op in (ops.POISSON_BLEND_OP_ID, ops.SCREENED_POISSON_OP_ID):
        m = ops.load_manifest(op)
        assert m["id"] == op
        assert m["implementations"] == ["cpu.reference@1", "hip.poisson@1"]
        assert m["determinism"] == "bounded" and m["roi"]["category"] == "full-domain"
    names = [p["name"] for p in ops.PoissonBlend.manifest()["params"]]
    assert names == ["max_iterations", "tolerance"]
    assert "lambda" in [p["name"] for p in ops.ScreenedPoisson.manifest()["params"]]


class _Stub:
    def __init__(self, tag):
        self.tag = tag

    def compute(self, inputs, params):
        return {"candidate": self.tag, "report": self.tag}


def test_backend_selection_mirrors_dispatch():
    op = ops.POISSON_BLEND_OP_ID
    man = ops.load_manifest(op)
    reg = executor.ImplRegistry()
    reg.register(op, _Stub("ref"))
    hip_b = executor.BackendId("hip", "poisson")
    # declared but not registered: prefer falls back to the reference (recorded)
    sel = executor.select_backend(op, man, reg, executor.BackendPolicy.prefer([hip_b]))
    assert str(sel.impl_id) == "cpu.reference@1" and sel.fell_back
    # require without a registered kernel: Policy error
    with pytest.raises(PaintopError) as ei:
        executor.select_backend(op, man, reg, executor.BackendPolicy.require(hip_b))
    assert ei.value.code == "E_BACKEND_UNSUPPORTED" and ei.value.error_class == ErrorClass.POLICY
    executor.register_hip_backend(reg)
    sel = executor.select_backend(op, man, reg, executor.BackendPolicy.prefer([hip_b]))
    assert str(sel.impl_id) == "hip.poisson@1" and not sel.fell_back
    # the reference policy keeps the reference
    sel = executor.select_backend(op, man, reg, executor.BackendPolicy.reference())
    assert str(sel.impl_id) == "cpu.reference@1" and not sel.fell_back
    # undeclared backend is never selected even when registered
    undeclared = dict(man, implementations=["cpu.reference@1"])
    sel = executor.select_backend(op, undeclared, reg, executor.BackendPolicy.prefer([hip_b]))
    assert str(sel.impl_id) == "cpu.reference@1" and sel.fell_back
    # duplicates and reference ids are rejected by register_backend
    with pytest.raises(PaintopError):
        executor.register_hip_backend(reg)
    with pytest.raises(PaintopError):
        reg.register_backend(op, executor.ImplId.parse("cpu.reference@1"), _Stub("x"))


def test_dispatch_node_traces_selection_and_checks_outputs():
    op = ops.POISSON_BLEND_OP_ID
    man = ops.load_manifest(op)
    reg = executor.ImplRegistry()
    reg.register(op, _Stub("ref"))
    reg.register_backend(op, executor.ImplId.parse("hip.poisson@1"), _Stub("gpu"))
    trace = []
    out = executor.dispatch_node("n1", op, man, reg,
                                 executor.BackendPolicy.prefer([executor.BackendId("hip", "poisson")]),
                                 {}, {}, trace)
    assert out["candidate"] == "gpu"
    ev = [t["event"] for t in trace]
    assert ev == ["implementation_selected", "dispatch_started", "cache_lookup",
                  "dispatch_completed"]
    assert trace[0]["implementation"] == "hip.poisson@1"

    class Missing:
        def compute(self, i, p):
            return {"candidate": 1}

    reg2 = executor.ImplRegistry()
    reg2.register(op, Missing())
    with pytest.raises(PaintopError) as ei:
        executor.dispatch_node("n2", op, man, reg2, executor.BackendPolicy.reference(), {}, {}, [])
    assert ei.value.code == "E_OUTPUT_NOT_PRODUCED"


def test_impl_id_grammar():
    assert str(executor.ImplId.parse("hip.poisson@1")) == "hip.poisson@1"
    for bad in ("Hip.poisson@1", "hip.poisson@0", "hip@1", "hip.poisson"):
        with pytest.raises(PaintopError):
            executor.ImplId.parse(bad)


def test_star_polygon_area_is_about_30_percent():
    from paintop_amd.inputs import star_polygon
    n = 4096
    v = np.array(star_polygon(n / 2, n / 2, np.sqrt(0.3 * n * n / 1.5), 0.5, 6, 0.1, 301))
    x, y = v[:, 0], v[:, 1]
    area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    assert 0.25 < area / (n * n) < 0.35


def test_integration_rust_binding_matches_the_header():
    """The Rust FFI block shown in INTEGRATION.md declares every entry point of
    include/paintop_hip.h, and its PhSolveDesc / PhReport / PhBatchItem fields
    are the header's, in order."""
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    rust_fns = set(re.findall(r"pub fn (ph_[a-z0-9_]+)\(", md))
    assert rust_fns == set(declared_symbols())
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "paintop_hip.h")).read(),
                 flags=re.S)

    def c_fields(name):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), hdr, re.S).group(1)
        return re.findall(r"\**([a-z_0-9]+)(?:\[[^\]]*\])*;", body)

    def rust_fields(name):
        body = re.search(r"pub struct %s \{(.*?)\n?\}" % name, md, re.S).group(1)
        return re.findall(r"pub ([a-z_0-9]+):", body)

    for c, r in (("ph_solve_desc", "PhSolveDesc"), ("ph_report", "PhReport"),
                 ("ph_batch_item", "PhBatchItem")):
        assert c_fields(c) == rust_fields(r), (c, c_fields(c), rust_fields(r))


def test_tuner_rows_ignore_timing_noise():
    """The sweep-configuration tuner's table (paintop_amd/tune.py): a
    configuration within the slack of a problem's fastest counts as equal, so
    a near-tie does not split a row; rows end at the geometric mean of the
    neighbouring problems' work; the last row is unbounded."""
    from paintop_amd import tune

    n = len(tune.CONFIGS)
    works = [1e5, 4e5, 1.6e6, 6.4e6, 2.6e7]

    def t(best, near=()):
        ts = [10.0] * n
        ts[best] = 5.0
        for c in near:
            ts[c] = 5.1  # within the slack
        return ts

    # config 4 is fastest or within 3 % of it up to 6.4e6; config 0 above
    times = [t(3, near=(4,)), t(4), t(4, near=(5,)), t(5, near=(4,)), t(0)]
    rows = tune.table_rows(works, times)
    assert [(k, hb) for _, k, hb in rows] == [tune.CONFIGS[4], tune.CONFIGS[0]]
    assert rows[0][0] == pytest.approx((6.4e6 * 2.6e7) ** 0.5)
    assert rows[-1][0] == 1e300
    # input order does not matter
    perm = [3, 0, 4, 2, 1]
    assert tune.table_rows([works[i] for i in perm], [times[i] for i in perm]) == rows
    # a real split stays
    rows2 = tune.table_rows(works[:2], [t(3), t(5)])
    assert [(k, hb) for _, k, hb in rows2] == [tune.CONFIGS[3], tune.CONFIGS[5]]


def test_tuning_header_matches_the_tuner_format():
    """The committed table (paintop_amd/csrc/sor_tuning.h) has its rows in
    increasing work bounds, the last unbounded, with configurations the
    sweep kernel is built for (K in 8, 12, 16; even tile rows >= 16)."""
    import re

    text = open(os.path.join(ROOT, "paintop_amd", "csrc", "sor_tuning.h")).read()
    for name in ("PH_SOR_TUNING",):
        body = re.search(name + r"\[\] = \{(.*?)\};", text, re.S).group(1)
        rows = [(float(a), int(b), int(c)) for a, b, c in
                re.findall(r"\{([0-9.e+]+), (\d+), (\d+)\}", body)]
        assert rows and rows[-1][0] >= 1e299
        assert all(rows[i][0] < rows[i + 1][0] for i in range(len(rows) - 1))
        assert all(k in (8, 12, 16) and hb % 2 == 0 and 16 <= hb <= 1024 for _, k, hb in rows)
