"""The operator runner's report formats (FlagGems-style summary.json,
performance_result.json, accuracy_result.json, per-phase detail files) --
the contract of the reference's tests/ci/test_runner_result_format.py --
and one end-to-end runner pass on the CPU (GPU phases report Skipped)."""

import csv
import json
import re

from flagsparse_amd import report

ACC_KEYS = {"total", "skipped", "failed", "passed", "details", "status", "duration", "exit_code", "data_file"}
PERF_KEYS = {"duration", "exit_code", "data", "status", "data_file", "test_case"}


def _env():
    return {"python": {"version": "3.test"}, "platform": {"system": "Linux", "release": "r", "machine": "x86_64"},
            "packages": {"torch": {"version": "2.test"}, "triton": {"version": "3.test"},
                         "flagsparse": {"version": "1.test"}},
            "cuda": {"available": True, "device_count": 1, "devices": [{"name": "Test GPU"}]}}


def _phase(phase, status, **kw):
    base = {"operator": "spmv_csr", "phase": phase, "configured": True, "status": status, "returncode": 0,
            "exit_code": 0, "duration_sec": 1.25, "duration": 1.25, "command": ["python", "-m", "pytest"],
            "log_path": f"spmv_csr/{phase}_stdout.log", "stdout_log_path": f"spmv_csr/{phase}_stdout.log",
            "stderr_log_path": f"spmv_csr/{phase}_stderr.log", "data_file": f"spmv_csr/{phase}_result.json"}
    base.update(kw)
    return base


def test_summary_schema(tmp_path):
    results = [{"operator": "spmv_csr", "gpu": 0, "customized": True, "labels": ["flagsparse", "sparse"],
                "accuracy": _phase("accuracy", "PASS", passed=3, failed=0, skipped=0, errors=0, total=3),
                "performance": _phase("performance", "SKIP", row_count=0, test_case="csv")}]
    report.write_summary(results, tmp_path, _env())
    summary = json.loads((tmp_path / "summary.json").read_text())
    assert list(summary) == ["timestamp", "env", "result"]
    assert re.fullmatch(r"\d{4}-\d{2}-\d{2} \d{2}:\d{2}:\d{2}", summary["timestamp"])
    assert set(summary["env"]) == {"architecture", "os_name", "os_release", "python", "torch", "flagtree",
                                   "triton", "flag_gems"}
    assert set(summary["env"]["torch"]) == {"version", "cuda_available", "device_name", "device_count"}
    assert set(summary["env"]["flag_gems"]) == {"version", "vendor", "device"}
    assert summary["env"]["torch"]["device_name"] == "Test GPU"
    op = summary["result"]["spmv_csr"]
    assert list(op) == ["customized", "accuracy", "performance", "labels"]
    assert set(op["accuracy"]) == ACC_KEYS and op["accuracy"]["status"] == "Passed"
    assert op["accuracy"]["duration"] == 1.25 and op["accuracy"]["passed"] == 3
    assert set(op["performance"]) == PERF_KEYS and op["performance"]["status"] == "Skipped"
    assert op["performance"]["test_case"] == "csv"
    flat = json.loads((tmp_path / "summary_flat.json").read_text())
    assert flat["totals"]["by_phase"]["accuracy"]["PASS"] == 1
    with (tmp_path / "summary.csv").open(newline="") as f:
        rows = list(csv.DictReader(f))
    assert rows[0]["operator"] == "spmv_csr" and rows[0]["phase"] == "accuracy"
    assert rows[0]["duration"] == "1.25" and rows[0]["stderr_log_path"] == "spmv_csr/accuracy_stderr.log"


def test_result_html(tmp_path):
    """result.html carries the anchors the reference's runner-format test
    checks (tests/ci/test_runner_result_format.py:166-177)."""
    (tmp_path / "spmv_csr").mkdir()
    (tmp_path / "spmv_csr" / "accuracy_result.json").write_text('{"x::y": "passed"}')
    perf = _phase("performance", "PASS", row_count=2, speedup=2.5,
                  records=[{"dtype": "float32", "speedup": "2.0"}, {"dtype": "float64", "speedup": "3.0"}])
    results = [{"operator": "spmv_csr", "gpu": 0, "customized": True, "labels": [],
                "accuracy": _phase("accuracy", "PASS", passed=3, failed=0, skipped=1, errors=0, total=4),
                "performance": perf}]
    report.write_summary(results, tmp_path, _env())
    html = (tmp_path / "result.html").read_text(encoding="utf-8")
    for anchor in ("FlagSparse Test Report", "<h3>Test Environment</h3>", "<h3>Test Result</h3>",
                   'class="table-stats"', "AccRes<br>", "PerfRes<br>", "OPAverageSpeedUp", "filterTable()",
                   "sortTable(5, 'asc')", "spmv_csr", "accuracy_result.json", "3/0/1", "2.500", "Test GPU"):
        assert anchor in html, anchor
    assert "<td>2.000</td>" in html and "<td>3.000</td>" in html  # fp32 / fp64 columns


def test_schema_fixed_on_errors_and_missing_phase():
    acc = report._flaggems_accuracy_result({"status": "CRASH", "errors": 2, "reason": "x", "exit_code": 17})
    perf = report._flaggems_performance_result({"status": "CRASH", "reason": "y", "exit_code": 18})
    assert set(acc) == ACC_KEYS and "errors" not in acc and acc["status"] == "Error"
    assert set(perf) == PERF_KEYS and "reason" not in perf
    s = report._operator_summary({"operator": "spmv_csr", "accuracy": {"phase": "accuracy", "status": "PASS"}})
    assert list(s) == ["customized", "accuracy", "performance", "labels"]
    assert s["performance"] == {"duration": 0.0, "exit_code": 0, "data_file": "", "data": {}, "status": "NotFound",
                                "test_case": "Unknown"}


def test_phase_detail_file(tmp_path):
    op_dir = tmp_path / "spmv_csr"
    op_dir.mkdir()
    (op_dir / "accuracy_result.json").write_text('{"raw": {"result": "passed"}}')
    r = _phase("accuracy", "FAIL", returncode=1, exit_code=1, passed=1, failed=1, total=2,
               failures=["FAILED tests/test_x.py::test_case"],
               tests=[{"nodeid": "tests/test_x.py::test_case[a]", "status": "Failed", "status_raw": "FAILED"}])
    path = report.write_phase_result(op_dir, "accuracy", r)
    assert json.loads(path.read_text()) == {"raw": {"result": "passed"}}
    d = json.loads((op_dir / "accuracy_detail.json").read_text())
    assert ACC_KEYS <= set(d) and d["status"] == "Failed" and d["status_raw"] == "FAIL"
    assert d["data_file"] == "spmv_csr/accuracy_result.json"
    assert d["details"]["failed"] == ["FAILED tests/test_x.py::test_case"]
    assert d["summary"]["failed"] == 1 and d["tests"][0]["nodeid"].endswith("[a]")


def test_pytest_output_parsing():
    out = """
tests/test_spmv_gpu.py::test_a[float32-32] PASSED
tests/test_spmv_gpu.py::test_b[float64] SKIPPED (no GPU)
tests/test_spmv_gpu.py::test_c FAILED
FAILED tests/test_spmv_gpu.py::test_c - AssertionError
===== 1 failed, 1 passed, 1 skipped, 2 deselected in 0.52s =====
"""
    cases = report.parse_pytest_cases(out)
    assert [c["status"] for c in cases] == ["Passed", "Skipped", "Failed"]
    assert cases[0]["parameters"] == {"param_0": "float32", "param_1": "32"}
    assert cases[1]["message"] == "(no GPU)"
    counts = report.pytest_counts(out)
    assert (counts["passed"], counts["failed"], counts["skipped"], counts["deselected"]) == (1, 1, 1, 2)
    assert report.status_from_pytest_counts(counts, 1) == "FAIL"
    assert report.status_from_pytest_counts(dict(counts, failed=0), 0) == "PASS"
    assert report.status_from_pytest_counts(dict.fromkeys(counts, 0), 5) == "NO_TESTS"
    assert report.status_from_pytest_counts(dict.fromkeys(counts, 0), 3) == "CRASH"


def test_accuracy_json_parsing(tmp_path):
    p = tmp_path / "accuracy_result.json"
    p.write_text(json.dumps({
        "tests/t.py::test_ok[d0]": {"params": {"dtype": "float32"}, "result": "passed", "opname": ["gpu"]},
        "tests/t.py::test_bad[d0]": {"params": {"dtype": "float64"}, "result": "failed", "reason": "AssertionError"},
        "tests/t.py::test_skip": {"params": {}, "result": "skipped", "reason": "no GPU"},
    }))
    r = report.parse_accuracy_json(p)
    assert (r["status"], r["total"], r["passed"], r["failed"], r["skipped"]) == ("Failed", 3, 1, 1, 1)
    assert r["details"]["failed"]["AssertionError"] == ["tests/t.py::test_bad:dtype=float64"]


def test_performance_csv_to_json(tmp_path):
    csv_path = tmp_path / "performance.csv"
    csv_path.write_text("dtype,shape,op,triton_ms,pytorch_ms,triton_speedup_vs_pytorch,status\n"
                        "float32,64x64x100,non,0.5,1.0,2.0,PASS\nfloat32,128x128x400,non,0.25,1.0,4.0,PASS\n")
    json_path = tmp_path / "performance_result.json"
    report.write_benchmark_json_from_csv("spmv_csr", csv_path, json_path)
    raw = json.loads(json_path.read_text())
    assert raw["spmv_csr"]["result"] == "passed" and raw["spmv_csr"]["test_case"] == "csv"
    assert raw["spmv_csr"]["details"][0]["dtype"] == "float32"
    parsed = report.parse_performance_json("spmv_csr", json_path)
    assert parsed["status"] == "Passed"
    d = parsed["data"]["float32"]
    assert d["details"]["64x64x100"] == {"base": 1.0, "gems": 0.5, "speedup": 2.0}
    assert d["speedup"] == 3.0
    s = report.summarize_performance_csv(csv_path)
    assert s["row_count"] == 2 and s["speedup"] == 3.0
    entry = s["benchmark"]["float32"]["64x64x100"][0]
    assert entry["metrics"]["triton_ms"] == 0.5 and entry["metadata"]["status"] == "PASS"
    strict = report._strict_perf_data(s["data"])
    assert set(strict) == {"fp32"} and strict["fp32"]["details"]["64x64x100|op=non"]["gems"] == 0.5


def test_should_fail():
    assert report._should_fail([{"operator": "x", "accuracy": {"phase": "accuracy", "status": "Failed"}}])
    assert not report._should_fail([{"operator": "x", "accuracy": {"phase": "accuracy", "status": "SKIP"}}])


def test_runner_end_to_end_on_cpu(tmp_path):
    # no GPU here: the parity tests are collected and skipped, the
    # performance child reports that no GPU is present
    results = report.run(["spmv_csr"], tmp_path, phase="both", quick=True, timeout=600)
    acc, perf = results[0]["accuracy"], results[0]["performance"]
    assert acc["status"] == "SKIP" and acc["skipped"] > 0 and acc["failed"] == 0
    assert perf["status"] == "SKIP"
    raw = json.loads((tmp_path / "spmv_csr" / "accuracy_result.json").read_text())
    assert raw and all(v["result"] == "skipped" for v in raw.values())
    summary = json.loads((tmp_path / "summary.json").read_text())
    assert summary["result"]["spmv_csr"]["accuracy"]["status"] == "Skipped"
    assert summary["result"]["spmv_csr"]["accuracy"]["data_file"] == "spmv_csr/accuracy_result.json"
    assert summary["result"]["spmv_csr"]["performance"]["status"] == "Skipped"
