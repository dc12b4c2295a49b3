"""The operator runner end to end on the device: ``python -m
flagsparse_amd.report`` for one operator (SDDMM: its GPU parity tests in a
child pytest, its quick benchmark cases in a child Python) writes the
reference's report files with Passed phases."""

import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_runner_one_operator(tmp_path):
    cmd = [sys.executable, "-m", "flagsparse_amd.report", "--ops", "sddmm_csr", "--phase", "both", "--quick",
           "--warmup", "2", "--iters", "5", "--timeout", "240", "--results-dir", str(tmp_path)]
    proc = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=540)
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-3000:]
    summary = json.loads((tmp_path / "summary.json").read_text())
    op = summary["result"]["sddmm_csr"]
    assert op["accuracy"]["status"] == "Passed", op["accuracy"]
    assert op["accuracy"]["passed"] > 0 and op["accuracy"]["failed"] == 0
    assert op["performance"]["status"] == "Passed", op["performance"]
    assert summary["env"]["torch"]["cuda_available"] is True
    acc = json.loads((tmp_path / "sddmm_csr" / "accuracy_result.json").read_text())
    assert acc and all(v["result"] == "passed" for v in acc.values())
    perf = json.loads((tmp_path / "sddmm_csr" / "performance_result.json").read_text())
    rows = perf["sddmm_csr"]["details"]
    assert rows and all(r["result"] for r in rows)
    html = (tmp_path / "result.html").read_text(encoding="utf-8")
    assert "FlagSparse Test Report" in html and "sddmm_csr" in html and "accuracy_result.json" in html
