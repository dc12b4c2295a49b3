"""pytest plugin that records per-test outcomes as the FlagGems-style
``accuracy_result.json`` (format of the reference's tests/pytest/conftest.py:
28-140): ``{nodeid: {"params": {...}, "result": "passed" | "failed" |
"skipped", "reason": str | None, "opname": [marks]}}``.

Load with ``-p flagsparse_amd._pytest_record --record json --output PATH``.
"""

import json
from pathlib import Path

import pytest

_BUILTIN_MARKS = {"filterwarnings", "parametrize", "skip", "skipif", "timeout", "tryfirst", "trylast",
                  "usefixtures", "xfail"}
_STATE = {"enabled": False, "path": "accuracy_result.json", "results": {}}


def pytest_addoption(parser):
    group = parser.getgroup("flagsparse_amd record")
    group.addoption("--record", action="store", default="none", choices=["none", "log", "json"],
                    help="record test results as JSON (accuracy_result.json format)")
    group.addoption("--output", default="accuracy_result.json", help="path of the JSON result file")


def pytest_configure(config):
    _STATE["enabled"] = config.getoption("--record") == "json"
    _STATE["path"] = config.getoption("--output") or "accuracy_result.json"
    _STATE["results"] = {}


def _jsonable(v):
    return v if isinstance(v, (bool, int, float, str, type(None))) else str(v)


def _reason(report):
    longrepr = getattr(report, "longrepr", None)
    if hasattr(longrepr, "reprcrash"):
        return longrepr.reprcrash.message
    if isinstance(longrepr, tuple):
        return str(longrepr[2])
    return str(longrepr)


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    outcome = yield
    if not _STATE["enabled"]:
        return
    report = outcome.get_result()
    callspec = getattr(item, "callspec", None)
    entry = _STATE["results"].setdefault(item.nodeid, {
        "params": {k: _jsonable(v) for k, v in (callspec.params.items() if callspec else [])},
        "result": None,
        "opname": [m.name for m in item.iter_markers() if m.name not in _BUILTIN_MARKS],
    })
    if report.when == "setup" and report.outcome == "skipped":
        entry["result"], entry["reason"] = "skipped", _reason(report)
    elif report.when == "setup" and report.outcome == "failed":
        entry["result"], entry["reason"] = "failed", _reason(report)
    elif report.when == "call":
        entry["result"] = report.outcome
        entry["reason"] = _reason(report) if report.outcome != "passed" else None
    elif report.when == "teardown" and report.outcome == "failed":
        entry["result"], entry["reason"] = "failed", _reason(report)


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not _STATE["enabled"]:
        return
    path = Path(_STATE["path"])
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(_STATE["results"], indent=2, sort_keys=True), encoding="utf-8")
