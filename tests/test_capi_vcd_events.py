"""C ABI exports, VCD writer details and the timing report (CPU only)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from jacquard_amd import _native as N
from jacquard_amd import events, sim, vcd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    with open(os.path.join(ROOT, "include", "jacquard_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)


def test_hip_library_exports_every_declared_symbol():
    names = [n for n in _declared_functions() if n not in ("if", "sizeof")]
    assert "simulate_v1_noninteractive_simple_scan_hip" in names
    assert "simulate_v1_noninteractive_timed_hip" in names
    lib = ctypes.CDLL(os.path.join(ROOT, "jacquard_amd", "lib", "libjacquard_hip.so"))
    missing = []
    for n in names:
        try:
            getattr(lib, n)
        except AttributeError:
            missing.append(n)
    assert not missing, missing
    assert N.missing_hip_symbols() == []


def test_id_codes():
    assert [vcd.id_code(i) for i in (0, 1, 93)] == ["!", '"', "~"]
    assert vcd.id_code(94) == "!!"
    assert vcd.id_code(95) == '"!'


class _FakeScript:
    def __init__(self, rio):
        self.reg_io_state_size = rio


def _timed(mapping, offs, states, arr, ts):
    import types
    s = _FakeScript(2)
    orig = vcd.output_mapping
    try:
        vcd.output_mapping = lambda script: mapping
        return vcd.write_output_vcd_timed(s, offs, np.asarray(states, np.uint32),
                                          np.asarray(arr, np.uint32), ts, "test")
    finally:
        vcd.output_mapping = orig


def test_timed_vcd_offsets():
    """src/sim/vcd_io.rs:1335-1400: A at 350 ps, B at 885 ps after 10000."""
    out = _timed([(2, 0, "A"), (4, 32, "B")], [(0, 0), (2, 10000)],
                 [0, 0, 1, 1], [0, 0, 350, 885], (1, "ps"))
    assert "#10350" in out and "#10885" in out and "#0" in out
    assert out.index("#10350") < out.index("#10885")


def test_timed_vcd_ns_rounding():
    """500 ps at a 1 ns timescale rounds to one unit: t = 11."""
    out = _timed([(2, 0, "A")], [(0, 0), (1, 10)], [0, 1], [0, 500], (1, "ns"))
    assert "#11\n" in out


def test_timed_vcd_zero_arrival():
    out = _timed([(2, 0, "A")], [(0, 0), (1, 5000)], [0, 1], [0, 0], (1, "ps"))
    assert "#5000\n1!" in out


def _ev(cycle, kind, word, slack, arrival, lim):
    e = np.zeros(1, events.np.dtype([("event_type", "<u4"), ("message_id", "<u4"),
                                     ("cycle", "<u4"), ("reserved", "<u4"),
                                     ("data", "<u4", (4,))]))
    e["event_type"] = kind
    e["message_id"] = word
    e["cycle"] = cycle
    e["data"] = [word, np.int32(slack).view(np.uint32), arrival, lim]
    return e


def test_timing_report_matches_reference_fixture():
    """The violation report built from raw device events (in a scrambled
    order, as atomics would leave them) against the reference's fixture
    tests/timing_ir/sample_reports/two_violations.json (copied verbatim to
    tests/golden/timing_ir/), field by field.  The fixture is pinned at
    schema 1.0.0; a 1.1.0 report differs from it only in schema_version and
    the 1.1.0 field stats.violations_truncated (0 here), which 1.0.0 readers
    default (src/timing_report.rs)."""
    with open(os.path.join(ROOT, "tests", "golden", "timing_ir", "two_violations.json")) as f:
        fix = json.load(f)
    ev = np.concatenate([_ev(91, 5, 12, -40, 10, 50), _ev(87, 4, 5, -150, 950, 200),
                         _ev(42, 4, 5, -100, 900, 200)])
    sites = {v["word_id"]: v["site"] for v in fix["violations"]}
    md = fix["metadata"]
    rep = events.build_report(ev, md["clock_period_ps"], md["cycles_run"],
                              design=md["design"], vector_source=md["vector_source"],
                              timing_source=md["timing_source"], describe=sites.get)
    assert fix["schema_version"] == "1.0.0" and rep["schema_version"] == "1.1.0"
    assert rep["metadata"] == md
    stats = dict(rep["stats"])
    assert stats.pop("violations_truncated") == 0
    assert stats == fix["stats"]
    for key in ("violations", "per_word", "worst_slack"):
        assert rep[key] == fix[key], key
    assert set(rep) == set(fix) - {"_comment"}
    # the report's JSON text parses back to the same report
    assert json.loads(events.report_json(rep)) == rep
    text = events.format_summary(rep)
    assert "Setup: 2" in text and "Hold:  1" in text


def test_report_truncation():
    ev = np.concatenate([_ev(c, 5, 1, -1, 0, 1) for c in range(5)])
    rep = events.build_report(ev, 5, 100, max_violations=3)
    assert len(rep["violations"]) == 3
    assert rep["stats"]["violations_truncated"] == 2
    assert rep["stats"]["hold_violations"] == 5


@pytest.mark.parametrize("n,seed,cap", [(0, 1, None), (1, 2, 0), (37, 3, 7), (5000, 4, None),
                                        (20000, 5, 1000), (20000, 6, events.DEFAULT_MAX_VIOLATIONS)])
def test_batch_report_equals_the_incremental_builder(n, seed, cap):
    """build_report (grouping over the whole array) against ReportBuilder
    fed one record at a time: the same JSON bytes, on random events with
    many equal slacks (the worst-slack lists' replacement rule decides
    between them), other event types mixed in, and record caps."""
    rng = np.random.default_rng(seed)
    ev = np.zeros(n, sim.EVENT_DTYPE)
    ev["event_type"] = rng.choice([1, 4, 4, 5, 5, 7], n)
    ev["cycle"] = rng.integers(0, 300, n)
    ev["message_id"] = rng.integers(0, 60, n)
    ev["data"][:, 0] = ev["message_id"]
    ev["data"][:, 1] = rng.integers(-40, 0, n).astype(np.int32).view(np.uint32)
    ev["data"][:, 2] = rng.integers(0, 5000, n)
    ev["data"][:, 3] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    kw = dict(describe=lambda w: f"w{w % 7}/{w}", design="d", events_dropped=3)
    got = events.build_report(ev, 300, 5000, max_violations=cap, **kw)
    want = events.build_report_sequential(ev, 300, 5000, max_violations=cap, **kw)
    assert events.report_json(got) == events.report_json(want)
