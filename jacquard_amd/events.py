"""Violation records and the end-of-run timing report.

Decodes the device's setup/hold event records (csrc/event_buffer.h:35-41:
type 4 = setup, 5 = hold, data = [word, slack as u32, arrival, limit]) into
``ViolationRecord``s and folds them into the JSON report of
src/timing_report.rs (schema 1.1.0): bounded per-cycle list, per-word
aggregates sorted by violation count then word id, top-N worst slack per
kind sorted by slack then cycle.  The device appends events with atomics in
arbitrary order; records are observed in canonical (cycle, word, kind)
order so the report is deterministic.
"""
import json
from dataclasses import dataclass, asdict, field
from typing import Dict, List, Optional

import numpy as np

SCHEMA_VERSION = "1.1.0"
DEFAULT_MAX_VIOLATIONS = 100_000
WORST_SLACK_TOP_N = 10
EVENT_SETUP, EVENT_HOLD = 4, 5


@dataclass
class ViolationRecord:
    cycle: int
    kind: str  # "setup" | "hold"
    word_id: int
    site: str
    arrival_ps: int
    constraint_ps: int
    slack_ps: int


@dataclass
class PerWordSummary:
    word_id: int
    site: str
    setup_violations: int = 0
    hold_violations: int = 0
    worst_setup_slack_ps: Optional[int] = None
    worst_hold_slack_ps: Optional[int] = None
    worst_arrival_ps: Optional[int] = None


class WordSymbolMap:
    """State word -> the DFFs packed into it (src/flatten.rs:125-178,
    1848-1879), for symbolic violation sites."""

    def __init__(self, word_to_dffs):
        # {word: [(name, bit)]} sorted by bit
        self.word_to_dffs = {int(w): sorted(v, key=lambda d: d[1])
                             for w, v in word_to_dffs.items()}

    @classmethod
    def from_script(cls, script):
        """DFF j sits at the state bit its data input is written to
        (output_map of its D literal); names are the netlist's DFF instance
        names."""
        A = script.netlist.arrays()
        names = script.netlist.dff_names()
        m = {}
        for j, d in enumerate(A["dff_d"].tolist()):
            pos = script.output_map.get(int(d))
            if pos is None:
                continue
            name = names[j] if j < len(names) and names[j] else f"dff#{j}"
            m.setdefault(pos // 32, []).append((name, pos % 32))
        return cls(m)

    def describe_word(self, word_id, max_names=4):
        """'name[bit b], ... [+N more] [word=W]' (src/flatten.rs:150-170)."""
        dffs = self.word_to_dffs.get(int(word_id))
        if not dffs:
            return f"<no DFF in word> [word={word_id}]"
        shown = min(len(dffs), max(max_names, 1))
        parts = [f"{n}[bit {b}]" for n, b in dffs[:shown]]
        if len(dffs) > shown:
            parts.append(f"+{len(dffs) - shown} more")
        return f"{', '.join(parts)} [word={word_id}]"

    def num_sites(self):
        return sum(len(v) for v in self.word_to_dffs.values())

    def num_words(self):
        return len(self.word_to_dffs)


def default_site(word_id):
    """Fallback site descriptor (src/event_buffer.rs:264-269)."""
    return f"word={word_id}"


def records_from_events(events, describe=default_site):
    """Canonically ordered ViolationRecords from raw event records."""
    from .sim import canonical_events
    out = []
    for e in canonical_events(events):
        t = int(e["event_type"])
        if t not in (EVENT_SETUP, EVENT_HOLD):
            continue
        d = e["data"]
        slack = int(np.array(d[1], np.uint32).view(np.int32))
        out.append(ViolationRecord(
            cycle=int(e["cycle"]), kind="setup" if t == EVENT_SETUP else "hold",
            word_id=int(d[0]), site=describe(int(d[0])), arrival_ps=int(d[2]),
            constraint_ps=int(d[3]), slack_ps=slack))
    return out


class ReportBuilder:
    def __init__(self, metadata: Dict, worst_slack_n=WORST_SLACK_TOP_N,
                 max_violations: Optional[int] = DEFAULT_MAX_VIOLATIONS):
        self.metadata = dict(metadata)
        self.violations: List[ViolationRecord] = []
        self.per_word: Dict[int, PerWordSummary] = {}
        self.worst = {"setup": [], "hold": []}
        self.n = max(1, worst_slack_n)
        self.max_violations = max_violations
        self.truncated = 0

    def _push_worst(self, v):
        items = self.worst[v.kind]
        if len(items) < self.n:
            items.append(v)
            return
        # replace the least negative slack; on ties the last one, as
        # Iterator::max_by returns the last of equal maxima
        mx = max(it.slack_ps for it in items)
        idx = max(i for i, it in enumerate(items) if it.slack_ps == mx)
        if v.slack_ps < items[idx].slack_ps:
            items[idx] = v

    def observe(self, v: ViolationRecord):
        self._push_worst(v)
        w = self.per_word.get(v.word_id)
        if w is None:
            w = self.per_word[v.word_id] = PerWordSummary(v.word_id, v.site)
        if v.kind == "setup":
            w.setup_violations += 1
            w.worst_setup_slack_ps = (v.slack_ps if w.worst_setup_slack_ps is None
                                      else min(w.worst_setup_slack_ps, v.slack_ps))
        else:
            w.hold_violations += 1
            w.worst_hold_slack_ps = (v.slack_ps if w.worst_hold_slack_ps is None
                                     else min(w.worst_hold_slack_ps, v.slack_ps))
        w.worst_arrival_ps = (v.arrival_ps if w.worst_arrival_ps is None
                              else max(w.worst_arrival_ps, v.arrival_ps))
        if self.max_violations is not None and len(self.violations) >= self.max_violations:
            self.truncated += 1
        else:
            self.violations.append(v)

    def finalize(self, cycles_run, events_dropped=0):
        md = dict(self.metadata)
        md["cycles_run"] = int(cycles_run)
        setup = sum(w.setup_violations for w in self.per_word.values())
        hold = sum(w.hold_violations for w in self.per_word.values())
        per_word = sorted(self.per_word.values(),
                          key=lambda w: (-(w.setup_violations + w.hold_violations), w.word_id))
        ws = {k: sorted(v, key=lambda r: (r.slack_ps, r.cycle)) for k, v in self.worst.items()}
        return {
            "schema_version": SCHEMA_VERSION,
            "metadata": md,
            "stats": {"setup_violations": setup, "hold_violations": hold,
                      "events_dropped": int(events_dropped),
                      "violations_truncated": self.truncated},
            "violations": [asdict(v) for v in self.violations],
            "per_word": [asdict(w) for w in per_word],
            "worst_slack": {k: [asdict(r) for r in v] for k, v in ws.items()},
        }


def _worst_of_kind(rec, slack, n_top):
    """ReportBuilder._push_worst over the records `rec` (indices, in
    canonical order) of one kind: the first n_top fill the list; afterwards
    a record replaces the last of the list's largest slacks when its slack
    is smaller.  That largest slack never rises, so once the list is full
    only records below its value at that point can change it: the others
    are skipped without changing the result."""
    items = [int(i) for i in rec[:n_top]]
    if len(rec) <= n_top:
        return items
    mx = max(int(slack[i]) for i in items)
    rest = rec[n_top:]
    for i in rest[slack[rest] < mx].tolist():
        s = int(slack[i])
        if s >= mx:
            continue
        j = max(k for k, it in enumerate(items) if int(slack[it]) == mx)
        items[j] = i
        mx = max(int(slack[it]) for it in items)
    return items


def build_report(events, cycles_run, clock_period_ps, design=None,
                 vector_source=None, timing_source=None, describe=default_site,
                 max_violations=DEFAULT_MAX_VIOLATIONS, events_dropped=0,
                 version="0.1.0"):
    """The report ReportBuilder gives after observing every record in
    canonical order, computed over the whole event array at once: per-word
    aggregates by grouping, the bounded record list from its first entries,
    and the worst-slack lists by the same replacement rule.  describe is
    called once per word."""
    from .sim import canonical_events
    md = dict(design=design, vector_source=vector_source, timing_source=timing_source,
              clock_period_ps=int(clock_period_ps), cycles_run=int(cycles_run),
              jacquard_version=version)
    ev = canonical_events(events)
    if len(ev):
        ev = ev[(ev["event_type"] == EVENT_SETUP) | (ev["event_type"] == EVENT_HOLD)]
    n = len(ev)
    cyc = ev["cycle"].astype(np.int64)
    word = ev["data"][:, 0].astype(np.int64)
    slack = np.ascontiguousarray(ev["data"][:, 1]).view(np.int32).astype(np.int64)
    arr = ev["data"][:, 2].astype(np.int64)
    lim = ev["data"][:, 3].astype(np.int64)
    setup = ev["event_type"] == EVENT_SETUP
    sites = {}

    def site(w):
        if w not in sites:
            sites[w] = describe(w)
        return sites[w]

    def rec_dict(i):
        w = int(word[i])
        return {"cycle": int(cyc[i]), "kind": "setup" if setup[i] else "hold",
                "word_id": w, "site": site(w), "arrival_ps": int(arr[i]),
                "constraint_ps": int(lim[i]), "slack_ps": int(slack[i])}

    keep = n if max_violations is None else min(n, max_violations)
    violations = [rec_dict(i) for i in range(keep)]
    per_word = []
    if n:
        uw, inv = np.unique(word, return_inverse=True)
        u = uw.size
        n_setup = np.bincount(inv[setup], minlength=u)
        n_hold = np.bincount(inv[~setup], minlength=u)
        big = np.iinfo(np.int64).max
        ws = np.full(u, big, np.int64)
        np.minimum.at(ws, inv[setup], slack[setup])
        wh = np.full(u, big, np.int64)
        np.minimum.at(wh, inv[~setup], slack[~setup])
        wa = np.full(u, -1, np.int64)
        np.maximum.at(wa, inv, arr)
        for k in np.lexsort((uw, -(n_setup + n_hold))).tolist():
            w = int(uw[k])
            per_word.append({
                "word_id": w, "site": site(w),
                "setup_violations": int(n_setup[k]), "hold_violations": int(n_hold[k]),
                "worst_setup_slack_ps": int(ws[k]) if n_setup[k] else None,
                "worst_hold_slack_ps": int(wh[k]) if n_hold[k] else None,
                "worst_arrival_ps": int(wa[k])})
    n_top = max(1, WORST_SLACK_TOP_N)
    worst = {}
    for kind, mask in (("setup", setup), ("hold", ~setup)):
        items = _worst_of_kind(np.flatnonzero(mask), slack, n_top)
        items.sort(key=lambda i: (int(slack[i]), int(cyc[i])))
        worst[kind] = [rec_dict(i) for i in items]
    return {
        "schema_version": SCHEMA_VERSION,
        "metadata": md,
        "stats": {"setup_violations": int(setup.sum()), "hold_violations": int(n - setup.sum()),
                  "events_dropped": int(events_dropped),
                  "violations_truncated": n - keep},
        "violations": violations,
        "per_word": per_word,
        "worst_slack": worst,
    }


def build_report_sequential(events, cycles_run, clock_period_ps, design=None,
                            vector_source=None, timing_source=None, describe=default_site,
                            max_violations=DEFAULT_MAX_VIOLATIONS, events_dropped=0,
                            version="0.1.0"):
    """build_report through ReportBuilder, one record at a time (the
    reference's incremental builder, src/timing_report.rs)."""
    rb = ReportBuilder(dict(design=design, vector_source=vector_source,
                            timing_source=timing_source,
                            clock_period_ps=int(clock_period_ps), cycles_run=0,
                            jacquard_version=version),
                       max_violations=max_violations)
    for r in records_from_events(events, describe):
        rb.observe(r)
    return rb.finalize(cycles_run, events_dropped)


def report_json(report):
    return json.dumps(report, indent=2)


def format_summary(report):
    """Human summary (TimingReport::format_summary, src/timing_report.rs:151)."""
    md, st = report["metadata"], report["stats"]
    lines = ["=== Jacquard Timing Summary ==="]
    if md.get("design"):
        lines.append(f"Design:        {md['design']}")
    if md.get("vector_source"):
        lines.append(f"Vectors:       {md['vector_source']} ({md['cycles_run']} cycles)")
    else:
        lines.append(f"Vectors:       {md['cycles_run']} cycles")
    lines.append(f"Clock period:  {md['clock_period_ps']} ps")
    if md.get("timing_source"):
        lines.append(f"Timing source: {md['timing_source']}")
    lines += ["", "Violations:", f"  Setup: {st['setup_violations']}",
              f"  Hold:  {st['hold_violations']}",
              f"  Total: {st['setup_violations'] + st['hold_violations']}"]
    if st["events_dropped"]:
        lines.append(f"  Note:  {st['events_dropped']} events dropped (buffer overflow)")
    if st.get("violations_truncated"):
        lines.append(f"  Note:  {st['violations_truncated']} violation records dropped from "
                     "the per-cycle list (cap reached); totals + worst-slack still "
                     "reflect all events")
    lines += ["", "Worst slack:"]
    for label, key in (("Setup", "setup"), ("Hold", "hold")):
        v = report["worst_slack"][key]
        if v:
            lines.append(f"  {label}: {v[0]['slack_ps']}ps  at {v[0]['site']}  "
                         f"(cycle {v[0]['cycle']})")
        else:
            lines.append(f"  {label}: (none)")
    pw = report["per_word"]
    if pw:
        shown = min(5, len(pw))
        lines += ["", f"Top {shown} by violation count (of {len(pw)} total words with violations):"]
        fmt = lambda x: "-" if x is None else f"{x}ps"
        for w in pw[:shown]:
            lines.append(f"  {w['site']} ({w['setup_violations'] + w['hold_violations']} "
                         f"violations): worst setup={fmt(w['worst_setup_slack_ps'])} "
                         f"hold={fmt(w['worst_hold_slack_ps'])} arrival={fmt(w['worst_arrival_ps'])}")
    return "\n".join(lines) + "\n"
