"""VCD stimulus reader and output writers with the reference's semantics.

Reader: ``parse_input_vcd`` follows src/sim/vcd_io.rs:347-455 -- a snapshot
is pushed at the first timestamp after one at which a clock edge happened,
data inputs change immediately, clock-pin changes are applied one
timestamp later, edge flags are cleared after every pushed snapshot, and
vector values are applied character by character from the MSB side
(``b11`` on a 4-bit vector sets bits 3 and 2 and leaves bits 1 and 0 as
they were, src/sim/vcd_io.rs:279-330, 403-443).

Writers: ``write_output_vcd`` (plain / X) and ``write_output_vcd_timed``
follow src/sim/vcd_io.rs:561-946 and the vcd-ng/rust-vcd text layout
(identifier codes '!'..'~' in base 94; beyond 94 wires this encoding is the
rust-vcd convention and is not pinned by any reference fixture).
"""
import ctypes
import re

import numpy as np

from . import _native as N

U64_MAX = (1 << 64) - 1
_PS_DIVISOR = {"s": 1, "ms": 10**3, "us": 10**6, "ns": 10**9, "ps": 10**12, "fs": 10**15}


# ---------------------------------------------------------------- header
class Var:
    def __init__(self, vtype, size, code, ref, index):
        self.vtype, self.size, self.code, self.ref, self.index = vtype, size, code, ref, index


class Scope:
    def __init__(self, kind, name, parent=None):
        self.kind, self.name, self.parent = kind, name, parent
        self.children = []  # Var or Scope

    def path(self):
        p, s = [], self
        while s is not None and s.name is not None:
            p.append(s.name)
            s = s.parent
        return "/".join(reversed(p))


class Header:
    def __init__(self):
        self.timescale = None  # (ratio, unit)
        self.root = Scope("root", None)


def _tokens(text):
    return text.split()


def parse_header(text):
    """Returns (Header, body_text)."""
    m = re.search(r"\$enddefinitions\s+\$end", text)
    if not m:
        raise ValueError("VCD: no $enddefinitions")
    head, body = text[:m.end()], text[m.end():]
    toks = _tokens(head)
    h = Header()
    cur = h.root
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == "$timescale":
            j = toks.index("$end", i)
            spec = "".join(toks[i + 1:j])
            mm = re.match(r"(\d+)\s*([a-z]+)", spec)
            h.timescale = (int(mm.group(1)), mm.group(2))
            i = j + 1
        elif t == "$scope":
            s = Scope(toks[i + 1], toks[i + 2], cur)
            cur.children.append(s)
            cur = s
            i = toks.index("$end", i) + 1
        elif t == "$upscope":
            cur = cur.parent
            i = toks.index("$end", i) + 1
        elif t == "$var":
            j = toks.index("$end", i)
            f = toks[i + 1:j]
            vtype, size, code, ref = f[0], int(f[1]), f[2], f[3]
            index = None
            rest = "".join(f[4:])
            if not rest and "[" in ref and ref.endswith("]"):
                ref, rest = ref[:ref.index("[")], ref[ref.index("["):]
            if rest:
                mm = re.match(r"\[(-?\d+)(?::(-?\d+))?\]", rest)
                if mm.group(2) is None:
                    index = ("bit", int(mm.group(1)))
                else:
                    index = ("range", int(mm.group(1)), int(mm.group(2)))
            cur.children.append(Var(vtype, size, code, ref, index))
            i = j + 1
        elif t.startswith("$"):
            try:
                i = toks.index("$end", i) + 1
            except ValueError:
                i += 1
        else:
            i += 1
    return h, body


def _scopes(s):
    for c in s.children:
        if isinstance(c, Scope):
            yield c
            yield from _scopes(c)


def find_scope(header, path):
    for s in _scopes(header.root):
        if s.path() == path.strip("/"):
            return s
    raise ValueError(f"VCD scope {path!r} not found")


def _range(a, b):
    step = 1 if b >= a else -1
    return list(range(a, b + step, step))


def match_vcd_inputs(scope, input_names):
    """{(code, pos): port bit name} for the scope's vars that name input
    ports (src/sim/vcd_io.rs:279-344)."""
    names = set(input_names)
    out = {}
    for v in scope.children:
        if not isinstance(v, Var):
            continue

        def one(bitname, pos):
            if bitname in names:
                out[(v.code, pos)] = bitname

        if v.index is None:
            if v.size == 1:
                one(v.ref, 0)
            else:
                for pos, i in enumerate(reversed(range(v.size))):
                    one(f"{v.ref}[{i}]", pos)
        elif v.index[0] == "bit":
            one(f"{v.ref}[{v.index[1]}]", 0)
        else:
            for pos, i in enumerate(_range(v.index[1], v.index[2])):
                one(f"{v.ref}[{i}]", pos)
    return out


def auto_scope(header, input_names):
    """The first scope (depth first) whose variables cover every input."""
    best = None
    for s in _scopes(header.root):
        m = match_vcd_inputs(s, input_names)
        covered = set(m.values())
        if covered >= set(input_names):
            return s
        if best is None or len(covered) > best[0]:
            best = (len(covered), s)
    if best is None:
        raise ValueError("VCD: no scope")
    return best[1]


# ---------------------------------------------------------------- reader
def parse_input_vcd(text, script, scope=None, max_clock_edges=None):
    """(input_states uint32[(C+1)*rio], offsets_timestamps [(offset, ts)],
    header)."""
    header, body, code_map, flag_bits, rio = input_plan(text, script, scope)
    states, offs = parse_body(body, code_map, flag_bits, rio, max_clock_edges)
    return states, offs, header


def input_plan(text, script, scope=None):
    """Header, value-change text, and the map from VCD codes to the state
    bits of the script's inputs: (header, body, code_map, flag_bits, rio)."""
    netlist = script.netlist
    ins, _ = netlist.ports()
    in_names = [n for (n, _, _) in ins]
    header, body = parse_header(text)
    sc = find_scope(header, scope) if scope else auto_scope(header, in_names)
    vcd2inp = match_vcd_inputs(sc, in_names)
    name2port = {n: (pin, flag) for (n, pin, flag) in ins}
    rio = script.reg_io_state_size
    # (code, pos) -> (state bit, flag bit or None)
    code_map = {}
    for (code, pos), name in vcd2inp.items():
        pin, flag = name2port[name]
        bit = script.input_map[pin]
        fbit = script.input_map[flag] if flag else None
        code_map.setdefault(code, {})[pos] = (bit, fbit)
    flag_bits = [script.input_map[f] for (_, _, f) in ins if f]
    return header, body, code_map, flag_bits, rio


def parse_body(body, code_map, flag_bits, rio, max_clock_edges=None):
    """Value-change section -> (snapshots uint32[(C+1)*rio], offsets
    [(word offset, previous active time)]), by the native streaming reader
    (csrc/host/vcd.cpp; src/sim/vcd_io.rs:347-455).  code_map: {code: {bit
    position in the value: (state bit, posedge flag bit or None)}}."""
    import ctypes
    L = N.host()
    codes = list(code_map)
    enc = [c.encode() for c in codes]
    code_off = np.zeros(len(codes) + 1, np.uint32)
    code_off[1:] = np.cumsum([len(e) for e in enc])
    blob = b"".join(enc) or b"\0"
    ent_start = np.zeros(len(codes) + 1, np.uint32)
    pos, bit, flag = [], [], []
    for i, c in enumerate(codes):
        for p_, (b_, f_) in sorted(code_map[c].items()):
            pos.append(p_)
            bit.append(b_)
            flag.append(0xFFFFFFFF if f_ is None else f_)
        ent_start[i + 1] = len(pos)
    arr = [np.ascontiguousarray(x, np.uint32) for x in (pos, bit, flag, flag_bits)]
    data = body.encode() if isinstance(body, str) else bytes(body)
    h = L.jqh_vcd_parse_body(data, len(data), blob, code_off.ctypes.data, len(codes),
                             ent_start.ctypes.data, arr[0].ctypes.data, arr[1].ctypes.data,
                             arr[2].ctypes.data, arr[3].ctypes.data, len(flag_bits), rio,
                             int(max_clock_edges or 0))
    if not h:
        raise ValueError(L.jqh_vcd_last_error().decode())
    try:
        nw = int(L.jqh_vcd_num_words(h))
        no = int(L.jqh_vcd_num_offsets(h))
        states = np.ctypeslib.as_array(
            ctypes.cast(L.jqh_vcd_states(h), ctypes.POINTER(ctypes.c_uint32)), (nw,)).copy()
        if no:
            off = np.ctypeslib.as_array(
                ctypes.cast(L.jqh_vcd_offsets(h), ctypes.POINTER(ctypes.c_uint64)), (no,))
            ts = np.ctypeslib.as_array(
                ctypes.cast(L.jqh_vcd_timestamps(h), ctypes.POINTER(ctypes.c_uint64)), (no,))
            offs = list(zip(off.tolist(), ts.tolist()))
        else:
            offs = []
    finally:
        L.jqh_vcd_free(h)
    return states, offs


# ---------------------------------------------------------------- writers
def id_code(i):
    """rust-vcd IdCode display: base-94 over '!'..'~'."""
    s = []
    while True:
        r = i % 94
        s.append(chr(33 + r))
        if i < 94:
            break
        i = i // 94 - 1
    return "".join(s)


def output_mapping(script):
    """[(literal, state bit or None, name)] in output port order, skipping
    outputs that pass an input port through (src/sim/vcd_io.rs:580-613)."""
    A = script.netlist.arrays()
    _, outs = script.netlist.ports()
    res = []
    for name, lit in outs:
        if lit > 1 and A["drv"][lit >> 1] == 2:  # input port pass-through
            continue
        if lit <= 1:
            res.append((lit, None, name))
        else:
            res.append((lit, script.output_map[lit], name))
    return res


def _header_text(timescale, scope, mapping):
    lines = []
    if timescale:
        lines.append(f"$timescale {timescale[0]} {timescale[1]} $end")
    parts = scope.split("/")
    for p in parts:
        lines.append(f"$scope module {p} $end")
    for i, (_, _, name) in enumerate(mapping):
        lines.append(f"$var wire 1 {id_code(i)} {name} $end")
    for _ in parts:
        lines.append("$upscope $end")
    lines.append("$enddefinitions $end")
    lines.append("$dumpvars")
    return lines


def _codes(n):
    """Identifier codes of columns 0..n-1, concatenated, and their offsets."""
    codes = [id_code(i) for i in range(n)]
    off = np.zeros(n + 1, np.uint64)
    if n:
        off[1:] = np.cumsum([len(c) for c in codes])
    return "".join(codes).encode(), off


def _native_text(h):
    L = N.host()
    if not h:
        raise RuntimeError("vcd writer: " + L.jqh_vcd_last_error().decode())
    try:
        return ctypes.string_at(L.jqh_text_data(h), L.jqh_text_size(h)).decode()
    finally:
        L.jqh_text_free(h)


def _text_from_values(header, vals, timestamps):
    """Header lines, then the value-change section written natively from
    vals[row, column] in {0, 1, 3}: '#t' per row, a column when it changes."""
    vals = np.ascontiguousarray(vals, np.uint8)
    ts = np.ascontiguousarray(np.asarray(timestamps, dtype=np.uint64).reshape(-1))
    if vals.ndim != 2 or vals.shape[0] != ts.size:
        raise ValueError("one row of values per timestamp")
    codes, off = _codes(vals.shape[1])
    h = N.host().jqh_vcd_write_values(vals.ctypes.data, vals.shape[0], vals.shape[1],
                                      ts.ctypes.data, codes, off.ctypes.data)
    return "\n".join(header) + "\n" + _native_text(h)


def write_output_vcd(script, offsets_timestamps, states, timescale=(1, "ps"),
                     scope="gem_top_module", xmask_states=None):
    """Text of the output VCD (write_output_vcd[_xprop],
    src/sim/vcd_io.rs:729-820).  states / xmask_states: value-only buffers
    laid out like parse_input_vcd's input_states.  Output values are
    gathered per column over all emitted snapshots; the value-change
    section is written by the host library."""
    mapping = output_mapping(script)
    rows = [(int(o), int(t)) for o, t in offsets_timestamps if t != U64_MAX]
    vals = _output_values(mapping, [o for o, _ in rows], states, xmask_states)
    return _text_from_values(_header_text(timescale, scope, mapping), vals,
                             [t for _, t in rows])


def _output_values(mapping, offsets, states, xmask_states):
    """vals[row, column] in {0, 1, 3 (X)}: each output's value in the
    snapshot at offsets[row] (value-only layout), gathered column by
    column."""
    offs = np.asarray(offsets, np.int64)
    st = np.asarray(states, np.uint32)
    xm = None if xmask_states is None else np.asarray(xmask_states, np.uint32)
    vals = np.zeros((offs.size, len(mapping)), np.uint8)
    for i, (lit, pos, _) in enumerate(mapping):
        if pos is None:
            vals[:, i] = lit
            continue
        idx = offs + (pos >> 5)
        sh = np.uint32(pos & 31)
        col = ((st[idx] >> sh) & np.uint32(1)).astype(np.uint8)
        if xm is not None:
            col[((xm[idx] >> sh) & np.uint32(1)) != 0] = 3
        vals[:, i] = col
    return vals


def write_lane_vcd(script, obs, obs_lits, timestamps, timescale=(1, "ps"),
                   scope="gem_top_module"):
    """Output VCD text of one batched lane: obs is the lane's
    uint8[cycles, len(obs_lits)] (BatchedRun.lane_outputs), row c the
    outputs after cycle c, emitted at timestamps[c].  The text is the one
    write_output_vcd writes for the same lane run through the plain path
    with snapshot c + 1 emitted at timestamps[c]; every output port that is
    not a constant or an input pass-through must be in obs_lits."""
    mapping = output_mapping(script)
    col = {lit: k for k, lit in enumerate(obs_lits)}
    n = len(timestamps)
    obs = np.asarray(obs)
    if obs.shape[0] < n:
        raise ValueError("fewer rows of outputs than timestamps")
    vals = np.zeros((n, len(mapping)), np.uint8)
    for i, (lit, pos, _) in enumerate(mapping):
        vals[:, i] = lit if pos is None else obs[:n, col[lit]]
    return _text_from_values(_header_text(timescale, scope, mapping), vals, timestamps)


CONST0_BIT = 0xFFFFFFFF
CONST1_BIT = 0xFFFFFFFE


def output_changes_device(script, d_states, offsets_timestamps, stream=None):
    """Per emitted snapshot, the outputs whose value changed, computed on the
    device by jq_output_changes over the device snapshot buffer (laid out
    [values | X | arrivals] per snapshot, effective_state_size words each).
    offsets_timestamps are parse_input_vcd's (value-only offsets); skipped
    timestamps are left out, as the writer skips them.  Returns (counts
    uint32[n], records uint32[sum]): record = output << 2 | value (3 = X)."""
    import ctypes
    import torch
    from . import _native as N
    from .sim import to_device
    mapping = output_mapping(script)
    rio = script.reg_io_state_size
    ess = script.effective_state_size()
    snaps = [off // rio for off, ts in offsets_timestamps if ts != U64_MAX]
    bits = np.asarray([CONST1_BIT if (pos is None and lit == 1) else
                       CONST0_BIT if pos is None else pos
                       for lit, pos, _ in mapping], np.uint32)
    n = len(snaps)
    if n == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    d_off = to_device(np.asarray(snaps, np.uint64) * np.uint64(ess))
    d_bits = to_device(bits if bits.size else np.zeros(1, np.uint32))
    d_cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    xoff = rio if script.xprop_enabled else 0
    st = stream if stream is not None else torch.cuda.current_stream()
    L = N.hip()
    N.check(L.jq_output_changes(d_states.data_ptr(), d_off.data_ptr(), n, d_bits.data_ptr(),
                                len(bits), xoff, d_cnt.data_ptr(), None, None,
                                ctypes.c_void_p(st.cuda_stream)))
    cnt64 = d_cnt.to(torch.int64)
    rec_off = torch.cumsum(cnt64, 0) - cnt64
    total = int(cnt64.sum().item())
    d_rec = torch.zeros(max(total, 1), dtype=torch.int32, device="cuda")
    N.check(L.jq_output_changes(d_states.data_ptr(), d_off.data_ptr(), n, d_bits.data_ptr(),
                                len(bits), xoff, None, d_rec.data_ptr(),
                                rec_off.data_ptr(), ctypes.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    return (d_cnt.cpu().numpy().view(np.uint32),
            d_rec[:total].cpu().numpy().view(np.uint32))


def write_output_vcd_changes(script, offsets_timestamps, counts, records,
                             timescale=(1, "ps"), scope="gem_top_module"):
    """Output VCD text from per-snapshot change records
    (output_changes_device); the same text write_output_vcd produces from
    the full snapshots."""
    mapping = output_mapping(script)
    ts = np.asarray([int(t) for _, t in offsets_timestamps if t != U64_MAX], np.uint64)
    counts = np.ascontiguousarray(counts, np.uint32)
    records = np.ascontiguousarray(records, np.uint32)
    if counts.size < ts.size:
        raise ValueError("one change count per emitted snapshot")
    codes, off = _codes(len(mapping))
    h = N.host().jqh_vcd_write_changes(counts.ctypes.data, records.ctypes.data, records.size,
                                       ts.size, ts.ctypes.data, len(mapping), codes,
                                       off.ctypes.data)
    return "\n".join(_header_text(timescale, scope, mapping)) + "\n" + _native_text(h)


def write_output_vcd_timed(script, offsets_timestamps, states, arrival_states,
                           timescale=(1, "ps"), scope="gem_top_module",
                           xmask_states=None):
    """Timed VCD (src/sim/vcd_io.rs:831-946): every transition is shifted by
    its word's arrival (ps, rounded half away from zero to VCD units) and
    transitions are stably sorted by time.  Values and arrivals are gathered
    here; the host library writes the value-change section."""
    mapping = output_mapping(script)
    if timescale:
        ps_per_unit = timescale[0] * (_PS_DIVISOR["ps"] / _PS_DIVISOR[timescale[1]])
    else:
        ps_per_unit = 1.0
    rows = [(int(o), int(t)) for o, t in offsets_timestamps if t != U64_MAX]
    offs = np.asarray([o for o, _ in rows], np.int64)
    vals = _output_values(mapping, offs, states, xmask_states)
    arr = np.zeros(vals.shape, np.uint16)
    ast = np.asarray(arrival_states, np.uint32)
    for i, (_, pos, _) in enumerate(mapping):
        if pos is not None:
            arr[:, i] = (ast[offs + (pos >> 5)] & np.uint32(0xFFFF)).astype(np.uint16)
    ts = np.asarray([t for _, t in rows], np.uint64)
    codes, off = _codes(len(mapping))
    h = N.host().jqh_vcd_write_timed(vals.ctypes.data, arr.ctypes.data, vals.shape[0],
                                     vals.shape[1], ts.ctypes.data, float(ps_per_unit),
                                     codes, off.ctypes.data)
    return "\n".join(_header_text(timescale, scope, mapping)) + "\n" + _native_text(h)


# ---------------------------------------------------------------- buffers
def expand_states_for_xprop(value_states, script, initial_xmask=None):
    """[values | X mask] per snapshot, template X (src/sim/vcd_io.rs:465-484),
    optionally an explicit X mask for snapshot 0."""
    rio = script.reg_io_state_size
    n = value_states.size // rio
    out = np.zeros((n, 2 * rio), np.uint32)
    out[:, :rio] = value_states.reshape(n, rio)
    out[:, rio:] = script.xmask_template()[None, :]
    if initial_xmask is not None:
        out[0, rio:] = initial_xmask
    return out.reshape(-1)


def expand_states_for_arrivals(states, script):
    """Append a zeroed arrival section per snapshot
    (src/sim/vcd_io.rs:511-530)."""
    old = script.reg_io_state_size * (2 if script.xprop_enabled else 1)
    rio = script.reg_io_state_size
    n = states.size // old
    out = np.zeros((n, old + rio), np.uint32)
    out[:, :old] = states.reshape(n, old)
    return out.reshape(-1)


def split_sections(gpu_states, script):
    """(values, xmask or None, arrivals or None), each laid out like the
    value-only buffer (src/sim/vcd_io.rs:491-550)."""
    ss = script.effective_state_size()
    rio = script.reg_io_state_size
    S = gpu_states.reshape(-1, ss)
    vals = S[:, :rio].reshape(-1).copy()
    xm = S[:, rio:2 * rio].reshape(-1).copy() if script.xprop_enabled else None
    arr = None
    if script.timing_arrivals_enabled:
        a0 = script.arrival_state_offset
        arr = S[:, a0:a0 + rio].reshape(-1).copy()
    return vals, xm, arr
