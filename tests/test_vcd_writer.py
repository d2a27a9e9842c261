"""The host library's output-VCD body writers against the restatement in
oracle/vcd_oracle.py, on random data: values with X, more columns than one
identifier character covers (94), empty tables and 64-bit timestamps.  The
whole output (header and body) is pinned byte for byte by the reference's
golden output VCDs in test_golden_vcd.py."""
import numpy as np
import pytest

import vcd_oracle
from jacquard_amd import vcd


def _body(text):
    return text[len("\n".join(["h"])) + 1:]


@pytest.mark.parametrize("rows,cols,seed", [(0, 5, 1), (7, 0, 2), (1, 1, 3), (300, 40, 4),
                                            (200, 300, 5), (50, 9000, 6)])
def test_value_rows_match_the_restatement(rows, cols, seed):
    rng = np.random.default_rng(seed)
    # sticky columns (few changes) and busy ones, X now and then
    vals = rng.choice(np.array([0, 1, 3], np.uint8), size=(rows, cols), p=[0.45, 0.45, 0.1])
    if rows > 1:
        keep = rng.random((rows, cols)) < 0.7
        for r in range(1, rows):
            vals[r][keep[r]] = vals[r - 1][keep[r]]
    ts = np.sort(rng.integers(0, 2**63, rows, dtype=np.uint64))
    if rows:
        ts[-1] = np.uint64(2**64 - 2)
    got = vcd._text_from_values(["h"], vals, ts)
    assert got.startswith("h\n")
    assert _body(got) == vcd_oracle.body_from_values(vals.tolist(), ts.tolist())


@pytest.mark.parametrize("rows,cols,seed", [(0, 3, 1), (5, 1, 2), (400, 200, 3)])
def test_change_records_match_the_restatement(rows, cols, seed):
    from jacquard_amd import _native as N
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 6, rows).astype(np.uint32)
    n = int(counts.sum())
    records = ((rng.integers(0, cols, n).astype(np.uint32) << 2)
               | rng.choice(np.array([0, 1, 3], np.uint32), n))
    ts = np.arange(rows, dtype=np.uint64) * np.uint64(1000) + np.uint64(17)
    codes, off = vcd._codes(cols)
    h = N.host().jqh_vcd_write_changes(counts.ctypes.data, records.ctypes.data, n, rows,
                                       ts.ctypes.data, cols, codes, off.ctypes.data)
    assert vcd._native_text(h) == vcd_oracle.body_from_changes(counts, records, ts.tolist())


def test_identifier_codes_match_the_restatement():
    for i in list(range(0, 200)) + [8835, 8836, 8929, 8930, 10**6]:
        assert vcd.id_code(i) == vcd_oracle.id_code(i)


def test_writers_reject_malformed_input():
    from jacquard_amd import _native as N
    with pytest.raises(RuntimeError, match="0, 1 or 3"):
        vcd._text_from_values(["h"], np.array([[0, 2]], np.uint8), [5])
    with pytest.raises(ValueError):
        vcd._text_from_values(["h"], np.zeros((2, 3), np.uint8), [1])
    codes, off = vcd._codes(4)
    counts = np.array([2, 3], np.uint32)
    records = np.array([0, 4, 8, 12], np.uint32)  # one record short
    ts = np.array([1, 2], np.uint64)
    h = N.host().jqh_vcd_write_changes(counts.ctypes.data, records.ctypes.data, 4, 2,
                                       ts.ctypes.data, 4, codes, off.ctypes.data)
    with pytest.raises(RuntimeError, match="exceed"):
        vcd._native_text(h)
    bad = np.array([(4 << 2) | 1], np.uint32)  # column 4 of 4
    h = N.host().jqh_vcd_write_changes(np.array([1], np.uint32).ctypes.data, bad.ctypes.data,
                                       1, 1, ts.ctypes.data, 4, codes, off.ctypes.data)
    with pytest.raises(RuntimeError, match="bad change record"):
        vcd._native_text(h)


def test_lane_writer_matches_the_snapshot_writer():
    """The batched form's writer (one row of observed outputs per cycle)
    gives the text the plain path's writer gives from full snapshots, with
    outputs read from an oracle run of the same stimulus (the GPU tests
    compare the same two writers on device results)."""
    import oracle
    from helpers import CONFIG1, design
    from jacquard_amd import stimulus
    _, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    c = 120
    st = stimulus.build_states(s, c, seed=5, p=0.4)
    oracle.run(s, st, c)
    rio = s.reg_io_state_size
    lits = [lit for lit, pos, _ in vcd.output_mapping(s) if pos is not None]
    assert lits, "the design has outputs computed by the script"
    snaps = st.reshape(c + 1, rio)[1:]
    obs = np.stack([(snaps[:, s.output_map[l] >> 5] >> np.uint32(s.output_map[l] & 31)) & 1
                    for l in lits], axis=1).astype(np.uint8)
    ts = [10 * (k + 1) for k in range(c)]
    want = vcd.write_output_vcd(s, [((k + 1) * rio, ts[k]) for k in range(c)], st)
    assert vcd.write_lane_vcd(s, obs, lits, ts) == want


@pytest.mark.parametrize("ps_per_unit", [1.0, 1000.0, 0.001, 333.0, 0.0])
def test_timed_body_matches_the_restatement(ps_per_unit):
    """Arrival-shifted changes, stably sorted by time (ties keep row and
    column order), with halves rounded away from zero and times past
    2^64 - 1 near the end of the range."""
    from jacquard_amd import _native as N
    rng = np.random.default_rng(int(ps_per_unit * 7) + 1)
    rows, cols = 150, 120
    vals = rng.choice(np.array([0, 1, 3], np.uint8), size=(rows, cols), p=[0.45, 0.45, 0.1])
    keep = rng.random((rows, cols)) < 0.6
    for r in range(1, rows):
        vals[r][keep[r]] = vals[r - 1][keep[r]]
    arr = rng.integers(0, 65536, (rows, cols)).astype(np.uint16)
    arr[:, ::7] = 500  # exact halves at 1000 ps per unit
    ts = np.arange(rows, dtype=np.uint64) * np.uint64(40)
    ts[-20:] = np.uint64(2**64 - 1) - np.arange(20, 0, -1, dtype=np.uint64)
    codes, off = vcd._codes(cols)
    h = N.host().jqh_vcd_write_timed(vals.ctypes.data, arr.ctypes.data, rows, cols,
                                     ts.ctypes.data, ps_per_unit, codes, off.ctypes.data)
    got = vcd._native_text(h)
    want = vcd_oracle.body_timed(vals.tolist(), arr.tolist(), ts.tolist(), ps_per_unit)
    assert got == want
    times = [int(line[1:]) for line in got.splitlines() if line.startswith("#")]
    assert times == sorted(times)
    if ps_per_unit > 0:
        assert max(times) > 2**64 - 1, "some shift passes the end of the u64 range"
