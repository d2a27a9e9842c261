"""The two measurement launches bench.py quotes beside the HBM roofline: the
LDS read probe (its checksum proves it reads every byte it is credited
with) and the scan kernel over an all-empty script (the grid-barrier
figure), which must leave every snapshot untouched."""
import ctypes

import numpy as np
import pytest

from jacquard_amd import _native as N
from jacquard_amd import sim

pytestmark = pytest.mark.gpu


def _probe_expected(blocks, iters):
    """Per-thread checksum of the probe, restated with numpy: thread t of
    block b reads vector t + 256 * (k mod 16) for k < 8 * iters, the vector
    at index i holding (i, 3i + b, 5i, 7i)."""
    t = np.arange(256, dtype=np.uint64)
    out = np.zeros((blocks, 256), np.uint32)
    for b in range(blocks):
        ax = np.zeros(256, np.uint64)
        ay = np.zeros(256, np.uint64)
        az = np.zeros(256, np.uint64)
        aw = np.zeros(256, np.uint64)
        for k in range(8 * iters):
            i = t + 256 * (k % 16)
            ax ^= i
            ay = (ay + 3 * i + b) & 0xFFFFFFFF
            az ^= 5 * i
            aw = (aw + 7 * i) & 0xFFFFFFFF
        out[b] = ((ax ^ ay ^ az ^ aw) & 0xFFFFFFFF).astype(np.uint32)
    return out


@pytest.mark.parametrize("blocks,iters", [(1, 1), (3, 5)])
def test_lds_probe_reads_what_it_counts(blocks, iters):
    import torch
    sink = torch.zeros(blocks * 256, dtype=torch.int32, device="cuda")
    N.check(N.hip().jq_lds_read_probe(blocks, iters, sink.data_ptr(), None))
    torch.cuda.synchronize()
    got = sink.cpu().numpy().view(np.uint32).reshape(blocks, 256)
    assert np.array_equal(got, _probe_expected(blocks, iters))


def test_lds_probe_rejects_bad_arguments():
    assert N.hip().jq_lds_read_probe(0, 1, None, None) != 0


def test_empty_script_leaves_states_and_barrier_clean():
    import torch
    B, M, cycles, ss = 37, 5, 200, 3
    bs = np.arange(B * M + 1, dtype=np.uint64) * 256
    d_bs = sim.to_device(bs)
    d_bd = torch.zeros(B * M * 256 + 2048, dtype=torch.int32, device="cuda")
    init = np.random.default_rng(5).integers(0, 2**32, (cycles + 1) * ss,
                                             dtype=np.uint64).astype(np.uint32)
    st = torch.from_numpy(init.view(np.int32)).to("cuda")
    one = torch.zeros(1, dtype=torch.int32, device="cuda")
    N.check(N.hip().jq_simulate_scan(B, M, d_bs.data_ptr(), d_bd.data_ptr(),
                                     one.data_ptr(), one.data_ptr(), cycles, ss,
                                     st.data_ptr(), None, None, 0, 0,
                                     ctypes.c_void_p(0)))
    torch.cuda.synchronize()
    assert sim.device_error() == 0
    assert np.array_equal(st.cpu().numpy().view(np.uint32), init)
