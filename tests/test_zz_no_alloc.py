"""Allocation-free MLP train step (jq_mlp_train_step_noalloc) against the
workspace-argument entry point (jq_mlp_train_step), through the C ABI.

Both run the same deterministic kernels, so parameters, gradients and the
batch loss must agree bit for bit.  Two calls in a row, the second with a
smaller batch, check that the shared static workspace carries nothing from
one call into the next; the size limit is checked on the host side."""
import ctypes

import numpy as np
import pytest

I, H1, H2 = 24, 64, 64


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, I)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2]).astype(np.float32)
    return x, t


@pytest.mark.gpu
def test_noalloc_train_step_matches_workspace_entry_point():
    import torch
    from jacquard_amd import _native as N
    L = N.hip()
    np_ = int(L.jq_mlp_num_params(I, H1, H2))
    rng = np.random.default_rng(11)
    p0 = (rng.normal(size=np_) * 0.1).astype(np.float32)
    pa = torch.from_numpy(p0.copy()).cuda()
    pb = torch.from_numpy(p0.copy()).cuda()
    ga = torch.zeros(np_, dtype=torch.float32, device="cuda")
    gb = torch.full((np_,), 7.0, dtype=torch.float32, device="cuda")
    la = torch.zeros(1, dtype=torch.float32, device="cuda")
    lb = torch.zeros(1, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert int(L.jq_mlp_noalloc_max_samples()) == 4096
    for call, n in enumerate((4096, 1000)):
        print("no-alloc train step: call", call, "batch", n, flush=True)
        x, t = _batch(n, 20 + call)
        xd = torch.from_numpy(x).cuda()
        td = torch.from_numpy(t).cuda()
        ws = torch.zeros(int(L.jq_mlp_workspace_floats(I, H1, H2, n)),
                         dtype=torch.float32, device="cuda")
        N.check(L.jq_mlp_train_step(pa.data_ptr(), ga.data_ptr(), ws.data_ptr(), I, H1,
                                    H2, xd.data_ptr(), td.data_ptr(), n,
                                    ctypes.c_float(0.05), la.data_ptr(), stream))
        N.check(L.jq_mlp_train_step_noalloc(pb.data_ptr(), gb.data_ptr(), I, H1, H2,
                                            xd.data_ptr(), td.data_ptr(), n,
                                            ctypes.c_float(0.05), lb.data_ptr(), stream))
        torch.cuda.synchronize()
        assert np.array_equal(ga.cpu().numpy().view(np.uint32),
                              gb.cpu().numpy().view(np.uint32)), f"gradients, call {call}"
        assert np.array_equal(pa.cpu().numpy().view(np.uint32),
                              pb.cpu().numpy().view(np.uint32)), f"parameters, call {call}"
        assert la.item() == lb.item() and np.isfinite(la.item())
    # a batch beyond the static workspace is refused before any launch
    rc = L.jq_mlp_train_step_noalloc(pb.data_ptr(), gb.data_ptr(), I, H1, H2,
                                     xd.data_ptr(), td.data_ptr(), 4097,
                                     ctypes.c_float(0.05), lb.data_ptr(), stream)
    assert rc != 0 and b"4096" in L.jq_last_error()
