"""Cost-model MLP: the CPU oracle's analytic gradient is pinned by finite
differences (CPU); the hand-written HIP forward/backward/SGD kernels are
checked against the oracle (GPU, float32 tolerance rtol 1e-4 / atol 1e-6)."""
import numpy as np
import pytest

import mlp_oracle as mo

I, H1, H2 = 24, 64, 64


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, I))
    t = np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2] - 0.1 * x[:, 3]
    return x, t


def test_oracle_gradient_matches_finite_differences():
    x, t = _data(40, 1)
    p = mo.init(I, H1, H2, seed=2, dtype=np.float64)
    g = mo.grad(p, x, t, I, H1, H2)
    rng = np.random.default_rng(3)
    idx = rng.choice(p.size, 60, replace=False)
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mo.loss(pp, x, t, I, H1, H2) - mo.loss(pm, x, t, I, H1, H2)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_oracle_closed_form_linear_case():
    """With ReLUs active and one hidden unit path, dL/db3 = 2 mean(y - t)."""
    x, t = _data(16, 4)
    p = mo.init(I, H1, H2, seed=5, dtype=np.float64)
    y, _ = mo.forward(p, x, I, H1, H2)
    g = mo.grad(p, x, t, I, H1, H2)
    assert np.isclose(g[-1], 2 * np.mean(y - t))


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    import torch
    from jacquard_amd.costmodel import MLP
    x, t = _data(1000, 6)
    p = mo.init(I, H1, H2, seed=7)
    m = MLP(params=p.copy())
    xd = torch.from_numpy(x.astype(np.float32)).cuda()
    td = torch.from_numpy(t.astype(np.float32)).cuda()
    y = torch.zeros(1000, dtype=torch.float32, device="cuda")
    m.forward_raw(xd, 1000, y)
    y_ref, _ = mo.forward(p.astype(np.float64), x.astype(np.float32).astype(np.float64), I, H1, H2)
    assert np.allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-5)
    from jacquard_amd import _native as N
    ws = torch.zeros(int(N.hip().jq_mlp_workspace_floats(I, H1, H2, 1000)),
                     dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    m.train_step_raw(xd, td, 1000, 0.0, ws, loss)
    g_ref = mo.grad(p.astype(np.float64), x.astype(np.float32).astype(np.float64),
                    t.astype(np.float32).astype(np.float64), I, H1, H2)
    g = m.grads.cpu().numpy()
    assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-6)
    assert np.isclose(loss.item(), mo.loss(p.astype(np.float64), x, t, I, H1, H2), rtol=1e-4)
    # an SGD step moves the parameters by exactly -lr * grad
    before = m.params.cpu().numpy().copy()
    m.train_step_raw(xd, td, 1000, 0.5, ws, loss)
    assert np.allclose(m.params.cpu().numpy(), before - 0.5 * m.grads.cpu().numpy(),
                       rtol=0, atol=1e-6)


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    from jacquard_amd.costmodel import MLP
    x, t = _data(5000, 8)
    m = MLP(seed=1)
    hist = m.fit(x, np.exp(t), epochs=60, lr=0.05, batch=512)
    assert hist[-1] < 0.5 * hist[0]


@pytest.mark.gpu
def test_learned_costs_against_heuristic():
    """The model trained on the batched kernel's own partition timings
    predicts held-out partition latency better than the reference heuristic
    (boomerang stages + 3, best linear fit), and the block assignment built
    from its predictions does not run slower than the heuristic's."""
    from helpers import design
    from jacquard_amd import costmodel as CM
    nl, comp = design("cm", level_split=4, seed=2, gates=300000, depth=48, dffs=20000,
                      inputs=256, outputs=128, geometric_p=0.3, invert_prob=0.5)
    W, G = 12, 2
    X, y, med = CM.training_set(comp, [W, 10, 8], G, 16)
    assert len(y) >= 5000
    s0 = comp.flatten(W)
    f0 = s0.part_features()
    truth = np.array([med[(m, i)] for (m, _b, i, _o, _l) in s0.parts.tolist()])
    h = f0[:, 0] + 3
    a, b = np.polyfit(h, truth, 1)
    mape_h = float(np.mean(np.abs(a * h + b - truth) / truth))
    model = CM.MLP(seed=1)
    model.fit(X, y, epochs=150, lr=0.05, batch=4096)
    mape_l = float(np.mean(np.abs(model.predict(f0) - truth) / truth))
    r = CM.compare_assignments(comp, W, G, {
        "reference": None, "learned": CM.stage_costs(comp, s0, model.predict(f0))},
        repeats=3, cycles=32)
    print(f"MAPE heuristic {mape_h:.3f} learned {mape_l:.3f}; us/cycle {r}", flush=True)
    assert mape_l < mape_h
    assert r["learned"] <= 1.03 * r["reference"]


@pytest.mark.gpu
def test_learned_merge_and_assignment_beat_the_heuristic():
    """On a script whose blocks run several partitions per major stage (12
    blocks for ~40 partitions per stage, as the 16M-gate workload runs about
    4.6), the learned model's merge decision plus its block assignment run
    faster than the heuristic assignment by more than the run-to-run spread
    of either, over 5 interleaved repeats; the simulated values of the
    merged design stay exact (checked against the oracle)."""
    import oracle
    from helpers import design
    from jacquard_amd import batched as B
    from jacquard_amd import costmodel as CM
    from jacquard_amd import stimulus
    nl, comp = design("cm", level_split=4, seed=2, gates=300000, depth=48, dffs=20000,
                      inputs=256, outputs=128, geometric_p=0.3, invert_prob=0.5)
    W, G = 12, 2
    costs, model, info = CM.learn_costs(comp, W, G)
    merged, mcosts, minfo = CM.learned_merge(comp, model, W, groups=G)
    assert merged is not comp and minfo["pairs_kept"] > 0
    spread = {}
    r = CM.compare_assignments(comp, W, G, {"heuristic": None, "learned+merge": (merged, mcosts)},
                               repeats=5, cycles=32, spread=spread)
    gain = r["heuristic"] - r["learned+merge"]
    print(f"us/cycle {r}, spread {spread}, merge {minfo}", flush=True)
    assert gain > max(spread["heuristic"], spread["learned+merge"]), (r, spread)
    # exact values: lanes of the merged, learned-assignment script against
    # the oracle on the original script
    s = merged.flatten(W, costs=mcosts)
    s0 = comp.flatten(W)
    seeds, probs, cyc = np.array([5, 9], np.uint32), np.array([0.2, 0.4]), np.array([24, 17])
    run = B.BatchedRun(s, seeds, probs, cyc, groups=1)
    out = run.run()
    for lane in range(2):
        st = stimulus.build_states(s0, int(cyc[lane]), seed=int(seeds[lane]), p=float(probs[lane]))
        oracle.run(s0, st, int(cyc[lane]))
        want = B.plain_observations(s0, st, int(cyc[lane]), run.obs_lits)
        assert np.array_equal(out[lane], want)
