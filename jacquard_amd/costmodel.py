"""Learned partition cost model.

A small FP32 MLP (24 script features -> 64 -> 64 -> 1) predicts the GPU
latency of one partition execution from its script section
(``FlattenedScriptV1.part_features``).  Training data comes from the
simulator itself: the profiled launches (jq_simulate_scan_profiled,
jq_simulate_batched_profiled) record the clock ticks of every partition
execution.  Training and inference run on the GPU through the hand-written
kernels in libjacquard_hip.so (jq_mlp_*).  The predictions replace the
reference's heuristic cost (boomerang stages + 3, src/flatten.rs:1207-1226)
in the block assignment (longest-predicted-first onto the least-loaded
block); the heuristic stays the fallback, and measured kernel time is the
judge (``compare_assignments``).  Features depend only on a partition's own
script section, not on which block it lands in, so costs predicted from one
flattening apply to any other.
"""
import ctypes

import numpy as np

from . import _native as N

IN, H1, H2 = 24, 64, 64


def _torch():
    import torch
    return torch


class MLP:
    """Device-resident MLP with hand-written forward/backward kernels."""

    def __init__(self, params=None, seed=0, i=IN, h1=H1, h2=H2):
        torch = _torch()
        self.i, self.h1, self.h2 = i, h1, h2
        L = N.hip()
        self.np = int(L.jq_mlp_num_params(i, h1, h2))
        if params is None:
            params = he_init(i, h1, h2, seed)
        assert params.size == self.np
        self.params = torch.from_numpy(np.ascontiguousarray(params, np.float32)).cuda()
        self.grads = torch.zeros(self.np, dtype=torch.float32, device="cuda")
        self.mu = np.zeros(i, np.float32)
        self.sd = np.ones(i, np.float32)
        self.y_mu, self.y_sd = 0.0, 1.0

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def forward_raw(self, x_dev, n, y_dev):
        N.check(N.hip().jq_mlp_forward(self.params.data_ptr(), self.i, self.h1, self.h2,
                                       x_dev.data_ptr(), n, y_dev.data_ptr(), self._stream()))

    def train_step_raw(self, x_dev, t_dev, n, lr, ws, loss_dev):
        N.check(N.hip().jq_mlp_train_step(
            self.params.data_ptr(), self.grads.data_ptr(), ws.data_ptr(), self.i,
            self.h1, self.h2, x_dev.data_ptr(), t_dev.data_ptr(), n, float(lr),
            loss_dev.data_ptr() if loss_dev is not None else None, self._stream()))

    # ------------------------------------------------------------------
    def fit(self, x, y, epochs=200, lr=0.05, batch=4096, seed=0, log=None):
        """Standardised features, standardised log-latency target, minibatch
        SGD.  Returns the per-epoch training loss."""
        torch = _torch()
        x = np.asarray(x, np.float32)
        y = np.log(np.maximum(np.asarray(y, np.float64), 1.0))
        self.mu = x.mean(0)
        self.sd = x.std(0) + 1e-6
        self.y_mu, self.y_sd = float(y.mean()), float(y.std() + 1e-9)
        xs = torch.from_numpy(((x - self.mu) / self.sd).astype(np.float32)).cuda()
        ys = torch.from_numpy(((y - self.y_mu) / self.y_sd).astype(np.float32)).cuda()
        n = x.shape[0]
        ws = torch.zeros(int(N.hip().jq_mlp_workspace_floats(self.i, self.h1, self.h2, batch)),
                         dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            xb_all, yb_all = xs[perm].contiguous(), ys[perm].contiguous()
            tot = 0.0
            for b0 in range(0, n, batch):
                nb = min(batch, n - b0)
                self.train_step_raw(xb_all[b0:b0 + nb], yb_all[b0:b0 + nb], nb, lr, ws, loss)
                tot += float(loss.item()) * nb
            hist.append(tot / n)
            if log and (ep % 50 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
        return hist

    def predict(self, x):
        """Predicted latency (same unit as the training target)."""
        torch = _torch()
        x = np.asarray(x, np.float32)
        xs = torch.from_numpy(((x - self.mu) / self.sd).astype(np.float32)).cuda()
        y = torch.zeros(x.shape[0], dtype=torch.float32, device="cuda")
        self.forward_raw(xs, x.shape[0], y)
        z = y.cpu().numpy().astype(np.float64) * self.y_sd + self.y_mu
        return np.exp(z)


def he_init(i, h1, h2, seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((i, h1), (h1, h2), (h2, 1)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.full(fan_out, 0.01 if fan_out > 1 else 0.0))
    return np.concatenate(parts).astype(np.float32)


# ----------------------------------------------------------------------
def profile_partitions(script, states_dev, num_cycles, sample_cycles):
    """Clock ticks of every partition execution of the first sample_cycles
    cycles: uint32[sample_cycles, num_parts] (blocks_data order)."""
    torch = _torch()
    from .sim import to_device
    L = N.hip()
    ss = script.effective_state_size()
    d_bs = to_device(np.ascontiguousarray(script.blocks_start, np.uint64))
    d_bd = to_device(np.ascontiguousarray(script.blocks_data, np.uint32))
    d_sram = torch.zeros(max(script.sram_storage_size, 1), dtype=torch.int32, device="cuda")
    pidx = to_device(script.part_index())
    samples = torch.zeros(sample_cycles * max(script.num_parts, 1), dtype=torch.int32,
                          device="cuda")
    N.check(L.jq_simulate_scan_profiled(
        script.num_blocks, script.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
        d_sram.data_ptr(), d_sram.data_ptr(), num_cycles, ss, states_dev.data_ptr(),
        pidx.data_ptr(), script.num_parts, samples.data_ptr(), sample_cycles,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return samples.cpu().numpy().view(np.uint32).reshape(sample_cycles, -1)


def stage_costs(compiled, script, per_part):
    """Per-partition costs [major][partition] from values given in the
    script's blocks_data order (script.parts rows)."""
    costs = [np.zeros(n) for n in compiled.parts_per_stage()]
    for k, (m, _b, idx, _o, _l) in enumerate(script.parts.tolist()):
        costs[m][idx] = per_part[k]
    return costs


# ----------------------------------------------------------------------
# batched form: samples, training set, assignment A/B

def profile_batched(compiled, blocks, groups, sample_cycles, costs=None, seed=0):
    """Flatten for `blocks` blocks per group, run `groups` full groups of
    lanes for sample_cycles + 2 cycles with per-partition clock sampling on
    group 0.  Returns (script, ticks float64[sample_cycles, num_parts]) in
    blocks_data partition order."""
    torch = _torch()
    from .batched import BatchedRun
    s = compiled.flatten_shared(blocks) if costs is None else compiled.flatten(blocks, costs=costs)
    n = 32 * groups
    rng = np.random.default_rng(seed)
    run = BatchedRun(s, rng.integers(1, 2**31, n), rng.uniform(0.02, 0.5, n),
                     np.full(n, sample_cycles + 2), obs_lits=[], groups=groups)
    run.prepare()
    torch.cuda.synchronize()
    samp = run.launch_profiled(sample_cycles)
    torch.cuda.synchronize()
    ticks = samp.cpu().numpy().view(np.uint32)[:sample_cycles * s.num_parts]
    return s, ticks.reshape(sample_cycles, s.num_parts).astype(np.float64)


def training_set(compiled, blocks_list, groups, sample_cycles, skip=2, seed=0):
    """(features float32[n, 24], ticks float64[n], per-partition median
    ticks by (major, index)) from profiled batched launches of several
    flattenings; every (cycle, partition) execution after the first `skip`
    cycles is one sample."""
    xs, ys = [], []
    med = {}
    for k, W in enumerate(blocks_list):
        s, ticks = profile_batched(compiled, W, groups, sample_cycles + skip, seed=seed + k)
        f = s.part_features()
        t = ticks[skip:]
        xs.append(np.repeat(f[None], t.shape[0], 0).reshape(-1, f.shape[1]))
        ys.append(t.reshape(-1))
        m = np.median(t, 0)
        for (mj, _b, idx, _o, _l), v in zip(s.parts.tolist(), m.tolist()):
            med.setdefault((mj, idx), []).append(v)
    med = {k: float(np.median(v)) for k, v in med.items()}
    return np.concatenate(xs).astype(np.float32), np.concatenate(ys), med


def batched_us_per_cycle(script, groups, cycles=48, seed=1):
    """Kernel time per group cycle of the batched launch of `script` with
    `groups` full groups of lanes."""
    torch = _torch()
    from .batched import BatchedRun
    n = 32 * groups
    rng = np.random.default_rng(seed)
    run = BatchedRun(script, rng.integers(1, 2**31, n), rng.uniform(0.02, 0.5, n),
                     np.full(n, cycles), obs_lits=[], groups=groups)
    run.prepare()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    run.launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / cycles


def compare_assignments(compiled, blocks, groups, variants, repeats=3, cycles=48,
                        spread=None):
    """Measured us/cycle of the batched kernel for each named variant
    ({name: costs or None, or (compiled design, costs or None)}); the
    variants' launches are interleaved `repeats` times and the median is
    returned per name.  With a dict as `spread`, each name's max - min over
    the repeats is stored in it."""
    def script(v):
        c, costs = v if isinstance(v, tuple) else (compiled, v)
        return c.flatten_shared(blocks) if costs is None else c.flatten(blocks, costs=costs)
    scripts = {k: script(v) for k, v in variants.items()}
    out = {k: [] for k in variants}
    for r in range(repeats):
        for k in variants:
            out[k].append(batched_us_per_cycle(scripts[k], groups, cycles, seed=r + 1))
    if spread is not None:
        spread.update({k: float(np.max(v) - np.min(v)) for k, v in out.items()})
    return {k: float(np.median(v)) for k, v in out.items()}


def lpt_makespan(costs, blocks):
    """Largest block total when partitions of these costs go, largest first,
    to the least-loaded of `blocks` blocks (the flatten's assignment)."""
    import heapq
    heap = [0.0] * blocks
    for x in sorted((float(c) for c in costs), reverse=True):
        heapq.heapreplace(heap, heap[0] + x)
    return max(heap) if heap else 0.0


def merge_candidates(pred, blocks):
    """Pairs (major, i, j) proposed for merging: in every major stage with
    n > blocks partitions, the 2k of smallest predicted latency, k =
    min(n - blocks, n // 2), paired smallest with largest among them so that
    the merged partitions' predicted latencies stay even (with k = n -
    blocks the stage then fits one partition per block)."""
    pairs = []
    for m, c in enumerate(pred):
        k = min(len(c) - blocks, len(c) // 2)
        if k <= 0:
            continue
        small = np.argsort(np.asarray(c), kind="stable")[:2 * k]
        pairs += [(m, int(small[a]), int(small[2 * k - 1 - a])) for a in range(k)]
    return pairs


def split_candidates(pred, endpoints, blocks):
    """Partitions (major, i) proposed for a cut: in every major stage with
    n < blocks partitions, the blocks - n of largest predicted latency that
    have more than one endpoint (the stage's time is its slowest block, and
    the free blocks can take the halves)."""
    out = []
    for m, c in enumerate(pred):
        k = blocks - len(c)
        if k <= 0:
            continue
        order = np.argsort(-np.asarray(c), kind="stable")
        out += [(m, int(i)) for i in order if endpoints[m][i] > 1][:k]
    return out


def learned_merge(compiled, model, blocks, groups=None, sample_cycles=16, epochs=100,
                  log=None, cut=True):
    """The partition cut and merge step driven by the learned latency model
    (the role of the merge acceptance of src/pe.rs:1033-1046, and of the
    partition sizing that precedes it).  In each major stage with more
    partitions than blocks, the candidates of merge_candidates are mapped
    as merged partitions; with cut, in each stage with fewer, the slowest
    predicted partitions of split_candidates are mapped as two halves each.
    The model predicts the new partitions' latencies from their script
    sections, and a stage keeps its changes when its predicted makespan
    (largest block total after the longest-first assignment of predicted
    costs) drops.

    The new partitions' sizes lie outside those the model was trained on,
    so with `groups` given (and the model's training set kept by
    learn_costs) the trial design is profiled first (sample_cycles cycles
    of `groups` lane groups) and the model refitted on the old and new
    samples before it predicts.  Returns (compiled design with the kept
    changes, its learned costs, info); with nothing kept, the design is
    `compiled` itself."""
    s0 = compiled.flatten_shared(blocks)
    pred0 = stage_costs(compiled, s0, model.predict(s0.part_features()))
    before = [lpt_makespan(c, blocks) for c in pred0]
    info = dict(candidates=0, merged=0, cut_candidates=0, cut=0, stages_kept=0,
                refit_samples=0, predicted_ticks_before=float(sum(before)))
    pairs = merge_candidates(pred0, blocks)
    cuts = split_candidates(pred0, compiled.endpoint_counts(), blocks) if cut else []
    info["candidates"], info["cut_candidates"] = len(pairs), len(cuts)
    if not pairs and not cuts:
        info["predicted_ticks_after"] = info["predicted_ticks_before"]
        return compiled, pred0, info

    def build(p, c):
        # cuts first (their stages have no merge pairs, so the merge pairs'
        # partition indices stay valid: untouched stages keep their order)
        d, okc = compiled.split(c) if c else (compiled, np.zeros(0, bool))
        d, okm = d.merged(p) if p else (d, np.zeros(0, bool))
        return d, okm, okc

    trial, ok, okc = build(pairs, cuts)
    s1 = trial.flatten_shared(blocks)
    if groups and (ok.any() or okc.any()) and getattr(model, "train_x", None) is not None:
        _s, ticks = profile_batched(trial, blocks, groups, sample_cycles + 2, seed=7)
        t = ticks[2:]
        f1 = s1.part_features()
        X = np.concatenate([model.train_x, np.repeat(f1[None], t.shape[0], 0).reshape(-1, f1.shape[1])])
        y = np.concatenate([model.train_y, t.reshape(-1)])
        model.fit(X.astype(np.float32), y, epochs=epochs, lr=0.05, batch=4096)
        model.train_x, model.train_y = X.astype(np.float32), y
        info["refit_samples"] = int(t.size)
        pred0 = stage_costs(compiled, s0, model.predict(s0.part_features()))
        before = [lpt_makespan(c, blocks) for c in pred0]
        info["predicted_ticks_before"] = float(sum(before))
    pred1 = stage_costs(trial, s1, model.predict(s1.part_features()))
    after = [lpt_makespan(c, blocks) for c in pred1]
    keep = {m for m in range(len(pred0)) if after[m] < before[m]}
    kept = [p for p, o in zip(pairs, ok) if o and p[0] in keep]
    kept_cuts = [c for c, o in zip(cuts, okc) if o and c[0] in keep]
    info.update(merged=int(ok.sum()), cut=int(okc.sum()), stages_kept=len(keep),
                pairs_kept=len(kept), cuts_kept=len(kept_cuts),
                predicted_ticks_after=float(sum(after[m] if m in keep else before[m]
                                                for m in range(len(pred0)))))
    if log:
        log(f"learned cut/merge: {len(pairs)} merge pairs ({int(ok.sum())} fit one "
            f"partition), {len(cuts)} cuts ({int(okc.sum())} fit); {len(keep)} of "
            f"{len(pred0)} major stages keep theirs; predicted ticks per cycle "
            f"{info['predicted_ticks_before']:.0f} -> {info['predicted_ticks_after']:.0f}")
    if not kept and not kept_cuts:
        return compiled, pred0, info
    if len(kept) == int(ok.sum()) and len(kept_cuts) == int(okc.sum()):
        return trial, pred1, info
    final, _, _ = build(kept, kept_cuts)
    sf = final.flatten_shared(blocks)
    return final, stage_costs(final, sf, model.predict(sf.part_features())), info


def learn_costs(compiled, blocks, groups, sample_cycles=24, epochs=150, log=None):
    """Train the MLP on profiled batched launches of the script flattened
    for `blocks`, `blocks`-8 and `blocks`-16 blocks per group (distinct
    partition-to-block placements), and return (costs [major][index],
    model, info)."""
    import time
    widths = sorted({blocks, max(4, blocks - 8), max(4, blocks - 16)}, reverse=True)
    t = time.time()
    X, y, med = training_set(compiled, widths, groups, sample_cycles)
    t_prof = time.time() - t
    t = time.time()
    model = MLP(seed=1)
    hist = model.fit(X, y, epochs=epochs, lr=0.05, batch=4096)
    model.train_x, model.train_y = X, y  # kept for refits on new samples
    t_fit = time.time() - t
    s0 = compiled.flatten_shared(blocks)
    f0 = s0.part_features()
    pred = model.predict(f0)
    truth = np.array([med[(m, i)] for (m, _b, i, _o, _l) in s0.parts.tolist()])
    h = f0[:, 0] + 3.0
    a, b = np.polyfit(h, truth, 1)
    info = dict(samples=int(len(y)), profile_s=t_prof, fit_s=t_fit,
                loss_first=float(hist[0]), loss_last=float(hist[-1]),
                mape_heuristic_fit=float(np.mean(np.abs(a * h + b - truth) / truth)),
                mape_learned=float(np.mean(np.abs(pred - truth) / truth)))
    if log:
        log(f"cost model: {info['samples']} samples, loss {info['loss_first']:.3f} -> "
            f"{info['loss_last']:.3f}, MAPE heuristic-fit {info['mape_heuristic_fit']:.3f} "
            f"learned {info['mape_learned']:.3f}")
    return stage_costs(compiled, s0, pred), model, info
