"""The proposal network's gfx950 MFMA kernels through the C ABI against the CPU
oracle (oracle/mlp_oracle.py, rounding to BF16 where the kernels do): forward
probabilities, loss and gradients, and Adam training step by step on the same
batches; then training on a learnable rule, and the proposal in the aligner
(a threshold nothing falls below leaves the GAF of the plain method)."""
import os
import sys

import numpy as np
import pytest

import graphaligner_amd as ga
import oracle_ffi as O
from workloads import workload

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

MASK64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def ctx():
    return ga.Context(0)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1.5, (n, 48)).astype(np.float32)
    y = (rng.uniform(0, 1, (n, 2)) < 0.3).astype(np.float32)
    return x, y


def _blocks():
    return {"W1": (M.W1, M.B1), "b1": (M.B1, M.W2), "W2": (M.W2, M.B2), "b2": (M.B2, M.W3), "W3": (M.W3, M.B3), "b3": (M.B3, M.NPARAMS)}


def test_gpu_mlp_params_roundtrip(ctx):
    net = ga.Mlp(ctx, seed=3)
    p = net.params()
    assert p.size == M.NPARAMS and np.isfinite(p).all()
    # padding is zero: the 16 padded inputs and the 14 padded outputs
    assert not p[M.W1 + 48 * 128:M.B1].any() and not p[M.W3:M.B3].reshape(128, 16)[:, 2:].any() and not p[M.B3 + 2:].any()
    q = M.init_params(11)
    net.set_params(q)
    assert np.array_equal(net.params(), q)
    bad = q.copy()
    bad[M.B3 + 5] = 1.0
    with pytest.raises(RuntimeError, match="GA_EBADARG"):
        net.set_params(bad)


@pytest.mark.parametrize("n", [1, 17, 1000, 5003])
def test_gpu_mlp_forward_matches_oracle(ctx, n):
    net = ga.Mlp(ctx)
    p = M.init_params(n)
    net.set_params(p)
    x, _ = _data(n, n)
    print(f"[gpu-mlp] forward n={n}", flush=True)
    got = net.forward(x)
    want = M.probabilities(p, x, True)
    assert got.shape == (n, 2)
    assert np.abs(got - want).max() < 2e-3, np.abs(got - want).max()


@pytest.mark.parametrize("n", [16, 333, 4096])
def test_gpu_mlp_gradients_match_oracle(ctx, n):
    """Same rounding points, float32 sums in another order: each block within 2%
    of its largest entry, the loss to 1e-4."""
    net = ga.Mlp(ctx)
    p = M.init_params(100 + n)
    net.set_params(p)
    x, y = _data(n, 200 + n)
    print(f"[gpu-mlp] gradients n={n}", flush=True)
    loss, g = net.gradients(x, y)
    wl, wg = M.gradients(p, x, y, True)
    assert abs(loss - wl) < 1e-4
    for name, (lo, hi) in _blocks().items():
        scale = max(np.abs(wg[lo:hi]).max(), 1e-12)
        err = np.abs(g[lo:hi] - wg[lo:hi]).max()
        assert err <= 0.02 * scale, (name, err, scale)
    # and close to the exact (float64, unrounded) gradient: BF16 keeps 8
    # significant bits, so a block's error norm stays within a few % of its norm
    _, exact = M.gradients(p, x, y, False)
    for name, (lo, hi) in _blocks().items():
        norm = max(np.linalg.norm(exact[lo:hi]), 1e-12)
        assert np.linalg.norm(g[lo:hi] - exact[lo:hi]) <= 0.05 * norm, name


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def _epoch_orders(n, epochs, seed):
    """The record order of each epoch (Fisher-Yates over splitmix64, as the library shuffles)."""
    st = (seed * 0x9E3779B97F4A7C15 + 7) & MASK64
    out = []
    for _ in range(epochs):
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            st, r = _splitmix(st)
            j = r % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        out.append(np.array(perm))
    return out


def test_gpu_mlp_training_steps_match_oracle(ctx):
    """Two epochs of batches 128, 128, 44 on the GPU and in the oracle (same record
    order, same Adam): the parameters agree. Adam's first steps move a weight by
    about lr whatever its gradient's size, so a weight whose tiny gradient has the
    other sign (float32 sum order, a BF16 tie) ends up to 2 lr per step apart."""
    n, batch, lr, epochs = 300, 128, 2e-3, 2
    x, y = _data(n, 77)
    p = M.init_params(5)
    net = ga.Mlp(ctx)
    net.set_params(p)
    losses, ms = net.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=9)
    got = net.params()
    ours = p.copy()
    m = np.zeros_like(ours)
    v = np.zeros_like(ours)
    t = 0
    want_losses = []
    for order in _epoch_orders(n, epochs, 9):
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = order[b0:b0 + batch]
            t += 1
            lb, g = M.gradients(ours, x[idx], y[idx], True)
            tot += lb * len(idx)
            M.adam(ours, g.astype(np.float32), m, v, t, lr)
        want_losses.append(tot / n)
    d = np.abs(got - ours)
    assert np.allclose(losses, want_losses, atol=2e-3), (losses, want_losses)
    assert np.median(d) < 0.05 * lr and np.quantile(d, 0.99) < 2 * lr and d.max() < 2 * lr * t + 1e-6, (np.median(d), d.max())
    assert ms > 0


def test_gpu_mlp_learns_a_rule(ctx):
    """Labels from a fixed nonlinear rule of the features: training on the GPU
    brings the loss well below the constant predictor's and classifies held-out
    records."""
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, (60000, 48)).astype(np.float32)
    y0 = ((x[:, 3] + x[:, 7] * x[:, 11]) > 0.8).astype(np.float32)
    y1 = (np.abs(x[:, 20] - 0.5) < 0.25).astype(np.float32)
    y = np.stack([y0, y1], 1)
    net = ga.Mlp(ctx, seed=2)
    losses, ms = net.train(x[:50000], y[:50000], epochs=20, batch=1024, lr=3e-3, seed=4)
    print(f"[gpu-mlp] losses {losses} ({ms:.1f} ms)", flush=True)
    assert losses[-1] < 0.3 * losses[0] and losses[-1] < 0.15
    pred = net.forward(x[50000:]) > 0.5
    acc = (pred == (y[50000:] > 0.5)).mean(axis=0)
    assert acc[0] > 0.95 and acc[1] > 0.95, acc


def test_gpu_proposal_threshold_zero_is_plain_and_records(ctx, workdir):
    gfa, names, seqs = workload("mini_sv", workdir)
    params = ga.params_preset("vg")
    g = ga.Graph(gfa)
    gi = ga.Index(g, params)
    ga.upload_reads(ctx, seqs)
    plain, pst, _ = ga.align_resident(ctx, g, gi, params, 0, names)
    feats, labels, reads = ga.cluster_records(ctx, g, gi, params, 0, len(seqs))
    assert len(feats) > len(seqs) and np.isfinite(feats).all()
    assert (labels[:, 0] <= labels[:, 1]).all()
    for i in range(len(seqs)):
        if (reads == i).any():
            assert labels[reads == i, 0].any() == (pst[i] == 0)
    net = ga.Mlp(ctx, seed=5)
    probs = net.forward(feats)
    for loop in (0, 1):
        params.deviceSeedLoop = loop
        ga.set_proposal(ctx, net, threshold=0.0)
        text, st, stats = ga.align_resident(ctx, g, gi, params, 0, names)
        assert text == plain and list(st) == list(pst)
        assert stats["proposalClusters"] == len(feats) and stats["prunedClusters"] == 0
        # a threshold above every score keeps only each read's first cluster
        ga.set_proposal(ctx, net, threshold=float(probs[:, 0].max()) + 1e-3)
        _, _, stats = ga.align_resident(ctx, g, gi, params, 0, names)
        first = len(np.unique(reads))
        assert stats["prunedClusters"] == len(feats) - first
        ga.set_proposal(ctx, None)
    params.deviceSeedLoop = 0
    text, _, stats = ga.align_resident(ctx, g, gi, params, 0, names)
    assert text == plain and stats["proposalClusters"] == 0

