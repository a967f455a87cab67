"""The CPU oracle of the proposal network (oracle/mlp_oracle.py), pinned by
independent computations: torch's own bfloat16 conversion for the rounding,
torch autograd in float64 and central finite differences for the chain-rule
gradients, torch.optim.Adam for the optimiser step."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

torch = pytest.importorskip("torch")


def _data(n, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 2, (n, 48)).astype(np.float32)
    y = (rng.uniform(0, 1, (n, 2)) < 0.3).astype(np.float32)
    return x, y


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 3, 20000).astype(np.float32),
                        # exact ties: 1 + 2^-8 rounds to 1 (even), 1 + 3 * 2^-8 to 1 + 2^-6
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, 2 ** -130], dtype=np.float32)])
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(v), want)
    assert M.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert M.bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)


def _torch_net(p):
    w1, b1, w2, b2, w3, b3 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in M.unpack(p.astype(np.float64))]
    return [w1, b1, w2, b2, w3, b3]


def test_chain_rule_matches_autograd():
    x, y = _data(300)
    p = M.init_params(5)
    loss, g = M.gradients(p, x, y, bf16_rounding=False)
    w1, b1, w2, b2, w3, b3 = ts = _torch_net(p)
    X = torch.zeros((len(x), 64), dtype=torch.float64)
    X[:, :48] = torch.from_numpy(x.astype(np.float64))
    z = torch.relu(torch.relu(X @ w1 + b1) @ w2 + b2) @ w3 + b3
    L = torch.nn.functional.binary_cross_entropy_with_logits(z[:, :2], torch.from_numpy(y.astype(np.float64)))
    L.backward()
    want = np.concatenate([t.grad.numpy().ravel() for t in ts])
    # autograd also differentiates the padding columns of W3/b3 (zero logits
    # outside the loss give zero gradient) and the padded inputs (zero)
    assert abs(loss - L.item()) < 1e-12
    assert np.allclose(g, want, rtol=1e-9, atol=1e-12)


def test_chain_rule_matches_finite_differences():
    x, y = _data(64, seed=9)
    p = M.init_params(7).astype(np.float64)
    _, g = M.gradients(p, x, y, bf16_rounding=False)
    rng = np.random.default_rng(1)
    used = [k for k in range(M.NPARAMS) if not (M.W1 + 48 * 128 <= k < M.B1) and not (M.W3 <= k < M.B3 and (k - M.W3) % 16 >= 2)
            and not (M.B3 + 2 <= k < M.NPARAMS)]
    for k in rng.choice(used, 60, replace=False):
        h = 1e-6
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        fd = (M.loss(a, x, y, False) - M.loss(b, x, y, False)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_bf16_gradients_close_to_exact():
    x, y = _data(2000, seed=4)
    p = M.init_params(2)
    l16, g16 = M.gradients(p, x, y, True)
    l64, g64 = M.gradients(p, x, y, False)
    assert abs(l16 - l64) < 5e-3
    for lo, hi in ((M.W1, M.B1), (M.W2, M.B2), (M.W3, M.B3), (M.B1, M.W2)):
        scale = np.abs(g64[lo:hi]).max()
        assert np.abs(g16[lo:hi] - g64[lo:hi]).max() < 0.05 * scale


def test_adam_matches_torch():
    rng = np.random.default_rng(8)
    p = M.init_params(3)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    t = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    ours = p.copy()
    for step in range(1, 6):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        M.adam(ours, g, m, v, step, 2e-3)
        t.grad = torch.from_numpy(g.copy())
        opt.step()
    assert np.allclose(ours, t.detach().numpy(), rtol=1e-5, atol=1e-7)
