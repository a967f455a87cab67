"""CPU oracle of the learned proposal's network (numpy), used ONLY as test
infrastructure: the checker of the GPU forward pass, gradients and training.
Nothing in the product imports it.

The network has no counterpart in the reference (it is the approximate mode
beside the plain method). Its definition: 48 features (padded to 64 with
zeros) -> 128 -> 128 -> 2 (padded to 16), ReLU hidden layers, sigmoid outputs,
loss = mean binary cross-entropy over records and both outputs. Parameters in
the product's flat layout: W1[64][128] (input-major), b1[128], W2[128][128],
b2[128], W3[128][16], b3[16].

bf16=True rounds at exactly the points the gfx950 kernels do (round to nearest
even): the inputs, the weights, each hidden layer's output after ReLU, the
output gradient and each hidden layer's pre-activation gradient; every sum
accumulates in float32. bf16=False is the same network in float64 without
rounding: the mathematical definition that finite differences pin
(tests/test_mlp_oracle.py).
"""
import numpy as np

IN, IN_USED, H, OUT, OUT_USED = 64, 48, 128, 16, 2
W1 = 0
B1 = W1 + IN * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * OUT
NPARAMS = B3 + OUT


def bf16(x):
    """float32 values rounded to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    return (p[W1:B1].reshape(IN, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, OUT), p[B3:NPARAMS])


def _inputs(features, dtype):
    x = np.zeros((len(features), IN), dtype=dtype)
    x[:, :IN_USED] = features
    return x


def forward(params, features, bf16_rounding=True):
    """(h1, h2, logits[:, :2]) for features [n, 48]."""
    if bf16_rounding:
        w1, b1, w2, b2, w3, b3 = [a.astype(np.float32) for a in unpack(params)]
        w1, w2, w3 = bf16(w1), bf16(w2), bf16(w3)
        x = bf16(_inputs(np.asarray(features, dtype=np.float32), np.float32))
        h1 = np.maximum(x @ w1 + b1, 0).astype(np.float32)
        h1b = bf16(h1)
        h2 = np.maximum(h1b @ w2 + b2, 0).astype(np.float32)
        h2b = bf16(h2)
        o = (h2b @ w3 + b3).astype(np.float32)
        return x, h1, h1b, h2, h2b, o
    w1, b1, w2, b2, w3, b3 = [a.astype(np.float64) for a in unpack(params)]
    x = _inputs(np.asarray(features, dtype=np.float64), np.float64)
    h1 = np.maximum(x @ w1 + b1, 0)
    h2 = np.maximum(h1 @ w2 + b2, 0)
    o = h2 @ w3 + b3
    return x, h1, h1, h2, h2, o


def probabilities(params, features, bf16_rounding=True):
    o = forward(params, features, bf16_rounding)[5][:, :OUT_USED]
    return 1.0 / (1.0 + np.exp(-o))


def loss(params, features, labels, bf16_rounding=True):
    z = forward(params, features, bf16_rounding)[5][:, :OUT_USED].astype(np.float64)
    y = np.asarray(labels, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def gradients(params, features, labels, bf16_rounding=True):
    """(mean loss, flat gradient) by the chain rule."""
    n = len(features)
    x, h1, h1b, h2, h2b, o = forward(params, features, bf16_rounding)
    ft = np.float32 if bf16_rounding else np.float64
    w1, b1, w2, b2, w3, b3 = [a.astype(ft) for a in unpack(params)]
    if bf16_rounding:
        w2, w3 = bf16(w2), bf16(w3)
    y = np.zeros((n, OUT), dtype=ft)
    y[:, :OUT_USED] = labels
    z = o.astype(ft)
    g = np.zeros((n, OUT), dtype=ft)
    g[:, :OUT_USED] = (1.0 / (1.0 + np.exp(-z[:, :OUT_USED])) - y[:, :OUT_USED]) * ft(1.0 / (2 * n))
    if bf16_rounding:
        g = bf16(g)
    dh2 = (g @ w3.T) * (h2 > 0)
    if bf16_rounding:
        dh2 = bf16(dh2.astype(np.float32))
    dh1 = (dh2 @ w2.T) * (h1 > 0)
    if bf16_rounding:
        dh1 = bf16(dh1.astype(np.float32))
    grads = np.zeros(NPARAMS, dtype=ft)
    grads[W1:B1] = (x.T @ dh1).ravel()
    grads[B1:W2] = dh1.sum(axis=0)
    grads[W2:B2] = (h1b.T @ dh2).ravel()
    grads[B2:W3] = dh2.sum(axis=0)
    grads[W3:B3] = (h2b.T @ g).ravel()
    grads[B3:NPARAMS] = g.sum(axis=0)
    return loss(params, features, labels, bf16_rounding), grads


def init_params(seed=0, scale=1.0):
    """Random parameters with zero padding (tests only)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    w1 = np.zeros((IN, H), np.float32)
    w1[:IN_USED] = rng.uniform(-1, 1, (IN_USED, H)) * np.sqrt(6 / IN_USED) * scale
    w3 = np.zeros((H, OUT), np.float32)
    w3[:, :OUT_USED] = rng.uniform(-1, 1, (H, OUT_USED)) * np.sqrt(6 / (H + OUT_USED)) * scale
    p[W1:B1] = w1.ravel()
    p[B1:W2] = rng.uniform(-0.1, 0.1, H)
    p[W2:B2] = (rng.uniform(-1, 1, (H, H)) * np.sqrt(6 / H) * scale).ravel()
    p[B2:W3] = rng.uniform(-0.1, 0.1, H)
    p[W3:B3] = w3.ravel()
    p[B3:B3 + OUT_USED] = rng.uniform(-0.1, 0.1, OUT_USED)
    return p


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba, 2015), float32, in place; returns params."""
    m[:] = beta1 * m + (1 - beta1) * grads
    v[:] = beta2 * v + (1 - beta2) * grads * grads
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    params -= (lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return params
