"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the surrogate MLP (forward, backward of
the masked MSE, one Adam step), the reference the tests hold kinetix_amd/csrc/kx_mlp.hip to.

Layer layout and parameter block as in kx_mlp.hip: W1[H][64] b1 W2[H][H] b2 W3 b3 W4[64][H]
b4 with H = HIDDEN = 256; GELU in its tanh form; loss = mean over the batch and the first 55
outputs of (Y - T)^2.

Pinned by tests/test_mlp_oracle.py, independently of this code: the forward pass and GELU
against torch's float64 operators, the gradient against fourth-order finite differences of the
loss (every parameter block) and torch.autograd, GELU's derivative against finite differences,
and adam_step against the closed form of Adam's first two steps.
"""
import numpy as np

HIDDEN = 256
SHAPES = ((HIDDEN, 64), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (64, HIDDEN))
VALID = 55


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for r, c in SHAPES:
        W = p[o:o + r * c].reshape(r, c); o += r * c
        b = p[o:o + r]; o += r
        out.append((W, b))
    assert o == p.size
    return out


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(0.7978845608028654 * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = 0.7978845608028654 * (x + 0.044715 * x ** 3)
    th = np.tanh(u)
    return 0.5 * (1 + th) + 0.5 * x * (1 - th ** 2) * 0.7978845608028654 * (1 + 3 * 0.044715 * x ** 2)


def forward(p, X):
    layers = unpack(p)
    H = np.asarray(X, np.float64)
    cache = []
    for i, (W, b) in enumerate(layers):
        Z = H @ W.T + b
        cache.append((H, Z))
        H = gelu(Z) if i < 3 else Z
    return H, cache


def loss_and_grad(p, X, T):
    Y, cache = forward(p, X)
    B = Y.shape[0]
    d = np.zeros_like(Y)
    d[:, :VALID] = Y[:, :VALID] - np.asarray(T, np.float64)[:, :VALID]
    loss = (d ** 2).sum() / (B * VALID)
    dY = 2.0 * d / (B * VALID)
    grads = []
    layers = unpack(p)
    g = dY
    for i in (3, 2, 1, 0):
        H, Z = cache[i]
        if i < 3:
            g = g * gelu_grad(Z)
        W, _ = layers[i]
        grads.append((g.T @ H, g.sum(0)))
        g = g @ W
    flat = []
    for (dW, db) in reversed(grads):
        flat += [dW.ravel(), db]
    return loss, Y, np.concatenate(flat)


def adam_step(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
