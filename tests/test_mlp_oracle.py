"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) -- the only check on the surrogate's
hand-written backward and Adam kernels -- to the mathematics, independently of its own code:

  * the forward pass and GELU against torch's float64 ``linear`` / ``gelu(approximate='tanh')``;
  * the gradient of the masked MSE against FP64 central finite differences of the loss, along
    a random direction restricted to each parameter block (every element of the block takes
    part) and on sampled single coordinates, and against torch.autograd;
  * GELU's derivative against finite differences;
  * ``adam_step`` against the closed form of Adam's first two steps on scalar sequences, and the
    invariant that a constant gradient g moves every step by exactly lr*g/(|g|+eps).

A planted mistake (a flipped sign in one term of the GELU derivative, planted in the oracle on
the CPU) must make the finite-difference check fail, so the check is shown to see such slips.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

BLOCKS = []
_o = 0
for _r, _c in mo.SHAPES:
    BLOCKS.append((f'W{len(BLOCKS) // 2 + 1}', _o, _r * _c)); _o += _r * _c
    BLOCKS.append((f'b{len(BLOCKS) // 2 + 1}', _o, _r)); _o += _r
NPAR = _o


def _case(seed=11, B=6):
    rng = np.random.default_rng(seed)
    p = np.empty(NPAR)
    for name, o, n in BLOCKS:
        fan_in = mo.SHAPES[int(name[1]) - 1][1]
        p[o:o + n] = rng.normal(0.0, 1.0 / np.sqrt(fan_in) if name[0] == 'W' else 0.3, size=n)
    X = rng.normal(size=(B, 64))
    T = rng.normal(size=(B, 64))
    return p, X, T


def _loss(p, X, T):
    """The masked MSE written from its definition: mean over the batch and the first 55 outputs."""
    Y, _ = mo.forward(p, X)
    return float(np.mean((Y[:, :mo.VALID] - T[:, :mo.VALID]) ** 2))


def _fd(p, X, T, d):
    """Fourth-order central difference of the loss along the step d (error O(|d|^4))."""
    f = lambda s: _loss(p + s * d, X, T)
    return (8.0 * (f(1) - f(-1)) - (f(2) - f(-2))) / 12.0


def _fd_errors(p, X, T, grad, h=1e-4, coords=6, seed=5):
    """Largest relative mismatch, per parameter block, between ``grad`` and central finite
    differences of the loss (directional along a random block-restricted vector, and on a few
    single coordinates)."""
    rng = np.random.default_rng(seed)
    errs = {}
    for name, o, n in BLOCKS:
        v = np.zeros(NPAR)
        v[o:o + n] = rng.normal(size=n)
        v /= np.linalg.norm(v)
        an = float(grad @ v)
        e = abs(_fd(p, X, T, h * v) / h - an) / max(abs(an), 1e-3 * np.linalg.norm(grad[o:o + n]))
        for i in rng.choice(n, size=min(coords, n), replace=False):
            d = np.zeros(NPAR)
            d[o + i] = h
            fdi = _fd(p, X, T, d) / h
            e = max(e, abs(fdi - grad[o + i]) / max(abs(grad[o + i]), 1e-3 * np.abs(grad[o:o + n]).max()))
        errs[name] = e
    return errs


def test_forward_and_gelu_match_torch():
    p, X, _ = _case()
    H = torch.from_numpy(X)
    for i, (W, b) in enumerate(mo.unpack(p)):
        Z = torch.nn.functional.linear(H, torch.from_numpy(W), torch.from_numpy(b))
        H = torch.nn.functional.gelu(Z, approximate='tanh') if i < 3 else Z
    Y, _ = mo.forward(p, X)
    assert np.max(np.abs(Y - H.numpy())) <= 1e-13 * np.max(np.abs(Y))
    x = np.linspace(-9, 9, 1001)
    ref = torch.nn.functional.gelu(torch.from_numpy(x), approximate='tanh').numpy()
    assert np.max(np.abs(mo.gelu(x) - ref)) <= 1e-15 * 9


def test_gelu_derivative_vs_finite_differences():
    x = np.linspace(-7, 7, 2801)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(mo.gelu_grad(x) - fd)) <= 1e-8
    # limits: GELU -> x for x >> 0 and -> 0 for x << 0
    assert abs(mo.gelu_grad(np.array([30.0]))[0] - 1.0) <= 1e-15
    assert abs(mo.gelu_grad(np.array([-30.0]))[0]) <= 1e-15
    assert abs(mo.gelu_grad(np.array([0.0]))[0] - 0.5) <= 1e-15


def test_loss_value_and_gradient_vs_finite_differences():
    p, X, T = _case()
    loss, Y, g = mo.loss_and_grad(p, X, T)
    assert abs(loss - _loss(p, X, T)) <= 1e-15 * loss
    errs = _fd_errors(p, X, T, g)
    assert max(errs.values()) <= 1e-7, errs


def test_gradient_vs_torch_autograd():
    p, X, T = _case(seed=12, B=9)
    _, _, g = mo.loss_and_grad(p, X, T)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    H, o = torch.from_numpy(X), 0
    for i, (r, c) in enumerate(mo.SHAPES):
        W = pt[o:o + r * c].view(r, c); o += r * c
        b = pt[o:o + r]; o += r
        Z = H @ W.T + b
        H = torch.nn.functional.gelu(Z, approximate='tanh') if i < 3 else Z
    Tt = torch.from_numpy(T)
    loss = ((H[:, :mo.VALID] - Tt[:, :mo.VALID]) ** 2).mean()
    loss.backward()
    ref = pt.grad.numpy()
    for name, o, n in BLOCKS:
        blk = slice(o, o + n)
        assert np.max(np.abs(g[blk] - ref[blk])) <= 1e-12 * np.max(np.abs(ref[blk])), name
    # outputs past the first 55 take no gradient: the last layer's rows 55..63 stay zero
    W4 = slice(BLOCKS[6][1] + mo.VALID * mo.HIDDEN, BLOCKS[6][1] + 64 * mo.HIDDEN)
    assert np.all(g[W4] == 0.0)


def test_planted_gelu_derivative_slip_is_caught(monkeypatch):
    """Flip the sign of the second term of the GELU derivative in the oracle (on the CPU): the
    finite-difference check must report it."""
    p, X, T = _case()

    def slipped(x):
        u = 0.7978845608028654 * (x + 0.044715 * x ** 3)
        th = np.tanh(u)
        return 0.5 * (1 + th) - 0.5 * x * (1 - th ** 2) * 0.7978845608028654 * (1 + 3 * 0.044715 * x ** 2)

    monkeypatch.setattr(mo, 'gelu_grad', slipped)
    _, _, g = mo.loss_and_grad(p, X, T)
    errs = _fd_errors(p, X, T, g)
    assert max(errs.values()) > 1e-2, errs


@pytest.mark.parametrize('g1,g2', [(0.3, -1.7), (-2e-3, 5e-3), (4.0, 4.0)])
def test_adam_first_two_steps_closed_form(g1, g2):
    lr, b1, b2, eps, p0 = 1e-2, 0.9, 0.999, 1e-8, 0.25
    p = np.array([p0]); m = np.zeros(1); v = np.zeros(1)
    p, m, v = mo.adam_step(p, np.array([g1]), m, v, lr, b1, b2, eps, 1)
    # step 1: the bias-corrected moments are g1 and g1^2 exactly
    p1 = p0 - lr * g1 / (abs(g1) + eps)
    assert abs(p[0] - p1) <= 1e-15
    assert abs(m[0] - (1 - b1) * g1) <= 1e-17 and abs(v[0] - (1 - b2) * g1 ** 2) <= 1e-17
    p, m, v = mo.adam_step(p, np.array([g2]), m, v, lr, b1, b2, eps, 2)
    # step 2: m_hat = (b1 g1 + g2)/(1 + b1), v_hat = (b2 g1^2 + g2^2)/(1 + b2)
    mh = (b1 * g1 + g2) / (1 + b1)
    vh = (b2 * g1 ** 2 + g2 ** 2) / (1 + b2)
    p2 = p1 - lr * mh / (np.sqrt(vh) + eps)
    assert abs(p[0] - p2) <= 1e-14


def test_adam_constant_gradient_invariant():
    lr, b1, b2, eps = 3e-3, 0.85, 0.995, 1e-8
    g = np.array([0.7, -0.02, 5.0, -300.0])
    p = np.zeros(4); m = np.zeros(4); v = np.zeros(4)
    for step in range(1, 40):
        q, m, v = mo.adam_step(p, g, m, v, lr, b1, b2, eps, step)
        assert np.max(np.abs((p - q) - lr * g / (np.abs(g) + eps))) <= 1e-15
        p = q
