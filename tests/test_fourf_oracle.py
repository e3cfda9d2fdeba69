"""Pins of the float64 4f-optimisation oracle (CPU only).

* the dense CZT operator reproduces the forward oracle on a random field
  (linearity: A u = CZT(u));
* the hand-written reverse-mode gradient equals central finite differences
  of the loss (eps = 1e-4 rad) on 24 seeded pixels of each phase mask;
* Adam's first two updates equal their closed forms (first step
  -lr g / (|g| + eps); second step from the two-term moment sums).
"""
import numpy as np

from oracle import fourf_ref as F
from oracle import xlumina_ref as R


def small_chain(n=12, seed=0):
    x, y = R.space(40.0, n)
    rng = np.random.default_rng(seed)
    e_in = np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / 30.0 ** 2).astype(np.complex128)
    T = (rng.random((n, n)) > 0.5).astype(np.float64)
    return F.FourF(x, y, 0.65, 2000.0, e_in, T), rng


def test_operator_is_the_forward_czt():
    n = 10
    x, y = R.space(40.0, n)
    A = F.czt_operator(x, y, 0.65, 1500.0)
    u = np.random.default_rng(1).standard_normal((n, n)) + 1j * np.random.default_rng(2).standard_normal((n, n))
    ref = R.czt(u, x, y, 0.65, 1500.0, x, y)
    assert np.allclose(A @ u.ravel(), ref.ravel(), rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_gradient_matches_finite_differences():
    ch, rng = small_chain()
    n = 12
    p1, p2 = rng.uniform(-np.pi, np.pi, (2, n, n))
    L, g1, g2 = ch.loss_and_grad(p1, p2)
    assert abs(L - ch.loss(p1, p2)) <= 1e-14 * abs(L)
    eps = 1e-4
    for which, g in ((0, g1), (1, g2)):
        idx = rng.choice(n * n, 24, replace=False)
        scale = np.abs(g).max()
        for j in idx:
            a, b = p1.copy(), p2.copy()
            tgt = a if which == 0 else b
            tgt.flat[j] += eps
            lp = ch.loss(a, b)
            tgt.flat[j] -= 2 * eps
            lm = ch.loss(a, b)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g.flat[j]) <= 1e-6 * scale, (which, j, fd, g.flat[j])


def test_adam_closed_forms():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(50)
    g1, g2 = rng.standard_normal((2, 50))
    lr, b1, b2, eps = 0.05, 0.9, 0.999, 1e-8
    st = (np.zeros(50), np.zeros(50), 0)
    p1, st = F.adam(p, g1, st, lr, b1, b2, eps)
    assert np.allclose(p1, p - lr * g1 / (np.abs(g1) + eps), rtol=1e-13, atol=1e-15)
    p2, st = F.adam(p1, g2, st, lr, b1, b2, eps)
    mu = (b1 * (1 - b1) * g1 + (1 - b1) * g2) / (1 - b1 ** 2)
    nu = (b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2) / (1 - b2 ** 2)
    assert np.allclose(p2, p1 - lr * mu / (np.sqrt(nu) + eps), rtol=1e-13, atol=1e-15)
    assert st[2] == 2
    # adamw: decoupled decay term lr * wd * p
    p3, _ = F.adam(p, g1, (np.zeros(50), np.zeros(50), 0), lr, b1, b2, eps, weight_decay=1e-4)
    assert np.allclose(p3, p - lr * (g1 / (np.abs(g1) + eps) + 1e-4 * p), rtol=1e-13, atol=1e-15)
