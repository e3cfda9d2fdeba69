"""Pins of the loss and beam-splitter oracles (CPU): closed forms, invariants
and finite differences, so that a dropped term or a wrong index in the oracle
fails here before it is used to check the device kernels."""
import numpy as np

from oracle import losses_ref as L
from oracle import xlumina_ref as R


def test_small_area_of_a_flat_spot_is_the_total_intensity():
    # D = 1 on K pixels, delta < 0.7 elsewhere: S = K + delta (n - K), the mask
    # holds the K spot pixels, sum m I = K / S, so the loss is S (up to 1e-8 terms)
    n, K, delta = 4096, 37, 0.05
    d = np.full(n, delta)
    d[np.random.default_rng(0).choice(n, K, replace=False)] = 1.0
    S = K + delta * (n - K)
    assert abs(L.small_area_hybrid(d.reshape(64, 64)) - S) < 1e-5 * S


def test_small_area_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    d = rng.random((24, 24)) ** 4
    g = L.small_area_hybrid_grad(d)
    I = d / d.sum()
    far = np.abs(I - 0.7 * I.max()) > 1e-3 * I.max()   # perturbations leave the mask as is
    for q in rng.choice(np.flatnonzero(far.ravel()), 12, replace=False):
        h = 1e-6
        dp, dm = d.copy(), d.copy()
        dp.flat[q] += h
        dm.flat[q] -= h
        fd = (L.small_area_hybrid(dp) - L.small_area_hybrid(dm)) / (2 * h)
        assert abs(fd - g.flat[q]) <= 1e-5 * max(1.0, abs(fd))


def test_mse_closed_forms():
    rng = np.random.default_rng(2)
    e = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    assert np.isclose(L.MSE_Intensity(e, np.zeros_like(e)), np.sum(np.abs(e) ** 4) / 256)
    assert np.isclose(L.MSE_Intensity(e, 2 * e), np.sum(9 * np.abs(e) ** 4) / 256)
    assert abs(L.MSE_Amplitude(e, e * np.exp(0.7j))) < 1e-28
    assert np.isclose(L.MSE_Amplitude(3 * e, e), np.sum(4 * np.abs(e) ** 2) / 256)
    p = np.abs(e) * np.exp(1j * rng.uniform(-1.2, 1.2, (16, 16)))
    assert np.isclose(L.MSE_Phase(p * np.exp(0.3j), p), 0.09)
    m, per = L.mean_batch_MSE_Intensity(np.stack([e, 2 * e]), np.stack([0 * e, e]))
    assert np.allclose(per, [np.sum(np.abs(e) ** 4) / 256, np.sum(9 * np.abs(e) ** 4) / 256]) and np.isclose(m, per.mean())


def test_softmin_limits():
    a = np.array([0.4, 0.2, 0.9, 0.2])
    assert np.isclose(L.softmin(a), 0.2 - np.log(2) / 90, atol=1e-9)
    assert np.isclose(L.softmin(np.full(5, 1.5)), 1.5 - np.log(5) / 90)


def test_beam_splitter_energy_and_ports():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 8, 8)) + 1j * rng.standard_normal((2, 8, 8))
    b = rng.standard_normal((2, 8, 8)) + 1j * rng.standard_normal((2, 8, 8))
    th = 0.6
    c, d = R.BS_symmetric(a, b, th)
    Rr, T = abs(np.sin(th)) - 0.01 * abs(np.cos(th)), 0.99 * abs(np.cos(th))
    e_in = np.sum(np.abs(a) ** 2) + np.sum(np.abs(b) ** 2)
    e_out = sum(np.sum(np.abs(v) ** 2) for v in c + d)
    assert np.isclose(e_out, (Rr ** 2 + T ** 2) * e_in)   # cross terms cancel: [[R, iT], [iT, R]]
    c0, d0 = R.BS_symmetric_SI(a, th, 0)
    c1, d1 = R.BS_symmetric_SI(a, th, 1)
    assert all(np.array_equal(x, y) for x, y in zip(c0, d1)) and all(np.array_equal(x, y) for x, y in zip(d0, c1))
    b1, b2 = R.BS(a, b, 0.3, 0.7, 1.1)                     # lossless for R + T = 1
    assert np.isclose(sum(np.sum(np.abs(v) ** 2) for v in b1 + b2), e_in)
    z1, _ = R.BS(a, b, 0.0, 1.0, 0.4)
    assert np.allclose(z1[0], a[0]) and np.allclose(z1[1], a[1])
