"""The oracle against what pins it independently (CPU only).

Pins: direct O(N^4) RS sums, the RS kernel as a numerical z-derivative of the
free-space Green's function, dense chirp-z DFT and scipy.signal.czt, the
reference's own value tests for the Jones masks, algebraic identities of the
Jones / Richards-Wolf matrices, closed forms of the sources and of the
quality factor, and the committed golden vectors.
"""
import numpy as np
import pytest
import scipy.signal as ssig

from oracle import brute as B
from oracle import xlumina_ref as R
from tests.conftest import rel_l2

RNG = np.random.default_rng(7)


def cfield(n, m=None):
    m = n if m is None else m
    return RNG.standard_normal((n, m)) + 1j * RNG.standard_normal((n, m))


# -- grids -----------------------------------------------------------------------

def test_space_and_lag_grid():
    x, y = R.space(60.0, 9)
    assert np.allclose(x, np.linspace(-60, 60, 9)) and x.dtype == np.float64
    nx, ny, dx, dy, Xext, Yext = R.build_grid(x, y)
    assert (nx, ny) == (9, 9) and Xext.shape == (17, 17)
    # monotone lags -(N-1)dx .. (N-1)dx along x, constant along y
    assert np.allclose(Xext[0], np.arange(-8, 9) * dx)
    assert np.allclose(Yext[:, 0], np.arange(-8, 9) * dy)


def test_np2_and_rejections():
    assert R.compute_np2(2559) == 4096 and R.compute_np2(2047) == 2048 and R.compute_np2(4095) == 4096
    with pytest.raises(ValueError):
        R.check_czt_lengths(2048, 2049)     # mp = 4096, a power of two
    with pytest.raises(ValueError):
        R.check_czt_lengths(64, 1)
    R.check_czt_lengths(2048, 512)


# -- RS --------------------------------------------------------------------------------

def test_rs_kernel_is_greens_function_derivative():
    k = 2 * np.pi / 0.65
    X = np.linspace(-50, 50, 11)[:, None]
    Y = np.linspace(-30, 30, 7)[None, :]
    for z in (150.0, 900.0):
        assert rel_l2(B.rs_kernel(X, Y, z, k), B.greens_dz_numeric(X, Y, z, k, 1e-4)) < 1e-6
        # the oracle's transfer function is the same kernel
        assert rel_l2(R.transfer_function_RS(z, X, Y, k), B.rs_kernel(X, Y, z, k)) < 1e-13


@pytest.mark.parametrize("n,z", [(16, 300.0), (17, -250.0), (24, 1000.0)])
def test_rs_equals_direct_sum(n, z):
    x, y = R.space(3.5 * n, n)
    u = cfield(n)
    out, _ = R.rs_propagation(u, x, y, 0.65, z)
    assert rel_l2(out, B.rs_direct(u, x, y, 0.65, z)) < 1e-11


def test_rs_linearity_and_shift():
    n = 16
    x, y = R.space(50.0, n)
    a, b = cfield(n), cfield(n)
    fa, _ = R.rs_propagation(a, x, y, 0.65, 400.0)
    fb, _ = R.rs_propagation(b, x, y, 0.65, 400.0)
    fab, _ = R.rs_propagation(2 * a - 3j * b, x, y, 0.65, 400.0)
    assert rel_l2(fab, 2 * fa - 3j * fb) < 1e-12
    # a point source gives the kernel itself (times dx dy) on the grid
    d = np.zeros((n, n), complex)
    d[0, 0] = 1.0
    out, _ = R.rs_propagation(d, x, y, 0.65, 400.0)
    dx = x[1] - x[0]
    X, Y = np.meshgrid(x - x[0], y - y[0])
    assert rel_l2(out, B.rs_kernel(X, Y, 400.0, 2 * np.pi / 0.65) * dx * dx) < 1e-12


def test_quality_factor_and_minimum_z():
    x, y = R.space(1200.0, 1024)
    qf = R.quality_factor(x, y, 0.65, 5e4)
    dx = x[1] - x[0]
    rmax = np.hypot(1200.0, 1200.0)
    expect = (np.sqrt(0.65 ** 2 + rmax ** 2 + 2 * 0.65 * np.sqrt(rmax ** 2 + 5e4 ** 2)) - rmax) / np.hypot(dx, dx)
    assert abs(qf - expect) < 1e-12 * expect
    # minimum z solves quality(z) = target for the reference's dx = range/N convention
    zmin = R.minimum_z(x, y, 0.65, quality_factor=1)
    rng_ = 2400.0
    dr = np.hypot(rng_ / 1024, rng_ / 1024)
    rm = np.hypot(rng_, rng_)
    q = (np.sqrt(0.65 ** 2 + rm ** 2 + 2 * 0.65 * np.sqrt(rm ** 2 + zmin ** 2)) - rm) / dr
    assert abs(q - 1.0) < 1e-6


def test_vrs_components():
    n = 16
    x, y = R.space(60.0, n)
    Ex, Ey = cfield(n), cfield(n)
    (ox, oy, oz), _ = R.vrs_propagation(Ex, Ey, x, y, 0.635, 500.0)
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + 500.0 ** 2)
    assert rel_l2(ox, B.rs_direct(Ex, x, y, 0.635, 500.0)) < 1e-11
    assert rel_l2(oz, B.rs_direct(Ex * X / r + Ey * Y / r, x, y, 0.635, 500.0)) < 1e-11


# -- CZT ---------------------------------------------------------------------------------

@pytest.mark.parametrize("m,M", [(32, 12), (33, 7), (20, 20), (16, 24)])
def test_bluestein_equals_direct_dft(m, M):
    x = cfield(m, 5)
    Dm = 37.0
    f1, f2 = -3 + Dm / 2, 4 + Dm / 2
    out = R.bluestein_method(x, f1, f2, Dm, M)
    assert out.shape == (5, M)
    assert rel_l2(out, B.czt_axis_direct(x, f1, f2, Dm, M)) < 1e-12
    if M < m:
        assert rel_l2(out, B.czt_axis_scipy(x, f1, f2, Dm, M)) < 1e-12


def test_czt_missing_tap_quirk_is_exactly_one_term():
    # M >= m: the reference's chirp table is one entry short, which removes
    # exactly the x_0 contribution from the last output sample.
    m = M = 20
    x = cfield(m, 3)
    Dm, f1, f2 = 37.0, 15.0, 26.0
    full = B.czt_matrix(m, f1, f2, Dm, M)
    full_nq = full.copy()
    D1, D2, delta, theta = B._czt_params(f1, f2, Dm, M)
    l = D1 + (M - 1) * delta
    post = np.exp(-1j * theta * (M - 0.5)) * np.exp(-1j * 2 * np.pi * l * (0.5 - m / 2) / Dm)
    full_nq[M - 1, 0] = post * 1.0
    out = R.bluestein_method(x, f1, f2, Dm, M)
    diff = (full_nq @ x).T - out
    assert np.allclose(diff[:, :M - 1], 0, atol=1e-12)
    assert np.allclose(diff[:, M - 1], post * x[0], atol=1e-12)


def test_czt_scipy_pin_on_dft_grid():
    # with M = m and the DFT frequency grid, the chirp-z is the FFT (scipy.signal.czt defaults)
    m = 24
    x = cfield(m, 2)
    np.testing.assert_allclose(ssig.czt(x, axis=0), np.fft.fft(x, axis=0), atol=1e-11)


def test_czt2d_equals_direct():
    n = 24
    x, y = R.space(100.0, n)
    xo = np.linspace(-20, 20, 10)
    yo = np.linspace(-15, 15, 12)
    u = cfield(n)
    assert rel_l2(R.czt(u, x, y, 0.65, 1000.0, xo, yo), B.czt2d_direct(u, x, y, 0.65, 1000.0, xo, yo)) < 1e-12


def test_czt_samples_rs_field_at_shifted_points():
    # Physical pin of CZT_jit: its magnitude matches the direct RS integral
    # sampled at out[0] + (k + 1.5) * Delta, Delta = (out[-1] - out[0]) / M
    # (the reference's one-sample window offset), in the far field of a small
    # aperture (the phase differs: F0 is evaluated at the nominal points).
    n = 48
    x, y = R.space(40.0, n)
    X, Y = np.meshgrid(x, y)
    u = np.exp(-(X ** 2 + Y ** 2) / 15.0 ** 2).astype(complex)
    z = 30000.0
    xo = np.linspace(-300, 300, 9)
    out = R.czt(u, x, y, 0.65, z, xo, xo)
    delta = (xo[-1] - xo[0]) / len(xo)
    xs = xo[0] + (np.arange(len(xo)) + 1.5) * delta
    k = 2 * np.pi / 0.65
    dx = x[1] - x[0]
    ref = np.array([[np.sum(u * B.rs_kernel(xp - X, yq - Y, z, k)) * dx * dx for xp in xs] for yq in xs])
    assert rel_l2(np.abs(out), np.abs(ref)) < 1e-3
    # unshifted sample positions do not match
    ref0 = np.array([[np.sum(u * B.rs_kernel(xp - X, yq - Y, z, k)) * dx * dx for xp in xo] for yq in xo])
    assert rel_l2(np.abs(out), np.abs(ref0)) > 0.1


def test_vczt_components():
    n = 16
    x, y = R.space(80.0, n)
    xo = np.linspace(-10, 10, 9)
    Ex, Ey = cfield(n), cfield(n)
    out = R.vczt(Ex, Ey, x, y, 0.65, 900.0, xo, xo)
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + 900.0 ** 2)
    ez = (Ex * X / r + Ey * Y / r) * 900.0 / r
    assert rel_l2(out[2], B.czt2d_direct(ez, x, y, 0.65, 900.0, xo, xo)) < 1e-12


# -- Jones masks (reference value tests + identities) -----------------------------------------

def test_jones_reference_value_tests():
    n = 32
    x, y = R.space(1500.0, n)
    Ex, Ey, Ez = R.polarized_gaussian_beam(x, y, 0.633, (1200, 1200), (1, 1))
    z0 = np.zeros((n, n))
    # test_optical_elements.py:32-38: SLM with zero phase is the identity
    u = R.gaussian_beam(x, y, 0.633, (1200, 1200), 1.0)
    assert np.allclose(R.slm(u, z0)[0], u)
    # :57-74 sSLM zero phase identity, pi phase -> -E
    ox, oy, oz = R.sSLM(Ex, Ey, Ez, z0, z0)
    assert np.allclose(ox, Ex) and np.allclose(oy, Ey) and np.allclose(oz, Ez)
    ox, oy, _ = R.sSLM(Ex, Ey, Ez, np.pi + z0, np.pi + z0)
    assert np.allclose(ox, -Ex) and np.allclose(oy, -Ey)
    # :76-85 LCD(0, 0) identity (horizontal input)
    Hx, Hy, Hz = R.polarized_gaussian_beam(x, y, 0.633, (1200, 1200), (1, 0))
    ox, oy, oz = R.LCD(Hx, Hy, Hz, 0.0, 0.0)
    assert np.allclose(ox, Hx) and np.allclose(oy, Hy) and np.allclose(oz, Hz)
    # :87-101 polariser aligned = identity, crossed = 0
    ox, oy, _ = R.linear_polarizer(Hx, Hy, Hz, z0)
    assert np.allclose(ox, Hx) and np.allclose(oy, Hy)
    ox, oy, _ = R.linear_polarizer(Hx, Hy, Hz, np.pi / 2 + z0)
    assert np.allclose(ox, 0) and np.allclose(oy, 0)


def test_jones_identities():
    for eta, th in RNG.uniform(0, 2 * np.pi, (6, 2)):
        J = R.jones_LCD(eta, th)
        assert np.allclose(J.conj().T @ J, np.eye(2), atol=1e-12)          # unitary retarder
        ev = np.sort(np.angle(np.linalg.eigvals(J)))
        expect = np.sort(np.angle(np.exp(np.array([-1j, 1j]) * eta / 2)))
        assert np.allclose(ev, expect, atol=1e-12)                          # eigenphases -+eta/2
        L = R.jones_LP(th)
        assert np.allclose(L @ L, L) and np.allclose(L, L.T)                 # symmetric projector
        assert np.allclose(L @ np.array([np.cos(th), np.sin(th)]), [np.cos(th), np.sin(th)])
    S = R.jones_sSLM(0.3, -1.1)
    assert np.allclose(S, np.diag(np.exp(1j * np.array([0.3, -1.1]))))


def _rot2(t):
    return np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])


def _retarder_axis_residual(jones_lcd):
    """Largest deviation of a retarder J(eta, theta) from the one whose fast
    axis is (cos theta, sin theta) (optical_elements.py:123-139, theta = angle
    of the fast axis w.r.t. horizontal): J (c, s) = e^{-i eta/2} (c, s),
    J (-s, c) = e^{+i eta/2} (-s, c), and J = R(theta) diag(e^{-i eta/2},
    e^{i eta/2}) R(-theta).  Eigenphases alone cannot see the axis direction."""
    worst = 0.0
    for eta, th in np.random.default_rng(11).uniform(0, 2 * np.pi, (16, 2)):
        J = np.asarray(jones_lcd(eta, th), dtype=np.complex128)
        fast = np.array([np.cos(th), np.sin(th)])
        slow = np.array([-np.sin(th), np.cos(th)])
        built = _rot2(th) @ np.diag(np.exp(np.array([-0.5j, 0.5j]) * eta)) @ _rot2(-th)
        worst = max(worst, np.abs(J @ fast - np.exp(-0.5j * eta) * fast).max(),
                    np.abs(J @ slow - np.exp(0.5j * eta) * slow).max(), np.abs(J - built).max())
    return worst


def test_retarder_fast_axis_orientation():
    """Pins the sign of theta in the retarder (the per-pixel LCD of the
    2048^2 VRS workload): the oracle's matrix has its fast axis along
    (cos theta, sin theta).  The same check applied to the oracle with theta's
    sign flipped (the mistake that unitarity and the eigenphases both miss)
    fails, which the second assertion demonstrates on the CPU."""
    assert _retarder_axis_residual(R.jones_LCD) < 1e-12
    assert _retarder_axis_residual(lambda eta, th: R.jones_LCD(eta, -th)) > 0.1


# -- high-NA -----------------------------------------------------------------------------------

def test_richards_wolf_matrix_is_rotation():
    th, ph = RNG.uniform(0, 1.2, 5), RNG.uniform(-np.pi, np.pi, 5)
    for t, p in zip(th, ph):
        M = np.array(R.richards_wolf_matrix(np.cos(t), np.sin(t), np.cos(p), np.sin(p)))
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(M) - 1) < 1e-12
    # phi = 0: rotation about y by theta
    t = 0.7
    M = np.array(R.richards_wolf_matrix(np.cos(t), np.sin(t), 1.0, 0.0))
    assert np.allclose(M, [[np.cos(t), 0, -np.sin(t)], [0, 1, 0], [np.sin(t), 0, np.cos(t)]])


def test_high_na_lens_norm_and_constants():
    n = 16
    x, y = R.space(1800.0, n)
    Ex, Ey = cfield(n), cfield(n)
    out, stm = R.high_NA_objective_lens(Ex, Ey, x, y, 1800.0, 2000.0)
    assert abs(stm - 1800.0 / np.hypot(1800.0, 2000.0)) < 1e-15
    X, Y = np.meshgrid(x, y)
    rho = np.hypot(X, Y)
    theta = rho / 2000.0
    G = np.where(rho < 1800.0, 1 / np.sqrt(np.abs(1 - (rho / 1800.0) ** 2 * stm ** 2)), 0)
    w = np.sqrt(np.abs(np.cos(theta))) * G
    ez = (Ex * X + Ey * Y) / rho
    inten_in = np.abs(Ex) ** 2 + np.abs(Ey) ** 2 + np.abs(ez) ** 2
    assert np.allclose(np.sum(np.abs(out) ** 2, 0), w ** 2 * inten_in)


def _rz(p):
    return np.array([[np.cos(p), -np.sin(p), 0.0], [np.sin(p), np.cos(p), 0.0], [0.0, 0.0, 1.0]])


def _ry(t):
    return np.array([[np.cos(t), 0.0, np.sin(t)], [0.0, 1.0, 0.0], [-np.sin(t), 0.0, np.cos(t)]])


def _hna_independent(Ex, Ey, x, y, radius, f):
    """The lens map built pixel by pixel from rotations: E_f = sqrt|cos t| G
    Rz(phi) Ry(-t) Rz(-phi) (Ex, Ey, Ez_in), phi the polar angle of (X, Y)
    taken from the complex number X + iY, t = rho / f, Ez_in = (Ex X + Ey Y) /
    rho, G and sin(theta_max) as optical_elements.py:528-545 define them
    (the rows at :573-581 are this product written out)."""
    stm = radius / np.hypot(radius, f)
    n = len(x)
    out = np.zeros((3, n, n), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            X, Y = x[j], y[i]
            rho = abs(complex(X, Y))
            phi = np.angle(complex(X, Y))
            t = rho / f
            G = 1.0 / np.sqrt(abs(1 - (rho / radius) ** 2 * stm ** 2)) if rho < radius else 0.0
            e = np.array([Ex[i, j], Ey[i, j], (Ex[i, j] * X + Ey[i, j] * Y) / rho])
            out[:, i, j] = np.sqrt(abs(np.cos(t))) * G * (_rz(phi) @ _ry(-t) @ _rz(-phi) @ e)
    return out


def _hna_orientation_residual(n=12):
    x, y = R.space(1800.0, n)
    rng = np.random.default_rng(12)
    Ex = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Ey = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    got, _ = R.high_NA_objective_lens(Ex, Ey, x, y, 1800.0, 2000.0)
    ref = _hna_independent(Ex, Ey, x, y, 1800.0, 2000.0)
    # azimuthal polarisation (-sin phi, cos phi) has Ez_in = 0 and is left
    # unchanged by the map apart from the factor sqrt|cos t| G
    X, Y = np.meshgrid(x, y)
    rho, ph = np.hypot(X, Y), np.arctan2(Y, X)
    g = np.exp(-rho ** 2 / 900.0 ** 2)
    az, _ = R.high_NA_objective_lens(-np.sin(ph) * g + 0j, np.cos(ph) * g + 0j, x, y, 1800.0, 2000.0)
    G = np.where(rho < 1800.0, 1 / np.sqrt(np.abs(1 - (rho / 1800.0) ** 2 * (1800.0 / np.hypot(1800.0, 2000.0)) ** 2)),
                 0.0)
    w = np.sqrt(np.abs(np.cos(rho / 2000.0))) * G
    az_ref = np.stack([-np.sin(ph) * g * w, np.cos(ph) * g * w, np.zeros_like(g)])
    return max(rel_l2(got, ref), rel_l2(az, az_ref))


def test_high_na_phi_orientation(monkeypatch):
    """Pins the orientation of phi in the Richards-Wolf map (the high-NA
    focusing workload): the oracle equals the rotation product Rz(phi) Ry(-t)
    Rz(-phi) built per pixel with phi taken independently, and passes an
    azimuthally polarised beam through unchanged apart from sqrt|cos t| G.
    Orthogonality, det = 1 and the phi = 0 case cannot see the sign of
    sin(phi); the second half plants that mistake in the oracle (sin(phi)
    replaced by -sin(phi)) and shows that this check fails on it."""
    assert _hna_orientation_residual() < 1e-12
    orig = R.richards_wolf_matrix
    monkeypatch.setattr(R, "richards_wolf_matrix", lambda c_t, s_t, c_p, s_p: orig(c_t, s_t, c_p, -s_p))
    assert _hna_orientation_residual() > 0.05


def test_hna_vczt_equals_dense_form(golden):
    c = golden["hna_n32"]
    out = R.vczt_objective_lens(c["Ex"], c["Ey"], c["x"], c["y"], float(c["wavelength"]), float(c["radius"]),
                                float(c["f"]), c["xout"], c["yout"])
    assert rel_l2(out, c["out"]) < 1e-12


# -- sources --------------------------------------------------------------------------------

def test_sources_closed_form():
    x, y = R.space(1500.0, 64)
    u = R.gaussian_beam(x, y, 0.633, (400.0, 400.0), 2.0)
    X, Y = np.meshgrid(x, y)
    assert np.allclose(np.abs(u), 2.0 * np.exp(-(X ** 2 + Y ** 2) / 400.0 ** 2))
    Ex, Ey, Ez = R.polarized_gaussian_beam(x, y, 0.633, (400.0, 400.0), (3, 4))
    assert np.allclose(np.abs(Ex) / np.abs(Ey), 0.75) and np.all(Ez == 0)
    pw = R.plane_wave(x, y, 0.633, 1.0, 0.0, 0.0, 0.0)
    assert np.allclose(pw, 1.0)


# -- golden vectors ---------------------------------------------------------------------------

def test_golden_vectors_reproduce(golden):
    for name, c in golden.items():
        if name.startswith("rs_"):
            out, qf = R.rs_propagation(c["field"], c["x"], c["y"], float(c["wavelength"]), float(c["z"]))
            assert rel_l2(out, c["out"]) < 1e-12 and abs(qf - float(c["qf"])) < 1e-12
            assert rel_l2(out, B.rs_direct(c["field"], c["x"], c["y"], float(c["wavelength"]), float(c["z"]))) < 1e-11
        elif name.startswith("czt_"):
            out = R.czt(c["field"], c["x"], c["y"], float(c["wavelength"]), float(c["z"]), c["xout"], c["yout"])
            assert rel_l2(out, c["out"]) < 1e-12
            d = B.czt2d_direct(c["field"], c["x"], c["y"], float(c["wavelength"]), float(c["z"]), c["xout"], c["yout"])
            assert rel_l2(out, d) < 1e-11
        elif name.startswith("vczt_"):
            out = np.stack(R.vczt(c["Ex"], c["Ey"], c["x"], c["y"], float(c["wavelength"]), float(c["z"]), c["xout"],
                                  c["yout"]))
            assert rel_l2(out, c["out"]) < 1e-12
        elif name.startswith("vrs_"):
            mx, my, _ = R.LCD(c["Ex"], c["Ey"], None, c["eta"].astype(np.float64), c["theta"].astype(np.float64))
            (px, py, pz), _ = R.vrs_propagation(mx, my, c["x"], c["y"], float(c["wavelength"]), float(c["z"]))
            assert rel_l2(np.stack([px, py, pz]), c["out"]) < 1e-12


def test_rs_kernel_z_derivative_vs_central_difference():
    """dH/dz (the derivative the reference's jax.grad takes through
    wave_optics.py:291-297 when z is a parameter) against a central
    difference of H in z, both signs of z."""
    X = np.array([0.0, 13.7, -250.0, 1200.0, 3.3])
    Y = np.array([5.0, -40.0, 300.0, 0.1, -700.0])
    k = 2 * np.pi / 0.65
    eps = 1e-4
    for z in (3e4, 731.0, 50.0, -500.0):
        d = R.transfer_function_RS_dz(z, X, Y, k)
        fd = (R.transfer_function_RS(z + eps, X, Y, k) - R.transfer_function_RS(z - eps, X, Y, k)) / (2 * eps)
        assert np.max(np.abs(fd - d) / np.abs(d)) < 1e-6


def test_rs_propagation_z_derivative_is_linear_in_h():
    """The propagation derivative is the convolution with dH/dz: it equals a
    central difference of the propagated field in z."""
    n = 24
    x, y = R.space(60.0, n)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z, eps = 900.0, 1e-4
    d = R.rs_propagation_dz(u, x, y, 0.65, z)
    fd = (R.rs_propagation(u, x, y, 0.65, z + eps)[0] - R.rs_propagation(u, x, y, 0.65, z - eps)[0]) / (2 * eps)
    assert np.linalg.norm(fd - d) / np.linalg.norm(d) < 1e-6


def test_amplitude_sslm_reduces_to_sslm_and_scales():
    rng = np.random.default_rng(30)
    Ex, Ey = rng.standard_normal((2, 6, 6)) + 1j * rng.standard_normal((2, 6, 6))
    a, p = rng.uniform(-np.pi, np.pi, (2, 6, 6))
    one = np.ones((6, 6))
    s1 = R.sSLM(Ex, Ey, None, a, p)
    s2 = R.sSLM_with_amplitude(Ex, Ey, None, a, p, one, one)
    assert np.allclose(s1[0], s2[0]) and np.allclose(s1[1], s2[1])
    a1, a2 = rng.uniform(0, 1, (2, 6, 6))
    s3 = R.sSLM_with_amplitude(Ex, Ey, None, a, p, a1, a2)
    assert np.allclose(np.abs(s3[0]), a1 * np.abs(Ex)) and np.allclose(np.abs(s3[1]), a2 * np.abs(Ey))


def test_lens_focuses_a_plane_wave():
    """A plane wave through lens(f) and RS-propagated by f concentrates at the
    axis: the on-axis intensity is many times the input's, and it is the
    maximum of the plane (a sign error in the lens phase would diverge)."""
    n = 256
    x, y = R.space(400.0, n)
    f = 20000.0
    m = R.lens_mask(x, y, 0.65, (350.0, 350.0), (f, f))
    assert abs(abs(m[n // 2, n // 2]) - 1.0) < 1e-12 and np.all(np.abs(m) <= 1.0 + 1e-12)
    out, _ = R.rs_propagation(m, x, y, 0.65, f)
    I = np.abs(out) ** 2
    c = I[n // 2 - 1:n // 2 + 1, n // 2 - 1:n // 2 + 1].max()
    assert c == I.max() and c > 50.0
    diverging = R.rs_propagation(np.conj(m), x, y, 0.65, f)[0]
    assert np.abs(diverging).max() ** 2 < 2.0


def test_plane_wave_phase_gradient():
    """The oracle's tilted plane wave has the closed-form phase step
    k sin(theta) (cos(phi) dx, sin(phi) dy) between neighbouring pixels and the
    constant phase k z0 cos(theta) at the origin (wave_optics.py:525-542)."""
    x, y = R.space(300.0, 33)
    wl, th, ph, z0 = 0.633, 0.03, 0.8, 5.0
    k = 2 * np.pi / wl
    pw = R.plane_wave(x, y, wl, 1.5, th, ph, z0)
    dx, dy = x[1] - x[0], y[1] - y[0]
    assert np.allclose(pw[:, 1:] / pw[:, :-1], np.exp(1j * k * np.sin(th) * np.cos(ph) * dx))
    assert np.allclose(pw[1:, :] / pw[:-1, :], np.exp(1j * k * np.sin(th) * np.sin(ph) * dy))
    assert np.allclose(pw[16, 16], 1.5 * np.exp(1j * k * z0 * np.cos(th)))     # x = y = 0 at the centre
    ex, ey, ez = R.polarized_plane_wave(x, y, wl, (3, 4), th, ph, z0)
    assert np.allclose(ex, 0.6 * pw / 1.5) and np.allclose(ey, 0.8 * pw / 1.5) and not ez.any()
