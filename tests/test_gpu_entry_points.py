"""The reference's module-level propagation kernels called with their
positional signatures, as a drop-in caller would (wave_optics.py:281-289,
333-357, 359-371; vectorized_optics.py:364-394), and the plane-wave sources
(wave_optics.py:525-542, vectorized_optics.py:467-490), each against the
float64 oracle called with the same arguments (GPU)."""
import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import vectorized_optics as V
from xlumina_amd import wave_optics as W
from xlumina_amd.vectorized_optics import PolarizedLightSource
from xlumina_amd.wave_optics import LightSource

pytestmark = pytest.mark.gpu
TOL = 1e-5
WL = 0.633


def H(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def cfield(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def grid():
    return R.space(400.0, 128)


def test_rs_propagation_jit_positional(grid, cuda_dev):
    x, y = grid
    u = R.gaussian_beam(x, y, WL, (150.0, 120.0), 1.0) * np.exp(1j * np.random.default_rng(1).uniform(-1, 1, (128, 128)))
    nx, ny, dx, dy, Xext, Yext = W.build_grid(x, y)
    k = 2 * np.pi / WL
    got = W.RS_propagation_jit(u, 2500.0, nx, ny, dx, dy, Xext, Yext, k)
    ref = R.RS_propagation_jit(u, 2500.0, *R.build_grid(x, y), k)
    assert tuple(got.shape) == (128, 128)
    assert rel_l2(H(got), ref) < TOL


def test_vrs_propagation_jit_three_given_components(grid, cuda_dev):
    """vmap of RS_propagation_jit over the 3 components: the given Ez is
    propagated as it is (no Ez construction on this path)."""
    x, y = grid
    rng = np.random.default_rng(2)
    E = np.stack([cfield(rng, 128, 128) * R.gaussian_beam(x, y, WL, (150.0, 150.0), 1.0) for _ in range(3)])
    nx, ny, dx, dy, Xext, Yext = W.build_grid(x, y)
    k = 2 * np.pi / WL
    got = V.VRS_propagation_jit(torch.from_numpy(E.astype(np.complex64)), 1800.0, nx, ny, dx, dy, Xext, Yext, k)
    ref = np.stack([R.RS_propagation_jit(E[c], 1800.0, *R.build_grid(x, y), k) for c in range(3)])
    assert tuple(got.shape) == (3, 128, 128)
    assert rel_l2(H(got), ref) < TOL


def _czt_args(x, y, xout, yout, z):
    nx, ny, dx, dy, Xout, Yout, Dm, fy_1, fy_2, fx_1, fx_2 = W.build_CZT_grid(z, WL, x, y, xout, yout)
    X, Y = np.meshgrid(x, y)
    return (z, WL, 2 * np.pi / WL, nx, ny, dx, dy, Xout, Yout, X, Y, Dm, fy_1, fy_2, fx_1, fx_2)


def test_czt_jit_positional(grid, cuda_dev):
    x, y = grid
    xo, yo = np.linspace(-60.0, 60.0, 40), np.linspace(-30.0, 50.0, 48)    # rectangular window
    u = R.gaussian_beam(x, y, WL, (150.0, 150.0), 1.0) * np.exp(1j * np.random.default_rng(3).uniform(-1, 1, (128, 128)))
    args = _czt_args(x, y, xo, yo, 3000.0)
    got = W.CZT_jit(u, *args)
    ref = R.CZT_jit(u, *args)
    assert tuple(got.shape) == (48, 40)
    assert rel_l2(H(got), ref) < TOL


def test_vczt_jit_three_given_components(grid, cuda_dev):
    """vmap of CZT_jit: the 3 given components, Ez included, with no Ez construction."""
    x, y = grid
    rng = np.random.default_rng(4)
    E = np.stack([cfield(rng, 128, 128) for _ in range(3)])
    xo, yo = np.linspace(-60.0, 60.0, 40), np.linspace(-30.0, 50.0, 48)
    args = _czt_args(x, y, xo, yo, 3000.0)
    got = V.VCZT_jit(torch.from_numpy(E.astype(np.complex64)), *args)
    ref = np.stack([R.CZT_jit(E[c], *args) for c in range(3)])
    assert tuple(got.shape) == (3, 48, 40)
    assert rel_l2(H(got), ref) < TOL


def _hna_args(x, xo, yo, wavelength=0.65, radius=400.0, f=400.0 / 0.9):
    return R.build_high_NA_VCZT_grid(f, radius, wavelength, x, xo, yo)


def test_czt_for_high_na_jit_positional(grid, cuda_dev):
    """The two Bluestein passes alone (XL_CZT_PLAIN forward)."""
    x, _ = grid
    xo, yo = np.linspace(-2.0, 2.0, 40), np.linspace(-1.5, 2.5, 48)
    nx, ny, Dm, fy_1, fy_2, fx_1, fx_2 = _hna_args(x, xo, yo)
    u = cfield(np.random.default_rng(5), 128, 128)
    got = W.CZT_for_high_NA_jit(u, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2)
    ref = R.CZT_for_high_NA_jit(u, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2)
    assert tuple(got.shape) == (48, 40)
    assert rel_l2(H(got), ref) < TOL


def test_vectorized_czt_for_high_na_positional(grid, cuda_dev):
    x, _ = grid
    xo, yo = np.linspace(-2.0, 2.0, 40), np.linspace(-1.5, 2.5, 48)
    nx, ny, Dm, fy_1, fy_2, fx_1, fx_2 = _hna_args(x, xo, yo)
    rng = np.random.default_rng(6)
    E = np.stack([cfield(rng, 128, 128) for _ in range(3)])
    got = V.vectorized_CZT_for_high_NA(torch.from_numpy(E.astype(np.complex64)), nx, ny, Dm, fy_1, fy_2, fx_1, fx_2)
    ref = np.stack([R.CZT_for_high_NA_jit(E[c], nx, ny, Dm, fy_1, fy_2, fx_1, fx_2) for c in range(3)])
    assert tuple(got.shape) == (3, 48, 40)
    assert rel_l2(H(got), ref) < TOL


@pytest.mark.parametrize("theta,phi,z0", [(0.0, 0.0, 0.0), (0.01, 0.7, 3.0), (0.02, -2.1, 0.0)])
def test_plane_wave_sources(grid, cuda_dev, theta, phi, z0):
    x, y = grid
    s = LightSource(x, y, WL)
    s.plane_wave(A=2.0, theta=theta, phi=phi, z0=z0)
    ref = R.plane_wave(x, y, WL, 2.0, theta, phi, z0)
    assert tuple(s.field.shape) == (128, 128)
    assert rel_l2(H(s.field), ref) < TOL
    p = PolarizedLightSource(x, y, WL)
    p.plane_wave(jones_vector=(1, -2), theta=theta, phi=phi, z0=z0)
    rx, ry, rz = R.polarized_plane_wave(x, y, WL, (1, -2), theta, phi, z0)
    assert rel_l2(H(p.Ex), rx) < TOL and rel_l2(H(p.Ey), ry) < TOL
    assert float(p.Ez.abs().max()) == 0.0
    # and the source feeds the propagation like the oracle's does
    prop, _ = s.RS_propagation(z=1000.0)
    rprop, _ = R.rs_propagation(ref, x, y, WL, 1000.0)
    assert rel_l2(H(prop.field), rprop) < TOL
