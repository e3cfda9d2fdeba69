"""Float64 NumPy restatement of the XLuminA propagation path (TEST ORACLE).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``; never by the product package.

Every function follows one reference function literally (same padding, same
FFT lengths, same slicing, same quirks) and cites it as
``<file>:<lines>`` relative to the reference repository root.  The arrays are
NumPy complex128 by default, which is what the reference computes in
(``xlumina/wave_optics.py:10-13`` forces ``jax_enable_x64``).  ``dtype`` lets
the CPU baseline run the same algorithm in complex64.

FFTs use ``scipy.fft`` (pocketfft) in place of ``jnp.fft`` (XLA/ducc in
jaxlib 0.4.33); both are exact-to-rounding DFTs of the same lengths.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as sfft

# units -- xlumina/__init__.py:15-20
um = 1.0
nm = 1e-3
mm = 1e3
cm = 1e4
radians = 1.0
degrees = 180.0 / np.pi

_WORKERS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1


def set_workers(n: int) -> None:
    """Thread count for scipy.fft inside the oracle (CPU-baseline knob)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def _fft(a, n=None, axis=-1):
    return sfft.fft(a, n=n, axis=axis, workers=_WORKERS)


def _ifft(a, n=None, axis=-1):
    return sfft.ifft(a, n=n, axis=axis, workers=_WORKERS)


def _fft2(a):
    return sfft.fft2(a, workers=_WORKERS)


def _ifft2(a):
    return sfft.ifft2(a, workers=_WORKERS)


# ----------------------------------------------------------------------------
# grids -- toolbox.py:49-61, wave_optics.py:265-279, :299-331
# ----------------------------------------------------------------------------

def space(x_total, num_pix):
    """toolbox.py:49-61: symmetric linspace, float64."""
    x = np.linspace(-x_total, x_total, num_pix, dtype=np.float64)
    y = np.linspace(-x_total, x_total, num_pix, dtype=np.float64)
    return x, y


def build_grid(x, y):
    """wave_optics.py:265-279: (2N-1)^2 lag grid, monotone lags -(N-1)d..(N-1)d.

    The sign vector ``III`` is built from ``x`` and applied to both axes, as in
    the reference (square grids only).
    """
    nx = len(x)
    ny = len(y)
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    x_padded = np.pad((x[0] - x[::-1]), (0, np.size(x) - 1), "reflect")
    y_padded = np.pad((y[0] - y[::-1]), (0, np.size(y) - 1), "reflect")
    I = np.ones((1, int(len(x_padded) / 2) + 1))
    II = -np.ones((1, int(len(x_padded) / 2)))
    III = np.ravel(np.concatenate((I, II), 1))
    Xext, Yext = np.meshgrid(x_padded * III, y_padded * III)
    return nx, ny, dx, dy, Xext, Yext


def transfer_function_RS(z, Xext, Yext, k):
    """wave_optics.py:291-297: sampled RS impulse response.

    H = (1/2pi) z/r^2 (1/r - i k) exp(+i k r) for z > 0, exp(-i k r) otherwise
    (the factor is not conjugated for z <= 0).
    """
    r = np.sqrt(Xext ** 2 + Yext ** 2 + z ** 2)
    factor = 1 / (2 * np.pi) * z / r ** 2 * (1 / r - 1j * k)
    if z > 0:
        return np.exp(1j * k * r) * factor
    return np.exp(-1j * k * r) * factor


def transfer_function_RS_dz(z, Xext, Yext, k):
    """d/dz of transfer_function_RS (the derivative JAX takes through
    wave_optics.py:291-297 when z is a parameter), from the product rule with
    dr/dz = z/r and s = +1 (z > 0) or -1:
        dH/dz = (1/2pi) e^{i s k r} [ (r^2 - 2 z^2)/r^4 (1/r - i k) - z^2/r^5
                                      + s (z^2/r^3)(1/r - i k)(i k) ].
    """
    r = np.sqrt(Xext ** 2 + Yext ** 2 + z ** 2)
    s = 1.0 if z > 0 else -1.0
    b = 1 / r - 1j * k
    coef = (r ** 2 - 2 * z ** 2) / r ** 4 * b - z ** 2 / r ** 5 + s * (z ** 2 / r ** 3) * b * (1j * k)
    return np.exp(1j * s * k * r) * coef / (2 * np.pi)


def rs_propagation_dz(field, x, y, wavelength, z, dtype=np.complex128):
    """d/dz of the RS propagation of a fixed field: the same zero-padded
    convolution (wave_optics.py:281-289) with dH/dz in place of H."""
    nx, ny, dx, dy, Xext, Yext = build_grid(np.asarray(x), np.asarray(y))
    k = 2 * np.pi / wavelength
    H = transfer_function_RS_dz(z, Xext, Yext, k).astype(dtype)
    U = np.zeros((2 * ny - 1, 2 * nx - 1), dtype=dtype)
    U[0:ny, 0:nx] = np.asarray(field, dtype=dtype)
    out = _ifft2(_fft2(U) * _fft2(H)) * dx * dy
    return out[ny - 1:, nx - 1:]


def RS_propagation_jit(input_field, z, nx, ny, dx, dy, Xext, Yext, k, dtype=np.complex128):
    """wave_optics.py:281-289: zero-pad to (2ny-1, 2nx-1), fft2 product, slice."""
    H = transfer_function_RS(z, Xext, Yext, k).astype(dtype)
    U = np.zeros((2 * ny - 1, 2 * nx - 1), dtype=dtype)
    U[0:ny, 0:nx] = input_field
    out = _ifft2(_fft2(U) * _fft2(H)) * dx * dy
    return out[ny - 1:, nx - 1:].astype(dtype)


def quality_factor(x, y, wavelength, z):
    """wave_optics.py:187-191 (and vectorized_optics.py:264-270)."""
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    dr_real = np.sqrt(dx ** 2 + dy ** 2)
    rmax = np.sqrt((x ** 2).max() + (y ** 2).max())
    dr_ideal = np.sqrt(wavelength ** 2 + rmax ** 2 + 2 * wavelength * np.sqrt(rmax ** 2 + z ** 2)) - rmax
    return dr_ideal / dr_real


def rs_propagation(field, x, y, wavelength, z, dtype=np.complex128):
    """ScalarLight.RS_propagation, wave_optics.py:173-196 -> (field, qf)."""
    nx, ny, dx, dy, Xext, Yext = build_grid(x, y)
    k = 2 * np.pi / wavelength
    out = RS_propagation_jit(np.asarray(field, dtype=dtype), z, nx, ny, dx, dy, Xext, Yext, k, dtype)
    return out, quality_factor(x, y, wavelength, z)


def minimum_z(x, y, wavelength, n=1, quality_factor=1):
    """wave_optics.py:198-231 (identical in vectorized_optics.py:286-319).

    The reference prints the value and returns None; this returns it.
    """
    range_x = x[-1] - x[0]
    range_y = y[-1] - y[0]
    dx = range_x / np.size(x)
    dy = range_y / np.size(y)
    dr_real = np.sqrt(dx ** 2 + dy ** 2)
    rmax = np.sqrt(range_x ** 2 + range_y ** 2)
    factor = (((quality_factor * dr_real + rmax) ** 2 - (wavelength / n) ** 2 - rmax ** 2)
              / (2 * wavelength / n)) ** 2 - rmax ** 2
    return float(np.sqrt(factor)) if factor > 0 else 0.0


# ----------------------------------------------------------------------------
# chirp z-transform -- wave_optics.py:299-460
# ----------------------------------------------------------------------------

def build_CZT_grid(z, wavelength, xin, yin, xout, yout):
    """wave_optics.py:299-331."""
    nx = len(xout)
    ny = len(yout)
    Xout, Yout = np.meshgrid(xout, yout)
    dx = xin[1] - xin[0]
    dy = yin[1] - yin[0]
    Dm = wavelength * z / dx
    fy_1 = yout[0] + Dm / 2
    fy_2 = yout[-1] + Dm / 2
    fx_1 = xout[0] + Dm / 2
    fx_2 = xout[-1] + Dm / 2
    return nx, ny, dx, dy, Xout, Yout, Dm, fy_1, fy_2, fx_1, fx_2


def compute_np2(x):
    """wave_optics.py:373-383."""
    return int(2 ** int(np.ceil(np.log2(x))))


def check_czt_lengths(m, M_out):
    """Shapes the reference cannot handle (wave_optics.py:440-448).

    ``mp = m + M_out - 1`` a power of two makes the window one row short and
    the reference's multiply raises; ``M_out == 1`` broadcasts to (n, n).
    """
    mp = m + M_out - 1
    if M_out < 2:
        raise ValueError("CZT output length must be >= 2 (the reference mis-broadcasts M_out=1)")
    if mp & (mp - 1) == 0:
        raise ValueError(
            f"CZT with m + M_out - 1 = {mp} a power of two is rejected (the reference's window "
            "slice is one sample short and raises)")


def compute_fft(x, D1, D2, Dm, m, n, mp, M_out, np2, dtype=np.complex128):
    """wave_optics.py:385-410: chirp premultiply, FFT conv with 1/h, length np2.

    The chirp tables are formed in float64 and cast to ``dtype`` for the FFTs."""
    A = np.exp(1j * 2 * np.pi * D1 / Dm)
    W = np.exp(-1j * 2 * np.pi * (D2 - D1) / (M_out * Dm))
    h = np.arange(-m + 1, max(M_out - 1, m - 1) + 1)
    h = W ** (h ** 2 / 2)
    h_sliced = h[:mp + 1]
    ft = _fft((1 / h_sliced).astype(dtype), np2)
    b = (A ** (-(np.arange(m))) * h[np.arange(m - 1, 2 * m - 1)]).astype(dtype)
    tmp = np.tile(b, (n, 1)).T
    b = _fft(x.astype(dtype, copy=False) * tmp, np2, axis=0)
    b = _ifft(b * np.tile(ft, (n, 1)).T, axis=0)
    return b, h


def bluestein_method(x, f1, f2, Dm, M_out, dtype=np.complex128):
    """wave_optics.py:412-460: 1-D CZT along axis 0; returns the TRANSPOSE (n, M_out)."""
    m, n = x.shape
    check_czt_lengths(m, M_out)
    D1 = f1 + (M_out * Dm + f2 - f1) / (2 * M_out)
    D2 = f2 + (M_out * Dm + f2 - f1) / (2 * M_out)
    mp = m + M_out - 1
    np2 = compute_np2(mp)
    b, h = compute_fft(x, D1, D2, Dm, m, n, mp, M_out, np2, dtype)
    b = b[m:mp + 1, 0:n].T * np.tile(h[m - 1:mp].astype(dtype), (n, 1))
    l = np.linspace(0, M_out - 1, M_out)
    l = l / M_out * (D2 - D1) + D1
    M_shift = -m / 2
    M_shift = np.tile(np.exp(-1j * 2 * np.pi * l * (M_shift + 1 / 2) / Dm).astype(dtype), (n, 1))
    return b * M_shift


def CZT_jit(field, z, wavelength, k, nx, ny, dx, dy, Xout, Yout, X, Y, Dm, fy_1, fy_2, fx_1, fx_2,
            dtype=np.complex128):
    """wave_optics.py:333-357: F0 * Bx(By(E*F)) * z*dx*dy*lambda."""
    F0 = transfer_function_RS(z, Xout, Yout, k).astype(dtype)
    F = transfer_function_RS(z, X, Y, k).astype(dtype)
    field = field * F
    U = bluestein_method(field, fy_1, fy_2, Dm, ny, dtype)
    U = bluestein_method(U, fx_1, fx_2, Dm, nx, dtype)
    return F0 * U * (z * dx * dy * wavelength)


def CZT_for_high_NA_jit(field, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2):
    """wave_optics.py:359-371: two Bluestein passes, no transfer functions."""
    U = bluestein_method(field, fy_1, fy_2, Dm, ny)
    return bluestein_method(U, fx_1, fx_2, Dm, nx)


def czt(field, x, y, wavelength, z, xout=None, yout=None):
    """ScalarLight.CZT, wave_optics.py:233-263."""
    if xout is None:
        xout = x
    if yout is None:
        yout = y
    X, Y = np.meshgrid(x, y)
    k = 2 * np.pi / wavelength
    nx, ny, dx, dy, Xout, Yout, Dm, fy_1, fy_2, fx_1, fx_2 = build_CZT_grid(z, wavelength, x, y, xout, yout)
    return CZT_jit(np.asarray(field, dtype=np.complex128), z, wavelength, k, nx, ny, dx, dy, Xout, Yout,
                   X, Y, Dm, fy_1, fy_2, fx_1, fx_2)


# ----------------------------------------------------------------------------
# vectorial propagation -- vectorized_optics.py:244-394
# ----------------------------------------------------------------------------

def vrs_propagation(Ex, Ey, x, y, wavelength, z, dtype=np.complex128):
    """VectorizedLight.VRS_propagation, vectorized_optics.py:244-284.

    The incoming Ez is discarded: Ez_in = (Ex X + Ey Y)/r, r = sqrt(X^2+Y^2+z^2).
    Returns ((Ex, Ey, Ez) propagated, qf).
    """
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + z ** 2)
    Ez = Ex * X / r + Ey * Y / r
    nx, ny, dx, dy, Xext, Yext = build_grid(x, y)
    k = 2 * np.pi / wavelength
    out = [RS_propagation_jit(np.asarray(c, dtype=dtype), z, nx, ny, dx, dy, Xext, Yext, k, dtype)
           for c in (Ex, Ey, Ez)]
    return tuple(out), quality_factor(x, y, wavelength, z)


def vczt(Ex, Ey, x, y, wavelength, z, xout=None, yout=None, dtype=np.complex128):
    """VectorizedLight.VCZT, vectorized_optics.py:321-361.

    Ez_in = (Ex X/r + Ey Y/r) z/r (differs from VRS by z/r); incoming Ez discarded.
    """
    if xout is None:
        xout = x
    if yout is None:
        yout = y
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + z ** 2)
    Ez = (Ex * X / r + Ey * Y / r) * z / r
    k = 2 * np.pi / wavelength
    nx, ny, dx, dy, Xout, Yout, Dm, fy_1, fy_2, fx_1, fx_2 = build_CZT_grid(z, wavelength, x, y, xout, yout)
    return tuple(CZT_jit(np.asarray(c, dtype=dtype), z, wavelength, k, nx, ny, dx, dy, Xout, Yout,
                         X, Y, Dm, fy_1, fy_2, fx_1, fx_2, dtype) for c in (Ex, Ey, Ez))


# ----------------------------------------------------------------------------
# Jones masks -- optical_elements.py:76-180, 186-332; toolbox.py:121-135
# ----------------------------------------------------------------------------

def slm(field, phase):
    """SLM, optical_elements.py:87-103 -> (field * e^{i phase}, mask)."""
    mask = np.exp(1j * np.asarray(phase, dtype=np.float64))
    return field * mask, mask


def jones_LP(alpha):
    """optical_elements.py:109-120."""
    c, s = np.cos(alpha), np.sin(alpha)
    return np.array([[c ** 2, c * s], [c * s, s ** 2]])


def jones_general_retarder(eta, theta, delta):
    """optical_elements.py:122-139."""
    em = np.exp(-(eta / 2) * 1j)
    ep = np.exp((eta / 2) * 1j)
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[em * c ** 2 + ep * s ** 2, (em - ep) * np.exp(-delta * 1j) * s * c],
                     [(em - ep) * np.exp(delta * 1j) * s * c, em * s ** 2 + ep * c ** 2]])


def jones_sSLM(alpha, phi):
    """optical_elements.py:141-152."""
    z = np.zeros_like(np.asarray(alpha, dtype=np.complex128))
    return np.array([[np.exp(1j * alpha), z], [z, np.exp(1j * phi)]])


def jones_LCD(eta, theta):
    """optical_elements.py:169-180 (general retarder with delta = 0)."""
    return jones_general_retarder(eta, theta, 0.0)


def _apply_jones(J, Ex, Ey):
    # J is (2, 2, ...) broadcast against the (N, N) fields; the reference does the
    # same per-pixel 2x2 product through reshape((N^2, 2, 2)) @ (N^2, 2, 1).
    return J[0, 0] * Ex + J[0, 1] * Ey, J[1, 0] * Ex + J[1, 1] * Ey


def sSLM(Ex, Ey, Ez, alpha, phi):
    """optical_elements.py:186-222: Ez kept."""
    ox, oy = _apply_jones(jones_sSLM(alpha, phi), Ex, Ey)
    return ox, oy, Ez


def sSLM_with_amplitude(Ex, Ey, Ez, alpha, phi, a1, a2):
    """optical_elements.py:155-168, 224-264: diag(a1 e^{i alpha}, a2 e^{i phi}), Ez kept."""
    return a1 * np.exp(1j * alpha) * Ex, a2 * np.exp(1j * phi) * Ey, Ez


def lens_mask(x, y, wavelength, radius, focal):
    """optical_elements.py:678-703 with circular_mask :781-795."""
    X, Y = np.meshgrid(x, y)
    rx, ry = radius
    fx, fy = focal
    k = 2 * np.pi / wavelength
    pupil = np.where((X ** 2 / rx ** 2 + Y ** 2 / ry ** 2) < 1, 1, 0)
    return pupil * np.exp(-1j * k * (X ** 2 / (2 * fx) + Y ** 2 / (2 * fy)))


def rotate_mask(X, Y, angle, origin=None):
    """toolbox.py:269-289."""
    if origin is None:
        x0, y0 = (X[-1] + X[0]) / 2, (Y[-1] + Y[0]) / 2
    else:
        x0, y0 = origin
    Xrot = (X - x0) * np.cos(angle) + (Y - y0) * np.sin(angle)
    Yrot = -(X - x0) * np.sin(angle) + (Y - y0) * np.cos(angle)
    return Xrot, Yrot


def cylindrical_lens_mask(x, y, wavelength, focal_length, refractive_index=1.5, angle=0.0):
    """optical_elements.py:705-744."""
    X, Y = np.meshgrid(x, y)
    k = 2 * np.pi / wavelength
    Xrot = X * np.cos(angle) + Y * np.sin(angle)
    R = focal_length * (refractive_index - 1)
    with np.errstate(invalid="ignore"):
        thickness = R - np.sqrt(R ** 2 - Xrot ** 2)
    phase = k * (refractive_index - 1) * (Xrot ** 2 / (2 * focal_length) + thickness)
    return np.exp(-1j * phase)


def axicon_mask(x, y, wavelength, alpha, n=1.5):
    """optical_elements.py:746-779."""
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2)
    return np.exp(-1j * (2 * np.pi / wavelength) * r * (n - 1) * np.sin(alpha) * np.tan(alpha))


def circular_mask(X, Y, r):
    """optical_elements.py:781-795."""
    rx, ry = r
    return np.where((X ** 2 / rx ** 2 + Y ** 2 / ry ** 2) < 1, 1, 0)


def triangular_mask(X, Y, r, angle, m, height):
    """optical_elements.py:797-817."""
    x0, y0 = r
    Xrot, Yrot = rotate_mask(X, Y, angle * (np.pi / 180), r)
    Yedge = -m * np.abs(Xrot - x0) + y0
    return np.where((Yrot < Yedge) & (Yrot > y0 - height), 1, 0)


def rectangular_mask(X, Y, center, width, height, angle):
    """optical_elements.py:819-836."""
    Xrot, Yrot = rotate_mask(X, Y, angle * (np.pi / 180), center)
    return np.where((Xrot < (width / 2)) & (Xrot > (-width / 2)) & (Yrot < (height / 2)) & (Yrot > (-height / 2)), 1, 0)


def annular_aperture(di, do, X, Y):
    """optical_elements.py:838-854."""
    di, do = di / 2, do / 2
    stop = np.where(((X ** 2 + Y ** 2) / di ** 2) < 1, 0, 1)
    ring = np.where(((X ** 2 + Y ** 2) / do ** 2) < 1, 1, 0)
    return stop * ring


def forked_grating(X, Y, center, angle, period, l, kind=""):
    """optical_elements.py:856-892 (None for a kind other than 'Amplitude'/'Phase')."""
    Xrot, Yrot = rotate_mask(X, Y, angle * (np.pi / 180), center)
    theta = np.arctan2(Xrot, Yrot)
    g = np.angle(np.exp(1j * np.cos(l * theta - 2 * np.pi / period * Xrot)))
    binary = np.where(g < 0, 0, 1)
    if kind == "Amplitude":
        return binary
    if kind == "Phase":
        return np.exp(1j * np.pi * binary)
    return None


def build_LCD_cell(eta, theta, shape):
    """toolbox.py:121-135: eta*ones -- scalars or per-pixel arrays both work."""
    return eta * np.ones((shape, shape)), theta * np.ones((shape, shape))


def LCD(Ex, Ey, Ez, eta, theta):
    """optical_elements.py:266-305: Ez kept."""
    eta_a, theta_a = build_LCD_cell(eta, theta, Ex.shape[0])
    ox, oy = _apply_jones(jones_LCD(eta_a, theta_a), Ex, Ey)
    return ox, oy, Ez


def linear_polarizer(Ex, Ey, Ez, alpha):
    """optical_elements.py:307-332: a fresh VectorizedLight, so Ez becomes 0."""
    ox, oy = _apply_jones(jones_LP(np.asarray(alpha, dtype=np.float64)), Ex, Ey)
    return ox, oy, np.zeros_like(ox)


# ----------------------------------------------------------------------------
# high-NA objective lens -- optical_elements.py:515-672
# ----------------------------------------------------------------------------

def high_NA_objective_lens(Ex, Ey, x, y, radius, f):
    """optical_elements.py:515-553 + _high_NA_objective_lens_ :555-594.

    Returns ((3, N, N) field after the lens, sin_theta_max).
    """
    sin_theta_max = radius / np.sqrt(radius ** 2 + f ** 2)
    X, Y = np.meshgrid(x, y)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.sqrt(X ** 2 + Y ** 2)
        phi = np.arctan2(Y, X)
        theta = r / f
        Ez = Ex * X / r + Ey * Y / r
        u = X / radius
        v = Y / radius
        pupil_mask = np.where((X ** 2 + Y ** 2) / radius ** 2 < 1, 1, 0)
        G = np.real(pupil_mask * (1 / np.sqrt(np.abs(1 - (u ** 2 + v ** 2) * sin_theta_max ** 2))))
    c_t, s_t = np.cos(theta), np.sin(theta)
    c_p, s_p = np.cos(phi), np.sin(phi)
    apod = np.sqrt(np.abs(c_t))
    RL = richards_wolf_matrix(c_t, s_t, c_p, s_p)
    E0 = [RL[i][0] * Ex + RL[i][1] * Ey + RL[i][2] * Ez for i in range(3)]
    return np.stack([apod * G * e for e in E0], axis=0), sin_theta_max


def richards_wolf_matrix(c_t, s_t, c_p, s_p):
    """optical_elements.py:573-581: rows of the 3x3 lens map."""
    return [[c_t * c_p ** 2 + s_p ** 2, c_t * c_p * s_p - s_p * c_p, -c_p * s_t],
            [s_p * c_t * c_p - c_p * s_p, c_t * s_p ** 2 + c_p ** 2, -s_p * s_t],
            [s_t * c_p, s_t * s_p, c_t]]


def build_high_NA_VCZT_grid(f, r, wavelength, xin, xout, yout):
    """optical_elements.py:640-672."""
    nx = len(xout)
    ny = len(yout)
    Din = len(xin)
    Dm = f * wavelength * (Din - 1) / (2 * r)
    return nx, ny, Dm, yout[0] + Dm / 2, yout[-1] + Dm / 2, xout[0] + Dm / 2, xout[-1] + Dm / 2


def vczt_objective_lens(Ex, Ey, x, y, wavelength, r, f, xout, yout):
    """VCZT_objective_lens, optical_elements.py:600-638 -> (3, My, Mx)."""
    field_in_lens, stm = high_NA_objective_lens(Ex, Ey, x, y, r, f)
    nx, ny, Dm, fy_1, fy_2, fx_1, fx_2 = build_high_NA_VCZT_grid(f, r, wavelength, x, xout, yout)
    U = np.stack([CZT_for_high_NA_jit(c, nx, ny, Dm, fy_1, fy_2, fx_1, fx_2) for c in field_in_lens])
    cte = -(1j * stm ** 2 / (f * wavelength))
    return cte * U


# ----------------------------------------------------------------------------
# sources -- wave_optics.py:468-542, vectorized_optics.py:402-490
# ----------------------------------------------------------------------------

def _gaussian_envelope(x, y, wavelength, w0, center, z_w0, alpha, n=1):
    k = 2 * np.pi / wavelength
    w0_x, w0_y = w0
    x0, y0 = center
    z_w0x, z_w0y = z_w0
    Rayleigh_x = k * w0_x ** 2 * n / 2
    Rayleigh_y = k * w0_y ** 2 * n / 2
    Gouy_phase_x = np.arctan2(z_w0x, Rayleigh_x)
    Gouy_phase_y = np.arctan2(z_w0y, Rayleigh_y)
    w_x = w0_x * np.sqrt(1 + (z_w0x / Rayleigh_x) ** 2)
    w_y = w0_y * np.sqrt(1 + (z_w0y / Rayleigh_y) ** 2)
    R_x = 1e12 if z_w0x == 0 else z_w0x * (1 + (Rayleigh_x / z_w0x) ** 2)
    # the reference tests z_w0x for R_y as well (wave_optics.py:507)
    R_y = 1e12 if z_w0x == 0 else z_w0y * (1 + (Rayleigh_y / z_w0y) ** 2)
    X, Y = np.meshgrid(x, y)
    x_rot = X * np.cos(alpha) + Y * np.sin(alpha)
    y_rot = -X * np.sin(alpha) + Y * np.cos(alpha)
    phase = np.exp(-1j * ((k * z_w0x + k * X ** 2 / (2 * R_x) - Gouy_phase_x)
                          + (k * z_w0y + k * Y ** 2 / (2 * R_y) - Gouy_phase_y)))
    amp = (w0_x / w_x) * (w0_y / w_y) * np.exp(-(x_rot - x0) ** 2 / (w_x ** 2) - (y_rot - y0) ** 2 / (w_y ** 2))
    return amp, phase


def gaussian_beam(x, y, wavelength, w0, E0, center=(0, 0), z_w0=(0, 0), alpha=0):
    """LightSource.gaussian_beam, wave_optics.py:468-523."""
    amp, phase = _gaussian_envelope(x, y, wavelength, w0, center, z_w0, alpha)
    return (E0 * amp) * phase


def plane_wave(x, y, wavelength, A=1, theta=0, phi=0, z0=0):
    """LightSource.plane_wave, wave_optics.py:525-542."""
    k = 2 * np.pi / wavelength
    X, Y = np.meshgrid(x, y)
    return A * np.exp(1j * k * (X * np.sin(theta) * np.cos(phi) + Y * np.sin(theta) * np.sin(phi)
                                + z0 * np.cos(theta)))


def polarized_gaussian_beam(x, y, wavelength, w0, jones_vector, center=(0, 0), z_w0=(0, 0), alpha=0):
    """PolarizedLightSource.gaussian_beam, vectorized_optics.py:402-465 -> (Ex, Ey, Ez)."""
    amp, phase = _gaussian_envelope(x, y, wavelength, w0, center, z_w0, alpha)
    nj = np.array(jones_vector) / np.linalg.norm(np.array(jones_vector))
    Ex = (nj[0] * amp) * phase
    Ey = (nj[1] * amp) * phase
    return Ex, Ey, np.zeros((len(x), len(x)))


def polarized_plane_wave(x, y, wavelength, jones_vector, theta=0, phi=0, z0=0):
    """PolarizedLightSource.plane_wave, vectorized_optics.py:467-490."""
    nj = np.array(jones_vector) / np.linalg.norm(np.array(jones_vector))
    pw = plane_wave(x, y, wavelength, 1, theta, phi, z0)
    return nj[0] * pw, nj[1] * pw, np.zeros((len(x), len(x)))


def radial_beam(x, y, w0):
    """Radially polarised Gaussian used by the 2048^2 VRS+VCZT workload.

    Not in the reference (which builds radial polarisation through a quadrant
    sSLM): Ex = G X/rho, Ey = G Y/rho, G = exp(-rho^2/w0^2).
    """
    X, Y = np.meshgrid(x, y)
    rho = np.sqrt(X ** 2 + Y ** 2)
    G = np.exp(-rho ** 2 / w0 ** 2)
    return (G * X / rho).astype(np.complex128), (G * Y / rho).astype(np.complex128)


def is_conserving_energy(src, prop):
    """toolbox.py:74-96 on tuples of component arrays."""
    return sum(np.sum(np.abs(c) ** 2) for c in prop) / sum(np.sum(np.abs(c) ** 2) for c in src)


def _bs_RT(theta):
    """optical_elements.py:376-383: T = |cos|, R = |sin|, both less 1% of T."""
    T = np.abs(np.cos(theta))
    R = np.abs(np.sin(theta))
    noise = T * 0.01
    return R - noise, T - noise


def BS_symmetric(a, b, theta):
    """optical_elements.py:334-392: a, b are (Ex, Ey) pairs; returns the
    (Ex, Ey, Ez=0) triples of ports c and d."""
    R, T = _bs_RT(theta)
    (ax, ay), (bx, by) = a, b
    z = np.zeros_like(ax)
    c = (R * ax + 1j * T * bx, R * ay + 1j * T * by, z)
    d = (1j * T * ax + R * bx, 1j * T * ay + R * by, z)
    return c, d


def BS_symmetric_SI(a, theta, port=0):
    """optical_elements.py:394-460 (bs_port0 / bs_port1)."""
    R, T = _bs_RT(theta)
    ax, ay = a
    z = np.zeros_like(ax)
    if port == 0:
        return (R * ax, R * ay, z), (1j * T * ax, 1j * T * ay, z)
    return (1j * T * ax, 1j * T * ay, z), (R * ax, R * ay, z)


def BS(a1, a2, R, T, phase):
    """optical_elements.py:462-513: lossless splitter with a phase."""
    (ax, ay), (bx, by) = a1, a2
    z = np.zeros_like(ax)
    st, sr = np.sqrt(T), np.sqrt(R)
    b1 = (st * ax + sr * np.exp(1j * phase) * bx, st * ay + sr * np.exp(1j * phase) * by, z)
    b2 = (-sr * np.exp(-1j * phase) * ax + st * bx, -sr * np.exp(-1j * phase) * ay + st * by, z)
    return b1, b2


def fluorescence(i_ex, i_dep, beta=1):
    """optical_elements.py:941-955."""
    eps = 1e-8
    return i_ex * (1 - beta * (1 - np.exp(-(i_dep / (i_ex + eps)))))
