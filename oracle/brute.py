"""Independent brute-force forms used to pin the oracle (TEST ORACLE).

Test infrastructure only.  Nothing here reuses the FFT/Bluestein structure of
``oracle.xlumina_ref``; these are the direct sums that the fast algorithms
must equal, built from the mathematics:

* ``rs_kernel``: the first Rayleigh-Sommerfeld kernel written as
  -(1/2pi) d/dz [exp(i k R)/R] (checked separately against a numerical
  derivative of the Green's function).  For z <= 0 the reference keeps the
  same real factor and flips only the phase (wave_optics.py:296); that branch
  is reproduced as stated there.
* ``rs_direct``: U(x_p, y_q) = sum_ij U(x_i, y_j) h(x_p-x_i, y_q-y_j) dx dy,
  an O(N^4) sum (the operator wave_optics.py:281-289 evaluates by FFT).
* ``czt_axis_direct``: the chirp-z output of wave_optics.py:412-460 as a dense
  DFT at frequencies nu_{k+1} = D1 + (k+1) Delta with the output phase factors
  exp(-i theta (k+1/2)) and exp(-i 2 pi l_k (1/2 - m/2)/Dm); when M_out >= m the
  reference's chirp table lacks its last tap (h has m + max(M_out, m) - 1
  entries, so h[:mp+1] is one entry short), which removes the j=0 term from
  output M_out-1.
* ``czt_axis_scipy``: the same via ``scipy.signal.czt`` (library chirp-z).
"""
from __future__ import annotations

import numpy as np
import scipy.signal as ssig


def rs_kernel(dx, dy, z, k):
    R = np.sqrt(dx ** 2 + dy ** 2 + z ** 2)
    # -(1/2pi) dG/dz with G = exp(ikR)/R and dR/dz = z/R
    g_over = (1j * k - 1.0 / R) * (z / R) / R          # (dG/dz) / exp(ikR)
    if z > 0:
        return -(1.0 / (2 * np.pi)) * g_over * np.exp(1j * k * R)
    return -(1.0 / (2 * np.pi)) * g_over * np.exp(-1j * k * R)


def greens_dz_numeric(dx, dy, z, k, eps):
    """-(1/2pi) d/dz exp(ikR)/R by a central difference (for pinning rs_kernel)."""
    def G(zz):
        R = np.sqrt(dx ** 2 + dy ** 2 + zz ** 2)
        return np.exp(1j * k * R) / R
    return -(G(z + eps) - G(z - eps)) / (2 * eps) / (2 * np.pi)


def rs_direct(field, x, y, wavelength, z):
    """O(N^4) Rayleigh-Sommerfeld sum on the input grid."""
    k = 2 * np.pi / wavelength
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    out = np.zeros((len(y), len(x)), dtype=np.complex128)
    for q in range(len(y)):
        for p in range(len(x)):
            ker = rs_kernel(x[p] - x[None, :], y[q] - y[:, None], z, k)
            out[q, p] = np.sum(field * ker) * dx * dy
    return out


def _czt_params(f1, f2, Dm, M):
    D1 = f1 + (M * Dm + f2 - f1) / (2 * M)
    D2 = f2 + (M * Dm + f2 - f1) / (2 * M)
    delta = (D2 - D1) / M
    theta = np.angle(np.exp(-1j * 2 * np.pi * (D2 - D1) / (M * Dm)))
    return D1, D2, delta, theta


def czt_matrix(m, f1, f2, Dm, M):
    """Dense (M, m) matrix K with out = K @ x for one axis (see module doc)."""
    D1, D2, delta, theta = _czt_params(f1, f2, Dm, M)
    k = np.arange(M)[:, None]
    j = np.arange(m)[None, :]
    nu = D1 + (k + 1) * delta
    K = np.exp(-1j * 2 * np.pi * j * nu / Dm)
    l = D1 + np.arange(M) * delta
    post = np.exp(-1j * theta * (np.arange(M) + 0.5)) * np.exp(-1j * 2 * np.pi * l * (0.5 - m / 2) / Dm)
    K = post[:, None] * K
    if M >= m:
        K[M - 1, 0] = 0.0
    return K


def czt_axis_direct(x, f1, f2, Dm, M):
    """Axis-0 chirp-z of x (m, n) -> (n, M), same layout as the reference."""
    K = czt_matrix(x.shape[0], f1, f2, Dm, M)
    return (K @ x).T


def czt_axis_scipy(x, f1, f2, Dm, M):
    """As czt_axis_direct through scipy.signal.czt (valid for M < m)."""
    m = x.shape[0]
    D1, D2, delta, theta = _czt_params(f1, f2, Dm, M)
    a = np.exp(1j * 2 * np.pi * (D1 + delta) / Dm)
    w = np.exp(-1j * 2 * np.pi * delta / Dm)
    X = ssig.czt(x, m=M, w=w, a=a, axis=0)
    l = D1 + np.arange(M) * delta
    post = np.exp(-1j * theta * (np.arange(M) + 0.5)) * np.exp(-1j * 2 * np.pi * l * (0.5 - m / 2) / Dm)
    return (post[:, None] * X).T


def czt2d_direct(field, x, y, wavelength, z, xout, yout):
    """Direct form of CZT_jit (wave_optics.py:333-357)."""
    k = 2 * np.pi / wavelength
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    Dm = wavelength * z / dx
    X, Y = np.meshgrid(x, y)
    Xo, Yo = np.meshgrid(xout, yout)
    F = rs_kernel(X, Y, z, k)
    F0 = rs_kernel(Xo, Yo, z, k)
    Ky = czt_matrix(len(y), yout[0] + Dm / 2, yout[-1] + Dm / 2, Dm, len(yout))
    Kx = czt_matrix(len(x), xout[0] + Dm / 2, xout[-1] + Dm / 2, Dm, len(xout))
    return F0 * (Ky @ (field * F) @ Kx.T) * z * dx * dy * wavelength
