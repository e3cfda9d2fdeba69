"""Float64 restatement of the batched 4f optimisation step (TEST ORACLE).

Test infrastructure only: imported by tests, never by the product package.

* Optical table: the dual-SLM 4f system of
  experiments/four_f_optical_table.py:36-83 (SLM, propagate, SLM, propagate)
  with ScalarLight.CZT (wave_optics.py:233-263, restated in
  ``oracle.xlumina_ref.czt``) as the propagator:
      a1 = CZT_f(E_in); a2 = CZT_2f(a1 e^{i phi1}); a3 = CZT_f(a2 e^{i phi2}).
* Loss: MSE_Intensity (xlumina/loss_functions.py:137-149),
  sum (|a3|^2 - |T|^2)^2 / npix.
* Optimiser: optax 0.2.3 (pinned in setup.py:33; a pip dependency, absent
  here).  Its published Adam (``scale_by_adam``): mu = b1 mu + (1-b1) g,
  nu = b2 nu + (1-b2) g^2, t = count + 1, update = (mu / (1-b1^t)) /
  (sqrt(nu / (1-b2^t) + eps_root) + eps) with eps_root = 0; adamw adds
  weight_decay * p (``add_decayed_weights``); ``scale_by_learning_rate``
  multiplies by -lr.  experiments/four_f_optimizer.py:112 uses adamw.
* Gradient: reverse mode written out by hand.  The adjoint of a CZT is the
  conjugate transpose of its dense operator matrix, built column by column
  from the forward oracle (small grids only).  Pinned in the tests by
  central finite differences of ``loss``.
"""
from __future__ import annotations

import numpy as np

from . import xlumina_ref as R


def czt_operator(x, y, wavelength, z):
    """Dense (n^2 x n^2) matrix of CZT(z) from the grid onto itself."""
    n = len(x)
    A = np.empty((n * n, n * n), dtype=np.complex128)
    e = np.zeros((n, n), dtype=np.complex128)
    for j in range(n * n):
        e.flat[j] = 1.0
        A[:, j] = R.czt(e, x, y, wavelength, z, x, y).ravel()
        e.flat[j] = 0.0
    return A


class FourF:
    """The chain for one candidate on an n x n grid (operators cached)."""

    def __init__(self, x, y, wavelength, f, e_in, target_intensity):
        self.A_f = czt_operator(x, y, wavelength, f)
        self.A_2f = czt_operator(x, y, wavelength, 2 * f)
        self.a1 = self.A_f @ np.asarray(e_in, dtype=np.complex128).ravel()
        self.T = np.asarray(target_intensity, dtype=np.float64).ravel()
        self.shape = np.shape(e_in)

    def forward(self, phi1, phi2):
        b1 = self.a1 * np.exp(1j * np.ravel(phi1))
        a2 = self.A_2f @ b1
        b2 = a2 * np.exp(1j * np.ravel(phi2))
        a3 = self.A_f @ b2
        return b1, a2, b2, a3

    def loss(self, phi1, phi2):
        a3 = self.forward(phi1, phi2)[3]
        return float(np.sum((np.abs(a3) ** 2 - self.T) ** 2) / a3.size)

    def loss_and_grad(self, phi1, phi2):
        """(loss, dL/dphi1, dL/dphi2); cotangents as dL/dRe + i dL/dIm."""
        b1, a2, b2, a3 = self.forward(phi1, phi2)
        d = np.abs(a3) ** 2 - self.T
        L = float(np.sum(d ** 2) / a3.size)
        g3 = 4.0 * d * a3 / a3.size
        gb2 = self.A_f.conj().T @ g3
        gphi2 = np.real(np.conj(gb2) * 1j * b2)
        ga2 = gb2 * np.exp(-1j * np.ravel(phi2))
        gb1 = self.A_2f.conj().T @ ga2
        gphi1 = np.real(np.conj(gb1) * 1j * b1)
        return L, gphi1.reshape(self.shape), gphi2.reshape(self.shape)


def adam(params, grads, state, lr, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
    """One optax adam / adamw update in float64; state = (mu, nu, count)."""
    mu, nu, count = state
    mu = b1 * mu + (1 - b1) * grads
    nu = b2 * nu + (1 - b2) * grads ** 2
    t = count + 1
    upd = (mu / (1 - b1 ** t)) / (np.sqrt(nu / (1 - b2 ** t)) + eps) + weight_decay * params
    return params - lr * upd, (mu, nu, t)
