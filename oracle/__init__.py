"""CPU oracle for the light-propagation hot path (test infrastructure only).

This package restates, in float64 NumPy, the arithmetic of the XLuminA
reference functions on the propagation path (RS / CZT / VRS / VCZT /
high-NA objective lens / Jones masks / sources / beam splitters, and in
``losses_ref`` and ``fourf_ref`` the losses and the 4f training step).  The
HIP kernels are compared against it, and ``bench.py`` times it as the CPU
baseline.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it.  The product package ``xlumina_amd`` never does: its
kernels run on the GPU or raise.

Pinning (what checks this restatement independently of retyping it):
  * RS / VRS: direct O(N^4) Rayleigh-Sommerfeld sum (``oracle.brute``), the
    kernel H against a finite-difference z-derivative of the free-space
    Green's function, linearity.
  * CZT / VCZT / high-NA CZT: direct DFT sum and ``scipy.signal.czt`` (a
    library chirp-z routine) with the reference's output phase factors.
  * Jones masks: the reference's own value tests (identity, pi, crossed
    polariser) plus unitarity / projector identities.
  * High-NA map: orthogonality of the Richards-Wolf 3x3 matrix and its
    phi=0 special case (rotation about y).
  * Sources: closed-form values (peak amplitude, 1/e radius, normalisation).
  * dH/dz: central differences of the RS kernel in z.
  * Losses: closed forms (a flat spot's small-area loss is the total
    intensity; MSE of scaled / rotated fields; softmin limits) and central
    differences of the hand-derived small-area gradient.
  * Beam splitters: energy identities ([[R, iT], [iT, R]] scales the energy
    by R^2 + T^2; BS is lossless for R + T = 1) and the port swap.
  * 4f step: central differences of the loss, Adam closed forms.
Parity unpinned: the z<0 branch of the RS kernel (the reference keeps the
1/r - i k factor unconjugated there) is pinned only by this restatement.
"""
from . import xlumina_ref, brute  # noqa: F401
