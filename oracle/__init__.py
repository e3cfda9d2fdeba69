"""CPU oracle for the light-propagation hot path (test infrastructure only).

This package restates, in float64 NumPy, the arithmetic of the XLuminA
reference functions on the propagation path (RS / CZT / VRS / VCZT /
high-NA objective lens / Jones masks / sources).  It is the checker that the
HIP kernels are compared against and the CPU baseline that ``bench.py`` times.

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
Parity unpinned: the z<0 branch of the RS kernel (the reference keeps the
1/r - i k factor unconjugated there) is pinned only by this restatement.
"""
from . import xlumina_ref, brute  # noqa: F401
