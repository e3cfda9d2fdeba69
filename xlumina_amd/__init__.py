"""MI355X-native light-propagation hot path of XLuminA.

Drop-in replacements for ScalarLight.RS_propagation / CZT,
VectorizedLight.VRS_propagation / VCZT and the high-NA objective-lens
focusing (VCZT_objective_lens), with the SLM / sSLM / LCD / polariser Jones
masks fused into the propagation kernels.  Fields are complex64 PyTorch-ROCm
tensors; the arithmetic runs in hand-written HIP kernels for gfx950
(libxlumina_hip.so, C ABI in include/xlumina_hip.h).

Units follow the reference (xlumina/__init__.py:15-20): lengths in microns.
"""
import math

um = 1
nm = 1e-3
mm = 1e3
cm = 1e4

radians = 1
degrees = 180 / math.pi

__version__ = "0.1.0"
