"""Split-BF16 MFMA zoom-DFT form of the CZT vs the float64 oracle (GPU).

Tolerance: relative L2 <= 1e-4 against float64 (bf16 hi + lo split,
FP32 accumulation), as for the chosen CZT form.
"""
import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import autograd as AG
from xlumina_amd import ops, workloads

pytestmark = pytest.mark.gpu
TOL = 1e-4


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)


def H(t):
    return t.detach().cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("n,mx,my", [(64, 40, 33), (200, 77, 50), (40, 70, 50), (1024, 256, 256)])
def test_gemm_czt_vs_oracle(n, mx, my, cuda_dev):
    rng = np.random.default_rng(n + mx)
    x, y = R.space(1200.0, n)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    xo = np.linspace(-100, 100, mx)
    yo = np.linspace(-80, 120, my)
    ref = R.czt(u, x, y, 0.635, 3e4, xo, yo)
    out = ops.czt(T(u, cuda_dev), ops.czt_geometry(x, y, xo, yo, 3e4, 0.635), form="gemm")
    assert rel_l2(H(out[0]), ref) < TOL


def test_gemm_matches_bluestein_config2_czt(cuda_dev):
    w = workloads.vrs_vczt_2048(cuda_dev)
    g = w.geometry()
    vrs, blue = w.run(g)
    gem = ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, form="gemm")
    assert float((gem - blue).norm() / blue.norm()) < TOL


def test_gemm_hna_vs_oracle(golden, cuda_dev):
    c = golden["hna_n32"]
    g = ops.hna_geometry(c["x"], c["y"], c["xout"], c["yout"], float(c["wavelength"]), float(c["radius"]),
                         float(c["f"]))
    out = ops.czt(T(np.stack([c["Ex"], c["Ey"]]), cuda_dev), g, nin=2, form="gemm")
    assert rel_l2(H(out), c["out"]) < TOL


def test_gemm_high_na_full_size_vs_bluestein(cuda_dev):
    w = workloads.high_na_2048(cuda_dev)
    g = w.geometry()
    a = ops.czt(w.field, g, masks=w.masks(), nin=2, form="bluestein")
    b = ops.czt(w.field, g, masks=w.masks(), nin=2, form="gemm")
    assert float((a - b).norm() / a.norm()) < TOL


def test_gemm_adjoint_identity(cuda_dev):
    n = 200
    x, y = R.space(300.0, n)
    xo, yo = np.linspace(-30, 30, 77), np.linspace(-20, 25, 50)
    g = ops.czt_geometry(x, y, xo, yo, 4000.0, 0.65)
    a = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    c = torch.randn((1, 50, 77), dtype=torch.complex64, device=cuda_dev)
    ya = ops.czt(a, g, form="gemm")
    ac = ops.czt(c, g, adjoint=True, form="gemm")
    lhs = torch.vdot(c.flatten().to(torch.complex128), ya.flatten().to(torch.complex128))
    rhs = torch.vdot(ac.flatten().to(torch.complex128), a.flatten().to(torch.complex128))
    assert abs(complex(lhs - rhs)) / abs(complex(lhs)) < TOL


def test_gemm_grad_matches_bluestein_grad(cuda_dev):
    w = workloads.high_na_2048(cuda_dev, n=256, m_out=64)
    g = w.geometry()
    grads = []
    for form in ("bluestein", "gemm"):
        alpha = w.alpha.clone().requires_grad_(True)
        E = AG.czt(w.field, g, masks=w.masks(alpha, w.phi), nin=2, form=form)
        L = workloads.focal_second_moment(E, w.xout, w.yout)
        (ga,) = torch.autograd.grad(L, (alpha,))
        grads.append(ga)
    assert float((grads[0] - grads[1]).norm() / grads[0].norm()) < 1e-3
