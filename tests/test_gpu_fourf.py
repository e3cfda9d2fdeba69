"""Batched CZT / mask adjoint / loss / Adam kernels and the batched 4f
optimisation (GPU).

Tolerances: batched launches vs one-problem launches <= 1e-6 relative L2
(same kernels on the same data; equality expected); MSE loss and cotangent
vs float64 <= 1e-5; Adam vs float64 optax restatement <= 1e-6; 4f loss and
phase gradients vs the float64 oracle <= 1e-5; a captured graph of S steps vs
S eager steps <= 1e-6.
"""
import math

import numpy as np
import pytest
import torch

from oracle import fourf_ref as F
from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import fourf, ops
from xlumina_amd.ops import MaskSpec

pytestmark = pytest.mark.gpu


def np_(t):
    return t.detach().cpu().numpy()


def test_batched_czt_matches_single(cuda_dev):
    n, B = 200, 3
    x, y = R.space(300.0, n)
    xo, yo = np.linspace(-40, 40, 90), np.linspace(-30, 35, 70)
    g = ops.czt_geometry(x, y, xo, yo, 4000.0, 0.65)
    rng = np.random.default_rng(0)
    shared = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    ph = torch.tensor(rng.uniform(-3, 3, (B, 2, n, n)), dtype=torch.float32, device=cuda_dev)
    out = ops.czt(shared, g, masks=[MaskSpec(N.XL_MASK_SLM, ph[:, 1])], batch=B)
    for b in range(B):
        ref = ops.czt(shared, g, masks=[MaskSpec(N.XL_MASK_SLM, ph[b, 1].contiguous())], form="bluestein")
        assert rel_l2(np_(out[b]), np_(ref)) <= 1e-6
    # per-member vectorial inputs with Ez construction and an sSLM
    src = torch.randn((B, 2, n, n), dtype=torch.complex64, device=cuda_dev)
    out = ops.czt(src, g, masks=[MaskSpec(N.XL_MASK_SSLM, ph[:, 0], ph[:, 1])], ez_mode=N.XL_EZ_VCZT, nin=2,
                  batch=B)
    assert out.shape == (B, 3, 70, 90)
    for b in range(B):
        ref = ops.czt(src[b], g, masks=[MaskSpec(N.XL_MASK_SSLM, ph[b, 0].contiguous(), ph[b, 1].contiguous())],
                      ez_mode=N.XL_EZ_VCZT, nin=2, form="bluestein")
        assert rel_l2(np_(out[b]), np_(ref)) <= 1e-6
    # adjoint
    cot = torch.randn((B, 1, 70, 90), dtype=torch.complex64, device=cuda_dev)
    adj = ops.czt(cot, g, adjoint=True, batch=B)
    for b in range(B):
        ref = ops.czt(cot[b], g, adjoint=True, form="bluestein")
        assert rel_l2(np_(adj[b]), np_(ref)) <= 1e-6


def test_batched_pre_backward_matches_single(cuda_dev):
    n, B = 96, 4
    x, y = R.space(100.0, n)
    rng = np.random.default_rng(1)
    ph = torch.tensor(rng.uniform(-3, 3, (B, 2, n, n)), dtype=torch.float32, device=cuda_dev)
    src = torch.randn((B, 1, n, n), dtype=torch.complex64, device=cuda_dev)
    gout = torch.randn((B, 1, n, n), dtype=torch.complex64, device=cuda_dev)
    grad = torch.zeros((B, 2, n, n), dtype=torch.float32, device=cuda_dev)
    gsrc, gp = ops.pre_backward_batched(src, gout, x, y, B, masks=[MaskSpec(N.XL_MASK_SLM, ph[:, 1])],
                                        need_params=((True, False), (False, False)),
                                        gparams=[[grad[:, 1], None], [None, None]])
    assert gp[0][0].data_ptr() == grad[:, 1].data_ptr()
    for b in range(B):
        rs, rp = ops.pre_backward(src[b], gout[b], x, y, masks=[MaskSpec(N.XL_MASK_SLM, ph[b, 1].contiguous())],
                                  need_params=((True, False), (False, False)))
        assert rel_l2(np_(gsrc[b]), np_(rs)) <= 1e-6
        assert rel_l2(np_(grad[b, 1]), np_(rp[0][0])) <= 1e-6
    assert float(grad[:, 0].abs().max()) == 0.0


def test_mse_intensity_kernel(cuda_dev):
    B, n = 3, 300
    E = torch.randn((B, n, n), dtype=torch.complex64, device=cuda_dev)
    T = (torch.rand((n, n), device=cuda_dev) > 0.5).float()
    loss, g = ops.mse_intensity(E, T)
    e = np_(E).astype(np.complex128)
    t = np_(T).astype(np.float64)
    d = np.abs(e) ** 2 - t
    ref_l = (d ** 2).sum(axis=(1, 2)) / (n * n)
    ref_g = 4 * d * e / (n * n)
    assert rel_l2(np_(loss), ref_l) <= 1e-5
    assert rel_l2(np_(g), ref_g) <= 1e-5
    # in place (grad aliases field) gives the same numbers
    E2 = E.clone()
    ops.mse_intensity(E2, T, grad=E2)
    assert rel_l2(np_(E2), np_(g)) <= 1e-7


@pytest.mark.parametrize("offset", [0, 1])
def test_adam_kernel_vs_float64(cuda_dev, offset):
    n = 10001
    rng = np.random.default_rng(4)
    base = torch.tensor(rng.standard_normal(4 * (n + 1)), dtype=torch.float32, device=cuda_dev)
    # offset 1 misaligns every array by 4 bytes (scalar path)
    p, g, m, v = (base[i * (n + 1) + offset: i * (n + 1) + offset + n] for i in range(4))
    m.zero_()
    v.zero_()
    count = torch.zeros(1, dtype=torch.int32, device=cuda_dev)
    pr = np_(p).astype(np.float64)
    st = (np.zeros(n), np.zeros(n), 0)
    for k in range(3):
        gk = rng.standard_normal(n)
        g.copy_(torch.tensor(gk, dtype=torch.float32))
        ops.adam_step(p, g, m, v, count, 0.05, weight_decay=1e-4)
        pr, st = F.adam(pr, np_(g).astype(np.float64), st, 0.05, weight_decay=1e-4)
    assert int(count.item()) == 3
    assert np.abs(np_(p) - pr).max() <= 1e-6 * max(1.0, np.abs(pr).max())


def test_fourf_loss_and_gradient_vs_oracle(cuda_dev):
    n, B = 24, 2
    opt = fourf.FourFOptimizer(cuda_dev, batch=B, n=n, half_width=60.0, f=2000.0, w0=40.0, steps_per_launch=1,
                               use_graph=False)
    loss = torch.zeros(B, dtype=torch.float32, device=cuda_dev)
    opt.forward_backward(loss)
    torch.cuda.synchronize()
    e_in = fourf.gaussian_field(opt.x, opt.y, 0.65, 40.0)
    ch = F.FourF(opt.x, opt.y, 0.65, 2000.0, e_in, np_(opt.target))
    ph = np_(opt.phi).astype(np.float64)
    for b in range(B):
        L, g1, g2 = ch.loss_and_grad(ph[b, 0], ph[b, 1])
        assert abs(float(loss[b]) - L) <= 1e-5 * L
        assert rel_l2(np_(opt.grad[b, 0]), g1) <= 1e-5
        assert rel_l2(np_(opt.grad[b, 1]), g2) <= 1e-5


def test_fourf_graph_equals_eager_steps(cuda_dev):
    n, B, S = 64, 2, 3
    kw = dict(batch=B, n=n, half_width=150.0, f=5000.0, w0=100.0, steps_per_launch=S)
    a = fourf.FourFOptimizer(cuda_dev, use_graph=True, **kw)
    b = fourf.FourFOptimizer(cuda_dev, use_graph=False, **kw)
    la = a.launch().clone()
    lb = b.launch().clone()
    torch.cuda.synchronize()
    assert rel_l2(np_(la), np_(lb)) <= 1e-6
    assert rel_l2(np_(a.phi), np_(b.phi)) <= 1e-6
    assert int(a.count.item()) == S and int(b.count.item()) == S
    la2 = a.launch().clone()                      # the replay continues from the updated state
    lb2 = b.launch().clone()
    torch.cuda.synchronize()
    assert rel_l2(np_(la2), np_(lb2)) <= 1e-6 and int(a.count.item()) == 2 * S


def test_fourf_config_size_descends(cuda_dev):
    """64 candidates x 1024^2: the loss of every candidate falls over 10 steps."""
    opt = fourf.FourFOptimizer(cuda_dev, steps_per_launch=10)
    L = np_(opt.launch())
    assert np.isfinite(L).all()
    assert (L[-1] < L[0]).all(), (L[0][:4], L[-1][:4])


def test_fused_mask_adjoint_adam_step_matches_unfused(cuda_dev):
    """step() (mask adjoints fused with the Adam update) against
    forward_backward() + xl_adam_step over three steps from one state."""
    from xlumina_amd.fourf import FourFOptimizer
    kw = dict(batch=3, n=96, half_width=300.0, f=5e3, w0=300.0, steps_per_launch=1, use_graph=False,
              weight_decay=0.01)
    fused, plain = FourFOptimizer(cuda_dev, **kw), FourFOptimizer(cuda_dev, **kw)
    lf = torch.zeros(3, device=cuda_dev)
    lp = torch.zeros(3, device=cuda_dev)
    for _ in range(3):
        fused.step(lf)
        plain.step_unfused(lp)
        torch.cuda.synchronize()
        assert torch.allclose(lf, lp, rtol=1e-6, atol=0)
    for a, b in ((fused.phi, plain.phi), (fused.m, plain.m), (fused.v, plain.v)):
        assert float((a - b).abs().max()) <= 1e-6 * float(b.abs().max())
    assert int(fused.count) == int(plain.count) == 3
