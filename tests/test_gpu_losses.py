"""Fused losses and beam splitters on the GPU against the float64 oracles.

Tolerances: losses and fields <= 1e-5 relative; field cotangents against
float64 torch autograd of the reference formula (CPU) <= 1e-5 relative L2;
beam-splitter parameter gradients against central differences of the oracle
<= 1e-4 relative.
"""
import numpy as np
import pytest
import torch

from oracle import losses_ref as LR
from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import loss_functions as LF
from xlumina_amd import optical_elements as OE
from xlumina_amd.toolbox import softmin
from xlumina_amd.vectorized_optics import VectorizedLight

pytestmark = pytest.mark.gpu


def cfield(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


REF = {"MSE_Intensity": LR.MSE_Intensity, "MSE_Amplitude": LR.MSE_Amplitude, "MSE_Phase": LR.MSE_Phase}


def torch_ref(name, e, t):
    """The reference formula in float64 torch (CPU), for autograd."""
    npix = e.shape[-1] * e.shape[-2]
    if name == "MSE_Intensity":
        d = e.abs() ** 2 - t.abs() ** 2
    elif name == "MSE_Amplitude":
        d = e.abs() - t.abs()
    else:
        d = torch.angle(e) - torch.angle(t)
    return (d ** 2).sum() / npix


@pytest.mark.parametrize("name", list(REF))
def test_field_losses_and_gradients(cuda_dev, name):
    rng = np.random.default_rng(60)
    e, t = cfield(rng, (200, 176)), cfield(rng, (200, 176))
    et = torch.tensor(e.astype(np.complex64), device=cuda_dev, requires_grad=True)
    loss = getattr(LF, name)(et, t)
    ref = REF[name](e, t)
    assert abs(float(loss) - ref) <= 1e-5 * abs(ref)
    (g,) = torch.autograd.grad(3.0 * loss, et)
    ec = torch.tensor(e, requires_grad=True)
    (gr,) = torch.autograd.grad(3.0 * torch_ref(name, ec, torch.tensor(t)), ec)
    assert rel_l2(g.cpu().numpy(), gr.numpy()) <= 1e-5


def test_vectorial_and_batched_mse(cuda_dev):
    rng = np.random.default_rng(61)
    x, y = R.space(50.0, 64)
    E, T = cfield(rng, (3, 64, 64)), cfield(rng, (3, 64, 64))
    li, lt = VectorizedLight(x, y, 0.6, cuda_dev), VectorizedLight(x, y, 0.6, cuda_dev)
    li.E, lt.E = E.astype(np.complex64), T.astype(np.complex64)
    for fn, ref in ((LF.vMSE_Intensity, LR.MSE_Intensity), (LF.vMSE_Amplitude, LR.MSE_Amplitude),
                    (LF.vMSE_Phase, LR.MSE_Phase)):
        got = fn(li, lt).cpu().numpy()
        assert rel_l2(got, LR.vMSE(ref, E, T)) <= 1e-5
    B = cfield(rng, (5, 64, 64))
    m, per = LF.mean_batch_MSE_Intensity(B, np.broadcast_to(T[0], B.shape))
    rm, rper = LR.mean_batch_MSE_Intensity(B, np.broadcast_to(T[0], B.shape))
    assert rel_l2(per.cpu().numpy(), rper) <= 1e-5 and abs(float(m) - rm) <= 1e-5 * rm


def test_small_area_losses_and_gradient(cuda_dev):
    rng = np.random.default_rng(62)
    x = np.linspace(-1, 1, 300)
    X, Y = np.meshgrid(x, x)
    dets = np.stack([np.exp(-((X - c) ** 2 + Y ** 2) / w) for c, w in ((0.0, 0.02), (0.3, 0.05), (-0.5, 0.01))])
    dets = dets + 0.01 * rng.random(dets.shape)
    I = dets / dets.sum(axis=(1, 2), keepdims=True)
    near = np.abs(I - 0.7 * I.max(axis=(1, 2), keepdims=True)) < 1e-5 * I.max(axis=(1, 2), keepdims=True)
    assert not near.any()   # no pixel sits on the mask threshold at float32 rounding
    d = torch.tensor(dets.astype(np.float32), device=cuda_dev, requires_grad=True)
    got = LF.vectorized_loss_hybrid(d)
    ref = LR.vectorized_loss_hybrid(dets.astype(np.float32).astype(np.float64))
    assert rel_l2(got.detach().cpu().numpy(), ref) <= 1e-5
    w = torch.tensor([1.0, -2.0, 0.5], device=cuda_dev)
    (g,) = torch.autograd.grad((w * got).sum(), d)
    gr = np.stack([wi * LR.small_area_hybrid_grad(di) for wi, di in zip((1.0, -2.0, 0.5),
                                                                         dets.astype(np.float32).astype(np.float64))])
    assert rel_l2(g.cpu().numpy(), gr) <= 1e-5
    one = LF.small_area_hybrid(dets[1])
    assert abs(float(one) - ref[1]) <= 1e-5 * ref[1]
    assert abs(float(softmin(got.detach().cpu().double())) - LR.softmin(ref)) <= 1e-5 * abs(LR.softmin(ref))


def lights(rng, dev, n=48):
    x, y = R.space(40.0, n)
    out = []
    for _ in range(2):
        L = VectorizedLight(x, y, 0.6, dev)
        e = cfield(rng, (3, n, n))
        L.E = e.astype(np.complex64)
        out.append((L, e.astype(np.complex64).astype(np.complex128)))
    return out


def stack(c):
    return np.stack([np.asarray(v) for v in c])


def test_beam_splitters_vs_oracle(cuda_dev):
    rng = np.random.default_rng(63)
    (la, ea), (lb, eb) = lights(rng, cuda_dev)
    c, d = OE.BS_symmetric(la, lb, 0.8)
    rc, rd = R.BS_symmetric(ea[:2], eb[:2], 0.8)
    assert rel_l2(c.E.cpu().numpy(), stack(rc)) <= 1e-6 and rel_l2(d.E.cpu().numpy(), stack(rd)) <= 1e-6
    for port in (0, 1):
        c, d = OE.BS_symmetric_SI(la, 2.1, port=port)
        rc, rd = R.BS_symmetric_SI(ea[:2], 2.1, port)
        assert rel_l2(c.E.cpu().numpy(), stack(rc)) <= 1e-6 and rel_l2(d.E.cpu().numpy(), stack(rd)) <= 1e-6
    b1, b2 = OE.BS(la, lb, 0.35, 0.65, 0.9)
    r1, r2 = R.BS(ea[:2], eb[:2], 0.35, 0.65, 0.9)
    assert rel_l2(b1.E.cpu().numpy(), stack(r1)) <= 1e-6 and rel_l2(b2.E.cpu().numpy(), stack(r2)) <= 1e-6


def test_beam_splitter_gradients(cuda_dev):
    rng = np.random.default_rng(64)
    (la, ea), (lb, eb) = lights(rng, cuda_dev)
    w1, w2 = rng.random((48, 48)), rng.random((48, 48))
    wt1, wt2 = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) for w in (w1, w2))

    def dev_loss(theta, a_E, b_E):
        la.E, lb.E = a_E, b_E
        c, d = OE.BS_symmetric(la, lb, theta)
        return (wt1 * (c.Ex.abs() ** 2 + c.Ey.abs() ** 2)).sum() + (wt2 * d.Ey.abs() ** 2).sum()

    def ref_loss(theta):
        c, d = R.BS_symmetric(ea[:2], eb[:2], theta)
        return float(np.sum(w1 * (np.abs(c[0]) ** 2 + np.abs(c[1]) ** 2)) + np.sum(w2 * np.abs(d[1]) ** 2))

    th = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    aE = torch.tensor(ea.astype(np.complex64), device=cuda_dev, requires_grad=True)
    bE = torch.tensor(eb.astype(np.complex64), device=cuda_dev, requires_grad=True)
    L = dev_loss(th, aE, bE)
    g_th, g_a, g_b = torch.autograd.grad(L, (th, aE, bE))
    fd = (ref_loss(0.7 + 1e-6) - ref_loss(0.7 - 1e-6)) / 2e-6
    assert abs(float(g_th) - fd) <= 1e-4 * abs(fd)
    # field cotangents against float64 torch autograd of the same splitter
    ta = torch.tensor(ea, requires_grad=True)
    tb = torch.tensor(eb, requires_grad=True)
    Rr, T = abs(np.sin(0.7)) - 0.01 * abs(np.cos(0.7)), 0.99 * abs(np.cos(0.7))
    cx, cy = Rr * ta[0] + 1j * T * tb[0], Rr * ta[1] + 1j * T * tb[1]
    dy = 1j * T * ta[1] + Rr * tb[1]
    Lr = (torch.tensor(w1) * (cx.abs() ** 2 + cy.abs() ** 2)).sum() + (torch.tensor(w2) * dy.abs() ** 2).sum()
    ra, rb = torch.autograd.grad(Lr, (ta, tb))
    assert rel_l2(g_a.cpu().numpy(), ra.numpy()) <= 1e-5 and rel_l2(g_b.cpu().numpy(), rb.numpy()) <= 1e-5


def test_focal_moment_kernel_and_gradient(cuda_dev):
    """The fused second-moment loss against float64 torch (CPU) and its
    autograd gradient, with a non-unit incoming gradient."""
    from xlumina_amd import workloads
    rng = np.random.default_rng(65)
    E = cfield(rng, (3, 96, 80))
    xo, yo = np.linspace(-2, 2, 80), np.linspace(-1.5, 1.5, 96)
    Et = torch.tensor(E.astype(np.complex64), device=cuda_dev, requires_grad=True)
    L = workloads.focal_second_moment(Et, xo, yo)
    (g,) = torch.autograd.grad(2.5 * L, Et)
    Ec = torch.tensor(E.astype(np.complex64).astype(np.complex128), requires_grad=True)
    yy, xx = np.meshgrid(yo, xo, indexing="ij")
    r2 = torch.tensor((xx ** 2 + yy ** 2).astype(np.float32).astype(np.float64))
    I = (Ec.abs() ** 2).sum(0)
    Lr = (I * r2).sum() / I.sum()
    (gr,) = torch.autograd.grad(2.5 * Lr, Ec)
    assert abs(float(L) - float(Lr)) <= 1e-6 * abs(float(Lr))
    assert rel_l2(g.cpu().numpy(), gr.numpy()) <= 1e-6


@pytest.mark.parametrize("port", [0, 1])
@pytest.mark.parametrize("theta", [0.7, 2.6, -1.9])
def test_single_input_splitter_gradients(cuda_dev, port, theta):
    """dL/dtheta of BS_symmetric_SI, formed by the dot-reduce kernel from the
    four coefficient sums, against central differences of the oracle splitter
    (theta in each quadrant's sign pattern of sin and cos)."""
    rng = np.random.default_rng(65 + port)
    (la, ea), _ = lights(rng, cuda_dev)
    w1, w2 = rng.random((48, 48)), rng.random((48, 48))
    wt1, wt2 = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) for w in (w1, w2))
    th = torch.tensor(theta, dtype=torch.float64, device=cuda_dev, requires_grad=True)
    c, d = OE.BS_symmetric_SI(la, th, port=port)
    L = (wt1 * (c.Ex.abs() ** 2 + c.Ey.abs() ** 2)).sum() + (wt2 * (d.Ex * d.Ey.conj()).real).sum()
    (g,) = torch.autograd.grad(L, (th,))

    def ref_loss(t):
        rc, rd = R.BS_symmetric_SI(ea[:2], t, port)
        return float(np.sum(w1 * (np.abs(rc[0]) ** 2 + np.abs(rc[1]) ** 2)) + np.sum(w2 * (rd[0] * np.conj(rd[1])).real))

    fd = (ref_loss(theta + 1e-6) - ref_loss(theta - 1e-6)) / 2e-6
    assert abs(float(g) - fd) <= 1e-4 * abs(fd) + 1e-6 * abs(ref_loss(theta)), (float(g), fd)


def test_quality_factor_from_device_distance(cuda_dev):
    """ops.quality_factor with z on the device (one launch) equals the float64
    host formula (wave_optics.py:187-191)."""
    from xlumina_amd import ops
    x, y = R.space(500.0, 64)
    for z in (150.0, 8.9e4, 2.3e6):
        got = ops.quality_factor(x, y, 0.635, torch.tensor([z], dtype=torch.float64, device=cuda_dev))
        assert got.device.type == "cuda"
        assert abs(float(got) - ops.quality_factor(x, y, 0.635, z)) <= 1e-12 * ops.quality_factor(x, y, 0.635, z)


def test_quality_factor_reported_by_the_propagation(cuda_dev):
    """RS_propagation / VRS_propagation to a distance held on the device
    return the quality factor written by the propagation launch itself,
    equal to the oracle's float64 formula (wave_optics.py:187-191)."""
    from xlumina_amd.vectorized_optics import VectorizedLight
    from xlumina_amd.wave_optics import ScalarLight
    n = 96
    x, y = R.space(300.0, n)
    rng = np.random.default_rng(61)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for z in (700.0, 3.1e4):
        zt = torch.tensor([z], dtype=torch.float64, device=cuda_dev, requires_grad=True)
        s = ScalarLight(x, y, 0.635, cuda_dev)
        s.field = torch.tensor(u.astype(np.complex64), device=cuda_dev)
        v = VectorizedLight(x, y, 0.635, cuda_dev)
        v.Ex = torch.tensor(u.astype(np.complex64), device=cuda_dev)
        want = R.quality_factor(x, y, 0.635, z)
        for light, qf in (s.RS_propagation(zt), v.VRS_propagation(zt)):
            assert isinstance(qf, torch.Tensor) and qf.device.type == "cuda"
            assert abs(float(qf) - want) <= 1e-12 * want
