"""Hand-written adjoints behind torch.autograd (GPU).

Pins: the dot-product identity <A x, y> = <x, A^H y> for every propagator
(with masks fused), and central finite differences of float64 oracle losses
(eps = 1e-3) on 64 seeded mask pixels.  The reference's jax.grad cannot run
here (JAX is absent); finite differences of the float64 restatement pin the
same derivative.
"""
import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import autograd as AG
from xlumina_amd import ops, workloads
from xlumina_amd.ops import MaskSpec

pytestmark = pytest.mark.gpu


def c128(t):
    return t.detach().to(torch.complex128).flatten()


def dot_identity(fn, x, cot):
    """Re<fn(x), cot> vs Re<x, grad> with grad from autograd (= A^H cot)."""
    x = x.clone().requires_grad_(True)
    y = fn(x)
    L = torch.real(torch.vdot(cot.flatten(), y.flatten()))
    (g,) = torch.autograd.grad(L, x)
    lhs = torch.vdot(c128(cot), c128(y))
    rhs = torch.vdot(c128(g), c128(x))
    return abs(complex(lhs) - complex(rhs)) / abs(complex(lhs))


def test_rs_adjoint_identity(cuda_dev):
    n = 256
    x, y = R.space(300.0, n)
    a = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    c = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    assert dot_identity(lambda u: AG.rs(u, x, y, 2000.0, 0.65), a, c) < 1e-5


# The adjoint identity holds to the form's accuracy: FP32 Bluestein 1e-5,
# split-BF16 zoom-GEMM 1e-4.
FORM_TOL = [("bluestein", 1e-5), ("gemm", 1e-4)]


@pytest.mark.parametrize("form,tol", FORM_TOL)
def test_czt_adjoint_identity(cuda_dev, form, tol):
    torch.manual_seed(5)
    n = 200
    x, y = R.space(300.0, n)
    xo, yo = np.linspace(-30, 30, 77), np.linspace(-20, 25, 50)
    g = ops.czt_geometry(x, y, xo, yo, 4000.0, 0.65)
    a = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    c = torch.randn((1, 50, 77), dtype=torch.complex64, device=cuda_dev)
    assert dot_identity(lambda u: AG.czt(u, g, form=form), a, c) < tol


def test_vrs_with_masks_adjoint_identity(cuda_dev):
    n = 128
    rng = np.random.default_rng(0)
    x, y = R.space(300.0, n)
    a, p = (torch.tensor(v, dtype=torch.float32, device=cuda_dev) for v in rng.uniform(-3, 3, (2, n, n)))
    masks = [MaskSpec(N.XL_MASK_SSLM, a, p), MaskSpec(N.XL_MASK_LCD_SCALAR, None, None, 0.3, 1.2)]
    E = torch.randn((2, n, n), dtype=torch.complex64, device=cuda_dev)
    c = torch.randn((3, n, n), dtype=torch.complex64, device=cuda_dev)
    assert dot_identity(lambda u: AG.rs(u, x, y, 2000.0, 0.65, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2), E, c) < 1e-5


@pytest.mark.parametrize("form,tol", FORM_TOL)
def test_hna_adjoint_identity(cuda_dev, form, tol):
    torch.manual_seed(6)
    n = 128
    x, y = R.space(1800.0, n)
    xo = np.linspace(-2, 2, 40)
    g = ops.hna_geometry(x, y, xo, xo, 0.65, 1800.0, 2000.0)
    E = torch.randn((2, n, n), dtype=torch.complex64, device=cuda_dev)
    c = torch.randn((3, 40, 40), dtype=torch.complex64, device=cuda_dev)
    assert dot_identity(lambda u: AG.czt(u, g, nin=2, form=form), E, c) < tol


def _oracle_hna_loss(w, alpha, phi):
    ex, ey = w.field[0].cpu().numpy().astype(np.complex128), w.field[1].cpu().numpy().astype(np.complex128)
    sx, sy, _ = R.sSLM(ex, ey, None, alpha, phi)
    lx, ly, _ = R.LCD(sx, sy, None, w.eta, w.theta)
    E = R.vczt_objective_lens(lx, ly, w.x, w.y, w.wavelength, w.radius, w.f, w.xout, w.yout)
    I = np.sum(np.abs(E) ** 2, 0)
    U, V = np.meshgrid(w.xout, w.yout)
    return float(np.sum(I * (U ** 2 + V ** 2)) / np.sum(I))


def test_high_na_grad_vs_central_fd_64_pixels(cuda_dev):
    """Second-moment loss through sSLM + LCD + VCZT_objective_lens (reduced
    128^2 pupil so 128 float64 oracle evaluations stay fast)."""
    w = workloads.high_na_2048(cuda_dev, n=128, m_out=64)
    alpha = w.alpha.clone().requires_grad_(True)
    phi = w.phi.clone().requires_grad_(True)
    L = w.loss(alpha, phi)
    ga, gp = torch.autograd.grad(L, (alpha, phi))
    a0 = w.alpha.cpu().numpy().astype(np.float64)
    p0 = w.phi.cpu().numpy().astype(np.float64)
    assert abs(L.item() - _oracle_hna_loss(w, a0, p0)) < 1e-5 * abs(L.item())
    rng = np.random.default_rng(4)
    pix = rng.choice(128 * 128, 64, replace=False)
    eps = 1e-3
    fd_a, fd_p = [], []
    for q in pix:
        i, j = divmod(int(q), 128)
        for arr, out in ((a0, fd_a), (p0, fd_p)):
            arr[i, j] += eps
            lp = _oracle_hna_loss(w, a0, p0)
            arr[i, j] -= 2 * eps
            lm = _oracle_hna_loss(w, a0, p0)
            arr[i, j] += eps
            out.append((lp - lm) / (2 * eps))
    got_a = ga.cpu().numpy().ravel()[pix]
    got_p = gp.cpu().numpy().ravel()[pix]
    assert rel_l2(got_a, np.array(fd_a)) < 2e-3
    assert rel_l2(got_p, np.array(fd_p)) < 2e-3


def test_high_na_grad_full_size_directional(cuda_dev):
    """Full 2048^2 pupil, 256^2 focal window: the autograd gradient along a
    random direction on 64 pixels vs a float64 central difference."""
    w = workloads.high_na_2048(cuda_dev)
    alpha = w.alpha.clone().requires_grad_(True)
    L = w.loss(alpha, w.phi)
    (ga,) = torch.autograd.grad(L, (alpha,))
    rng = np.random.default_rng(4)
    pix = rng.choice(2048 * 2048, 64, replace=False)
    v = np.zeros(2048 * 2048)
    v[pix] = rng.standard_normal(64)
    v = v.reshape(2048, 2048)
    a0 = w.alpha.cpu().numpy().astype(np.float64)
    p0 = w.phi.cpu().numpy().astype(np.float64)
    eps = 1e-3
    fd = (_oracle_hna_loss(w, a0 + eps * v, p0) - _oracle_hna_loss(w, a0 - eps * v, p0)) / (2 * eps)
    got = float((ga.double().cpu().numpy() * v).sum())
    assert abs(got - fd) < 5e-3 * abs(fd) + 1e-9


def test_rs_slm_grad_vs_fd(cuda_dev):
    n = 64
    x, y = R.space(100.0, n)
    rng = np.random.default_rng(9)
    u = R.gaussian_beam(x, y, 0.65, (40.0, 40.0), 1.0)
    phi0 = rng.uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    tgt = rng.random((n, n))
    field = torch.from_numpy(u.astype(np.complex64))[None].to(cuda_dev)
    phi = torch.from_numpy(phi0).to(cuda_dev).requires_grad_(True)
    out = AG.rs(field, x, y, 500.0, 0.65, masks=[MaskSpec(N.XL_MASK_SLM, phi)])
    L = ((out[0].abs() ** 2 - torch.from_numpy(tgt).float().to(cuda_dev)) ** 2).mean()
    (g,) = torch.autograd.grad(L, phi)

    def oracle_loss(ph):
        o, _ = R.rs_propagation(u * np.exp(1j * ph), x, y, 0.65, 500.0)
        return float(np.mean((np.abs(o) ** 2 - tgt) ** 2))

    p = phi0.astype(np.float64)
    pix = rng.choice(n * n, 32, replace=False)
    fd = []
    for q in pix:
        i, j = divmod(int(q), n)
        p[i, j] += 1e-3
        lp = oracle_loss(p)
        p[i, j] -= 2e-3
        lm = oracle_loss(p)
        p[i, j] += 1e-3
        fd.append((lp - lm) / 2e-3)
    assert rel_l2(g.cpu().numpy().ravel()[pix], np.array(fd)) < 2e-3


def test_rs_kernel_dz_propagation_vs_oracle(cuda_dev):
    """The dH/dz convolution (kernel_dz) against the float64 oracle's."""
    n = 96
    x, y = R.space(150.0, n)
    rng = np.random.default_rng(21)
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for z in (800.0, -600.0):
        got = ops.rs_propagate(torch.from_numpy(u.astype(np.complex64))[None].to(cuda_dev), x, y, z, 0.65,
                               kernel_dz=True)[0].cpu().numpy()
        ref = R.rs_propagation_dz(u, x, y, 0.65, z)
        assert rel_l2(got, ref) <= 1e-5, z


def test_rs_grad_wrt_z_vs_fd(cuda_dev):
    """d loss / d z through ScalarLight-style RS with a fused SLM, against a
    central difference of the float64 oracle loss in z."""
    n = 64
    x, y = R.space(100.0, n)
    rng = np.random.default_rng(22)
    u = R.gaussian_beam(x, y, 0.65, (40.0, 40.0), 1.0)
    phi = rng.uniform(-np.pi, np.pi, (n, n))
    tgt = rng.random((n, n))
    field = torch.from_numpy(u.astype(np.complex64))[None].to(cuda_dev)
    z0 = 700.0
    z = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
    out = AG.rs(field, x, y, z, 0.65, masks=[MaskSpec(N.XL_MASK_SLM, torch.from_numpy(phi).float().to(cuda_dev))])
    L = ((out[0].abs() ** 2 - torch.from_numpy(tgt).float().to(cuda_dev)) ** 2).mean()
    (gz,) = torch.autograd.grad(L, z)

    def oracle_loss(zz):
        o, _ = R.rs_propagation(u * np.exp(1j * phi), x, y, 0.65, zz)
        return float(np.mean((np.abs(o) ** 2 - tgt) ** 2))

    eps = 1e-3
    fd = (oracle_loss(z0 + eps) - oracle_loss(z0 - eps)) / (2 * eps)
    assert abs(float(gz) - fd) <= 2e-3 * abs(fd)


def test_vrs_grad_wrt_z_vs_fd(cuda_dev):
    """VRS with an LCD and the constructed Ez (whose z-dependence enters through
    r = sqrt(X^2 + Y^2 + z^2)) against a central difference of the oracle."""
    n = 64
    x, y = R.space(100.0, n)
    rng = np.random.default_rng(23)
    ex, ey, _ = R.polarized_gaussian_beam(x, y, 0.65, (40.0, 40.0), (1.0, 1j))
    eta, th = rng.uniform(0, 2 * np.pi, (2, n, n))
    w = rng.random((n, n))
    src = torch.from_numpy(np.stack([ex, ey]).astype(np.complex64)).to(cuda_dev)
    masks = [MaskSpec(N.XL_MASK_LCD, torch.from_numpy(eta).float().to(cuda_dev),
                      torch.from_numpy(th).float().to(cuda_dev))]
    z0 = 600.0
    z = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
    out = AG.rs(src, x, y, z, 0.65, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2)
    wt = torch.from_numpy(w).float().to(cuda_dev)
    L = (wt * (out.abs() ** 2).sum(0)).sum()
    (gz,) = torch.autograd.grad(L, z)

    def oracle_loss(zz):
        lx, ly, _ = R.LCD(ex, ey, None, eta, th)
        (vx, vy, vz), _ = R.vrs_propagation(lx, ly, x, y, 0.65, zz)
        return float(np.sum(w * (np.abs(vx) ** 2 + np.abs(vy) ** 2 + np.abs(vz) ** 2)))

    eps = 1e-3
    fd = (oracle_loss(z0 + eps) - oracle_loss(z0 - eps)) / (2 * eps)
    assert abs(float(gz) - fd) <= 2e-3 * abs(fd)


@pytest.mark.parametrize("form", ["bluestein", "gemm"])
def test_captured_high_na_step_matches_eager(cuda_dev, form):
    """A HIP-graph capture of the high-NA forward + backward: every replay
    (with parameter updates and eager steps in between) matches the eager step."""
    from xlumina_amd import ops, workloads
    h = workloads.high_na_2048(cuda_dev, n=256, m_out=64)
    geom = h.geometry()
    prev = ops.get_czt_form()
    ops.set_czt_form(form)
    try:
        alpha = h.alpha.clone().requires_grad_(True)
        phi = h.phi.clone().requires_grad_(True)
        step = workloads.CapturedStep(lambda: h.loss(alpha, phi, geom), [alpha, phi])
        for it in range(3):
            L, (ga, gp) = step.replay()
            L, ga, gp = float(L.detach()), ga.clone(), gp.clone()
            a2 = alpha.detach().clone().requires_grad_(True)
            p2 = phi.detach().clone().requires_grad_(True)
            Le = h.loss(a2, p2, geom)
            ea, ep = torch.autograd.grad(Le, (a2, p2))
            torch.cuda.synchronize()
            assert abs(L - float(Le)) <= 1e-6 * abs(float(Le)), (it, L, float(Le))
            assert float((ga - ea).norm() / ea.norm()) <= 1e-6 and float((gp - ep).norm() / ep.norm()) <= 1e-6
            with torch.no_grad():
                alpha.add_(0.05 * ea.sign())
                phi.sub_(0.03)
    finally:
        ops.set_czt_form(prev)


def test_dz_dot_vs_torch_reference(cuda_dev):
    """xl_dz_dot (the reduction of the distance gradient) against the same sums
    written in plain torch float32 / float64: the field term alone (RS) and
    with the VRS Ez_in term, z from the host and from device memory."""
    from xlumina_amd import ops
    rng = np.random.default_rng(91)
    rows, cols = 96, 80
    x = np.linspace(-300.0, 300.0, cols)
    y = np.linspace(-250.0, 250.0, rows)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    d, g, pre, gez = c64(3, rows, cols), c64(3, rows, cols), c64(2, rows, cols), c64(rows, cols)
    z = 850.0
    base = torch.real(torch.vdot(d.reshape(-1), g.reshape(-1))).double()
    X = torch.tensor(x, device=cuda_dev).float()[None, :]
    Y = torch.tensor(y, device=cuda_dev).float()[:, None]
    w = -z * torch.rsqrt(X * X + Y * Y + z * z) ** 3
    q = torch.real(gez.conj() * pre[0]) * (X * w) + torch.real(gez.conj() * pre[1]) * (Y * w)
    ref_vrs = base + q.sum(dtype=torch.float64)
    got = ops.dz_dot(d, g)
    assert abs(float(got) - float(base)) <= 1e-5 * float(torch.linalg.vector_norm(d) * torch.linalg.vector_norm(g))
    got_vrs = ops.dz_dot(d, g, pre=pre, gez=gez, x=x, y=y, z=z)
    scale = float(torch.linalg.vector_norm(d) * torch.linalg.vector_norm(g)) + float(q.abs().sum())
    assert abs(float(got_vrs) - float(ref_vrs)) <= 1e-5 * scale
    zt = torch.tensor([z], dtype=torch.float64, device=cuda_dev)
    got_dev = ops.dz_dot(d, g, pre=pre, gez=gez, x=x, y=y, z=0.0, z_dev=zt)
    assert float(got_dev) == float(got_vrs)
    assert float(ops.dz_dot(d, g)) == float(got)                 # fixed-order sums: repeatable bit for bit


@pytest.mark.parametrize("n,ez", [(64, N.XL_EZ_NONE), (96, N.XL_EZ_VRS)])
def test_stored_kernel_spectrum_reuse(cuda_dev, n, ez):
    """A forward that keeps the kernel spectrum in a caller buffer (hhat_mode 1)
    and an adjoint that reuses it (mode 2, conjugated on read) give exactly
    the results of building the spectrum in the workspace each time."""
    rng = np.random.default_rng(92)
    x, y = R.space(200.0, n)
    nin = 1 if ez == N.XL_EZ_NONE else 2
    src = torch.tensor(rng.standard_normal((nin, n, n)) + 1j * rng.standard_normal((nin, n, n)),
                       dtype=torch.complex64, device=cuda_dev)
    ncomp = 3 if ez == N.XL_EZ_VRS else 1
    g = torch.tensor(rng.standard_normal((ncomp, n, n)) + 1j * rng.standard_normal((ncomp, n, n)),
                     dtype=torch.complex64, device=cuda_dev)
    hhat = torch.empty(ops.rs_hhat_bytes(n), dtype=torch.uint8, device=cuda_dev)
    fwd0 = ops.rs_propagate(src, x, y, 700.0, 0.6, ez_mode=ez)
    fwd1 = ops.rs_propagate(src, x, y, 700.0, 0.6, ez_mode=ez, hhat=hhat, hhat_mode=1)
    assert torch.equal(fwd0, fwd1)
    adj0 = ops.rs_propagate(g, x, y, 700.0, 0.6, adjoint=True)
    adj2 = ops.rs_propagate(g, x, y, 700.0, 0.6, adjoint=True, hhat=hhat, hhat_mode=2)
    assert torch.equal(adj0, adj2)
    fwd2 = ops.rs_propagate(src, x, y, 700.0, 0.6, ez_mode=ez, hhat=hhat, hhat_mode=2)
    assert torch.equal(fwd0, fwd2)
    with pytest.raises(Exception):
        ops.rs_propagate(src, x, y, 700.0, 0.6, ez_mode=ez, hhat=hhat, hhat_mode=1, kernel_dz=True)


@pytest.mark.parametrize("form", ["bluestein", "gemm"])
def test_czt_grad_wrt_z_vs_fd(cuda_dev, form):
    """dL/dz of ScalarLight.CZT after an SLM (the distance enters the
    transfer functions on both grids, the constant z dx dy lambda and both
    axes' kernels through Dm = lambda z / dx) against central differences of
    the float64 oracle CZT."""
    from xlumina_amd import optical_elements as OE
    from xlumina_amd.wave_optics import ScalarLight
    n, m = 64, 40
    x, y = R.space(150.0, n)
    xo, yo = R.space(40.0, m)
    rng = np.random.default_rng(93)
    f = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ph = rng.uniform(-np.pi, np.pi, (n, n))
    w = rng.random((m, m))
    z0 = 2500.0
    prev = ops.get_czt_form()
    ops.set_czt_form(form)
    try:
        light = ScalarLight(x, y, 0.6, cuda_dev)
        light.field = f.astype(np.complex64)
        zt = torch.tensor([z0], dtype=torch.float64, device=cuda_dev, requires_grad=True)
        lit, _ = OE.SLM(light, ph, n)
        out = lit.CZT(zt, xo, yo)
        L = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) * out.field.abs() ** 2).sum()
        (gz,) = torch.autograd.grad(L, (zt,))
    finally:
        ops.set_czt_form(prev)

    def ref(z):
        return float(np.sum(w * np.abs(R.czt(f * np.exp(1j * ph), x, y, 0.6, z, xo, yo)) ** 2))

    h = 0.05
    fd = (ref(z0 + h) - ref(z0 - h)) / (2 * h)
    assert abs(float(gz) - fd) <= 2e-3 * abs(fd), (float(gz), fd)


def test_vczt_grad_wrt_z_vs_fd(cuda_dev):
    """dL/dz of VectorizedLight.VCZT after a per-pixel LCD, including the
    z-dependence of the constructed Ez_in = (Ex X + Ey Y) z / r^2, against
    central differences of the float64 oracle."""
    from xlumina_amd import optical_elements as OE
    from xlumina_amd.vectorized_optics import VectorizedLight
    n, m = 64, 40
    x, y = R.space(150.0, n)
    xo, yo = R.space(40.0, m)
    rng = np.random.default_rng(94)
    ex, ey = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    eta, th = rng.uniform(0, 2 * np.pi, (2, n, n))
    w = rng.random((3, m, m))
    z0 = 900.0
    light = VectorizedLight(x, y, 0.6, cuda_dev)
    light.Ex = torch.tensor(ex.astype(np.complex64), device=cuda_dev)
    light.Ey = torch.tensor(ey.astype(np.complex64), device=cuda_dev)
    zt = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
    out = OE.LCD(light, eta, th).VCZT(zt, xo, yo)
    E = torch.stack([out.Ex, out.Ey, out.Ez])
    L = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) * E.abs() ** 2).sum()
    (gz,) = torch.autograd.grad(L, (zt,))

    def ref(z):
        lx, ly, _ = R.LCD(ex, ey, None, eta, th)
        o = R.vczt(lx, ly, x, y, 0.6, z, xo, yo)
        return float(sum(np.sum(w[c] * np.abs(o[c]) ** 2) for c in range(3)))

    h = 0.02
    fd = (ref(z0 + h) - ref(z0 - h)) / (2 * h)
    assert abs(float(gz) - fd) <= 2e-3 * abs(fd), (float(gz), fd)


@pytest.mark.parametrize("n,vectorial", [(20, False), (64, True), (300, False), (1024, True), (2048, False)])
def test_rs_spectral_grads_match_derivative_propagation(cuda_dev, n, vectorial):
    """The z and k gradients of the adjoint's column pass (Parseval over the
    kept input spectra, spec_mode 1 / 2) against the derivative propagation
    (kernel_dz 1 / 2) dotted with the same cotangent by xl_dz_dot; the forward
    and adjoint fields are unchanged by keeping / dotting the spectra.  Sizes
    cover P = 64 (16 columns per block) up to 4096; z > 0 and z < 0."""
    x, y = R.space(3.0 * n, n)
    rng = np.random.default_rng(1000 + n)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    nin = 2 if vectorial else 1
    ncomp = 3 if vectorial else 1
    src = c64(nin, n, n)
    p0 = torch.tensor(rng.uniform(-np.pi, np.pi, (n, n)), dtype=torch.float32, device=cuda_dev)
    p1 = torch.tensor(rng.uniform(-np.pi, np.pi, (n, n)), dtype=torch.float32, device=cuda_dev)
    masks = [MaskSpec(N.XL_MASK_LCD, p0, p1)] if vectorial else [MaskSpec(N.XL_MASK_SLM, p0)]
    ez = N.XL_EZ_VRS if vectorial else N.XL_EZ_NONE
    g = c64(ncomp, n, n)
    for z in (40.0 * n, -25.0 * n):
        kw = dict(masks=masks, ez_mode=ez, nin=nin)
        hhat = torch.empty(ops.rs_hhat_bytes(n), dtype=torch.uint8, device=cuda_dev)
        spec = torch.empty(ops.rs_spec_bytes(n, ncomp), dtype=torch.uint8, device=cuda_dev)
        out = ops.rs_propagate(src, x, y, z, 0.6, hhat=hhat, hhat_mode=1, spec=spec, spec_mode=1, **kw)
        plain = ops.rs_propagate(src, x, y, z, 0.6, **kw)
        assert float((out - plain).norm() / plain.norm()) <= 1e-6
        grads = torch.full((2,), float("nan"), dtype=torch.float64, device=cuda_dev)
        gp = ops.rs_propagate(g, x, y, z, 0.6, adjoint=True, hhat=hhat, hhat_mode=2, spec=spec, spec_mode=2,
                              spec_grads=3, grad_out=grads)
        plain = ops.rs_propagate(g, x, y, z, 0.6, adjoint=True, hhat=hhat, hhat_mode=2)
        assert float((gp - plain).norm() / plain.norm()) <= 1e-6
        for i, kd in enumerate((1, 2)):
            d = ops.rs_propagate(src, x, y, z, 0.6, kernel_dz=kd, **kw)
            ref = float(ops.dz_dot(d, g))
            scale = float(torch.linalg.vector_norm(d) * torch.linalg.vector_norm(g))
            assert abs(float(grads[i]) - ref) <= 2e-5 * scale, (n, z, kd, float(grads[i]), ref, scale)
        # one derivative at a time leaves the other slot alone
        one = torch.full((2,), 7.0, dtype=torch.float64, device=cuda_dev)
        ops.rs_propagate(g, x, y, z, 0.6, adjoint=True, hhat=hhat, hhat_mode=2, spec=spec, spec_mode=2,
                         spec_grads=2, grad_out=one)
        assert float(one[0]) == 7.0 and float(one[1]) == float(grads[1])
        # derivative spectra built by the forward call from its own kernel rows
        # (each subset of {dH/dz, dH/dk}), taken by the adjoint as prebuilt
        for bits in (3, 1, 2):
            spec2 = torch.empty_like(spec)
            hhat2 = torch.empty_like(hhat)
            out2 = ops.rs_propagate(src, x, y, z, 0.6, hhat=hhat2, hhat_mode=1, spec=spec2, spec_mode=1,
                                    spec_grads=bits, **kw)
            assert float((out2 - out).norm() / out.norm()) <= 1e-6
            pre = torch.full((2,), 7.0, dtype=torch.float64, device=cuda_dev)
            ops.rs_propagate(g, x, y, z, 0.6, adjoint=True, hhat=hhat2, hhat_mode=2, spec=spec2, spec_mode=2,
                             spec_grads=bits | N.XL_RS_SPEC_PREBUILT, grad_out=pre)
            for i in range(2):
                if bits >> i & 1:
                    scale = abs(float(grads[i])) + 1e-30
                    assert abs(float(pre[i]) - float(grads[i])) <= 1e-6 * scale, (n, z, bits, i, float(pre[i]))
                else:
                    assert float(pre[i]) == 7.0
            # XL_RS_SPEC_PARTIALS: the per-unit partial sums stay in the spec
            # buffer (grad_out untouched); their sum is the reduced value
            ops.rs_propagate(g, x, y, z, 0.6, adjoint=True, hhat=hhat2, hhat_mode=2, spec=spec2, spec_mode=2,
                             spec_grads=bits | N.XL_RS_SPEC_PREBUILT | N.XL_RS_SPEC_PARTIALS, grad_out=None)
            for i in range(2):
                if bits >> i & 1:
                    part = float(ops.rs_spec_partials(spec2, n, ncomp, i).sum())
                    assert abs(part - float(pre[i])) <= 1e-9 * abs(float(pre[i])), (n, z, bits, i, part)


def test_dz_dot_ez_term_alone(cuda_dev):
    """xl_dz_dot with no propagated planes (the VRS Ez_in term of the
    distance gradient alone) equals the full sum minus its field term."""
    rng = np.random.default_rng(97)
    rows, cols = 70, 90
    x, y = np.linspace(-200.0, 200.0, cols), np.linspace(-150.0, 150.0, rows)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    d, g, pre, gez = c64(3, rows, cols), c64(3, rows, cols), c64(2, rows, cols), c64(rows, cols)
    full = float(ops.dz_dot(d, g, pre=pre, gez=gez, x=x, y=y, z=500.0))
    alone = float(ops.dz_dot(None, None, pre=pre, gez=gez, x=x, y=y, z=500.0))
    field = float(ops.dz_dot(d, g))
    scale = float(torch.linalg.vector_norm(d) * torch.linalg.vector_norm(g)) + abs(alone)
    assert abs(alone + field - full) <= 1e-6 * scale


@pytest.mark.parametrize("vectorial", [False, True])
def test_rs_grad_wrt_wavelength_vs_fd(cuda_dev, vectorial):
    """dL/dlambda of RS / VRS (after a per-pixel mask): the convolution with
    dH/dk times dk/dlambda = -k/lambda, against central differences of the
    float64 oracle in the wavelength."""
    n = 64
    x, y = R.space(120.0, n)
    rng = np.random.default_rng(95)
    ex, ey = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    a, b = rng.uniform(-np.pi, np.pi, (2, n, n))
    w = rng.random((3 if vectorial else 1, n, n))
    wl0, z = 0.6, 600.0
    wl = torch.tensor([wl0], dtype=torch.float64, device=cuda_dev, requires_grad=True)
    if vectorial:
        src = torch.tensor(np.stack([ex, ey]).astype(np.complex64), device=cuda_dev)
        masks = (MaskSpec(N.XL_MASK_LCD, torch.tensor(a, dtype=torch.float32, device=cuda_dev),
                          torch.tensor(b, dtype=torch.float32, device=cuda_dev)),)
        out = AG.rs(src, x, y, z, wl, masks=masks, ez_mode=N.XL_EZ_VRS)
    else:
        src = torch.tensor(ex[None].astype(np.complex64), device=cuda_dev)
        masks = (MaskSpec(N.XL_MASK_SLM, torch.tensor(a, dtype=torch.float32, device=cuda_dev)),)
        out = AG.rs(src, x, y, z, wl, masks=masks)
    L = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) * out.abs() ** 2).sum()
    (g,) = torch.autograd.grad(L, (wl,))

    def ref(lam):
        if vectorial:
            lx, ly, _ = R.LCD(ex, ey, None, a, b)
            (ox, oy, oz), _ = R.vrs_propagation(lx, ly, x, y, lam, z)
            o = np.stack([ox, oy, oz])
        else:
            o = R.rs_propagation(ex * np.exp(1j * a), x, y, lam, z)[0][None]
        return float(np.sum(w * np.abs(o) ** 2))

    h = 1e-6
    fd = (ref(wl0 + h) - ref(wl0 - h)) / (2 * h)
    assert abs(float(g) - fd) <= 2e-3 * abs(fd), (float(g), fd)


@pytest.mark.parametrize("form,vectorial", [("bluestein", False), ("gemm", False), ("bluestein", True)])
def test_czt_grad_wrt_wavelength_vs_fd(cuda_dev, form, vectorial):
    """dL/dlambda of CZT / VCZT after a per-pixel mask: lambda enters the
    transfer functions on both grids (k = 2 pi / lambda), the constant
    z dx dy lambda and both axes' kernels through Dm = lambda z / dx (not the
    VCZT Ez_in); against central differences of the float64 oracle in the
    wavelength (step 3e-8 um: the phase rate in lambda is ~k z / lambda, so a
    wider step is dominated by truncation)."""
    n, m = 64, 40
    x, y = R.space(150.0, n)
    xo, yo = R.space(40.0, m)
    rng = np.random.default_rng(96)
    ex, ey = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    a, b = rng.uniform(-np.pi, np.pi, (2, n, n))
    w = rng.random((3 if vectorial else 1, m, m))
    wl0, z = 0.6, 2500.0
    wl = torch.tensor([wl0], dtype=torch.float64, device=cuda_dev, requires_grad=True)
    geom = ops.czt_geometry(x, y, xo, yo, z, wl0)
    if vectorial:
        src = torch.tensor(np.stack([ex, ey]).astype(np.complex64), device=cuda_dev)
        masks = (MaskSpec(N.XL_MASK_LCD, torch.tensor(a, dtype=torch.float32, device=cuda_dev),
                          torch.tensor(b, dtype=torch.float32, device=cuda_dev)),)
        out = AG.czt(src, geom, masks=masks, ez_mode=N.XL_EZ_VCZT, nin=2, form=form, wavelength=wl)
    else:
        src = torch.tensor(ex[None].astype(np.complex64), device=cuda_dev)
        masks = (MaskSpec(N.XL_MASK_SLM, torch.tensor(a, dtype=torch.float32, device=cuda_dev)),)
        out = AG.czt(src, geom, masks=masks, form=form, wavelength=wl)
    L = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) * out.abs() ** 2).sum()
    (g,) = torch.autograd.grad(L, (wl,))

    def ref(lam):
        if vectorial:
            lx, ly, _ = R.LCD(ex, ey, None, a, b)
            o = np.stack(R.vczt(lx, ly, x, y, lam, z, xo, yo))
        else:
            o = R.czt(ex * np.exp(1j * a), x, y, lam, z, xo, yo)[None]
        return float(np.sum(w * np.abs(o) ** 2))

    h = 3e-8
    fd = (ref(wl0 + h) - ref(wl0 - h)) / (2 * h)
    assert abs(float(g) - fd) <= 2e-3 * abs(fd), (float(g), fd)


@pytest.mark.parametrize("stages", [0, 1, 2])
def test_pre_backward_ez_distance_term(cuda_dev, stages):
    """The VRS Ez_in distance term formed in the pre-op adjoint's pass (ez_dz)
    equals xl_dz_dot's `pre` term over the separately masked field, with no,
    one and two mask stages."""
    rng = np.random.default_rng(400 + stages)
    n = 90
    x = np.linspace(-150.0, 150.0, n)
    src = torch.tensor(rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n)), dtype=torch.complex64,
                       device=cuda_dev)
    g = torch.tensor(rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n)), dtype=torch.complex64,
                     device=cuda_dev)
    t = [torch.tensor(rng.uniform(-np.pi, np.pi, (n, n)), dtype=torch.float32, device=cuda_dev) for _ in range(4)]
    masks = [MaskSpec(N.XL_MASK_SSLM, t[0], t[1]), MaskSpec(N.XL_MASK_LCD, t[2], t[3])][:stages]
    z = 700.0
    buf = torch.full((ops.pre_scalar_grad_elems(n, n),), float("nan"), dtype=torch.float64, device=cuda_dev)
    ops.pre_backward(src, g, x, x, z, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2, ez_dz=buf)
    pre = ops.pre_forward(src, x, x, z, masks=masks, ncomp=2, nin=2) if masks else src
    ref = float(ops.dz_dot(None, None, pre=pre, gez=g[2], x=x, y=x, z=z))
    # sum of the terms' magnitudes: |g_Ez| (|Ex'| + |Ey'|) max|X w|, |X w| <= 150 z / z^3
    scale = float((g[2].abs() * pre[:2].abs().sum(0)).sum()) * 150.0 * z / z ** 3
    assert abs(float(buf[0]) - ref) <= 1e-5 * scale, (float(buf[0]), ref, scale)
    # with a base value (the propagation's own term) added by the same reduction
    base = torch.tensor([0.375], dtype=torch.float64, device=cuda_dev)
    ops.pre_backward(src, g, x, x, z, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2, ez_dz=buf, ez_dz_base=base)
    assert abs(float(buf[0]) - (ref + 0.375)) <= 1e-5 * scale + 1e-15


@pytest.mark.parametrize("float_z", [False, True])
def test_same_distance_propagations_share_spectra(cuda_dev, float_z):
    """Two differentiated VRS over one distance (one tensor, or one float):
    the second reuses the first's kernel spectrum and derivative spectra
    (xl_rs_params.spec_dg).  Outputs and every gradient equal those of two
    propagations over separate, equal distance tensors, which build their own;
    the distance gradient of the shared form is the sum of the two."""
    n = 96
    x, y = R.space(150.0, n)
    rng = np.random.default_rng(41)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    a, b = c64(2, n, n), c64(2, n, n)
    eta = torch.tensor(rng.uniform(0, 2 * np.pi, (n, n)), dtype=torch.float32, device=cuda_dev)
    th = torch.tensor(rng.uniform(0, 2 * np.pi, (n, n)), dtype=torch.float32, device=cuda_dev)
    w = torch.tensor(rng.random((2, n, n)), dtype=torch.float32, device=cuda_dev)
    z0 = 700.0

    def run(z1, z2, share=True):
        AG._SHARED_SPECTRA.clear()
        sa, sb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
        e = eta.clone().requires_grad_(True)
        masks = [MaskSpec(N.XL_MASK_LCD, e, th)]
        o1 = AG.rs(sa, x, y, z1, 0.65, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2)
        if not share:
            AG._SHARED_SPECTRA.clear()   # the second builds its own spectra
        o2 = AG.rs(sb, x, y, z2, 0.65, masks=masks, ez_mode=N.XL_EZ_VRS, nin=2)
        entries = len(AG._SHARED_SPECTRA)
        L = (w[0] * (o1.abs() ** 2).sum(0)).sum() + (w[1] * (o2.abs() ** 2).sum(0)).sum()
        wrt = [sa, sb, e] + [t for t in (z1, z2) if isinstance(t, torch.Tensor)]
        grads = torch.autograd.grad(L, wrt)
        return o1.detach(), o2.detach(), grads, entries

    if float_z:
        s1, s2, gs, entries = run(z0, z0)
        u1, u2, gu, _ = run(z0, z0, share=False)
        assert entries == 1
        for p, q in zip(gs, gu):
            assert torch.equal(p, q)
    else:
        z = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
        s1, s2, gs, entries = run(z, z)
        assert entries == 1   # the second propagation found the first's spectra
        za = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
        zb = torch.tensor(z0, dtype=torch.float64, device=cuda_dev, requires_grad=True)
        u1, u2, gu, entries_u = run(za, zb)
        assert entries_u == 2   # separate tensors: each builds its own
        for p, q in zip(gs[:3], gu[:3]):
            assert torch.equal(p, q)
        ref = float(gu[3]) + float(gu[4])
        assert abs(float(gs[3]) - ref) <= 1e-12 * abs(ref), (float(gs[3]), ref)
    assert torch.equal(s1, u1) and torch.equal(s2, u2)


def test_czt_calls_of_one_geometry_share_staged_tables(cuda_dev):
    """Two differentiated high-NA VCZT over equal geometries (separate
    objects, as each lens call builds): the second forward and adjoint use
    the axis tables the first staged (tables_mode 2).  Outputs and gradients
    equal those of calls that stage their own."""
    n = 128
    x, y = R.space(1800.0, n)
    xo = np.linspace(-2, 2, 40)
    rng = np.random.default_rng(17)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    a, b = c64(2, n, n), c64(2, n, n)
    w = torch.tensor(rng.random((2, 40, 40)), dtype=torch.float32, device=cuda_dev)

    def run(share):
        AG._STAGED_TABLES.clear()
        sa, sb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
        o1 = AG.czt(sa, ops.hna_geometry(x, y, xo, xo, 0.65, 1800.0, 2000.0), nin=2, form="bluestein")
        if not share:
            AG._STAGED_TABLES.clear()
        o2 = AG.czt(sb, ops.hna_geometry(x, y, xo, xo, 0.65, 1800.0, 2000.0), nin=2, form="bluestein")
        L = (w[0] * (o1.abs() ** 2).sum(0)).sum() + (w[1] * (o2.abs() ** 2).sum(0)).sum()
        return o1.detach(), o2.detach(), torch.autograd.grad(L, (sa, sb))

    s1, s2, gs = run(True)
    u1, u2, gu = run(False)
    assert torch.equal(s1, u1) and torch.equal(s2, u2)
    assert torch.equal(gs[0], gu[0]) and torch.equal(gs[1], gu[1])
    # the shared run really shared: one live entry while both calls' state lived
    sa = a.clone().requires_grad_(True)
    AG._STAGED_TABLES.clear()
    o1 = AG.czt(sa, ops.hna_geometry(x, y, xo, xo, 0.65, 1800.0, 2000.0), nin=2, form="bluestein")
    o2 = AG.czt(sa, ops.hna_geometry(x, y, xo, xo, 0.65, 1800.0, 2000.0), nin=2, form="bluestein")
    assert len(AG._STAGED_TABLES) == 1
    share = next(iter(AG._STAGED_TABLES.values()))[1]()
    assert share is not None and share.fwd is not None
    del o1, o2


@pytest.mark.parametrize("mode", [1, 2, 1 | 4])
def test_splitter_fused_backward_matches_separate(cuda_dev, mode):
    """A splitter whose angle and input fields both need gradients forms them
    in one pass (xl_bs_dots with ga / gb): the same values as the separate
    adjoint (xl_beam_splitter) and angle-gradient (xl_bs_dots) passes,
    including the zeroed third plane.  Modes: symmetric, single input, and
    symmetric with a raw angle."""
    from xlumina_amd import optical_elements as OE
    n = 96
    rng = np.random.default_rng(53 + mode)

    def c64(*shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), dtype=torch.complex64,
                            device=cuda_dev)

    a0, b0 = c64(3, n, n), c64(3, n, n)
    th0 = torch.tensor([0.37], dtype=torch.float64, device=cuda_dev)
    w = torch.tensor(rng.random((2, 3, n, n)), dtype=torch.float32, device=cuda_dev)
    two = mode & 3 == 1

    def run(fields, angle):
        a = a0.clone().requires_grad_(fields)
        b = b0.clone().requires_grad_(fields) if two else None
        th = th0.clone().requires_grad_(angle)
        c, d = OE._SplitterTheta.apply(a, b, th, mode)
        L = (w[0] * c.abs() ** 2).sum() + (w[1] * d.abs() ** 2).sum()
        wrt = ([a] + ([b] if two else []) if fields else []) + ([th] if angle else [])
        return torch.autograd.grad(L, wrt)

    fused = run(True, True)
    fields_only = run(True, False)
    angle_only = run(False, True)
    for p, q in zip(fused[:-1], fields_only):
        assert torch.equal(p, q)
        assert float(p[2].abs().max()) == 0.0
    assert torch.equal(fused[-1], angle_only[0])


def test_three_component_plain_rs_rows_per_component(cuda_dev):
    """Three plain components at P = 2048 (n = 600) take the row pass with one
    block per (row, component): each output component against the float64
    oracle's scalar propagation, and the adjoint (the same kernel on the
    cotangent) through the dot-product identity."""
    n = 600
    x, y = R.space(900.0, n)
    rng = np.random.default_rng(71)
    u = rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))
    src = torch.from_numpy(u.astype(np.complex64)).to(cuda_dev)
    got = ops.rs_propagate(src, x, y, 5000.0, 0.65).cpu().numpy()
    for c in range(3):
        ref, _ = R.rs_propagation(u[c], x, y, 0.65, 5000.0)
        assert rel_l2(got[c], ref) <= 1e-5, c
    cot = torch.randn((3, n, n), dtype=torch.complex64, device=cuda_dev)
    assert dot_identity(lambda v: AG.rs(v, x, y, 5000.0, 0.65), src, cot) < 1e-5


def test_vrs_without_masks_at_p2048(cuda_dev):
    """Ex, Ey with the constructed Ez at P = 2048 (n = 600), no masks (the row
    pass keeps one block per row: the Ez construction couples the
    components).  All three output components against the float64 oracle."""
    n = 600
    x, y = R.space(900.0, n)
    rng = np.random.default_rng(73)
    ex, ey = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    src = torch.from_numpy(np.stack([ex, ey]).astype(np.complex64)).to(cuda_dev)
    got = ops.rs_propagate(src, x, y, 5000.0, 0.65, ez_mode=N.XL_EZ_VRS, nin=2).cpu().numpy()
    (vx, vy, vz), _ = R.vrs_propagation(ex, ey, x, y, 0.65, 5000.0)
    for c, ref in enumerate((vx, vy, vz)):
        assert rel_l2(got[c], ref) <= 1e-5, c
