"""HIP kernels vs the float64 oracle (GPU).

Tolerance: relative L2 <= 1e-5 against float64 for every FP32 path (the
kernels compute in FP32 with float64 phase generation).
"""
import numpy as np
import pytest
import torch

from oracle import brute as B
from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import ops

pytestmark = pytest.mark.gpu
TOL = 1e-5


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.complex64)).to(dev)


def H(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def F32(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def cfield(rng, n, m=None):
    m = n if m is None else m
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


# -- golden vectors ------------------------------------------------------------------------

def test_golden_rs(golden, cuda_dev):
    for name, c in golden.items():
        if not name.startswith("rs_"):
            continue
        out = ops.rs_propagate(T(c["field"], cuda_dev), c["x"], c["y"], float(c["z"]), float(c["wavelength"]))
        assert rel_l2(H(out[0]), c["out"]) < TOL, name


def test_golden_vrs_lcd(golden, cuda_dev):
    c = golden["vrs_lcd_n20"]
    src = T(np.stack([c["Ex"], c["Ey"]]), cuda_dev)
    m = ops.MaskSpec(N.XL_MASK_LCD, F32(c["eta"], cuda_dev), F32(c["theta"], cuda_dev))
    out = ops.rs_propagate(src, c["x"], c["y"], float(c["z"]), float(c["wavelength"]), masks=[m],
                           ez_mode=N.XL_EZ_VRS, nin=2)
    assert rel_l2(H(out), c["out"]) < TOL


def test_golden_czt(golden, cuda_dev):
    for name, c in golden.items():
        if not name.startswith("czt_"):
            continue
        g = ops.czt_geometry(c["x"], c["y"], c["xout"], c["yout"], float(c["z"]), float(c["wavelength"]))
        out = ops.czt(T(c["field"], cuda_dev), g)
        assert out.shape[1:] == c["out"].shape
        assert rel_l2(H(out[0]), c["out"]) < TOL, name


def test_golden_vczt_and_hna(golden, cuda_dev):
    c = golden["vczt_n24"]
    g = ops.czt_geometry(c["x"], c["y"], c["xout"], c["yout"], float(c["z"]), float(c["wavelength"]))
    out = ops.czt(T(np.stack([c["Ex"], c["Ey"]]), cuda_dev), g, ez_mode=N.XL_EZ_VCZT, nin=2)
    assert rel_l2(H(out), c["out"]) < TOL
    c = golden["hna_n32"]
    g = ops.hna_geometry(c["x"], c["y"], c["xout"], c["yout"], float(c["wavelength"]), float(c["radius"]),
                         float(c["f"]))
    src = T(np.stack([c["Ex"], c["Ey"]]), cuda_dev)
    out = ops.czt(src, g, nin=2)
    assert rel_l2(H(out), c["out"]) < TOL
    lens = ops.pre_forward(src, c["x"], c["y"], hna=True, nin=2, hna_radius=float(c["radius"]), hna_f=float(c["f"]))
    assert rel_l2(H(lens), c["lens"]) < TOL


# -- oracle at moderate and workload sizes --------------------------------------------------

@pytest.mark.parametrize("n,z", [(20, 150.0), (37, 400.0), (250, 2000.0), (300, 1500.0), (1000, 5e4), (1024, -3e4)])
def test_rs_vs_oracle(n, z, cuda_dev):
    rng = np.random.default_rng(n)
    x, y = R.space(1200.0 * n / 1024, n)
    u = R.gaussian_beam(x, y, 0.65, (400.0 * n / 1024, 400.0 * n / 1024), 1.0) * np.exp(
        1j * rng.uniform(-np.pi, np.pi, (n, n)))
    ref, _ = R.rs_propagation(u, x, y, 0.65, z)
    out, inten = ops.rs_propagate(T(u, cuda_dev), x, y, z, 0.65, intensity=True)
    assert rel_l2(H(out[0]), ref) < TOL
    assert rel_l2(inten.cpu().numpy(), np.abs(ref) ** 2) < 2 * TOL


@pytest.mark.parametrize("n", [20, 300])
def test_vrs_lcd_vs_oracle_small_and_1024_fft(n, cuda_dev):
    """VRS with a per-pixel LCD at FFT lengths 64 and 1024 (the smallest plan
    and the first that runs the one-row-per-block row passes)."""
    rng = np.random.default_rng(500 + n)
    x, y = R.space(2.0 * n, n)
    ex, ey = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    eta, th = rng.uniform(0, 2 * np.pi, (2, n, n))
    lx, ly, _ = R.LCD(ex, ey, None, eta, th)
    (ox, oy, oz), _ = R.vrs_propagation(lx, ly, x, y, 0.6, 3.0 * n)
    src = torch.tensor(np.stack([ex, ey]).astype(np.complex64), device=cuda_dev)
    m = ops.MaskSpec(N.XL_MASK_LCD, torch.tensor(eta, dtype=torch.float32, device=cuda_dev),
                     torch.tensor(th, dtype=torch.float32, device=cuda_dev))
    out = ops.rs_propagate(src, x, y, 3.0 * n, 0.6, masks=[m], ez_mode=N.XL_EZ_VRS, nin=2)
    assert rel_l2(out.cpu().numpy(), np.stack([ox, oy, oz])) < TOL


def test_config1_slm_fused(cuda_dev):
    """1024^2, x,y in [-1.2, 1.2] mm, lambda 650 nm, Gaussian w0 0.4 mm, seeded SLM phase, z = 50 mm."""
    n = 1024
    x, y = R.space(1200.0, n)
    u = R.gaussian_beam(x, y, 0.65, (400.0, 400.0), 1.0)
    phi = np.random.default_rng(1).uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    ref, qf = R.rs_propagation(R.slm(u, phi.astype(np.float64))[0], x, y, 0.65, 5e4)
    m = ops.MaskSpec(N.XL_MASK_SLM, F32(phi, cuda_dev))
    out, inten = ops.rs_propagate(T(u, cuda_dev), x, y, 5e4, 0.65, masks=[m], intensity=True)
    assert rel_l2(H(out[0]), ref) < TOL
    assert ops.quality_factor(x, y, 0.65, 5e4) == pytest.approx(qf, rel=1e-14)


def test_vrs_masks_chain_and_lp(cuda_dev):
    rng = np.random.default_rng(3)
    n = 128
    x, y = R.space(300.0, n)
    Ex, Ey = cfield(rng, n), cfield(rng, n)
    a, p = rng.uniform(-np.pi, np.pi, (2, n, n)).astype(np.float32)
    eta, th = 0.7, 1.9
    sx, sy, sz = R.sSLM(Ex, Ey, None, a.astype(np.float64), p.astype(np.float64))
    lx, ly, _ = R.LCD(sx, sy, None, eta, th)
    (ox, oy, oz), _ = R.vrs_propagation(lx, ly, x, y, 0.65, 3000.0)
    masks = [ops.MaskSpec(N.XL_MASK_SSLM, F32(a, cuda_dev), F32(p, cuda_dev)),
             ops.MaskSpec(N.XL_MASK_LCD_SCALAR, None, None, eta, th)]
    out = ops.rs_propagate(T(np.stack([Ex, Ey]), cuda_dev), x, y, 3000.0, 0.65, masks=masks, ez_mode=N.XL_EZ_VRS)
    assert rel_l2(H(out), np.stack([ox, oy, oz])) < TOL
    # polariser: masked Ex/Ey, Ez -> 0 (materialised by the standalone pre-op)
    alpha = rng.uniform(0, np.pi, (n, n)).astype(np.float32)
    px, py, pz = R.linear_polarizer(Ex, Ey, Ex, alpha.astype(np.float64))
    got = ops.pre_forward(T(np.stack([Ex, Ey, Ex]), cuda_dev), x, y, masks=[ops.MaskSpec(N.XL_MASK_LP,
                                                                                         F32(alpha, cuda_dev))])
    assert rel_l2(H(got[:2]), np.stack([px, py])) < TOL and float(got[2].abs().max()) == 0.0


@pytest.mark.parametrize("n,mx,my", [(1024, 512, 512), (500, 123, 77), (64, 64, 64), (40, 70, 50)])
def test_czt_vs_oracle(n, mx, my, cuda_dev):
    rng = np.random.default_rng(mx)
    x, y = R.space(1200.0, n)
    u = cfield(rng, n)
    xo = np.linspace(-100, 100, mx)
    yo = np.linspace(-80, 120, my)
    ref = R.czt(u, x, y, 0.635, 3e4, xo, yo)
    out = ops.czt(T(u, cuda_dev), ops.czt_geometry(x, y, xo, yo, 3e4, 0.635))
    assert rel_l2(H(out[0]), ref) < TOL


@pytest.mark.parametrize("n,mx,my", [(1100, 200, 300), (700, 60, 100), (2000, 600, 700)])
def test_czt_three_smooth_lengths(n, mx, my, cuda_dev):
    """Bluestein lengths 3 * 2^k (768, 1536, 3072; chosen when shorter than the
    next power of two) give the same operator as the oracle's power-of-two
    length, forward and through the adjoint identity."""
    rng = np.random.default_rng(n)
    x, y = R.space(1200.0, n)
    u = cfield(rng, n)
    xo = np.linspace(-100, 100, mx)
    yo = np.linspace(-80, 120, my)
    g = ops.czt_geometry(x, y, xo, yo, 3e4, 0.635)
    ref = R.czt(u, x, y, 0.635, 3e4, xo, yo)
    out = ops.czt(T(u, cuda_dev), g, form="bluestein")
    assert rel_l2(H(out[0]), ref) < TOL
    c = torch.randn((1, my, mx), dtype=torch.complex64, device=cuda_dev)
    back = ops.czt(c, g, adjoint=True, form="bluestein")
    lhs = complex(torch.vdot(c.flatten().to(torch.complex128), out.flatten().to(torch.complex128)))
    rhs = complex(torch.vdot(back.flatten().to(torch.complex128), T(u, cuda_dev).flatten().to(torch.complex128)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("n,my", [(512, 257), (1024, 513), (2048, 1025)])
def test_czt_length_equal_to_support(n, my, cuda_dev):
    """Column axis whose Bluestein length 3 * 2^k equals m + M - 1 exactly (no
    spare sample on the circle: every lag of the kernel is used once), with a
    row axis in the zero-half form (input and output in half a power-of-two
    length); forward against the oracle, adjoint through the inner-product
    identity."""
    rng = np.random.default_rng(my)
    x, y = R.space(1200.0, n)
    u = cfield(rng, n)
    xo = np.linspace(-100, 100, 100)
    yo = np.linspace(-80, 120, my)
    g = ops.czt_geometry(x, y, xo, yo, 3e4, 0.635)
    ref = R.czt(u, x, y, 0.635, 3e4, xo, yo)
    out = ops.czt(T(u, cuda_dev), g, form="bluestein")
    assert rel_l2(H(out[0]), ref) < TOL
    # the last output row is the one a layout one sample short would miss
    assert rel_l2(H(out[0])[-1], ref[-1]) < TOL
    c = torch.randn((1, my, 100), dtype=torch.complex64, device=cuda_dev)
    back = ops.czt(c, g, adjoint=True, form="bluestein")
    lhs = complex(torch.vdot(c.flatten().to(torch.complex128), out.flatten().to(torch.complex128)))
    rhs = complex(torch.vdot(back.flatten().to(torch.complex128), T(u, cuda_dev).flatten().to(torch.complex128)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_czt_rejections(cuda_dev):
    x, y = R.space(100.0, 64)
    with pytest.raises(ValueError, match="power of two"):
        ops.czt(T(np.ones((64, 64)), cuda_dev), ops.czt_geometry(x, y, np.linspace(-1, 1, 65), y, 1e3, 0.65))
    with pytest.raises(ValueError):
        ops.czt(T(np.ones((64, 64)), cuda_dev), ops.czt_geometry(x, y, np.array([0.0]), y, 1e3, 0.65))


def test_config2_full_size(cuda_dev):
    """2048^2 radial beam, lambda 635 nm, window 2.4 mm, seeded per-pixel LCD, VRS z = 30 mm,
    then VCZT to a 512^2 ROI of 0.2 mm (full workload size against float64)."""
    n = 2048
    x, y = R.space(1200.0, n)
    Ex, Ey = R.radial_beam(x, y, 1200.0)
    eta, th = np.random.default_rng(2).uniform(0, 2 * np.pi, (2, n, n)).astype(np.float32)
    mx_, my_, _ = R.LCD(Ex, Ey, None, eta.astype(np.float64), th.astype(np.float64))
    (vx, vy, vz), _ = R.vrs_propagation(mx_, my_, x, y, 0.635, 3e4)
    xo = np.linspace(-100.0, 100.0, 512)
    ref_czt = np.stack(R.vczt(vx, vy, x, y, 0.635, 3e4, xo, xo))
    src = T(np.stack([Ex, Ey]), cuda_dev)
    m = ops.MaskSpec(N.XL_MASK_LCD, F32(eta, cuda_dev), F32(th, cuda_dev))
    vrs = ops.rs_propagate(src, x, y, 3e4, 0.635, masks=[m], ez_mode=N.XL_EZ_VRS, nin=2)
    assert rel_l2(H(vrs), np.stack([vx, vy, vz])) < TOL
    out = ops.czt(vrs, ops.czt_geometry(x, y, xo, xo, 3e4, 0.635), ez_mode=N.XL_EZ_VCZT, nin=2)
    assert rel_l2(H(out), ref_czt) < TOL


def test_high_na_odd_grid_nan_like_reference(cuda_dev):
    # odd N puts a sample at rho = 0: the reference's Ez = .../rho is NaN there
    n = 33
    x, y = R.space(1800.0, n)
    Ex = np.ones((n, n), complex)
    src = T(np.stack([Ex, Ex]), cuda_dev)
    lens = ops.pre_forward(src, x, y, hna=True, nin=2, hna_radius=1800.0, hna_f=2000.0)
    ref, _ = R.high_NA_objective_lens(Ex, Ex, x, y, 1800.0, 2000.0)
    assert np.array_equal(np.isnan(H(lens)), np.isnan(ref))


# -- full-size properties --------------------------------------------------------------------------

def test_linearity_and_adjoint_at_2048(cuda_dev):
    n = 2048
    rng = np.random.default_rng(11)
    x, y = R.space(1200.0, n)
    a = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    b = torch.randn((1, n, n), dtype=torch.complex64, device=cuda_dev)
    A = lambda u: ops.rs_propagate(u, x, y, 3e4, 0.635)
    lhs = A(2 * a - 3j * b)
    rhs = 2 * A(a) - 3j * A(b)
    assert float((lhs - rhs).norm() / rhs.norm()) < TOL
    # <A a, b> = <a, A^H b>
    AH = lambda u: ops.rs_propagate(u, x, y, 3e4, 0.635, adjoint=True)
    l = torch.vdot(A(a).flatten().to(torch.complex128), b.flatten().to(torch.complex128))
    r = torch.vdot(a.flatten().to(torch.complex128), AH(b).flatten().to(torch.complex128))
    assert abs(complex(l - r)) / abs(complex(l)) < 1e-4
    xo = np.linspace(-100.0, 100.0, 512)
    g = ops.czt_geometry(x, y, xo, xo, 3e4, 0.635)
    c = torch.randn((1, 512, 512), dtype=torch.complex64, device=cuda_dev)
    l = torch.vdot(ops.czt(a, g).flatten().to(torch.complex128), c.flatten().to(torch.complex128))
    r = torch.vdot(a.flatten().to(torch.complex128), ops.czt(c, g, adjoint=True).flatten().to(torch.complex128))
    assert abs(complex(l - r)) / abs(complex(l)) < 1e-4


def test_energy_conservation_rs(cuda_dev):
    # toolbox.is_conserving_energy ~ 1 for a beam well inside the window when the
    # sampled kernel is resolved (quality factor > 1; at z = 50 mm it is 5.7)
    n = 1024
    x, y = R.space(1200.0, n)
    u = R.gaussian_beam(x, y, 0.65, (200.0, 200.0), 1.0)
    assert ops.quality_factor(x, y, 0.65, 5e4) > 1
    out = ops.rs_propagate(T(u, cuda_dev), x, y, 5e4, 0.65)
    ratio = float((out.abs() ** 2).sum() / (torch.from_numpy(np.abs(u) ** 2).sum()))
    assert abs(ratio - 1) < 1e-3


def test_transfer_function(cuda_dev):
    x, y = R.space(100.0, 40)
    X, Y = np.meshgrid(x, y)
    got = ops.transfer_function_rs(x, y, 500.0, 0.65, cuda_dev)
    assert rel_l2(H(got), R.transfer_function_RS(500.0, X, Y, 2 * np.pi / 0.65)) < TOL
