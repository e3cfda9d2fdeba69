"""Host-side logic of the drop-in API (CPU only: no kernels run here)."""
import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from xlumina_amd import _native as N
from xlumina_amd import ops, toolbox
from xlumina_amd import wave_optics as W
from xlumina_amd import optical_elements as OE
from xlumina_amd.vectorized_optics import VectorizedLight, PolarizedLightSource


def test_units_and_space():
    import xlumina_amd as X
    assert (X.um, X.nm, X.mm, X.cm) == (1, 1e-3, 1e3, 1e4)
    x, y = toolbox.space(1200.0, 7)
    assert np.array_equal(x, R.space(1200.0, 7)[0]) and x.dtype == np.float64


def test_build_grid_matches_oracle():
    x, y = R.space(50.0, 9)
    a = W.build_grid(x, y)
    b = R.build_grid(x, y)
    assert a[:4] == b[:4]
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])
    assert np.allclose(W._grid_from_lags(a[4][0], 9), x - x[0])


def test_czt_grid_and_geometry():
    x, y = R.space(100.0, 32)
    xo = np.linspace(-20, 20, 10)
    yo = np.linspace(-10, 15, 12)
    a = W.build_CZT_grid(1000.0, 0.65, x, y, xo, yo)
    b = R.build_CZT_grid(1000.0, 0.65, x, y, xo, yo)
    for u, v in zip(a, b):
        assert np.allclose(u, v)
    g = ops.czt_geometry(x, y, xo, yo, 1000.0, 0.65)
    assert g.dm == b[6] and g.fy == (b[7], b[8]) and g.fx == (b[9], b[10])
    assert g.cte == pytest.approx(1000.0 * (x[1] - x[0]) * (y[1] - y[0]) * 0.65)
    h = ops.hna_geometry(x, y, xo, yo, 0.65, 1800.0, 2000.0)
    nb = R.build_high_NA_VCZT_grid(2000.0, 1800.0, 0.65, x, xo, yo)
    assert h.dm == nb[2] and h.fx == (nb[5], nb[6])
    stm = 1800.0 / np.hypot(1800.0, 2000.0)
    assert h.cte == pytest.approx(-1j * stm ** 2 / (2000.0 * 0.65))
    assert W.compute_np2(2559) == 4096


def test_quality_factor_matches_oracle():
    x, y = R.space(1200.0, 1024)
    assert ops.quality_factor(x, y, 0.65, 5e4) == pytest.approx(R.quality_factor(x, y, 0.65, 5e4), rel=1e-14)
    assert W.minimum_z(x, y, 0.65) == pytest.approx(R.minimum_z(x, y, 0.65), rel=1e-14)


def test_lights_on_cpu_keep_reference_attributes():
    x, y = R.space(100.0, 16)
    l = W.LightSource(x, y, 0.633, device="cpu")
    assert l.info == "Wave optics light source" and l.n == 1 and l.k == pytest.approx(2 * np.pi / 0.633)
    l.gaussian_beam(w0=(40.0, 40.0), E0=1)
    ref = R.gaussian_beam(x, y, 0.633, (40.0, 40.0), 1.0)
    assert np.allclose(l.field.numpy(), ref, atol=1e-7)
    v = PolarizedLightSource(x, y, 0.633, device="cpu")
    v.gaussian_beam(w0=(40.0, 40.0), jones_vector=(1, 1))
    ex, ey, ez = R.polarized_gaussian_beam(x, y, 0.633, (40.0, 40.0), (1, 1))
    assert np.allclose(v.Ex.numpy(), ex, atol=1e-7) and np.allclose(v.Ey.numpy(), ey, atol=1e-7)
    assert v.Ez.abs().max() == 0
    assert isinstance(VectorizedLight(x, y, 0.633, device="cpu").Ex, torch.Tensor)


def test_jones_matrices_match_oracle():
    for a, b in ((0.3, 1.2), (np.pi / 2, np.pi / 4)):
        assert np.allclose(OE.jones_LP(a).numpy(), R.jones_LP(a))
        assert np.allclose(OE.jones_LCD(a, b).numpy(), R.jones_LCD(a, b))
        assert np.allclose(OE.jones_sSLM(a, b).numpy(), R.jones_sSLM(a, b))
        assert np.allclose(OE.jones_general_retarder(a, b, 0.4).numpy(), R.jones_general_retarder(a, b, 0.4))


def test_masks_stay_pending_and_chain():
    x, y = R.space(100.0, 16)
    v = PolarizedLightSource(x, y, 0.633, device="cpu")
    v.gaussian_beam(w0=(40.0, 40.0), jones_vector=(1, 0))
    a = OE.sSLM(v, np.zeros((16, 16)), np.zeros((16, 16)))
    b = OE.LCD(a, 0.5, 0.2)
    assert len(b._masks) == 2 and b._masks[1].kind == N.XL_MASK_LCD_SCALAR
    assert b._src is v._E


def test_product_path_fails_loudly_without_gpu():
    x, y = R.space(100.0, 16)
    l = W.LightSource(x, y, 0.633, device="cpu")
    l.gaussian_beam(w0=(40.0, 40.0), E0=1)
    with pytest.raises(N.NativeUnavailable):
        l.RS_propagation(100.0)
    with pytest.raises(N.NativeUnavailable):
        l.CZT(1000.0)


# -- argument validation before any pointer reaches a kernel (CPU: the checks are host logic) --

def test_rs_argument_checks_reject_bad_buffers():
    cpu = torch.device("cpu")
    n = 8
    good_out = torch.empty((3, n, n), dtype=torch.complex64)
    good_i = torch.empty((n, n), dtype=torch.float32)
    ops.check_rs_args((2, n, n), 2, 3, n, n, good_out, good_i, cpu)      # accepted
    bad = [
        dict(nin=3),                                                   # more components than the source holds
        dict(nin=0),
        dict(nx=n + 1),                                                # grid length differs from the field
        dict(ny=n - 1),
        dict(out=torch.empty((2, n, n), dtype=torch.complex64)),       # too few planes
        dict(out=torch.empty((3, n, n), dtype=torch.complex128)),      # wrong dtype
        dict(out=torch.empty((3, n, 2 * n), dtype=torch.complex64)[..., ::2]),   # not contiguous
        dict(inten=torch.empty((n, n), dtype=torch.float64)),
        dict(inten=torch.empty((n, n + 1), dtype=torch.float32)),
    ]
    for kw in bad:
        args = dict(nin=2, nx=n, ny=n, out=good_out, inten=good_i)
        args.update(kw)
        with pytest.raises(ValueError):
            ops.check_rs_args((2, n, n), args["nin"], 3, args["nx"], args["ny"], args["out"], args["inten"], cpu)
    with pytest.raises(ValueError):       # a buffer on another device than the field
        ops.check_rs_args((2, n, n), 2, 3, n, n, good_out, good_i, torch.device("meta"))


def test_czt_argument_checks_reject_bad_buffers():
    cpu = torch.device("cpu")
    out = torch.empty((3, 5, 6), dtype=torch.complex64)
    ops.check_czt_args((2, 8, 9), 2, 3, (8, 9), (5, 6), out, cpu)
    with pytest.raises(ValueError):
        ops.check_czt_args((2, 8, 9), 2, 3, (9, 8), (5, 6), out, cpu)      # input plane mismatch
    with pytest.raises(ValueError):
        ops.check_czt_args((2, 8, 9), 3, 3, (8, 9), (5, 6), out, cpu)      # nin above the source's components
    with pytest.raises(ValueError):
        ops.check_czt_args((2, 8, 9), 2, 3, (8, 9), (6, 5), out, cpu)      # output plane mismatch
    with pytest.raises(ValueError):
        ops.check_czt_args((2, 8, 9), 2, 2, (8, 9), (5, 6), out, cpu)      # component count mismatch
    with pytest.raises(ValueError):
        ops.check_czt_args((2, 8, 9), 2, 3, (8, 9), (5, 6), out.to(torch.complex128), cpu)
    with pytest.raises(ValueError):
        ops.check_workspace(torch.empty(16, dtype=torch.uint8), torch.device("meta"))


def test_detector_intensity_stack_matches_torch_chain():
    """The fused |E_j[2]|^2 stack and its backward (cotangent 2 g E in plane 2,
    zeros in planes 0 and 1) against the select / abs / pow / stack chain."""
    torch.manual_seed(3)
    fields = [torch.randn(3, 5, 4, dtype=torch.complex64, requires_grad=True) for _ in range(3)]
    w = torch.randn(3, 5, 4)
    got = OE._ComponentIntensity.apply(2, *fields)
    ref = torch.stack([e[2].abs() ** 2 for e in fields])
    assert torch.allclose(got, ref, rtol=1e-6, atol=0)
    g1 = torch.autograd.grad((got * w).sum(), fields)
    g2 = torch.autograd.grad((ref * w).sum(), fields)
    for a, b in zip(g1, g2):
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-6)
        assert torch.count_nonzero(a[:2]) == 0


def test_distances_mapped_together_match_one_by_one():
    """_distances (one chain of launches for a table's distances) equals the
    per-distance map (|p| 100 + offset) 1e4 in values and gradients."""
    raw = [torch.tensor([v], dtype=torch.float64, requires_grad=True) for v in (0.3, -0.7, 0.05, 1.2)]
    w = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64)
    a = OE._distances(raw, 8.9)
    b = [OE._distance(p, 8.9) for p in raw]
    assert all(x.shape == p.shape and torch.equal(x, y) for x, y, p in zip(a, b, raw))
    ga = torch.autograd.grad(sum(wi * x.sum() for wi, x in zip(w, a)), raw)
    gb = torch.autograd.grad(sum(wi * x.sum() for wi, x in zip(w, b)), raw)
    assert all(torch.equal(x, y) for x, y in zip(ga, gb))
