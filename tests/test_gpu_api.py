"""The drop-in API, following the reference's own tests
(tests/test_wave_optics.py, test_vectorized_optics.py, test_optical_elements.py)
and checking the values against the float64 oracle as well (GPU)."""
import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import optical_elements as OE
from xlumina_amd.vectorized_optics import PolarizedLightSource, VectorizedLight
from xlumina_amd.wave_optics import LightSource, ScalarLight

pytestmark = pytest.mark.gpu
TOL = 1e-5
WL = 633e-3


def H(t):
    return t.detach().cpu().numpy().astype(np.complex128)


@pytest.fixture(scope="module")
def grid1024():
    return np.linspace(-1500, 1500, 1024), np.linspace(-1500, 1500, 1024)


@pytest.fixture(scope="module")
def grid512():
    return np.linspace(-1500, 1500, 512), np.linspace(-1500, 1500, 512)


def test_scalar_light(grid1024, cuda_dev):
    x, y = grid1024
    light = ScalarLight(x, y, WL)
    assert light.wavelength == WL and light.k == 2 * np.pi / WL
    assert tuple(light.field.shape) == (1024, 1024)


def test_rs_propagation(grid1024, cuda_dev):
    x, y = grid1024
    light = LightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), E0=1)
    propagated, qf = light.RS_propagation(z=1000)
    assert tuple(propagated.field.shape) == (1024, 1024)
    ref, rqf = R.rs_propagation(R.gaussian_beam(x, y, WL, (1200, 1200), 1), x, y, WL, 1000.0)
    assert rel_l2(H(propagated.field), ref) < TOL and qf == pytest.approx(rqf, rel=1e-14)


def test_czt(grid1024, cuda_dev):
    x, y = grid1024
    light = LightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), E0=1)
    propagated = light.CZT(z=1000)
    assert tuple(propagated.field.shape) == (1024, 1024)
    ref = R.czt(R.gaussian_beam(x, y, WL, (1200, 1200), 1), x, y, WL, 1000.0)
    assert rel_l2(H(propagated.field), ref) < TOL


def test_vrs_propagation(grid1024, cuda_dev):
    x, y = grid1024
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 1))
    propagated, _ = light.VRS_propagation(z=1000)
    for c in (propagated.Ex, propagated.Ey, propagated.Ez):
        assert tuple(c.shape) == (1024, 1024)
    assert float((light.Ex.abs() ** 2).sum()) > 0 and float((light.Ey.abs() ** 2).sum()) > 0
    ex, ey, _ = R.polarized_gaussian_beam(x, y, WL, (1200, 1200), (1, 1))
    (rx, ry, rz), _ = R.vrs_propagation(ex, ey, x, y, WL, 1000.0)
    assert rel_l2(H(propagated.E), np.stack([rx, ry, rz])) < TOL


def test_vczt(grid1024, cuda_dev):
    x, y = grid1024
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 1))
    propagated = light.VCZT(1000, x, y)
    assert tuple(propagated.Ez.shape) == (1024, 1024)
    ex, ey, _ = R.polarized_gaussian_beam(x, y, WL, (1200, 1200), (1, 1))
    ref = np.stack(R.vczt(ex, ey, x, y, WL, 1000.0, x, y))
    assert rel_l2(H(propagated.E), ref) < TOL


def test_slm_identity(grid512, cuda_dev):
    x, y = grid512
    light = LightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), E0=1)
    out, mask = OE.SLM(light, torch.zeros((512, 512)), 512)
    assert tuple(out.field.shape) == (512, 512)
    assert torch.allclose(out.field, light.field) and torch.allclose(mask, torch.ones_like(mask))


def test_polarization_devices(grid512, cuda_dev):
    x, y = grid512
    n = 512
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 1))
    z = torch.zeros((n, n))
    o = OE.sSLM(light, z, z)
    assert torch.allclose(o.Ex, light.Ex) and torch.allclose(o.Ey, light.Ey) and torch.allclose(o.Ez, light.Ez)
    pi = torch.full((n, n), np.pi)
    o = OE.sSLM(light, pi, pi)
    assert torch.allclose(o.Ex, light.Ex * np.exp(1j * np.pi), atol=1e-6)
    assert torch.allclose(o.Ey, light.Ey * np.exp(1j * np.pi), atol=1e-6)
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 0))
    o = OE.LCD(light, 0, 0)
    assert torch.allclose(o.Ex, light.Ex) and torch.allclose(o.Ey, light.Ey) and torch.allclose(o.Ez, light.Ez)
    o = OE.linear_polarizer(light, z)
    assert torch.allclose(o.Ex, light.Ex) and torch.allclose(o.Ey, light.Ey) and torch.allclose(o.Ez, light.Ez)
    o = OE.linear_polarizer(light, torch.full((n, n), np.pi / 2))
    assert torch.allclose(o.Ex, torch.zeros_like(o.Ex), atol=1e-6)
    assert torch.allclose(o.Ey, torch.zeros_like(o.Ey), atol=1e-6)


def test_high_na_objective_lens(grid512, cuda_dev):
    x, y = grid512
    radius = 3.6 * 1e3 / 2
    f = radius / 0.9
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 0))
    out, stm = OE.high_NA_objective_lens(light, radius, f)
    assert tuple(out.shape) == (3, 512, 512)
    ex, ey, _ = R.polarized_gaussian_beam(x, y, WL, (1200, 1200), (1, 0))
    ref, rstm = R.high_NA_objective_lens(ex, ey, x, y, radius, f)
    assert rel_l2(H(out), ref) < TOL and stm == pytest.approx(rstm)


def test_vczt_objective_lens(grid512, cuda_dev):
    x, y = grid512
    radius = 3.6 * 1e3 / 2
    f = radius / 0.9
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 0))
    xo = np.linspace(-2, 2, 200)
    out = OE.VCZT_objective_lens(light, radius, f, xo, xo)
    assert tuple(out.Ex.shape) == (200, 200)
    ex, ey, _ = R.polarized_gaussian_beam(x, y, WL, (1200, 1200), (1, 0))
    ref = R.vczt_objective_lens(ex, ey, x, y, WL, radius, f, xo, xo)
    assert rel_l2(H(out.E), ref) < TOL


def test_masks_fuse_into_propagation(grid512, cuda_dev):
    x, y = grid512
    n = 512
    rng = np.random.default_rng(5)
    light = PolarizedLightSource(x, y, WL)
    light.gaussian_beam(w0=(1200, 1200), jones_vector=(1, 1))
    a, p = rng.uniform(-np.pi, np.pi, (2, n, n))
    lit = OE.LCD(OE.sSLM(light, torch.tensor(a, dtype=torch.float32), torch.tensor(p, dtype=torch.float32)), 0.4, 1.1)
    fused, _ = lit.VRS_propagation(2000.0)
    # same chain materialised step by step
    stepped = VectorizedLight(x, y, WL)
    stepped.E = lit.E
    ref, _ = stepped.VRS_propagation(2000.0)
    assert rel_l2(H(fused.E), H(ref.E)) < 1e-6
    ex, ey, _ = R.polarized_gaussian_beam(x, y, WL, (1200, 1200), (1, 1))
    sx, sy, _ = R.sSLM(ex, ey, None, a.astype(np.float32).astype(np.float64), p.astype(np.float32).astype(np.float64))
    lx, ly, _ = R.LCD(sx, sy, None, 0.4, 1.1)
    (rx, ry, rz), _ = R.vrs_propagation(lx, ly, x, y, WL, 2000.0)
    assert rel_l2(H(fused.E), np.stack([rx, ry, rz])) < TOL


def test_training_pair_loader_delivers_device_batches(cuda_dev, tmp_path):
    from xlumina_amd import data as D
    rng = np.random.default_rng(70)
    xin = (rng.standard_normal((6, 32, 32)) + 1j * rng.standard_normal((6, 32, 32))).astype(np.complex64)
    xout = rng.standard_normal((6, 32, 32)).astype(np.float32)
    D.save_training_pairs(str(tmp_path / "pairs.npz"), xin, xout)
    L = D.MultiHDF5DataLoader(str(tmp_path), batch_size=3, device=cuda_dev, seed=0)
    a, b = next(L)
    torch.cuda.synchronize()
    assert a.device.type == "cuda" and a.shape == (3, 32, 32)
    np.testing.assert_array_equal(a.cpu().numpy(), xin[L.last_indices])
    np.testing.assert_array_equal(b.cpu().numpy().real, xout[L.last_indices])


def test_staged_czt_tables_and_overlapped_chain(cuda_dev):
    """CZT axis tables staged by one call (tables_mode 1, here also on another
    stream) and used by the next (mode 2) give exactly the plain call's output."""
    import torch
    from xlumina_amd import _native as N
    from xlumina_amd import ops, workloads
    w = workloads.vrs_vczt_2048(cuda_dev, n=256, m_out=64)
    g = w.geometry()
    vrs, ref = w.run(g, czt_fn=lambda v, gg, **kw: ops.czt(v, gg, form="bluestein", **kw))
    ws = torch.empty(ops.czt_workspace_bytes(g, ncomp=3, nin=2, ez_mode=N.XL_EZ_VCZT), dtype=torch.uint8,
                     device=cuda_dev)
    out = torch.empty_like(ref)
    ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, out=out, workspace=ws, form="bluestein", tables_mode=1)
    assert float(out.abs().sum()) >= 0.0                # staging leaves `out` alone (any values)
    ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, out=out, workspace=ws, form="bluestein", tables_mode=2)
    assert torch.equal(out, ref)
    side = torch.cuda.Stream(device=cuda_dev)          # staged on another stream, then joined
    for _ in range(3):                                   # the workspace is reused call after call
        out2 = torch.empty_like(ref)
        side.wait_stream(torch.cuda.current_stream(cuda_dev))
        with torch.cuda.stream(side):
            ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, out=out2, workspace=ws, form="bluestein", tables_mode=1)
        torch.cuda.current_stream(cuda_dev).wait_stream(side)
        ops.czt(vrs, g, ez_mode=N.XL_EZ_VCZT, nin=2, out=out2, workspace=ws, form="bluestein", tables_mode=2)
        assert torch.equal(out2, ref)
