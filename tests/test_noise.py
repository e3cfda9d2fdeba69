"""Noise for noise-aware tables (optical_elements.py:2127-2217).

The reference draws from JAX's threefry streams, which are not available
here, so the values are not reproducible bit for bit; these tests pin what
the reference's code fixes instead: shapes, the bounds of each noise level,
the sign conventions, the shared magnitudes of SLM phase and amplitude noise
(both drawn from key2 at optical_elements.py:2213-2214), the uint8 cast of
the sign under 'discretize' (-1 -> 255), and uniformity of the magnitudes
(Kolmogorov-Smirnov, fixed seeds)."""
import numpy as np
import pytest
import torch
from scipy import stats

from xlumina_amd import optical_elements as OE


def ns(level='mild', discretize=False, **kw):
    d = {'n_tables': 2, 'number of distances': 8, 'number of sSLM': 3, 'number of wps': 3,
         'noise_level': level, 'discretize': discretize}
    d.update(kw)
    return d


def test_shapes_and_determinism():
    d, s, w, a, k0, k = OE.shake_setup(11, 24, ns())
    assert d.shape == (8,) and s.shape == (6, 24, 24) and w.shape == (6,) and a.shape == (6, 24, 24)
    assert all(t.dtype == torch.float64 for t in (d, s, w, a))
    assert k == 11 and k0 != 11
    again = OE.shake_setup(11, 24, ns())
    assert all(torch.equal(x, y) for x, y in zip((d, s, w, a), again[:4]))
    other = OE.shake_setup(k0, 24, ns())
    assert not torch.equal(other[1], s)


@pytest.mark.parametrize("level,dlim,plim", [('low', (10, 50), (0.01, 0.05)), ('mild', (50, 500), (0.05, 0.5)),
                                             ('high', (500, 1000), (0.5, 1.0)), ('all', (10, 1000), (0.01, 1.0))])
def test_level_bounds(level, dlim, plim):
    d, s, w, a, _, _ = OE.shake_setup(3, 32, ns(level))
    assert float(d.min()) >= dlim[0] and float(d.max()) < dlim[1]      # distances carry no sign
    for t in (s, w, a):
        m = t.abs()
        assert float(m.min()) >= plim[0] and float(m.max()) < plim[1]
    assert (s < 0).any() and (s > 0).any() and (a < 0).any() and (a > 0).any()


def test_tunable_and_unknown_level():
    d, s, w, a, _, _ = OE.shake_setup(5, 16, ns('tunable', misalignment=(1.0, 2.0), phase_noise=(0.2, 0.3)))
    assert 1.0 <= float(d.min()) and float(d.max()) < 2.0
    assert 0.2 <= float(s.abs().min()) and float(s.abs().max()) < 0.3
    with pytest.raises(ValueError):
        OE.shake_setup(5, 16, ns('extreme'))


def test_shared_magnitudes_independent_signs():
    _, s, _, a, _, _ = OE.shake_setup(21, 64, ns())
    assert torch.equal(s.abs(), a.abs())
    agree = float(((s > 0) == (a > 0)).double().mean())
    assert abs(agree - 0.5) < 0.02                                        # 24576 pixels
    assert abs(float((s > 0).double().mean()) - 0.5) < 0.02


def test_discretize_wraps_negative_sign():
    _, s, w, a, _, _ = OE.shake_setup(9, 64, ns('low', discretize=True))
    m = s.abs()
    ratio_s = (s / OE.shake_setup(9, 64, ns('low'))[1].abs()).numpy()
    assert set(np.unique(np.round(ratio_s, 9))) == {1.0, 255.0}
    assert (s > 0).all() and (a > 0).all()                                # uint8 signs are never negative
    assert (w < 0).any()                                                  # waveplates keep int8 signs
    assert float(m.max()) < 255 * 0.05


def test_uniform_magnitudes():
    d, s, w, a, _, _ = OE.shake_setup(1234, 64, ns('mild', **{'number of distances': 4000}))
    lo, hi = 0.05, 0.5
    assert stats.kstest((s.abs().flatten().numpy() - lo) / (hi - lo), 'uniform').pvalue > 1e-3
    assert stats.kstest((d.numpy() - 50) / 450, 'uniform').pvalue > 1e-3


def test_batch_stacks():
    d, s, w, a, k0, k = OE.batch_shake_setup([1, 2, 3], 16, ns())
    assert d.shape == (3, 8) and s.shape == (3, 6, 16, 16) and w.shape == (3, 6) and a.shape == (3, 6, 16, 16)
    assert k == [1, 2, 3] and len(set(k0)) == 3
    assert torch.equal(s[1], OE.shake_setup(2, 16, ns())[1])
