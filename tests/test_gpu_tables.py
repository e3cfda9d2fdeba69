"""A pre-built optical table run device-resident (hybrid_setup_sharp_focus,
optical_elements.py:1503-1649) against the same table composed from the
float64 oracle, and its parameter gradients against central differences.

Tolerances: detector intensities <= 2e-5 relative L2 per detector;
gradients of a smooth detector loss <= 1e-3 relative.
"""
import math

import numpy as np
import pytest
import torch

from oracle import losses_ref as LR
from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import loss_functions as LF
from xlumina_amd import optical_elements as OE
from xlumina_amd.toolbox import softmin
from xlumina_amd.vectorized_optics import PolarizedLightSource

pytestmark = pytest.mark.gpu

N, M = 64, 48
WL = 0.635
X, Y = R.space(2500.0, N)
XO, YO = R.space(10.0, M)
RL = 1800.0
FL = RL / 0.9


def raw_parameters(seed=80):
    rng = np.random.default_rng(seed)
    p = []
    for _ in range(3):
        p += [rng.uniform(0, 1, (N, N)), rng.uniform(0, 1, (N, N))] + [rng.uniform(0, 1, 1) for _ in range(4)]
    p += [rng.uniform(0, 1, 1) for _ in range(11)]
    return p


def oracle_table(p, offset=8.9):
    """The table of optical_elements.py:1503-1649 composed from oracle functions."""
    ang = lambda v: v * 2 * np.pi - np.pi
    dist = lambda v: (np.abs(v) * 100 + offset) * 1e4
    src = R.polarized_gaussian_beam(X, Y, WL, (1200.0, 1200.0), (1, 1))[:2]
    blk = []
    for i in range(3):
        q = p[6 * i: 6 * i + 6]
        blk.append((ang(q[0]), ang(q[1]), float(ang(q[2])[0]), float(ang(q[3])[0]), float(dist(q[4])[0]),
                    float(dist(q[5])[0])))
    bs = [float(ang(v)[0]) for v in p[18:27]]
    z4, z5 = float(dist(p[27])[0]), float(dist(p[28])[0])

    def vrs(e, z):
        (ex, ey, ez), _ = R.vrs_propagation(e[0], e[1], X, Y, WL, z)
        return ex, ey

    def bb(e, a, ph, z, eta, th):
        sx, sy, _ = R.sSLM(e[0], e[1], None, a, ph)
        vx, vy = vrs((sx, sy), z)
        lx, ly, _ = R.LCD(vx, vy, None, eta, th)
        return lx, ly

    def bsym(a, b, th):
        c, d = R.BS_symmetric(a[:2], b[:2], th)
        return c[:2], d[:2]

    def ol(e):
        return R.vczt_objective_lens(e[0], e[1], X, Y, WL, RL, FL, XO, YO)

    (p11, p12, e1, t1, z11, z12), (p21, p22, e2, t2, z21, z22), (p31, p32, e3, t3, z31, z32) = blk
    c1, d1 = bsym(src, src, bs[0])
    b2 = vrs(bb(c1, p11, p12, z11, e1, t1), z12)
    c2, d2 = bsym(src, b2, bs[1])
    b3 = vrs(c2, z21 + z22)
    c3, d3 = bsym(src, b3, bs[2])
    det1 = ol(vrs(c3, z31 + z32))
    a5, a6 = vrs(d2, z4), vrs(d3, z4)
    c4, d4 = bsym(d1, src, bs[3])
    b5 = vrs(c4, z11 + z12)
    c5, d5 = bsym(a5, b5, bs[4])
    b6 = vrs(bb(c5, p21, p22, z21, e2, t2), z22)
    c6, d6 = bsym(a6, b6, bs[5])
    det2 = ol(vrs(c6, z31 + z32))
    a8, a9 = vrs(d5, z5), vrs(d6, z5)
    c7, d7 = bsym(d4, src, bs[6])
    b8 = vrs(c7, z11 + z12)
    c8, d8 = bsym(a8, b8, bs[7])
    b9 = vrs(c8, z21 + z22)
    c9, d9 = bsym(a9, b9, bs[8])
    det3 = ol(vrs(bb(c9, p31, p32, z31, e3, t3), z32))
    dets = [det1, det2, det3, ol(d7), ol(d8), ol(d9)]
    return np.stack([np.abs(d[2]) ** 2 for d in dets])


def device_table(p, dev):
    ls = PolarizedLightSource(X, Y, WL, dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
    return OE.hybrid_setup_sharp_focus(ls, ls, ls, ls, ls, ls, p, [RL, FL, XO, YO])


def test_sharp_focus_table_vs_oracle(cuda_dev):
    p = raw_parameters()
    inten, dets = device_table([torch.tensor(v, device=cuda_dev) for v in p], cuda_dev)
    ref = oracle_table(p)
    got = inten.cpu().numpy()
    assert got.shape == (6, M, M) and len(dets) == 6
    for d in range(6):
        assert rel_l2(got[d], ref[d]) <= 2e-5, d
    # the experiment's loss (hybrid_sharp_optical_table.py:50-58) on these detectors
    loss = softmin(LF.vectorized_loss_hybrid(inten))
    assert abs(float(loss) - LR.softmin(LR.vectorized_loss_hybrid(ref))) <= 1e-4 * abs(float(loss))


def test_sharp_focus_table_gradients(cuda_dev):
    p = raw_parameters(81)
    w = np.random.default_rng(82).random((6, M, M))
    tp = [torch.tensor(v, device=cuda_dev, requires_grad=True) for v in p]
    inten, _ = device_table(tp, cuda_dev)
    L = (torch.tensor(w, dtype=torch.float32, device=cuda_dev) * inten).sum()
    grads = torch.autograd.grad(L, tp)

    def ref_loss(q):
        return float(np.sum(w * oracle_table(q)))

    # z1_1, one pixel of phase3_1, bs1, z4.  Distances are (|p| 100 + 8.9) cm and
    # the ports interfere with period lambda in z, so their step is 1e-9 (1 nm).
    for idx in (4, 12, 18, 27):
        q = [v.copy() for v in p]
        h = 1e-9 if idx in (4, 27) else 1e-5
        if idx == 12:
            q[idx][20, 30] += h
            lp = ref_loss(q)
            q[idx][20, 30] -= 2 * h
            fd = (lp - ref_loss(q)) / (2 * h)
            g = float(grads[idx][20, 30])
        else:
            q[idx] = q[idx] + h
            lp = ref_loss(q)
            q[idx] = q[idx] - 2 * h
            fd = (lp - ref_loss(q)) / (2 * h)
            g = float(grads[idx].sum())
        assert abs(g - fd) <= 1e-3 * abs(fd) + 1e-7 * abs(lp), (idx, g, fd)
