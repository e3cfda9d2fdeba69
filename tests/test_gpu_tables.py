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


WL_DEP, WL_EX = 0.65, 0.532


def oracle_fluorophores(p, offset=8.9):
    """optical_elements.py:1651-1853 composed from oracle functions."""
    ang = lambda v: v * 2 * np.pi - np.pi
    dist = lambda v: (np.abs(v) * 100 + offset) * 1e4
    red = R.polarized_gaussian_beam(X, Y, WL_DEP, (1200.0, 1200.0), (1, 1))[:2]
    grn = R.polarized_gaussian_beam(X, Y, WL_EX, (1200.0, 1200.0), (1, 1))[:2]
    blk = []
    for i in range(3):
        q = p[6 * i: 6 * i + 6]
        blk.append((ang(q[0]), ang(q[1]), float(ang(q[2])[0]), float(ang(q[3])[0]), float(dist(q[4])[0]),
                    float(dist(q[5])[0])))
    bs = [float(ang(v)[0]) for v in p[18:27]]
    z4, z5 = float(dist(p[27])[0]), float(dist(p[28])[0])

    def vrs(e, z, wl):
        (ex, ey, _), _ = R.vrs_propagation(e[0], e[1], X, Y, wl, z)
        return ex, ey

    def bb(e, a, ph, z, eta, th, wl):
        sx, sy, _ = R.sSLM(e[0], e[1], None, a, ph)
        vx, vy = vrs((sx, sy), z, wl)
        lx, ly, _ = R.LCD(vx, vy, None, eta, th)
        return lx, ly

    def bsym(a, b, th):
        c, d = R.BS_symmetric(a[:2], b[:2], th)
        return c[:2], d[:2]

    def bsi(a, th, port):
        c, d = R.BS_symmetric_SI(a[:2], th, port)
        return c[:2], d[:2]

    def ol(e, wl):
        return R.vczt_objective_lens(e[0], e[1], X, Y, wl, RL, FL, XO, YO)

    (p11, p12, e1, t1, z11, z12), (p21, p22, e2, t2, z21, z22), (p31, p32, e3, t3, z31, z32) = blk
    R_, G_ = WL_DEP, WL_EX
    c1r, d1r = bsi(red, bs[0], 0)
    c1g, d1g = bsi(grn, bs[0], 1)
    b2r, b2g = vrs(bb(c1r, p11, p12, z11, e1, t1, R_), z12, R_), vrs(bb(c1g, p11, p12, z11, e1, t1, G_), z12, G_)
    c2r, d2r = bsi(b2r, bs[1], 1)
    c2g, d2g = bsym(grn, b2g, bs[1])
    b3r, b3g = vrs(c2r, z21 + z22, R_), vrs(c2g, z21 + z22, G_)
    c3r, d3r = bsym(red, b3r, bs[2])
    c3g, d3g = bsi(b3g, bs[2], 1)
    det1r, det1g = ol(vrs(c3r, z31 + z32, R_), R_), ol(vrs(c3g, z31 + z32, G_), G_)
    a5r, a5g, a6r, a6g = vrs(d2r, z4, R_), vrs(d2g, z4, G_), vrs(d3r, z4, R_), vrs(d3g, z4, G_)
    c4r, d4r = bsym(d1r, red, bs[3])
    c4g, d4g = bsi(d1g, bs[3], 0)
    b5r, b5g = vrs(c4r, z11 + z12, R_), vrs(c4g, z11 + z12, G_)
    c5r, d5r = bsym(a5r, b5r, bs[4])
    c5g, d5g = bsym(a5g, b5g, bs[4])
    b6r, b6g = vrs(bb(c5r, p21, p22, z21, e2, t2, R_), z22, R_), vrs(bb(c5g, p21, p22, z21, e2, t2, G_), z22, G_)
    c6r, d6r = bsym(a6r, b6r, bs[5])
    c6g, d6g = bsym(a6g, b6g, bs[5])
    det2r, det2g = ol(vrs(c6r, z31 + z32, R_), R_), ol(vrs(c6g, z31 + z32, G_), G_)
    a8r, a8g, a9r, a9g = vrs(d5r, z5, R_), vrs(d5g, z5, G_), vrs(d6r, z5, R_), vrs(d6g, z5, G_)
    c7r, d7r = bsi(d4r, bs[6], 0)
    c7g, d7g = bsym(d4g, grn, bs[6])
    b8r, b8g = vrs(c7r, z11 + z12, R_), vrs(c7g, z11 + z12, G_)
    c8r, d8r = bsym(a8r, b8r, bs[7])
    c8g, d8g = bsym(a8g, b8g, bs[7])
    b9r, b9g = vrs(c8r, z21 + z22, R_), vrs(c8g, z21 + z22, G_)
    c9r, d9r = bsym(a9r, b9r, bs[8])
    c9g, d9g = bsym(a9g, b9g, bs[8])
    det3r = ol(vrs(bb(c9r, p31, p32, z31, e3, t3, R_), z32, R_), R_)
    det3g = ol(vrs(bb(c9g, p31, p32, z31, e3, t3, G_), z32, G_), G_)
    dep = [det1r, det2r, det3r, ol(d7r, R_), ol(d8r, R_), ol(d9r, R_)]
    exc = [det1g, det2g, det3g, ol(d7g, G_), ol(d8g, G_), ol(d9g, G_)]
    norm = lambda ds: np.stack([(np.abs(d[0]) ** 2 + np.abs(d[1]) ** 2) / np.sum(np.abs(d[0]) ** 2 + np.abs(d[1]) ** 2)
                                for d in ds])
    return R.fluorescence(norm(exc), norm(dep))


def test_sted_table_vs_oracle(cuda_dev):
    p = raw_parameters(83)
    ls = []
    for wl in (WL_DEP, WL_EX):
        s = PolarizedLightSource(X, Y, wl, cuda_dev)
        s.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
        ls.append(s)
    red, grn = ls
    tp = [torch.tensor(v, device=cuda_dev, requires_grad=True) for v in p]
    got = OE.hybrid_setup_fluorophores(red, grn, red, grn, red, grn, tp, [RL, FL, XO, YO])
    ref = oracle_fluorophores(p)
    assert got.shape == (6, M, M)
    g = got.detach().cpu().numpy()
    for d in range(6):
        assert rel_l2(g[d], ref[d]) <= 2e-5, d
    (gz,) = torch.autograd.grad(got.sum() * 1e3, [tp[27]])     # the whole chain is differentiable
    assert torch.isfinite(gz).all()


def oracle_fixed_slms(p, fixed, offset=8.9):
    """optical_elements.py:1347-1501 composed from oracle functions."""
    ang = lambda v: float(v[0]) * 2 * np.pi - np.pi
    dist = lambda v: (abs(float(v[0])) * 100 + offset) * 1e4
    src = R.polarized_gaussian_beam(X, Y, WL, (1200.0, 1200.0), (1, 1))[:2]
    z11, z12, z21, z22, z31, z32, z41, z42 = [dist(v) for v in p[0:8]]
    bs = [ang(v) for v in p[8:17]]
    e1, t1, e2, t2, e3, t3, e4, t4 = [ang(v) for v in p[17:25]]
    ph11, ph12, ph2, ph3, ph41, ph42 = fixed

    def vrs(e, z):
        (ex, ey, _), _ = R.vrs_propagation(e[0], e[1], X, Y, WL, z)
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

    c1, d1 = bsym(src, src, bs[0])
    b2 = vrs(bb(c1, ph11, ph12, z11, e1, t1), z12)
    c2, d2 = bsym(src, b2, bs[1])
    b3 = vrs(c2, z21 + z22)
    c3, d3 = bsym(src, b3, bs[2])
    det1 = ol(c3)
    a5 = vrs(d2, z31 + z32)
    a6 = vrs(bb(d3, ph3, ph3, z31, e3, t3), z32)
    c4, d4 = bsym(d1, src, bs[3])
    b5 = vrs(c4, z11 + z12)
    c5, d5 = bsym(a5, b5, bs[4])
    b6 = vrs(bb(c5, ph2, ph2, z21, e2, t2), z22)
    c6, d6 = bsym(a6, b6, bs[5])
    det2 = ol(c6)
    a7 = vrs(bb(d4, ph41, ph42, z41, e4, t4), z42)
    a8, a9 = vrs(d5, z41 + z42), vrs(d6, z41 + z42)
    c7, d7 = bsym(a7, src, bs[6])
    b8 = vrs(c7, z11 + z12)
    c8, d8 = bsym(a8, b8, bs[7])
    b9 = vrs(c8, z21 + z22)
    c9, d9 = bsym(a9, b9, bs[8])
    det3 = ol(c9)
    dets = [det1, det2, det3, ol(d7), ol(d8), ol(d9)]
    return np.stack([np.abs(d[2]) ** 2 for d in dets])


def test_fixed_slm_table_vs_oracle(cuda_dev):
    rng = np.random.default_rng(84)
    p = [rng.uniform(0, 1, 1) for _ in range(25)]
    fixed = [rng.uniform(-np.pi, np.pi, (N, N)) for _ in range(6)]
    ls = PolarizedLightSource(X, Y, WL, cuda_dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
    tp = [torch.tensor(v, device=cuda_dev) for v in p]
    got, dets = OE.hybrid_setup_fixed_slms(ls, ls, ls, ls, ls, ls, tp, [RL, FL, XO, YO] + fixed)
    ref = oracle_fixed_slms(p, fixed)
    g = got.cpu().numpy()
    assert g.shape == (6, M, M) and len(dets) == 6
    for d in range(6):
        assert rel_l2(g[d], ref[d]) <= 2e-5, d


def oracle_robust(p, noise, offset=15.0):
    """optical_elements.py:964-1117 composed from oracle functions."""
    n_d, n_s, n_w, n_a = noise
    ang = lambda v: v * 2 * np.pi - np.pi
    dist = lambda v: (np.abs(float(v[0])) * 100 + offset) * 1e4
    src = R.polarized_gaussian_beam(X, Y, WL, (1200.0, 1200.0), (1, 1))[:2]
    blk, zs = [], []
    for i in range(3):
        q = p[8 * i: 8 * i + 8]
        blk.append((ang(q[0]) + n_s[2 * i], ang(q[1]) + n_s[2 * i + 1], q[2] + n_a[2 * i], q[3] + n_a[2 * i + 1],
                    float(ang(q[4])[0]) + n_w[2 * i], float(ang(q[5])[0]) + n_w[2 * i + 1]))
        zs.append((dist(q[6]) + n_d[2 * i], dist(q[7]) + n_d[2 * i + 1]))
    bs = [float(ang(v)[0]) for v in p[24:33]]
    z4, z5 = dist(p[33]) + n_d[6], dist(p[34]) + n_d[7]

    def vrs(e, z):
        (ex, ey, _), _ = R.vrs_propagation(e[0], e[1], X, Y, WL, z)
        return ex, ey

    def bb(e, i):
        ph1, ph2, a1, a2, eta, th = blk[i]
        sx, sy, _ = R.sSLM_with_amplitude(e[0], e[1], None, ph1, ph2, a1, a2)
        vx, vy = vrs((sx, sy), zs[i][0])
        lx, ly, _ = R.LCD(vx, vy, None, eta, th)
        return lx, ly

    def bsym(a, b, th):
        c, d = R.BS_symmetric(a[:2], b[:2], th)
        return c[:2], d[:2]

    def ol(e):
        return R.vczt_objective_lens(e[0], e[1], X, Y, WL, RL, FL, XO, YO)

    (z11, z12), (z21, z22), (z31, z32) = zs
    c1, d1 = bsym(src, src, bs[0])
    b2 = vrs(bb(c1, 0), z12)
    c2, d2 = bsym(src, b2, bs[1])
    b3 = vrs(c2, z21 + z22)
    c3, d3 = bsym(src, b3, bs[2])
    det1 = ol(vrs(c3, z31 + z32))
    a5, a6 = vrs(d2, z4), vrs(d3, z4)
    c4, d4 = bsym(d1, src, bs[3])
    b5 = vrs(c4, z11 + z12)
    c5, d5 = bsym(a5, b5, bs[4])
    b6 = vrs(bb(c5, 1), z22)
    c6, d6 = bsym(a6, b6, bs[5])
    det2 = ol(vrs(c6, z31 + z32))
    a8, a9 = vrs(d5, z5), vrs(d6, z5)
    c7, d7 = bsym(d4, src, bs[6])
    b8 = vrs(c7, z11 + z12)
    c8, d8 = bsym(a8, b8, bs[7])
    b9 = vrs(c8, z21 + z22)
    c9, d9 = bsym(a9, b9, bs[8])
    det3 = ol(vrs(bb(c9, 2), z32))
    dets = [det1, det2, det3, ol(d7), ol(d8), ol(d9)]
    return np.stack([np.abs(d[2]) ** 2 for d in dets])


def test_robust_discovery_table_vs_oracle(cuda_dev):
    rng = np.random.default_rng(85)
    p = []
    for _ in range(3):
        p += [rng.uniform(0, 1, (N, N)), rng.uniform(0, 1, (N, N)), rng.uniform(0.3, 1, (N, N)),
              rng.uniform(0.3, 1, (N, N))] + [rng.uniform(0, 1, 1) for _ in range(4)]
    p += [rng.uniform(0, 1, 1) for _ in range(11)]
    noise = (rng.uniform(-50, 50, 8), [rng.uniform(-0.05, 0.05, (N, N)) for _ in range(6)],
             rng.uniform(-0.05, 0.05, 6), [rng.uniform(-0.05, 0.05, (N, N)) for _ in range(6)])
    ls = PolarizedLightSource(X, Y, WL, cuda_dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
    tp = [torch.tensor(v, device=cuda_dev) for v in p]
    n_d, n_s, n_w, n_a = noise
    got = OE.robust_discovery(ls, ls, ls, ls, ls, ls, tp, [RL, FL, XO, YO], list(n_d), n_s, list(n_w), n_a)
    ref = oracle_robust(p, noise)
    g = got.cpu().numpy()
    assert g.shape == (6, M, M)
    for d in range(6):
        assert rel_l2(g[d], ref[d]) <= 2e-5, d
