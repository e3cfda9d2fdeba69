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


def oracle_fixed_sted(p, fixed, offset=8.9):
    """optical_elements.py:1119-1345 composed from oracle functions, one colour
    at a time (depletion: ls1, ls3, ls5; excitation: ls2, ls4, ls6)."""
    ang = lambda v: float(v[0]) * 2 * np.pi - np.pi
    dist = lambda v: (abs(float(v[0])) * 100 + offset) * 1e4
    z11, z12, z21, z22, z31, z32, z41, z42 = [dist(v) for v in p[0:8]]
    bs = [ang(v) for v in p[8:17]]
    e1, t1, e2, t2, e3, t3, e4, t4 = [ang(v) for v in p[17:25]]
    ph11, ph12, ph2, ph3, ph41, ph42 = fixed

    def colour(wl, dep):
        src = R.polarized_gaussian_beam(X, Y, wl, (1200.0, 1200.0), (1, 1))[:2]

        def vrs(e, z):
            (ex, ey, _), _ = R.vrs_propagation(e[0], e[1], X, Y, wl, z)
            return ex, ey

        def bb(e, a, ph, z, eta, th):
            sx, sy, _ = R.sSLM(e[0], e[1], None, a, ph)
            lx, ly, _ = R.LCD(*vrs((sx, sy), z), None, eta, th)
            return lx, ly

        def bsym(a, b, th):
            c, d = R.BS_symmetric(a[:2], b[:2], th)
            return c[:2], d[:2]

        def bsi(a, th, port):
            c, d = R.BS_symmetric_SI(a[:2], th, port)
            return c[:2], d[:2]

        def ol(e):
            return R.vczt_objective_lens(e[0], e[1], X, Y, wl, RL, FL, XO, YO)

        c1, d1 = bsi(src, bs[0], 0 if dep else 1)
        b2 = vrs(bb(c1, ph11, ph12, z11, e1, t1), z12)
        c2, d2 = bsi(b2, bs[1], 1) if dep else bsym(src, b2, bs[1])
        b3 = vrs(c2, z21 + z22)
        c3, d3 = bsym(src, b3, bs[2]) if dep else bsi(b3, bs[2], 1)
        det1 = ol(c3)
        a5 = vrs(d2, z31 + z32)
        a6 = vrs(bb(d3, ph3, ph3, z31, e3, t3), z32)
        c4, d4 = bsym(d1, src, bs[3]) if dep else bsi(d1, bs[3], 0)
        b5 = vrs(c4, z11 + z12)
        c5, d5 = bsym(a5, b5, bs[4])
        b6 = vrs(bb(c5, ph2, ph2, z21, e2, t2), z22)
        c6, d6 = bsym(a6, b6, bs[5])
        det2 = ol(c6)
        a7 = vrs(bb(d4, ph41, ph42, z41, e4, t4), z42)
        a8, a9 = vrs(d5, z41 + z42), vrs(d6, z41 + z42)
        c7, d7 = bsi(a7, bs[6], 0) if dep else bsym(a7, src, bs[6])
        b8 = vrs(c7, z11 + z12)
        c8, d8 = bsym(a8, b8, bs[7])
        b9 = vrs(c8, z21 + z22)
        c9, d9 = bsym(a9, b9, bs[8])
        det3 = ol(c9)
        i = [np.abs(d[0]) ** 2 + np.abs(d[1]) ** 2 for d in (det1, det2, det3, ol(d7), ol(d8), ol(d9))]
        return np.stack([v / v.sum() for v in i])

    return R.fluorescence(colour(WL_EX, False), colour(WL_DEP, True))


def two_colour_sources(dev):
    out = []
    for wl in (WL_DEP, WL_EX):
        s = PolarizedLightSource(X, Y, wl, dev)
        s.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
        out.append(s)
    return out


def test_fixed_slm_sted_table_vs_oracle(cuda_dev):
    rng = np.random.default_rng(86)
    p = [rng.uniform(0, 1, 1) for _ in range(25)]
    fixed = [rng.uniform(-np.pi, np.pi, (N, N)) for _ in range(6)]
    red, grn = two_colour_sources(cuda_dev)
    tp = [torch.tensor(v, device=cuda_dev, requires_grad=True) for v in p]
    got = OE.hybrid_setup_fixed_slms_fluorophores(red, grn, red, grn, red, grn, tp, [RL, FL, XO, YO] + fixed)
    ref = oracle_fixed_sted(p, fixed)
    g = got.detach().cpu().numpy()
    assert g.shape == (6, M, M)
    for d in range(6):
        assert rel_l2(g[d], ref[d]) <= 2e-5, d
    (gz,) = torch.autograd.grad(got.sum() * 1e3, [tp[3]])
    assert torch.isfinite(gz).all()


def oracle_six_by_six(p, fixed, offset=7.6):
    """optical_elements.py:1855-2125 composed from oracle functions, splitter by
    splitter as the reference lists them."""
    ang = lambda v: float(v[0]) * 2 * np.pi - np.pi
    dist = lambda v: (abs(float(v[0])) * 100 + offset) * 1e4
    z1, z2, z3, z4, z5, z6, z7, z8, z9, z10, z11, z12 = [dist(v) for v in p[0:12]]
    bs = [ang(v) for v in p[12:48]]
    e1, t1, e2, t2 = [ang(v) for v in p[48:52]]
    ph11, ph12, ph21, ph22 = fixed
    s = R.polarized_gaussian_beam(X, Y, WL, (1200.0, 1200.0), (1, 1))[:2]

    def v(e, z):
        (ex, ey, _), _ = R.vrs_propagation(e[0], e[1], X, Y, WL, z)
        return ex, ey

    def bb(e, a, ph, z, eta, th):
        sx, sy, _ = R.sSLM(e[0], e[1], None, a, ph)
        lx, ly, _ = R.LCD(*v((sx, sy), z), None, eta, th)
        return lx, ly

    def B(a, b, k):
        c, d = R.BS_symmetric(a[:2], b[:2], bs[k - 1])
        return c[:2], d[:2]

    def ol(e):
        return R.vczt_objective_lens(e[0], e[1], X, Y, WL, RL, FL, XO, YO)

    c1, d1 = B(s, s, 1); b2 = v(c1, z1); a7 = v(d1, z8)                           # noqa: E702
    c2, d2 = B(s, b2, 2); b3 = v(c2, z2 + z3); a8 = v(d2, z8)                     # noqa: E702
    c3, d3 = B(s, b3, 3); b4 = v(c3, z4); a9 = v(d3, z8)                          # noqa: E702
    c4, d4 = B(s, b4, 4); b5 = v(c4, z5 + z6); a10 = v(d4, z8)                    # noqa: E702
    c5, d5 = B(s, b5, 5); b6 = v(c5, z7); a11 = v(d5, z8)                         # noqa: E702
    c6, d6 = B(s, b6, 6); a12 = v(d6, z8); det1 = ol(c6)                          # noqa: E702
    c7, d7 = B(a7, s, 7); b8 = v(c7, z1); a13 = v(d7, z9)                         # noqa: E702
    c8, d8 = B(a8, b8, 8); b9 = v(bb(c8, ph11, ph12, z2, e1, t1), z3); a14 = v(d8, z9)  # noqa: E702
    c9, d9 = B(a9, b9, 9); b10 = v(c9, z4); a15 = v(d9, z9)                       # noqa: E702
    c10, d10 = B(a10, b10, 10); b11 = v(c10, z5 + z6); a16 = v(d10, z9)           # noqa: E702
    c11, d11 = B(a11, b11, 11); b12 = v(c11, z7); a17 = v(d11, z9)                # noqa: E702
    c12, d12 = B(a12, b12, 12); a18 = v(d12, z9); det2 = ol(c12)                  # noqa: E702
    c13, d13 = B(a13, s, 13); b14 = v(c13, z1); a19 = v(d13, z10)                 # noqa: E702
    c14, d14 = B(a14, b14, 14); b15 = v(c14, z2 + z3); a20 = v(d14, z10)          # noqa: E702
    c15, d15 = B(a15, b15, 15); b16 = v(c15, z4); a21 = v(d15, z10)               # noqa: E702
    c16, d16 = B(a16, b16, 16); b17 = v(c16, z5 + z6); a22 = v(d16, z10)          # noqa: E702
    c17, d17 = B(a17, b17, 17); b18 = v(c17, z7); a23 = v(d17, z10)               # noqa: E702
    c18, d18 = B(a18, b18, 18); a24 = v(d18, z10); det3 = ol(c18)                 # noqa: E702
    c19, d19 = B(a19, s, 19); b20 = v(c19, z1); a25 = v(d19, z11)                 # noqa: E702
    c20, d20 = B(a20, b20, 20); b21 = v(c20, z2 + z3); a26 = v(d20, z11)          # noqa: E702
    c21, d21 = B(a21, b21, 21); b22 = v(c21, z4); a27 = v(d21, z11)               # noqa: E702
    c22, d22 = B(a22, b22, 22); b23 = v(bb(c22, ph21, ph22, z5, e2, t2), z6); a28 = v(d22, z11)  # noqa: E702
    c23, d23 = B(a23, b23, 23); b24 = v(c23, z7); a29 = v(d23, z11)               # noqa: E702
    c24, d24 = B(a24, b24, 24); a30 = v(d24, z11); det4 = ol(c24)                 # noqa: E702
    c25, d25 = B(a25, s, 25); b26 = v(c25, z1); a31 = v(d25, z12)                 # noqa: E702
    c26, d26 = B(a26, b26, 26); b27 = v(c26, z2 + z3); a32 = v(d26, z12)          # noqa: E702
    c27, d27 = B(a27, b27, 27); b28 = v(c27, z4); a33 = v(d27, z12)               # noqa: E702
    c28, d28 = B(a28, b28, 28); b29 = v(c28, z5 + z6); a34 = v(d28, z12)          # noqa: E702
    c29, d29 = B(a29, b29, 29); b30 = v(c29, z7); a35 = v(d29, z12)               # noqa: E702
    c30, d30 = B(a30, b30, 30); a36 = v(d30, z12); det5 = ol(c30)                 # noqa: E702
    c31, d31 = B(a31, s, 31); b32 = v(c31, z1); det7 = ol(c31)                    # noqa: E702
    c32, d32 = B(a32, b32, 32); b33 = v(c32, z2 + z3); det8 = ol(c32)             # noqa: E702
    c33, d33 = B(a33, b33, 33); b34 = v(c33, z4); det9 = ol(c33)                  # noqa: E702
    c34, d34 = B(a34, b34, 34); b35 = v(c34, z5 + z6); det10 = ol(c34)            # noqa: E702
    c35, d35 = B(a35, b35, 35); b36 = v(c35, z7); det11 = ol(c35)                 # noqa: E702
    c36, d36 = B(a36, b36, 36); det6 = ol(c36); det12 = ol(d36)                   # noqa: E702
    dets = [det1, det2, det3, det4, det5, det6, det7, det8, det9, det10, det11, det12]
    return np.stack([np.abs(d[2]) ** 2 for d in dets])


def test_six_by_six_table_vs_oracle(cuda_dev):
    rng = np.random.default_rng(87)
    p = [rng.uniform(0, 1, 1) for _ in range(52)]
    fixed = [rng.uniform(-np.pi, np.pi, (N, N)) for _ in range(4)]
    ls = PolarizedLightSource(X, Y, WL, cuda_dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
    tp = [torch.tensor(v, device=cuda_dev, requires_grad=True) for v in p]
    got = OE.six_times_six_ansatz(*([ls] * 12), tp, [RL, FL, XO, YO] + fixed)
    ref = oracle_six_by_six(p, fixed)
    g = got.detach().cpu().numpy()
    assert g.shape == (12, M, M)
    for d in range(12):
        assert rel_l2(g[d], ref[d]) <= 2e-5, d
    gz = torch.autograd.grad(got.sum() * 1e3, [tp[7], tp[20], tp[48]])   # z8, bs9, eta1
    assert all(torch.isfinite(v).all() for v in gz)


def test_noise_aware_batch_vs_oracle(cuda_dev):
    """Two tables of robust_discovery with shake_setup noise (batch over the
    noise axis, noisy_optimization.ipynb cell 7) against the oracle table per
    noise draw; the experiment's batch loss (cell 11) on top."""
    rng = np.random.default_rng(88)
    p = []
    for _ in range(3):
        p += [rng.uniform(0, 1, (N, N)), rng.uniform(0, 1, (N, N)), rng.uniform(0.3, 1, (N, N)),
              rng.uniform(0.3, 1, (N, N))] + [rng.uniform(0, 1, 1) for _ in range(4)]
    p += [rng.uniform(0, 1, 1) for _ in range(11)]
    NS = {'n_tables': 2, 'number of distances': 8, 'number of sSLM': 3, 'number of wps': 3,
          'noise_level': 'low', 'discretize': False}
    nd, nsl, nw, na, _, _ = OE.batch_shake_setup([5, 6], N, NS, device=cuda_dev)
    ls = PolarizedLightSource(X, Y, WL, cuda_dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))
    tp = [torch.tensor(v, device=cuda_dev) for v in p]
    got = OE.batch_robust_discovery(ls, ls, ls, ls, ls, ls, tp, [RL, FL, XO, YO], nd, nsl, nw, na)
    assert got.shape == (2, 6, M, M)
    g = got.cpu().numpy()
    refs = []
    for t in range(2):
        noise = (nd[t].cpu().numpy(), list(nsl[t].cpu().numpy()), nw[t].cpu().numpy(), list(na[t].cpu().numpy()))
        refs.append(oracle_robust(p, noise))
        for d in range(6):
            assert rel_l2(g[t, d], refs[t][d]) <= 2e-5, (t, d)
    loss = softmin(torch.stack([LF.vectorized_loss_hybrid(got[t]) for t in range(2)]).mean(0, keepdim=True))
    ref_loss = LR.softmin(np.mean([LR.vectorized_loss_hybrid(r) for r in refs], axis=0, keepdims=True))
    assert abs(float(loss) - float(ref_loss)) <= 1e-4 * abs(float(ref_loss))


def test_captured_table_step_matches_eager(cuda_dev):
    """The whole sharp-focus table step (forward, loss, backward) captured in
    a HIP graph: three replays, with parameter updates and eager steps between
    them, against the eager step on the same parameters."""
    from xlumina_amd import workloads
    p = raw_parameters(89)
    tp = [torch.tensor(v, device=cuda_dev, requires_grad=True) for v in p]
    ls = PolarizedLightSource(X, Y, WL, cuda_dev)
    ls.gaussian_beam(w0=(1200.0, 1200.0), jones_vector=(1, 1))

    def loss():
        inten, _ = OE.hybrid_setup_sharp_focus(ls, ls, ls, ls, ls, ls, tp, [RL, FL, XO, YO])
        return softmin(LF.vectorized_loss_hybrid(inten))

    cap = workloads.CapturedStep(loss, tp)
    gen = torch.Generator(device=cuda_dev).manual_seed(5)
    for _ in range(3):
        Lg, gg = cap.replay()
        Lg, gg = float(Lg.detach()), [g.clone() for g in gg]
        for q in tp:
            q.grad = None
        Le = loss()
        Le.backward()
        assert abs(Lg - float(Le.detach())) <= 1e-6 * abs(float(Le.detach()))
        for a, b in zip(gg, (q.grad for q in tp)):
            assert float((a - b).norm()) <= 1e-6 * float(b.norm()) + 1e-30
        with torch.no_grad():
            for q in tp:
                q.add_(0.01 * torch.randn(q.shape, generator=gen, device=cuda_dev, dtype=q.dtype))
