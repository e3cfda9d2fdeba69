"""Serving loop (GPU): each result is the same alone and in the mix (2048^2
requests included), the same from batched graphs as from eager batched
launches, and the same as the eager autograd path; a 512^2 scalar request
matches the float64 oracle and a vectorial gradient request matches central
differences of it; a request with NaN inputs is quarantined without touching
the results of the requests batched with it.

Tolerances: alone vs mix, batched vs one-by-one and graph vs eager launches
<= 1e-6 relative L2 (members of a batch run in separate blocks on separate
workspace; no atomics, so bit equality is expected); served vs autograd path
<= 1e-5; intensity vs the float64 oracle <= 1e-5; the gradient against
float64 central differences (eps = 1e-3 rad on 8 pixels) <= 2e-3.
"""
import math

import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import autograd as AG
from xlumina_amd import ops
from xlumina_amd import serving as S
from xlumina_amd.ops import MaskSpec

pytestmark = pytest.mark.gpu


def small_stream(count, rate=None, seed=11):
    return S.make_requests(count, rate, seed, sizes=(512, 1024), size_p=(0.6, 0.4), p_vectorial=0.4, p_grad=0.4)


def eager_reference(server, r, dev):
    """The request through the differentiable drop-in path (no graphs, no batch)."""
    sh = server.shared(r.n)
    p = S.request_phases(r, dev)
    m0 = p[0].clone().requires_grad_(r.grad)
    if r.vectorial:
        masks = [MaskSpec(N.XL_MASK_SSLM, m0, p[1])]
        E = AG.rs(sh.beam_v, sh.x, sh.y, r.z, server.wavelength, masks=masks, ez_mode=N.XL_EZ_VRS)
    else:
        masks = [MaskSpec(N.XL_MASK_SLM, m0)]
        E = AG.rs(sh.beam_s, sh.x, sh.y, r.z, server.wavelength, masks=masks)
    inten = (E.real ** 2 + E.imag ** 2).sum(0)
    g = None
    if r.grad:
        L = (0.5 * sh.w2 * inten).sum()
        (g,) = torch.autograd.grad(L, m0)
    return inten.detach(), g


@pytest.fixture(scope="module")
def server(cuda_dev):
    return S.Server(cuda_dev, slots_per_bucket=2)


def _compare(a_results, b_results, reqs):
    worst = 0.0
    for r in reqs:
        ia, ga = a_results[r.rid]
        ib, gb = b_results[r.rid]
        worst = max(worst, rel_l2(ib.cpu().numpy(), ia.cpu().numpy()))
        if r.grad:
            worst = max(worst, rel_l2(gb.cpu().numpy(), ga.cpu().numpy()))
    return worst


def test_alone_vs_mix(server):
    reqs = small_stream(24)
    server.run(reqs, keep_results=True, serial=True)
    alone = dict(server.results)
    rep = server.run(reqs, keep_results=True)            # all arrive at t = 0: maximal batching and overlap
    mix = dict(server.results)
    assert rep.completed == len(reqs) and rep.batches < len(reqs)
    worst = _compare(alone, mix, reqs)
    print(f"alone vs mix: worst rel L2 {worst:.3e} over {len(reqs)} requests in {rep.batches} batches")
    assert worst <= 1e-6


def test_alone_vs_mix_with_2048(server):
    reqs = S.make_requests(10, None, 21, sizes=(512, 2048), size_p=(0.5, 0.5), p_vectorial=0.5, p_grad=0.5)
    assert any(r.n == 2048 and r.vectorial for r in reqs) and any(r.n == 2048 and r.grad for r in reqs)
    server.run(reqs, keep_results=True, serial=True)
    alone = dict(server.results)
    rep = server.run(reqs, keep_results=True)
    assert rep.completed == len(reqs)
    worst = _compare(alone, dict(server.results), reqs)
    print(f"alone vs mix with 2048^2: worst rel L2 {worst:.3e}")
    assert worst <= 1e-6


def test_served_matches_autograd_path(server, cuda_dev):
    reqs = small_stream(10, seed=12)
    assert any(r.grad and r.vectorial for r in reqs) or any(r.grad for r in reqs)
    server.run(reqs, keep_results=True)
    for r in reqs:
        inten, g = server.results[r.rid]
        ref_i, ref_g = eager_reference(server, r, cuda_dev)
        assert rel_l2(inten.cpu().numpy(), ref_i.cpu().numpy()) <= 1e-5, r
        if r.grad:
            assert rel_l2(g.cpu().numpy(), ref_g.cpu().numpy()) <= 1e-5, r


def test_graphs_match_eager_launches(server, cuda_dev):
    reqs = small_stream(8, seed=13)
    server.run(reqs, keep_results=True)
    with_graphs = dict(server.results)
    plain = S.Server(cuda_dev, slots_per_bucket=1, use_graphs=False)
    plain.run(reqs, keep_results=True)
    assert _compare(with_graphs, plain.results, reqs) <= 1e-6


def test_scalar_512_vs_oracle(server, cuda_dev):
    r = S.Request(0, 512, False, 37000.0, False, 123456789)
    server.run([r], keep_results=True)
    inten = server.results[0][0].cpu().numpy().astype(np.float64)
    sh = server.shared(512)
    phi = S.request_phases(r, cuda_dev)[0].cpu().numpy().astype(np.float64)
    u, _ = R.slm(sh.beam_s[0].cpu().numpy().astype(np.complex128), phi)
    E, _ = R.rs_propagation(u, sh.x, sh.y, server.wavelength, r.z)
    assert rel_l2(inten, np.abs(E) ** 2) <= 1e-5


def test_vectorial_gradient_request_vs_float64_central_differences(server, cuda_dev):
    """dL/dalpha of a vectorial 512^2 request (sSLM -> VRS, L = sum w |E|^2)
    against central differences of the float64 oracle at 8 seeded pixels."""
    r = S.Request(0, 512, True, 21000.0, True, 424242)
    server.run([r], keep_results=True)
    inten, g = server.results[0]
    sh = server.shared(512)
    ph = S.request_phases(r, cuda_dev).cpu().numpy().astype(np.float64)
    bx, by = (sh.beam_v[i].cpu().numpy().astype(np.complex128) for i in range(2))
    w = sh.w2.cpu().numpy().astype(np.float64) / 2.0

    def loss(alpha):
        ex, ey, _ = R.sSLM(bx, by, None, alpha, ph[1])
        (vx, vy, vz), _ = R.vrs_propagation(ex, ey, sh.x, sh.y, server.wavelength, r.z)
        return float(np.sum(w * (np.abs(vx) ** 2 + np.abs(vy) ** 2 + np.abs(vz) ** 2)))

    rng = np.random.default_rng(5)
    # pixels inside the beam, where the gradient is not negligible
    pix = [(int(i), int(j)) for i, j in rng.integers(160, 352, (8, 2))]
    eps = 1e-3
    fd = []
    for i, j in pix:
        a = ph[0].copy()
        a[i, j] += eps
        lp = loss(a)
        a[i, j] -= 2 * eps
        lm = loss(a)
        fd.append((lp - lm) / (2 * eps))
    got = np.array([float(g[i, j]) for i, j in pix])
    fd = np.array(fd)
    err = np.linalg.norm(got - fd) / np.linalg.norm(fd)
    print(f"vectorial gradient vs float64 central differences: rel L2 {err:.3e}")
    assert err <= 2e-3


def test_nan_request_is_quarantined_without_touching_its_batch(server, cuda_dev):
    reqs = [S.Request(i, 512, False, 30000.0 + 1000.0 * i, False, 1000 + i) for i in range(4)]
    server.run(reqs, keep_results=True, serial=True)
    alone = dict(server.results)
    bad = np.full((1, 512, 512), np.nan, dtype=np.float32)
    mixed = reqs[:2] + [S.Request(99, 512, False, 33000.0, False, 7, phases=bad)] + reqs[2:]
    rep = server.run(mixed, keep_results=True)
    assert rep.batches == 1                                  # the five requests shared one batch
    assert rep.status[99] == S.NONFINITE and 99 not in server.results
    assert all(rep.status[r.rid] == S.OK for r in reqs)
    assert _compare(alone, server.results, reqs) == 0.0      # bit for bit as served alone


def test_poisson_stream_report(server):
    rep = server.run(small_stream(60, rate=2000.0, seed=14), offered_rate=2000.0)
    assert rep.completed == 60 and 0 < rep.p50_ms <= rep.p99_ms
    assert server.used_bytes <= server.hbm_budget
    d = rep.summary()
    assert d["offered_rate_req_s"] == 2000.0 and d["throughput_req_s"] > 0


def test_seeded_phase_masks(cuda_dev):
    """xl_uniform_phases: U[-pi, pi) statistics, the same masks for the same
    seed in any batch position, different masks for different seeds, and
    members whose flag is 0 left untouched."""
    n = 256
    out = torch.full((3, 2, n, n), 7.0, dtype=torch.float32, device=cuda_dev)
    seeds = torch.tensor([11, 12, 11], dtype=torch.int64, device=cuda_dev)
    enable = torch.tensor([1, 1, 0], dtype=torch.int32, device=cuda_dev)
    ops.uniform_phases(out, seeds, enable)
    a = out.cpu().numpy()
    assert np.all(a[2] == 7.0)                                   # disabled member untouched
    v = a[:2].ravel()
    assert v.min() >= -math.pi and v.max() < math.pi
    assert abs(v.mean()) < 0.02 and abs(v.var() - math.pi ** 2 / 3) < 0.03     # about 5 standard errors
    hist = np.histogram(v, bins=16, range=(-math.pi, math.pi))[0] / v.size
    assert np.abs(hist - 1 / 16).max() < 0.002
    again = torch.empty((1, 2, n, n), dtype=torch.float32, device=cuda_dev)
    ops.uniform_phases(again, seeds[2:3])                        # seed 11 alone
    assert np.array_equal(again.cpu().numpy()[0], a[0])
    assert not np.array_equal(a[0], a[1])
    assert abs(np.corrcoef(a[0].ravel(), a[1].ravel())[0, 1]) < 0.015
