"""Serving loop (GPU): each result is the same alone and in the mix, and the
same as the eager autograd path; a 512^2 request matches the float64 oracle.

Tolerances: alone vs mix and graph vs eager launches <= 1e-6 relative L2
(identical kernels on identical inputs; no atomics, so bit equality is
expected); served vs autograd path <= 1e-5; intensity vs the float64 oracle
<= 1e-5.
"""
import math

import numpy as np
import pytest
import torch

from oracle import xlumina_ref as R
from tests.conftest import rel_l2
from xlumina_amd import _native as N
from xlumina_amd import autograd as AG
from xlumina_amd import serving as S
from xlumina_amd.ops import MaskSpec

pytestmark = pytest.mark.gpu


def small_stream(count, rate=None, seed=11):
    return S.make_requests(count, rate, seed, sizes=(512, 1024), size_p=(0.6, 0.4), p_vectorial=0.4, p_grad=0.4)


def phases(r, dev):
    nin = 2 if r.vectorial else 1
    gen = torch.Generator(device=dev)
    gen.manual_seed(r.seed)
    p = torch.rand((nin, r.n, r.n), generator=gen, device=dev)
    return p.mul_(2 * math.pi).sub_(math.pi)


def eager_reference(server, r, dev):
    """The request through the differentiable drop-in path (no graphs)."""
    sh = server.shared(r.n)
    p = phases(r, dev)
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


def test_alone_vs_mix(server):
    reqs = small_stream(24)
    server.run(reqs, keep_results=True, serial=True)
    alone = dict(server.results)
    rep = server.run(reqs, keep_results=True)            # all arrive at t = 0: maximal overlap
    mix = dict(server.results)
    assert rep.completed == len(reqs)
    worst = 0.0
    for r in reqs:
        ia, ga = alone[r.rid]
        im, gm = mix[r.rid]
        worst = max(worst, rel_l2(im.cpu().numpy(), ia.cpu().numpy()))
        if r.grad:
            worst = max(worst, rel_l2(gm.cpu().numpy(), ga.cpu().numpy()))
    print(f"alone vs mix: worst rel L2 {worst:.3e} over {len(reqs)} requests")
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
    for r in reqs:
        a, ga = with_graphs[r.rid]
        b, gb = plain.results[r.rid]
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) <= 1e-6
        if r.grad:
            assert rel_l2(ga.cpu().numpy(), gb.cpu().numpy()) <= 1e-6


def test_scalar_512_vs_oracle(server, cuda_dev):
    r = S.Request(0, 512, False, 37000.0, False, 123456789)
    server.run([r], keep_results=True)
    inten = server.results[0][0].cpu().numpy().astype(np.float64)
    sh = server.shared(512)
    phi = phases(r, cuda_dev)[0].cpu().numpy().astype(np.float64)
    u, _ = R.slm(sh.beam_s[0].cpu().numpy().astype(np.complex128), phi)
    E, _ = R.rs_propagation(u, sh.x, sh.y, server.wavelength, r.z)
    assert rel_l2(inten, np.abs(E) ** 2) <= 1e-5


def test_poisson_stream_report(server):
    rep = server.run(small_stream(60, rate=2000.0, seed=14), offered_rate=2000.0)
    assert rep.completed == 60 and 0 < rep.p50_ms <= rep.p99_ms
    assert server.used_bytes <= server.hbm_budget
    d = rep.summary()
    assert d["offered_rate_req_s"] == 2000.0 and d["throughput_req_s"] > 0
