"""xl_rs_propagate_noalloc (the library's per-device scratch instead of a
workspace argument) against xl_rs_propagate with a caller workspace: the same
kernels, so the outputs must be bit-identical.  Two calls in a row, the
second on a smaller grid, so a stale scratch from the first call would show.
Both paths are called through the C ABI (ops.rs_propagate binds both)."""
import ctypes as C

import numpy as np
import pytest
import torch

from xlumina_amd import _native as N
from xlumina_amd import ops
from xlumina_amd.ops import MaskSpec

pytestmark = pytest.mark.gpu


def _case(n, vectorial, seed, dev):
    rng = np.random.default_rng(seed)
    half = 1.2e3 * n / 1024
    x = np.linspace(-half, half, n)
    nin = 2 if vectorial else 1
    src = torch.from_numpy((rng.standard_normal((nin, n, n)) + 1j * rng.standard_normal((nin, n, n)))
                           .astype(np.complex64)).to(dev)
    if vectorial:
        eta, th = rng.uniform(0, 2 * np.pi, (2, n, n)).astype(np.float32)
        masks = [MaskSpec(N.XL_MASK_LCD, torch.from_numpy(eta).to(dev), torch.from_numpy(th).to(dev))]
        ez = N.XL_EZ_VRS
    else:
        phi = rng.uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
        masks = [MaskSpec(N.XL_MASK_SLM, torch.from_numpy(phi).to(dev))]
        ez = N.XL_EZ_NONE
    return src, x, masks, ez


def _both(n, vectorial, seed, z, dev):
    src, x, masks, ez = _case(n, vectorial, seed, dev)
    ncomp = 3 if vectorial else 1
    ref = torch.empty((ncomp, n, n), dtype=torch.complex64, device=dev)
    ref_i = torch.empty((n, n), dtype=torch.float32, device=dev)
    got = torch.full((ncomp, n, n), float("nan"), dtype=torch.complex64, device=dev)
    got_i = torch.full((n, n), float("nan"), dtype=torch.float32, device=dev)
    ops.rs_propagate(src, x, x, z, 0.635, masks=masks, ez_mode=ez, intensity=True, out=ref, inten=ref_i)
    ops.rs_propagate(src, x, x, z, 0.635, masks=masks, ez_mode=ez, intensity=True, out=got, inten=got_i,
                     noalloc=True)
    torch.cuda.synchronize()
    return ref, ref_i, got, got_i


def test_noalloc_matches_workspace_entry_point_twice_shrinking():
    dev = torch.device("cuda", 0)
    N.ensure_device(dev)
    print("noalloc: VRS n=1024 then RS n=256", flush=True)
    for n, vectorial, seed, z in ((1024, True, 11, 3.0e4), (256, False, 12, 5.0e3)):
        ref, ref_i, got, got_i = _both(n, vectorial, seed, z, dev)
        assert torch.isfinite(got).all() and torch.isfinite(got_i).all()
        assert torch.equal(got, ref), f"n={n}: max |diff| {float((got - ref).abs().max())}"
        assert torch.equal(got_i, ref_i)


def test_noalloc_rejects_grids_above_its_scratch():
    dev = torch.device("cuda", 0)
    N.ensure_device(dev)
    lib = N.load()
    p = N.RsParams()
    p.n = N.XL_RS_NOALLOC_MAX_N + 1
    p.nin = 1
    assert lib.xl_rs_propagate_noalloc(C.byref(p), None) == N.XL_EINVAL
    assert b"up to 1024" in lib.xl_last_error()
    assert lib.xl_rs_workspace_bytes(N.XL_RS_NOALLOC_MAX_N, 3) <= N.XL_RS_NOALLOC_SCRATCH_BYTES
