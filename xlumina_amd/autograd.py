"""torch.autograd.Function wrappers with hand-written adjoints.

Each propagator is linear in the field, so its backward is the adjoint
operator run by the same kernels:
  * RS / VRS: convolution with conj(h) (the RS kernel array is even, so the
    correlation is a convolution with the conjugated transform);
  * CZT / VCZT / high-NA: conj(F) (K_y^H (conj(F0 c) G) conj(K_x)) through the
    Bluestein kernels with adjoint tables;
  * fused Jones masks / Ez construction / lens map: xl_pre_backward, which
    also returns the cotangents of the per-pixel mask parameters.
Only the inputs (field, mask parameters) are saved; masked intermediates are
recomputed in the backward pass, so a chain keeps one field per stage alive.
"""
from __future__ import annotations

import math
import weakref
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import ops
from .ops import MaskSpec


def _mask_structure(masks: Sequence[MaskSpec]):
    """Static part of a mask chain (kinds and scalar values)."""
    return tuple((m.kind, m.s0, m.s1, m.unit) for m in masks)


def _rebuild(struct, tensors):
    out = []
    for i, (kind, s0, s1, unit) in enumerate(struct):
        out.append(MaskSpec(kind, tensors[2 * i], tensors[2 * i + 1], s0, s1, unit))
    return out


def _flat_params(masks: Sequence[MaskSpec]):
    flat = [None, None, None, None]
    for i, m in enumerate(masks[:2]):
        flat[2 * i], flat[2 * i + 1] = m.p0, m.p1
    return flat


def _full_grad(ctx, src, nin):
    """A complex64 buffer of src's shape for the adjoint of a pre-op that
    reads only src[:nin]: xl_pre_backward fills its first nin planes and the
    rest become zeros (no zero-filled copy of the whole field); None where
    src is not a contiguous complex64 (C > nin, rows, cols) field."""
    if (not ctx.needs_input_grad[1] or src.dim() != 3 or src.shape[0] <= nin or src.dtype != torch.complex64
            or not src.is_contiguous()):
        return None
    return torch.empty(src.shape, dtype=torch.complex64, device=src.device)


def _param_grads(ctx, struct, params, gp):
    """Reduce per-pixel parameter cotangents to each parameter's shape."""
    outs = []
    for i in range(2):
        for j in range(2):
            p = params[2 * i + j]
            if p is None or not ctx.needs_input_grad[2 + 2 * i + j]:
                outs.append(None)
                continue
            g = gp[i][j]
            if g is None:
                outs.append(None)
                continue
            if g.shape != p.shape:
                if g.numel() == p.numel():     # a one-value parameter's gradient, summed by the kernels
                    g = g.reshape(p.shape)
                else:
                    g = g.sum() if p.dim() == 0 else g.sum_to_size(p.shape)
            outs.append(g.to(p.dtype))
    return outs


@dataclass
class RSConfig:
    x: object
    y: object
    z: float
    wavelength: float
    ez_mode: int
    nin: int
    mask_struct: tuple = field(default_factory=tuple)
    z_dev: Optional[torch.Tensor] = None   # float64 device scalar read by the kernels instead of z
    qf_out: Optional[torch.Tensor] = None  # 0-d float64 device tensor the forward launch fills with the quality factor


# Kernel and derivative spectra of the differentiated forward propagations,
# by geometry and distance: a later propagation over the same distance (one
# distance tensor, unchanged, or one float) reuses them instead of building
# its own.  An entry holds its key's tensors (so their addresses cannot be
# reused by others while it lives) and weak references to the spectra, which
# live as long as the first propagation's saved state.
_SHARED_SPECTRA: dict = {}


def _spectra_key(cfg: RSConfig, z_t, n: int, ncomp: int, dev):
    xd, yd = ops.device_grid(cfg.x, dev), ops.device_grid(cfg.y, dev)
    zk = ("t", z_t.data_ptr(), z_t._version, str(z_t.device)) if z_t is not None else ("f", float(cfg.z))
    key = (zk, xd.data_ptr(), yd.data_ptr(), n, ncomp, float(cfg.wavelength), str(dev))
    return key, (None if z_t is None else z_t.detach(), xd, yd)


class _RSFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg: RSConfig, src, p00, p01, p10, p11, z_t, wl_t):
        params = (p00, p01, p10, p11)
        masks = _rebuild(cfg.mask_struct, params)
        # a step that will run the adjoint keeps the kernel spectrum for it, and,
        # when z or lambda receive gradients, the input spectra they are dotted with
        n = src.shape[-1]
        ncomp = 3 if cfg.ez_mode == N.XL_EZ_VRS else cfg.nin
        hhat = spec = spec_dg = None
        hhat_mode = 0
        # the derivative spectra come from the forward's own kernel rows
        spec_grads = (1 if ctx.needs_input_grad[6] else 0) | (2 if ctx.needs_input_grad[7] else 0)
        key = None
        if any(ctx.needs_input_grad[1:]):
            key, keep = _spectra_key(cfg, z_t, n, ncomp, src.device)
            hit = _SHARED_SPECTRA.get(key)
            if hit is not None and hit["hhat"]() is not None:
                # an earlier propagation over this distance built the spectra
                hhat, hhat_mode = hit["hhat"](), 2
                if spec_grads and (spec_grads & ~hit["bits"]) == 0 and hit["spec"] is not None:
                    spec_dg = hit["spec"]()
                key = None
            else:
                hhat, hhat_mode = torch.empty(ops.rs_hhat_bytes(n), dtype=torch.uint8, device=src.device), 1
        if spec_grads:
            spec = torch.empty(ops.rs_spec_bytes(n, ncomp), dtype=torch.uint8, device=src.device)
        out = ops.rs_propagate(src, cfg.x, cfg.y, cfg.z, cfg.wavelength, masks=masks, ez_mode=cfg.ez_mode,
                               nin=cfg.nin, z_dev=cfg.z_dev, hhat=hhat, hhat_mode=hhat_mode,
                               spec=spec, spec_mode=1 if spec is not None else 0, spec_grads=spec_grads,
                               quality_factor=cfg.qf_out, spec_dg=spec_dg)
        if key is not None:
            if len(_SHARED_SPECTRA) > 64:
                _SHARED_SPECTRA.clear()
            _SHARED_SPECTRA[key] = {"hhat": weakref.ref(hhat), "spec": weakref.ref(spec) if spec is not None else None,
                                    "bits": spec_grads, "keep": keep}
        cfg.qf_out = None   # the backward's launches report nothing
        ctx.cfg = cfg
        ctx.spec_dg = spec_dg
        ctx.save_for_backward(src, *[p if p is not None else torch.empty(0) for p in params],
                              hhat if hhat is not None else torch.empty(0),
                              spec if spec is not None else torch.empty(0))
        ctx.has = [p is not None for p in params]
        ctx.z_meta = None if z_t is None else (z_t.shape, z_t.dtype, z_t.device)
        ctx.wl_meta = None if wl_t is None else (wl_t.shape, wl_t.dtype, wl_t.device)
        return out

    @staticmethod
    def backward(ctx, gout):
        cfg = ctx.cfg
        saved = ctx.saved_tensors
        src = saved[0]
        params = [saved[1 + i] if ctx.has[i] else None for i in range(4)]
        hhat = saved[5] if saved[5].numel() else None
        spec = saved[6] if saved[6].numel() else None
        gout = gout.contiguous()
        want_z, want_wl = ctx.needs_input_grad[6], ctx.needs_input_grad[7]
        grads = None
        # with the VRS Ez_in term the pre-op adjoint's reduction also sums the
        # column pass's partial sums (one reduction launch instead of two)
        ez_path = want_z and cfg.ez_mode == N.XL_EZ_VRS
        partials = spec is not None and ez_path and not want_wl
        if spec is not None:
            # Re<g, conv(E', dH/dz)> and Re<g, conv(E', dH/dk)> by Parseval in the adjoint's column pass
            grads = None if partials else torch.empty(2, dtype=torch.float64, device=gout.device)
            bits = (1 if want_z else 0) | (2 if want_wl else 0) | N.XL_RS_SPEC_PREBUILT
            g_pre = ops.rs_propagate(gout, cfg.x, cfg.y, cfg.z, cfg.wavelength, adjoint=True, z_dev=cfg.z_dev,
                                     hhat=hhat, hhat_mode=2 if hhat is not None else 0, spec=spec, spec_mode=2,
                                     spec_grads=bits | (N.XL_RS_SPEC_PARTIALS if partials else 0),
                                     grad_out=grads, spec_dg=ctx.spec_dg)
        else:
            g_pre = ops.rs_propagate(gout, cfg.x, cfg.y, cfg.z, cfg.wavelength, adjoint=True, z_dev=cfg.z_dev,
                                     hhat=hhat, hhat_mode=2 if hhat is not None else 0)
        masks = _rebuild(cfg.mask_struct, params)
        need_params = ((ctx.needs_input_grad[2] and params[0] is not None, ctx.needs_input_grad[3] and params[1] is not None),
                       (ctx.needs_input_grad[4] and params[2] is not None, ctx.needs_input_grad[5] and params[3] is not None))
        # the VRS Ez_in construction depends on z too: its term comes from the pre-op adjoint's pass
        ez_dz = ez_base = None
        if ez_path:
            ez_dz = torch.empty(ops.pre_scalar_grad_elems(*src.shape[-2:]), dtype=torch.float64, device=src.device)
            ez_base = ops.rs_spec_partials(spec, src.shape[-1], 3, 0) if partials else grads[:1]
        if not masks and cfg.ez_mode == N.XL_EZ_NONE:
            g_src, gp = g_pre, [[None, None], [None, None]]
        else:
            g_src, gp = ops.pre_backward(src, g_pre, cfg.x, cfg.y, cfg.z, masks=masks, ez_mode=cfg.ez_mode,
                                         nin=cfg.nin, need_src=ctx.needs_input_grad[1], need_params=need_params,
                                         z_dev=cfg.z_dev, out=_full_grad(ctx, src, cfg.nin), ez_dz=ez_dz,
                                         ez_dz_base=ez_base)
        g_z = None
        if want_z:
            # dL/dz = Re<g, conv(E', dH/dz)> (+ the Ez_in term, vectorized_optics.py:262-267)
            shape, dtype, device = ctx.z_meta
            g_z = grads[0] if ez_dz is None else ez_dz[0]   # ez_dz[0] holds both terms
            g_z = g_z.to(device=device, dtype=dtype).reshape(shape)
        g_wl = None
        if want_wl:
            # dL/dlambda = -(k/lambda) Re<g, conv(E', dH/dk)>; Ez_in does not depend on lambda
            shape, dtype, device = ctx.wl_meta
            k = 2 * math.pi / cfg.wavelength
            g_wl = (-(k / cfg.wavelength) * grads[1]).to(device=device, dtype=dtype).reshape(shape)
        if g_src is not None and ctx.needs_input_grad[1]:
            if g_src.shape[0] != src.shape[0]:
                full = torch.zeros_like(src)
                full[: g_src.shape[0]] = g_src
                g_src = full
        else:
            g_src = None
        return (None, g_src, *_param_grads(ctx, cfg.mask_struct, params, gp), g_z, g_wl)


def rs(src: torch.Tensor, x, y, z, wavelength: float, masks: Sequence[MaskSpec] = (),
       ez_mode: int = N.XL_EZ_NONE, nin: Optional[int] = None,
       quality_factor: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable RS / VRS propagation of src (C, n, n).  `z` may be a
    float or a one-element tensor; with requires_grad the distance receives
    dL/dz (Parseval over the forward's and the adjoint's spectra with dH/dz).
    A one-element `wavelength` tensor with requires_grad receives dL/dlambda
    (the same with dH/dk).  `quality_factor`: optional 0-d float64 device
    tensor filled by the forward launch (wave_optics.py:187-191)."""
    nin = src.shape[0] if nin is None else nin
    wl_t = wavelength if isinstance(wavelength, torch.Tensor) and wavelength.requires_grad else None
    if isinstance(wavelength, torch.Tensor):
        wavelength = float(wavelength.detach().reshape(-1)[0])
    z_t = z if isinstance(z, torch.Tensor) else None
    if z_t is not None:
        # the kernels read the distance from device memory: no host round trip
        if z_t.numel() != 1:
            raise ValueError("the propagation distance must be a single value")
        z_dev = z_t.detach().to(device=src.device, dtype=torch.float64).reshape(1)
        cfg = RSConfig(x, y, 0.0, float(wavelength), ez_mode, nin, _mask_structure(masks), z_dev, quality_factor)
    else:
        cfg = RSConfig(x, y, float(z), float(wavelength), ez_mode, nin, _mask_structure(masks),
                       qf_out=quality_factor)
    return _RSFunction.apply(cfg, src, *_flat_params(masks), z_t, wl_t)


@dataclass
class CZTConfig:
    geom: ops.CztGeometry
    ez_mode: int
    nin: int
    mode: int
    mask_struct: tuple = field(default_factory=tuple)
    form: str = "bluestein"   # resolved implementation, used for forward and adjoint


class _StagedTables:
    """Workspaces holding the staged Bluestein tables (chirps and their
    spectra, functions of the geometry only) of one CZT geometry: one for the
    forward calls, one for the adjoints.  Every differentiated call of the
    geometry holds it, so it lives until the last of their backward passes;
    the first call of each direction stages the tables, the others use them."""
    __slots__ = ("fwd", "adj", "__weakref__")

    def __init__(self):
        self.fwd = None
        self.adj = None


_STAGED_TABLES: dict = {}


def _staged_tables(cfg: CZTConfig):
    """The _StagedTables of cfg's geometry, found by content (each call
    builds its own geometry object); None for the GEMM form."""
    if cfg.form != "bluestein":
        return None
    g = cfg.geom
    grids = (g.x.tobytes(), g.y.tobytes(), g.xout.tobytes(), g.yout.tobytes())
    key = (cfg.mode, float(g.z), float(g.wavelength), float(g.dm), int(g.mode), float(g.hna_radius), float(g.hna_f),
           *(ops.content_key(b) for b in grids))
    hit = _STAGED_TABLES.get(key)
    share = hit[1]() if hit is not None and hit[0] == grids else None
    if share is None:
        share = _StagedTables()
        if len(_STAGED_TABLES) > 64:
            _STAGED_TABLES.clear()
        _STAGED_TABLES[key] = (grids, weakref.ref(share))
    return share


def _czt_staged(share, attr: str, src, cfg: CZTConfig, **kw):
    """ops.czt on the workspace share.<attr>: the first call stages the axis
    tables there, later calls with the same arguments use them."""
    if share is None:
        return ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, **kw)
    ws = getattr(share, attr)
    adjoint, ez_mode = kw.get("adjoint", False), kw.get("ez_mode", N.XL_EZ_NONE)
    nin = kw.get("nin") or src.shape[-3]
    ncomp = 3 if (ez_mode == N.XL_EZ_VCZT or cfg.mode == N.XL_CZT_HNA) and not adjoint else nin
    if cfg.mode != cfg.geom.mode:
        return ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, **kw)
    try:
        need = ops.czt_workspace_bytes(cfg.geom, ncomp=ncomp, nin=nin, adjoint=adjoint, ez_mode=ez_mode)
    except ValueError:
        return ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, **kw)   # raises with the plan error
    if ws is not None and ws.numel() >= need and ws.device == src.device:
        return ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, workspace=ws, tables_mode=2, **kw)
    ws = torch.empty(need, dtype=torch.uint8, device=src.device)
    ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, workspace=ws, tables_mode=1, **kw)   # stage only
    setattr(share, attr, ws)
    return ops.czt(src, cfg.geom, mode=cfg.mode, form=cfg.form, workspace=ws, tables_mode=2, **kw)


class _CZTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg: CZTConfig, src, p00, p01, p10, p11, z_t, wl_t):
        params = (p00, p01, p10, p11)
        masks = _rebuild(cfg.mask_struct, params)
        # the calls of one geometry in a differentiated step stage its tables once
        share = _staged_tables(cfg) if any(ctx.needs_input_grad[1:]) else None
        out = _czt_staged(share, "fwd", src, cfg, masks=masks, ez_mode=cfg.ez_mode, nin=cfg.nin)
        ctx.tables = share
        ctx.cfg = cfg
        keep_out = (z_t is not None and ctx.needs_input_grad[6]) or (wl_t is not None and ctx.needs_input_grad[7])
        ctx.save_for_backward(src, *[p if p is not None else torch.empty(0) for p in params],
                              out if keep_out else torch.empty(0))
        ctx.has = [p is not None for p in params]
        ctx.z_meta = None if z_t is None else (z_t.shape, z_t.dtype, z_t.device)
        ctx.wl_meta = None if wl_t is None else (wl_t.shape, wl_t.dtype, wl_t.device)
        return out

    @staticmethod
    def backward(ctx, gout):
        cfg = ctx.cfg
        g = cfg.geom
        saved = ctx.saved_tensors
        src = saved[0]
        params = [saved[1 + i] if ctx.has[i] else None for i in range(4)]
        g_pre = _czt_staged(ctx.tables, "adj", gout.contiguous(), cfg, adjoint=True)
        masks = _rebuild(cfg.mask_struct, params)
        hna = cfg.mode == N.XL_CZT_HNA
        g_z = None
        if ctx.z_meta is not None and ctx.needs_input_grad[6]:
            shape, dtype, device = ctx.z_meta
            g_z = _czt_z_grad(cfg, src, masks, gout.contiguous(), g_pre, saved[5]).to(device=device, dtype=dtype)
            g_z = g_z.reshape(shape)
        g_wl = None
        if ctx.wl_meta is not None and ctx.needs_input_grad[7]:
            shape, dtype, device = ctx.wl_meta
            g_wl = _czt_z_grad(cfg, src, masks, gout.contiguous(), g_pre, saved[5], wrt="wavelength")
            g_wl = g_wl.to(device=device, dtype=dtype).reshape(shape)
        need_params = ((ctx.needs_input_grad[2] and params[0] is not None, ctx.needs_input_grad[3] and params[1] is not None),
                       (ctx.needs_input_grad[4] and params[2] is not None, ctx.needs_input_grad[5] and params[3] is not None))
        if not masks and cfg.ez_mode == N.XL_EZ_NONE and not hna:
            g_src, gp = g_pre, [[None, None], [None, None]]
        else:
            g_src, gp = ops.pre_backward(src, g_pre, g.x, g.y, g.z, masks=masks, ez_mode=cfg.ez_mode, hna=hna,
                                         nin=cfg.nin, need_src=ctx.needs_input_grad[1], need_params=need_params,
                                         hna_radius=g.hna_radius, hna_f=g.hna_f, out=_full_grad(ctx, src, cfg.nin))
        if g_src is not None and ctx.needs_input_grad[1]:
            if g_src.shape[0] != src.shape[0]:
                full = torch.zeros_like(src)
                full[: g_src.shape[0]] = g_src
                g_src = full
        else:
            g_src = None
        return (None, g_src, *_param_grads(ctx, cfg.mask_struct, params, gp), g_z, g_wl)


def _dlnh_dz(X, Y, z: float, k: float):
    """d/dz log h for h = transfer_function_RS (wave_optics.py:291-297):
    h = (1/2pi) (z/r^2)(1/r - i k) e^{i s k r}, s = sign(z), r^2 = X^2 + Y^2 + z^2."""
    r2 = X * X + Y * Y + z * z
    r = torch.sqrt(r2)
    s = 1.0 if z > 0 else -1.0
    return 1.0 / z - 2.0 * z / r2 - (z / (r2 * r)) / (1.0 / r - 1j * k) + 1j * s * k * z / r


def _czt_axis_weights(out_grid, m: int, dm: float, dev):
    """The z-derivative of the CZT kernel of one axis (the dense form of
    Bluestein_method, wave_optics.py:412-460): with W = out[-1] - out[0],
    Delta = W / M and D1 = C + Dm (C = out[0] + W / (2M)), the kernel phase is
    psi(k, j) - 2 pi j' with psi = (p_k - j' q_k), p_k = 2 pi Delta (k + 1/2 -
    (m-1)/2) / Dm, q_k = 2 pi (C + (k+1) Delta) / Dm, j' = j - (m-1)/2, and
    Dm = lambda z / dx makes d phase / dz = -psi / z."""
    o = np.asarray(out_grid, dtype=np.float64)
    M = len(o)
    W = o[-1] - o[0]
    delta = W / M
    C = o[0] + W / (2 * M)
    k = np.arange(M, dtype=np.float64)
    p = 2 * np.pi * delta * (k + 0.5 - (m - 1) / 2) / dm
    q = 2 * np.pi * (C + (k + 1) * delta) / dm
    jp = np.arange(m, dtype=np.float64) - (m - 1) / 2
    t = lambda a: torch.from_numpy(a).to(dev)
    return t(p), t(q), t(jp)


def _dlnh_dk(X, Y, z: float, k: float):
    """d/dk log h for h = transfer_function_RS (wave_optics.py:291-297):
    -i / (1/r - i k) + i s r, s = sign(z)."""
    r = torch.sqrt(X * X + Y * Y + z * z)
    s = 1.0 if z > 0 else -1.0
    return -1j / (1.0 / r - 1j * k) + 1j * s * r


def _czt_z_grad(cfg: CZTConfig, src, masks, gout, g_pre, out, wrt: str = "z"):
    """dL/dz (wrt = "z") or dL/dlambda (wrt = "wavelength") of CZT / VCZT
    (wave_optics.py:333-357, vectorized_optics.py:321-361),
    out = c F0 (K_y (E' F) K_x^T) with c = z dx dy lambda, F, F0 the transfer
    function on the input / output grids, K the per-axis kernels of Dm =
    lambda z / dx and E' the pre-op'd input (for VCZT with Ez_in = (Ex X +
    Ey Y) z / r^2).  With t the parameter:  Re<g, out>/t  +  Re<g, (dF0/F0) out>
    +  the two kernel terms Re<g, (-i/t)(p (.) out - q (.) CZT(j' E'))> per axis
    (output rows / columns weighted by p, q; the input by j'; the kernels
    depend on t only through Dm, which is proportional to both z and lambda)
    +  Re<CZT^H g, (dF/F) E'>  +  (VCZT, t = z) Re<g_pre[2], dEz_in/dz>.
    For lambda, d log F / d lambda = (d log h / dk)(-k / lambda).  Sums in
    float64."""
    g = cfg.geom
    if cfg.mode != N.XL_CZT_RS:
        raise ValueError("distance and wavelength gradients exist for the RS-mode CZT only")
    dev = src.device
    z, k = g.z, 2 * math.pi / g.wavelength
    t = z if wrt == "z" else g.wavelength
    ncomp = out.shape[0]
    if masks or cfg.ez_mode != N.XL_EZ_NONE:
        pre = ops.pre_forward(src, g.x, g.y, z, masks=masks, ez_mode=cfg.ez_mode, ncomp=ncomp, nin=cfg.nin)
    else:
        pre = src[:ncomp]

    def rdot(a, b):
        return torch.real(torch.sum(a.conj().to(torch.complex128) * b.to(torch.complex128)))

    py, qy, jy = _czt_axis_weights(g.yout, len(g.y), g.dm, dev)
    px, qx, jx = _czt_axis_weights(g.xout, len(g.x), g.dm, dev)
    pre_y = (pre * jy.to(torch.float32)[None, :, None]).contiguous()
    pre_x = (pre * jx.to(torch.float32)[None, None, :]).contiguous()
    wy = ops.czt(pre_y, g, nin=ncomp, mode=N.XL_CZT_RS, form=cfg.form)
    wx = ops.czt(pre_x, g, nin=ncomp, mode=N.XL_CZT_RS, form=cfg.form)
    o = out.to(torch.complex128)
    total = rdot(gout, o) / t
    Xo = torch.from_numpy(np.asarray(g.xout, dtype=np.float64)).to(dev)[None, :]
    Yo = torch.from_numpy(np.asarray(g.yout, dtype=np.float64)).to(dev)[:, None]
    if wrt == "z":
        dln = lambda X, Y: _dlnh_dz(X, Y, z, k)
    else:
        dln = lambda X, Y: _dlnh_dk(X, Y, z, k) * (-k / t)
    total = total + rdot(gout, o * dln(Xo, Yo))
    total = total + rdot(gout, (-1j / t) * (py[None, :, None] * o - qy[None, :, None] * wy.to(torch.complex128)))
    total = total + rdot(gout, (-1j / t) * (px[None, None, :] * o - qx[None, None, :] * wx.to(torch.complex128)))
    Xi = torch.from_numpy(np.asarray(g.x, dtype=np.float64)).to(dev)[None, :]
    Yi = torch.from_numpy(np.asarray(g.y, dtype=np.float64)).to(dev)[:, None]
    # Re<CZT^H g, (dF/F) E'>: dF/F holds a large near-constant phase rate (i s k
    # z / r ~ i k in z; -i s k r / lambda ~ -i s k z / lambda in lambda), whose
    # product with sum conj(g_pre) E' cancels only in the sum (adjointness makes
    # that sum <g, out>).  The constant is taken out and applied to <g, out>, so
    # the per-pixel float32 rounding of g_pre meets the varying remainder only.
    s = 1.0 if z > 0 else -1.0
    c0 = 1j * s * k if wrt == "z" else -1j * s * k * z / t
    total = total + rdot(g_pre[:ncomp], pre.to(torch.complex128) * (dln(Xi, Yi) - c0))
    total = total + torch.real(c0 * torch.sum(gout.conj().to(torch.complex128) * o))
    if cfg.ez_mode == N.XL_EZ_VCZT and wrt == "z":
        r2 = Xi * Xi + Yi * Yi + z * z
        dez = (pre[0].to(torch.complex128) * Xi + pre[1].to(torch.complex128) * Yi) * ((r2 - 2 * z * z) / (r2 * r2))
        total = total + rdot(g_pre[2], dez)
    return total


def czt(src: torch.Tensor, geom: ops.CztGeometry, masks: Sequence[MaskSpec] = (), ez_mode: int = N.XL_EZ_NONE,
        nin: Optional[int] = None, mode: Optional[int] = None, form: Optional[str] = None,
        z: Optional[torch.Tensor] = None, wavelength: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable CZT / VCZT / high-NA VCZT of src (C, ny, nx); `form` as
    ops.czt.  `z` / `wavelength`: the distance / wavelength as one-element
    tensors (geom holds their values) when dL/dz / dL/dlambda is wanted
    (RS-mode CZT only)."""
    nin = src.shape[0] if nin is None else nin
    mode = geom.mode if mode is None else mode
    if wavelength is not None and abs(float(wavelength.detach().reshape(-1)[0]) - geom.wavelength) > 1e-12 * geom.wavelength:
        raise ValueError("the wavelength tensor must hold the geometry's wavelength")
    cfg = CZTConfig(geom, ez_mode, nin, mode, _mask_structure(masks), ops.resolve_czt_form(geom, form))
    return _CZTFunction.apply(cfg, src, *_flat_params(masks), z, wavelength)


@dataclass
class PreConfig:
    x: object
    y: object
    z: float
    ez_mode: int
    hna: bool
    nin: int
    ncomp: int
    hna_radius: float
    hna_f: float
    mask_struct: tuple = field(default_factory=tuple)


class _PreFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg: PreConfig, src, p00, p01, p10, p11):
        params = (p00, p01, p10, p11)
        masks = _rebuild(cfg.mask_struct, params)
        out = ops.pre_forward(src, cfg.x, cfg.y, cfg.z, masks=masks, ez_mode=cfg.ez_mode, hna=cfg.hna,
                              ncomp=cfg.ncomp, nin=cfg.nin, hna_radius=cfg.hna_radius, hna_f=cfg.hna_f)
        ctx.cfg = cfg
        ctx.save_for_backward(src, *[p if p is not None else torch.empty(0) for p in params])
        ctx.has = [p is not None for p in params]
        return out

    @staticmethod
    def backward(ctx, gout):
        cfg = ctx.cfg
        saved = ctx.saved_tensors
        src = saved[0]
        params = [saved[1 + i] if ctx.has[i] else None for i in range(4)]
        masks = _rebuild(cfg.mask_struct, params)
        need_params = ((ctx.needs_input_grad[2] and params[0] is not None, ctx.needs_input_grad[3] and params[1] is not None),
                       (ctx.needs_input_grad[4] and params[2] is not None, ctx.needs_input_grad[5] and params[3] is not None))
        g_src, gp = ops.pre_backward(src, gout.contiguous(), cfg.x, cfg.y, cfg.z, masks=masks, ez_mode=cfg.ez_mode,
                                     hna=cfg.hna, nin=cfg.nin, need_src=ctx.needs_input_grad[1],
                                     need_params=need_params, hna_radius=cfg.hna_radius, hna_f=cfg.hna_f)
        if g_src is not None and ctx.needs_input_grad[1]:
            if g_src.shape[0] != src.shape[0]:
                full = torch.zeros_like(src)
                full[: g_src.shape[0]] = g_src
                g_src = full
        else:
            g_src = None
        return (None, g_src, *_param_grads(ctx, cfg.mask_struct, params, gp))


def pre(src: torch.Tensor, x, y, z: float = 0.0, masks: Sequence[MaskSpec] = (), ez_mode: int = N.XL_EZ_NONE,
        hna: bool = False, nin: Optional[int] = None, ncomp: Optional[int] = None, hna_radius: float = 0.0,
        hna_f: float = 0.0) -> torch.Tensor:
    """Differentiable standalone Jones masks / Ez construction / lens map."""
    nin = src.shape[0] if nin is None else nin
    if ncomp is None:
        ncomp = 3 if (hna or ez_mode != N.XL_EZ_NONE) else src.shape[0]
    cfg = PreConfig(x, y, float(z), ez_mode, hna, nin, ncomp, hna_radius, hna_f, _mask_structure(masks))
    return _PreFunction.apply(cfg, src, *_flat_params(masks))
