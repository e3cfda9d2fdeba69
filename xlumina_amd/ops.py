"""Torch-facing calls into libxlumina_hip.so (no autograd here).

Fields are complex64 CUDA(ROCm) tensors shaped (C, rows, cols), row-major
[y][x]; mask parameters float32 (rows, cols); grids float64.  Outputs and the
workspace come from PyTorch's caching allocator, the launch goes on the
current stream.  There is no CPU path: non-GPU tensors raise.
"""
from __future__ import annotations

import ctypes as C
import functools
import math
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as N


@dataclass
class MaskSpec:
    """One Jones-mask stage applied before a propagation (see XL_MASK_*).

    p0 / p1 are float32 per-pixel arrays, float64 per-pixel arrays (read by
    the kernels as they are) or one-element float64 tensors (one value for
    every pixel, read from device memory).  Bit j of `unit` marks parameter j
    as a raw value p of the angle 2 pi p - pi, the map the reference's
    optimisers apply (optical_elements.py:1556-1559), formed in the kernels
    for float64 parameters; gradients then reach the raw values."""
    kind: int
    p0: Optional[torch.Tensor] = None
    p1: Optional[torch.Tensor] = None
    s0: float = 0.0
    s1: float = 0.0
    unit: int = 0

    def tensors(self):
        return (self.p0, self.p1)


def _ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(device) -> int:
    """Handle of torch's current stream on `device` (a tensor's device): the
    direct query where torch has it (a C call instead of a Stream object)."""
    if _raw_stream is not None and device.index is not None:
        return _raw_stream(device.index)
    return torch.cuda.current_stream(device).cuda_stream


try:
    from xxhash import xxh3_64_intdigest as _digest
except ImportError:
    _digest = None


def content_key(b: bytes):
    """Dictionary key of a byte string: its 64-bit digest where xxhash is
    present (about 1 us for a 1024-point grid, against 5 us for hashing the
    bytes; entries keep the bytes and a hit compares them), else the bytes."""
    return _digest(b) if _digest is not None else b


_grid_cache: dict = {}
_grid_ids: dict = {}   # (id(array), device) -> (weak reference, its bytes, device copy)


def device_grid(arr, device) -> torch.Tensor:
    """float64 device copy of a 1-D grid, cached by content (a grid array
    seen before is recognised by identity and a comparison of its bytes)."""
    if isinstance(arr, torch.Tensor):
        if arr.device == device and arr.dtype == torch.float64 and arr.is_contiguous():
            return arr
        arr = arr.detach().cpu().numpy()
    if isinstance(arr, np.ndarray):
        ent = _grid_ids.get((id(arr), device))
        if ent is not None and ent[0]() is arr and ent[1] == arr.tobytes():
            return ent[2]
    t = _device_grid_by_content(arr, device)
    if isinstance(arr, np.ndarray):
        if len(_grid_ids) > 256:
            _grid_ids.clear()
        _grid_ids[(id(arr), device)] = (weakref.ref(arr), arr.tobytes(), t)
    return t


def _device_grid_by_content(arr, device) -> torch.Tensor:
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
    ab = a.tobytes()
    key = (content_key(ab), str(device))
    hit = _grid_cache.get(key)
    if hit is not None and hit[0] == ab:
        return hit[1]
    if len(_grid_cache) > 256:
        _grid_cache.clear()
    t = torch.from_numpy(a.copy()).to(device)
    _grid_cache[key] = (ab, t)
    return t


def _prep_field(src: torch.Tensor) -> torch.Tensor:
    if not isinstance(src, torch.Tensor) or src.device.type != "cuda":
        raise N.NativeUnavailable("fields must be complex64 tensors on a ROCm GPU")
    N.ensure_device(src.device)
    if src.dtype != torch.complex64:
        src = src.to(torch.complex64)
    if src.dim() == 2:
        src = src.unsqueeze(0)
    return src.contiguous()


def _prep_param(p, shape, device) -> Optional[torch.Tensor]:
    if p is None:
        return None
    if not isinstance(p, torch.Tensor):
        p = torch.as_tensor(np.asarray(p, dtype=np.float32), device=device)
    p = p.to(device=device, dtype=torch.float32)
    if tuple(p.shape) != tuple(shape):
        p = p.expand(shape)
    return p.contiguous()


def _param_with_form(p, shape, device, unit: bool, allow_f64: bool):
    """(tensor, XL_PF_* form) of one mask parameter: float64 device tensors
    are passed as they are (per pixel, or one value for every pixel), the
    rest as float32 per-pixel arrays (a raw angle then mapped here)."""
    if p is None:
        return None, 0
    if allow_f64 and isinstance(p, torch.Tensor) and p.dtype == torch.float64 and p.device == device:
        if p.numel() == 1:
            t, f = p.reshape(1), N.XL_PF_F64 | N.XL_PF_SCALAR
        else:
            t, f = (p if tuple(p.shape) == tuple(shape) else p.expand(shape)).contiguous(), N.XL_PF_F64
        return t, f | (N.XL_PF_UNIT_ANGLE if unit else 0)
    if unit:
        p = torch.as_tensor(p, dtype=torch.float64, device=device) * (2 * math.pi) - math.pi
    return _prep_param(p, shape, device), 0


def fill_masks(dst, masks: Sequence[MaskSpec], shape, device, nin: int, batch: Optional[int] = None,
               allow_f64: bool = True):
    """Fills the two xl_mask_stage slots; returns the tensors kept alive.

    With `batch`, a parameter may also be a (batch, rows, cols) float32 view
    whose rows are contiguous; every such parameter must share one batch
    stride, returned as the second value ((keep, mask_batch_stride)).
    Batched launches take float32 parameters only (allow_f64 False)."""
    if len(masks) > 2:
        raise ValueError("at most two fused mask stages")
    if batch is not None:
        return _fill_masks_batched(dst, masks, shape, device, batch)
    keep = []
    for i in range(2):
        st = dst[i]
        if i >= len(masks) or masks[i] is None:
            st.kind = N.XL_MASK_NONE
            st.forms = 0
            st.m0 = st.m1 = None
            st.s0 = st.s1 = 0.0
            continue
        m = masks[i]
        p0, f0 = _param_with_form(m.p0, shape, device, bool(m.unit & 1), allow_f64)
        p1, f1 = _param_with_form(m.p1, shape, device, bool(m.unit & 2), allow_f64)
        keep += [p0, p1]
        st.kind = m.kind
        st.forms = f0 | (f1 << 3)
        st.m0 = _ptr(p0) or None
        st.m1 = _ptr(p1) or None
        st.s0 = float(m.s0)
        st.s1 = float(m.s1)
    return keep


def _fill_masks_batched(dst, masks, shape, device, batch):
    keep, strides = [], set()
    simple = []
    for m in masks:
        ps = []
        for p in (m.p0, m.p1):
            if isinstance(p, torch.Tensor) and p.dim() == 3:
                if (tuple(p.shape) != (batch,) + tuple(shape) or p.dtype != torch.float32 or p.device != device
                        or p.stride(2) != 1 or p.stride(1) != shape[1]):
                    raise ValueError(f"batched mask parameter must be float32 {(batch,) + tuple(shape)} with "
                                     "contiguous rows on the field's device")
                strides.add(p.stride(0))
                keep.append(p)
                ps.append(("batched", p))
            else:
                ps.append(("shared", p))
        simple.append(ps)
    if len(strides) > 1:
        raise ValueError("batched mask parameters must share one batch stride")
    if strides and any(kind == "shared" and p is not None for ps in simple for kind, p in ps):
        raise ValueError("mix of batched and shared per-pixel mask parameters")
    flat = [MaskSpec(m.kind, None if isinstance(m.p0, torch.Tensor) and m.p0.dim() == 3 else m.p0,
                     None if isinstance(m.p1, torch.Tensor) and m.p1.dim() == 3 else m.p1, m.s0, m.s1, m.unit)
            for m in masks]
    keep += fill_masks(dst, flat, shape, device, 1, allow_f64=False)
    for i, ps in enumerate(simple):
        for j, (kind, p) in enumerate(ps):
            if kind == "batched":
                setattr(dst[i], "m0" if j == 0 else "m1", p.data_ptr())
    return keep, (strides.pop() if strides else 0)


_qf_cache: dict = {}


def _qf_grid(x, y):
    """(sqrt(dx^2 + dy^2), sqrt(max x^2 + max y^2)) of a grid pair, cached by content."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xb, yb = x.tobytes(), y.tobytes()
    key = (content_key(xb), content_key(yb))
    hit = _qf_cache.get(key)
    if hit is None or hit[0] != xb or hit[1] != yb:
        dx = x[1] - x[0]
        dy = y[1] - y[0]
        hit = (xb, yb, (math.sqrt(dx ** 2 + dy ** 2), math.sqrt(float((x ** 2).max()) + float((y ** 2).max()))))
        if len(_qf_cache) > 64:
            _qf_cache.clear()
        _qf_cache[key] = hit
    return hit[2]


def quality_factor_buffer(z) -> Optional[torch.Tensor]:
    """A 0-d float64 tensor for a propagation launch to fill with the quality
    factor when z is a device tensor (no launch of its own); None otherwise."""
    if isinstance(z, torch.Tensor) and z.device.type == "cuda":
        return torch.empty((), dtype=torch.float64, device=z.device)
    return None


def quality_factor(x, y, wavelength, z):
    """wave_optics.py:187-191 in float64: a float for a float z; for a tensor z
    a tensor on z's device (computed there, so no host synchronisation)."""
    dr_real, rmax = _qf_grid(x, y)
    if isinstance(z, torch.Tensor):
        zt = z.detach().to(torch.float64).reshape(())
        if zt.device.type == "cuda":       # one launch (xl_quality_factor)
            out = torch.empty((), dtype=torch.float64, device=zt.device)
            N.check(N.load().xl_quality_factor(zt.contiguous().data_ptr(), float(rmax), float(dr_real),
                                               float(wavelength), out.data_ptr(), _stream(zt.device)))
            return out
        dr_ideal = torch.sqrt(wavelength ** 2 + rmax ** 2 + 2 * wavelength * torch.sqrt(rmax ** 2 + zt ** 2)) - rmax
        return dr_ideal / dr_real
    dr_ideal = math.sqrt(wavelength ** 2 + rmax ** 2 + 2 * wavelength * math.sqrt(rmax ** 2 + z ** 2)) - rmax
    return dr_ideal / dr_real


def check_buffer(t, shape, dtype, dev, name: str) -> None:
    """Raises ValueError unless `t` is a contiguous `dtype` tensor of `shape`
    on `dev`: the kernels write whole planes through its pointer."""
    if (not isinstance(t, torch.Tensor) or t.dtype != dtype or t.device != torch.device(dev)
            or tuple(t.shape) != tuple(shape) or not t.is_contiguous()):
        got = (f"{tuple(t.shape)} {t.dtype} on {t.device}" if isinstance(t, torch.Tensor) else type(t).__name__)
        raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {tuple(shape)} on {dev} (got {got})")


def check_workspace(ws, dev) -> None:
    if ws is not None and (not isinstance(ws, torch.Tensor) or ws.device != torch.device(dev)
                           or not ws.is_contiguous()):
        raise ValueError(f"workspace must be a contiguous device buffer on {dev}")


def check_rs_args(src_shape, nin: int, ncomp: int, nx: int, ny: int, out, inten, dev) -> None:
    """Shape checks of rs_propagate before any pointer reaches a kernel: the
    component count read, the grid lengths and the caller's output buffers."""
    n = src_shape[-1]
    if not 1 <= nin <= src_shape[0]:
        raise ValueError(f"nin = {nin} but the source holds {src_shape[0]} component(s)")
    if nx != n or ny != n:
        raise ValueError(f"grid lengths ({nx}, {ny}) do not match the {n} x {n} field")
    if out is not None:
        check_buffer(out, (ncomp, n, n), torch.complex64, dev, "out")
    if inten is not None:
        check_buffer(inten, (n, n), torch.float32, dev, "inten")


def _grid_len(g) -> int:
    return int(g.numel()) if isinstance(g, torch.Tensor) else len(np.asarray(g).ravel())


def _z_ptr(z_dev: Optional[torch.Tensor], dev):
    if z_dev is None:
        return None
    if z_dev.dtype != torch.float64 or z_dev.device != dev or z_dev.numel() != 1:
        raise ValueError("z_dev must be a one-element float64 tensor on the field's device")
    return z_dev.data_ptr()


# ---------------------------------------------------------------------------
# RS / VRS
# ---------------------------------------------------------------------------

@functools.lru_cache(maxsize=256)
def rs_workspace_bytes(n: int, ncomp: int) -> int:
    return int(N.load().xl_rs_workspace_bytes(n, ncomp))


@functools.lru_cache(maxsize=256)
def rs_hhat_bytes(n: int) -> int:
    """Bytes of the kernel-spectrum buffer of an n x n RS propagation (hhat)."""
    return int(N.load().xl_rs_hhat_bytes(n))


def rs_spec_partials(spec: torch.Tensor, n: int, ncomp: int, i: int = 0) -> torch.Tensor:
    """float64 view of the per-unit partial sums of derivative i (0: z, 1: k)
    that a spec_mode 2 call writes into `spec` (with XL_RS_SPEC_PARTIALS they
    are left there for a later reduction)."""
    count = C.c_int32(0)
    off = int(N.load().xl_rs_spec_partials(n, ncomp, i, C.byref(count)))
    if off < 0:
        raise ValueError(N.load().xl_last_error().decode())
    start = (-spec.data_ptr()) % 256 + off   # the library aligns the buffer to 256 bytes
    return spec[start:start + 8 * count.value].view(torch.float64)


@functools.lru_cache(maxsize=256)
def rs_spec_bytes(n: int, ncomp: int) -> int:
    """Bytes of the spectrum buffer of an n x n, ncomp-component RS propagation
    whose gradients in z / lambda come from the column pass (spec_mode)."""
    return int(N.load().xl_rs_spec_bytes(n, ncomp))


def rs_propagate(src: torch.Tensor, x, y, z: float, wavelength: float, masks: Sequence[MaskSpec] = (),
                 ez_mode: int = N.XL_EZ_NONE, nin: Optional[int] = None, intensity: bool = False,
                 adjoint: bool = False, out: Optional[torch.Tensor] = None,
                 workspace: Optional[torch.Tensor] = None, z_dev: Optional[torch.Tensor] = None,
                 inten: Optional[torch.Tensor] = None, kernel_dz: int = 0,
                 hhat: Optional[torch.Tensor] = None, hhat_mode: int = 0, noalloc: bool = False,
                 spec: Optional[torch.Tensor] = None, spec_mode: int = 0, spec_grads: int = 0,
                 grad_out: Optional[torch.Tensor] = None, quality_factor: Optional[torch.Tensor] = None,
                 spec_dg: Optional[torch.Tensor] = None):
    """RS_propagation_jit / VRS_propagation_jit (wave_optics.py:281-289,
    vectorized_optics.py:364-373) with fused masks and Ez construction.

    src: (C, n, n) complex64; `nin` components are read (default C).
    Returns the (ncomp, n, n) field, plus |E|^2 if `intensity`.
    z_dev: optional float64 device scalar read by the kernels instead of `z`
    (a captured graph then replays with whatever distance it holds).
    kernel_dz: 1 (or True) convolves with dH/dz instead of H (the z-derivative
    of the propagation of the pre-op'd field), 2 with dH/dk (k = 2 pi/lambda).
    hhat, hhat_mode: 1 builds the kernel spectrum into `hhat` (a device buffer
    of rs_hhat_bytes(n)) and keeps it; 2 reuses what a mode-1 call of the same
    geometry and distance left there (an adjoint call conjugates it).
    spec, spec_mode: 1 keeps the spectra of the padded, masked input in `spec`
    (a device buffer of rs_spec_bytes(n, ncomp)), and builds there the
    derivative spectra named by spec_grads from the kernel rows it evaluates
    for H; 2, on the adjoint call of the same geometry and distance, writes
    grad_out[0] = Re<gout, conv(E', dH/dz)> (spec_grads bit 0) and
    grad_out[1] = Re<gout, conv(E', dH/dk)> (bit 1) into the float64 device
    tensor grad_out (2 elements) by Parseval over those spectra, with no
    derivative propagation; spec_grads | XL_RS_SPEC_PREBUILT takes the
    derivative spectra the mode-1 call built instead of building them.
    spec_dg: the spec buffer of another call (same n, ncomp, grid, distance
    and wavelength, spec_grads covering these bits) whose derivative spectra
    this call uses: a spec_mode 1 call then builds none, a spec_mode 2 call
    (with XL_RS_SPEC_PREBUILT) reads them there.
    quality_factor: optional 0-d float64 device tensor that the launch fills
    with the quality factor of the distance propagated (wave_optics.py:187-191).
    noalloc: run on the library's own per-device scratch
    (xl_rs_propagate_noalloc, n <= 1024) instead of a workspace; with `out`
    (and `inten`) given the call allocates nothing.  Such calls share one
    scratch per device, so they must not overlap on a device.
    """
    lib = N.load()
    src = _prep_field(src)
    dev = src.device
    n = src.shape[-1]
    if src.shape[-2] != n:
        raise ValueError("RS propagation needs a square grid (the reference builds one lag vector for both axes)")
    nin = src.shape[0] if nin is None else nin
    ncomp = 3 if ez_mode == N.XL_EZ_VRS else nin
    check_rs_args(tuple(src.shape), nin, ncomp, _grid_len(x), _grid_len(y), out, inten if intensity else None, dev)
    check_workspace(workspace, dev)
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    if out is None:
        out = torch.empty((ncomp, n, n), dtype=torch.complex64, device=dev)
    if intensity and inten is None:
        inten = torch.empty((n, n), dtype=torch.float32, device=dev)
    if not intensity:
        inten = None
    if noalloc:
        if workspace is not None:
            raise ValueError("noalloc runs on the library's scratch; pass no workspace")
        if n > N.XL_RS_NOALLOC_MAX_N:
            raise ValueError(f"noalloc propagation takes grids up to {N.XL_RS_NOALLOC_MAX_N} per side, got {n}")
    else:
        need = rs_workspace_bytes(n, ncomp)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    p = N.RsParams()
    p.inp = src.data_ptr()
    p.in_comp_stride = n * n
    p.nin = nin
    p.ez_mode = ez_mode
    p.out = out.data_ptr()
    p.out_comp_stride = n * n
    p.intensity = _ptr(inten) or None
    p.n = n
    p.adjoint = 1 if adjoint else 0
    p.x = xd.data_ptr()
    p.y = yd.data_ptr()
    p.z = float(z)
    p.k = 2 * math.pi / wavelength
    xa = np.asarray(x, dtype=np.float64) if not isinstance(x, torch.Tensor) else x.detach().cpu().double().numpy()
    ya = np.asarray(y, dtype=np.float64) if not isinstance(y, torch.Tensor) else y.detach().cpu().double().numpy()
    p.dx = float(xa[1] - xa[0])
    p.dy = float(ya[1] - ya[0])
    p.z_dev = _z_ptr(z_dev, dev)
    p.kernel_dz = int(kernel_dz)
    if hhat_mode:
        if hhat is None or hhat.device != dev or hhat.numel() * hhat.element_size() < rs_hhat_bytes(n):
            raise ValueError("hhat must be a device buffer of rs_hhat_bytes(n) bytes on the field's device")
        p.hhat_mode, p.hhat = int(hhat_mode), hhat.data_ptr()
    if spec_mode:
        if spec is None or spec.device != dev or spec.numel() * spec.element_size() < rs_spec_bytes(n, ncomp):
            raise ValueError("spec must be a device buffer of rs_spec_bytes(n, ncomp) bytes on the field's device")
        p.spec_mode, p.spec, p.spec_grads = int(spec_mode), spec.data_ptr(), int(spec_grads)
        if spec_dg is not None:
            if spec_dg.device != dev or spec_dg.numel() * spec_dg.element_size() < rs_spec_bytes(n, ncomp):
                raise ValueError("spec_dg must be a device buffer of rs_spec_bytes(n, ncomp) bytes on the field's device")
            p.spec_dg = spec_dg.data_ptr()
        if spec_mode == 2 and not (grad_out is None and spec_grads & N.XL_RS_SPEC_PARTIALS):
            if (grad_out is None or grad_out.device != dev or grad_out.dtype != torch.float64
                    or grad_out.numel() < 2 or not grad_out.is_contiguous()):
                raise ValueError("grad_out must be a contiguous float64 device tensor of 2 elements")
            p.grad_out = grad_out.data_ptr()
    if quality_factor is not None:
        check_buffer(quality_factor, (), torch.float64, dev, "quality_factor")
        dr_real, rmax = _qf_grid(x, y)
        p.quality_factor, p.qf_rmax, p.qf_dr_real, p.qf_wavelength = (quality_factor.data_ptr(), float(rmax),
                                                                      float(dr_real), float(wavelength))
    keep = fill_masks(p.mask, masks, (n, n), dev, nin)
    if noalloc:
        N.check(lib.xl_rs_propagate_noalloc(C.byref(p), _stream(dev)))
    else:
        N.check(lib.xl_rs_propagate(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    del keep
    return (out, inten) if intensity else out


def rs_batch_workspace_bytes(n: int, ncomp: int, batch: int) -> int:
    return int(N.load().xl_rs_batch_workspace_bytes(n, ncomp, batch))


def rs_propagate_batched(src: torch.Tensor, x, y, wavelength: float, z_dev: torch.Tensor,
                         masks: Sequence[MaskSpec] = (), ez_mode: int = N.XL_EZ_NONE, nin: Optional[int] = None,
                         intensity: bool = False, adjoint: bool = False, batch: Optional[int] = None,
                         out: Optional[torch.Tensor] = None, inten: Optional[torch.Tensor] = None,
                         workspace: Optional[torch.Tensor] = None, hhat: Optional[torch.Tensor] = None,
                         hhat_mode: int = 0):
    """B independent RS / VRS propagations of one grid in the same launches
    (rs_propagate over a batch, each member with its own distance).

    src: (B, C, n, n) complex64, or one (C, n, n) source shared by the batch
    (then `batch` gives B); z_dev: (B,) float64 device distances, member b
    propagates over z_dev[b]; mask parameters: (B, n, n) float32 views with
    contiguous rows sharing one batch stride, or (n, n) shared by the batch;
    hhat / hhat_mode as in rs_propagate with B spectra (B x rs_hhat_bytes(n)).
    Returns out (B, ncomp, n, n), plus inten (B, n, n) if `intensity`.
    """
    lib = N.load()
    if not isinstance(src, torch.Tensor) or src.device.type != "cuda" or src.dtype != torch.complex64:
        raise N.NativeUnavailable("fields must be complex64 tensors on a ROCm GPU")
    dev = src.device
    N.ensure_device(dev)
    src = src.contiguous()
    shared = src.dim() == 3
    if src.dim() not in (3, 4):
        raise ValueError("src must be (B, C, n, n) or a shared (C, n, n)")
    if batch is None:
        if shared:
            raise ValueError("a source shared by the batch needs batch=")
        batch = src.shape[0]
    if batch < 1 or (not shared and src.shape[0] != batch):
        raise ValueError(f"source holds {src.shape[0]} members, batch is {batch}")
    n = src.shape[-1]
    if src.shape[-2] != n:
        raise ValueError("RS propagation needs a square grid (the reference builds one lag vector for both axes)")
    ncin = src.shape[-3]
    nin = ncin if nin is None else nin
    ncomp = 3 if ez_mode == N.XL_EZ_VRS else nin
    check_rs_args((ncin, n, n), nin, ncomp, _grid_len(x), _grid_len(y), None, None, dev)
    if (not isinstance(z_dev, torch.Tensor) or z_dev.dtype != torch.float64 or z_dev.device != dev
            or z_dev.numel() != batch or not z_dev.is_contiguous()):
        raise ValueError(f"z_dev must be a contiguous float64 tensor of {batch} distances on the field's device")
    if out is None:
        out = torch.empty((batch, ncomp, n, n), dtype=torch.complex64, device=dev)
    check_buffer(out, (batch, ncomp, n, n), torch.complex64, dev, "out")
    if intensity:
        if inten is None:
            inten = torch.empty((batch, n, n), dtype=torch.float32, device=dev)
        check_buffer(inten, (batch, n, n), torch.float32, dev, "inten")
    else:
        inten = None
    check_workspace(workspace, dev)
    need = rs_batch_workspace_bytes(n, ncomp, batch)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    p = N.RsParams()
    p.inp = src.data_ptr()
    p.in_comp_stride = n * n
    p.nin = nin
    p.ez_mode = ez_mode
    p.out = out.data_ptr()
    p.out_comp_stride = n * n
    p.intensity = _ptr(inten) or None
    p.n = n
    p.adjoint = 1 if adjoint else 0
    p.x, p.y = xd.data_ptr(), yd.data_ptr()
    p.k = 2 * math.pi / wavelength
    xa = np.asarray(x, dtype=np.float64) if not isinstance(x, torch.Tensor) else x.detach().cpu().double().numpy()
    ya = np.asarray(y, dtype=np.float64) if not isinstance(y, torch.Tensor) else y.detach().cpu().double().numpy()
    p.dx = float(xa[1] - xa[0])
    p.dy = float(ya[1] - ya[0])
    p.z = 0.0
    p.z_dev = z_dev.data_ptr()
    if hhat_mode:
        if hhat is None or hhat.device != dev or hhat.numel() * hhat.element_size() < batch * rs_hhat_bytes(n):
            raise ValueError("hhat must be a device buffer of batch * rs_hhat_bytes(n) bytes on the field's device")
        p.hhat_mode, p.hhat = int(hhat_mode), hhat.data_ptr()
    keep, p.mask_batch_stride = fill_masks(p.mask, masks, (n, n), dev, nin, batch=batch)
    p.batch = batch
    p.in_batch_stride = 0 if shared else ncin * n * n
    p.out_batch_stride = ncomp * n * n
    p.intensity_batch_stride = n * n
    N.check(lib.xl_rs_propagate(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    del keep
    return (out, inten) if intensity else out


# ---------------------------------------------------------------------------
# CZT / VCZT / high-NA
# ---------------------------------------------------------------------------

@dataclass
class CztGeometry:
    """Host float64 set-up of one CZT (build_CZT_grid wave_optics.py:299-331 or
    build_high_NA_VCZT_grid optical_elements.py:640-672)."""
    x: np.ndarray
    y: np.ndarray
    xout: np.ndarray
    yout: np.ndarray
    z: float
    wavelength: float
    dm: float
    mode: int
    cte: complex
    hna_radius: float = 0.0
    hna_f: float = 0.0

    @property
    def fy(self):
        return self.yout[0] + self.dm / 2, self.yout[-1] + self.dm / 2

    @property
    def fx(self):
        return self.xout[0] + self.dm / 2, self.xout[-1] + self.dm / 2


def _f64(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def czt_geometry(x, y, xout, yout, z: float, wavelength: float) -> CztGeometry:
    """build_CZT_grid (wave_optics.py:299-331) and CZT_jit's output constant."""
    x, y, xout, yout = _f64(x), _f64(y), _f64(xout), _f64(yout)
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    dm = wavelength * z / dx
    return CztGeometry(x, y, xout, yout, float(z), float(wavelength), float(dm), N.XL_CZT_RS,
                       complex(z * dx * dy * wavelength))


def hna_geometry(x, y, xout, yout, wavelength: float, radius: float, f: float) -> CztGeometry:
    """build_high_NA_VCZT_grid (optical_elements.py:640-672) and the constant
    -i sin^2(theta_max)/(f lambda) of VCZT_objective_lens (:627)."""
    x, y, xout, yout = _f64(x), _f64(y), _f64(xout), _f64(yout)
    dm = f * wavelength * (len(x) - 1) / (2 * radius)
    stm = radius / math.sqrt(radius ** 2 + f ** 2)
    return CztGeometry(x, y, xout, yout, 0.0, float(wavelength), float(dm), N.XL_CZT_HNA,
                       complex(-(1j * stm ** 2 / (f * wavelength))), float(radius), float(f))


def _czt_params(g: CztGeometry, dev, ncomp, nin, adjoint, ez_mode, mode=None):
    p = N.CztParams()
    p.nin = nin
    p.ncomp = ncomp
    p.ny_in, p.nx_in = len(g.y), len(g.x)
    p.my, p.mx = len(g.yout), len(g.xout)
    p.mode = g.mode if mode is None else mode
    p.adjoint = 1 if adjoint else 0
    p.ez_mode = ez_mode
    p.z = g.z
    p.wavelength = g.wavelength
    p.k = 2 * math.pi / g.wavelength
    p.dm = g.dm
    p.fy1, p.fy2 = g.fy
    p.fx1, p.fx2 = g.fx
    p.cte_re, p.cte_im = g.cte.real, g.cte.imag
    p.hna_radius, p.hna_f = g.hna_radius, g.hna_f
    keep = [device_grid(g.x, dev), device_grid(g.y, dev), device_grid(g.xout, dev), device_grid(g.yout, dev)]
    p.x, p.y, p.xout, p.yout = (t.data_ptr() for t in keep)
    return p, keep


CZT_FORMS = ("auto", "bluestein", "gemm")
_czt_form = "auto"
# Output-window size (per axis) below which "auto" picks the split-BF16 MFMA
# zoom-DFT over the FP32 Bluestein FFT form (measured on MI355X; see bench.py).
GEMM_CROSSOVER = 160


def set_czt_form(form: str) -> None:
    """Select the CZT implementation: 'bluestein' (FP32 FFT), 'gemm' (split-BF16
    MFMA zoom-DFT) or 'auto' (gemm when both output axes are short)."""
    global _czt_form
    if form not in CZT_FORMS:
        raise ValueError(f"CZT form must be one of {CZT_FORMS}")
    _czt_form = form


def get_czt_form() -> str:
    return _czt_form


def resolve_czt_form(g: "CztGeometry", form: Optional[str] = None) -> str:
    form = form or _czt_form
    if form == "auto":
        small = max(len(g.xout), len(g.yout)) <= GEMM_CROSSOVER
        return "gemm" if small else "bluestein"
    return form


def czt(src: torch.Tensor, g: CztGeometry, masks: Sequence[MaskSpec] = (), ez_mode: int = N.XL_EZ_NONE,
        nin: Optional[int] = None, ncomp: Optional[int] = None, adjoint: bool = False, mode: Optional[int] = None,
        out: Optional[torch.Tensor] = None, workspace: Optional[torch.Tensor] = None,
        form: Optional[str] = None, batch: Optional[int] = None, tables_mode: int = 0) -> torch.Tensor:
    """CZT_jit / VCZT_jit / CZT_for_high_NA_jit / VCZT_objective_lens core.

    Forward: src (C, ny, nx) -> (ncomp, my, mx).  Adjoint: src (ncomp, my, mx)
    cotangent -> (ncomp, ny, nx) cotangent of the pre-op output.  `form`:
    'bluestein', 'gemm' or 'auto' (default: the module setting).
    `batch`: B independent problems of this geometry in the same launches
    (Bluestein form): src (B, C, ...) or one shared (C, ...), mask parameters
    (B, rows, cols) views or shared, output (B, ncomp, ...).
    `tables_mode` (Bluestein form, with a caller `workspace`): 1 only stages the
    axis tables into the workspace (src and out are not touched), 2 uses the
    tables a mode-1 call with the same arguments staged there (a mode-0 call
    with a caller workspace leaves the same tables there too).
    """
    if batch is not None:
        return _czt_batched(src, g, masks, ez_mode, nin, ncomp, adjoint, mode, out, workspace, batch)
    lib = N.load()
    src = _prep_field(src)
    dev = src.device
    nin = src.shape[0] if nin is None else nin
    mode_eff = g.mode if mode is None else mode
    if ncomp is None:
        ncomp = 3 if (ez_mode == N.XL_EZ_VCZT or mode_eff == N.XL_CZT_HNA) and not adjoint else nin
    p, keep = _czt_params(g, dev, ncomp, nin, adjoint, ez_mode, mode)
    rows_in, cols_in = (p.ny_in, p.nx_in) if not adjoint else (p.my, p.mx)
    rows_out, cols_out = (p.my, p.mx) if not adjoint else (p.ny_in, p.nx_in)
    check_czt_args(tuple(src.shape), nin, ncomp, (rows_in, cols_in), (rows_out, cols_out), out, dev)
    check_workspace(workspace, dev)
    if out is None:
        out = torch.empty((ncomp, rows_out, cols_out), dtype=torch.complex64, device=dev)
    p.inp = src.data_ptr()
    p.in_comp_stride = rows_in * cols_in
    p.out = out.data_ptr()
    p.out_comp_stride = rows_out * cols_out
    keep += fill_masks(p.mask, masks, (rows_in, cols_in), dev, nin)
    use_gemm = resolve_czt_form(g, form) == "gemm"
    wsfn = lib.xl_czt_gemm_workspace_bytes if use_gemm else lib.xl_czt_workspace_bytes
    runfn = lib.xl_czt_gemm if use_gemm else lib.xl_czt
    need = wsfn(C.byref(p))
    if need == 0:
        N.check(runfn(C.byref(p), None, 0, _stream(dev)))   # raises with the plan error
    if tables_mode:
        if use_gemm or workspace is None or workspace.numel() < need:
            raise ValueError("staged CZT tables need the Bluestein form and a caller workspace of "
                             "czt_workspace_bytes")
        p.tables_mode = int(tables_mode)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    N.check(runfn(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    del keep
    return out


def check_czt_args(src_shape, nin: int, ncomp: int, plane_in, plane_out, out, dev) -> None:
    """Shape checks of a single CZT: input plane, components read, output buffer."""
    if tuple(src_shape[-2:]) != tuple(plane_in):
        raise ValueError(f"CZT input is {tuple(src_shape[-2:])}, expected {tuple(plane_in)}")
    if not 1 <= nin <= src_shape[0]:
        raise ValueError(f"nin = {nin} but the source holds {src_shape[0]} component(s)")
    if not 1 <= ncomp <= 3:
        raise ValueError(f"ncomp must be 1..3 (got {ncomp})")
    if out is not None:
        check_buffer(out, (ncomp,) + tuple(plane_out), torch.complex64, dev, "out")


def _czt_batched(src, g, masks, ez_mode, nin, ncomp, adjoint, mode, out, workspace, batch):
    lib = N.load()
    if not isinstance(src, torch.Tensor) or src.device.type != "cuda" or src.dtype != torch.complex64:
        raise N.NativeUnavailable("fields must be complex64 tensors on a ROCm GPU")
    N.ensure_device(src.device)
    dev = src.device
    src = src.contiguous()
    shared = src.dim() == 3
    if not shared and (src.dim() != 4 or src.shape[0] != batch):
        raise ValueError(f"batched CZT input must be (C, rows, cols) or ({batch}, C, rows, cols)")
    cdim = src.shape[-3]
    nin = cdim if nin is None else nin
    mode_eff = g.mode if mode is None else mode
    if ncomp is None:
        ncomp = 3 if (ez_mode == N.XL_EZ_VCZT or mode_eff == N.XL_CZT_HNA) and not adjoint else nin
    p, keep = _czt_params(g, dev, ncomp, nin, adjoint, ez_mode, mode)
    rows_in, cols_in = (p.ny_in, p.nx_in) if not adjoint else (p.my, p.mx)
    rows_out, cols_out = (p.my, p.mx) if not adjoint else (p.ny_in, p.nx_in)
    if tuple(src.shape[-2:]) != (rows_in, cols_in):
        raise ValueError(f"CZT input is {tuple(src.shape[-2:])}, expected {(rows_in, cols_in)}")
    if not 1 <= nin <= cdim:
        raise ValueError(f"nin = {nin} but the source holds {cdim} component(s)")
    check_workspace(workspace, dev)
    if out is None:
        out = torch.empty((batch, ncomp, rows_out, cols_out), dtype=torch.complex64, device=dev)
    check_buffer(out, (batch, ncomp, rows_out, cols_out), torch.complex64, dev, "batched CZT output")
    p.inp = src.data_ptr()
    p.in_comp_stride = rows_in * cols_in
    p.out = out.data_ptr()
    p.out_comp_stride = rows_out * cols_out
    p.batch = batch
    p.in_batch_stride = 0 if shared else cdim * rows_in * cols_in
    p.out_batch_stride = ncomp * rows_out * cols_out
    mk, p.mask_batch_stride = fill_masks(p.mask, masks, (rows_in, cols_in), dev, nin, batch=batch)
    keep += mk
    need = lib.xl_czt_workspace_bytes(C.byref(p))
    if need == 0:
        N.check(lib.xl_czt(C.byref(p), None, 0, _stream(dev)))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    N.check(lib.xl_czt(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    del keep
    return out


def czt_workspace_bytes(g: CztGeometry, ncomp: int = 1, nin: Optional[int] = None, adjoint: bool = False,
                        batch: int = 1, ez_mode: int = N.XL_EZ_NONE) -> int:
    """Workspace of the Bluestein CZT of geometry g (for preallocation)."""
    p = N.CztParams()
    p.nin, p.ncomp = (ncomp if nin is None else nin), ncomp
    p.ny_in, p.nx_in = len(g.y), len(g.x)
    p.my, p.mx = len(g.yout), len(g.xout)
    p.mode, p.adjoint, p.ez_mode = g.mode, 1 if adjoint else 0, ez_mode
    p.dm = g.dm
    p.fy1, p.fy2 = g.fy
    p.fx1, p.fx2 = g.fx
    p.batch = batch
    need = N.load().xl_czt_workspace_bytes(C.byref(p))
    if need == 0:
        raise ValueError(N.load().xl_last_error().decode())
    return int(need)


# ---------------------------------------------------------------------------
# standalone pre-ops
# ---------------------------------------------------------------------------

def _pre_params(src, x, y, z, masks, ez_mode, hna, ncomp, nin, hna_radius, hna_f, z_dev=None):
    dev = src.device
    rows, cols = src.shape[-2:]
    p = N.PreParams()
    p.inp = src.data_ptr()
    p.in_comp_stride = rows * cols
    p.nin = nin
    p.ncomp = ncomp
    p.comp_stride = rows * cols
    p.rows, p.cols = rows, cols
    p.ez_mode = ez_mode
    p.hna = 1 if hna else 0
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    p.x, p.y = xd.data_ptr(), yd.data_ptr()
    p.z = float(z)
    p.hna_radius, p.hna_f = float(hna_radius), float(hna_f)
    p.z_dev = _z_ptr(z_dev, dev)
    keep = [xd, yd] + fill_masks(p.mask, masks, (rows, cols), dev, nin)
    return p, keep


def pre_forward(src: torch.Tensor, x, y, z: float = 0.0, masks: Sequence[MaskSpec] = (),
                ez_mode: int = N.XL_EZ_NONE, hna: bool = False, ncomp: Optional[int] = None,
                nin: Optional[int] = None, hna_radius: float = 0.0, hna_f: float = 0.0) -> torch.Tensor:
    lib = N.load()
    src = _prep_field(src)
    nin = src.shape[0] if nin is None else nin
    if ncomp is None:
        ncomp = 3 if (hna or ez_mode != N.XL_EZ_NONE) else nin
    p, keep = _pre_params(src, x, y, z, masks, ez_mode, hna, ncomp, nin, hna_radius, hna_f)
    out = torch.empty((ncomp,) + tuple(src.shape[-2:]), dtype=torch.complex64, device=src.device)
    p.out = out.data_ptr()
    N.check(lib.xl_pre_forward(C.byref(p), _stream(src.device)))
    del keep
    return out


def pre_backward(src: torch.Tensor, gout: torch.Tensor, x, y, z: float = 0.0, masks: Sequence[MaskSpec] = (),
                 ez_mode: int = N.XL_EZ_NONE, hna: bool = False, nin: Optional[int] = None,
                 need_src: bool = True, need_params=((False, False), (False, False)),
                 hna_radius: float = 0.0, hna_f: float = 0.0, z_dev: Optional[torch.Tensor] = None,
                 gparams=None, out: Optional[torch.Tensor] = None, ez_dz: Optional[torch.Tensor] = None,
                 ez_dz_base: Optional[torch.Tensor] = None):
    """Adjoint of pre_forward: returns (grad of src[:nin], [[g00, g01], [g10, g11]]).
    `gparams` may hold preallocated (rows, cols) float32 outputs in the
    [[g00, g01], [g10, g11]] slots that are needed.  `out`: a contiguous
    complex64 (C >= nin, rows, cols) tensor that receives the gradient of src
    in its first nin planes and zeros in the others (the components the
    operator does not read); it is then returned whole.  `ez_dz` (XL_EZ_VRS):
    a float64 device buffer of pre_scalar_grad_elems(rows, cols) whose element 0
    receives the distance dependence of the constructed Ez_in through its
    cotangent (the `pre` term of dz_dot), formed in the same pass, plus the
    sum of the float64 device values `ez_dz_base` (fixed order) when given --
    one value, or the partial sums an adjoint propagation left in its spec
    buffer (rs_spec_partials)."""
    lib = N.load()
    src = _prep_field(src)
    gout = _prep_field(gout)
    nin = src.shape[0] if nin is None else nin
    ncomp = gout.shape[0]
    p, keep = _pre_params(src, x, y, z, masks, ez_mode, hna, ncomp, nin, hna_radius, hna_f, z_dev)
    p.gout = gout.data_ptr()
    rows, cols = src.shape[-2:]
    if out is not None and need_src:
        if (out.dtype != torch.complex64 or out.device != src.device or out.dim() != 3 or out.shape[0] < nin
                or tuple(out.shape[1:]) != (rows, cols) or not out.is_contiguous()):
            raise ValueError(f"out must be a contiguous complex64 (>= {nin}, {rows}, {cols}) tensor on the field's device")
        gsrc = out
        p.out_planes = out.shape[0]      # the kernel writes zeros beyond nin
    else:
        gsrc = torch.empty((nin, rows, cols), dtype=torch.complex64, device=src.device) if need_src else None
    p.out = _ptr(gsrc) or None
    gp = [[None, None], [None, None]]
    for i in range(2):
        for j in range(2):
            if need_params[i][j]:
                # the gradient of the parameter as stored (see MaskSpec): float32 or float64
                # per pixel, or for a one-value parameter the sum at element 0 of its buffer
                form = (p.mask[i].forms >> (3 * j)) & 7
                given = gparams[i][j] if gparams is not None else None
                if form & N.XL_PF_SCALAR:
                    buf = torch.empty(pre_scalar_grad_elems(rows, cols), dtype=torch.float64, device=src.device)
                    gp[i][j] = buf[:1]
                elif given is not None:
                    if form and given.dtype != torch.float64:
                        raise ValueError("a float64 mask parameter takes a float64 gradient buffer")
                    buf = gp[i][j] = given
                else:
                    buf = gp[i][j] = torch.empty((rows, cols), dtype=torch.float64 if form else torch.float32,
                                                 device=src.device)
                p.gparam[i][j] = buf.data_ptr()
            else:
                p.gparam[i][j] = None
    if ez_dz is not None:
        if (ez_dz.dtype != torch.float64 or ez_dz.device != src.device or not ez_dz.is_contiguous()
                or ez_dz.numel() < pre_scalar_grad_elems(rows, cols)):
            raise ValueError("ez_dz must be a contiguous float64 device buffer of pre_scalar_grad_elems(rows, cols)")
        p.ez_dz = ez_dz.data_ptr()
        if ez_dz_base is not None:
            if (ez_dz_base.dtype != torch.float64 or ez_dz_base.device != src.device or ez_dz_base.numel() < 1
                    or not ez_dz_base.is_contiguous()):
                raise ValueError("ez_dz_base must be a contiguous float64 device tensor")
            p.ez_dz_base, p.ez_dz_base_n = ez_dz_base.data_ptr(), ez_dz_base.numel()
    N.check(lib.xl_pre_backward(C.byref(p), _stream(src.device)))
    del keep
    return gsrc, gp


def pre_scalar_grad_elems(rows: int, cols: int) -> int:
    """float64 elements of a pre_backward buffer summed over the pixels (a
    one-value parameter's gradient, the Ez_in distance term)."""
    return int(N.load().xl_pre_scalar_grad_elems(rows, cols))


def uniform_phases(out: torch.Tensor, seeds: torch.Tensor, enable: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Seeded uniform phase masks in [-pi, pi) (xl_uniform_phases): out is a
    (B, ...) float32 device tensor, member b filled from seeds[b] (a (B,)
    int64 device tensor) where enable[b] != 0 (a (B,) int32 device tensor, or
    None for every member).  Counter-based, so a captured graph regenerates
    new masks from new seeds."""
    lib = N.load()
    if out.dtype != torch.float32 or out.device.type != "cuda" or out.dim() < 2 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 (B, ...) device tensor")
    dev = out.device
    N.ensure_device(dev)
    B = out.shape[0]
    for t, dt, name in ((seeds, torch.int64, "seeds"), (enable, torch.int32, "enable")):
        if t is not None and (t.dtype != dt or t.device != dev or t.numel() != B or not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous {dt} tensor of {B} values on the output's device")
    count = out[0].numel()
    N.check(lib.xl_uniform_phases(out.data_ptr(), seeds.data_ptr(), _ptr(enable) or None, count, B, count,
                                  _stream(dev)))
    return out


def transfer_function_rs(x, y, z: float, wavelength: float, device) -> torch.Tensor:
    """transfer_function_RS (wave_optics.py:291-297) on meshgrid(x, y)."""
    lib = N.load()
    N.ensure_device(device)
    xd, yd = device_grid(x, device), device_grid(y, device)
    out = torch.empty((len(yd), len(xd)), dtype=torch.complex64, device=device)
    N.check(lib.xl_transfer_function_rs(out.data_ptr(), xd.data_ptr(), yd.data_ptr(), len(xd), len(yd), float(z),
                                        2 * math.pi / wavelength, _stream(device)))
    return out


def pre_backward_batched(src: torch.Tensor, gout: torch.Tensor, x, y, batch: int, z: float = 0.0,
                         masks: Sequence[MaskSpec] = (), ez_mode: int = N.XL_EZ_NONE, nin: Optional[int] = None,
                         need_src: bool = True, need_params=((False, False), (False, False)),
                         gparams=None, out: Optional[torch.Tensor] = None, z_dev: Optional[torch.Tensor] = None):
    """pre_backward over a batch in one launch.  src: (B, nin, rows, cols) or one
    shared (nin, rows, cols); gout: (B, ncomp, rows, cols); mask parameters
    (B, rows, cols) views or shared.  Returns (grad of src (B, nin, ...) or
    None, [[g00, g01], [g10, g11]] each (B, rows, cols) or None); `gparams`
    may hold preallocated (B, rows, cols) views sharing one batch stride.
    z_dev: optional (B,) float64 device distances of the Ez construction, one
    per member (instead of `z`)."""
    lib = N.load()
    for t in (src, gout):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.complex64:
            raise N.NativeUnavailable("fields must be complex64 tensors on a ROCm GPU")
    dev = src.device
    N.ensure_device(dev)
    src, gout = src.contiguous(), gout.contiguous()
    rows, cols = src.shape[-2:]
    shared = src.dim() == 3
    nin = src.shape[-3] if nin is None else nin
    if gout.dim() != 4 or gout.shape[0] != batch:
        raise ValueError(f"cotangent must be ({batch}, ncomp, rows, cols)")
    ncomp = gout.shape[1]
    p = N.PreParams()
    p.inp = src.data_ptr()
    p.in_comp_stride = rows * cols
    p.nin, p.ncomp = nin, ncomp
    p.comp_stride = rows * cols
    p.rows, p.cols = rows, cols
    p.ez_mode = ez_mode
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    p.x, p.y = xd.data_ptr(), yd.data_ptr()
    p.z = float(z)
    p.gout = gout.data_ptr()
    keep, p.mask_batch_stride = fill_masks(p.mask, masks, (rows, cols), dev, nin, batch=batch)
    p.batch = batch
    p.in_batch_stride = 0 if shared else src.shape[-3] * rows * cols
    p.gout_batch_stride = ncomp * rows * cols
    gsrc = None
    if need_src:
        gsrc = out if out is not None else torch.empty((batch, nin, rows, cols), dtype=torch.complex64, device=dev)
        if tuple(gsrc.shape) != (batch, nin, rows, cols) or not gsrc.is_contiguous():
            raise ValueError("source-gradient output must be a contiguous (B, nin, rows, cols) tensor")
        p.out = gsrc.data_ptr()
        p.out_batch_stride = nin * rows * cols
    gp = [[None, None], [None, None]]
    strides = set()
    for i in range(2):
        for j in range(2):
            if need_params[i][j]:
                t = gparams[i][j] if gparams is not None and gparams[i][j] is not None else \
                    torch.empty((batch, rows, cols), dtype=torch.float32, device=dev)
                if (tuple(t.shape) != (batch, rows, cols) or t.dtype != torch.float32 or t.device != dev
                        or t.stride(2) != 1 or t.stride(1) != cols or (batch > 1 and t.stride(0) < rows * cols)):
                    raise ValueError("parameter-gradient outputs must be float32 (B, rows, cols) on the field's "
                                     "device with contiguous rows and one array per member")
                strides.add(t.stride(0))
                gp[i][j] = t
                p.gparam[i][j] = t.data_ptr()
            else:
                p.gparam[i][j] = None
    if len(strides) > 1:
        raise ValueError("parameter-gradient outputs must share one batch stride")
    p.gparam_batch_stride = strides.pop() if strides else 0
    if z_dev is not None:
        if (z_dev.dtype != torch.float64 or z_dev.device != dev or z_dev.numel() != batch
                or not z_dev.is_contiguous()):
            raise ValueError(f"z_dev must be a contiguous float64 tensor of {batch} distances on the field's device")
        p.z_dev = z_dev.data_ptr()
    N.check(lib.xl_pre_backward(C.byref(p), _stream(dev)))
    del keep
    return gsrc, gp


# ---------------------------------------------------------------------------
# loss and optimiser steps of batched optimisation
# ---------------------------------------------------------------------------

def mse_intensity(field: torch.Tensor, target_intensity: torch.Tensor, grad: Optional[torch.Tensor] = None,
                  loss: Optional[torch.Tensor] = None, want_grad: bool = True, scale: float = 1.0,
                  workspace: Optional[torch.Tensor] = None):
    """MSE_Intensity (loss_functions.py:137-149) per batch member of `field`
    (B, rows, cols) against |T|^2 = target_intensity (rows, cols).  Returns
    (loss (B,), cotangent (B, rows, cols) or None); `grad` may alias `field`."""
    lib = N.load()
    if field.dtype != torch.complex64 or field.device.type != "cuda" or not field.is_contiguous():
        raise N.NativeUnavailable("field must be a contiguous complex64 tensor on a ROCm GPU")
    dev = field.device
    N.ensure_device(dev)
    B = field.shape[0]
    npix = field[0].numel()
    tgt = target_intensity.to(device=dev, dtype=torch.float32).contiguous()
    if tgt.numel() != npix:
        raise ValueError("target must have one value per pixel")
    if loss is None:
        loss = torch.empty(B, dtype=torch.float32, device=dev)
    if want_grad and grad is None:
        grad = torch.empty_like(field)
    need = lib.xl_mse_workspace_bytes(npix, B)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    p = N.MseParams()
    p.field = field.data_ptr()
    p.field_batch_stride = npix
    p.target = tgt.data_ptr()
    p.grad = grad.data_ptr() if want_grad else None
    p.grad_batch_stride = npix
    p.loss = loss.data_ptr()
    p.npix = npix
    p.batch = B
    p.scale = scale
    N.check(lib.xl_mse_intensity(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    return loss, (grad if want_grad else None)


def adam_step(param: torch.Tensor, grad: torch.Tensor, m: torch.Tensor, v: torch.Tensor, count: torch.Tensor,
              lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0) -> None:
    """In-place Adam / AdamW step (optax.adam / optax.adamw) over contiguous
    float32 tensors; `count` is a one-element int32 device tensor, advanced."""
    lib = N.load()
    for t in (param, grad, m, v):
        if t.dtype != torch.float32 or t.device.type != "cuda" or not t.is_contiguous() or t.numel() != param.numel():
            raise ValueError("Adam tensors must be contiguous float32 GPU tensors of one size")
    if count.dtype != torch.int32 or count.numel() != 1 or count.device != param.device:
        raise ValueError("count must be a one-element int32 tensor on the parameters' device")
    N.ensure_device(param.device)
    p = N.AdamParams()
    p.param, p.grad, p.m, p.v = param.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr()
    p.n = param.numel()
    p.count = count.data_ptr()
    p.lr, p.b1, p.b2, p.eps, p.weight_decay = lr, b1, b2, eps, weight_decay
    N.check(lib.xl_adam_step(C.byref(p), _stream(param.device)))


def _gpu_c64(t: torch.Tensor, what: str) -> torch.Tensor:
    if t.dtype != torch.complex64 or t.device.type != "cuda":
        raise N.NativeUnavailable(f"{what} must be a complex64 tensor on a ROCm GPU")
    return t.contiguous()


def field_loss(field: torch.Tensor, target: torch.Tensor, mode: int, want_grad: bool = False,
               scale: float = 1.0, grad: Optional[torch.Tensor] = None):
    """MSE_Intensity / MSE_Amplitude / MSE_Phase (loss_functions.py:107-149)
    of each member of `field` (B, rows, cols) against `target` (B, rows, cols)
    or one shared (rows, cols) target.  Returns (loss (B,), cotangent or None)."""
    lib = N.load()
    field = _gpu_c64(field, "field")
    dev = field.device
    N.ensure_device(dev)
    B = field.shape[0]
    npix = field[0].numel()
    tgt = _gpu_c64(target.to(dev), "target")
    if tgt.numel() == npix:
        tbs = 0
    elif tgt.numel() == B * npix:
        tbs = npix
    else:
        raise ValueError("target must be one field or one field per batch member")
    loss = torch.empty(B, dtype=torch.float32, device=dev)
    if want_grad and grad is None:
        grad = torch.empty_like(field)
    need = lib.xl_field_loss_workspace_bytes(npix, B)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    p = N.FieldLossParams()
    p.field, p.field_batch_stride = field.data_ptr(), npix
    p.target, p.target_batch_stride = tgt.data_ptr(), tbs
    p.grad = grad.data_ptr() if want_grad else None
    p.grad_batch_stride = npix
    p.loss = loss.data_ptr()
    p.npix, p.batch, p.mode, p.scale = npix, B, mode, scale
    N.check(lib.xl_field_loss(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
    return loss, (grad if want_grad else None)


def small_area(intensity: torch.Tensor, want_grad: bool = False, scale: float = 1.0):
    """small_area_hybrid (loss_functions.py:24-40) of each detector of
    `intensity` (M, rows, cols) float32.  Returns (loss (M,), dloss/dI or None)."""
    lib = N.load()
    if intensity.dtype != torch.float32 or intensity.device.type != "cuda":
        raise N.NativeUnavailable("intensity must be a float32 tensor on a ROCm GPU")
    d = intensity.contiguous()
    dev = d.device
    N.ensure_device(dev)
    M = d.shape[0]
    npix = d[0].numel()
    loss = torch.empty(M, dtype=torch.float32, device=dev)
    grad = torch.empty_like(d) if want_grad else None
    need = lib.xl_small_area_workspace_bytes(npix, M)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    p = N.SmallAreaParams()
    p.intensity, p.batch_stride = d.data_ptr(), npix
    p.grad = _ptr(grad)
    p.grad_batch_stride = npix
    p.loss, p.npix, p.batch, p.scale = loss.data_ptr(), npix, M, scale
    N.check(lib.xl_small_area(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
    return loss, grad


def _theta_ptr(theta: Optional[torch.Tensor], dev):
    if theta is None:
        return None
    if theta.dtype != torch.float64 or theta.device != dev or theta.numel() != 1:
        raise ValueError("theta must be a one-element float64 tensor on the field's device")
    return theta.data_ptr()


def beam_splitter(a: torch.Tensor, b: Optional[torch.Tensor], coef, want_c: bool = True,
                  want_d: bool = True, ncomp_out: int = 3, theta: Optional[torch.Tensor] = None,
                  theta_mode: int = 0, adjoint: bool = False):
    """c = ca a + cb b, d = da a + db b on components 0, 1 of `a`, `b`
    (k, rows, cols) complex64, k >= 2; coef = (ca, cb, da, db): complex
    numbers, or a complex tensor of 4 on the GPU (then read by the kernel, with
    no host synchronisation).  With theta_mode 1 (symmetric), 2 / 3 (single
    input at port 0 / 1) the kernel forms the coefficients from the float64
    device angle `theta` instead (coef is ignored), and `adjoint` applies
    their conjugate transpose.  Outputs are (ncomp_out, rows, cols) with
    component 2 (if any) 0."""
    lib = N.load()
    a = _gpu_c64(a, "a")
    dev = a.device
    N.ensure_device(dev)
    if b is not None:
        b = _gpu_c64(b.to(dev), "b")
        if b.shape[1:] != a.shape[1:] or b.shape[0] < 2:
            raise ValueError("beam-splitter inputs must have the same plane shape and >= 2 components")
    if a.shape[0] < 2:
        raise ValueError("beam-splitter inputs need Ex and Ey")
    npix = a[0].numel()
    shape = (ncomp_out,) + tuple(a.shape[1:])
    c = torch.empty(shape, dtype=torch.complex64, device=dev) if want_c else None
    d = torch.empty(shape, dtype=torch.complex64, device=dev) if want_d else None
    p = N.BsParams()
    p.a, p.a_comp_stride = a.data_ptr(), npix
    p.b, p.b_comp_stride = _ptr(b), npix
    p.c, p.d = _ptr(c), _ptr(d)
    p.c_comp_stride = p.d_comp_stride = npix
    keep = None
    if theta_mode:
        p.theta_mode, p.theta, p.adjoint = int(theta_mode), _theta_ptr(theta, dev), 1 if adjoint else 0
    elif isinstance(coef, torch.Tensor) and coef.device.type == "cuda":
        if coef.numel() != 4 or not coef.is_complex():
            raise ValueError("device coefficients must be 4 complex numbers (ca, cb, da, db)")
        keep = coef.to(device=dev, dtype=torch.complex64).contiguous()
        p.coef = keep.data_ptr()       # read by the kernel: no host round trip
    else:
        vals = coef.tolist() if isinstance(coef, torch.Tensor) else coef
        for name, k in zip(("ca", "cb", "da", "db"), vals):
            k = complex(k)
            getattr(p, name)[0], getattr(p, name)[1] = k.real, k.imag
    p.npix = npix
    p.zero_ez = 1 if ncomp_out > 2 else 0
    N.check(lib.xl_beam_splitter(C.byref(p), _stream(dev)))
    del keep
    return c, d


def dz_dot(d: Optional[torch.Tensor], gout: Optional[torch.Tensor], pre: Optional[torch.Tensor] = None,
           gez: Optional[torch.Tensor] = None, x=None, y=None, z: float = 0.0,
           z_dev: Optional[torch.Tensor] = None) -> torch.Tensor:
    """0-d float64 device tensor: sum_c Re(conj(d_c) gout_c) over the planes
    of d and gout (C, rows, cols), plus, with `pre` (Ex', Ey') and `gez` (the
    cotangent of the VRS Ez_in), sum Re(conj(gez) Ex') X w + Re(conj(gez) Ey') Y w,
    w = -z / (X^2 + Y^2 + z^2)^{3/2} (xl_dz_dot: one pass, fixed-order sums).
    d = gout = None leaves the Ez_in term alone."""
    lib = N.load()
    p = N.DzDotParams()
    if d is None:
        if gout is not None or pre is None:
            raise ValueError("without d, pass gout=None and pre / gez (the Ez_in term alone)")
        pre = _gpu_c64(pre, "pre")
        dev = pre.device
        N.ensure_device(dev)
        rows, cols = pre.shape[-2:]
        p.ncomp, p.rows, p.cols = 0, rows, cols
    else:
        d = _gpu_c64(d, "d")
        g = _gpu_c64(gout, "gout")
        dev = d.device
        N.ensure_device(dev)
        if d.shape != g.shape or d.dim() < 2:
            raise ValueError("d and gout must have the same (..., rows, cols) shape")
        rows, cols = d.shape[-2:]
        d3, g3 = d.reshape(-1, rows, cols), g.reshape(-1, rows, cols)
        if d3.shape[0] > 3:
            raise ValueError("at most 3 planes")
        p.d, p.g = d3.data_ptr(), g3.data_ptr()
        p.comp_stride, p.ncomp, p.rows, p.cols = rows * cols, d3.shape[0], rows, cols
    if pre is not None:
        pre = _gpu_c64(pre, "pre")
        gez = _gpu_c64(gez, "gez")
        if pre.shape[0] < 2 or tuple(pre.shape[-2:]) != (rows, cols) or tuple(gez.shape[-2:]) != (rows, cols):
            raise ValueError("pre must hold (Ex', Ey') and gez one plane of the field's shape")
        xd, yd = device_grid(x, dev), device_grid(y, dev)
        if xd.numel() != cols or yd.numel() != rows:
            raise ValueError("grid lengths do not match the field")
        p.pre, p.pre_comp_stride, p.gez = pre.data_ptr(), rows * cols, gez.data_ptr()
        p.x, p.y = xd.data_ptr(), yd.data_ptr()
    p.z = float(z)
    p.z_dev = _z_ptr(z_dev, dev)
    out = torch.empty((), dtype=torch.float64, device=dev)
    ws = torch.empty(lib.xl_dz_dot_workspace_bytes(), dtype=torch.uint8, device=dev)
    p.out = out.data_ptr()
    N.check(lib.xl_dz_dot(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
    return out


def bs_dots(gc: Optional[torch.Tensor], gd: Optional[torch.Tensor], a: torch.Tensor,
            b: Optional[torch.Tensor], theta: Optional[torch.Tensor] = None, theta_mode: int = 0,
            want_a: bool = False, want_b: bool = False, ncomp_out: int = 3):
    """(4,) complex64 device tensor: sums over Ex, Ey and all pixels of
    conj(gc) a, conj(gc) b, conj(gd) a, conj(gd) b (absent terms are 0).
    With theta_mode 1-3 (the splitters of beam_splitter), the 0-d float64
    device tensor dL/dtheta instead; with want_a / want_b as well, the tuple
    (dL/dtheta, ga, gb) with the input cotangents of beam_splitter(...,
    adjoint=True) formed in the same pass ((ncomp_out, rows, cols) each, or
    None where not wanted)."""
    lib = N.load()
    a = _gpu_c64(a, "a")
    dev = a.device
    npix = a[0].numel()
    gc = None if gc is None else _gpu_c64(gc, "gc")
    gd = None if gd is None else _gpu_c64(gd, "gd")
    b = None if b is None else _gpu_c64(b, "b")
    dots = None if theta_mode else torch.empty(8, dtype=torch.float32, device=dev)
    need = lib.xl_bs_dots_workspace_bytes(npix)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    p = N.BsDotsParams()
    p.gc, p.gd = _ptr(gc), _ptr(gd)
    p.gc_comp_stride = p.gd_comp_stride = npix
    p.a, p.b = a.data_ptr(), _ptr(b)
    p.a_comp_stride = p.b_comp_stride = npix
    p.dots, p.npix = _ptr(dots), npix
    if theta_mode:
        dtheta = torch.empty((), dtype=torch.float64, device=dev)
        p.theta, p.theta_mode, p.dtheta = _theta_ptr(theta, dev), int(theta_mode), dtheta.data_ptr()
        ga = gb = None
        if want_a or want_b:
            if not 2 <= ncomp_out <= 3:
                raise ValueError("ncomp_out must be 2 or 3")
            shape = (ncomp_out,) + tuple(a.shape[1:])
            ga = torch.empty(shape, dtype=torch.complex64, device=dev) if want_a else None
            gb = torch.empty(shape, dtype=torch.complex64, device=dev) if want_b else None
            p.ga, p.gb = _ptr(ga), _ptr(gb)
            p.ga_comp_stride = p.gb_comp_stride = npix
            p.g_planes = ncomp_out
        N.check(lib.xl_bs_dots(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
        return (dtheta, ga, gb) if (want_a or want_b) else dtheta
    N.check(lib.xl_bs_dots(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
    return torch.view_as_complex(dots.view(4, 2))


def slm_backward_adam(src: torch.Tensor, gout: torch.Tensor, x, y, batch: int, phase: torch.Tensor,
                      m: torch.Tensor, v: torch.Tensor, count: torch.Tensor, lr: float, b1: float = 0.9,
                      b2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0,
                      out: Optional[torch.Tensor] = None) -> Optional[torch.Tensor]:
    """Adjoint of a lone SLM on one component fused with the Adam / AdamW step
    of its phases (xl_slm_backward_adam).  src: (B, 1, rows, cols) or shared
    (1, rows, cols); gout: (B, 1, rows, cols); phase, m, v: (B, rows, cols)
    float32 views with one batch stride and contiguous rows (phase is updated
    in place).  `count` is read, not advanced (adam_tick).  Returns the field
    cotangent into `out` (B, 1, rows, cols) when given."""
    lib = N.load()
    dev = src.device
    N.ensure_device(dev)
    src, gout = src.contiguous(), gout.contiguous()
    rows, cols = src.shape[-2:]
    for t in (phase, m, v):
        if (tuple(t.shape) != (batch, rows, cols) or t.dtype != torch.float32 or t.device != dev
                or t.stride(2) != 1 or t.stride(1) != cols or t.stride(0) != phase.stride(0)):
            raise ValueError("phase, m and v must be float32 (B, rows, cols) views with one batch stride")
    if batch > 1 and phase.stride(0) < rows * cols:
        raise ValueError("the fused Adam step needs one phase array per batch member (members sharing phases "
                         "would race on them)")
    if count.dtype != torch.int32 or count.numel() != 1 or count.device != dev:
        raise ValueError("count must be a one-element int32 tensor on the field's device")
    p = N.PreParams()
    p.inp, p.in_comp_stride, p.nin, p.ncomp = src.data_ptr(), rows * cols, 1, 1
    p.comp_stride = rows * cols
    p.rows, p.cols = rows, cols
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    p.x, p.y = xd.data_ptr(), yd.data_ptr()
    p.gout = gout.data_ptr()
    p.mask[0].kind = N.XL_MASK_SLM
    p.mask[0].m0 = phase.data_ptr()
    p.mask[1].kind = N.XL_MASK_NONE
    p.batch = batch
    p.mask_batch_stride = phase.stride(0)
    p.in_batch_stride = 0 if src.dim() == 3 else rows * cols
    p.gout_batch_stride = rows * cols
    if out is not None:
        if tuple(out.shape) != (batch, 1, rows, cols) or not out.is_contiguous():
            raise ValueError("out must be a contiguous (B, 1, rows, cols) tensor")
        p.out = out.data_ptr()
        p.out_batch_stride = rows * cols
    a = N.AdamParams()
    a.param, a.grad, a.m, a.v = phase.data_ptr(), None, m.data_ptr(), v.data_ptr()
    a.n = rows * cols
    a.count = count.data_ptr()
    a.lr, a.b1, a.b2, a.eps, a.weight_decay = lr, b1, b2, eps, weight_decay
    N.check(lib.xl_slm_backward_adam(C.byref(p), C.byref(a), _stream(dev)))
    return out


def adam_tick(count: torch.Tensor) -> None:
    """count += 1 on the current stream (after the last fused update of a step)."""
    N.check(N.load().xl_adam_tick(count.data_ptr(), _stream(count.device)))


def focal_moment(E: torch.Tensor, r2: torch.Tensor):
    """L = sum I r2 / sum I, I = sum_c |E_c|^2 (xl_moment_forward): returns
    (L as a 0-d float32 tensor, stats (sum I, L)) with no host synchronisation."""
    lib = N.load()
    E = _gpu_c64(E, "field")
    dev = E.device
    N.ensure_device(dev)
    nc = E.shape[0]
    npix = E[0].numel()
    r2 = r2.to(device=dev, dtype=torch.float32).contiguous()
    if r2.numel() != npix:
        raise ValueError("r2 must hold one weight per pixel")
    stats = torch.empty(2, dtype=torch.float32, device=dev)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    ws = torch.empty(lib.xl_moment_workspace_bytes(npix), dtype=torch.uint8, device=dev)
    p = N.MomentParams()
    p.field, p.comp_stride, p.ncomp = E.data_ptr(), npix, nc
    p.r2, p.npix = r2.data_ptr(), npix
    p.stats, p.loss = stats.data_ptr(), loss.data_ptr()
    N.check(lib.xl_moment_forward(C.byref(p), ws.data_ptr(), ws.numel(), _stream(dev)))
    return loss.reshape(()), stats


def focal_moment_grad(E: torch.Tensor, r2: torch.Tensor, stats: torch.Tensor,
                      gscale: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Field cotangent of focal_moment scaled by the device scalar `gscale`."""
    lib = N.load()
    E = _gpu_c64(E, "field")
    dev = E.device
    nc = E.shape[0]
    npix = E[0].numel()
    r2 = r2.to(device=dev, dtype=torch.float32).contiguous()
    g = None if gscale is None else gscale.to(device=dev, dtype=torch.float32).reshape(1).contiguous()
    grad = torch.empty_like(E)
    p = N.MomentParams()
    p.field, p.comp_stride, p.ncomp = E.data_ptr(), npix, nc
    p.r2, p.npix = r2.data_ptr(), npix
    p.stats = stats.data_ptr()
    p.gscale = _ptr(g)
    p.grad = grad.data_ptr()
    N.check(lib.xl_moment_backward(C.byref(p), _stream(dev)))
    return grad
