"""Torch-facing calls into libxlumina_hip.so (no autograd here).

Fields are complex64 CUDA(ROCm) tensors shaped (C, rows, cols), row-major
[y][x]; mask parameters float32 (rows, cols); grids float64.  Outputs and the
workspace come from PyTorch's caching allocator, the launch goes on the
current stream.  There is no CPU path: non-GPU tensors raise.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as N


@dataclass
class MaskSpec:
    """One Jones-mask stage applied before a propagation (see XL_MASK_*)."""
    kind: int
    p0: Optional[torch.Tensor] = None
    p1: Optional[torch.Tensor] = None
    s0: float = 0.0
    s1: float = 0.0

    def tensors(self):
        return (self.p0, self.p1)


def _ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


_grid_cache: dict = {}


def device_grid(arr, device) -> torch.Tensor:
    """float64 device copy of a 1-D grid, cached by content."""
    if isinstance(arr, torch.Tensor):
        if arr.device == device and arr.dtype == torch.float64 and arr.is_contiguous():
            return arr
        arr = arr.detach().cpu().numpy()
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
    key = (a.tobytes(), str(device))
    t = _grid_cache.get(key)
    if t is None:
        if len(_grid_cache) > 256:
            _grid_cache.clear()
        t = torch.from_numpy(a.copy()).to(device)
        _grid_cache[key] = t
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


def fill_masks(dst, masks: Sequence[MaskSpec], shape, device, nin: int):
    """Fills the two xl_mask_stage slots; returns the tensors kept alive."""
    if len(masks) > 2:
        raise ValueError("at most two fused mask stages")
    keep = []
    for i in range(2):
        st = dst[i]
        if i >= len(masks) or masks[i] is None:
            st.kind = N.XL_MASK_NONE
            st.m0 = st.m1 = None
            st.s0 = st.s1 = 0.0
            continue
        m = masks[i]
        p0 = _prep_param(m.p0, shape, device)
        p1 = _prep_param(m.p1, shape, device)
        keep += [p0, p1]
        st.kind = m.kind
        st.m0 = _ptr(p0) or None
        st.m1 = _ptr(p1) or None
        st.s0 = float(m.s0)
        st.s1 = float(m.s1)
    return keep


def quality_factor(x, y, wavelength, z) -> float:
    """wave_optics.py:187-191 (host float64, as the reference)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    dr_real = math.sqrt(dx ** 2 + dy ** 2)
    rmax = math.sqrt(float((x ** 2).max()) + float((y ** 2).max()))
    dr_ideal = math.sqrt(wavelength ** 2 + rmax ** 2 + 2 * wavelength * math.sqrt(rmax ** 2 + z ** 2)) - rmax
    return dr_ideal / dr_real


# ---------------------------------------------------------------------------
# RS / VRS
# ---------------------------------------------------------------------------

def rs_workspace_bytes(n: int, ncomp: int) -> int:
    return int(N.load().xl_rs_workspace_bytes(n, ncomp))


def rs_propagate(src: torch.Tensor, x, y, z: float, wavelength: float, masks: Sequence[MaskSpec] = (),
                 ez_mode: int = N.XL_EZ_NONE, nin: Optional[int] = None, intensity: bool = False,
                 adjoint: bool = False, out: Optional[torch.Tensor] = None,
                 workspace: Optional[torch.Tensor] = None):
    """RS_propagation_jit / VRS_propagation_jit (wave_optics.py:281-289,
    vectorized_optics.py:364-373) with fused masks and Ez construction.

    src: (C, n, n) complex64; `nin` components are read (default C).
    Returns the (ncomp, n, n) field, plus |E|^2 if `intensity`.
    """
    lib = N.load()
    src = _prep_field(src)
    dev = src.device
    n = src.shape[-1]
    if src.shape[-2] != n:
        raise ValueError("RS propagation needs a square grid (the reference builds one lag vector for both axes)")
    nin = src.shape[0] if nin is None else nin
    ncomp = 3 if ez_mode == N.XL_EZ_VRS else nin
    xd, yd = device_grid(x, dev), device_grid(y, dev)
    if out is None:
        out = torch.empty((ncomp, n, n), dtype=torch.complex64, device=dev)
    inten = torch.empty((n, n), dtype=torch.float32, device=dev) if intensity else None
    need = lib.xl_rs_workspace_bytes(n, ncomp)
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
    keep = fill_masks(p.mask, masks, (n, n), dev, nin)
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
GEMM_CROSSOVER = 300


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
        form: Optional[str] = None) -> torch.Tensor:
    """CZT_jit / VCZT_jit / CZT_for_high_NA_jit / VCZT_objective_lens core.

    Forward: src (C, ny, nx) -> (ncomp, my, mx).  Adjoint: src (ncomp, my, mx)
    cotangent -> (ncomp, ny, nx) cotangent of the pre-op output.  `form`:
    'bluestein', 'gemm' or 'auto' (default: the module setting).
    """
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
    if tuple(src.shape[-2:]) != (rows_in, cols_in):
        raise ValueError(f"CZT input is {tuple(src.shape[-2:])}, expected {(rows_in, cols_in)}")
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
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    N.check(runfn(C.byref(p), workspace.data_ptr(), workspace.numel(), _stream(dev)))
    del keep
    return out


# ---------------------------------------------------------------------------
# standalone pre-ops
# ---------------------------------------------------------------------------

def _pre_params(src, x, y, z, masks, ez_mode, hna, ncomp, nin, hna_radius, hna_f):
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
                 hna_radius: float = 0.0, hna_f: float = 0.0):
    """Adjoint of pre_forward: returns (grad of src[:nin], [[g00, g01], [g10, g11]])."""
    lib = N.load()
    src = _prep_field(src)
    gout = _prep_field(gout)
    nin = src.shape[0] if nin is None else nin
    ncomp = gout.shape[0]
    p, keep = _pre_params(src, x, y, z, masks, ez_mode, hna, ncomp, nin, hna_radius, hna_f)
    p.gout = gout.data_ptr()
    rows, cols = src.shape[-2:]
    gsrc = torch.empty((nin, rows, cols), dtype=torch.complex64, device=src.device) if need_src else None
    p.out = _ptr(gsrc) or None
    gp = [[None, None], [None, None]]
    for i in range(2):
        for j in range(2):
            if need_params[i][j]:
                gp[i][j] = torch.empty((rows, cols), dtype=torch.float32, device=src.device)
                p.gparam[i][j] = gp[i][j].data_ptr()
            else:
                p.gparam[i][j] = None
    N.check(lib.xl_pre_backward(C.byref(p), _stream(src.device)))
    del keep
    return gsrc, gp


def transfer_function_rs(x, y, z: float, wavelength: float, device) -> torch.Tensor:
    """transfer_function_RS (wave_optics.py:291-297) on meshgrid(x, y)."""
    lib = N.load()
    N.ensure_device(device)
    xd, yd = device_grid(x, device), device_grid(y, device)
    out = torch.empty((len(yd), len(xd)), dtype=torch.complex64, device=device)
    N.check(lib.xl_transfer_function_rs(out.data_ptr(), xd.data_ptr(), yd.data_ptr(), len(xd), len(yd), float(z),
                                        2 * math.pi / wavelength, _stream(device)))
    return out
