"""ctypes binding of libxlumina_hip.so (include/xlumina_hip.h).

The product path has no fallback: if the library is missing or the device is
not a ROCm GPU, calls raise ``NativeUnavailable``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

from . import build_lib


class NativeUnavailable(RuntimeError):
    pass


class XlError(RuntimeError):
    pass


XL_OK, XL_EINVAL, XL_EHIP, XL_EWORKSPACE = range(4)
XL_MASK_NONE, XL_MASK_SLM, XL_MASK_SSLM, XL_MASK_LCD, XL_MASK_LCD_SCALAR, XL_MASK_LP, XL_MASK_AMP = range(7)
XL_EZ_NONE, XL_EZ_VRS, XL_EZ_VCZT, XL_EZ_HNA = range(4)
XL_PF_F64, XL_PF_SCALAR, XL_PF_UNIT_ANGLE = 1, 2, 4
XL_RS_SPEC_PREBUILT = 4   # spec_grads bit of an adjoint call: the forward built the derivative spectra
XL_RS_SPEC_PARTIALS = 8   # spec_grads bit of an adjoint call: partial sums stay in spec (xl_rs_spec_partials)
XL_CZT_RS, XL_CZT_HNA, XL_CZT_PLAIN = range(3)
XL_LOSS_INTENSITY, XL_LOSS_AMPLITUDE, XL_LOSS_PHASE = range(3)
XL_RS_NOALLOC_MAX_N = 1024
XL_RS_NOALLOC_SCRATCH_BYTES = 67125504


class MaskStage(C.Structure):
    _fields_ = [("kind", C.c_int32), ("forms", C.c_int32), ("m0", C.c_void_p), ("m1", C.c_void_p),
                ("s0", C.c_double), ("s1", C.c_double)]


class RsParams(C.Structure):
    _fields_ = [("inp", C.c_void_p), ("in_comp_stride", C.c_int64), ("nin", C.c_int32), ("ez_mode", C.c_int32),
                ("out", C.c_void_p), ("out_comp_stride", C.c_int64), ("intensity", C.c_void_p),
                ("n", C.c_int32), ("adjoint", C.c_int32), ("x", C.c_void_p), ("y", C.c_void_p),
                ("z", C.c_double), ("k", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("mask", MaskStage * 2), ("z_dev", C.c_void_p), ("kernel_dz", C.c_int32), ("hhat_mode", C.c_int32),
                ("hhat", C.c_void_p), ("batch", C.c_int32), ("reserved2", C.c_int32), ("in_batch_stride", C.c_int64),
                ("out_batch_stride", C.c_int64), ("intensity_batch_stride", C.c_int64),
                ("mask_batch_stride", C.c_int64), ("spec_mode", C.c_int32), ("spec_grads", C.c_int32),
                ("spec", C.c_void_p), ("grad_out", C.c_void_p), ("quality_factor", C.c_void_p),
                ("qf_rmax", C.c_double), ("qf_dr_real", C.c_double), ("qf_wavelength", C.c_double),
                ("spec_dg", C.c_void_p)]


class CztParams(C.Structure):
    _fields_ = [("inp", C.c_void_p), ("in_comp_stride", C.c_int64), ("nin", C.c_int32), ("ncomp", C.c_int32),
                ("out", C.c_void_p), ("out_comp_stride", C.c_int64),
                ("ny_in", C.c_int32), ("nx_in", C.c_int32), ("my", C.c_int32), ("mx", C.c_int32),
                ("mode", C.c_int32), ("adjoint", C.c_int32), ("ez_mode", C.c_int32), ("reserved", C.c_int32),
                ("x", C.c_void_p), ("y", C.c_void_p), ("xout", C.c_void_p), ("yout", C.c_void_p),
                ("z", C.c_double), ("wavelength", C.c_double), ("k", C.c_double), ("dm", C.c_double),
                ("fy1", C.c_double), ("fy2", C.c_double), ("fx1", C.c_double), ("fx2", C.c_double),
                ("cte_re", C.c_double), ("cte_im", C.c_double), ("hna_radius", C.c_double), ("hna_f", C.c_double),
                ("mask", MaskStage * 2), ("batch", C.c_int32), ("tables_mode", C.c_int32),
                ("in_batch_stride", C.c_int64), ("out_batch_stride", C.c_int64), ("mask_batch_stride", C.c_int64)]


class PreParams(C.Structure):
    _fields_ = [("inp", C.c_void_p), ("in_comp_stride", C.c_int64), ("nin", C.c_int32), ("ncomp", C.c_int32),
                ("gout", C.c_void_p), ("out", C.c_void_p), ("comp_stride", C.c_int64),
                ("gparam", (C.c_void_p * 2) * 2), ("rows", C.c_int32), ("cols", C.c_int32),
                ("ez_mode", C.c_int32), ("hna", C.c_int32), ("x", C.c_void_p), ("y", C.c_void_p),
                ("z", C.c_double), ("hna_radius", C.c_double), ("hna_f", C.c_double), ("mask", MaskStage * 2),
                ("z_dev", C.c_void_p), ("batch", C.c_int32), ("out_planes", C.c_int32),
                ("in_batch_stride", C.c_int64), ("gout_batch_stride", C.c_int64), ("out_batch_stride", C.c_int64),
                ("mask_batch_stride", C.c_int64), ("gparam_batch_stride", C.c_int64), ("ez_dz", C.c_void_p),
                ("ez_dz_base", C.c_void_p), ("ez_dz_base_n", C.c_int32), ("reserved3", C.c_int32)]


class MseParams(C.Structure):
    _fields_ = [("field", C.c_void_p), ("field_batch_stride", C.c_int64), ("target", C.c_void_p),
                ("grad", C.c_void_p), ("grad_batch_stride", C.c_int64), ("loss", C.c_void_p),
                ("npix", C.c_int64), ("batch", C.c_int32), ("scale", C.c_float)]


class AdamParams(C.Structure):
    _fields_ = [("param", C.c_void_p), ("grad", C.c_void_p), ("m", C.c_void_p), ("v", C.c_void_p),
                ("n", C.c_int64), ("count", C.c_void_p), ("lr", C.c_float), ("b1", C.c_float), ("b2", C.c_float),
                ("eps", C.c_float), ("weight_decay", C.c_float), ("reserved", C.c_int32)]


class FieldLossParams(C.Structure):
    _fields_ = [("field", C.c_void_p), ("field_batch_stride", C.c_int64), ("target", C.c_void_p),
                ("target_batch_stride", C.c_int64), ("grad", C.c_void_p), ("grad_batch_stride", C.c_int64),
                ("loss", C.c_void_p), ("npix", C.c_int64), ("batch", C.c_int32), ("mode", C.c_int32),
                ("scale", C.c_float), ("reserved", C.c_int32)]


class SmallAreaParams(C.Structure):
    _fields_ = [("intensity", C.c_void_p), ("batch_stride", C.c_int64), ("grad", C.c_void_p),
                ("grad_batch_stride", C.c_int64), ("loss", C.c_void_p), ("npix", C.c_int64), ("batch", C.c_int32),
                ("scale", C.c_float)]


class BsParams(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("a_comp_stride", C.c_int64), ("b_comp_stride", C.c_int64),
                ("c", C.c_void_p), ("d", C.c_void_p), ("c_comp_stride", C.c_int64), ("d_comp_stride", C.c_int64),
                ("ca", C.c_float * 2), ("cb", C.c_float * 2), ("da", C.c_float * 2), ("db", C.c_float * 2),
                ("npix", C.c_int64), ("zero_ez", C.c_int32), ("theta_mode", C.c_int32), ("coef", C.c_void_p),
                ("theta", C.c_void_p), ("adjoint", C.c_int32), ("reserved", C.c_int32)]


class BsDotsParams(C.Structure):
    _fields_ = [("gc", C.c_void_p), ("gd", C.c_void_p), ("gc_comp_stride", C.c_int64),
                ("gd_comp_stride", C.c_int64), ("a", C.c_void_p), ("b", C.c_void_p), ("a_comp_stride", C.c_int64),
                ("b_comp_stride", C.c_int64), ("dots", C.c_void_p), ("npix", C.c_int64), ("theta", C.c_void_p),
                ("theta_mode", C.c_int32), ("reserved", C.c_int32), ("dtheta", C.c_void_p),
                ("ga", C.c_void_p), ("gb", C.c_void_p), ("ga_comp_stride", C.c_int64), ("gb_comp_stride", C.c_int64),
                ("g_planes", C.c_int32), ("reserved2", C.c_int32)]


class MomentParams(C.Structure):
    _fields_ = [("field", C.c_void_p), ("comp_stride", C.c_int64), ("ncomp", C.c_int32), ("reserved", C.c_int32),
                ("r2", C.c_void_p), ("npix", C.c_int64), ("stats", C.c_void_p), ("loss", C.c_void_p),
                ("gscale", C.c_void_p), ("grad", C.c_void_p)]


class DzDotParams(C.Structure):
    _fields_ = [("d", C.c_void_p), ("g", C.c_void_p), ("comp_stride", C.c_int64), ("ncomp", C.c_int32),
                ("rows", C.c_int32), ("cols", C.c_int32), ("reserved", C.c_int32), ("pre", C.c_void_p),
                ("pre_comp_stride", C.c_int64), ("gez", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p),
                ("z", C.c_double), ("z_dev", C.c_void_p), ("out", C.c_void_p)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("total_ms", C.c_double)]


# symbol -> (restype, argtypes); every declaration of include/xlumina_hip.h
SIGNATURES = {
    "xl_last_error": (C.c_char_p, []),
    "xl_abi_version": (C.c_int, []),
    "xl_max_fft_length": (C.c_int, []),
    "xl_init": (C.c_int, []),
    "xl_rs_workspace_bytes": (C.c_size_t, [C.c_int32, C.c_int32]),
    "xl_rs_batch_workspace_bytes": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32]),
    "xl_rs_hhat_bytes": (C.c_size_t, [C.c_int32]),
    "xl_rs_spec_bytes": (C.c_size_t, [C.c_int32, C.c_int32]),
    "xl_rs_spec_partials": (C.c_int64, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "xl_pre_scalar_grad_elems": (C.c_size_t, [C.c_int32, C.c_int32]),
    "xl_rs_propagate": (C.c_int, [C.POINTER(RsParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_rs_propagate_noalloc": (C.c_int, [C.POINTER(RsParams), C.c_void_p]),
    "xl_czt_workspace_bytes": (C.c_size_t, [C.POINTER(CztParams)]),
    "xl_czt": (C.c_int, [C.POINTER(CztParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_czt_gemm_workspace_bytes": (C.c_size_t, [C.POINTER(CztParams)]),
    "xl_czt_gemm": (C.c_int, [C.POINTER(CztParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_pre_forward": (C.c_int, [C.POINTER(PreParams), C.c_void_p]),
    "xl_pre_backward": (C.c_int, [C.POINTER(PreParams), C.c_void_p]),
    "xl_transfer_function_rs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                          C.c_double, C.c_double, C.c_void_p]),
    "xl_mse_workspace_bytes": (C.c_size_t, [C.c_int64, C.c_int32]),
    "xl_mse_intensity": (C.c_int, [C.POINTER(MseParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_adam_step": (C.c_int, [C.POINTER(AdamParams), C.c_void_p]),
    "xl_slm_backward_adam": (C.c_int, [C.POINTER(PreParams), C.POINTER(AdamParams), C.c_void_p]),
    "xl_adam_tick": (C.c_int, [C.c_void_p, C.c_void_p]),
    "xl_field_loss_workspace_bytes": (C.c_size_t, [C.c_int64, C.c_int32]),
    "xl_field_loss": (C.c_int, [C.POINTER(FieldLossParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_small_area_workspace_bytes": (C.c_size_t, [C.c_int64, C.c_int32]),
    "xl_small_area": (C.c_int, [C.POINTER(SmallAreaParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_moment_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "xl_moment_forward": (C.c_int, [C.POINTER(MomentParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_moment_backward": (C.c_int, [C.POINTER(MomentParams), C.c_void_p]),
    "xl_quality_factor": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "xl_dz_dot_workspace_bytes": (C.c_size_t, []),
    "xl_dz_dot": (C.c_int, [C.POINTER(DzDotParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_beam_splitter": (C.c_int, [C.POINTER(BsParams), C.c_void_p]),
    "xl_bs_dots_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "xl_bs_dots": (C.c_int, [C.POINTER(BsDotsParams), C.c_void_p, C.c_size_t, C.c_void_p]),
    "xl_uniform_phases": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int64,
                                    C.c_void_p]),
    "xl_timing_enable": (C.c_int, [C.c_int32, C.c_int32]),
    "xl_timing_collect": (C.c_int, [C.POINTER(KernelTime), C.c_int32]),
}

_lib = None
_lock = threading.Lock()
_inited_devices: set = set()


def load(path: str | None = None):
    """Loads the shared library (no GPU needed)."""
    global _lib
    lib = _lib
    if lib is not None and path is None:   # loaded: no lock on the per-call path
        return lib
    with _lock:
        if _lib is not None:
            return _lib
        path = path or os.environ.get("XL_LIBRARY") or build_lib.lib_path()
        if not os.path.exists(path):
            raise NativeUnavailable(
                f"{path} is missing: build it with `python -m xlumina_amd.build_lib` (hipcc, gfx950)")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = _lib.xl_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(msg)
        raise XlError(f"libxlumina_hip status {rc}: {msg}")


def ensure_device(device) -> None:
    """xl_init on `device` (float64 twiddle table); torch selects the device."""
    import torch
    if device.type != "cuda":
        raise NativeUnavailable(f"the propagation kernels run on a ROCm GPU, got device {device}")
    if not torch.version.hip:
        raise NativeUnavailable("PyTorch is not a ROCm build")
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx in _inited_devices:
        return
    lib = load()
    with torch.cuda.device(idx):
        check(lib.xl_init())
    _inited_devices.add(idx)


def timing_enable(capacity: int = 8192) -> None:
    """Record HIP events around every kernel launch from now on."""
    check(load().xl_timing_enable(1, capacity))


def timing_disable() -> None:
    check(load().xl_timing_enable(0, 0))


def timing_collect() -> dict:
    """{kernel name: (launches, total ms)} of the launches since the last collect."""
    lib = load()
    buf = (KernelTime * 32)()
    n = lib.xl_timing_collect(buf, 32)
    if n < 0:
        check(-n)
    return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(min(n, 32))}
