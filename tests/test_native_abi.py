"""The C ABI library: loads without a GPU, exports every declared symbol, and
the ctypes structures match the C layout (host-only calls, no kernels)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from tests.conftest import ROOT
from xlumina_amd import _native as N
from xlumina_amd import build_lib

HEADER = os.path.join(ROOT, "include", "xlumina_hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xl_\w+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(build_lib.lib_path()):
        pytest.fail("libxlumina_hip.so is not built (run __graft_entry__.build())")
    return N.load()


def test_every_header_symbol_is_exported_and_bound(lib):
    names = header_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"
    assert set(N.SIGNATURES) == set(names)


def test_struct_layout_matches_c():
    code = r'''
#include <stdio.h>
#include <stddef.h>
#include "xlumina_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(xl_mask_stage), sizeof(xl_rs_params), sizeof(xl_czt_params), sizeof(xl_pre_params));
  printf("%zu %zu %zu\n", offsetof(xl_rs_params, mask), offsetof(xl_czt_params, mask), offsetof(xl_pre_params, mask));
  printf("%zu %zu %zu\n", offsetof(xl_rs_params, dx), offsetof(xl_czt_params, hna_f), offsetof(xl_pre_params, gparam));
  printf("%zu %zu %zu %zu %zu\n", offsetof(xl_rs_params, z_dev), offsetof(xl_czt_params, mask_batch_stride),
         offsetof(xl_pre_params, z_dev), offsetof(xl_pre_params, gparam_batch_stride), sizeof(xl_mse_params));
  printf("%zu %zu %zu %zu\n", offsetof(xl_mse_params, scale), sizeof(xl_adam_params), offsetof(xl_adam_params, count),
         offsetof(xl_adam_params, weight_decay));
  printf("%zu\n", offsetof(xl_rs_params, kernel_dz));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write(code)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), src, "-o", exe], check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = [int(v) for v in lines[0].split()]
    offs = [int(v) for v in lines[1].split()]
    offs2 = [int(v) for v in lines[2].split()]
    assert sizes == [C.sizeof(N.MaskStage), C.sizeof(N.RsParams), C.sizeof(N.CztParams), C.sizeof(N.PreParams)]
    assert offs == [N.RsParams.mask.offset, N.CztParams.mask.offset, N.PreParams.mask.offset]
    assert offs2 == [N.RsParams.dx.offset, N.CztParams.hna_f.offset, N.PreParams.gparam.offset]
    assert [int(v) for v in lines[3].split()] == [N.RsParams.z_dev.offset, N.CztParams.mask_batch_stride.offset,
                                                  N.PreParams.z_dev.offset, N.PreParams.gparam_batch_stride.offset,
                                                  C.sizeof(N.MseParams)]
    assert [int(v) for v in lines[4].split()] == [N.MseParams.scale.offset, C.sizeof(N.AdamParams),
                                                  N.AdamParams.count.offset, N.AdamParams.weight_decay.offset]
    assert int(lines[5]) == N.RsParams.kernel_dz.offset


STRUCTS = {"xl_mask_stage": "MaskStage", "xl_rs_params": "RsParams", "xl_czt_params": "CztParams",
           "xl_pre_params": "PreParams", "xl_mse_params": "MseParams", "xl_adam_params": "AdamParams",
           "xl_field_loss_params": "FieldLossParams", "xl_small_area_params": "SmallAreaParams",
           "xl_bs_params": "BsParams", "xl_bs_dots_params": "BsDotsParams", "xl_kernel_time": "KernelTime",
           "xl_moment_params": "MomentParams", "xl_dz_dot_params": "DzDotParams"}


def test_every_struct_field_matches_c():
    """sizeof and the offset of every field, for every parameter struct of the header."""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    assert set(re.findall(r"typedef struct (xl_\w+)", src)) == set(STRUCTS)
    lines, expect = [], []
    for cname, pname in STRUCTS.items():
        cls = getattr(N, pname)
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for fname, *_ in cls._fields_:
            cfield = {"inp": "in"}.get(fname, fname)   # `in` is a Python keyword
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {cfield}));')
            expect.append(getattr(cls, fname).offset)
    code = ("#include <stdio.h>\n#include <stddef.h>\n#include \"xlumina_hip.h\"\nint main(void) {\n"
            + "\n".join(lines) + "\n  return 0;\n}\n")
    with tempfile.TemporaryDirectory() as d:
        c_src = os.path.join(d, "fields.c")
        exe = os.path.join(d, "fields")
        open(c_src, "w").write(code)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), c_src, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect


def test_host_side_queries(lib):
    assert lib.xl_abi_version() == 10
    assert lib.xl_max_fft_length() == 4096
    # RS workspace: ncomp * round4(n) * P + (P/2) * P + P complex64 (+ alignment slack)
    n, P = 1024, 2048
    assert lib.xl_rs_workspace_bytes(n, 1) == (n * P + (P // 2) * P + P) * 8 + 256
    assert lib.xl_rs_workspace_bytes(1, 1) == 0
    assert lib.xl_rs_workspace_bytes(4096, 1) > 0   # rejected later by the propagate call (P > 4096)
    # spectrum buffer of the z / lambda gradients: ncomp P^2 input spectra, two
    # kernel spectra, two float64 partials per (column, component) unit
    hh = ((P // 2) * P + P) * 8
    for nc in (1, 3):
        assert lib.xl_rs_spec_bytes(n, nc) == nc * P * P * 8 + 2 * hh + 2 * nc * P * 8 + 256
    # P = 64 (every region already a multiple of 256 bytes): 192 units in 12 blocks of 16
    assert lib.xl_rs_spec_bytes(20, 3) == 3 * 64 * 64 * 8 + 2 * (32 * 64 + 64) * 8 + 2 * 192 * 8 + 256
    assert lib.xl_rs_spec_bytes(20, 1) == 64 * 64 * 8 + 2 * (32 * 64 + 64) * 8 + 2 * 64 * 8 + 256
    assert lib.xl_rs_spec_bytes(1, 1) == 0 and lib.xl_rs_spec_bytes(64, 4) == 0 and lib.xl_rs_spec_bytes(4096, 1) == 0
    # the partial sums are the last two regions: units float64 each, ending at the buffer's end
    cnt = C.c_int32(0)
    for nn, nc in ((1024, 3), (20, 3), (20, 1)):
        total = lib.xl_rs_spec_bytes(nn, nc) - 256
        o0 = lib.xl_rs_spec_partials(nn, nc, 0, C.byref(cnt))
        units = cnt.value
        o1 = lib.xl_rs_spec_partials(nn, nc, 1, C.byref(cnt))
        assert cnt.value == units and o1 == o0 + (units * 8 + 255) // 256 * 256
        assert total == o1 + (units * 8 + 255) // 256 * 256
    assert units == 64 and lib.xl_rs_spec_partials(1024, 3, 2, C.byref(cnt)) == -1
    assert lib.xl_rs_spec_partials(1, 1, 0, C.byref(cnt)) == -1


def _czt_params(m, M):
    p = N.CztParams()
    p.nin = p.ncomp = 1
    p.ny_in = p.nx_in = m
    p.my = p.mx = M
    p.dm = 100.0
    p.fy1 = p.fx1 = 40.0
    p.fy2 = p.fx2 = 60.0
    return p


def test_czt_plan_rejects_reference_failures(lib):
    assert lib.xl_czt_workspace_bytes(C.byref(_czt_params(2048, 512))) > 0
    p = _czt_params(2048, 2049)           # m + M - 1 = 4096: the reference raises
    assert lib.xl_czt_workspace_bytes(C.byref(p)) == 0
    assert b"power of two" in lib.xl_last_error()
    assert lib.xl_czt_workspace_bytes(C.byref(_czt_params(64, 1))) == 0   # M_out = 1
    assert lib.xl_czt_workspace_bytes(C.byref(_czt_params(4000, 200))) == 0   # FFT length > 4096


def test_entry_points_refuse_without_init(lib):
    # no kernel is launched: the call stops at the initialisation check
    p = N.RsParams()
    p.n = 64
    rc = lib.xl_rs_propagate(C.byref(p), None, 0, None)
    assert rc != 0


FAKE = 1 << 20     # a non-null address the host checks never dereference (no kernel is launched)


def test_gemm_entry_rejects_tables_mode(lib):
    """xl_czt_gemm has no staged-table mode: tables_mode != 0 is an error
    (checked on the host, before the initialisation check)."""
    p = _czt_params(256, 64)
    p.inp = p.out = p.x = p.y = p.xout = p.yout = FAKE
    p.tables_mode = 1
    assert lib.xl_czt_gemm(C.byref(p), None, 0, None) == N.XL_EINVAL
    assert b"tables_mode" in lib.xl_last_error()


def _slm_pre(batch, mask_stride, gparam_stride=0):
    p = N.PreParams()
    p.inp = p.x = p.y = p.gout = FAKE
    p.nin = p.ncomp = 1
    p.rows = p.cols = 8
    p.comp_stride = p.in_comp_stride = 64
    p.mask[0].kind = N.XL_MASK_SLM
    p.mask[0].m0 = FAKE
    p.batch = batch
    p.mask_batch_stride = mask_stride
    p.gparam_batch_stride = gparam_stride
    return p


def test_batched_adjoints_reject_shared_outputs(lib):
    """Members of a batch must not share the arrays a kernel writes: a fused
    Adam step on one shared phase array, or one shared parameter-gradient
    array, would race (host checks; nothing is launched)."""
    a = N.AdamParams()
    a.param = a.m = a.v = FAKE
    a.count = FAKE
    a.lr, a.b1, a.b2, a.eps = 0.05, 0.9, 0.999, 1e-8
    assert lib.xl_slm_backward_adam(C.byref(_slm_pre(4, 0)), C.byref(a), None) == N.XL_EINVAL
    assert b"race" in lib.xl_last_error()
    p = _slm_pre(4, 64)
    p.gparam[0][0] = FAKE
    assert lib.xl_pre_backward(C.byref(p), None) == N.XL_EINVAL
    assert b"race" in lib.xl_last_error()
