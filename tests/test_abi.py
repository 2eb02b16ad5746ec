"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/svx.h declares, and the ctypes mirrors have the C layouts."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

from solvix_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert "svx_solve_model" in declared and "svx_csr_pcg" in declared
    missing = [s for s in declared if not hasattr(L, s)]
    assert missing == []


def test_every_declared_symbol_has_a_ctypes_signature():
    sigs = {s[0] for s in _lib.SIGNATURES}
    assert set(_lib.declared_symbols()) <= sigs


def test_version_string_needs_no_device():
    assert b"gfx950" in _lib.lib().svx_version()


def test_struct_layouts_match_the_c_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "svx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(svx_error), sizeof(svx_model_desc), sizeof(svx_solution),
         offsetof(svx_model_desc, thickness), offsetof(svx_model_desc, cload_value), offsetof(svx_solution, recover_ms),
         sizeof(svx_solve_stats), offsetof(svx_solve_stats, status));
  printf("%zu %zu %zu %zu %zu\n", sizeof(svx_server_config), offsetof(svx_server_config, chunk),
         sizeof(svx_request_info), sizeof(svx_server_summary), offsetof(svx_server_summary, p50_ms));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [C.sizeof(_abi.SvxError), C.sizeof(_abi.SvxModelDesc), C.sizeof(_abi.SvxSolution),
           _abi.SvxModelDesc.thickness.offset, _abi.SvxModelDesc.cload_value.offset,
           _abi.SvxSolution.recover_ms.offset, C.sizeof(_abi.SvxSolveStats), _abi.SvxSolveStats.status.offset,
           C.sizeof(_abi.SvxServerConfig), _abi.SvxServerConfig.chunk.offset, C.sizeof(_abi.SvxRequestInfo),
           C.sizeof(_abi.SvxServerSummary), _abi.SvxServerSummary.p50_ms.offset]
    assert [int(x) for x in out] == got


def test_header_compiles_as_c_and_cpp():
    for compiler, ext in (("gcc", "c"), ("g++", "cpp")):
        with tempfile.TemporaryDirectory() as d:
            f = os.path.join(d, "h." + ext)
            open(f, "w").write('#include "svx.h"\nint main(void){return SVX_OK;}\n')
            subprocess.run([compiler, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), f, "-o",
                            os.path.join(d, "h")], check=True)


def test_context_creation_fails_cleanly_without_a_device():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = _lib.lib().svx_ctx_create(0, C.byref(h))
    assert rc == _abi.SVX_E_DEVICE and not h.value
    assert len(_lib.lib().svx_last_error()) > 0
