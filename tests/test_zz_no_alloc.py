"""svx_csr_pcg_noalloc (FP64 PCG with its state in the library's fixed
per-GPU scratch, no device allocation) against svx_csr_pcg on the same
matrices: two calls in a row, the second on a smaller system, each equal to
the allocating entry point's result (same kernels and launch shape, so the
iterates agree bit for bit; the bar is the 1e-8 displacement tolerance)."""
import os

import numpy as np
import pytest

import oracle
from solvix_amd import decks
from solvix_amd.core import CsrMatrix
from solvix_amd.inp import read_inp
from solvix_amd.model import Precision, SolveError

from helpers import dirichlet, rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _system(model):
    r = oracle.solve(model, stop_after_assembly=True)
    assert r.status == 0
    rows = _constrained_rows(model)
    v, f = dirichlet(r.row_ptr, r.col_idx, r.vals, r.f, rows)
    return r.row_ptr.astype(np.int32), r.col_idx.astype(np.int32), v, f


def _constrained_rows(model):
    from solvix_amd.core import constrained_dofs
    return constrained_dofs(model)[0]


def test_noalloc_pcg_equals_svx_csr_pcg_twice_second_smaller(ctx):
    big = decks.cfg1()  # 32,882 rows
    small = read_inp(os.path.join(GOLDEN, "cantilever.inp"))  # 410 rows
    for model in (big, small):
        rp, ci, v, f = _system(model)
        A = CsrMatrix(rp, ci, v, ctx)
        u0, its0, res0 = A.pcg(f, 1e-10, 0, Precision.FP64)
        u1, its1, res1 = A.pcg_noalloc(f, 1e-10, 0)
        same = np.array_equal(u0.view(np.uint64), u1.view(np.uint64))
        print(f"{model.name}: n {A.n}, iterations {its1} (allocating {its0}), residual {res1:.3e} ({res0:.3e}), "
              f"rel diff {rel_err(u1, u0):.1e}, bitwise {same}", flush=True)
        assert res1 <= 1e-10 and its1 == its0
        assert rel_err(u1, u0) <= 1e-8
        A.close()


def test_noalloc_pcg_refuses_systems_beyond_its_scratch(ctx):
    n = (1 << 19) + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    A = CsrMatrix(rp, ci, np.ones(n), ctx)
    with pytest.raises(SolveError) as e:
        A.pcg_noalloc(np.ones(n))
    assert e.value.kind == "Capacity"
    A.close()
