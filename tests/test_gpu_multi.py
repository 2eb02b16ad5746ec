"""Multi-RHS block solve: each load case matches its own single-RHS solve
and the oracle within the north-star tolerance."""
import numpy as np
import pytest

import oracle
from solvix_amd import decks
from solvix_amd.core import assemble, constrained_dofs
from solvix_amd.model import Precision

from helpers import rel_err

pytestmark = pytest.mark.gpu


def test_multi_rhs_matches_single_solves(ctx):
    model, cases = decks.cfg4(nrhs=12, nx=24, ny=4, nz=4)
    mesh, sys_ = assemble(model, ctx)
    rows, _, _ = constrained_dofs(model)
    sys_.apply_dirichlet(rows)
    n = mesh.n
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    F = np.zeros((len(cases), n))
    for k, (nodes, dofs, vals) in enumerate(cases):
        for nid, d, v in zip(nodes, dofs, vals):
            F[k, pos[int(nid)] * 3 + int(d) - 1] += v
    U, its, res, ok, ms = sys_.solve_multi(F)
    assert ok.all() and (res <= 1e-10).all()
    for k in range(len(cases)):
        m = decks.cfg4(nrhs=12, nx=24, ny=4, nz=4)[0]
        m.cload_node, m.cload_dof, m.cload_value = cases[k]
        ref = oracle.solve(m)
        e = rel_err(U[k], ref.u)
        print(f"case {k}: its {its[k]} (oracle {ref.iterations}) res {res[k]:.2e} err {e:.1e}")
        assert e <= 1e-8


def test_multi_rhs_fp32_ir_matches_oracle_and_single_solves(ctx):
    """SVX_FP32_IR: every load case runs the FP32-IR solve of a single
    svx_solve (FP32 inner iterations, FP64 refinement and final test) over
    the shared matrix; each result equals that case's own FP32-IR solve bit
    for bit and the oracle within 1e-8."""
    model, cases = decks.cfg4(nrhs=12, nx=24, ny=4, nz=4)
    mesh, sys_ = assemble(model, ctx)
    rows, _, _ = constrained_dofs(model)
    sys_.apply_dirichlet(rows)
    n = mesh.n
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    F = np.zeros((len(cases), n))
    for k, (nodes, dofs, vals) in enumerate(cases):
        for nid, d, v in zip(nodes, dofs, vals):
            F[k, pos[int(nid)] * 3 + int(d) - 1] += v
    U, its, res, ok, ms = sys_.solve_multi(F, precision=Precision.FP32_IR)
    assert ok.all() and (res <= 1e-10).all()
    for k in range(len(cases)):
        sys_.set_rhs(F[k])
        u1, its1, _ = sys_.solve(1e-10, 0, Precision.FP32_IR)
        m = decks.cfg4(nrhs=12, nx=24, ny=4, nz=4)[0]
        m.cload_node, m.cload_dof, m.cload_value = cases[k]
        ref = oracle.solve(m)
        e = rel_err(U[k], ref.u)
        print(f"case {k}: its {its[k]} (single {its1}, oracle {ref.iterations}) res {res[k]:.2e} err {e:.1e}")
        assert e <= 1e-8 and its[k] == its1
        assert np.array_equal(U[k].view(np.uint64), u1.view(np.uint64))



def test_multi_rhs_tile_sizes():
    """Other tile sizes of the block solve regroup the per-tile partial sums
    (rows per tile rounded to whole nodes), so they agree with the default to
    the stopping tolerance and with the same iteration counts within 2%."""
    from solvix_amd.core import Context
    model, cases = decks.cfg4(nrhs=8, nx=24, ny=4, nz=4)
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    out = {}
    for rpt in (0, 12, 50, 96, 192):
        ctx = Context(0)
        try:
            ctx.set_tuning(rows_per_tile=rpt)
            mesh, sys_ = assemble(model, ctx)
            sys_.apply_dirichlet(constrained_dofs(model)[0])
            F = np.zeros((len(cases), mesh.n))
            for k, (nodes, dofs, vals) in enumerate(cases):
                for nid, d, v in zip(nodes, dofs, vals):
                    F[k, pos[int(nid)] * 3 + int(d) - 1] += v
            U, its, res, ok, _ = sys_.solve_multi(F)
            assert ok.all()
            out[rpt] = (U, its)
        finally:
            ctx.close()
    U0, its0 = out[0]
    for rpt, (U, its) in out.items():
        assert (np.abs(its - its0) <= np.maximum(2, 0.02 * its0)).all() and rel_err(U, U0) <= 1e-8, rpt
