"""Parity of the HIP path (through the C ABI) with the oracle.

Bars (north star): CSR pattern and DOF numbering bit-exact; assembled values
within 1e-12 relative; displacements and von Mises within 1e-8 relative at
rel. residual 1e-10; the measured error is printed per result.
"""
import os

import numpy as np
import pytest

import oracle
from solvix_amd import _abi, decks
from solvix_amd.core import CsrMatrix, Mesh, assemble, constrained_dofs, solve_with_method
from solvix_amd.inp import read_inp
from solvix_amd.model import Method, Precision, SolveError

from helpers import dirichlet, eval_floor, rel_err, small_models

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _models():
    ms = [read_inp(os.path.join(GOLDEN, f"{n}.inp")) for n in ("cantilever", "cantilever3d", "plate_hole")]
    return ms + small_models()


MODELS = _models()
IDS = [m.name for m in MODELS]


@pytest.mark.parametrize("model", MODELS, ids=IDS)
def test_pattern_and_values(model, ctx):
    r = oracle.solve(model, stop_after_assembly=True)
    mesh, sys_ = assemble(model, ctx)
    rp, ci = mesh.pattern()
    assert np.array_equal(rp, r.row_ptr) and np.array_equal(ci, r.col_idx)
    vals, f = sys_.values()
    err = np.abs(vals - r.vals) / np.maximum(np.abs(r.vals), 1e-300 + np.abs(r.vals).max() * 1e-30)
    exact = np.count_nonzero(vals == r.vals)
    print(f"{model.name}: nnz {vals.size}, bit-exact {exact}, max rel {np.abs(vals - r.vals).max() / np.abs(r.vals).max():.2e}")
    assert np.abs(vals - r.vals).max() <= 1e-12 * np.abs(r.vals).max()
    assert np.array_equal(f, r.f)


def test_spmv_ordered_is_bit_exact_and_tuned_is_close(ctx):
    m = decks.hex_model(40, 12, 12, 2.0, 0.1, 0.1, E=1.0, nu=0.25, jitter=0.1, rng=decks.SplitMix64(3))
    m.bcs = [(np.array([1], np.uint32), np.array([1, 2], np.int32))]
    r = oracle.solve(m, stop_after_assembly=True)
    A = CsrMatrix(r.row_ptr, r.col_idx, r.vals, ctx)
    x = oracle.lcg_vector(A.n)
    y_ref = oracle.spmv(r.row_ptr, r.col_idx, r.vals, x)
    y_ord = A.mat_vec(x, ordered=True)
    assert np.array_equal(y_ord.view(np.uint64), y_ref.view(np.uint64))
    y = A.mat_vec(x)
    absrow = oracle.spmv(r.row_ptr, r.col_idx, np.abs(r.vals), np.abs(x))
    assert np.all(np.abs(y - y_ref) <= 1e-14 * absrow + 1e-300)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
@pytest.mark.parametrize("model", MODELS, ids=IDS)
def test_raw_pcg_matches_oracle(model, precision, ctx):
    r = oracle.solve(model)
    assert r.status == 0
    v, f = dirichlet(r.row_ptr, r.col_idx, r.vals, r.f, r.reaction_rows)
    A = CsrMatrix(r.row_ptr, r.col_idx, v, ctx)
    u, its, res = A.pcg(f, 1e-10, 0, precision)
    e = rel_err(u, r.u)
    fl = eval_floor(r.row_ptr, r.col_idx, v, f, u)
    print(f"{model.name} [{precision}]: iterations {its} (oracle {r.iterations}), residual {res:.3e} "
          f"(floor {fl:.1e}), rel err u {e:.2e}")
    # Both precisions stop only at residual <= tol: FP64 on its recursive
    # residual (solver.rs:131), FP32-IR on the FP64 true residual or, from the
    # FP64 evaluation floor on, on the recursive residual of FP64 PCG.
    assert res <= 1e-10
    assert e <= 1e-8


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
@pytest.mark.parametrize("model", MODELS, ids=IDS)
def test_solve_model_matches_oracle(model, precision, ctx):
    r = oracle.solve(model)
    s = solve_with_method(model, Method.AUTO, precision, ctx)
    eu = rel_err(s.displacements, r.displacements)
    ev = rel_err(s.von_mises, r.von_mises)
    en = rel_err(s.nodal_von_mises, r.nodal_von_mises)
    er = rel_err([x.value for x in s.reactions], r.reaction_value)
    print(f"{model.name} [{precision}]: its {s.iterations} (oracle {r.iterations}) res {s.final_residual:.2e} "
          f"err u {eu:.1e} vm {ev:.1e} nodal {en:.1e} react {er:.1e} cond {s.cond_estimate:.3e}/{r.cond_estimate:.3e}")
    assert s.final_residual <= 1e-10
    assert eu <= 1e-8 and ev <= 1e-8 and en <= 1e-8
    assert er <= 1e-8
    assert [x.node_id for x in s.reactions] == list(r.reaction_node)
    assert [x.dof for x in s.reactions] == list(r.reaction_dof)
    assert s.cond_estimate == pytest.approx(r.cond_estimate, rel=1e-12)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
@pytest.mark.parametrize("model", MODELS, ids=IDS)
def test_stopping_rule_and_true_residual(model, precision, ctx):
    """Every converged solve reached tol by a stated test, and the true FP64
    residual of its u is reported; both sit at or below the FP64 evaluation
    floor of f - K u, recomputed here on the host from the oracle's K."""
    r = oracle.solve(model)
    mesh, sy = assemble(model, ctx)
    sy.apply_dirichlet(constrained_dofs(model)[0])
    u, its, res = sy.solve(1e-10, 0, precision)
    st = sy.stats()
    v, f = dirichlet(r.row_ptr, r.col_idx, r.vals, r.f, r.reaction_rows)
    host_true = float(np.linalg.norm(f - _csr_mul(r.row_ptr, r.col_idx, v, u)) / np.linalg.norm(f))
    fl = eval_floor(r.row_ptr, r.col_idx, v, f, u)
    print(f"{model.name} [{precision}]: stop {st['converged_by']} fallback {st['fallback_reason']} residual "
          f"{st['residual']:.2e} true {st['true_residual']:.2e} (host {host_true:.2e}) floor {st['true_floor']:.2e} "
          f"(host {fl:.2e}) fp64 its {st['fp64_iterations']} of {st['iterations']}")
    assert st["status"] == 0 and st["residual"] <= 1e-10
    if precision == Precision.FP64:
        assert st["converged_by"] == _abi.SVX_STOP_RECURSIVE and st["fallback_reason"] == 0
    else:
        # an FP64 true residual at tol, or FP64 PCG (which the FP32 phase hands
        # over to at the floor or on a stall) on its recursive residual
        assert st["converged_by"] in (_abi.SVX_STOP_TRUE, _abi.SVX_STOP_RECURSIVE)
        if st["converged_by"] == _abi.SVX_STOP_RECURSIVE:
            assert st["fallback_reason"] in (1, 2) and st["fp64_iterations"] > 0
    # at the floor the residual is rounding noise: two FP64 evaluations in
    # different orders (device rows, scipy) may differ by up to a floor
    assert abs(st["true_residual"] - host_true) <= 1e-6 * host_true + fl
    assert st["true_floor"] == pytest.approx(fl, rel=1e-6)
    assert st["true_residual"] <= 1e-10 + 4.0 * st["true_floor"]
    assert rel_err(u, r.u) <= 1e-8


def _csr_mul(rp, ci, v, x):
    import scipy.sparse as sp
    n = rp.size - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n)) @ x


def test_known_answers_on_device(ctx):
    import json
    ka = json.load(open(os.path.join(GOLDEN, "known_answers.json")))
    for name in ("cantilever", "cantilever3d", "plate_hole"):
        m = read_inp(os.path.join(GOLDEN, ka[name]["deck"]))
        for prec in (Precision.FP64, Precision.FP32_IR):
            s = solve_with_method(m, Method.AUTO, prec, ctx)
            assert float(f"{s.max_disp() * 1e3:.3f}") == ka[name]["max_disp_mm"]
            assert float(f"{s.max_von_mises() / 1e6:.3f}") == ka[name]["max_von_mises_mpa"]


def test_patch_tests_on_device(ctx):
    m = decks.hex_model(1, 1, 1, 1, 1, 1, E=210e9, nu=0.3)
    m.bcs = [(np.array([1, 2, 3, 4], np.uint32), np.array([3], np.int32)),
             (np.array([1], np.uint32), np.array([1], np.int32)),
             (np.array([1, 2], np.uint32), np.array([2], np.int32))]
    m.cload_node, m.cload_dof, m.cload_value = np.array([5, 6, 7, 8], np.uint32), np.full(4, 3, np.int32), np.full(4, 2.5e5)
    for prec in (Precision.FP64, Precision.FP32_IR):
        s = solve_with_method(m, Method.AUTO, prec, ctx)
        assert np.all(np.abs(s.von_mises - 1e6) < 1e-6 * 1e6)
        assert np.all(np.abs(s.nodal_von_mises - 1e6) < 1e-6 * 1e6)
        assert abs(sum(x.value for x in s.reactions if x.dof == 3) + 1e6) < 1.0


def test_errors_match_reference(ctx):
    q = decks.quad_model(1, 1, 1, 1)
    q.cload_node, q.cload_dof, q.cload_value = np.array([2], np.uint32), np.array([1], np.int32), np.array([1.0])
    with pytest.raises(SolveError) as e:
        solve_with_method(q, Method.AUTO, Precision.FP64, ctx)
    assert e.value.kind == "NoConstraints"
    h = decks.hex_model(2, 1, 1, 2, 1, 1)
    h.conn_ids = h.conn_ids.copy()
    h.conn_ids[8 + 1], h.conn_ids[8 + 2] = h.conn_ids[8 + 2], h.conn_ids[8 + 1]  # element 2 inverted
    h.bcs = [(np.array([1], np.uint32), np.array([1, 2, 3], np.int32))]
    r = oracle.solve(h)
    with pytest.raises(SolveError) as e:
        solve_with_method(h, Method.AUTO, Precision.FP64, ctx)
    assert e.value.kind == "InvalidModel" and str(e.value.message) == r.message
    assert e.value.element_index == 1 and e.value.reason == _abi.SVX_ELEM_NONPOSITIVE_JACOBIAN
    q = decks.quad_model(2, 1, 2, 1)
    q.bcs = [(np.array([99], np.uint32), np.array([1], np.int32))]
    with pytest.raises(SolveError) as e:
        solve_with_method(q, Method.AUTO, Precision.FP64, ctx)
    assert e.value.message == oracle.solve(q).message
    q.bcs = [(np.array([1], np.uint32), np.array([1], np.int32))]
    q.materials = [("MAT", 70e9, 1.2)]
    with pytest.raises(SolveError) as e:
        solve_with_method(q, Method.AUTO, Precision.FP64, ctx)
    assert e.value.message == "element 1: invalid material constants"


def test_not_converged_budget(ctx):
    m = read_inp(os.path.join(GOLDEN, "cantilever.inp"))
    mesh, sys_ = assemble(m, ctx)
    rows, _, _ = constrained_dofs(m)
    sys_.apply_dirichlet(rows)
    with pytest.raises(SolveError) as e:
        sys_.solve(1e-10, 25, Precision.FP64)
    assert e.value.kind == "NotConverged" and e.value.iterations == 25
    r = oracle.solve(m, max_iter=25)
    assert e.value.residual == pytest.approx(r.residual, rel=1e-6)


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_every_launch_shape_matches_oracle(precision, ctx):
    """Blocks in flight per lane 1..4 in the ELL SpMV, one-chunk and
    multi-chunk tiles (48 and 576 rows round up to whole 64-node-row chunks),
    and 2 or 64 lanes per CSR row in the FP64 residual / warm-start passes, on
    a 3D and a 2D model. The unroll only changes how many loads are in flight,
    not the order of the sums: at a fixed tile size every unroll gives the
    same bits."""
    models = [m for m in small_models() if m.name.startswith(("cfg2_c3d8", "cps4_24x6"))]
    assert len(models) == 2
    try:
        for m in models:
            r = oracle.solve(m)
            for rpt in (48, 576):
                for lanes in (2, 64):
                    first = None
                    for unroll in (1, 2, 3, 4):
                        ctx.set_tuning(lanes=lanes, unroll=unroll, rows_per_tile=rpt)
                        s = solve_with_method(m, Method.AUTO, precision, ctx)
                        e = rel_err(s.displacements, r.displacements)
                        print(f"{m.name} [{precision}] unroll {unroll} lanes {lanes} rows/tile {rpt}: its "
                              f"{s.iterations} (oracle {r.iterations}) err u {e:.1e}")
                        assert e <= 1e-8
                        if first is None:
                            first = s.displacements
                        else:
                            np.testing.assert_array_equal(s.displacements, first)
    finally:
        ctx.set_tuning()


def test_spmv_rotation_hook(ctx):
    """The rotation timing hook launches each system's own SpMV: it reports
    the mean of the systems' algorithmic bytes, leaves their solutions as
    they were, and refuses a system not solved with that precision."""
    systems, meshes = [], []
    for k in range(2):
        m = decks.cfg2(seed=decks.SEED_BASE + 2 + 0x100 * k, nx=20, ny=4, nz=4)
        mesh, s = assemble(m, ctx)
        rows, _, _ = constrained_dofs(m)
        s.apply_dirichlet(rows)
        meshes.append(mesh)
        systems.append(s)
    us = [s.solve(1e-10, 0, Precision.FP64)[0].copy() for s in systems]
    single = [s.time_spmv(Precision.FP64, 3)[1] for s in systems]
    ms, by = type(systems[0]).time_spmv_rotation(systems, Precision.FP64, 5)
    assert ms > 0.0 and by == pytest.approx(sum(single) / 2, rel=1e-12)
    for s, u in zip(systems, us):
        assert np.array_equal(s.solution(), u)
    with pytest.raises(SolveError):
        type(systems[0]).time_spmv_rotation(systems, Precision.FP32_IR, 5)
    for s in systems:
        s.close()
    for mesh in meshes:
        mesh.close()
