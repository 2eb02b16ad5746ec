"""The benchmark workloads at their full sizes on the device, against the
oracle: displacements, centroid and nodal von Mises, reactions (each within
1e-8 relative to its largest magnitude), the conditioning indicator, and the
reference's published column / gate decks.

Sizes: cfg1 400x40 CPS4 (32,882 DOF), cfg2 200x20x20 C3D8 (265,923 DOF), a
sample of the cfg4 load cases on the 120x12x12 cantilever (61,347 DOF) and a
seeded sample of cfg3 instances in one batch. The oracle solves each on the
host (about 2 s for cfg1, 15 s for cfg2).
"""
import numpy as np
import pytest

import oracle
from solvix_amd import decks
from solvix_amd.core import Batch, Context, assemble, constrained_dofs, solve_with_method
from solvix_amd.model import Method, Precision

from helpers import rel_err

pytestmark = pytest.mark.gpu
PRECISIONS = [Precision.FP64, Precision.FP32_IR]
PIDS = ["fp64", "fp32ir"]


def _check(name, precision, s, r):
    eu = rel_err(s.displacements, r.displacements)
    ev = rel_err(s.von_mises, r.von_mises)
    en = rel_err(s.nodal_von_mises, r.nodal_von_mises)
    er = rel_err([x.value for x in s.reactions], r.reaction_value)
    print(f"{name} [{precision}]: iterations {s.iterations} (oracle {r.iterations}), residual {s.final_residual:.2e}, "
          f"err u {eu:.1e} vm {ev:.1e} nodal {en:.1e} reactions {er:.1e}, cond {s.cond_estimate:.4e} "
          f"(oracle {r.cond_estimate:.4e}), solve {s.timings_ms}", flush=True)
    assert s.final_residual <= 1e-10
    assert eu <= 1e-8 and ev <= 1e-8 and en <= 1e-8 and er <= 1e-8
    assert [x.node_id for x in s.reactions] == list(r.reaction_node)
    assert s.cond_estimate == pytest.approx(r.cond_estimate, rel=1e-12)


@pytest.fixture(scope="module")
def cfg1_ref():
    m = decks.cfg1()
    return m, oracle.solve(m, want_csr=False)


@pytest.fixture(scope="module")
def cfg2_ref():
    m = decks.cfg2()
    print("cfg2: oracle solve (about 15 s on 8 host threads)", flush=True)
    return m, oracle.solve(m, want_csr=False)


@pytest.mark.parametrize("precision", PRECISIONS, ids=PIDS)
def test_cfg1_full_size_matches_oracle(precision, ctx, cfg1_ref):
    m, r = cfg1_ref
    assert m.n_dofs() == 32882
    _check("cfg1 CPS4 400x40", precision, solve_with_method(m, Method.AUTO, precision, ctx), r)


@pytest.mark.parametrize("precision", PRECISIONS, ids=PIDS)
def test_cfg2_full_size_matches_oracle(precision, ctx, cfg2_ref):
    m, r = cfg2_ref
    assert m.n_dofs() == 265923
    _check("cfg2 C3D8 200x20x20", precision, solve_with_method(m, Method.AUTO, precision, ctx), r)


def test_cfg4_load_cases_match_oracle(ctx):
    """Eight of the 64 load cases (every eighth) through the block solve of
    all 64, each against its own oracle solve."""
    model, cases = decks.cfg4()
    mesh, sy = assemble(model, ctx)
    sy.apply_dirichlet(constrained_dofs(model)[0])
    n = mesh.n
    assert n == 61347
    pos = {int(nid): i for i, nid in enumerate(model.node_ids)}
    F = np.zeros((len(cases), n))
    for k, (nodes, dofs, vals) in enumerate(cases):
        for nid, d, v in zip(nodes, dofs, vals):
            F[k, pos[int(nid)] * 3 + int(d) - 1] += v
    U, its, res, ok, ms = sy.solve_multi(F)
    assert ok.all() and (res <= 1e-10).all()
    for k in range(0, len(cases), 8):
        m = decks.cfg4()[0]
        m.cload_node, m.cload_dof, m.cload_value = cases[k]
        r = oracle.solve(m, want_csr=False)
        e = rel_err(U[k], r.u)
        print(f"cfg4 case {k}: iterations {its[k]} (oracle {r.iterations}), err u {e:.1e}", flush=True)
        assert e <= 1e-8


@pytest.mark.parametrize("precision,operator", [(Precision.FP64, "assembled"), (Precision.FP32_IR, "assembled"),
                                                (Precision.FP64, "matrix_free")],
                         ids=["fp64", "fp32ir", "fp64_matrix_free"])
def test_cfg3_sample_batched_matches_oracle(precision, operator):
    """Sixteen seeded cfg3 instances (every 32nd of the 512, CPS3 and CPS4,
    up to 2e5 DOF) solved in one batch, each against the oracle."""
    models = decks.cfg3()[::32]
    ctx = Context(0)
    ctx.set_operator(operator)
    try:
        keep, b = [], Batch(ctx)
        for m in models:
            mesh, sy = assemble(m, ctx)
            sy.apply_dirichlet(constrained_dofs(m)[0])
            keep.append((mesh, sy))
            b.add(sy)
        results = b.solve(precision)
        worst = 0.0
        for (mesh, sy), (its, res, ok), m in zip(keep, results, models):
            assert ok and res <= 1e-10, m.name
            r = oracle.solve(m, want_csr=False)
            e = rel_err(sy.solution(), r.u)
            worst = max(worst, e)
            assert e <= 1e-8, (m.name, e)
        print(f"cfg3 sample [{precision}, {operator}]: {len(models)} instances, "
              f"{sum(m.n_dofs() for m in models)} DOF, worst err u {worst:.1e}", flush=True)
    finally:
        ctx.close()


def test_cond_estimate_and_gate_decks_on_device(ctx):
    """cond_estimate on the device equals the closed forms of the column and
    the unit cube (lib.rs:634-644); the v0.1.0 gate deck's printed answers."""
    E, nu = 210e9, 0.3
    D11, D33 = E / (1 - nu * nu), E / (2 * (1 + nu))
    col = decks.quad_model(1, 5, 0.1, 1.0, E, nu, 1.0)
    col.bcs = [(np.array([1, 2], np.uint32), np.array([1, 2], np.int32))]
    col.cload_node, col.cload_dof, col.cload_value = np.array([11], np.uint32), np.array([2], np.int32), np.array([-2500.0])
    cube = decks.hex_model(1, 1, 1, 1, 1, 1, E=E, nu=nu)
    cube.bcs = [(np.array([1, 2, 3, 4], np.uint32), np.array([3], np.int32)),
                (np.array([1], np.uint32), np.array([1], np.int32)),
                (np.array([1, 2], np.uint32), np.array([2], np.int32))]
    cube.cload_node, cube.cload_dof, cube.cload_value = (np.array([5, 6, 7, 8], np.uint32), np.full(4, 3, np.int32),
                                                         np.full(4, 2.5e5))
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    gate = decks.readme_cantilever(80, 8)
    gate.cload_value = np.full(gate.cload_value.size, -111.11111111)
    for prec in PRECISIONS:
        s = solve_with_method(col, Method.AUTO, prec, ctx)
        assert s.cond_estimate == pytest.approx(2 * (D11 * 0.2 / 0.3 + D33 * 0.1 / 0.6), rel=1e-13)
        assert float(f"{s.max_von_mises() / 1e6:.3f}") == 0.025
        s = solve_with_method(cube, Method.AUTO, prec, ctx)
        assert s.cond_estimate == pytest.approx((lam + 4 * mu) / 9, rel=1e-13)
        s = solve_with_method(gate, Method.AUTO, prec, ctx)
        assert float(f"{s.max_disp() * 1e3:.3f}") == 14.877
        assert float(f"{s.max_von_mises() / 1e6:.3f}") == 100.344
