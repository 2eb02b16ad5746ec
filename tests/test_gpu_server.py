"""Request server: continuous batching with footprint admission.

Each result must equal the same deck solved alone (solve_with_method with the
same precision) within the displacement / stress bar (1e-8 relative); with the
fixed per-block-size launch shape the engine runs every instance exactly as
alone, so the results are expected to be bit-identical and the test reports
whether they are. Requests arrive while others are on the device (staggered
submission, a slot limit below the request count).
"""
import time

import numpy as np
import pytest

from solvix_amd import decks
from solvix_amd.core import Context, Server, solve_with_method
from solvix_amd.model import Method, Precision, SolveError

from helpers import rel_err, small_models

pytestmark = pytest.mark.gpu


def _decks():
    ms = small_models()
    ms += [decks.solid_instance(decks.SplitMix64(0x5E + i), 3e3, 2e4, name=f"blk{i}") for i in range(3)]
    ms += [decks.plane_instance(decks.SplitMix64(0x77 + i), 2e3, 1e4, name=f"pl{i}") for i in range(3)]
    return ms


@pytest.mark.parametrize("precision,operator,graph", [(Precision.FP64, "assembled", False),
                                                      (Precision.FP32_IR, "assembled", False),
                                                      (Precision.FP64, "matrix_free", False),
                                                      (Precision.FP32_IR, "matrix_free", False),
                                                      (Precision.FP64, "assembled", True),
                                                      (Precision.FP32_IR, "assembled", True)],
                         ids=["fp64", "fp32ir", "fp64_matrix_free", "fp32ir_matrix_free", "fp64_graph",
                              "fp32ir_graph"])
def test_results_alone_equal_results_in_the_mix(precision, operator, graph):
    """matrix_free: a server created on a context with that operator runs the
    FP64 iterations of its instances (plane meshes and C3D8 blocks of one
    material; FP32-IR instances in their FP64 phase) on it, as the same
    context's solve does alone. graph: the server's chunks run as
    captured hipGraphs (svx_server_config.use_graph = 1)."""
    models = _decks()
    actx, sctx = Context(0), Context(0)
    actx.set_operator(operator)
    sctx.set_operator(operator)
    try:
        alone = [solve_with_method(m, Method.AUTO, precision, actx) for m in models]
        srv = Server(sctx, precision, hbm_budget=1 << 30, max_active=4, use_graph=graph)
        sctx.set_operator("assembled")  # the server keeps the operator it was created with
        _check_mix(models, alone, srv, precision)
    finally:
        actx.close()
        sctx.close()


def _check_mix(models, alone, srv, precision):
    try:
        tickets = [srv.submit(m) for m in models[:4]]
        time.sleep(0.05)  # the rest arrives while the first ones are on the device
        tickets += [srv.submit(m) for m in models[4:]]
        bitwise = 0
        for m, t, a in zip(models, tickets, alone):
            s, info = srv.wait(t, 60000.0)
            eu = rel_err(s.displacements, a.displacements)
            ev = rel_err(s.von_mises, a.von_mises)
            en = rel_err(s.nodal_von_mises, a.nodal_von_mises)
            same = np.array_equal(s.displacements, a.displacements)
            bitwise += same
            print(f"{m.name} [{precision}]: its {s.iterations} (alone {a.iterations}) err u {eu:.1e} vm {ev:.1e} "
                  f"nodal {en:.1e} bitwise {same} latency {info['latency_ms']:.1f} ms queue {info['queue_ms']:.1f} ms")
            assert eu <= 1e-8 and ev <= 1e-8 and en <= 1e-8
            assert s.iterations == a.iterations
            assert [r.node_id for r in s.reactions] == [r.node_id for r in a.reactions]
            assert rel_err([r.value for r in s.reactions], [r.value for r in a.reactions]) <= 1e-8
        st = srv.stats()
        print(f"bit-identical {bitwise}/{len(models)}; stats {st}")
        assert st["completed"] == len(models) and st["failed"] == 0
        assert 1 <= st["peak_active"] <= 4
        assert st["p99_ms"] >= st["p50_ms"] > 0
        # device-thread accounting: every request was admitted and retired,
        # and the engines ran at least one chunk per request's join
        assert st["chunks"] >= 1 and st["retiles"] >= 1 and st["retiles"] <= st["chunks"]
        assert st["admit_ms"] > 0 and st["solve_ms"] > 0 and st["retire_ms"] > 0 and st["idle_ms"] >= 0
    finally:
        srv.close()


def test_admission_respects_the_budget_and_reports_errors(ctx):
    models = _decks()[:4]
    sctx = Context(0)
    # budget for about one deck at a time: requests queue and run one by one
    srv = Server(sctx, Precision.FP64, hbm_budget=40 << 20, max_active=64)
    try:
        tickets = [srv.submit(m) for m in models]
        for m, t in zip(models, tickets):
            s, info = srv.wait(t, 60000.0)
            a = solve_with_method(m, Method.AUTO, Precision.FP64, ctx)
            assert rel_err(s.displacements, a.displacements) <= 1e-8
            assert info["device_bytes"] <= 40 << 20
        st = srv.stats()
        assert st["peak_resident_bytes"] <= 40 << 20
        # a deck larger than the whole budget is refused
        big = decks.cfg2(nx=80, ny=16, nz=16)
        t = srv.submit(big)
        with pytest.raises(SolveError) as e:
            srv.wait(t, 60000.0)
        assert e.value.kind == "Capacity"
        # model errors come back as solve_with_method reports them
        q = decks.quad_model(2, 1, 2, 1)
        q.cload_node, q.cload_dof, q.cload_value = np.array([2], np.uint32), np.array([1], np.int32), np.array([1.0])
        t = srv.submit(q)
        with pytest.raises(SolveError) as e:
            srv.wait(t, 60000.0)
        assert e.value.kind == "NoConstraints"
        h = decks.hex_model(2, 1, 1, 2, 1, 1)
        h.conn_ids = h.conn_ids.copy()
        h.conn_ids[8 + 1], h.conn_ids[8 + 2] = h.conn_ids[8 + 2], h.conn_ids[8 + 1]
        h.bcs = [(np.array([1], np.uint32), np.array([1, 2, 3], np.int32))]
        with pytest.raises(SolveError) as e_alone:
            solve_with_method(h, Method.AUTO, Precision.FP64, ctx)
        t = srv.submit(h)
        with pytest.raises(SolveError) as e:
            srv.wait(t, 60000.0)
        assert e.value.kind == "InvalidModel" and e.value.message == e_alone.value.message
    finally:
        srv.close()
        sctx.close()


def test_wait_refuses_short_reaction_arrays_and_keeps_the_ticket():
    """svx_server_wait writes reactions only into arrays that hold them all;
    a short capacity returns SVX_E_ARGUMENT and the ticket stays valid."""
    import ctypes as C
    from solvix_amd import _abi
    from solvix_amd._lib import lib
    m = small_models()[0]
    sctx = Context(0)
    srv = Server(sctx, Precision.FP64, hbm_budget=1 << 30, max_active=4)
    try:
        t = srv.submit(m)
        nr = m.reaction_count()
        assert nr > 1
        disp = np.zeros((m.n_nodes, 3))
        rv = np.zeros(nr)
        sol = _abi.SvxSolution(disp.ctypes.data_as(C.POINTER(C.c_double)), None, None, None, None,
                               rv.ctypes.data_as(C.POINTER(C.c_double)), nr - 1)
        err = _abi.SvxError()
        rc = lib().svx_server_wait(srv._h, t, -1.0, C.byref(sol), None, C.byref(err))
        assert rc == _abi.SVX_E_ARGUMENT and not rv.any()
        s, info = srv.wait(t, 60000.0)  # same ticket, full arrays
        assert len(s.reactions) == nr
    finally:
        srv.close()
        sctx.close()


def test_destroy_fails_pending_requests_and_releases_waiters():
    """Destroying a server with requests queued and in flight returns every
    blocked waiter with a Device error instead of leaving it waiting."""
    import threading
    models = _decks()
    sctx = Context(0)
    srv = Server(sctx, Precision.FP64, hbm_budget=1 << 30, max_active=1)
    tickets = [srv.submit(m) for m in models]
    outcomes = []

    def waiter(t):
        try:
            srv.wait(t, -1.0)
            outcomes.append("done")
        except SolveError as e:
            outcomes.append(e.kind)

    threads = [threading.Thread(target=waiter, args=(t,)) for t in tickets[-3:]]
    for th in threads:
        th.start()
    time.sleep(0.2)  # the waiters are blocked inside svx_server_wait
    srv.close()  # returns once every waiter has left the server
    for th in threads:
        th.join(timeout=60.0)
        assert not th.is_alive()
    print("outcomes", outcomes)
    assert len(outcomes) == 3 and all(o in ("done", "Device") for o in outcomes)
    sctx.close()


def test_deck_text_submissions_equal_solves_of_the_parsed_decks(ctx):
    """svx_server_submit_deck: the server's host workers read the deck text;
    each result equals solve_with_method on the deck read by inp.py, and a
    deck the reader rejects comes back as a Parse error with its line."""
    import os
    from solvix_amd.inp import parse_inp, write_inp
    golden = os.path.join(os.path.dirname(__file__), "golden")
    texts = [open(os.path.join(golden, f"{n}.inp")).read() for n in ("cantilever", "cantilever3d", "plate_hole")]
    texts += [write_inp(m) for m in small_models()[:3]]
    sctx = Context(0)
    srv = Server(sctx, Precision.FP64, hbm_budget=1 << 30, max_active=4)
    try:
        tickets = [srv.submit_deck(t) for t in texts]
        bad = srv.submit_deck("*NODE\n1, 0.0\n")
        for text, t in zip(texts, tickets):
            s, info = srv.wait(t, 60000.0)
            a = solve_with_method(parse_inp(text), Method.AUTO, Precision.FP64, ctx)
            print(f"deck: {info['n_nodes']} nodes, {info['n_elem']} elements, its {s.iterations} "
                  f"(alone {a.iterations}), bitwise {np.array_equal(s.displacements, a.displacements)}")
            assert rel_err(s.displacements, a.displacements) <= 1e-8
            assert rel_err(s.von_mises, a.von_mises) <= 1e-8
            assert [r.node_id for r in s.reactions] == [r.node_id for r in a.reactions]
        with pytest.raises(SolveError) as e:
            srv.wait(bad, 60000.0)
        assert e.value.kind == "Parse" and e.value.message.startswith("line 2: *NODE expects")
        assert srv.stats()["failed"] == 1
    finally:
        srv.close()
        sctx.close()
