"""Batched solves on the device: every instance's result equals its solo
solve bit for bit (alone vs in the mix), for FP64 and FP32-IR, and matches
the oracle within the north-star tolerance."""
import numpy as np
import pytest

import oracle
from solvix_amd import decks
from solvix_amd.core import Batch, assemble, constrained_dofs
from solvix_amd.model import Precision

from helpers import rel_err, small_models

pytestmark = pytest.mark.gpu


def _system(model, ctx):
    mesh, sys_ = assemble(model, ctx)
    rows, _, _ = constrained_dofs(model)
    sys_.apply_dirichlet(rows)
    return mesh, sys_


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_batch_equals_solo_bitwise(precision, ctx):
    models = small_models() + [decks.plane_instance(decks.SplitMix64(0x53170003 ^ i), 2e3, 2e4) for i in range(6)]
    solo = []
    for m in models:
        mesh, s = _system(m, ctx)
        u, its, res = s.solve(1e-10, 0, precision)
        solo.append((u, its, res))
    keep = []
    b = Batch(ctx)
    for m in models:
        mesh, s = _system(m, ctx)
        keep.append((mesh, s))
        b.add(s)
    results = b.solve(precision)
    for (mesh, s), (u0, its0, res0), (its, res, ok), m in zip(keep, solo, results, models):
        assert ok, m.name
        u = s.solution()
        print(f"{m.name} [{precision}]: its {its} res {res:.2e} solo its {its0}")
        assert its == its0 and res == res0, m.name
        assert np.array_equal(u.view(np.uint64), u0.view(np.uint64)), m.name


@pytest.mark.parametrize("precision,operator", [(Precision.FP64, "assembled"), (Precision.FP32_IR, "assembled"),
                                                (Precision.FP64, "matrix_free"), (Precision.FP32_IR, "matrix_free")],
                         ids=["fp64", "fp32ir", "fp64-mf", "fp32ir-mf"])
def test_batch_streams_equal_one_stream_bitwise(precision, operator):
    """A batch split over 3 engines on 3 streams (svx_set_batch_streams), 2D
    and 3D instances of uneven sizes, gives every instance the bits of the
    one-engine solve, in both precisions and operators (FP32-IR instances
    that fall back finish in FP64 within their own lane)."""
    from solvix_amd.core import Context
    models = small_models() + [decks.plane_instance(decks.SplitMix64(0x53170005 ^ i), 2e3, 3e4) for i in range(7)]
    out = {}
    for streams in (1, 3):
        ctx = Context(0)
        try:
            ctx.set_batch_streams(streams)
            ctx.set_operator(operator)
            keep = []
            b = Batch(ctx)
            for m in models:
                mesh, s = _system(m, ctx)
                keep.append((mesh, s))
                b.add(s)
            res = b.solve(precision)
            out[streams] = [(r, s.solution()) for r, (mesh, s) in zip(res, keep)]
            b.close()
            for mesh, s in keep:
                s.close()
                mesh.close()
        finally:
            ctx.close()
    for m, ((r1, u1), (r3, u3)) in zip(models, zip(out[1], out[3])):
        assert r1[2] and r3[2], m.name
        assert r1[0] == r3[0] and r1[1] == r3[1], m.name
        assert np.array_equal(u1.view(np.uint64), u3.view(np.uint64)), m.name


def test_batch_matches_oracle(ctx):
    models = [decks.plane_instance(decks.SplitMix64(0x53170003 ^ i), 2e3, 1.5e4) for i in range(8)]
    models.append(decks.cfg2(nx=12, ny=3, nz=3))
    b = Batch(ctx)
    keep = []
    for m in models:
        mesh, s = _system(m, ctx)
        keep.append((mesh, s))
        b.add(s)
    for prec in (Precision.FP64, Precision.FP32_IR):
        res = b.solve(prec)
        for (mesh, s), m, (its, r, ok) in zip(keep, models, res):
            ref = oracle.solve(m)
            vm, vn, _, _ = s.recover([])
            e = rel_err(vm, ref.von_mises)
            print(f"{m.name} [{prec}] its {its} (oracle {ref.iterations}) res {r:.2e} err vm {e:.1e}")
            assert ok and e <= 1e-8


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_autotuned_batch_matches_oracle(precision, ctx):
    """The autotuner's shape changes summation orders only: results stay
    within the tolerance of the oracle, 2D and 3D in one batch."""
    models = [decks.plane_instance(decks.SplitMix64(0x53170003 ^ i), 2e3, 1.5e4) for i in range(5)]
    models.append(decks.cfg2(nx=12, ny=3, nz=3))
    b = Batch(ctx)
    keep = []
    for m in models:
        mesh, s = _system(m, ctx)
        keep.append((mesh, s))
        b.add(s)
    tuned = b.autotune(precision, 2)
    print(f"[{precision}] tuned {tuned}")
    assert tuned["spmv_ms"] > 0
    for bs in (2, 3):
        assert tuned["unroll"][bs] in (1, 2, 3, 4)
        assert tuned["rows_per_tile"][bs] % (64 * bs) == 0  # whole ELL chunks
        assert tuned["tile_order"][bs] in (0, 1)
    res = b.solve(precision)
    for (mesh, s), m, (its, r, ok) in zip(keep, models, res):
        ref = oracle.solve(m)
        e = rel_err(s.solution(), ref.u)
        print(f"{m.name} [{precision}] its {its} (oracle {ref.iterations}) err u {e:.1e}")
        assert ok and e <= 1e-8


@pytest.mark.parametrize("rows_per_tile", [128, 1024, 2048], ids=["one_chunk", "8_chunks", "16_chunks"])
@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_many_tiles_equal_solo(precision, rows_per_tile, ctx):
    """A batch with far more workgroup tiles than the device holds at once
    (so tiles of one instance start after others of that instance have
    finished) and tiles longer than a workgroup's preloaded rows: every
    result equals the instance's solo solve with the same shape bit for bit.
    Regression test for a kernel whose tiles read a scalar its own leader
    tile writes in the same launch."""
    models = [decks.plane_instance(decks.SplitMix64(0x7113 ^ i), 2e4, 6e4) for i in range(24)]
    try:
        ctx.set_tuning(rows_per_tile=rows_per_tile)
        solo = []
        for m in models:
            mesh, s = _system(m, ctx)
            solo.append(s.solve(1e-10, 0, precision))
        keep = []
        b = Batch(ctx)
        for m in models:
            mesh, s = _system(m, ctx)
            keep.append((mesh, s))
            b.add(s)
        results = b.solve(precision)
    finally:
        ctx.set_tuning()
    for (mesh, s), (u0, its0, res0), (its, res, ok), m in zip(keep, solo, results, models):
        print(f"{m.name} [{precision}] rows/tile {rows_per_tile}: its {its} (solo {its0}) res {res:.2e}")
        assert ok and its == its0 and res == res0, m.name
        assert np.array_equal(s.solution().view(np.uint64), u0.view(np.uint64)), m.name


def test_step_kernel_changes_no_bits():
    """Instances with more row tiles than the threshold take each pending
    direction step in the per-instance step kernel (and, past four times the
    threshold, each alpha in the per-instance alpha kernel): the same
    reductions over the same partial sums as every SpMV and update tile takes
    otherwise, so the bits and iteration counts do not change (one-chunk
    tiles give the 265,923-DOF block 1,385 tiles and a 1000x80 plane
    cantilever 1,267; a threshold of 300 puts both past the alpha kernel's
    1,200)."""
    from solvix_amd.core import Context
    models = [decks.cfg2(), decks.cfg1(nx=1000, ny=80)]
    runs = [(m, Precision.FP64) for m in models] + [(models[0], Precision.FP32_IR)]
    out = {}
    for tiles in (0, 300, -1):
        ctx = Context(0)
        ctx.set_tuning(rows_per_tile=1)  # one ELL chunk per tile
        ctx.set_step_kernel_tiles(tiles)
        try:
            res = []
            for m, prec in runs:
                mesh, s = _system(m, ctx)
                u, its, r = s.solve(1e-10, 0, prec)
                res.append((u, its, r))
                s.close()
                mesh.close()
            out[tiles] = res
        finally:
            ctx.close()
    for tiles in (0, 300):
        for (ua, ia, ra), (ub, ib, rb), (m, prec) in zip(out[tiles], out[-1], runs):
            print(f"{m.name} [{prec}] threshold {tiles}: {ia} / {ib} iterations, residual {ra:.3e} / {rb:.3e}")
            assert ia == ib and ra == rb, m.name
            assert np.array_equal(ua.view(np.uint64), ub.view(np.uint64)), m.name


def test_alpha_kernel_changes_no_bits_on_the_matrix_free_operator():
    """Matrix-free instances with more patches than four times the threshold
    (4 here) take alpha in the per-instance alpha kernel: the same bits and iteration
    counts as every update tile reducing the patch partials itself, for a
    plane cantilever and a jittered plane instance, in FP64 and in FP32-IR's
    FP64 phase."""
    from solvix_amd.core import Context
    models = [decks.cfg1(nx=400, ny=40), decks.plane_instance(decks.SplitMix64(0x91), 2e4, 6e4)]
    out = {}
    for tiles in (4, -1):
        ctx = Context(0)
        ctx.set_operator("matrix_free")
        ctx.set_step_kernel_tiles(tiles)
        try:
            res = []
            for m in models:
                for prec in (Precision.FP64, Precision.FP32_IR):
                    mesh, s = _system(m, ctx)
                    u, its, r = s.solve(1e-10, 0, prec)
                    assert s.operator == "matrix_free"
                    res.append((u, its, r))
                    s.close()
                    mesh.close()
            out[tiles] = res
        finally:
            ctx.close()
    for (ua, ia, ra), (ub, ib, rb) in zip(out[4], out[-1]):
        print(f"{ia} / {ib} iterations, residual {ra:.3e} / {rb:.3e}")
        assert ia == ib and ra == rb
        assert np.array_equal(ua.view(np.uint64), ub.view(np.uint64))


@pytest.mark.parametrize("operator", ["assembled", "matrix_free"])
@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_load_hint_changes_no_bits(precision, operator):
    """Nontemporal and ordinary matrix loads, the update kernels' vector loads
    and stores, and the matrix-free products' patch-list loads move the same
    values in the same order: single solves and a batch of 2D and 3D systems
    on either operator give identical bits and iteration counts under either
    hint."""
    from solvix_amd.core import Context
    models = small_models() + [decks.plane_instance(decks.SplitMix64(0x78 ^ i), 2e4, 6e4) for i in range(2)]
    models.append(decks.cfg2(nx=40, ny=6, nz=6))
    results = {}
    for hint in ("cached", "streaming"):
        ctx = Context(0)
        ctx.set_load_hint(hint)
        ctx.set_operator(operator)
        try:
            solo = []
            for m in models:
                mesh, s = _system(m, ctx)
                solo.append(s.solve(1e-10, 0, precision))
            keep, b = [], Batch(ctx)
            for m in models:
                mesh, s = _system(m, ctx)
                keep.append((mesh, s))
                b.add(s)
            its = [r[0] for r in b.solve(precision)]
            results[hint] = ([u for u, _, _ in solo], [k for _, k, _ in solo], [s.solution() for _, s in keep], its)
        finally:
            ctx.close()
    a, c = results["cached"], results["streaming"]
    assert a[1] == c[1] and a[3] == c[3]
    for x, y in zip(a[0] + a[2], c[0] + c[2]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_tile_order_changes_no_bits(precision):
    """XCD-contiguous and linear tile orders run the same tiles with the same
    partial-sum columns: single solves, a batch and the matrix-free operator
    give identical bits either way."""
    from solvix_amd.core import Context
    models = small_models() + [decks.plane_instance(decks.SplitMix64(0x77 ^ i), 2e4, 6e4) for i in range(3)]
    results = {}
    for order in ("xcd", "linear"):
        ctx = Context(0)
        ctx.set_tile_order(order)
        if precision == Precision.FP64:
            ctx.set_operator("matrix_free")
        try:
            solo = []
            for m in models:
                mesh, s = _system(m, ctx)
                solo.append(s.solve(1e-10, 0, precision)[0])
            keep, b = [], Batch(ctx)
            for m in models:
                mesh, s = _system(m, ctx)
                keep.append((mesh, s))
                b.add(s)
            b.solve(precision)
            results[order] = (solo, [s.solution() for _, s in keep])
        finally:
            ctx.close()
    for a, c in zip(results["xcd"][0] + results["xcd"][1], results["linear"][0] + results["linear"][1]):
        assert np.array_equal(a.view(np.uint64), c.view(np.uint64))
