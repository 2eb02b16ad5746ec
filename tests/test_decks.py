"""The synthetic workloads match their stated sizes and distributions
(seeded, deterministic), on the CPU."""
import numpy as np

import oracle
from solvix_amd import decks


def test_splitmix_is_deterministic_and_uniform():
    a = decks.SplitMix64(123)
    b = decks.SplitMix64(123)
    xa = np.array([a.u() for _ in range(2000)])
    xb = np.array([b.u() for _ in range(2000)])
    assert np.array_equal(xa, xb)
    assert 0.0 <= xa.min() and xa.max() < 1.0 and abs(xa.mean() - 0.5) < 0.03


def test_single_plane_cantilever_sizes_and_pattern():
    m = decks.cfg1()
    assert m.n_elem == 16000 and m.n_nodes == 16441 and m.n_dofs() == 32882
    r = oracle.solve(m, stop_after_assembly=True)
    nx, ny = 400, 40
    assert r.row_ptr[-1] == (3 * nx + 1) * (3 * ny + 1) * 4 == 581284


def test_single_solid_block_sizes():
    m = decks.cfg2()
    assert m.n_elem == 80000 and m.n_nodes == 88641 and m.n_dofs() == 265923
    total = float(np.sum(m.cload_value))
    assert abs(total + 1e6 * 2.0 * 2.0) < 1e-6 * 4e6  # uniform -1e6 Pa over the 2 x 2 m face


def test_batched_plane_distribution():
    ms = decks.cfg3(count=64)
    n = np.array([m.n_dofs() for m in ms])
    assert n.min() >= 1500 and n.max() <= 2.2e5
    quads = sum("cps4" in m.name for m in ms)
    assert 0.4 * 64 <= quads <= 0.8 * 64
    again = decks.cfg3(count=4)
    assert [m.n_dofs() for m in again] == [m.n_dofs() for m in ms[:4]]


def test_multi_rhs_cases():
    model, cases = decks.cfg4()
    assert model.n_dofs() == 61347 and len(cases) == 64
    for nodes, dofs, vals in cases:
        assert 3 <= nodes.size <= 24 and nodes.size % 3 == 0
        mag = np.linalg.norm(vals.reshape(-1, 3), axis=1)
        assert np.all((mag >= 1e3 * (1 - 1e-12)) & (mag <= 1e5 * (1 + 1e-12)))


def test_served_stream_rate_and_mix():
    reqs = list(decks.cfg5_stream(duration_s=2.0, rate=200.0))
    assert 300 <= len(reqs) <= 500
    t = np.array([r.arrival_s for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] <= 2.0
    solids = sum(r.model.ndof() == 3 for r in reqs)
    assert 0.7 * len(reqs) <= solids <= 0.9 * len(reqs)
