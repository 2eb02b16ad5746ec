"""CPU checks of the multi-RHS block SpMV's node groups (host planner, no
device): every node in exactly one group, every CSR entry packed exactly
once, and the group-by-group product indexed the way the kernel indexes it
against the plain CSR product (svx_blk_plan_check)."""
import numpy as np
import pytest

from solvix_amd import decks
from solvix_amd.core import blk_plan_check


def _positions(model):
    pos = {int(i): k for k, i in enumerate(model.node_ids)}
    return np.array([pos[int(i)] for i in model.conn_ids], np.int32)


@pytest.mark.parametrize("group", [0, 4, 8])
def test_solid_groups_reproduce_the_csr_product(group):
    m = decks.hex_model(14, 5, 4, 7.0, 2.5, 2.0, jitter=0.1, rng=decks.SplitMix64(11))
    st, d = blk_plan_check(3, m.coords, m.conn_ptr, _positions(m), group)
    g = group or 4
    assert st["groups"] == -(-m.coords.shape[0] // g) and d <= 1e-14


def test_cfg4_mesh_union_columns_per_node():
    """The 120x12x12 C3D8 cantilever: groups of 8 Morton-consecutive nodes
    (2x2x2 cubes of the structured grid) gather 8.7 union columns per node,
    against 24.2 blocks per node (27 inside) when each block gathers its own
    column."""
    m, _ = decks.cfg4(nrhs=1)
    st, d = blk_plan_check(3, m.coords, m.conn_ptr, _positions(m), 8)
    nn = m.coords.shape[0]
    assert d <= 1e-14
    assert st["blocks"] / nn > 24 and st["cols"] / nn < 9.0
    # groups of 4 (2x2x1): 12.3 columns per node
    st4, d4 = blk_plan_check(3, m.coords, m.conn_ptr, _positions(m), 4)
    assert d4 <= 1e-14 and st4["cols"] / nn < 13.0


def test_plane_groups():
    m = decks.quad_model(60, 20, 6.0, 2.0, jitter=0.15, rng=decks.SplitMix64(5))
    st, d = blk_plan_check(2, m.coords, m.conn_ptr, _positions(m), 8)
    assert d <= 1e-14
    # 4x2 node groups of a quad grid: a 6x4 union, 3 columns per node against 9
    assert st["cols"] / m.coords.shape[0] < 4.0


def test_node_numbering_does_not_matter():
    m = decks.hex_model(9, 4, 4, 3.0, 1.0, 1.0, jitter=0.1, rng=decks.SplitMix64(2))
    rng = np.random.default_rng(3)
    perm = rng.permutation(m.coords.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    coords = m.coords[perm]
    conn = inv[_positions(m)].astype(np.int32)
    st, d = blk_plan_check(3, coords, m.conn_ptr, conn)
    assert st["groups"] > 0 and d <= 1e-14


def test_isolated_node_packs_its_diagonal():
    """A node of no element has diagonal-only rows (assembly.rs:300-302); its
    packed block is the diagonal with zeros elsewhere."""
    m = decks.quad_model(6, 3, 3.0, 1.0)
    coords = np.vstack([m.coords, [[10.0, 10.0, 0.0]]])
    st, d = blk_plan_check(2, coords, m.conn_ptr, _positions(m))
    assert st["groups"] > 0 and d <= 1e-14


def test_group_size_is_checked():
    from solvix_amd.model import SolveError
    m = decks.quad_model(4, 2, 2.0, 1.0)
    with pytest.raises(SolveError):
        blk_plan_check(2, m.coords, m.conn_ptr, _positions(m), 5)
