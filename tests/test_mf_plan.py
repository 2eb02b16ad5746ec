"""CPU checks of the matrix-free operator's node patches (host planner, no
device): limits, single ownership and the patch-by-patch product, evaluated
the way the kernel does, against the element-by-element product."""
import numpy as np
import pytest

from solvix_amd import decks
from solvix_amd.core import mf_plan_check


def _positions(model):
    """Connectivity as node positions (the ABI's conn_idx)."""
    pos = {int(i): k for k, i in enumerate(model.node_ids)}
    return np.array([pos[int(i)] for i in model.conn_ids], np.int32)


def _check(model, ndof=2):
    return mf_plan_check(ndof, model.coords, model.kinds, model.conn_ptr, _positions(model))


def test_jittered_cps4_patches_reproduce_the_element_product():
    m = decks.quad_model(37, 23, 3.0, 1.0, 1.0, 0.3, 1.0, jitter=0.15, rng=decks.SplitMix64(3))
    st, d = _check(m)
    assert st["kind"] == 4
    assert d <= 1e-15
    assert st["cap_nodes"] <= 512 and st["cap_elem"] <= 256
    assert st["patches"] >= (38 * 24) // 256


def test_cps3_random_diagonals():
    m = None
    for s in range(60):
        cand = decks.plane_instance(decks.SplitMix64(200 + s), 2e4, 3e4)
        if cand.kinds[0] == 0:
            m = cand
            break
    assert m is not None
    st, d = _check(m)
    assert st["kind"] == 3
    assert d <= 1e-15
    assert st["cap_elem"] <= 512


def test_patch_overlap_is_small_on_a_structured_grid():
    m = decks.quad_model(300, 120, 3.0, 1.0, 1.0, 0.3, 1.0)
    st, d = _check(m)
    assert d <= 1e-15
    # every element is evaluated by each patch that owns one of its nodes
    assert st["elem_slots"] <= 1.35 * 300 * 120


def test_node_numbering_does_not_matter():
    rng = np.random.default_rng(7)
    m = decks.quad_model(40, 12, 4.0, 1.0, 1.0, 0.3, 1.0, jitter=0.1, rng=decks.SplitMix64(9))
    perm = rng.permutation(m.coords.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    coords = m.coords[perm]  # new node k is old node perm[k]
    conn = inv[_positions(m)].astype(np.int32)
    st, d = mf_plan_check(2, coords, m.kinds, m.conn_ptr, conn)
    assert st["kind"] == 4 and d <= 1e-15


def test_lone_node_and_cfg3_instance():
    m = decks.plane_instance(decks.SplitMix64(5), 5e3, 6e3)
    coords = np.vstack([m.coords, [[99.0, 99.0, 0.0]]])  # a node no element uses
    st, d = mf_plan_check(2, coords, m.kinds, m.conn_ptr, _positions(m))
    assert st["kind"] in (3, 4) and d <= 1e-15


def test_meshes_without_a_matrix_free_operator():
    solid = decks.hex_model(3, 2, 2, 3.0, 1.0, 1.0)
    st, d = _check(solid, ndof=3)
    assert st["kind"] == 0 and d is None
    # two element kinds in one mesh
    ids, coords, conn = decks.quad_grid(2, 1, 2.0, 1.0)
    kinds = np.array([1, 0, 0], np.uint8)
    tri = [conn[1][[0, 1, 2]], conn[1][[0, 2, 3]]]
    conn_ids = np.concatenate([conn[0], tri[0], tri[1]]).astype(np.int64) - 1
    conn_ptr = np.array([0, 4, 7, 10], np.int64)
    st, d = mf_plan_check(2, coords, kinds, conn_ptr, conn_ids.astype(np.int32))
    assert st["kind"] == 0
    # five quads around one node (more than four elements at a node)
    ang = np.linspace(0.0, 2 * np.pi, 6)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), 0 * ang], 1)
    mid = 0.5 * (ring + np.roll(ring, -1, 0))
    coords5 = np.vstack([[[0.0, 0.0, 0.0]], ring, 1.4 * mid])
    conn5 = []
    for k in range(5):
        conn5 += [0, 1 + k, 6 + k, 1 + (k + 1) % 5]
    st, d = mf_plan_check(2, coords5, np.ones(5, np.uint8), np.arange(6, dtype=np.int64) * 4,
                          np.array(conn5, np.int32))
    assert st["kind"] == 0
