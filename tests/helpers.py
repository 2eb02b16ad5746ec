"""Shared helpers for the parity tests."""
import numpy as np

import oracle
from solvix_amd import decks


def dirichlet(row_ptr, col_idx, vals, f, rows):
    """apply_dirichlet (solver.rs:45-62) on host arrays."""
    n = row_ptr.size - 1
    mask = np.zeros(n, bool)
    mask[np.asarray(rows, np.int64)] = True
    rr = np.repeat(np.arange(n), np.diff(row_ptr))
    hit = mask[rr] | mask[col_idx]
    v = vals.copy()
    v[hit] = np.where(rr[hit] == col_idx[hit], 1.0, 0.0)
    f2 = f.copy()
    f2[mask] = 0.0
    return v, f2


def rel_err(a, b):
    """max |a - b| / max |b| (the per-result error measure)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(b).max()
    return float(np.abs(a - b).max() / d) if d > 0 else float(np.abs(a - b).max())


def small_models():
    """Small jittered meshes of every element kind, quick on the oracle."""
    out = []
    rng = decks.SplitMix64(0x5EED)
    m = decks.quad_model(24, 6, 4.0, 1.0, 210e9, 0.3, 0.01, jitter=0.15, rng=rng, name="cps4_24x6")
    m.bcs = [(decks.left_edge_ids(24, 6), np.array([1, 2], np.int32))]
    tip = 3 * 25 + 25
    m.cload_node, m.cload_dof, m.cload_value = np.array([tip, tip], np.uint32), np.array([1, 2], np.int32), np.array([300.0, -1e4])
    out.append(m)
    out.append(decks.plane_instance(decks.SplitMix64(17), 2e3, 3e3, name="cps3_or_4"))
    tri = None
    for s in range(40):  # find a CPS3 instance
        cand = decks.plane_instance(decks.SplitMix64(100 + s), 1e3, 2e3, name="tri")
        if cand.kinds[0] == 0:
            tri = cand
            break
    out.append(tri)
    out.append(decks.cfg2(nx=16, ny=3, nz=3))
    out.append(decks.solid_instance(decks.SplitMix64(5), 3e3, 5e3, name="blk"))
    return [x for x in out if x is not None]


def eval_floor(row_ptr, col_idx, vals, f, u):
    """sqrt(m) * eps * || |K||u| + |f| || / ||f||: the rounding level of an FP64
    evaluation of f - K u (m = longest row). The reference's own solutions of
    ill-conditioned meshes have true residuals at this level."""
    import scipy.sparse as sp
    n = row_ptr.size - 1
    A = sp.csr_matrix((np.abs(vals), col_idx, row_ptr), shape=(n, n))
    m = np.diff(row_ptr).max()
    return float(np.sqrt(m) * np.finfo(float).eps * np.linalg.norm(A @ np.abs(u) + np.abs(f)) / np.linalg.norm(f))
