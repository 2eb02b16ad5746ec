"""The thermal solve's PCG call (heat.rs:84-87: solver::pcg on a scalar,
one-DOF-per-node conduction matrix after apply_temperatures, heat.rs:361-405)
served by the same C ABI entry point (svx_csr_create + svx_csr_pcg).

The system is the 5-point graph Laplacian of an nx x ny node grid with
prescribed temperatures on the left (0) and right (100) columns and insulated
top and bottom edges; a temperature linear in x is discrete-harmonic there,
so the exact solution is T = 100 i / (nx - 1) whatever the solver. The
jittered variant (edge weights from a seeded generator) is pinned against
the oracle's PCG on the same matrix.
"""
import numpy as np
import pytest

import oracle
from solvix_amd.core import CsrMatrix
from solvix_amd.model import Precision

from helpers import rel_err


def grid_laplacian(nx, ny, weights=None):
    """Sorted CSR (explicit diagonal) of sum_edges w (T_a - T_b)^2 / 2 on an
    nx x ny grid, node (i, j) at position j * nx + i."""
    rows = [dict() for _ in range(nx * ny)]
    k = 0
    for j in range(ny):
        for i in range(nx):
            a = j * nx + i
            rows[a].setdefault(a, 0.0)
            for di, dj in ((1, 0), (0, 1)):
                if i + di < nx and j + dj < ny:
                    b = (j + dj) * nx + i + di
                    w = 1.0 if weights is None else weights[k]
                    k += 1
                    rows[a][a] += w
                    rows[b][b] = rows[b].get(b, 0.0) + w
                    rows[a][b] = rows[a].get(b, 0.0) - w
                    rows[b][a] = rows[b].get(a, 0.0) - w
    rp, ci, v = [0], [], []
    for r in rows:
        for c in sorted(r):
            ci.append(c)
            v.append(r[c])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


def apply_temperatures(rp, ci, vals, f, prescribed):
    """heat.rs:361-405: free rows move K[a, r] T_r to the right-hand side and
    zero the column entry; prescribed rows become identity rows with f = T."""
    vals = vals.copy()
    f = f.copy()
    fixed = ~np.isnan(prescribed)
    for r in range(len(f)):
        if fixed[r]:
            continue
        for ii in range(rp[r], rp[r + 1]):
            c = ci[ii]
            if fixed[c]:
                f[r] -= vals[ii] * prescribed[c]
                vals[ii] = 0.0
    for r in np.nonzero(fixed)[0]:
        for ii in range(rp[r], rp[r + 1]):
            vals[ii] = 1.0 if ci[ii] == r else 0.0
        f[r] = prescribed[r]
    return vals, f


def plate(nx, ny, weights=None):
    rp, ci, v = grid_laplacian(nx, ny, weights)
    T = np.full(nx * ny, np.nan)
    for j in range(ny):
        T[j * nx] = 0.0
        T[j * nx + nx - 1] = 100.0
    return (rp, ci) + apply_temperatures(rp, ci, v, np.zeros(nx * ny), T)


def test_oracle_pcg_thermal_linear_field_is_exact():
    nx, ny = 60, 12
    rp, ci, v, f = plate(nx, ny)
    st, T, its, res = oracle.pcg(rp, ci, v, f, 1e-10)
    exact = np.tile(100.0 * np.arange(nx) / (nx - 1), ny)
    assert st == 0 and res <= 1e-10 and rel_err(T, exact) <= 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32_IR], ids=["fp64", "fp32ir"])
def test_thermal_pcg_linear_field_is_exact(precision, ctx):
    nx, ny = 60, 12
    rp, ci, v, f = plate(nx, ny)
    A = CsrMatrix(rp, ci, v, ctx)
    T, its, res = A.pcg(f, 1e-10, 0, precision)
    exact = np.tile(100.0 * np.arange(nx) / (nx - 1), ny)
    e = rel_err(T, exact)
    print(f"thermal {nx}x{ny} [{precision}]: iterations {its}, residual {res:.2e}, rel err vs exact {e:.2e}")
    assert e <= 1e-8


@pytest.mark.gpu
def test_thermal_pcg_matches_oracle_on_jittered_conductances(ctx):
    nx, ny = 80, 20
    rng = np.random.default_rng(7)
    n_edges = (nx - 1) * ny + nx * (ny - 1)
    rp, ci, v, f = plate(nx, ny, weights=rng.uniform(0.5, 2.0, n_edges))
    st, u_ref, its_ref, res_ref = oracle.pcg(rp, ci, v, f, 1e-10)
    assert st == 0
    A = CsrMatrix(rp, ci, v, ctx)
    T, its, res = A.pcg(f, 1e-10, 0, Precision.FP64)
    e = rel_err(T, u_ref)
    print(f"thermal jittered {nx}x{ny}: iterations {its} (oracle {its_ref}), residual {res:.2e} "
          f"(oracle {res_ref:.2e}), rel err {e:.2e}")
    assert res <= 1e-10 and e <= 1e-8
    assert abs(its - its_ref) <= max(3, its_ref // 50)
