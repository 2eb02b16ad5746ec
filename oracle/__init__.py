"""Oracle (test infrastructure only): ctypes binding to liboracle.so, the CPU
restatement of the reference's linear-static path in solvix_oracle.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package. The product (solvix_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

from solvix_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OracleOpts(C.Structure):
    _fields_ = [
        ("threads", C.c_int32),
        ("faithful", C.c_int32),
        ("tol", C.c_double),
        ("max_iter", C.c_int64),
        ("stop_after_assembly", C.c_int32),
        ("dup_seed", C.c_uint64),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
        L.oracle_solve.restype = vp
        L.oracle_solve.argtypes = [C.POINTER(_abi.SvxModelDesc), C.POINTER(OracleOpts)]
        for nm, rt in [("oracle_status", C.c_int), ("oracle_ndof", C.c_int), ("oracle_rows", i64),
                       ("oracle_nnz", i64), ("oracle_n_reactions", i64), ("oracle_iterations", i64),
                       ("oracle_residual", C.c_double), ("oracle_cond", C.c_double),
                       ("oracle_message", C.c_char_p)]:
            getattr(L, nm).restype = rt
            getattr(L, nm).argtypes = [vp]
        L.oracle_free.argtypes = [vp]
        L.oracle_times.argtypes = [vp, dp]
        L.oracle_get_csr.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp]
        L.oracle_get_f.argtypes = [vp, dp]
        L.oracle_get_u.argtypes = [vp, dp]
        L.oracle_get_solution.argtypes = [vp, dp, dp, dp]
        L.oracle_get_reactions.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_int32), dp]
        L.oracle_spmv.argtypes = [i64, C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp, dp, dp, C.c_int]
        L.oracle_pcg.restype = C.c_int
        L.oracle_pcg.argtypes = [i64, C.POINTER(C.c_int64), C.POINTER(C.c_int64), dp, dp, C.c_double, i64,
                                 C.c_int, dp, C.POINTER(C.c_int64), dp]
        L.oracle_element_stiffness.restype = C.c_int
        L.oracle_element_stiffness.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, dp]
        L.oracle_lcg_vector.argtypes = [i64, dp]
        L.oracle_worker_threads.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class OracleResult:
    status: int
    message: str
    ndof: int
    iterations: int
    residual: float
    cond_estimate: float
    times_ms: dict
    row_ptr: Optional[np.ndarray] = None
    col_idx: Optional[np.ndarray] = None
    vals: Optional[np.ndarray] = None
    f: Optional[np.ndarray] = None
    u: Optional[np.ndarray] = None
    displacements: Optional[np.ndarray] = None
    von_mises: Optional[np.ndarray] = None
    nodal_von_mises: Optional[np.ndarray] = None
    reaction_rows: Optional[np.ndarray] = None
    reaction_node: Optional[np.ndarray] = None
    reaction_dof: Optional[np.ndarray] = None
    reaction_value: Optional[np.ndarray] = None

    def max_disp(self) -> float:
        d = self.displacements
        return float(np.sqrt((d * d).sum(axis=1)).max(initial=0.0))

    def max_von_mises(self) -> float:
        return float(np.max(self.von_mises, initial=0.0))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def solve(model, threads: int = 0, faithful: bool = False, tol: float = 0.0, max_iter: int = 0,
          stop_after_assembly: bool = False, want_csr: bool = True, dup_seed: int = 0) -> OracleResult:
    """solve_with_method(model, Method::Pcg) restated on the CPU. dup_seed != 0
    sums each row's duplicate contributions in a seeded shuffled order (the
    order the reference leaves to its unstable sort, assembly.rs:97)."""
    L = lib()
    desc, keep = model.to_desc()
    opts = OracleOpts(threads, 1 if faithful else 0, tol, max_iter, 1 if stop_after_assembly else 0, dup_seed)
    h = L.oracle_solve(C.byref(desc), C.byref(opts))
    try:
        t = np.zeros(4)
        L.oracle_times(h, _dp(t))
        st = L.oracle_status(h)
        r = OracleResult(st, L.oracle_message(h).decode(), L.oracle_ndof(h), L.oracle_iterations(h),
                         L.oracle_residual(h), L.oracle_cond(h),
                         dict(assemble=t[0], pcg=t[1], recover=t[2], total=t[3]))
        n, nnz = L.oracle_rows(h), L.oracle_nnz(h)
        if n > 0 and want_csr:
            r.row_ptr = np.zeros(n + 1, np.int64)
            r.col_idx = np.zeros(nnz, np.int64)
            r.vals = np.zeros(nnz)
            L.oracle_get_csr(h, _ip(r.row_ptr), _ip(r.col_idx), _dp(r.vals))
            r.f = np.zeros(n)
            L.oracle_get_f(h, _dp(r.f))
        if st == _abi.SVX_OK and not stop_after_assembly and n > 0:
            r.u = np.zeros(n)
            L.oracle_get_u(h, _dp(r.u))
            r.displacements = np.zeros((model.n_nodes, 3))
            r.von_mises = np.zeros(model.n_elem)
            r.nodal_von_mises = np.zeros(model.n_nodes)
            L.oracle_get_solution(h, _dp(r.displacements), _dp(r.von_mises), _dp(r.nodal_von_mises))
            m = L.oracle_n_reactions(h)
            r.reaction_rows = np.zeros(m, np.int64)
            r.reaction_node = np.zeros(m, np.uint32)
            r.reaction_dof = np.zeros(m, np.int32)
            r.reaction_value = np.zeros(m)
            L.oracle_get_reactions(h, _ip(r.reaction_rows), r.reaction_node.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   r.reaction_dof.ctypes.data_as(C.POINTER(C.c_int32)), _dp(r.reaction_value))
        return r
    finally:
        L.oracle_free(h)
        del keep


def spmv(row_ptr, col_idx, vals, x, threads: int = 1) -> np.ndarray:
    """Csr::mat_vec_with (assembly.rs:46-72)."""
    rp = np.ascontiguousarray(row_ptr, np.int64)
    ci = np.ascontiguousarray(col_idx, np.int64)
    v = np.ascontiguousarray(vals, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(rp.shape[0] - 1)
    lib().oracle_spmv(rp.shape[0] - 1, _ip(rp), _ip(ci), _dp(v), _dp(x), _dp(y), threads)
    return y


def pcg(row_ptr, col_idx, vals, f, tol=1e-10, max_iter=0, threads: int = 1):
    """solver::pcg_with_threads (solver.rs:82-149) -> (status, u, iterations, residual)."""
    rp = np.ascontiguousarray(row_ptr, np.int64)
    ci = np.ascontiguousarray(col_idx, np.int64)
    v = np.ascontiguousarray(vals, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    n = rp.shape[0] - 1
    if max_iter <= 0:
        max_iter = max(1000, 10 * n)
    u = np.zeros(n)
    its = C.c_int64(0)
    res = C.c_double(0)
    st = lib().oracle_pcg(n, _ip(rp), _ip(ci), _dp(v), _dp(f), tol, max_iter, threads, _dp(u), C.byref(its),
                          C.byref(res))
    return st, u, its.value, res.value


def element_stiffness(kind: int, coords, E: float, nu: float, t: float = 1.0):
    """Element Ke (cps3/cps4/c3d8_stiffness) -> (reason, ke[N, N])."""
    nen = _abi.NODES_PER_KIND[kind]
    N = nen * (3 if kind == _abi.SVX_C3D8 else 2)
    x = np.zeros((nen, 3))
    c = np.asarray(coords, np.float64)
    x[:, :c.shape[1]] = c
    ke = np.zeros((N, N))
    rc = lib().oracle_element_stiffness(kind, _dp(np.ascontiguousarray(x)), E, nu, t, _dp(ke))
    return rc, ke


def lcg_vector(n: int) -> np.ndarray:
    """The LCG test vector of assembly.rs:800-808."""
    out = np.zeros(n)
    lib().oracle_lcg_vector(n, _dp(out))
    return out


def worker_threads() -> int:
    return lib().oracle_worker_threads()
