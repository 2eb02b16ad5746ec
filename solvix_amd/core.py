"""Host API of the device path, mirroring solvix-core's linear-static seam.

    solve(model)                       lib.rs:556   (linear models)
    solve_with_method(model, method)   lib.rs:586
    Mesh.assemble(...)                 assembly.rs:281  assemble
    System.apply_dirichlet(rows)       solver.rs:45     apply_dirichlet
    System.solve / pcg(...)            solver.rs:66     pcg
    System.recover(rows)               stress.rs:22,53,103
    CsrMatrix.mat_vec(x)               assembly.rs:38   Csr::mat_vec
    constrained_dofs(model)            solver.rs:15

Everything runs through libsolvix_hip.so (include/svx.h); there is no CPU
fallback. Arguments and errors keep the reference's meaning: errors raise
SolveError with the reference's kinds and messages.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._lib import lib
from .model import Method, Model, Precision, Reaction, Solution, SolveError, error_from_code

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


def _d(a):
    return a.ctypes.data_as(_dp)


def _raise(rc: int, err: Optional[_abi.SvxError] = None):
    if rc == _abi.SVX_OK:
        return
    if err is not None and err.code:
        raise error_from_code(err.code, err.message.decode(errors="replace"), err.iterations, err.residual,
                              err.index, err.reason)
    raise error_from_code(rc, lib().svx_last_error().decode(errors="replace"))


class Context:
    """One device context (svx_ctx): a device, a HIP stream, tuning knobs."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _raise(lib().svx_ctx_create(device, C.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        lib().svx_device_name(self._h, buf, 256)
        return buf.value.decode()

    def set_tuning(self, lanes=0, unroll=0, rows_per_tile=0, period=0, chunk=0, use_graph=True, ru_delta=0.0):
        _raise(lib().svx_set_tuning(self._h, lanes, unroll, rows_per_tile, period, chunk, int(use_graph), ru_delta))

    def set_tile_order(self, order: str):
        """'xcd' (default: each XCD takes a contiguous run of tiles) or 'linear'
        (tile = workgroup) for the gathering kernels of subsequent solves."""
        code = {"xcd": _abi.SVX_TILES_XCD, "linear": _abi.SVX_TILES_LINEAR}[order]
        _raise(lib().svx_set_tile_order(self._h, code))

    def set_step_kernel_tiles(self, tiles: int):
        """Systems with more FP64 row tiles than `tiles` take each pending
        direction step in one per-system kernel (0: the default 1024,
        negative: never); same bits either way."""
        _raise(lib().svx_set_step_kernel_tiles(self._h, int(tiles)))

    def set_multi_group(self, nodes: int = 0):
        """Nodes per group of the multi-RHS block SpMV: 4 or 8 (0: default
        4), negative: no groups (the gathering SpMV)."""
        _raise(lib().svx_set_multi_group(self._h, int(nodes)))

    def set_batch_streams(self, streams: int = 0):
        """Engines a batch solve runs side by side, each on its own stream
        (0: the default); the same bits for every setting."""
        _raise(lib().svx_set_batch_streams(self._h, int(streams)))

    def set_load_hint(self, hint: str):
        """'auto' (default: nontemporal matrix loads once a launch's FP64 matrix
        values exceed 256 MiB), 'cached' or 'streaming' for the FP64 assembled
        SpMV of subsequent solves; results are the same bit for bit."""
        code = {"auto": _abi.SVX_LOADS_AUTO, "cached": _abi.SVX_LOADS_CACHED,
                "streaming": _abi.SVX_LOADS_STREAMING}[hint]
        _raise(lib().svx_set_load_hint(self._h, code))

    def set_operator(self, op: str):
        """'assembled' (default) or 'matrix_free': the operator of subsequent
        FP64 iterations (FP64 solves, and FP32-IR's FP64 phase) on plane meshes
        of one element kind and C3D8 meshes, each of one material."""
        code = {"assembled": _abi.SVX_OP_ASSEMBLED, "matrix_free": _abi.SVX_OP_MATRIX_FREE}[op]
        _raise(lib().svx_set_operator(self._h, code))

    def close(self):
        if self._h:
            lib().svx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ------------------------------------------------------------ whole model
def solve_with_method(model: Model, method: int = Method.AUTO, precision: int = Precision.FP64,
                      ctx: Optional[Context] = None) -> Solution:
    """solve_with_method (lib.rs:586) on the device."""
    ctx = ctx or default_context()
    desc, keep = model.to_desc()
    nn, ne = model.n_nodes, model.n_elem
    nr = max(0, int(lib().svx_model_reaction_count(C.byref(desc))))
    disp = np.zeros((nn, 3))
    vm = np.zeros(ne)
    nvm = np.zeros(nn)
    rn = np.zeros(nr, np.uint32)
    rd = np.zeros(nr, np.int32)
    rv = np.zeros(nr)
    sol = _abi.SvxSolution(_d(disp), _d(vm), _d(nvm), rn.ctypes.data_as(C.POINTER(C.c_uint32)),
                           rd.ctypes.data_as(_i32p), _d(rv), nr)
    err = _abi.SvxError()
    rc = lib().svx_solve_model(ctx.handle, C.byref(desc), method, precision, C.byref(sol), C.byref(err))
    del keep
    _raise(rc, err)
    return Solution(
        displacements=disp, von_mises=vm, nodal_von_mises=nvm,
        reactions=[Reaction(int(a), int(b), float(c)) for a, b, c in zip(rn, rd, rv)],
        iterations=int(sol.iterations), final_residual=float(sol.final_residual), solver_method="pcg",
        cond_estimate=float(sol.cond_estimate),
        timings_ms=dict(assemble=sol.assemble_ms, solve=sol.solve_ms, recover=sol.recover_ms))


def solve(model: Model, precision: int = Precision.FP64, ctx: Optional[Context] = None) -> Solution:
    """solve (lib.rs:556) for linear-static models."""
    return solve_with_method(model, Method.AUTO, precision, ctx)


def parse_deck(text: str, name: str = "model") -> Model:
    """parse_inp (crates/solvix-io/src/parser.rs:87) for the linear path's
    keyword subset, read by the library's C++ reader (svx_deck_parse; no
    device needed); raises SolveError kind "Parse" with "line N: ..."."""
    raw = text.encode()
    h = C.c_void_p()
    err = _abi.SvxError()
    _raise(lib().svx_deck_parse(raw, len(raw), C.byref(h), C.byref(err)), err)
    try:
        return Model.from_desc(lib().svx_deck_model(h).contents, name)
    finally:
        lib().svx_deck_destroy(h)


def constrained_dofs(model: Model) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """constrained_dofs (solver.rs:15-41) -> (rows, node_ids, dofs), BC order."""
    ndof = model.ndof()
    pos = {}
    for i, nid in enumerate(model.node_ids.tolist()):
        pos.setdefault(nid, i)
    rows, nodes, dofs = [], [], []
    for nids, ds in model.bcs:
        for nid in nids.tolist():
            if nid not in pos:
                raise SolveError("InvalidModel", f"boundary condition on unknown node {nid}")
            for d in ds.tolist():
                if d == 0 or d > ndof:
                    raise SolveError("InvalidModel", f"boundary condition on node {nid} has unsupported dof {d}")
                rows.append(pos[nid] * ndof + d - 1)
                nodes.append(nid)
                dofs.append(d)
    return np.array(rows, np.int64), np.array(nodes, np.uint32), np.array(dofs, np.int32)


def node_positions(model: Model) -> np.ndarray:
    """Element connectivity as node positions (Mesh::node_index, lib.rs:106)."""
    order = np.argsort(model.node_ids, kind="stable")
    sorted_ids = model.node_ids[order]
    idx = np.searchsorted(sorted_ids, model.conn_ids)
    if np.any(idx >= sorted_ids.size) or np.any(sorted_ids[np.minimum(idx, sorted_ids.size - 1)] != model.conn_ids):
        bad = model.conn_ids[(idx >= sorted_ids.size) | (sorted_ids[np.minimum(idx, sorted_ids.size - 1)] != model.conn_ids)][0]
        raise SolveError("InvalidModel", f"UnknownNode({bad})")
    return order[idx].astype(np.int32)


def element_materials(model: Model) -> Tuple[np.ndarray, np.ndarray]:
    """Per-element (E, nu) by material_for_element (lib.rs:357-367); NaN E
    marks an element without material."""
    E = np.full(model.n_elem, np.nan)
    nu = np.zeros(model.n_elem)
    done = np.zeros(model.n_elem, bool)
    es = {name: ids for name, ids in model.elsets}
    mats = {m[0]: (m[1], m[2]) for m in model.materials}
    first_es = {}
    for name, ids in model.elsets:
        first_es.setdefault(name, ids)
    for es_name, mat_name in model.assignments:
        if es_name not in first_es:
            continue
        hit = np.isin(model.elem_ids, first_es[es_name]) & ~done
        if mat_name in mats:
            E[hit], nu[hit] = mats[mat_name]
        done |= hit
    return E, nu


# --------------------------------------------------------------- the seam
def mf_plan_check(ndof: int, coords, kinds, conn_ptr, conn_idx):
    """Host-only check of the matrix-free operator's node patches (see
    svx_mf_plan_check): returns (stats dict, max relative difference of the
    patch-wise product against the element-by-element one, or None when the
    mesh has no matrix-free operator)."""
    coords = np.ascontiguousarray(coords, np.float64)
    kinds = np.ascontiguousarray(kinds, np.uint8)
    conn_ptr = np.ascontiguousarray(conn_ptr, np.int64)
    conn_idx = np.ascontiguousarray(conn_idx, np.int32)
    stats = np.zeros(5, np.int64)
    diff = C.c_double(-1.0)
    rc = lib().svx_mf_plan_check(ndof, coords.shape[0], _d(coords), kinds.size,
                                 kinds.ctypes.data_as(C.POINTER(C.c_uint8)), _abi.ptr(conn_ptr, C.c_int64),
                                 _abi.ptr(conn_idx, C.c_int32), _abi.ptr(stats, C.c_int64), C.byref(diff))
    _raise(rc)
    keys = ("kind", "patches", "cap_nodes", "cap_elem", "elem_slots")
    st = dict(zip(keys, (int(v) for v in stats)))
    return st, (diff.value if st["kind"] else None)


def blk_plan_check(ndof: int, coords, conn_ptr, conn_idx, group: int = 0):
    """Host-only check of the multi-RHS block SpMV's node groups (see
    svx_blk_plan_check): returns (stats dict, max relative difference of the
    group-wise product against the CSR product, or None when the mesh has no
    plan)."""
    coords = np.ascontiguousarray(coords, np.float64)
    conn_ptr = np.ascontiguousarray(conn_ptr, np.int64)
    conn_idx = np.ascontiguousarray(conn_idx, np.int32)
    stats = np.zeros(3, np.int64)
    diff = C.c_double(-1.0)
    rc = lib().svx_blk_plan_check(ndof, coords.shape[0], _d(coords), conn_ptr.size - 1,
                                  _abi.ptr(conn_ptr, C.c_int64), _abi.ptr(conn_idx, C.c_int32), int(group),
                                  _abi.ptr(stats, C.c_int64), C.byref(diff))
    _raise(rc)
    st = dict(zip(("groups", "cols", "blocks"), (int(v) for v in stats)))
    return st, (diff.value if st["groups"] else None)


class Mesh:
    """Device-resident mesh with its CSR pattern (built once per mesh)."""

    def __init__(self, ndof: int, coords: np.ndarray, kinds: np.ndarray, conn_ptr: np.ndarray,
                 conn_pos: np.ndarray, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.ndof = ndof
        self._coords = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
        kinds = np.ascontiguousarray(kinds, np.uint8)
        conn_ptr = np.ascontiguousarray(conn_ptr, np.int64)
        conn_pos = np.ascontiguousarray(conn_pos, np.int32)
        self.n_nodes = self._coords.shape[0]
        self.n_elem = kinds.shape[0]
        self._h = C.c_void_p()
        err = _abi.SvxError()
        rc = lib().svx_mesh_create(self.ctx.handle, ndof, self.n_nodes, _d(self._coords), self.n_elem,
                                   kinds.ctypes.data_as(C.POINTER(C.c_uint8)), conn_ptr.ctypes.data_as(_i64p),
                                   conn_pos.ctypes.data_as(_i32p), C.byref(self._h), C.byref(err))
        _raise(rc, err)
        self.n = int(lib().svx_mesh_rows(self._h))
        self.nnz = int(lib().svx_mesh_nnz(self._h))

    @classmethod
    def from_model(cls, model: Model, ctx: Optional[Context] = None) -> "Mesh":
        return cls(model.ndof(), model.coords, model.kinds, model.conn_ptr, node_positions(model), ctx)

    def pattern(self) -> Tuple[np.ndarray, np.ndarray]:
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        _raise(lib().svx_mesh_pattern(self._h, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p)))
        return rp, ci

    def assemble(self, E, nu, thickness: float, load_rows=(), load_values=()) -> "System":
        """assemble (assembly.rs:281): element pass + CSR values + load vector."""
        E = np.ascontiguousarray(np.broadcast_to(np.asarray(E, np.float64), (self.n_elem,)))
        nu = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, np.float64), (self.n_elem,)))
        lr = np.ascontiguousarray(load_rows, np.int64)
        lv = np.ascontiguousarray(load_values, np.float64)
        h = C.c_void_p()
        err = _abi.SvxError()
        rc = lib().svx_assemble(self._h, _d(E), _d(nu), float(thickness), lr.shape[0], lr.ctypes.data_as(_i64p),
                                _d(lv), C.byref(h), C.byref(err))
        _raise(rc, err)
        return System(self, h)

    def close(self):
        if self._h:
            lib().svx_mesh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class System:
    """An assembled K, f on a mesh (device resident)."""

    def __init__(self, mesh: Mesh, handle):
        self.mesh = mesh
        self._h = handle
        self.iterations = 0
        self.residual = 0.0

    @property
    def handle(self):
        return self._h

    def values(self) -> Tuple[np.ndarray, np.ndarray]:
        vals = np.zeros(self.mesh.nnz)
        f = np.zeros(self.mesh.n)
        _raise(lib().svx_system_values(self._h, _d(vals), _d(f)))
        return vals, f

    def solve_multi(self, F: np.ndarray, tol: float = 1e-10, max_iter: int = 0,
                    precision: int = Precision.FP64, raise_on_failure: bool = True):
        """Solve for every row of F ([nrhs, n], nrhs <= 64) on this system.
        Returns (U [nrhs, n], iterations, residuals, converged, device ms)."""
        F = np.ascontiguousarray(F, np.float64).reshape(-1, self.mesh.n)
        R = F.shape[0]
        U = np.zeros_like(F)
        its = np.zeros(R, np.int64)
        res = np.zeros(R)
        ok = np.zeros(R, np.int32)
        ms = C.c_double()
        err = _abi.SvxError()
        rc = lib().svx_solve_multi(self._h, R, _d(F), tol, max_iter, precision, _d(U), its.ctypes.data_as(_i64p),
                                   _d(res), ok.ctypes.data_as(_i32p), C.byref(ms), C.byref(err))
        if rc != _abi.SVX_OK and (raise_on_failure or rc != _abi.SVX_E_NOT_CONVERGED):
            _raise(rc, err)
        return U, its, res, ok.astype(bool), ms.value

    def multi_spmv_timing(self, nrhs: int = 64, reps: int = 100):
        """(average us per launch, algorithmic bytes per launch) of the FP64
        block solve's SpMV kernel for `nrhs` load cases (svx_multi_spmv_timing:
        with node groups the product of the direction the step pass wrote)."""
        us, by = C.c_double(), C.c_double()
        _raise(lib().svx_multi_spmv_timing(self._h, int(nrhs), int(reps), C.byref(us), C.byref(by)))
        return us.value, by.value

    @property
    def operator(self) -> str:
        """'matrix_free' once a matrix-free solve prepared this system's operator, else 'assembled'."""
        return "matrix_free" if lib().svx_system_operator(self._h) == _abi.SVX_OP_MATRIX_FREE else "assembled"

    def solution(self) -> np.ndarray:
        """The current solution u (after solve() or a batched solve)."""
        u = np.zeros(self.mesh.n)
        _raise(lib().svx_system_solution(self._h, _d(u)))
        return u

    def set_rhs(self, f):
        f = np.ascontiguousarray(f, np.float64)
        _raise(lib().svx_system_set_rhs(self._h, _d(f)))

    def apply_dirichlet(self, rows) -> None:
        rows = np.ascontiguousarray(rows, np.int64)
        _raise(lib().svx_apply_dirichlet(self._h, rows.shape[0], rows.ctypes.data_as(_i64p)))

    def solve(self, tol: float = 1e-10, max_iter: int = 0, precision: int = Precision.FP64,
              want_u: bool = True) -> Tuple[Optional[np.ndarray], int, float]:
        """pcg (solver.rs:66) -> (u, iterations, residual); raises NotConverged."""
        u = np.zeros(self.mesh.n) if want_u else None
        its = C.c_int64()
        res = C.c_double()
        err = _abi.SvxError()
        rc = lib().svx_solve(self._h, tol, max_iter, precision, _d(u) if want_u else None, C.byref(its),
                             C.byref(res), C.byref(err))
        self.iterations, self.residual = its.value, res.value
        _raise(rc, err)
        return u, its.value, res.value

    def stats(self) -> dict:
        """What the last solve did: iterations, FP64 refreshes and fallback
        iterations (FP32-IR), final residual and its FP64 evaluation floor."""
        st = _abi.SvxSolveStats()
        _raise(lib().svx_system_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _abi.SvxSolveStats._fields_}

    def recover(self, rows=()) -> Tuple[np.ndarray, np.ndarray, np.ndarray, float]:
        """(centroid von Mises, nodal von Mises, reactions at rows, cond_estimate)."""
        rows = np.ascontiguousarray(rows, np.int64)
        vm = np.zeros(self.mesh.n_elem)
        vn = np.zeros(self.mesh.n_nodes)
        re = np.zeros(rows.shape[0])
        cond = C.c_double()
        _raise(lib().svx_recover(self._h, _d(vm), _d(vn), rows.shape[0], rows.ctypes.data_as(_i64p), _d(re),
                                 C.byref(cond)))
        return vm, vn, re, cond.value

    def prepare_solver(self, precision: int):
        """The solver set-up of solve() without solving (svx_system_prepare_solver),
        so that time_spmv can time that precision's SpMV alone."""
        _raise(lib().svx_system_prepare_solver(self._h, precision))

    def time_spmv(self, precision: int, reps: int = 50) -> Tuple[float, float]:
        """(average ms, algorithmic bytes) of the solver's SpMV kernel."""
        ms = C.c_double()
        by = C.c_double()
        _raise(lib().svx_system_time_spmv(self._h, precision, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    @staticmethod
    def time_spmv_rotation(systems, precision: int, reps: int = 20) -> Tuple[float, float]:
        """(average ms, average algorithmic bytes) of one SpMV launch when the
        systems' SpMV kernels run in turn, reps rounds, on their shared
        context's stream: with matrices together larger than the Infinity
        Cache, each launch streams its matrix from HBM."""
        arr = (C.c_void_p * len(systems))(*[s._h for s in systems])
        ms = C.c_double()
        by = C.c_double()
        _raise(lib().svx_time_spmv_rotation(arr, len(systems), precision, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def time_phases(self, precision: int, reps: int = 20) -> Tuple[np.ndarray, np.ndarray]:
        """(average ms, algorithmic bytes) of [SpMV, update, direction update,
        FP32-IR refresh], each phase launched alone. Overwrites the solver
        state and the solution: solve again before reading them."""
        ms = np.zeros(4)
        by = np.zeros(4)
        _raise(lib().svx_system_time_phases(self._h, precision, reps, _d(ms), _d(by)))
        return ms, by

    def close(self):
        if self._h:
            lib().svx_system_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble(model: Model, ctx: Optional[Context] = None) -> Tuple[Mesh, System]:
    """assemble (assembly.rs:281) for a whole model: (mesh, system)."""
    mesh = Mesh.from_model(model, ctx)
    E, nu = element_materials(model)
    ndof = model.ndof()
    pos = {}
    for i, nid in enumerate(model.node_ids.tolist()):
        pos.setdefault(nid, i)
    rows = np.array([pos[int(n)] * ndof + int(d) - 1 for n, d in zip(model.cload_node, model.cload_dof)], np.int64)
    return mesh, mesh.assemble(E, nu, model.thickness, rows, model.cload_value)


class Batch:
    """Many systems solved together in one graph-captured launch sequence
    (svx_batch_*). Systems must be assembled and constrained."""

    def __init__(self, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self._h = C.c_void_p()
        _raise(lib().svx_batch_create(self.ctx.handle, C.byref(self._h)))
        self.systems: List[System] = []

    def add(self, system: System, tol: float = 1e-10, max_iter: int = 0) -> int:
        slot = C.c_int64()
        _raise(lib().svx_batch_add(self._h, system.handle, tol, max_iter, C.byref(slot)))
        self.systems.append(system)
        return slot.value

    def solve(self, precision: int = Precision.FP64, raise_on_failure: bool = True):
        err = _abi.SvxError()
        rc = lib().svx_batch_solve(self._h, precision, C.byref(err))
        if rc != _abi.SVX_OK and (raise_on_failure or rc != _abi.SVX_E_NOT_CONVERGED):
            _raise(rc, err)
        return [self.result(i) for i in range(len(self.systems))]

    def result(self, slot: int) -> Tuple[int, float, bool]:
        its = C.c_int64()
        res = C.c_double()
        ok = C.c_int32()
        _raise(lib().svx_batch_result(self._h, slot, C.byref(its), C.byref(res), C.byref(ok)))
        return its.value, res.value, bool(ok.value)

    def time_spmv(self, reps: int = 20) -> Tuple[float, float]:
        ms = C.c_double()
        by = C.c_double()
        _raise(lib().svx_batch_time_spmv(self._h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def prepare(self, precision: int = Precision.FP64):
        """The solver set-up of solve() without solving (svx_batch_prepare),
        so that time_spmv can time the batch's SpMV alone."""
        _raise(lib().svx_batch_prepare(self._h, precision))

    def autotune(self, precision: int = Precision.FP64, reps: int = 5) -> dict:
        """Pick the fastest SpMV launch shape for this batch (svx_batch_autotune);
        later solves with this precision use it."""
        unroll = np.zeros(4, np.int32)
        rows = np.zeros(4, np.int32)
        order = np.zeros(4, np.int32)
        ms = C.c_double()
        _raise(lib().svx_batch_autotune(self._h, precision, reps, unroll.ctypes.data_as(_i32p),
                                        rows.ctypes.data_as(_i32p), order.ctypes.data_as(_i32p), C.byref(ms)))
        return {"unroll": unroll.tolist(), "rows_per_tile": rows.tolist(), "tile_order": order.tolist(),
                "spmv_ms": ms.value}

    def close(self):
        if self._h:
            lib().svx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrMatrix:
    """A caller-provided CSR matrix on the device (Csr, assembly.rs:29-35)."""

    def __init__(self, row_ptr, col_idx, vals, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self._rp = np.ascontiguousarray(row_ptr, np.int32)
        self._ci = np.ascontiguousarray(col_idx, np.int32)
        self._v = np.ascontiguousarray(vals, np.float64)
        self.n = self._rp.shape[0] - 1
        self._h = C.c_void_p()
        _raise(lib().svx_csr_create(self.ctx.handle, self.n, self._rp.ctypes.data_as(_i32p),
                                    self._ci.ctypes.data_as(_i32p), _d(self._v), C.byref(self._h)))

    def mat_vec(self, x, ordered: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        _raise(lib().svx_csr_spmv(self._h, _d(x), _d(y), int(ordered)))
        return y

    def pcg(self, f, tol: float = 1e-10, max_iter: int = 0, precision: int = Precision.FP64):
        f = np.ascontiguousarray(f, np.float64)
        u = np.zeros(self.n)
        its = C.c_int64()
        res = C.c_double()
        err = _abi.SvxError()
        rc = lib().svx_csr_pcg(self._h, _d(f), tol, max_iter, precision, _d(u), C.byref(its), C.byref(res),
                               C.byref(err))
        _raise(rc, err)
        return u, its.value, res.value

    def pcg_noalloc(self, f, tol: float = 1e-10, max_iter: int = 0):
        """FP64 pcg without device allocation (svx_csr_pcg_noalloc): the
        library's fixed per-GPU scratch holds the solver state, so calls on one
        GPU must not overlap."""
        f = np.ascontiguousarray(f, np.float64)
        u = np.zeros(self.n)
        its = C.c_int64()
        res = C.c_double()
        err = _abi.SvxError()
        rc = lib().svx_csr_pcg_noalloc(self._h, _d(f), tol, max_iter, _d(u), C.byref(its), C.byref(res),
                                       C.byref(err))
        _raise(rc, err)
        return u, its.value, res.value

    def close(self):
        if self._h:
            lib().svx_csr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ request server
class Server:
    """Continuous-batching request server (svx_server_*): whole-model solves
    submitted at any time, admitted by device-memory footprint, solved
    together, each result equal to its solo solve_with_method result. The
    server owns its context: use a context of its own."""

    def __init__(self, ctx: Context, precision: int = Precision.FP32_IR, hbm_budget: int = 200 << 30,
                 max_active: int = 256, tol: float = 1e-10, chunk: int = 0, use_graph: bool = False):
        self.ctx = ctx
        cfg = _abi.SvxServerConfig(precision, max_active, hbm_budget, tol, chunk, int(use_graph))
        self._h = C.c_void_p()
        _raise(lib().svx_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)))
        self._pending = {}

    def submit(self, model: Model) -> int:
        desc, keep = model.to_desc()
        t = C.c_int64()
        _raise(lib().svx_server_submit(self._h, C.byref(desc), C.byref(t)))
        del keep  # the server copied the deck
        self._pending[t.value] = model
        return t.value

    def submit_deck(self, text: str) -> int:
        """Deck text (svx_server_submit_deck): a host worker of the server reads
        it; read errors come back from wait() as SolveError kind "Parse"."""
        raw = text.encode()
        t = C.c_int64()
        _raise(lib().svx_server_submit_deck(self._h, raw, len(raw), C.byref(t)))
        self._pending[t.value] = None
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> Tuple[Solution, dict]:
        """(Solution, request info); raises SolveError like solve_with_method
        (kind "Timeout" if the request is not done within timeout_ms)."""
        model = self._pending[ticket]
        if model is None:  # deck text: the server reports the result's sizes
            info = _abi.SvxRequestInfo()
            err = _abi.SvxError()
            rc = lib().svx_server_wait(self._h, ticket, timeout_ms, None, C.byref(info), C.byref(err))
            if rc not in (_abi.SVX_OK, _abi.SVX_E_TIMEOUT, _abi.SVX_E_ARGUMENT):
                del self._pending[ticket]
                lib().svx_server_wait(self._h, ticket, 0.0, C.byref(_abi.SvxSolution()), None, None)  # release
            _raise(rc, err)
            nn, ne, nr = int(info.n_nodes), int(info.n_elem), int(info.n_reactions)
        else:
            nn, ne = model.n_nodes, model.n_elem
            nr = max(0, int(model.reaction_count()))
        disp = np.zeros((nn, 3))
        vm = np.zeros(ne)
        nvm = np.zeros(nn)
        rn = np.zeros(nr, np.uint32)
        rd = np.zeros(nr, np.int32)
        rv = np.zeros(nr)
        sol = _abi.SvxSolution(_d(disp), _d(vm), _d(nvm), rn.ctypes.data_as(C.POINTER(C.c_uint32)),
                               rd.ctypes.data_as(_i32p), _d(rv), nr)
        info = _abi.SvxRequestInfo()
        err = _abi.SvxError()
        rc = lib().svx_server_wait(self._h, ticket, timeout_ms, C.byref(sol), C.byref(info), C.byref(err))
        if rc not in (_abi.SVX_E_TIMEOUT, _abi.SVX_E_ARGUMENT):
            del self._pending[ticket]
        _raise(rc, err)
        out = Solution(
            displacements=disp, von_mises=vm, nodal_von_mises=nvm,
            reactions=[Reaction(int(a), int(b), float(c)) for a, b, c in zip(rn, rd, rv)],
            iterations=int(sol.iterations), final_residual=float(sol.final_residual), solver_method="pcg",
            cond_estimate=float(sol.cond_estimate), timings_ms=dict(solve=sol.solve_ms))
        return out, {k: getattr(info, k) for k, _ in _abi.SvxRequestInfo._fields_}

    def stats(self) -> dict:
        st = _abi.SvxServerSummary()
        _raise(lib().svx_server_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _abi.SvxServerSummary._fields_}

    def close(self):
        if self._h:
            lib().svx_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

