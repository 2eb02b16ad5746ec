"""Host-side model, solution and error types of the linear-static path.

Mirrors the public contract of solvix-core (crates/solvix-core/src/lib.rs):
`Model` (lib.rs:39-330, the fields the linear path reads), `Solution`
(lib.rs:486-513), `SolveError` (lib.rs:515-521) and `Method` (lib.rs:523-533).
The model is held as flat numpy arrays (structure of arrays) so that meshes of
hundreds of thousands of elements cross the C ABI without per-node objects.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi


class SolveError(Exception):
    """SolveError (lib.rs:515-521). `kind` is one of EmptyMesh, NoConstraints,
    InvalidModel, NotConverged; NotConverged carries iterations/residual."""

    def __init__(self, kind: str, message: str = "", iterations: int = 0,
                 residual: float = 0.0, element_index: int = -1, reason: int = 0):
        super().__init__(f"{kind}: {message}" if message else kind)
        self.kind = kind
        self.message = message
        self.iterations = iterations
        self.residual = residual
        self.element_index = element_index
        self.reason = reason


_KIND_OF_CODE = {
    _abi.SVX_E_EMPTY_MESH: "EmptyMesh",
    _abi.SVX_E_NO_CONSTRAINTS: "NoConstraints",
    _abi.SVX_E_INVALID_MODEL: "InvalidModel",
    _abi.SVX_E_NOT_CONVERGED: "NotConverged",
    _abi.SVX_E_DEVICE: "Device",
    _abi.SVX_E_ARGUMENT: "Argument",
    _abi.SVX_E_CAPACITY: "Capacity",
    _abi.SVX_E_TIMEOUT: "Timeout",
    _abi.SVX_E_PARSE: "Parse",
}


def error_from_code(code: int, message: str = "", iterations: int = 0,
                    residual: float = 0.0, index: int = -1, reason: int = 0) -> SolveError:
    return SolveError(_KIND_OF_CODE.get(code, f"code{code}"), message, iterations,
                      residual, index, reason)


class Method:
    AUTO = _abi.SVX_METHOD_AUTO
    PCG = _abi.SVX_METHOD_PCG
    LDLT = _abi.SVX_METHOD_LDLT


class Precision:
    FP64 = _abi.SVX_FP64
    FP32_IR = _abi.SVX_FP32_IR


@dataclass
class Reaction:
    node_id: int
    dof: int
    value: float


@dataclass
class Solution:
    """Solution (lib.rs:486-513); arrays parallel to the model's nodes/elements."""
    displacements: np.ndarray          # [n_nodes, 3]
    von_mises: np.ndarray              # [n_elem]
    nodal_von_mises: np.ndarray        # [n_nodes]
    reactions: List[Reaction]
    iterations: int
    final_residual: float
    solver_method: str
    cond_estimate: float
    timings_ms: dict = field(default_factory=dict)

    def max_disp(self) -> float:
        """max over nodes of |u| (crates/solvix-cli/src/main.rs:386-393)."""
        d = self.displacements
        return float(np.sqrt((d * d).sum(axis=1)).max(initial=0.0))

    def max_von_mises(self) -> float:
        return float(np.max(self.von_mises, initial=0.0))


@dataclass
class Model:
    """Flat Model. Ids are deck ids; connectivity holds node ids (lib.rs:88-95).

    bcs: list of (node_ids, dofs); cloads: arrays (node_id, dof, value) —
    the effective LoadState of the last step (lib.rs:388-403)."""
    node_ids: np.ndarray                       # uint32 [n_nodes]
    coords: np.ndarray                         # float64 [n_nodes, 3]
    elem_ids: np.ndarray                       # uint32 [n_elem]
    kinds: np.ndarray                          # uint8 [n_elem]
    conn_ptr: np.ndarray                       # int64 [n_elem + 1]
    conn_ids: np.ndarray                       # uint32
    materials: List[Tuple[str, float, float]]  # (name, E, nu)
    elsets: List[Tuple[str, np.ndarray]]       # (name, uint32 element ids)
    assignments: List[Tuple[str, str]]         # (elset name, material name)
    thickness: float = 1.0
    bcs: List[Tuple[np.ndarray, np.ndarray]] = field(default_factory=list)
    cload_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    cload_dof: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    cload_value: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    name: str = "model"

    @property
    def n_nodes(self) -> int:
        return int(self.node_ids.shape[0])

    @property
    def n_elem(self) -> int:
        return int(self.elem_ids.shape[0])

    def ndof(self) -> int:
        """ndof_per_node (lib.rs:725-752) for the CPS3/CPS4/C3D8 subset."""
        return 3 if np.any(self.kinds == _abi.SVX_C3D8) else 2

    def n_dofs(self) -> int:
        return self.n_nodes * self.ndof()

    def reaction_count(self) -> int:
        return int(sum(len(n) * len(d) for n, d in self.bcs))

    # ---------------------------------------------------------------- C ABI
    def to_desc(self):
        """Build an svx_model_desc; returns (desc, keepalive)."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        node_ids = arr(self.node_ids, np.uint32)
        coords = arr(self.coords.reshape(-1, 3), np.float64)
        elem_ids = arr(self.elem_ids, np.uint32)
        kinds = arr(self.kinds, np.uint8)
        conn_ptr = arr(self.conn_ptr, np.int64)
        conn_ids = arr(self.conn_ids, np.uint32)
        youngs = arr([m[1] for m in self.materials], np.float64)
        poisson = arr([m[2] for m in self.materials], np.float64)
        es_ptr = arr(np.concatenate([[0], np.cumsum([len(s[1]) for s in self.elsets])]), np.int64)
        es_mem = arr(np.concatenate([s[1] for s in self.elsets]) if self.elsets else [], np.uint32)
        es_names = [s[0] for s in self.elsets]
        mat_names = [m[0] for m in self.materials]

        def first(names, key):
            return names.index(key) if key in names else -1

        a_es = arr([first(es_names, a[0]) for a in self.assignments], np.int32)
        a_mat = arr([first(mat_names, a[1]) for a in self.assignments], np.int32)
        bc_nptr = arr(np.concatenate([[0], np.cumsum([len(b[0]) for b in self.bcs])]), np.int64)
        bc_nodes = arr(np.concatenate([b[0] for b in self.bcs]) if self.bcs else [], np.uint32)
        bc_dptr = arr(np.concatenate([[0], np.cumsum([len(b[1]) for b in self.bcs])]), np.int64)
        bc_dofs = arr(np.concatenate([b[1] for b in self.bcs]) if self.bcs else [], np.int32)
        cl_node = arr(self.cload_node, np.uint32)
        cl_dof = arr(self.cload_dof, np.int32)
        cl_val = arr(self.cload_value, np.float64)
        P = _abi.ptr
        d = _abi.SvxModelDesc(
            self.n_nodes, P(node_ids, C.c_uint32), P(coords, C.c_double),
            self.n_elem, P(elem_ids, C.c_uint32), P(kinds, C.c_uint8),
            P(conn_ptr, C.c_int64), P(conn_ids, C.c_uint32),
            len(self.materials), P(youngs, C.c_double), P(poisson, C.c_double),
            len(self.elsets), P(es_ptr, C.c_int64), P(es_mem, C.c_uint32),
            len(self.assignments), P(a_es, C.c_int32), P(a_mat, C.c_int32),
            float(self.thickness),
            len(self.bcs), P(bc_nptr, C.c_int64), P(bc_nodes, C.c_uint32),
            P(bc_dptr, C.c_int64), P(bc_dofs, C.c_int32),
            int(cl_node.shape[0]), P(cl_node, C.c_uint32), P(cl_dof, C.c_int32),
            P(cl_val, C.c_double),
        )
        return d, keep

    @classmethod
    def from_desc(cls, d, name: str = "model") -> "Model":
        """A Model holding copies of an svx_model_desc's arrays. The
        descriptor carries no names: element sets and materials are named by
        position ("ES0", "MAT0", ...), and an assignment index of -1 by a name
        that resolves to nothing, so resolution is unchanged."""
        def get(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(int(n),)).astype(dt, copy=True) if n > 0 else np.zeros(0, dt)
        n_nodes, n_elem = int(d.n_nodes), int(d.n_elem)
        conn_ptr = get(d.conn_ptr, n_elem + 1, np.int64)
        es_ptr = get(d.elset_ptr, d.n_elsets + 1, np.int64)
        es_mem = get(d.elset_members, es_ptr[-1] if es_ptr.size else 0, np.uint32)
        E, nu = get(d.youngs, d.n_materials, np.float64), get(d.poisson, d.n_materials, np.float64)
        a_es, a_mat = get(d.assign_elset, d.n_assign, np.int32), get(d.assign_material, d.n_assign, np.int32)
        bn = get(d.bc_node_ptr, d.n_bcs + 1, np.int64)
        bd = get(d.bc_dof_ptr, d.n_bcs + 1, np.int64)
        bc_nodes = get(d.bc_nodes, bn[-1] if bn.size else 0, np.uint32)
        bc_dofs = get(d.bc_dofs, bd[-1] if bd.size else 0, np.int32)
        return cls(
            node_ids=get(d.node_ids, n_nodes, np.uint32),
            coords=get(d.coords, 3 * n_nodes, np.float64).reshape(-1, 3),
            elem_ids=get(d.elem_ids, n_elem, np.uint32),
            kinds=get(d.kinds, n_elem, np.uint8),
            conn_ptr=conn_ptr,
            conn_ids=get(d.conn_ids, conn_ptr[-1] if conn_ptr.size else 0, np.uint32),
            materials=[(f"MAT{k}", float(E[k]), float(nu[k])) for k in range(E.size)],
            elsets=[(f"ES{k}", es_mem[es_ptr[k]:es_ptr[k + 1]]) for k in range(es_ptr.size - 1)],
            assignments=[(f"ES{e}" if e >= 0 else "?", f"MAT{m}" if m >= 0 else "?") for e, m in zip(a_es, a_mat)],
            thickness=float(d.thickness),
            bcs=[(bc_nodes[bn[k]:bn[k + 1]], bc_dofs[bd[k]:bd[k + 1]]) for k in range(bn.size - 1)],
            cload_node=get(d.cload_node, d.n_cloads, np.uint32),
            cload_dof=get(d.cload_dof, d.n_cloads, np.int32),
            cload_value=get(d.cload_value, d.n_cloads, np.float64),
            name=name,
        )


def simple_model(node_ids, coords, elem_ids, kinds, conn, E, nu, thickness=1.0,
                 bcs=(), cloads=(), name="model") -> Model:
    """One material on one elset "ALL" covering every element.

    conn: list of per-element node-id sequences; cloads: (node, dof, value)."""
    conn_ptr = np.zeros(len(conn) + 1, np.int64)
    conn_ptr[1:] = np.cumsum([len(c) for c in conn])
    conn_ids = np.concatenate([np.asarray(c, np.uint32) for c in conn]) if conn else np.zeros(0, np.uint32)
    cl = list(cloads)
    return Model(
        node_ids=np.asarray(node_ids, np.uint32),
        coords=np.asarray(coords, np.float64).reshape(-1, 3),
        elem_ids=np.asarray(elem_ids, np.uint32),
        kinds=np.asarray(kinds, np.uint8),
        conn_ptr=conn_ptr, conn_ids=conn_ids,
        materials=[("MAT", float(E), float(nu))],
        elsets=[("ALL", np.asarray(elem_ids, np.uint32))],
        assignments=[("ALL", "MAT")],
        thickness=thickness,
        bcs=[(np.asarray(n, np.uint32), np.asarray(d, np.int32)) for n, d in bcs],
        cload_node=np.asarray([c[0] for c in cl], np.uint32),
        cload_dof=np.asarray([c[1] for c in cl], np.int32),
        cload_value=np.asarray([c[2] for c in cl], np.float64),
        name=name,
    )
