"""ctypes mirror of include/svx.h (structs, enums) shared by every binding."""
import ctypes as C

SVX_OK = 0
SVX_E_EMPTY_MESH = 1
SVX_E_NO_CONSTRAINTS = 2
SVX_E_INVALID_MODEL = 3
SVX_E_NOT_CONVERGED = 4
SVX_E_DEVICE = 5
SVX_E_ARGUMENT = 6
SVX_E_CAPACITY = 7
SVX_E_TIMEOUT = 8
SVX_E_PARSE = 9

SVX_CPS3, SVX_CPS4, SVX_C3D8 = 0, 1, 2
KIND_NAMES = {"CPS3": SVX_CPS3, "CPS4": SVX_CPS4, "C3D8": SVX_C3D8}
NODES_PER_KIND = {SVX_CPS3: 3, SVX_CPS4: 4, SVX_C3D8: 8}

SVX_ELEM_OK = 0
SVX_ELEM_NO_MATERIAL = 1
SVX_ELEM_BAD_MATERIAL = 2
SVX_ELEM_BAD_POISSON_SOLID = 3
SVX_ELEM_DEGENERATE_TRIANGLE = 4
SVX_ELEM_NONPOSITIVE_JACOBIAN = 5

SVX_METHOD_AUTO, SVX_METHOD_PCG, SVX_METHOD_LDLT = 0, 1, 2
SVX_FP64, SVX_FP32_IR = 0, 1
SVX_OP_ASSEMBLED, SVX_OP_MATRIX_FREE = 0, 1
SVX_STOP_NONE, SVX_STOP_RECURSIVE, SVX_STOP_TRUE = 0, 1, 2
SVX_TILES_LINEAR, SVX_TILES_XCD = 0, 1
SVX_LOADS_AUTO, SVX_LOADS_CACHED, SVX_LOADS_STREAMING = 0, 1, 2

_p = C.POINTER


class SvxError(C.Structure):
    _fields_ = [
        ("code", C.c_int32),
        ("reason", C.c_int32),
        ("index", C.c_int64),
        ("iterations", C.c_int64),
        ("residual", C.c_double),
        ("message", C.c_char * 256),
    ]


class SvxModelDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64),
        ("node_ids", _p(C.c_uint32)),
        ("coords", _p(C.c_double)),
        ("n_elem", C.c_int64),
        ("elem_ids", _p(C.c_uint32)),
        ("kinds", _p(C.c_uint8)),
        ("conn_ptr", _p(C.c_int64)),
        ("conn_ids", _p(C.c_uint32)),
        ("n_materials", C.c_int32),
        ("youngs", _p(C.c_double)),
        ("poisson", _p(C.c_double)),
        ("n_elsets", C.c_int32),
        ("elset_ptr", _p(C.c_int64)),
        ("elset_members", _p(C.c_uint32)),
        ("n_assign", C.c_int32),
        ("assign_elset", _p(C.c_int32)),
        ("assign_material", _p(C.c_int32)),
        ("thickness", C.c_double),
        ("n_bcs", C.c_int32),
        ("bc_node_ptr", _p(C.c_int64)),
        ("bc_nodes", _p(C.c_uint32)),
        ("bc_dof_ptr", _p(C.c_int64)),
        ("bc_dofs", _p(C.c_int32)),
        ("n_cloads", C.c_int64),
        ("cload_node", _p(C.c_uint32)),
        ("cload_dof", _p(C.c_int32)),
        ("cload_value", _p(C.c_double)),
    ]


class SvxServerConfig(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("max_active", C.c_int32),
        ("hbm_budget", C.c_int64),
        ("tol", C.c_double),
        ("chunk", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class SvxRequestInfo(C.Structure):
    _fields_ = [
        ("latency_ms", C.c_double),
        ("queue_ms", C.c_double),
        ("prepare_ms", C.c_double),
        ("iterations", C.c_int64),
        ("refreshes", C.c_int64),
        ("fp64_iterations", C.c_int64),
        ("device_bytes", C.c_int64),
        ("n_dof", C.c_int64),
        ("n_nodes", C.c_int64),
        ("n_elem", C.c_int64),
        ("n_reactions", C.c_int64),
    ]


class SvxServerSummary(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("completed", C.c_int64),
        ("failed", C.c_int64),
        ("queued", C.c_int64),
        ("resident_bytes", C.c_int64),
        ("peak_resident_bytes", C.c_int64),
        ("peak_active", C.c_int64),
        ("solved_dof", C.c_int64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("dof_per_s", C.c_double),
        ("admit_ms", C.c_double),
        ("solve_ms", C.c_double),
        ("retire_ms", C.c_double),
        ("idle_ms", C.c_double),
        ("chunks", C.c_int64),
        ("retiles", C.c_int64),
    ]


class SvxSolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("refreshes", C.c_int64),
        ("fp64_iterations", C.c_int64),
        ("residual_evals", C.c_int64),
        ("residual", C.c_double),
        ("floor", C.c_double),
        ("status", C.c_int32),
        ("precision", C.c_int32),
        ("true_residual", C.c_double),
        ("true_floor", C.c_double),
        ("converged_by", C.c_int32),
        ("fallback_reason", C.c_int32),
    ]


class SvxSolution(C.Structure):
    _fields_ = [
        ("displacements", _p(C.c_double)),
        ("von_mises", _p(C.c_double)),
        ("nodal_von_mises", _p(C.c_double)),
        ("reaction_node", _p(C.c_uint32)),
        ("reaction_dof", _p(C.c_int32)),
        ("reaction_value", _p(C.c_double)),
        ("reaction_capacity", C.c_int64),
        ("iterations", C.c_int64),
        ("final_residual", C.c_double),
        ("cond_estimate", C.c_double),
        ("solver_method", C.c_int32),
        ("ndof", C.c_int32),
        ("assemble_ms", C.c_double),
        ("solve_ms", C.c_double),
        ("recover_ms", C.c_double),
    ]


def ptr(a, ctype):
    """Pointer to a contiguous numpy array's data (None for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_p(ctype))
