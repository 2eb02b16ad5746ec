/*
 * svx.h — C ABI of the MI355X (gfx950) linear-static hot path of Solvix:
 * global stiffness assembly (CPS3 / CPS4 / C3D8 into CSR), homogeneous
 * Dirichlet constraints, Jacobi-preconditioned conjugate gradient, stress
 * recovery and reactions.
 *
 * Every entry point is plain C: opaque handles, caller-owned host buffers,
 * library-owned device buffers behind the handles, integer status codes.
 * No C++ exception crosses this boundary. One svx_ctx per host thread.
 *
 * Reference seam (typezo/Solvix, crates/solvix-core/src):
 *   lib.rs:586        solve_with_method(&Model, Method) -> Solution
 *   lib.rs:593-716    solve_with_thermal (the pipeline these calls replace)
 *   assembly.rs:281   assemble(model, ndof, loads) -> (Csr, f, Vec<ElemInfo>)
 *   assembly.rs:38-72 Csr::mat_vec / mat_vec_with
 *   solver.rs:15      constrained_dofs
 *   solver.rs:45      apply_dirichlet
 *   solver.rs:66      pcg(k, f, tol, max_iter) -> (u, iterations, residual)
 *   stress.rs:22,53,103  von_mises_at_centroids / nodal_von_mises / reactions
 */
#ifndef SVX_H
#define SVX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Status codes. SVX_E_EMPTY_MESH .. SVX_E_NOT_CONVERGED mirror the variants of
 * SolveError (lib.rs:515-521); the rest are boundary/runtime failures.
 * ------------------------------------------------------------------------- */
enum {
    SVX_OK = 0,
    SVX_E_EMPTY_MESH = 1,
    SVX_E_NO_CONSTRAINTS = 2,
    SVX_E_INVALID_MODEL = 3,
    SVX_E_NOT_CONVERGED = 4,
    SVX_E_DEVICE = 5,   /* HIP runtime failure; message holds the HIP string */
    SVX_E_ARGUMENT = 6, /* bad handle, size, option or index */
    SVX_E_CAPACITY = 7, /* batch / server budget or slot limit exceeded */
    SVX_E_TIMEOUT = 8,  /* svx_server_wait: the request is not done yet */
    SVX_E_PARSE = 9     /* deck text the .inp reader rejects: message "line N: ...",
                           svx_error.index = N */
};

/* Element kinds on the path (lib.rs:53-73, subset CPS3 / CPS4 / C3D8). */
enum { SVX_CPS3 = 0, SVX_CPS4 = 1, SVX_C3D8 = 2 };

/* Element-level reasons for SVX_E_INVALID_MODEL, one per message of
 * assembly.rs:145-175,194-196,259-261. The lowest failing element index wins
 * (assembly.rs:354-362). */
enum {
    SVX_ELEM_OK = 0,
    SVX_ELEM_NO_MATERIAL = 1,           /* "no material assigned" */
    SVX_ELEM_BAD_MATERIAL = 2,          /* E <= 0 or |nu| >= 1 */
    SVX_ELEM_BAD_POISSON_SOLID = 3,     /* nu >= 0.5 on a solid */
    SVX_ELEM_DEGENERATE_TRIANGLE = 4,   /* CPS3 with zero signed area */
    SVX_ELEM_NONPOSITIVE_JACOBIAN = 5   /* det J <= 0 at a Gauss point */
};

/* Linear solver selection (lib.rs:523-533). LDLT is not on the device path:
 * SVX_METHOD_LDLT returns SVX_E_ARGUMENT and SVX_METHOD_AUTO reports
 * SVX_E_NOT_CONVERGED where the reference would fall back to LDLT. */
enum { SVX_METHOD_AUTO = 0, SVX_METHOD_PCG = 1, SVX_METHOD_LDLT = 2 };

/* Arithmetic of the Krylov iterations.
 * SVX_FP64:    plain FP64 Jacobi-PCG, the recurrence of solver.rs:82-149.
 * SVX_FP32_IR: FP32 matrix/vectors inside the Krylov loop (dot products
 *              accumulated in FP64) with FP64 iterative refinement: the true
 *              residual f - K u, the update of u and the convergence test run
 *              in FP64 on the FP64 matrix. */
enum { SVX_FP64 = 0, SVX_FP32_IR = 1 };

/* Operator of the FP64 Krylov iterations (svx_set_operator).
 * SVX_OP_ASSEMBLED:   the assembled K (Csr::mat_vec, assembly.rs:38-72); the
 *                     drop-in, whose products sum in the reference's order.
 * SVX_OP_MATRIX_FREE: systems of one material on meshes of one element kind,
 *                     plane (CPS3 or CPS4, at most 8 / 4 elements per node)
 *                     or solid (C3D8), multiply by K without its values: K p
 *                     is evaluated element by element from the coordinates
 *                     with the formulas of element.rs (91-211 plane, 427-561
 *                     solid, the reference's J^-1 sign pattern included), the
 *                     same product up to rounding, so iterates differ from the
 *                     assembled ones in the last bits. It serves FP64 solves
 *                     and the FP64 phase (finish, fallback) of SVX_FP32_IR
 *                     solves; other systems, the FP32 inner iterations and
 *                     residual evaluations keep the assembled K. */
enum { SVX_OP_ASSEMBLED = 0, SVX_OP_MATRIX_FREE = 1 };

/* Order in which the workgroups of the gathering kernels (SpMV, FP64
 * residual, matrix-free product) take their tiles (svx_set_tile_order).
 * SVX_TILES_XCD:    each of the eight XCDs takes a contiguous run of tiles, so
 *                   tiles whose rows gather each other's vectors share an L2
 *                   (workgroups are dealt to the XCDs round robin);
 * SVX_TILES_LINEAR: tile = workgroup index.
 * Either order gives the same results bit for bit. */
enum { SVX_TILES_LINEAR = 0, SVX_TILES_XCD = 1 };

/* How the assembled ELL SpMVs (FP64, and FP32 in FP32-IR) load the matrix
 * values (svx_set_load_hint).
 * SVX_LOADS_STREAMING: nontemporal loads; the values, read once per product,
 *                      do not evict the gathered vectors from the caches;
 * SVX_LOADS_CACHED:    ordinary loads, so a matrix that fits the 256 MB
 *                      last-level cache is read from there on the next product;
 * SVX_LOADS_AUTO:      streaming when the matrix values one launch reads in
 *                      its precision (all its instances of one node-block
 *                      size) exceed 256 MiB, else cached.
 * The hint also covers the vectors of the PCG updates (x, r, z written; p,
 * q, the diagonal, x, r read): STREAMING and CACHED set them alike, AUTO
 * streams them once the FP64 update's vectors (64 bytes per row over the
 * launch's instances of one node-block size) exceed 256 MiB.
 * Every hint gives the same results bit for bit. */
enum { SVX_LOADS_AUTO = 0, SVX_LOADS_CACHED = 1, SVX_LOADS_STREAMING = 2 };

typedef struct svx_error {
    int32_t code;       /* SVX_* status */
    int32_t reason;     /* SVX_ELEM_* for element errors, else 0 */
    int64_t index;      /* element index (position) for element errors, else -1 */
    int64_t iterations; /* SVX_E_NOT_CONVERGED: iterations done */
    double residual;    /* SVX_E_NOT_CONVERGED: relative residual reached */
    char message[256];  /* human-readable text in the reference's wording */
} svx_error;

/* ---------------------------------------------------------------------------
 * Flat model descriptor: the fields of Model (lib.rs:39-330) that the linear
 * path reads, as parallel arrays. Ids are the deck's ids; positions are
 * 0-based indices into the node / element arrays.
 * ------------------------------------------------------------------------- */
typedef struct svx_model_desc {
    int64_t n_nodes;
    const uint32_t* node_ids; /* [n_nodes] */
    const double* coords;     /* [n_nodes * 3], x y z per node */

    int64_t n_elem;
    const uint32_t* elem_ids; /* [n_elem] */
    const uint8_t* kinds;     /* [n_elem] SVX_CPS3 / SVX_CPS4 / SVX_C3D8 */
    const int64_t* conn_ptr;  /* [n_elem + 1] */
    const uint32_t* conn_ids; /* [conn_ptr[n_elem]] node ids */

    int32_t n_materials;
    const double* youngs;  /* [n_materials] */
    const double* poisson; /* [n_materials] */

    int32_t n_elsets;
    const int64_t* elset_ptr;      /* [n_elsets + 1] */
    const uint32_t* elset_members; /* element ids, duplicates allowed */

    /* (elset, material) assignments in deck order (Model::mat_assignments);
     * an index of -1 stands for a name that resolves to nothing. */
    int32_t n_assign;
    const int32_t* assign_elset;
    const int32_t* assign_material;

    double thickness; /* model-global plane thickness */

    /* Effective boundary conditions (LoadState::bcs): each one constrains
     * every listed dof (1-based) of every listed node. */
    int32_t n_bcs;
    const int64_t* bc_node_ptr; /* [n_bcs + 1] */
    const uint32_t* bc_nodes;
    const int64_t* bc_dof_ptr; /* [n_bcs + 1] */
    const int32_t* bc_dofs;

    /* Effective concentrated loads (LoadState::cloads). */
    int64_t n_cloads;
    const uint32_t* cload_node;
    const int32_t* cload_dof; /* 1-based */
    const double* cload_value;
} svx_model_desc;

/* Caller-owned output buffers for svx_solve_model (Solution, lib.rs:486-513).
 * Array lengths: displacements n_nodes*3, von_mises n_elem, nodal_von_mises
 * n_nodes, reactions svx_model_reaction_count(). A NULL array is skipped.
 * reaction_capacity is the number of entries each non-NULL reaction array
 * holds; a model with more reactions than that is refused with
 * SVX_E_ARGUMENT before anything is written. */
typedef struct svx_solution {
    double* displacements;
    double* von_mises;
    double* nodal_von_mises;
    uint32_t* reaction_node;
    int32_t* reaction_dof;
    double* reaction_value;
    int64_t reaction_capacity;
    int64_t iterations;
    double final_residual;
    double cond_estimate;
    int32_t solver_method; /* 0 = "pcg" */
    int32_t ndof;          /* 2 for plane models, 3 for solid models */
    double assemble_ms;    /* device time of assembly (pattern + values + f) */
    double solve_ms;       /* Dirichlet + Krylov solve */
    double recover_ms;     /* stress recovery + reactions */
} svx_solution;

typedef struct svx_ctx svx_ctx;
typedef struct svx_deck svx_deck;
typedef struct svx_mesh svx_mesh;
typedef struct svx_system svx_system;
typedef struct svx_csr svx_csr;
typedef struct svx_batch svx_batch;
typedef struct svx_server svx_server;

/* ---------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------- */
int svx_ctx_create(int device, svx_ctx** out);
void svx_ctx_destroy(svx_ctx* ctx);
/* Library version string and the device's name; never NULL. */
const char* svx_version(void);
int svx_device_name(svx_ctx* ctx, char* buf, size_t len);
/* Message of the last failure on this host thread (also for calls that
 * take no svx_error). No reference counterpart. */
const char* svx_last_error(void);
/* Launch-shape and refinement knobs of the Krylov engine for this context
 * (0 / <= 0 keeps the automatic choice): lanes per CSR row (raw CSR systems,
 * and the CSR passes of node-block systems: FP64 residual, warm start),
 * blocks in flight per lane in the SpMV of node-block systems (their matrix
 * is multiplied in a node-row sliced ELL layout, 64 node rows per
 * wavefront), rows per workgroup tile (node-block systems: rounded up to
 * whole 64-node-row chunks), FP32-IR refresh period, iterations per captured
 * graph, graph capture on/off, and the FP32-IR refresh threshold (replace the
 * residual when ||r|| < ru_delta * max ||r|| since the last refresh). Lane
 * counts must be 0 or a power of two from 2 to 64, unroll 0..4, and
 * 0 <= ru_delta < 1; anything else returns SVX_E_ARGUMENT. No reference
 * counterpart: the autotuner drives it. */
int svx_set_tuning(svx_ctx* ctx, int lanes, int unroll, int rows_per_tile,
                   int period, int chunk, int use_graph, double ru_delta);
/* Operator of this context's subsequent FP64 iterations (svx_solve,
 * svx_batch_solve, and the engines of a server created on the context
 * afterwards; for SVX_FP32_IR, its FP64 phase): SVX_OP_ASSEMBLED (default) or
 * SVX_OP_MATRIX_FREE. No reference counterpart: the reference multiplies its
 * assembled CSR only. */
int svx_set_operator(svx_ctx* ctx, int op);
/* Tile order of this context's subsequent solves (default SVX_TILES_XCD).
 * No reference counterpart: a launch-shape knob the autotuner also sets. */
int svx_set_tile_order(svx_ctx* ctx, int order);
/* Load hint of this context's subsequent solves' assembled SpMVs (default
 * SVX_LOADS_AUTO). No reference counterpart: a launch knob. */
int svx_set_load_hint(svx_ctx* ctx, int hint);
/* Assembled 2D / 3D systems whose FP64 solve has more row tiles than this
 * take the pending direction step of each iteration (r.r and r.z over all
 * their tiles, the convergence and budget tests of solver.rs:129-148, beta)
 * in one workgroup per system, where smaller ones take it in every SpMV tile
 * (whose reduction grows with the square of the tile count). Systems with
 * more than four times this many p.q partials (row tiles, or matrix-free
 * patches) also take alpha and the breakdown test of solver.rs:116-123 in
 * one workgroup per system instead of in every update tile. 0: the default,
 * 1024; negative: never. Either way gives the same results bit for bit. No
 * reference counterpart: a launch knob. */
int svx_set_step_kernel_tiles(svx_ctx* ctx, int32_t tiles);
/* Node groups of the FP64 multi-RHS block SpMV (svx_solve_multi): one
 * wavefront multiplies the block rows of `nodes` consecutive nodes in Morton
 * order (4 or 8; 0: the default, 4), gathering each column of their union
 * once; negative: no groups, the gathering SpMV that loads each column once
 * per block referencing it. Every setting adds each row's entries in its
 * column order. No reference counterpart: a launch knob. */
int svx_set_multi_group(svx_ctx* ctx, int32_t nodes);
/* Streams of a batch solve (svx_batch_solve): the instances are split over
 * `streams` engines (1..8; 0: the default, 2), each on its own stream, and one
 * chunk of iterations is enqueued on every stream before any is polled, so
 * the engines' kernels run side by side (one set's latency-bound products
 * beside another's bandwidth-bound updates). Every instance runs the same
 * operations as with one engine: the same bits. No reference counterpart:
 * a launch knob. */
int svx_set_batch_streams(svx_ctx* ctx, int32_t streams);
/* SVX_OP_MATRIX_FREE once a matrix-free solve has prepared the system's
 * operator (and until its constraints change), else SVX_OP_ASSEMBLED; -1 for
 * NULL. No reference counterpart. */
int svx_system_operator(const svx_system* sys);
/* Host-only check of the plane matrix-free operator's node patches for a
 * mesh (no device use; solid meshes report kind 0): plans the patches, checks their limits and that every node
 * has exactly one owner, and evaluates K x patch by patch as the kernel
 * does against the element-by-element product (seeded x, unit material).
 * stats[5] = {kind (4 CPS4, 3 CPS3, 0 none), patches, largest patch nodes,
 * largest patch element slots, element evaluations per product};
 * *max_rel_diff = max |difference| / max |K x| (unset when kind is 0, the
 * reason in svx_last_error). SVX_E_INVALID_MODEL if a check fails. No
 * reference counterpart: a test and diagnostics entry. */
int svx_mf_plan_check(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem,
                      const uint8_t* kinds, const int64_t* conn_ptr, const int32_t* conn_idx,
                      int64_t* stats, double* max_rel_diff);

/* Host-only check of the multi-RHS block SpMV's node groups for a mesh (no
 * device use): plans the groups (`group` 4 or 8; 0 for the default), checks
 * that every node is in exactly one group and every CSR entry is packed
 * exactly once, and evaluates the product of a seeded matrix on the mesh's
 * pattern with 3 seeded columns group by group, indexed as the kernel indexes
 * it, against the plain CSR product. stats[3] = {groups, union columns over
 * all groups, node blocks}; *max_rel_diff = max |difference| / max |product|
 * (unset, and the reason in svx_last_error, when the mesh has no plan:
 * groups 0). SVX_E_INVALID_MODEL if a check fails. No reference counterpart:
 * a test and diagnostics entry. */
int svx_blk_plan_check(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem,
                       const int64_t* conn_ptr, const int32_t* conn_idx, int32_t group,
                       int64_t* stats, double* max_rel_diff);

/* ---------------------------------------------------------------------------
 * Deck text — replaces solvix_io::parse_inp (crates/solvix-io/src/parser.rs:87)
 * for the keywords the linear path's decks use: *NODE (146-173), *ELEMENT
 * TYPE=CPS3|CPS4|C3D8 [ELSET=] with comma continuation (175-256), *NSET /
 * *ELSET [GENERATE] (258-312, names upper-cased, duplicates kept), *MATERIAL
 * + *ELASTIC (496-535), *SOLID SECTION (537-584, model-global thickness),
 * *STEP [OP=NEW] / *END STEP (1114-1150), *BOUNDARY node|set, dof[, dof2]
 * (742-796), *CLOAD node|set, dof, value (798-844); the last step's
 * constraints and loads, else those outside any step (1325-1336). Other
 * keywords are skipped. Errors: SVX_E_PARSE.
 * ------------------------------------------------------------------------- */
int svx_deck_parse(const char* text, size_t len, svx_deck** out, svx_error* err);
/* The parsed model; valid until svx_deck_destroy. */
const svx_model_desc* svx_deck_model(const svx_deck* deck);
void svx_deck_destroy(svx_deck* deck);

/* ---------------------------------------------------------------------------
 * Whole pipeline — replaces solve_with_method(&Model, Method) (lib.rs:586),
 * i.e. solve_with_thermal without thermal loads (lib.rs:593-716).
 * ------------------------------------------------------------------------- */
/* Number of Reaction entries the model produces (one per constrained
 * node/dof pair, duplicates kept; solver.rs:15-41). */
int64_t svx_model_reaction_count(const svx_model_desc* model);
int svx_solve_model(svx_ctx* ctx, const svx_model_desc* model, int method,
                    int precision, svx_solution* out, svx_error* err);

/* ---------------------------------------------------------------------------
 * Mesh: node coordinates and connectivity resident on the device, with the
 * CSR pattern and DOF numbering (row = node_position * ndof + dof - 1,
 * assembly.rs:2-3) built once per mesh.
 * ------------------------------------------------------------------------- */
int svx_mesh_create(svx_ctx* ctx, int ndof, int64_t n_nodes,
                    const double* coords /* [n_nodes*3] */, int64_t n_elem,
                    const uint8_t* kinds, const int64_t* conn_ptr,
                    const int32_t* conn_idx /* node positions */,
                    svx_mesh** out, svx_error* err);
void svx_mesh_destroy(svx_mesh* mesh);
int64_t svx_mesh_rows(const svx_mesh* mesh); /* n = n_nodes * ndof */
int64_t svx_mesh_nnz(const svx_mesh* mesh);
/* The pattern of the Csr built by csr_from_triplets (assembly.rs:87-116). */
int svx_mesh_pattern(const svx_mesh* mesh, int32_t* row_ptr /* [n+1] */,
                     int32_t* col_idx /* [nnz] */);

/* ---------------------------------------------------------------------------
 * Assembled system on a mesh — replaces assembly::assemble (assembly.rs:281).
 * E / nu per element (NaN E = "no material assigned"); loads are (row, value)
 * pairs accumulated into f in the given order (assembly.rs:365-380).
 * ------------------------------------------------------------------------- */
int svx_assemble(svx_mesh* mesh, const double* youngs, const double* poisson,
                 double thickness, int64_t n_loads, const int64_t* load_rows,
                 const double* load_values, svx_system** out, svx_error* err);
void svx_system_destroy(svx_system* sys);
/* Copy the current CSR values and load vector to the host (before
 * svx_apply_dirichlet: the assembled K and f; after: the constrained ones). */
int svx_system_values(const svx_system* sys, double* vals /* [nnz] */,
                      double* f /* [n] or NULL */);
/* Replace f (the assembled vector is kept for reactions only if this is
 * called before svx_apply_dirichlet). */
int svx_system_set_rhs(svx_system* sys, const double* f /* [n] */);
/* solver::apply_dirichlet (solver.rs:45-62): rows may repeat. */
int svx_apply_dirichlet(svx_system* sys, int64_t n_rows, const int64_t* rows);
/* Copy the system's current solution u (after svx_solve or a batched
 * solve) to the host. */
int svx_system_solution(const svx_system* sys, double* u /* [n] */);
/* solver::pcg (solver.rs:66) on the system's (constrained) K and f. */
int svx_solve(svx_system* sys, double tol, int64_t max_iter, int precision,
              double* u /* [n] or NULL */, int64_t* iterations,
              double* residual, svx_error* err);
/* What the last svx_solve or svx_batch_solve of this system did (no
 * reference counterpart: the reference reports iterations and residual
 * only, lib.rs:706-716).
 *
 * Stopping rules. A solve converges only at residual <= tol, on the
 * recursive residual of FP64 PCG, the test of solver.rs:129-131
 * (converged_by SVX_STOP_RECURSIVE), or, for a warm start whose FP64 true
 * residual ||f - K u|| / ||f|| is already there, on that (SVX_STOP_TRUE).
 * FP64 cannot evaluate f - K u below its floor sqrt(max row length) * eps64 *
 * || |K||u| + |f| || / ||f||, and on ill-conditioned systems a true residual
 * <= tol does not give the accuracy of the reference's test. So an SVX_FP32_IR
 * solve never stops in its FP32 phase: once its FP64 true residual is within
 * the floor of tol, or below it (fallback_reason 2), FP64 PCG finishes from
 * the current u to min(tol, that residual / 10); once it stops improving
 * above the floor (fallback_reason 1), to tol. After every solve the library
 * evaluates the true residual once more in FP64 (true_residual, true_floor),
 * whatever stopped it, so both precisions report the same quantity. */
typedef struct svx_solve_stats {
    int64_t iterations;      /* all Krylov iterations, FP64 fallback included */
    int64_t refreshes;       /* FP32-IR: FP64 residual replacements */
    int64_t fp64_iterations; /* FP32-IR: iterations run by the FP64 fallback */
    int64_t residual_evals;  /* FP32-IR: FP64 residuals f - K u evaluated */
    double residual;         /* the residual the stopping test read */
    double floor;            /* FP32-IR: evaluation floor at the last refresh */
    int32_t status;    /* 0 converged, 1 breakdown, 2 iteration budget spent */
    int32_t precision; /* SVX_FP64 / SVX_FP32_IR; -1 before any solve */
    double true_residual;    /* ||f - K u|| / ||f|| of the final u, in FP64 */
    double true_floor;       /* its FP64 evaluation floor */
    int32_t converged_by;    /* SVX_STOP_* */
    int32_t fallback_reason; /* FP32-IR: 0 none, 1 stalled above the floor,
                                2 at the floor above tol */
} svx_solve_stats;

/* Stopping test that ended a converged solve (svx_solve_stats). */
enum { SVX_STOP_NONE = 0, SVX_STOP_RECURSIVE = 1, SVX_STOP_TRUE = 2 };
int svx_system_stats(const svx_system* sys, svx_solve_stats* out);

/* Recovery from the last solution: centroid von Mises (stress.rs:22-42),
 * nodal von Mises (stress.rs:53-98), reactions K_orig u - f_orig at the
 * given rows (stress.rs:103-118) and the Jacobi conditioning indicator
 * max|K_ii| / min|K_ii| of the constrained K (lib.rs:634-644). */
int svx_recover(svx_system* sys, double* vm_elem /* [n_elem] */,
                double* vm_node /* [n_nodes] */, int64_t n_rows,
                const int64_t* rows, double* reactions /* [n_rows] */,
                double* cond_estimate);

/* ---------------------------------------------------------------------------
 * Raw CSR operator — Csr::mat_vec (assembly.rs:38-72) and solver::pcg
 * (solver.rs:66) on caller-provided matrices.
 * ------------------------------------------------------------------------- */
int svx_csr_create(svx_ctx* ctx, int64_t n, const int32_t* row_ptr,
                   const int32_t* col_idx, const double* vals, svx_csr** out);
void svx_csr_destroy(svx_csr* csr);
/* y = K x. ordered != 0: one lane per row, serial column order, no fused
 * multiply-add (bit-identical to the reference); ordered == 0: the tuned
 * wavefront kernel. */
int svx_csr_spmv(svx_csr* csr, const double* x, double* y, int ordered);
int svx_csr_pcg(svx_csr* csr, const double* f, double tol, int64_t max_iter,
                int precision, double* u, int64_t* iterations,
                double* residual, svx_error* err);
/* svx_csr_pcg in FP64 without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops). The Krylov vectors, scalar
 * state and launch tables live in a fixed device array of the library,
 * about 38 MB per GPU, held until the process ends; matrices of more than
 * 524,288 rows return SVX_E_CAPACITY. Every call on one GPU uses the same
 * array: calls that overlap on a GPU (from several threads or contexts)
 * race and must be serialised by the caller. Same kernels and launch shape
 * as svx_csr_pcg with SVX_FP64, run kernel by kernel. */
int svx_csr_pcg_noalloc(svx_csr* csr, const double* f, double tol,
                        int64_t max_iter, double* u, int64_t* iterations,
                        double* residual, svx_error* err);

/* ---------------------------------------------------------------------------
 * Batched solves (no reference counterpart: the reference solves one model
 * per call, lib.rs:586). Systems of any size and element family, assembled
 * and constrained, are solved together: each PCG phase is one launch over the
 * row tiles of every running instance, chunks of iterations run as one
 * captured hipGraph, every instance stops on its own convergence test. Each
 * instance's result equals its solo solve bit for bit (per-instance
 * reductions in a fixed order). The batch does not own the systems; after
 * svx_batch_solve each system holds its solution for svx_recover.
 * ------------------------------------------------------------------------- */
int svx_batch_create(svx_ctx* ctx, svx_batch** out);
void svx_batch_destroy(svx_batch* batch);
/* tol > 0; max_iter <= 0 selects max(1000, 10 n) (lib.rs:632). */
int svx_batch_add(svx_batch* batch, svx_system* sys, double tol,
                  int64_t max_iter, int64_t* slot);
int64_t svx_batch_size(const svx_batch* batch);
/* SVX_E_NOT_CONVERGED if any instance failed; see svx_batch_result. */
int svx_batch_solve(svx_batch* batch, int precision, svx_error* err);
int svx_batch_result(const svx_batch* batch, int64_t slot,
                     int64_t* iterations, double* residual,
                     int32_t* converged);
/* SpMV autotuner: times one solver iteration (SpMV with its direction step,
 * update, and the FP32-IR period end per iteration) over the whole batch for
 * blocks in flight per lane in {1, 2, 3, 4} x 1, 2 or 4 ELL chunks (64 node
 * rows each) per workgroup tile x both tile orders, one node-block size at
 * a time, and keeps the fastest shape for later svx_batch_solve calls with
 * this precision (until svx_batch_add). unroll[bs] / rows_per_tile[bs] /
 * tile_order[bs] (arrays of 4, indexed by block size 1..3, or NULL;
 * rows_per_tile 0 = each instance's default) receive the choice, spmv_ms the
 * iteration time.
 * The unroll does not change results; a tile size other than the default
 * changes the order of the partial sums of the dot products, so results then
 * differ from solo solves at rounding level instead of bit for bit. */
int svx_batch_autotune(svx_batch* batch, int precision, int reps,
                       int32_t* unroll /* [4] */, int32_t* rows_per_tile /* [4] */,
                       int32_t* tile_order /* [4] */, double* spmv_ms);
/* Solver set-up of svx_batch_solve without the solve (vectors, instance
 * state, tile tables with the tuned or default launch shape and the
 * operator in force), so that svx_batch_time_spmv can time the batch's SpMV
 * alone. No reference counterpart: a measurement entry. */
int svx_batch_prepare(svx_batch* batch, int precision);
/* Average duration and algorithmic bytes of one SpMV phase over the whole
 * batch (all instances), timed like svx_system_time_spmv, after a solve or
 * svx_batch_prepare with the same precision. */
int svx_batch_time_spmv(svx_batch* batch, int reps, double* avg_ms,
                        double* bytes_per_launch);

/* ---------------------------------------------------------------------------
 * Multi-RHS solve (no reference counterpart: the reference solves one load
 * vector per call). nrhs <= 64 load vectors F[k*n .. k*n+n) on one assembled,
 * constrained system; constrained rows of every F are zeroed as
 * apply_dirichlet does (solver.rs:59-61). Each load case follows the
 * recurrence of solver.rs:82-149 with its own scalars and convergence test;
 * the matrix is read once per iteration for all of them. U[k*n ..] receives
 * the solutions; iterations / residuals / converged are per load case;
 * elapsed_ms is device time from the upload of F to the download of U.
 * SVX_FP64: one block PCG, vectors [n][R], the matrix read once per
 * iteration for all cases. SVX_FP32_IR on a mesh with node groups
 * (svx_set_multi_group): a block FP32 phase (FP32 matrix and vectors, FP64
 * dots) from u = 0 with an FP64 refresh every 32 iterations (u += the FP32
 * correction, r = f - K u in FP64); a case leaves it at the first refresh
 * whose true residual reaches tol or does not halve its best, or on a
 * breakdown, and the FP64 block PCG finishes every case from its u to tol on
 * the recursive residual (the FP32 phase never converges a case). Without
 * node groups every case runs the FP32-IR solve of svx_solve as one instance
 * of a batched engine over shared matrix copies, bit for bit that solve.
 * ------------------------------------------------------------------------- */
int svx_solve_multi(svx_system* sys, int32_t nrhs, const double* F, double tol,
                    int64_t max_iter, int precision, double* U,
                    int64_t* iterations, double* residuals, int32_t* converged,
                    double* elapsed_ms, svx_error* err);
/* Timing of the FP64 block solve's SpMV kernel for `nrhs` load cases on
 * this system, in the launch shape svx_solve_multi uses: `reps` launches on
 * the context's stream after one warm iteration's SpMV, *avg_us the average
 * per launch (HIP events), *bytes its algorithmic bytes. With node groups
 * (svx_set_multi_group) the kernel is the group product of a direction the
 * step pass already wrote: 8 nnz + 4 nblk + 16 n nrhs (matrix values, block
 * columns, p' in, q out). The gathering SpMV evaluates the step itself:
 * 8 nnz + 4 nblk + 32 n nrhs + 8 n (r and the previous direction in, the
 * diagonal, p' and q out). No reference counterpart: a measurement entry. */
int svx_multi_spmv_timing(svx_system* sys, int32_t nrhs, int32_t reps, double* avg_us, double* bytes);

/* ---------------------------------------------------------------------------
 * Measurement hook (no reference counterpart): average duration of the
 * solver's own SpMV kernel for `precision` on this system's matrix over
 * `reps` launches, timed with HIP events on the context's stream, and the
 * algorithmic bytes one launch moves (matrix + indices + one read of x +
 * one write of y). The system must have been solved with that precision.
 * ------------------------------------------------------------------------- */
int svx_system_time_spmv(svx_system* sys, int precision, int reps,
                         double* avg_ms, double* bytes_per_launch);
/* Solver set-up of svx_solve for `precision` (vectors, scalar state, tile
 * tables, with the context's operator) without solving, after which
 * svx_system_time_spmv times that precision's SpMV. No reference
 * counterpart: a measurement entry. */
int svx_system_prepare_solver(svx_system* sys, int precision);

/* Measurement hook (no reference counterpart): the SpMV of every system in
 * systems[0..count) launched in turn, reps rounds, on their shared context's
 * stream, timed with HIP events. avg_ms receives the average duration of one
 * launch and bytes_per_launch the average algorithmic bytes of one launch.
 * With systems whose matrices together exceed the 256 MB Infinity Cache,
 * every launch streams its matrix from HBM. Every system must have been
 * solved with that precision on the same context. */
int svx_time_spmv_rotation(svx_system* const* systems, int count, int precision,
                           int reps, double* avg_ms, double* bytes_per_launch);

/* Measurement hook (no reference counterpart): average device time of each
 * phase of one solver iteration, each launched alone on this system with the
 * solve's launch shape, reps times: avg_ms[0] SpMV, [1] update (alpha, u, r,
 * z), [2] direction update (beta, p), [3] FP32-IR refresh (fold, FP64
 * residual, repoint; 0 for FP64). bytes_per_launch[4] (or NULL) receives the
 * algorithmic bytes of each phase. The system must have been solved with that
 * precision; its solver state and solution are overwritten (solve again
 * before reading either). */
int svx_system_time_phases(svx_system* sys, int precision, int reps,
                           double* avg_ms /* [4] */,
                           double* bytes_per_launch /* [4] */);

/* ---------------------------------------------------------------------------
 * Request server (no reference counterpart: the reference solves one deck per
 * process). Whole-model solves (the svx_solve_model pipeline, Method::Auto
 * without the LDLT fallback) submitted from any host thread are validated and
 * given their CSR pattern on a host worker, then admitted to the device while
 * their estimated footprint fits hbm_budget and fewer than max_active are in
 * flight; admitted systems join running Krylov engines at the next chunk
 * boundary (continuous batching) and leave as they converge, freeing their
 * memory. A request's result equals its svx_solve_model result with the same
 * precision. The server owns ctx while it exists: no other calls on ctx.
 * ------------------------------------------------------------------------- */
typedef struct svx_server_config {
    int32_t precision;  /* SVX_FP64 or SVX_FP32_IR */
    int32_t max_active; /* requests in flight on the device at most */
    int64_t hbm_budget; /* device bytes the admitted requests may hold */
    double tol;         /* relative residual; 1e-10 is the reference's */
    int32_t chunk;      /* iterations between admissions; 0 = context default */
    int32_t use_graph;  /* 1: each chunk as a captured hipGraph; 0: kernel by
                           kernel (membership changes at most chunk boundaries
                           of a busy server, so a capture is rarely reused) */
} svx_server_config;

typedef struct svx_request_info {
    double latency_ms; /* submit to result */
    double queue_ms;   /* submit to admission */
    double prepare_ms; /* submit to end of host preparation */
    int64_t iterations, refreshes, fp64_iterations;
    int64_t device_bytes; /* admitted footprint estimate */
    int64_t n_dof;
    /* array lengths the result needs (for decks submitted as text): nodes,
     * elements, reactions; 0 before the deck is read */
    int64_t n_nodes, n_elem, n_reactions;
} svx_request_info;

typedef struct svx_server_summary {
    int64_t submitted, completed, failed, queued;
    int64_t resident_bytes, peak_resident_bytes, peak_active;
    int64_t solved_dof;        /* DOF of completed requests */
    double p50_ms, p99_ms, max_ms; /* latency of completed requests */
    double dof_per_s;          /* solved DOF / (last completion - first submit) */
    /* device thread's wall time (ms) by activity: admission (upload,
     * assembly, constraints, solver set-up), Krylov chunks (tile tables and
     * graph capture included), retirement (results, recovery, freeing) and
     * waiting for work; chunk launches and tile-table rebuilds */
    double admit_ms, solve_ms, retire_ms, idle_ms;
    int64_t chunks, retiles;
} svx_server_summary;

int svx_server_create(svx_ctx* ctx, const svx_server_config* cfg,
                      svx_server** out);
/* Stops the server. Requests still in flight are dropped: every waiter
 * blocked in svx_server_wait returns SVX_E_DEVICE, and destroy returns once
 * no waiter is inside the server. */
void svx_server_destroy(svx_server* srv);
/* Copies the deck; returns at once with a ticket. */
int svx_server_submit(svx_server* srv, const svx_model_desc* model,
                      int64_t* ticket);
/* Copies the deck text; a host worker reads it (svx_deck_parse) before
 * preparing the request. Read errors come back from svx_server_wait as
 * SVX_E_PARSE. */
int svx_server_submit_deck(svx_server* srv, const char* text, size_t len,
                           int64_t* ticket);
/* Waits for a ticket (timeout_ms < 0: no limit; SVX_E_TIMEOUT when it
 * expires, the ticket stays valid). On completion fills out (arrays sized as
 * for svx_solve_model) and info, returns the request's status (errors as
 * svx_solve_model reports them, SVX_E_CAPACITY for a deck larger than the
 * budget, SVX_E_DEVICE for requests the server dropped when it failed or was
 * destroyed), and releases the ticket. A done request whose reactions exceed
 * out->reaction_capacity returns SVX_E_ARGUMENT, writes nothing and keeps
 * the ticket, so the caller can wait again with larger arrays. With out NULL
 * the call waits, fills info (with the result's array lengths) and returns
 * the status, and keeps the ticket. */
int svx_server_wait(svx_server* srv, int64_t ticket, double timeout_ms,
                    svx_solution* out, svx_request_info* info,
                    svx_error* err);
int svx_server_stats(svx_server* srv, svx_server_summary* out);

#ifdef __cplusplus
}
#endif

#endif /* SVX_H */
