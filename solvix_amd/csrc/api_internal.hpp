// Handle types behind the C ABI and the host-side model preparation.
#pragma once
#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"
#include "pattern.hpp"
#include "svx_internal.hpp"

namespace svx {

// Launch-shape and refinement knobs (0 = automatic choice).
constexpr int kBatchStreamsDefault = 2;
constexpr int kBatchStreamsMax = 8;

struct TuneParams {
    int lanes = 0;          // lanes per row, raw CSR systems
    int unroll = 0;         // blocks in flight per lane, node-block (ELL) systems
    int rows_per_tile = 0;  // rows per workgroup tile
    int period = 8;         // FP32-IR refresh period (iterations)
    int chunk = 32;         // iterations per captured graph
    bool use_graph = true;
    double ru_delta = 0.1;  // FP32-IR: refresh when ||r|| < ru_delta * ||r_last||
    bool matrix_free = false;  // FP64 iterations on the matrix-free operator where a system has one
    int tile_order = SVX_TILES_XCD;  // order of the gathering kernels' tiles over the XCDs
    int load_hint = SVX_LOADS_AUTO;  // FP64 ELL SpMV matrix loads
    int step_tiles = 0;              // EngineShape::step_tiles; 0: the engine's default
    int multi_group = 0;  // multi-RHS block SpMV: nodes per group (0: default, -1: gathering SpMV)
    int batch_streams = 0;  // engines a batch solve runs side by side (0: default)
};

// Grow-only device scratch of one context for the temporaries of a call
// (element matrices, masks, recovery buffers), so that assembling or
// recovering a system allocates and frees no device memory. Every use is on
// the context's stream: a later call's temporaries are ordered after the
// kernels of an earlier call that used the same bytes.
struct Scratch {
    DevBuf<char> buf;
    // Room for `sizes` (each 256-byte aligned) carved in order; pointers out[i].
    void carve(hipStream_t s, std::initializer_list<size_t> sizes, void** out) {
        size_t total = 0;
        for (size_t b : sizes) total += (b + 255) / 256 * 256;
        if (total > buf.n) {
            SVX_HIP(hipStreamSynchronize(s));  // the old bytes may still be in use
            buf.alloc(total + total / 2);
        }
        size_t off = 0, i = 0;
        for (size_t b : sizes) {
            out[i++] = buf.p + off;
            off += (b + 255) / 256 * 256;
        }
    }
};

}  // namespace svx

struct svx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    int32_t* h_flag = nullptr;  // pinned
    int32_t* d_flag = nullptr;
    svx::TuneParams tune;
    svx::Scratch scratch;
};

struct svx_mesh {
    svx_ctx* ctx = nullptr;
    int ndof = 0;
    int64_t n_nodes = 0, n_elem = 0, n = 0, nnz = 0, nblk = 0, max_row_len = 0;
    svx::HostPattern hp;
    std::vector<int64_t> h_conn_ptr;
    std::vector<int32_t> h_conn;
    std::vector<uint8_t> h_kinds;
    svx::DevBuf<double> coords;
    svx::DevBuf<int64_t> conn_ptr;
    svx::DevBuf<int32_t> conn;
    svx::DevBuf<uint8_t> kinds;
    svx::DevBuf<int32_t> row_ptr, col, brow_ptr, bcol;
    svx::DevBuf<int32_t> ell_off, ell_nb, ell_col;  // node-row sliced ELL pattern
    int64_t n_chunks = 0, ell_blocks = 0;
    svx::DevBuf<int64_t> inc_ptr;
    svx::DevBuf<int32_t> inc_elem;
    svx::DevBuf<int8_t> inc_slot;
    svx::DevBuf<int32_t> sel[3];
    int64_t nsel[3] = {0, 0, 0};
    svx::DevBuf<int64_t> ke_off, ns_off;
    int64_t ke_total = 0, ns_total = 0;
    std::vector<double> h_coords;  // [n_nodes][3]
    // matrix-free operator, planned on first use: 0 not yet, 2 planned on the
    // host (not uploaded), 1 ready, -1 none
    int mf_state = 0;
    std::unique_ptr<svx::MfPlan> mf_plan;  // plane meshes, while mf_state == 2
    std::vector<uint8_t> mf_rank_host;     // C3D8 meshes, while mf_state == 2
    int mf_kind = 0;
    int32_t mf_patches = 0, mf_cap_nodes = 0, mf_cap_elem = 0;
    int64_t mf_elem_slots = 0;
    svx::DevBuf<svx::MfPatch> mf_patch;
    svx::DevBuf<int32_t> mf_nodes;
    svx::DevBuf<uint16_t> mf_econn, mf_adj;
    svx::DevBuf<double> mf_xy;
    svx::DevBuf<uint8_t> mf_rank;  // C3D8: rank of each (element, corner) among the corner node's elements
    int32_t mf_width = 0;          // C3D8: most elements at one node
    // multi-RHS block SpMV node groups, planned on the first block solve for
    // the group size in force (blk_group); blk_groups 0: no plan
    int32_t blk_group = 0, blk_groups = 0;
    svx::DevBuf<int32_t> blk_grp, blk_nodes, blk_ucol, blk_pk_src, blk_pk_stride;
};

struct svx_system {
    svx_mesh* mesh = nullptr;
    svx::DevBuf<double> E, nu;
    double thickness = 1.0;
    bool mat_uniform = false;  // one (E, nu) for every element
    double E0 = 0.0, nu0 = 0.0;
    svx::DevBuf<uint8_t> mf_mask;  // constrained dofs per node, for the matrix-free operator
    svx::DevBuf<double> mf_fe;  // C3D8 matrix-free product: forces per (node, incident element)
    bool mf_ready = false;
    svx::DevBuf<double> vals, f;
    std::vector<double> h_f;  // current load vector (host copy)
    bool dirichlet = false;
    std::vector<int64_t> cons_rows;  // unique, ascending
    std::vector<double> f_orig_cons;
    svx::DevBuf<int64_t> d_cons_rows, d_cons_off;
    svx::DevBuf<double> d_cons_vals;  // original values of the constrained rows
    svx::DevBuf<double> dinv;
    svx::DevBuf<float> dinv32;
    svx::DevBuf<double> ell64;  // K in the ELL layout (FP64 PCG, FP32-IR fallback)
    svx::DevBuf<float> ell32;   // K in the ELL layout, FP32 (FP32-IR)
    svx::DevBuf<double> u, r, p, z, q, p_alt;
    svx::DevBuf<float> x32, r32, p32, z32, q32, p32_alt;
    svx::DevBuf<svx::InstScal> scal;
    std::unique_ptr<svx::Engine> engine;
    int engine_precision = -1;  // precision the engine's descriptors were built for
    bool engine_mf = false;     // the engine's FP64 products use the matrix-free operator
    double cond = 0.0;
    bool solved = false, has_u = false;
    svx::InstScal last{};    // scalars of the last solve
    int last_precision = -1;
    double true_res = 0.0, true_floor = 0.0;  // FP64 ||f - K u|| / ||f|| of the last solve's u, its floor
};

struct svx_csr {
    svx_ctx* ctx = nullptr;
    int64_t n = 0, nnz = 0, max_row_len = 0;
    svx::DevBuf<int32_t> row_ptr, col;
    svx::DevBuf<double> vals;
};

namespace svx {

int nodes_of(uint8_t kind);
// Message text of an SVX_ELEM_* reason, in the reference's wording.
const char* reason_text(int reason);
// The per-block-size launch shape every batched solve uses by default.
EngineShape batch_shape(const std::vector<svx_system*>& systems, const TuneParams& tp, int precision);
// The load hint (EngineShape::nt, nt32) of the FP64 (value_bytes 8) or
// FP32 (4) ELL SpMV over `systems`' instances of node-block size bs:
// tp.load_hint, or for SVX_LOADS_AUTO streaming once their ELL values in
// that precision exceed kStreamBytes.
constexpr int64_t kStreamBytes = int64_t(256) << 20;
int streaming_loads(const std::vector<svx_system*>& systems, int bs, const TuneParams& tp, int value_bytes);
// The vector hint (EngineShape::ntv) of the FP64 and FP32 updates over
// `systems`' instances of node-block size bs: tp.load_hint, or for
// SVX_LOADS_AUTO streaming once the FP64 update's vectors (five reads and
// three writes of 8 bytes per row) exceed kStreamBytes.
int streaming_vectors(const std::vector<svx_system*>& systems, int bs, const TuneParams& tp);

// A model validated and resolved on the host (lib.rs:593-632 up to the
// element pass): node positions, per-element material constants, the load
// vector and the constrained rows, with deferred load / BC errors that the
// reference reports only after the element pass.
struct PreparedModel {
    int ndof = 2;
    std::vector<int32_t> conn_pos;
    std::vector<double> E, nu;  // NaN E = no material
    std::vector<double> f;
    std::vector<int64_t> cons_rows;  // BC order, duplicates kept
    std::vector<uint32_t> cons_node;
    std::vector<int32_t> cons_dof;
    std::string load_error, bc_error;
};
int prepare_model(const svx_model_desc* m, PreparedModel& out, svx_error* err);

void mesh_build(svx_mesh* m, int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const uint8_t* kinds,
                const int64_t* conn_ptr, const int32_t* conn_idx);
// The host half of mesh_build (pattern, incidence, offsets; no device calls)
// and the device half (uploads; needs m->ctx).
void mesh_build_host(svx_mesh* m, int ndof, int64_t n_nodes, int64_t n_elem, const uint8_t* kinds,
                     const int64_t* conn_ptr, const int32_t* conn_idx);
void mesh_upload(svx_mesh* m, const double* coords);
void system_assemble(svx_system* sys, const double* E, const double* nu, double thickness, int64_t* bad_index,
                     int* bad_reason);
void system_set_f(svx_system* sys, const double* f);
void system_dirichlet(svx_system* sys, int64_t n_rows, const int64_t* rows);
// matrix_free: also prepare the matrix-free operator (FP64, plane meshes of
// one element kind and one material); system_desc then describes it.
void system_prepare(svx_system* sys, int precision, bool matrix_free = false);
InstDesc system_desc(svx_system* sys);
// Plan and upload the mesh's matrix-free patches once; false if it has none.
bool mesh_mf_prepare(svx_mesh* m);
// The host half of mesh_mf_prepare (no device calls; coords [n_nodes][3]),
// for callers that plan off the device thread; false if the mesh has no
// matrix-free operator.
bool mesh_mf_plan(svx_mesh* m, const double* coords);
double spmv_bytes(const svx_mesh* m, int precision);
// Algorithmic bytes of one matrix-free product (direction step fused in).
double mf_spmv_bytes(const svx_mesh* m);
InstScal initial_scal(double tol, int64_t max_iter, double ru_delta, int64_t max_row_len);
// The set-up half of system_solve: vectors, scalar state, engine tables.
void system_engine_setup(svx_system* sys, double tol, int64_t max_iter, int precision, const TuneParams& tp);
void system_solve(svx_system* sys, double tol, int64_t max_iter, int precision, InstScal* out_scal,
                  const TuneParams& tp);
void system_recover(svx_system* sys, double* vm_elem, double* vm_node, int64_t n_rows, const int64_t* rows,
                    double* reactions);
// ||f - K u|| / ||f|| of the system's current u on its constrained K and f,
// evaluated in FP64, and its FP64 evaluation floor; stored in sys->true_res /
// sys->true_floor.
void system_true_residual(svx_system* sys);
// Resolve this device's copy of the no-allocation scratch array (with the
// device current); svx_ctx_create calls it once per context.
void noalloc_register_device(int device);
int64_t noalloc_max_rows();

}  // namespace svx
