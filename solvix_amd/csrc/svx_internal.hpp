// Internal declarations shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/svx.h"

namespace svx {

// ----------------------------------------------------------------- errors
struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};

#define SVX_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::svx::DeviceError(std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

// ------------------------------------------------------------ device memory
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool own = true;  // false: a view of memory someone else holds (never freed here)
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), own(o.own) { o.p = nullptr, o.n = 0, o.own = true; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p, n = o.n, own = o.own;
            o.p = nullptr, o.n = 0, o.own = true;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) SVX_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    // Use `count` elements at `mem` (device memory owned elsewhere).
    void view(T* mem, size_t count) {
        release();
        p = mem, n = count, own = false;
    }
    void release() {
        if (p && own) (void)hipFree(p);
        p = nullptr, n = 0, own = true;
    }
    size_t bytes() const { return n * sizeof(T); }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count) SVX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) SVX_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ------------------------------------------------------- Krylov instances
// Instance status values written by the device.
// ST_FALLBACK: an FP32-IR instance whose FP64 residual stopped improving; it
// continues as FP64 PCG warm-started from its current u.
enum : int32_t { ST_RUNNING = 0, ST_CONVERGED = 1, ST_BREAKDOWN = 2, ST_MAXITER = 3, ST_IDLE = 4, ST_FALLBACK = 5 };
// Instance flags. FL_FRESH marks work the next init phase must initialise.
// FL_WARM (with FL_FRESH): initialise the FP64 recurrence from the current u
// (r = f - K u) instead of u = 0.
enum : int32_t { FL_REFRESH = 1, FL_FRESH = 4, FL_WARM = 8 };

// Per-instance scalar state of a (preconditioned) CG solve, device resident.
// Within a kernel only the instance's leader tile (its first row tile)
// writes the live fields; every tile re-derives the scalars it needs from the
// per-tile partials of the previous phase (fixed order, identical in every
// tile) and from the snap_* fields, which the leader writes one phase earlier
// and nobody writes in the phase that reads them.
struct alignas(16) InstScal {
    double bnorm;       // ||f||
    double rz;          // r . z of the current iterate
    double pap;         // p . K p
    double alpha;
    double beta;
    double rr;          // r . r of the current (recursive) residual
    double res;         // reported relative residual
    double rmax;        // FP32-IR: max ||r|| since the latest refresh
    double floor;       // FP32-IR: FP64 evaluation floor at the latest refresh
    double res_best;    // FP32-IR: best FP64 relative residual at a refresh
    int64_t it;         // iterations done
    int64_t refreshes;  // FP32-IR residual replacements
    int64_t it_refresh; // FP32-IR: iteration of the latest refresh
    int64_t fp64_it;    // FP32-IR fallback: -(iterations before FP64 took over)
    int64_t evals;      // FP32-IR: FP64 residual evaluations (refresh steps run)
    int32_t status;     // ST_*
    int32_t flags;      // FL_*
    int32_t stall;      // FP32-IR refreshes without halving res_best
    int32_t dir;        // 1: the direction in P[0] is complete (after init / refresh); 0: the next SpMV takes the step
    int32_t stop_rule;  // SVX_STOP_TRUE when an FP64 true residual <= tol ended the solve; 0 otherwise
    int32_t fb_reason;  // FP32-IR: why FP64 took over (0 none, 1 stalled above the floor, 2 at the floor above tol)
    // snapshots
    double snap_rz;
    double snap_rmax;
    double snap_res_best;
    int64_t snap_it;
    int32_t snap_flags;
    int32_t snap_stall;
    // configuration (host)
    double tol;
    double ru_delta;     // FP32-IR refresh when ||r|| < ru_delta * rmax
    double floor_scale;  // sqrt(max row length) * eps64
    int64_t max_iter;
};

// One node patch of the matrix-free operator: up to kBlock owned nodes (a
// run of the mesh's Morton order), the elements incident to them and the
// nodes of those elements.
struct MfPatch {
    int32_t node_off, n_nodes, n_owned, elem_off, n_elem, adj_off, n_adj, pad;
};

// Device pointers of one instance. FP64 CSR always (FP32-IR residuals,
// warm starts). Node-block instances (bs = 2, 3) multiply in the node-row
// sliced ELL layout, FP64 values for FP64 PCG and FP32 values for the FP32-IR
// inner iteration; raw CSR instances (bs = 1) multiply the CSR directly, and
// FP32-IR reads an FP32 copy of the CSR values (bval).
struct InstDesc {
    int32_t n;           // rows
    int32_t tile_begin;  // first tile of this instance in the tile table
    int32_t tile_count;
    int32_t bs;          // node-block size (ndof); 1 = raw CSR
    // set by the engine: the FP64 SpMV of this (large, assembled) instance
    // reads the pending direction step that cg64_mf_step_kernel took once
    // for it, instead of every tile reducing the instance's partial sums
    int32_t step_kernel;
    // set by the engine: the FP64 update of this (large) instance reads the
    // alpha that cg64_alpha_kernel took once for it (the p.q reduction and
    // the breakdown test), instead of every row tile reducing the partials
    int32_t alpha_kernel;
    const int32_t* row_ptr;
    const int32_t* col;
    const double* val;
    const int32_t* brow_ptr;  // bs = 1: row_ptr
    const int32_t* bcol;      // bs = 1: col
    const float* bval;        // bs = 1: FP32 CSR values
    const int32_t* ell_off;   // [n/bs/64 + 1] chunk offsets in blocks
    const int32_t* ell_nb;    // padded blocks per row, per chunk
    const int32_t* ell_col;   // block columns, [k][lane] per chunk
    const double* ell_val;    // [block][a*bs+b][lane] per chunk
    const float* ell_val32;
    const double* dinv;
    const float* dinv32;
    const double* f;
    double* u;   // FP64 solution
    double* r;   // FP64 residual (FP64 CG) / true residual (FP32-IR)
    double* p;
    double* z;
    double* q;   // K p
    float* x32;  // FP32-IR: correction accumulated since the last replacement
    float* r32;
    float* p32;
    float* z32;
    float* q32;
    double* p_alt;    // second direction buffer (the SpMV writes the new direction there)
    float* p32_alt;
    InstScal* s;
    // Matrix-free plane-stress operator (mf_kind 3 = CPS3, 4 = CPS4, 0 =
    // none): q = K p from the element formulas over the mesh's node patches,
    // instead of the assembled values (FP64 iterations only).
    const MfPatch* mf_patch;
    const int32_t* mf_nodes;    // patch node lists: owned nodes, then halo nodes
    const uint16_t* mf_econn;   // [element slot][4] patch-local node positions
    const uint16_t* mf_adj;     // [patch][k][kBlock] (slot << 2 | corner) of owned nodes
    const double* mf_xy;        // [n_nodes][2] coordinates
    const uint8_t* mf_mask;     // [n_nodes] constrained dofs (bit d: dof d + 1)
    double mf_c11, mf_c12, mf_c33, mf_t;  // plane-stress D (element.rs:10-17) and thickness;
                                          // C3D8: lambda + 2 mu, lambda, mu (element.rs:33-47)
    // C3D8 (mf_kind 8): connectivity [n_elem][8]; for each (element, corner)
    // the element's rank among the corner node's incident elements; the
    // force buffer of the two-pass product, node-major [n_nodes][mf_width][3]
    // (slot k of node i: the force of its k-th incident element, in element
    // order; unused slots stay zero); mf_xy is the mesh coordinates
    // [n_nodes][3]
    const int32_t* mf_conn;
    const uint8_t* mf_rank;
    double* mf_fe;
    int32_t mf_n_elem;
    int32_t mf_width;
    int32_t mf_kind;
    int32_t mf_patches;
    int32_t mf_cap_nodes;  // largest patch (nodes, elements)
    int32_t mf_cap_elem;
};


// A tile is a contiguous row range of one instance handled by one workgroup.
// It carries its instance's scalar-state pointer and tile range, so a kernel
// can issue the loads of the instance state, of the partials and of the
// descriptor together, one dependent step after the tile itself.
struct alignas(16) Tile {
    int32_t inst;
    int32_t row_begin;   // matrix-free SpMV tiles: the patch
    int32_t row_end;
    int32_t tile_begin;  // the instance's first row tile (in its group)
    int32_t tile_count;  // the instance's row tiles
    int32_t pq_slot;     // partial slot and tile range of the instance's p.q sums
    int32_t pq_begin;
    int32_t pq_count;
    InstScal* s;
};

constexpr int kBlock = 256;
constexpr int kMaxUnroll = 4;  // ELL SpMV: blocks in flight per lane, 1..kMaxUnroll
constexpr int kEllRows = 64;   // node rows per ELL chunk (one wavefront)
// Per-tile partial sums, one slot per reduced quantity.
// P_MPQ: p.q of the matrix-free SpMV, indexed by its patch tiles.
enum : int { P_PQ = 0, P_RR = 1, P_RZ = 2, P_TRR = 3, P_TRZ = 4, P_TFL = 5, P_IFF = 6, P_IRZ = 7, P_MPQ = 8, kPartialSlots = 9 };
// Matrix-free operator limits: nodes per patch (two per thread), element
// slots per patch (one per thread for CPS4, two for CPS3) and incident
// elements per owned node (CPS4 4, CPS3 8).
constexpr int kMfMaxNodes = 2 * 256;
constexpr int kMf3Elems = 32;  // C3D8 matrix-free product: elements per workgroup (eight threads each)
inline int mf_max_elem(int kind) { return kind == 3 ? 2 * 256 : 256; }
inline int mf_max_adj(int kind) { return kind == 3 ? 8 : 4; }

// SpMV launch shape chosen per instance set: lanes per row (CSR) or per
// block row (BSR) and rows per tile.
struct SpmvShape {
    int lanes = 16;
    int rows_per_tile = 64;
};

// ------------------------------------------------------ kernel launchers
// assembly.hip
void launch_element_ke(int kind, int64_t n_sel, const int32_t* sel, const int64_t* conn_ptr,
                       const int32_t* conn, const double* coords, const double* E, const double* nu,
                       double thickness, int ndof, double* ke, const int64_t* ke_off, uint8_t* reason,
                       unsigned long long* first_bad, hipStream_t s);
void launch_gather_assembly(int64_t n_rows, int ndof, const int32_t* row_ptr, const int32_t* col,
                            const int64_t* inc_ptr, const int32_t* inc_elem, const int8_t* inc_slot,
                            const int64_t* conn_ptr, const int32_t* conn, const uint8_t* kinds,
                            const double* ke, const int64_t* ke_off, double* vals, hipStream_t s);
void launch_dirichlet(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, const uint8_t* mask,
                      double* vals, double* f, hipStream_t s);
void launch_gather_rows(int64_t n_sel, const int64_t* rows, const int32_t* row_ptr, const double* vals,
                        const int64_t* out_off, double* out, hipStream_t s);
void launch_diag_inv(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, const double* vals,
                     double* dinv, float* dinv32, unsigned long long* absmax, unsigned long long* absmin,
                     hipStream_t s);
void launch_fill_bsr(int64_t n_rows, int bs, const int32_t* row_ptr, const int32_t* col, const double* vals,
                     const int32_t* brow_ptr, float* bval, hipStream_t s);
// CSR values of a node-block matrix (bs = 2, 3) into the node-row sliced ELL
// layout of the mesh (ell_off / ell_nb over n_nodes node rows), FP64 or FP32.
void launch_ell_pack(int bs, int64_t n_nodes, const int32_t* ell_off, const int32_t* ell_nb, const int32_t* row_ptr,
                     const int32_t* brow_ptr, const double* vals, double* out64, float* out32, hipStream_t s);
void launch_reaction_rows(int64_t n_sel, const int64_t* rows, const int32_t* row_ptr, const int32_t* col,
                          const int64_t* off, const double* orig_vals, const double* u, double* out,
                          hipStream_t s);

// recover.hip
void launch_element_stress(int kind, int64_t n_sel, const int32_t* sel, const int64_t* conn_ptr,
                           const int32_t* conn, const double* coords, const double* E, const double* nu,
                           int ndof, const double* u, double* vm_elem, double* nodal_stress,
                           const int64_t* ns_off, hipStream_t s);
void launch_nodal_average(int64_t n_nodes, const int64_t* inc_ptr, const int32_t* inc_elem,
                          const int8_t* inc_slot, const int64_t* ns_off, const double* nodal_stress,
                          double* vm_node, hipStream_t s);

// krylov.hip
void launch_spmv_ordered(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                         const double* x, double* y, hipStream_t s);
// ||f - K u||^2, || |K||u| + |f| ||^2 and ||f||^2 over the CSR rows, in FP64
// with a fixed summation order: per-block partials (3 * true_residual_blocks(n)
// doubles of `partials`) reduced by one block into out[3].
int64_t true_residual_blocks(int64_t n);
void launch_true_residual(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* u,
                          const double* f, double* partials, double* out, hipStream_t s);
void launch_spmv_csr64(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                       const double* x, double* y, int lanes, hipStream_t s);

// One CG iteration phase over a set of instances. `tiles` / `n_tiles`
// describe the work; `descs` the instances; `partials` holds kPartialSlots
// doubles per tile. Every launcher is graph-capturable (no allocation, no memset).
struct KrylovSet {
    const Tile* tiles = nullptr;
    int32_t n_tiles = 0;
    const InstDesc* descs = nullptr;
    int32_t n_inst = 0;
    double* partials = nullptr;  // [kPartialSlots * pstride]
    int32_t pstride = 0;         // partial slot stride, max(n_tiles, n_mf_tiles)
    int lanes = 16;              // lanes per row (raw CSR instances, bs = 1)
    int unroll = 2;              // blocks in flight per lane (ELL instances, bs = 2, 3)
    int bs = 0;                  // node-block size shared by the set's instances (1, 2, 3)
    int xcd = 1;                 // gathering kernels: 1 = XCD-contiguous tile order, 0 = tile = workgroup
    int nt = 0;                  // FP64 ELL SpMV: matrix values loaded nontemporal (1) or cached (0)
    int nt32 = 0;                // FP32 ELL SpMV: the same
    int ntv = 0;                 // FP64 and FP32 updates: vectors loaded and stored nontemporal (1) or cached (0)
    // FP64 SpMV: row tiles [0, n_ell_tiles) multiply the assembled matrix;
    // instances with matrix-free tiles (mf_tiles, CPS4 first, then CPS3) own
    // the row tiles after them. FP32 SpMV launches cover every row tile.
    int32_t n_ell_tiles = 0;
    const Tile* mf_tiles = nullptr;
    int32_t n_mf4 = 0, n_mf3 = 0;
    // one per instance whose FP64 pending step runs in cg64_mf_step_kernel:
    // the matrix-free instances, then the large assembled ones
    const Tile* mf_leaders = nullptr;
    int32_t n_mf_inst = 0;
    // the leaders of the instances whose alpha runs in cg64_alpha_kernel
    // (more than four times the step-kernel threshold of p.q partials)
    const Tile* alpha_leaders = nullptr;
    int32_t n_alpha = 0;
    const Tile* mf3_tiles = nullptr;  // C3D8 matrix-free element tiles (row_begin / row_end: elements)
    int32_t n_mf3_tiles = 0;
    int32_t mf_cap_nodes[2] = {0, 0};  // CPS4, CPS3: largest patch of the set
    int32_t mf_cap_elem[2] = {0, 0};
};
// An iteration is two launches: the direction step fused into the SpMV, then
// the update. The direction is double-buffered: iteration parity par reads
// the current direction from P[par] and writes the new one to P[1 - par]
// (P[0] = p / p32, P[1] = p_alt / p32_alt); a chunk starts with par = 0 and
// holds an even number of iterations, so the current direction is in P[0]
// at every chunk and period boundary.
void cg64_init(const KrylovSet& k, hipStream_t s);
void cg64_iteration(const KrylovSet& k, hipStream_t s, int par);
// After the last iteration of an FP64 chunk: the pending direction step in
// place, so that nothing is pending across a chunk boundary.
void cg64_chunk_end(const KrylovSet& k, hipStream_t s);
void ir32_init(const KrylovSet& k, hipStream_t s);
void ir32_iteration(const KrylovSet& k, hipStream_t s, int par);
// After the last iteration of an FP32-IR period: refresh decision, and the
// direction step of instances that do not refresh (a chunk is whole periods,
// so nothing is pending across a chunk boundary either).
void ir32_period_end(const KrylovSet& k, hipStream_t s);
void ir32_refresh(const KrylovSet& k, hipStream_t s);
// out[0] = running instances, out[1] = their tiles in the current tables.
void launch_count_running(const InstDesc* descs, int32_t n_inst, int32_t* out /* [2] */, hipStream_t s);
// Move per-tile partials (all slots) to a new tile layout of the same group:
// new tile g takes old tile src[g] (src[g] < 0: left untouched).
void launch_carry_partials(const double* old_p, int old_stride, double* new_p, int new_stride, const int32_t* src,
                           int n_new, hipStream_t s);
// The SpMV phase of one iteration alone (cg64 / ir32, direction step with
// beta = 1, no reduction, no scalar writes), for timing.
void spmv_phase(const KrylovSet& k, int precision, hipStream_t s);
// Set every instance of the set back to ST_RUNNING (timing helper).
void launch_set_running(const InstDesc* descs, int32_t n_inst, hipStream_t s, int32_t flags = 0);
// The update phase of one iteration, and the FP32-IR period end, alone, for
// timing.
void update_phase(const KrylovSet& k, int precision, hipStream_t s);
void period_end_phase(const KrylovSet& k, int precision, hipStream_t s);
// ST_FALLBACK instances -> ST_RUNNING with FL_FRESH | FL_WARM.
void launch_promote_fallback(const InstDesc* descs, int32_t n_inst, hipStream_t s);
// Copy every instance's scalars into a contiguous array.
void launch_gather_scalars(const InstDesc* descs, int32_t n_inst, InstScal* out, hipStream_t s);

}  // namespace svx
