// C ABI of the device path (include/svx.h): handles, the assemble / Dirichlet
// / solve / recover seam, the raw CSR operator, and the whole-model pipeline
// that mirrors solve_with_thermal (crates/solvix-core/src/lib.rs:593-716).
// No C++ exception crosses this boundary.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "api_internal.hpp"
#include "block.hpp"

using namespace svx;

namespace {

thread_local std::string g_last_error;

void set_err(svx_error* err, int code, const std::string& msg, int reason = 0, int64_t index = -1,
             int64_t iterations = 0, double residual = 0.0) {
    g_last_error = msg;
    if (!err) return;
    err->code = code;
    err->reason = reason;
    err->index = index;
    err->iterations = iterations;
    err->residual = residual;
    snprintf(err->message, sizeof err->message, "%s", msg.c_str());
}

void clear_err(svx_error* err) {
    if (!err) return;
    std::memset(err, 0, sizeof *err);
    err->index = -1;
}

template <typename F>
int guarded(svx_error* err, F&& fn) {
    try {
        return fn();
    } catch (const DeviceError& e) {
        set_err(err, SVX_E_DEVICE, e.what());
        return SVX_E_DEVICE;
    } catch (const std::bad_alloc&) {
        set_err(err, SVX_E_DEVICE, "host allocation failed");
        return SVX_E_DEVICE;
    } catch (const std::exception& e) {
        set_err(err, SVX_E_ARGUMENT, e.what());
        return SVX_E_ARGUMENT;
    }
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace svx {

int nodes_of(uint8_t kind) { return kind == SVX_CPS3 ? 3 : kind == SVX_CPS4 ? 4 : 8; }

const char* reason_text(int reason) {
    switch (reason) {
        case SVX_ELEM_NO_MATERIAL: return "no material assigned";
        case SVX_ELEM_BAD_MATERIAL: return "invalid material constants";
        case SVX_ELEM_BAD_POISSON_SOLID: return "poisson ratio must be < 0.5 for solids and plane strain";
        case SVX_ELEM_DEGENERATE_TRIANGLE: return "degenerate (zero-area) triangle";
        case SVX_ELEM_NONPOSITIVE_JACOBIAN: return "non-positive Jacobian determinant at a Gauss point";
        default: return "invalid element";
    }
}


// ------------------------------------------------------------------ mesh
void mesh_build(svx_mesh* m, int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const uint8_t* kinds,
                const int64_t* conn_ptr, const int32_t* conn_idx) {
    mesh_build_host(m, ndof, n_nodes, n_elem, kinds, conn_ptr, conn_idx);
    mesh_upload(m, coords);
}

void mesh_build_host(svx_mesh* m, int ndof, int64_t n_nodes, int64_t n_elem, const uint8_t* kinds,
                     const int64_t* conn_ptr, const int32_t* conn_idx) {
    m->ndof = ndof;
    m->n_nodes = n_nodes;
    m->n_elem = n_elem;
    m->h_conn_ptr.assign(conn_ptr, conn_ptr + n_elem + 1);
    m->h_conn.assign(conn_idx, conn_idx + conn_ptr[n_elem]);
    m->h_kinds.assign(kinds, kinds + n_elem);
    build_pattern(ndof, n_nodes, n_elem, conn_ptr, conn_idx, m->hp);
    m->n = n_nodes * ndof;
    m->nnz = m->hp.row_ptr[m->n];
    m->nblk = m->hp.brow_ptr[n_nodes];
    m->max_row_len = 0;
    for (int64_t r = 0; r < m->n; ++r)
        m->max_row_len = std::max<int64_t>(m->max_row_len, m->hp.row_ptr[r + 1] - m->hp.row_ptr[r]);
    int64_t ke = 0, ns = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        int nen = nodes_of(kinds[e]);
        ke += (int64_t)nen * ndof * nen * ndof;
        ns += (int64_t)nen * 6;
    }
    m->ke_total = ke;
    m->ns_total = ns;
}

void mesh_upload(svx_mesh* m, const double* coords) {
    hipStream_t s = m->ctx->stream;
    const int64_t n_nodes = m->n_nodes, n_elem = m->n_elem;
    const int ndof = m->ndof;
    const int64_t* conn_ptr = m->h_conn_ptr.data();
    const uint8_t* kinds = m->h_kinds.data();
    m->coords.alloc(n_nodes * 3);
    m->coords.upload(coords, n_nodes * 3, s);
    m->h_coords.assign(coords, coords + n_nodes * 3);
    m->conn_ptr.alloc(n_elem + 1);
    m->conn_ptr.upload(conn_ptr, n_elem + 1, s);
    m->conn.alloc(m->h_conn.size());
    m->conn.upload(m->h_conn.data(), m->h_conn.size(), s);
    m->kinds.alloc(n_elem);
    m->kinds.upload(kinds, n_elem, s);
    m->row_ptr.alloc(m->n + 1);
    m->row_ptr.upload(m->hp.row_ptr.data(), m->n + 1, s);
    m->col.alloc(m->nnz);
    m->col.upload(m->hp.col.data(), m->nnz, s);
    m->brow_ptr.alloc(n_nodes + 1);
    m->brow_ptr.upload(m->hp.brow_ptr.data(), n_nodes + 1, s);
    m->bcol.alloc(m->nblk);
    m->bcol.upload(m->hp.bcol.data(), m->nblk, s);
    m->n_chunks = (int64_t)m->hp.ell_nb.size();
    m->ell_blocks = m->hp.ell_off.empty() ? 0 : m->hp.ell_off.back();
    m->ell_off.alloc(m->n_chunks + 1);
    m->ell_off.upload(m->hp.ell_off.data(), m->n_chunks + 1, s);
    m->ell_nb.alloc(std::max<int64_t>(1, m->n_chunks));
    m->ell_nb.upload(m->hp.ell_nb.data(), m->n_chunks, s);
    m->ell_col.alloc(std::max<int64_t>(1, m->ell_blocks));
    m->ell_col.upload(m->hp.ell_col.data(), m->ell_blocks, s);
    m->inc_ptr.alloc(n_nodes + 1);
    m->inc_ptr.upload(m->hp.inc_ptr.data(), n_nodes + 1, s);
    m->inc_elem.alloc(m->hp.inc_elem.size());
    m->inc_elem.upload(m->hp.inc_elem.data(), m->hp.inc_elem.size(), s);
    m->inc_slot.alloc(m->hp.inc_slot.size());
    m->inc_slot.upload(m->hp.inc_slot.data(), m->hp.inc_slot.size(), s);
    std::vector<int32_t> sel[3];
    std::vector<int64_t> ke_off(n_elem), ns_off(n_elem);
    int64_t ke = 0, ns = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
        int nen = nodes_of(kinds[e]);
        sel[kinds[e]].push_back((int32_t)e);
        ke_off[e] = ke;
        ns_off[e] = ns;
        ke += (int64_t)nen * ndof * nen * ndof;
        ns += (int64_t)nen * 6;
    }
    for (int k = 0; k < 3; ++k) {
        m->nsel[k] = (int64_t)sel[k].size();
        m->sel[k].alloc(sel[k].size());
        m->sel[k].upload(sel[k].data(), sel[k].size(), s);
    }
    m->ke_off.alloc(n_elem);
    m->ke_off.upload(ke_off.data(), n_elem, s);
    m->ns_off.alloc(n_elem);
    m->ns_off.upload(ns_off.data(), n_elem, s);
    SVX_HIP(hipStreamSynchronize(s));
}

// Element pass + gather. Returns SVX_OK or the first bad element (index,
// reason) through *bad_index / *bad_reason.
void system_assemble(svx_system* sys, const double* E, const double* nu, double thickness, int64_t* bad_index,
                     int* bad_reason) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    sys->E.alloc(m->n_elem);
    sys->E.upload(E, m->n_elem, s);
    sys->nu.alloc(m->n_elem);
    sys->nu.upload(nu, m->n_elem, s);
    sys->thickness = thickness;
    sys->mat_uniform = m->n_elem > 0;
    for (int64_t e = 1; e < m->n_elem && sys->mat_uniform; ++e) sys->mat_uniform = E[e] == E[0] && nu[e] == nu[0];
    sys->E0 = m->n_elem > 0 ? E[0] : 0.0;
    sys->nu0 = m->n_elem > 0 ? nu[0] : 0.0;
    sys->mf_ready = false;
    void* tmp[3];
    m->ctx->scratch.carve(s, {sizeof(double) * std::max<int64_t>(1, m->ke_total), (size_t)std::max<int64_t>(1, m->n_elem),
                              sizeof(unsigned long long)}, tmp);
    double* ke = (double*)tmp[0];
    uint8_t* reason = (uint8_t*)tmp[1];
    unsigned long long* first_bad = (unsigned long long*)tmp[2];
    unsigned long long init = ~0ull;
    SVX_HIP(hipMemcpyAsync(first_bad, &init, sizeof init, hipMemcpyHostToDevice, s));
    for (int k = 0; k < 3; ++k)
        launch_element_ke(k, m->nsel[k], m->sel[k].p, m->conn_ptr.p, m->conn.p, m->coords.p, sys->E.p, sys->nu.p,
                          thickness, m->ndof, ke, m->ke_off.p, reason, first_bad, s);
    unsigned long long fb = 0;
    SVX_HIP(hipMemcpyAsync(&fb, first_bad, sizeof fb, hipMemcpyDeviceToHost, s));
    SVX_HIP(hipStreamSynchronize(s));
    *bad_index = -1;
    *bad_reason = 0;
    if (fb != ~0ull) {
        uint8_t r = 0;
        SVX_HIP(hipMemcpy(&r, reason + fb, 1, hipMemcpyDeviceToHost));
        *bad_index = (int64_t)fb;
        *bad_reason = r;
        return;
    }
    sys->vals.alloc(m->nnz);
    launch_gather_assembly(m->n, m->ndof, m->row_ptr.p, m->col.p, m->inc_ptr.p, m->inc_elem.p, m->inc_slot.p,
                           m->conn_ptr.p, m->conn.p, m->kinds.p, ke, m->ke_off.p, sys->vals.p, s);
}

void system_set_f(svx_system* sys, const double* f) {
    svx_mesh* m = sys->mesh;
    sys->h_f.assign(f, f + m->n);
    sys->f.alloc(m->n);
    sys->f.upload(sys->h_f.data(), m->n, m->ctx->stream);
    SVX_HIP(hipStreamSynchronize(m->ctx->stream));  // a later call may rewrite h_f
}

void system_dirichlet(svx_system* sys, int64_t n_rows, const int64_t* rows) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    std::vector<uint8_t> mask(m->n, 0);
    for (int64_t i = 0; i < n_rows; ++i) {
        if (rows[i] < 0 || rows[i] >= m->n) throw std::invalid_argument("constrained row out of range");
        mask[rows[i]] = 1;
    }
    sys->cons_rows.clear();
    for (int64_t r = 0; r < m->n; ++r)
        if (mask[r]) sys->cons_rows.push_back(r);
    std::vector<int64_t> off(sys->cons_rows.size());
    int64_t tot = 0;
    sys->f_orig_cons.resize(sys->cons_rows.size());
    for (size_t i = 0; i < sys->cons_rows.size(); ++i) {
        int64_t r = sys->cons_rows[i];
        off[i] = tot;
        tot += m->hp.row_ptr[r + 1] - m->hp.row_ptr[r];
        sys->f_orig_cons[i] = sys->h_f[r];
    }
    sys->d_cons_rows.alloc(sys->cons_rows.size());
    sys->d_cons_rows.upload(sys->cons_rows.data(), sys->cons_rows.size(), s);
    sys->d_cons_off.alloc(off.size());
    sys->d_cons_off.upload(off.data(), off.size(), s);
    sys->d_cons_vals.alloc(std::max<int64_t>(1, tot));
    launch_gather_rows((int64_t)sys->cons_rows.size(), sys->d_cons_rows.p, m->row_ptr.p, sys->vals.p,
                       sys->d_cons_off.p, sys->d_cons_vals.p, s);
    void* tmp[1];
    m->ctx->scratch.carve(s, {(size_t)m->n}, tmp);
    uint8_t* dmask = (uint8_t*)tmp[0];
    SVX_HIP(hipMemcpyAsync(dmask, mask.data(), m->n, hipMemcpyHostToDevice, s));
    launch_dirichlet(m->n, m->row_ptr.p, m->col.p, dmask, sys->vals.p, sys->f.p, s);
    for (int64_t r : sys->cons_rows) sys->h_f[r] = 0.0;
    sys->dirichlet = true;
    // K changed: the preconditioner and the ELL copies are rebuilt on the next solve
    sys->dinv.release();
    sys->dinv32.release();
    sys->ell64.release();
    sys->ell32.release();
    sys->mf_mask.release();
    sys->mf_ready = false;
    SVX_HIP(hipStreamSynchronize(s));
}

bool mesh_mf_plan(svx_mesh* m, const double* coords) {
    if (m->mf_state != 0) return m->mf_state > 0;
    if (m->ndof == 3) {  // C3D8 meshes: ranks of the elements at their nodes
        bool hex = m->n_elem > 0;
        for (int64_t e = 0; e < m->n_elem && hex; ++e)
            hex = m->h_kinds[e] == SVX_C3D8 && m->h_conn_ptr[e] == 8 * e;
        int32_t width = 0;
        for (int64_t i = 0; i < m->n_nodes && hex; ++i)
            width = std::max<int32_t>(width, (int32_t)(m->hp.inc_ptr[i + 1] - m->hp.inc_ptr[i]));
        if (width > 255 || m->n_nodes * width > (int64_t)INT32_MAX / 3) hex = false;
        m->mf_state = hex ? 2 : -1;
        m->mf_kind = hex ? 8 : 0;
        if (!hex) return false;
        m->mf_rank_host.assign((size_t)m->n_elem * 8, 0);
        for (int64_t i = 0; i < m->n_nodes; ++i)  // incidences are in (element, corner) order
            for (int64_t q = m->hp.inc_ptr[i]; q < m->hp.inc_ptr[i + 1]; ++q)
                m->mf_rank_host[(size_t)m->hp.inc_elem[q] * 8 + m->hp.inc_slot[q]] = (uint8_t)(q - m->hp.inc_ptr[i]);
        m->mf_width = width;
        return true;
    }
    auto plan = std::make_unique<MfPlan>();
    build_mf_plan(m->ndof, m->n_nodes, coords, m->n_elem, m->h_kinds.data(), m->h_conn_ptr.data(), m->h_conn.data(),
                  m->hp, *plan);
    if (plan->kind == 0) {
        m->mf_state = -1;
        return false;
    }
    m->mf_kind = plan->kind;
    m->mf_patches = (int32_t)plan->patches.size();
    m->mf_cap_nodes = plan->cap_nodes;
    m->mf_cap_elem = plan->cap_elem;
    m->mf_elem_slots = plan->elem_slots;
    m->mf_plan = std::move(plan);
    m->mf_state = 2;
    return true;
}

bool mesh_mf_prepare(svx_mesh* m) {
    if (m->mf_state == 0) mesh_mf_plan(m, m->h_coords.data());
    if (m->mf_state != 2) return m->mf_state > 0;
    hipStream_t s = m->ctx->stream;
    if (m->mf_kind == 8) {
        m->mf_rank.alloc(m->mf_rank_host.size());
        m->mf_rank.upload(m->mf_rank_host.data(), m->mf_rank_host.size(), s);
        SVX_HIP(hipStreamSynchronize(s));
        std::vector<uint8_t>().swap(m->mf_rank_host);
        m->mf_state = 1;
        return true;
    }
    const MfPlan& plan = *m->mf_plan;
    m->mf_patch.alloc(plan.patches.size());
    m->mf_patch.upload(plan.patches.data(), plan.patches.size(), s);
    m->mf_nodes.alloc(plan.nodes.size());
    m->mf_nodes.upload(plan.nodes.data(), plan.nodes.size(), s);
    m->mf_econn.alloc(plan.econn.size());
    m->mf_econn.upload(plan.econn.data(), plan.econn.size(), s);
    m->mf_adj.alloc(std::max<size_t>(1, plan.adj.size()));
    m->mf_adj.upload(plan.adj.data(), plan.adj.size(), s);
    std::vector<double> xy((size_t)m->n_nodes * 2);
    for (int64_t i = 0; i < m->n_nodes; ++i) xy[2 * i] = m->h_coords[3 * i], xy[2 * i + 1] = m->h_coords[3 * i + 1];
    m->mf_xy.alloc(xy.size());
    m->mf_xy.upload(xy.data(), xy.size(), s);
    SVX_HIP(hipStreamSynchronize(s));
    m->mf_plan.reset();
    m->mf_state = 1;
    return true;
}

// Per-system preparation of the Krylov data for one precision.
void system_prepare(svx_system* sys, int precision, bool matrix_free) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n;
    (void)precision;  // FP64 solves multiply on the operator, FP32-IR solves in their FP64 phase
    if (matrix_free && !sys->mf_ready && sys->mat_uniform && sys->dirichlet &&
        mesh_mf_prepare(m)) {
        std::vector<uint8_t> mask(m->n_nodes, 0);
        const int nd = m->ndof;
        for (int64_t r : sys->cons_rows) mask[r / nd] |= (uint8_t)(1u << (r % nd));
        sys->mf_mask.alloc(m->n_nodes);
        sys->mf_mask.upload(mask.data(), m->n_nodes, s);
        if (m->mf_kind == 8) {  // unused slots must read as zero
            sys->mf_fe.alloc((size_t)m->n_nodes * m->mf_width * 3);
            SVX_HIP(hipMemsetAsync(sys->mf_fe.p, 0, sys->mf_fe.bytes(), s));
        }
        sys->mf_ready = true;
    }
    if (!sys->dinv.p) {
        sys->dinv.alloc(n);
        sys->dinv32.alloc(n);
        // the diagonal's extremes in the context's scratch: a temporary
        // buffer's free would wait for every stream of the device
        void* tmp[1];
        m->ctx->scratch.carve(s, {2 * sizeof(unsigned long long)}, tmp);
        unsigned long long* ext = (unsigned long long*)tmp[0];
        unsigned long long init[2] = {0ull, ~0ull};
        SVX_HIP(hipMemcpyAsync(ext, init, sizeof init, hipMemcpyHostToDevice, s));
        launch_diag_inv(n, m->row_ptr.p, m->col.p, sys->vals.p, sys->dinv.p, sys->dinv32.p, ext, ext + 1, s);
        unsigned long long h[2];
        SVX_HIP(hipMemcpyAsync(h, ext, sizeof h, hipMemcpyDeviceToHost, s));
        SVX_HIP(hipStreamSynchronize(s));
        double hi, lo;
        std::memcpy(&hi, &h[0], 8);
        std::memcpy(&lo, &h[1], 8);
        sys->cond = (h[1] == ~0ull) ? std::numeric_limits<double>::quiet_NaN() : hi / lo;
    }
    if (!sys->u.p) sys->u.alloc(n);
    if (!sys->scal.p) sys->scal.alloc(1);
    if (!sys->r.p) {  // FP64 vectors: the FP64 arm, and the FP32-IR fallback
        sys->r.alloc(n), sys->p.alloc(n), sys->z.alloc(n), sys->q.alloc(n), sys->p_alt.alloc(n);
    }
    const size_t ell_vals = (size_t)std::max<int64_t>(1, m->ell_blocks) * m->ndof * m->ndof;
    if (!sys->ell64.p) {  // FP64 PCG, and the FP64 fallback of FP32-IR
        sys->ell64.alloc(ell_vals);
        launch_ell_pack(m->ndof, m->n_nodes, m->ell_off.p, m->ell_nb.p, m->row_ptr.p, m->brow_ptr.p, sys->vals.p,
                        sys->ell64.p, nullptr, s);
    }
    if (precision != SVX_FP64) {
        if (!sys->ell32.p) {
            sys->ell32.alloc(ell_vals);
            launch_ell_pack(m->ndof, m->n_nodes, m->ell_off.p, m->ell_nb.p, m->row_ptr.p, m->brow_ptr.p, sys->vals.p,
                            nullptr, sys->ell32.p, s);
        }
        if (!sys->x32.p) {
            sys->x32.alloc(n), sys->r32.alloc(n), sys->p32.alloc(n), sys->z32.alloc(n), sys->q32.alloc(n);
            sys->p32_alt.alloc(n);
        }
    }
    SVX_HIP(hipStreamSynchronize(s));
}

InstDesc system_desc(svx_system* sys) {
    svx_mesh* m = sys->mesh;
    InstDesc d{};
    d.n = (int32_t)m->n;
    d.bs = m->ndof;
    d.row_ptr = m->row_ptr.p;
    d.col = m->col.p;
    d.val = sys->vals.p;
    d.brow_ptr = m->brow_ptr.p;
    d.bcol = m->bcol.p;
    d.ell_off = m->ell_off.p;
    d.ell_nb = m->ell_nb.p;
    d.ell_col = m->ell_col.p;
    d.ell_val = sys->ell64.p;
    d.ell_val32 = sys->ell32.p;
    d.dinv = sys->dinv.p;
    d.dinv32 = sys->dinv32.p;
    d.f = sys->f.p;
    d.u = sys->u.p;
    d.r = sys->r.p;
    d.p = sys->p.p;
    d.z = sys->z.p;
    d.q = sys->q.p;
    d.x32 = sys->x32.p;
    d.r32 = sys->r32.p;
    d.p32 = sys->p32.p;
    d.z32 = sys->z32.p;
    d.q32 = sys->q32.p;
    d.p_alt = sys->p_alt.p;
    d.p32_alt = sys->p32_alt.p;
    d.s = sys->scal.p;
    if (sys->mf_ready && m->mf_kind == 8) {
        d.mf_kind = 8;
        d.mf_xy = m->coords.p;
        d.mf_conn = m->conn.p;
        d.mf_rank = m->mf_rank.p;
        d.mf_width = m->mf_width;
        d.mf_fe = sys->mf_fe.p;
        d.mf_mask = sys->mf_mask.p;
        d.mf_n_elem = (int32_t)m->n_elem;
        // d3_matrix (element.rs:33-47)
        const double E = sys->E0, nu = sys->nu0;
        const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
        d.mf_c11 = lam + 2.0 * mu;
        d.mf_c12 = lam;
        d.mf_c33 = mu;
    } else if (sys->mf_ready) {
        d.mf_kind = m->mf_kind;
        d.mf_patch = m->mf_patch.p;
        d.mf_nodes = m->mf_nodes.p;
        d.mf_econn = m->mf_econn.p;
        d.mf_adj = m->mf_adj.p;
        d.mf_xy = m->mf_xy.p;
        d.mf_mask = sys->mf_mask.p;
        const double E = sys->E0, nu = sys->nu0, c = E / (1.0 - nu * nu);  // element.rs:10-17
        d.mf_c11 = c;
        d.mf_c12 = nu * c;
        d.mf_c33 = c * (1.0 - nu) / 2.0;
        d.mf_t = sys->thickness;
        d.mf_patches = m->mf_patches;
        d.mf_cap_nodes = m->mf_cap_nodes;
        d.mf_cap_elem = m->mf_cap_elem;
    }
    return d;
}

// Algorithmic HBM bytes of one solver SpMV launch on this mesh's matrix (the
// direction step fused in): the matrix values (nnz, without the ELL padding),
// one block column per block, the chunk table, one read of z and of the old
// direction, one write of the new direction and of q.
double spmv_bytes(const svx_mesh* m, int precision) {
    const double n = (double)m->n, nnz = (double)m->nnz, nblk = (double)m->nblk, ch = (double)m->n_chunks;
    if (precision == SVX_FP64) return 8.0 * nnz + 4.0 * nblk + 8.0 * ch + 32.0 * n;
    return 4.0 * nnz + 4.0 * nblk + 8.0 * ch + 16.0 * n;
}

// Algorithmic HBM bytes of one matrix-free product: coordinates, z and the
// old direction of every node read, the new direction and q written,
// connectivity (four bytes per element node) and the constrained-dof mask.
double mf_spmv_bytes(const svx_mesh* m) {
    if (m->mf_kind == 8)  // two passes: forces per (element, corner) written, read back per node
        return 24.0 * m->n_nodes + 48.0 * m->n_nodes + 40.0 * m->n_elem + 192.0 * m->n_elem  // element pass
               + 192.0 * m->n_elem + 1.0 * m->n_nodes + 96.0 * m->n_nodes;                  // node pass
    return 81.0 * (double)m->n_nodes + 4.0 * (double)m->mf_kind * (double)m->n_elem;
}

InstScal initial_scal(double tol, int64_t max_iter, double ru_delta, int64_t max_row_len) {
    InstScal sc{};
    sc.tol = tol;
    sc.max_iter = max_iter;
    sc.status = ST_RUNNING;
    sc.flags = FL_FRESH;
    sc.ru_delta = ru_delta;
    sc.floor_scale = std::sqrt((double)std::max<int64_t>(1, max_row_len)) * std::numeric_limits<double>::epsilon();
    return sc;
}

int status_to_code(const InstScal& sc) {
    if (sc.status == ST_CONVERGED) return SVX_OK;
    return SVX_E_NOT_CONVERGED;
}

// Solver set-up of system_solve without the solve: vectors, scalar state
// and the engine's tile tables for `precision`.
void system_engine_setup(svx_system* sys, double tol, int64_t max_iter, int precision, const TuneParams& tp) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    system_prepare(sys, precision, tp.matrix_free);
    InstScal sc = initial_scal(tol, max_iter, tp.ru_delta, m->max_row_len);
    sys->scal.upload(&sc, 1, s);
    std::vector<InstDesc> descs{system_desc(sys)};
    if (!sys->engine) sys->engine = std::make_unique<Engine>(s);
    EngineShape shape;
    const int bs = m->ndof;
    shape.rows_per_tile[bs] = tp.rows_per_tile;  // 0: the instance's default
    shape.unroll[bs] = tp.unroll > 0 ? tp.unroll : unroll_for(bs, precision);
    shape.matrix_free = tp.matrix_free;
    shape.xcd[bs] = tp.tile_order == SVX_TILES_XCD;
    shape.nt[bs] = streaming_loads({sys}, bs, tp, 8);
    shape.nt32[bs] = streaming_loads({sys}, bs, tp, 4);
    shape.ntv[bs] = streaming_vectors({sys}, bs, tp);
    if (tp.step_tiles != 0) shape.step_tiles = tp.step_tiles;
    sys->engine->set_instances(descs, {1}, shape);
    sys->engine_precision = precision;
    sys->engine_mf = shape.matrix_free && sys->mf_ready && precision == SVX_FP64;  // the timed SpMV's operator
}

void system_solve(svx_system* sys, double tol, int64_t max_iter, int precision, InstScal* out_scal,
                  const TuneParams& tp) {
    svx_mesh* m = sys->mesh;
    svx_ctx* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    system_engine_setup(sys, tol, max_iter, precision, tp);
    EngineOptions opt;
    opt.precision = precision;
    opt.period = tp.period;
    opt.chunk = tp.chunk;
    opt.use_graph = tp.use_graph;
    sys->engine->run(opt, ctx->h_flag, ctx->d_flag);
    sys->scal.download(out_scal, 1, s);
    SVX_HIP(hipStreamSynchronize(s));
    sys->last = *out_scal;
    sys->last_precision = precision;
    sys->solved = out_scal->status == ST_CONVERGED;
    sys->has_u = true;
    system_true_residual(sys);
}

void system_true_residual(svx_system* sys) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    const int64_t nb = true_residual_blocks(m->n);
    void* tmp[2];
    m->ctx->scratch.carve(s, {sizeof(double) * 3 * (size_t)std::max<int64_t>(1, nb), 4 * sizeof(double)}, tmp);
    double* d_out = static_cast<double*>(tmp[1]);
    launch_true_residual(m->n, m->row_ptr.p, m->col.p, sys->vals.p, sys->u.p, sys->f.p, static_cast<double*>(tmp[0]),
                         d_out, s);
    double h[3] = {0.0, 0.0, 0.0};
    SVX_HIP(hipMemcpyAsync(h, d_out, sizeof h, hipMemcpyDeviceToHost, s));
    SVX_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(h[2]);
    const double scale = std::sqrt((double)std::max<int64_t>(1, m->max_row_len)) * std::numeric_limits<double>::epsilon();
    sys->true_res = bnorm > 0.0 ? std::sqrt(h[0]) / bnorm : 0.0;
    sys->true_floor = bnorm > 0.0 ? scale * std::sqrt(h[1]) / bnorm : 0.0;
}

void system_recover(svx_system* sys, double* vm_elem, double* vm_node, int64_t n_rows, const int64_t* rows,
                    double* reactions) {
    svx_mesh* m = sys->mesh;
    hipStream_t s = m->ctx->stream;
    if (!sys->has_u) throw std::invalid_argument("no solution to recover from");
    if (n_rows > 0 && reactions && !sys->dirichlet) throw std::invalid_argument("reactions need svx_apply_dirichlet first");
    const size_t nc = sys->cons_rows.size();
    void* tmp[4];
    m->ctx->scratch.carve(s, {sizeof(double) * std::max<int64_t>(1, m->n_elem), sizeof(double) * std::max<int64_t>(1, m->ns_total),
                              sizeof(double) * std::max<int64_t>(1, m->n_nodes), sizeof(double) * std::max<size_t>(1, nc)},
                          tmp);
    double *d_vm = (double*)tmp[0], *d_ns = (double*)tmp[1], *d_vn = (double*)tmp[2], *d_ku = (double*)tmp[3];
    if (vm_elem || vm_node) {
        for (int k = 0; k < 3; ++k)
            launch_element_stress(k, m->nsel[k], m->sel[k].p, m->conn_ptr.p, m->conn.p, m->coords.p, sys->E.p,
                                  sys->nu.p, m->ndof, sys->u.p, d_vm, d_ns, m->ns_off.p, s);
        if (vm_elem) SVX_HIP(hipMemcpyAsync(vm_elem, d_vm, sizeof(double) * m->n_elem, hipMemcpyDeviceToHost, s));
        if (vm_node) {
            launch_nodal_average(m->n_nodes, m->inc_ptr.p, m->inc_elem.p, m->inc_slot.p, m->ns_off.p, d_ns, d_vn, s);
            SVX_HIP(hipMemcpyAsync(vm_node, d_vn, sizeof(double) * m->n_nodes, hipMemcpyDeviceToHost, s));
        }
    }
    std::vector<double> ku;
    if (n_rows > 0 && reactions) {
        launch_reaction_rows((int64_t)nc, sys->d_cons_rows.p, m->row_ptr.p, m->col.p, sys->d_cons_off.p,
                             sys->d_cons_vals.p, sys->u.p, d_ku, s);
        ku.resize(nc);
        if (nc) SVX_HIP(hipMemcpyAsync(ku.data(), d_ku, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
    }
    SVX_HIP(hipStreamSynchronize(s));
    if (n_rows > 0 && reactions) {
        for (int64_t i = 0; i < n_rows; ++i) {
            auto it = std::lower_bound(sys->cons_rows.begin(), sys->cons_rows.end(), rows[i]);
            if (it == sys->cons_rows.end() || *it != rows[i])
                throw std::invalid_argument("reaction row is not constrained");
            size_t q = it - sys->cons_rows.begin();
            reactions[i] = ku[q] - sys->f_orig_cons[q];
        }
    }
}

}  // namespace svx

// ======================================================================= ABI
extern "C" {

const char* svx_version(void) { return "solvix-mi355x 0.1 (gfx950)"; }

int svx_ctx_create(int device, svx_ctx** out) {
    if (!out) return SVX_E_ARGUMENT;
    *out = nullptr;
    return guarded(nullptr, [&] {
        auto c = std::make_unique<svx_ctx>();
        c->device = device;
        SVX_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        SVX_HIP(hipGetDeviceProperties(&prop, device));
        c->name = prop.gcnArchName;
        c->name += " / ";
        c->name += prop.name;
        SVX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        SVX_HIP(hipHostMalloc((void**)&c->h_flag, sizeof(int32_t) * 4, hipHostMallocDefault));
        SVX_HIP(hipMalloc((void**)&c->d_flag, sizeof(int32_t) * 4));
        noalloc_register_device(device);
        *out = c.release();
        return SVX_OK;
    });
}

void svx_ctx_destroy(svx_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int svx_device_name(svx_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !buf || len == 0) return SVX_E_ARGUMENT;
    snprintf(buf, len, "%s", ctx->name.c_str());
    return SVX_OK;
}

const char* svx_last_error(void) { return g_last_error.c_str(); }

int svx_mesh_create(svx_ctx* ctx, int ndof, int64_t n_nodes, const double* coords, int64_t n_elem,
                    const uint8_t* kinds, const int64_t* conn_ptr, const int32_t* conn_idx, svx_mesh** out,
                    svx_error* err) {
    clear_err(err);
    if (!ctx || !out || (ndof != 2 && ndof != 3) || n_nodes <= 0 || n_elem < 0 || !coords ||
        (n_elem > 0 && (!kinds || !conn_ptr || !conn_idx))) {
        set_err(err, SVX_E_ARGUMENT, "svx_mesh_create: bad argument");
        return SVX_E_ARGUMENT;
    }
    *out = nullptr;
    return guarded(err, [&] {
        if (conn_ptr[0] != 0) throw std::invalid_argument("conn_ptr[0] must be 0");
        for (int64_t e = 0; e < n_elem; ++e) {
            if (kinds[e] > SVX_C3D8) throw std::invalid_argument("unsupported element kind");
            if ((kinds[e] == SVX_C3D8) != (ndof == 3)) throw std::invalid_argument("element kind does not match ndof");
            if (conn_ptr[e + 1] - conn_ptr[e] != nodes_of(kinds[e]))
                throw std::invalid_argument("connectivity length does not match the element kind");
            for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k)
                if (conn_idx[k] < 0 || conn_idx[k] >= n_nodes) throw std::invalid_argument("node position out of range");
        }
        SVX_HIP(hipSetDevice(ctx->device));
        auto m = std::make_unique<svx_mesh>();
        m->ctx = ctx;
        mesh_build(m.get(), ndof, n_nodes, coords, n_elem, kinds, conn_ptr, conn_idx);
        *out = m.release();
        return SVX_OK;
    });
}

void svx_mesh_destroy(svx_mesh* mesh) { delete mesh; }
int64_t svx_mesh_rows(const svx_mesh* mesh) { return mesh ? mesh->n : -1; }
int64_t svx_mesh_nnz(const svx_mesh* mesh) { return mesh ? mesh->nnz : -1; }

int svx_mesh_pattern(const svx_mesh* mesh, int32_t* row_ptr, int32_t* col_idx) {
    if (!mesh) return SVX_E_ARGUMENT;
    if (row_ptr) std::copy(mesh->hp.row_ptr.begin(), mesh->hp.row_ptr.end(), row_ptr);
    if (col_idx) std::copy(mesh->hp.col.begin(), mesh->hp.col.end(), col_idx);
    return SVX_OK;
}

int svx_assemble(svx_mesh* mesh, const double* youngs, const double* poisson, double thickness, int64_t n_loads,
                 const int64_t* load_rows, const double* load_values, svx_system** out, svx_error* err) {
    clear_err(err);
    if (!mesh || !youngs || !poisson || !out || n_loads < 0 || (n_loads > 0 && (!load_rows || !load_values))) {
        set_err(err, SVX_E_ARGUMENT, "svx_assemble: bad argument");
        return SVX_E_ARGUMENT;
    }
    *out = nullptr;
    return guarded(err, [&] {
        SVX_HIP(hipSetDevice(mesh->ctx->device));
        auto sys = std::make_unique<svx_system>();
        sys->mesh = mesh;
        int64_t bad = -1;
        int why = 0;
        system_assemble(sys.get(), youngs, poisson, thickness, &bad, &why);
        if (bad >= 0) {
            char buf[256];
            snprintf(buf, sizeof buf, "element at position %lld: %s", (long long)bad, reason_text(why));
            set_err(err, SVX_E_INVALID_MODEL, buf, why, bad);
            return SVX_E_INVALID_MODEL;
        }
        std::vector<double> f(mesh->n, 0.0);
        for (int64_t l = 0; l < n_loads; ++l) {
            if (load_rows[l] < 0 || load_rows[l] >= mesh->n) throw std::invalid_argument("load row out of range");
            f[load_rows[l]] += load_values[l];
        }
        system_set_f(sys.get(), f.data());
        *out = sys.release();
        return SVX_OK;
    });
}

void svx_system_destroy(svx_system* sys) { delete sys; }

int svx_system_values(const svx_system* sys, double* vals, double* f) {
    if (!sys) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        hipStream_t s = sys->mesh->ctx->stream;
        if (vals) sys->vals.download(vals, sys->mesh->nnz, s);
        if (f) sys->f.download(f, sys->mesh->n, s);
        SVX_HIP(hipStreamSynchronize(s));
        return SVX_OK;
    });
}

int svx_system_solution(const svx_system* sys, double* u) {
    if (!sys || !u || !sys->has_u) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        hipStream_t s = sys->mesh->ctx->stream;
        sys->u.download(u, sys->mesh->n, s);
        SVX_HIP(hipStreamSynchronize(s));
        return SVX_OK;
    });
}

int svx_system_set_rhs(svx_system* sys, const double* f) {
    if (!sys || !f) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        system_set_f(sys, f);
        return SVX_OK;
    });
}

int svx_apply_dirichlet(svx_system* sys, int64_t n_rows, const int64_t* rows) {
    if (!sys || n_rows < 0 || (n_rows > 0 && !rows)) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        if (sys->dirichlet) throw std::invalid_argument("Dirichlet constraints already applied");
        system_dirichlet(sys, n_rows, rows);
        return SVX_OK;
    });
}

int svx_solve(svx_system* sys, double tol, int64_t max_iter, int precision, double* u, int64_t* iterations,
              double* residual, svx_error* err) {
    clear_err(err);
    if (!sys || (precision != SVX_FP64 && precision != SVX_FP32_IR) || !(tol > 0.0)) {
        set_err(err, SVX_E_ARGUMENT, "svx_solve: bad argument");
        return SVX_E_ARGUMENT;
    }
    return guarded(err, [&] {
        svx_mesh* m = sys->mesh;
        int64_t mi = max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * m->n);
        InstScal sc{};
        system_solve(sys, tol, mi, precision, &sc, m->ctx->tune);
        if (u) {
            sys->u.download(u, m->n, m->ctx->stream);
            SVX_HIP(hipStreamSynchronize(m->ctx->stream));
        }
        if (iterations) *iterations = sc.it;
        if (residual) *residual = sc.res;
        if (sc.status != ST_CONVERGED) {
            char buf[256];
            snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)sc.it, sc.res);
            set_err(err, SVX_E_NOT_CONVERGED, buf, 0, -1, sc.it, sc.res);
            return SVX_E_NOT_CONVERGED;
        }
        return SVX_OK;
    });
}

int svx_system_stats(const svx_system* sys, svx_solve_stats* out) {
    if (!sys || !out) return SVX_E_ARGUMENT;
    const InstScal& sc = sys->last;
    *out = svx_solve_stats{};
    out->precision = sys->last_precision;
    if (sys->last_precision < 0) return SVX_OK;
    out->iterations = sc.it;
    out->refreshes = sc.refreshes;
    out->fp64_iterations = sc.fp64_it != 0 ? sc.it + sc.fp64_it : 0;
    out->residual_evals = sc.evals;
    out->residual = sc.res;
    out->floor = sc.floor;
    out->status = sc.status == ST_CONVERGED ? 0 : sc.status == ST_BREAKDOWN ? 1 : 2;
    out->true_residual = sys->true_res;
    out->true_floor = sys->true_floor;
    out->converged_by = sc.status != ST_CONVERGED ? SVX_STOP_NONE
                        : sc.stop_rule == SVX_STOP_TRUE ? SVX_STOP_TRUE : SVX_STOP_RECURSIVE;
    out->fallback_reason = sc.fb_reason;
    return SVX_OK;
}

int svx_recover(svx_system* sys, double* vm_elem, double* vm_node, int64_t n_rows, const int64_t* rows,
                double* reactions, double* cond_estimate) {
    if (!sys || n_rows < 0 || (n_rows > 0 && (!rows || !reactions))) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        system_recover(sys, vm_elem, vm_node, n_rows, rows, reactions);
        if (cond_estimate) *cond_estimate = sys->cond;
        return SVX_OK;
    });
}

// ----------------------------------------------------------------- raw CSR
int svx_csr_create(svx_ctx* ctx, int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                   svx_csr** out) {
    if (!ctx || !out || n <= 0 || n >= INT32_MAX || !row_ptr || !col_idx || !vals) return SVX_E_ARGUMENT;
    *out = nullptr;
    return guarded(nullptr, [&] {
        if (row_ptr[0] != 0) throw std::invalid_argument("row_ptr[0] must be 0");
        int64_t mrl = 0;
        for (int64_t r = 0; r < n; ++r) {
            if (row_ptr[r + 1] < row_ptr[r]) throw std::invalid_argument("row_ptr must be non-decreasing");
            mrl = std::max<int64_t>(mrl, row_ptr[r + 1] - row_ptr[r]);
        }
        const int64_t nnz = row_ptr[n];
        for (int64_t k = 0; k < nnz; ++k)
            if (col_idx[k] < 0 || col_idx[k] >= n) throw std::invalid_argument("column index out of range");
        SVX_HIP(hipSetDevice(ctx->device));
        auto c = std::make_unique<svx_csr>();
        c->ctx = ctx;
        c->n = n;
        c->nnz = nnz;
        c->max_row_len = mrl;
        hipStream_t s = ctx->stream;
        c->row_ptr.alloc(n + 1);
        c->row_ptr.upload(row_ptr, n + 1, s);
        c->col.alloc(std::max<int64_t>(1, nnz));
        c->col.upload(col_idx, nnz, s);
        c->vals.alloc(std::max<int64_t>(1, nnz));
        c->vals.upload(vals, nnz, s);
        SVX_HIP(hipStreamSynchronize(s));
        *out = c.release();
        return SVX_OK;
    });
}

void svx_csr_destroy(svx_csr* csr) { delete csr; }

int svx_csr_spmv(svx_csr* csr, const double* x, double* y, int ordered) {
    if (!csr || !x || !y) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        hipStream_t s = csr->ctx->stream;
        DevBuf<double> dx(csr->n), dy(csr->n);
        dx.upload(x, csr->n, s);
        if (ordered)
            launch_spmv_ordered(csr->n, csr->row_ptr.p, csr->col.p, csr->vals.p, dx.p, dy.p, s);
        else
            launch_spmv_csr64(csr->n, csr->row_ptr.p, csr->col.p, csr->vals.p, dx.p, dy.p,
                              default_lanes((double)csr->nnz / csr->n), s);
        dy.download(y, csr->n, s);
        SVX_HIP(hipStreamSynchronize(s));
        return SVX_OK;
    });
}

int svx_csr_pcg(svx_csr* csr, const double* f, double tol, int64_t max_iter, int precision, double* u,
                int64_t* iterations, double* residual, svx_error* err) {
    clear_err(err);
    if (!csr || !f || (precision != SVX_FP64 && precision != SVX_FP32_IR) || !(tol > 0.0)) {
        set_err(err, SVX_E_ARGUMENT, "svx_csr_pcg: bad argument");
        return SVX_E_ARGUMENT;
    }
    return guarded(err, [&] {
        svx_ctx* ctx = csr->ctx;
        hipStream_t s = ctx->stream;
        const int64_t n = csr->n;
        int64_t mi = max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * n);
        DevBuf<double> df(n), dinv(n), du(n), dr(n), dp(n), dz(n), dq(n), dp_alt(n);
        DevBuf<float> dinv32(n), bval, x32, r32, p32, z32, q32, p32_alt;
        df.upload(f, n, s);
        launch_diag_inv(n, csr->row_ptr.p, csr->col.p, csr->vals.p, dinv.p, dinv32.p, nullptr, nullptr, s);
        InstDesc d{};
        d.n = (int32_t)n;
        d.bs = 1;
        d.row_ptr = csr->row_ptr.p;
        d.col = csr->col.p;
        d.val = csr->vals.p;
        d.dinv = dinv.p;
        d.dinv32 = dinv32.p;
        d.f = df.p;
        d.u = du.p;
        d.r = dr.p, d.p = dp.p, d.z = dz.p, d.q = dq.p, d.p_alt = dp_alt.p;
        if (precision == SVX_FP32_IR) {
            bval.alloc(std::max<int64_t>(1, csr->nnz));
            launch_fill_bsr(n, 1, csr->row_ptr.p, csr->col.p, csr->vals.p, csr->row_ptr.p, bval.p, s);
            x32.alloc(n), r32.alloc(n), p32.alloc(n), z32.alloc(n), q32.alloc(n), p32_alt.alloc(n);
            d.brow_ptr = csr->row_ptr.p;
            d.bcol = csr->col.p;
            d.bval = bval.p;
            d.x32 = x32.p, d.r32 = r32.p, d.p32 = p32.p, d.z32 = z32.p, d.q32 = q32.p, d.p32_alt = p32_alt.p;
        }
        DevBuf<InstScal> scal(1);
        InstScal sc = initial_scal(tol, mi, ctx->tune.ru_delta, csr->max_row_len);
        scal.upload(&sc, 1, s);
        d.s = scal.p;
        std::vector<InstDesc> descs{d};
        Engine eng(s);
        const TuneParams& tp = ctx->tune;
        double avg = (double)csr->nnz / n;
        EngineShape shape;
        shape.rows_per_tile[1] = tp.rows_per_tile > 0 ? tp.rows_per_tile : 256;
        shape.lanes[1] = tp.lanes > 0 ? tp.lanes : default_lanes(avg);
        eng.set_instances(descs, {1}, shape);
        EngineOptions opt;
        opt.precision = precision;
        opt.period = tp.period;
        opt.chunk = tp.chunk;
        opt.use_graph = tp.use_graph;
        eng.run(opt, ctx->h_flag, ctx->d_flag);
        scal.download(&sc, 1, s);
        if (u) du.download(u, n, s);
        SVX_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = sc.it;
        if (residual) *residual = sc.res;
        if (sc.status != ST_CONVERGED) {
            char buf[256];
            snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)sc.it, sc.res);
            set_err(err, SVX_E_NOT_CONVERGED, buf, 0, -1, sc.it, sc.res);
            return SVX_E_NOT_CONVERGED;
        }
        return SVX_OK;
    });
}


int svx_system_prepare_solver(svx_system* sys, int precision) {
    if (!sys || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        if (!sys->dirichlet) throw std::invalid_argument("svx_system_prepare_solver needs svx_apply_dirichlet first");
        system_engine_setup(sys, 1e-10, std::max<int64_t>(1000, 10 * sys->mesh->n), precision, sys->mesh->ctx->tune);
        return SVX_OK;
    });
}

int svx_system_time_spmv(svx_system* sys, int precision, int reps, double* avg_ms, double* bytes_per_launch) {
    if (!sys || reps <= 0 || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        if (!sys->engine || sys->engine_precision != precision)
            throw std::invalid_argument("solve the system with this precision first");
        svx_mesh* m = sys->mesh;
        double ms = sys->engine->time_spmv(precision, reps);
        if (avg_ms) *avg_ms = ms;
        if (bytes_per_launch) *bytes_per_launch = sys->engine_mf ? mf_spmv_bytes(m) : spmv_bytes(m, precision);
        return SVX_OK;
    });
}

int svx_time_spmv_rotation(svx_system* const* systems, int count, int precision, int reps, double* avg_ms,
                           double* bytes_per_launch) {
    if (!systems || count <= 0 || reps <= 0 || (precision != SVX_FP64 && precision != SVX_FP32_IR))
        return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        svx_ctx* ctx = nullptr;
        double bytes = 0.0;
        for (int i = 0; i < count; ++i) {
            svx_system* s = systems[i];
            if (!s || !s->engine || s->engine_precision != precision)
                throw std::invalid_argument("solve every system with this precision first");
            if (ctx && s->mesh->ctx != ctx) throw std::invalid_argument("systems of different contexts");
            ctx = s->mesh->ctx;
            bytes += s->engine_mf ? mf_spmv_bytes(s->mesh) : spmv_bytes(s->mesh, precision);
        }
        for (int i = 0; i < count; ++i) systems[i]->engine->spmv_prepare();
        for (int i = 0; i < count; ++i) systems[i]->engine->spmv_launch(precision);  // warm
        hipEvent_t a, b;
        SVX_HIP(hipEventCreate(&a));
        SVX_HIP(hipEventCreate(&b));
        SVX_HIP(hipEventRecord(a, ctx->stream));
        for (int r = 0; r < reps; ++r)
            for (int i = 0; i < count; ++i) systems[i]->engine->spmv_launch(precision);
        SVX_HIP(hipEventRecord(b, ctx->stream));
        SVX_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        SVX_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (avg_ms) *avg_ms = (double)ms / ((double)reps * count);
        if (bytes_per_launch) *bytes_per_launch = bytes / count;
        return SVX_OK;
    });
}

int svx_system_time_phases(svx_system* sys, int precision, int reps, double* avg_ms, double* bytes_per_launch) {
    if (!sys || reps <= 0 || !avg_ms || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded(nullptr, [&] {
        if (!sys->engine || sys->engine_precision != precision)
            throw std::invalid_argument("solve the system with this precision first");
        svx_mesh* m = sys->mesh;
        sys->engine->time_phases(precision, reps, avg_ms);
        if (bytes_per_launch) {
            const double n = (double)m->n, nnz = (double)m->nnz;
            const double nblk = (double)m->nblk, ch = (double)m->n_chunks;
            bytes_per_launch[0] = sys->engine_mf ? mf_spmv_bytes(m) : spmv_bytes(m, precision);
            if (precision == SVX_FP64) {
                bytes_per_launch[1] = 64.0 * n;  // p q r u dinv read, u r z written
                bytes_per_launch[2] = 0.0;
                bytes_per_launch[3] = 0.0;
            } else {
                bytes_per_launch[1] = 32.0 * n;  // p q r x dinv read, x r z written (FP32)
                bytes_per_launch[2] = 12.0 * n;  // period end: z p read, p written
                // fold (u, x32 read+write) + FP64 residual (ELL values and columns, u, f, dinv32; r32 z32
                // written) + repoint (z32 p32 read, p32 written)
                bytes_per_launch[3] = 24.0 * n + (8.0 * nnz + 4.0 * nblk + 8.0 * ch + 28.0 * n) + 12.0 * n;
            }
        }
        return SVX_OK;
    });
}

// --------------------------------------------------------------- tuning knobs
int svx_set_tuning(svx_ctx* ctx, int lanes, int unroll, int rows_per_tile, int period, int chunk, int use_graph,
                   double ru_delta) {
    auto lanes_ok = [](int l) { return l == 0 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64; };
    if (!ctx || !lanes_ok(lanes) || unroll < 0 || unroll > kMaxUnroll || rows_per_tile < 0 || period < 0 ||
        chunk < 0 || ru_delta < 0.0 || ru_delta >= 1.0)
        return SVX_E_ARGUMENT;
    TuneParams& t = ctx->tune;
    t.lanes = lanes;
    t.unroll = unroll;
    t.rows_per_tile = rows_per_tile;
    if (period > 0) t.period = period;
    if (chunk > 0) t.chunk = chunk;
    t.use_graph = use_graph != 0;
    if (ru_delta > 0) t.ru_delta = ru_delta;
    return SVX_OK;
}

int svx_set_tile_order(svx_ctx* ctx, int order) {
    if (!ctx || (order != SVX_TILES_XCD && order != SVX_TILES_LINEAR)) return SVX_E_ARGUMENT;
    ctx->tune.tile_order = order;
    return SVX_OK;
}

int svx_set_step_kernel_tiles(svx_ctx* ctx, int32_t tiles) {
    if (!ctx) return SVX_E_ARGUMENT;
    ctx->tune.step_tiles = tiles;
    return SVX_OK;
}

int svx_set_multi_group(svx_ctx* ctx, int32_t nodes) {
    if (!ctx || !(nodes <= 0 || nodes == 4 || nodes == 8)) return SVX_E_ARGUMENT;
    ctx->tune.multi_group = nodes < 0 ? -1 : nodes;
    return SVX_OK;
}

int svx_set_load_hint(svx_ctx* ctx, int hint) {
    if (!ctx || (hint != SVX_LOADS_AUTO && hint != SVX_LOADS_CACHED && hint != SVX_LOADS_STREAMING))
        return SVX_E_ARGUMENT;
    ctx->tune.load_hint = hint;
    return SVX_OK;
}

int svx_set_operator(svx_ctx* ctx, int op) {
    if (!ctx || (op != SVX_OP_ASSEMBLED && op != SVX_OP_MATRIX_FREE)) return SVX_E_ARGUMENT;
    ctx->tune.matrix_free = op == SVX_OP_MATRIX_FREE;
    return SVX_OK;
}

int svx_mf_plan_check(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const uint8_t* kinds,
                      const int64_t* conn_ptr, const int32_t* conn_idx, int64_t* stats, double* max_rel_diff) {
    if (ndof < 1 || n_nodes <= 0 || n_elem < 0 || !coords || !kinds || !conn_ptr || !conn_idx || !stats)
        return SVX_E_ARGUMENT;
    try {
        HostPattern hp;
        build_pattern(ndof, n_nodes, n_elem, conn_ptr, conn_idx, hp);
        MfPlan plan;
        build_mf_plan(ndof, n_nodes, coords, n_elem, kinds, conn_ptr, conn_idx, hp, plan);
        stats[0] = plan.kind;
        stats[1] = (int64_t)plan.patches.size();
        stats[2] = plan.cap_nodes;
        stats[3] = plan.cap_elem;
        stats[4] = plan.elem_slots;
        if (plan.kind == 0) {
            g_last_error = plan.why;
            return SVX_OK;
        }
        const double d = verify_mf_plan(plan, n_nodes, coords, n_elem, conn_ptr, conn_idx);
        if (max_rel_diff) *max_rel_diff = d;
        return SVX_OK;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SVX_E_INVALID_MODEL;
    }
}

int svx_blk_plan_check(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const int64_t* conn_ptr,
                       const int32_t* conn_idx, int32_t group, int64_t* stats, double* max_rel_diff) {
    if (ndof < 1 || n_nodes <= 0 || n_elem < 0 || !coords || !conn_ptr || !conn_idx || !stats ||
        !(group == 0 || group == 4 || group == 8))
        return SVX_E_ARGUMENT;
    try {
        HostPattern hp;
        build_pattern(ndof, n_nodes, n_elem, conn_ptr, conn_idx, hp);
        BlkPlan plan;
        build_blk_plan(n_nodes, coords, hp, group > 0 ? group : kGroupDefault, plan);
        stats[0] = (int64_t)(plan.grp.size() / 4);
        stats[1] = plan.cols;
        stats[2] = (int64_t)hp.bcol.size();
        if (*plan.why) {
            stats[0] = 0;
            g_last_error = plan.why;
            return SVX_OK;
        }
        const double d = verify_blk_plan(plan, hp, 3);
        if (max_rel_diff) *max_rel_diff = d;
        return SVX_OK;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SVX_E_INVALID_MODEL;
    }
}

int svx_system_operator(const svx_system* sys) {
    if (!sys) return -1;
    return sys->mf_ready ? SVX_OP_MATRIX_FREE : SVX_OP_ASSEMBLED;
}

// ------------------------------------------------------------ whole model
int64_t svx_model_reaction_count(const svx_model_desc* m) {
    if (!m) return -1;
    int64_t c = 0;
    for (int32_t b = 0; b < m->n_bcs; ++b)
        c += (m->bc_node_ptr[b + 1] - m->bc_node_ptr[b]) * (m->bc_dof_ptr[b + 1] - m->bc_dof_ptr[b]);
    return c;
}

int svx_solve_model(svx_ctx* ctx, const svx_model_desc* md, int method, int precision, svx_solution* out,
                    svx_error* err) {
    clear_err(err);
    if (!ctx || !md || !out) {
        set_err(err, SVX_E_ARGUMENT, "svx_solve_model: bad argument");
        return SVX_E_ARGUMENT;
    }
    if (method == SVX_METHOD_LDLT) {
        set_err(err, SVX_E_ARGUMENT, "Method::Ldlt is not on the device path");
        return SVX_E_ARGUMENT;
    }
    if (precision != SVX_FP64 && precision != SVX_FP32_IR) {
        set_err(err, SVX_E_ARGUMENT, "unknown precision");
        return SVX_E_ARGUMENT;
    }
    if ((out->reaction_node || out->reaction_dof || out->reaction_value) &&
        svx_model_reaction_count(md) > out->reaction_capacity) {
        set_err(err, SVX_E_ARGUMENT, "svx_solve_model: reaction arrays smaller than the model's reaction count");
        return SVX_E_ARGUMENT;
    }
    return guarded(err, [&]() -> int {
        SVX_HIP(hipSetDevice(ctx->device));
        PreparedModel pm;
        int rc = prepare_model(md, pm, err);
        if (rc != SVX_OK) return rc;
        double t0 = now_ms();
        svx_mesh mesh;
        mesh.ctx = ctx;
        mesh_build(&mesh, pm.ndof, md->n_nodes, md->coords, md->n_elem, md->kinds, md->conn_ptr, pm.conn_pos.data());
        svx_system sys;
        sys.mesh = &mesh;
        int64_t bad = -1;
        int why = 0;
        system_assemble(&sys, pm.E.data(), pm.nu.data(), md->thickness, &bad, &why);
        if (bad >= 0) {
            char buf[256];
            snprintf(buf, sizeof buf, "element %u: %s", md->elem_ids[bad], reason_text(why));
            set_err(err, SVX_E_INVALID_MODEL, buf, why, bad);
            return SVX_E_INVALID_MODEL;
        }
        if (!pm.load_error.empty()) {
            set_err(err, SVX_E_INVALID_MODEL, pm.load_error);
            return SVX_E_INVALID_MODEL;
        }
        if (!pm.bc_error.empty()) {
            set_err(err, SVX_E_INVALID_MODEL, pm.bc_error);
            return SVX_E_INVALID_MODEL;
        }
        system_set_f(&sys, pm.f.data());
        double t1 = now_ms();
        system_dirichlet(&sys, (int64_t)pm.cons_rows.size(), pm.cons_rows.data());
        const int64_t n = mesh.n;
        InstScal sc{};
        system_solve(&sys, 1e-10, std::max<int64_t>(1000, 10 * n), precision, &sc, ctx->tune);
        double t2 = now_ms();
        out->iterations = sc.it;
        out->final_residual = sc.res;
        out->ndof = pm.ndof;
        out->solver_method = 0;
        if (sc.status != ST_CONVERGED) {
            char buf[256];
            snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)sc.it, sc.res);
            set_err(err, SVX_E_NOT_CONVERGED, buf, 0, -1, sc.it, sc.res);
            return SVX_E_NOT_CONVERGED;
        }
        std::vector<double> u(n);
        sys.u.download(u.data(), n, ctx->stream);
        SVX_HIP(hipStreamSynchronize(ctx->stream));
        if (out->displacements)
            for (int64_t i = 0; i < md->n_nodes; ++i) {
                out->displacements[i * 3] = u[i * pm.ndof];
                out->displacements[i * 3 + 1] = u[i * pm.ndof + 1];
                out->displacements[i * 3 + 2] = pm.ndof == 3 ? u[i * pm.ndof + 2] : 0.0;
            }
        std::vector<double> react(pm.cons_rows.size());
        system_recover(&sys, out->von_mises, out->nodal_von_mises, (int64_t)pm.cons_rows.size(), pm.cons_rows.data(),
                       react.data());
        for (size_t i = 0; i < pm.cons_rows.size(); ++i) {
            if (out->reaction_node) out->reaction_node[i] = pm.cons_node[i];
            if (out->reaction_dof) out->reaction_dof[i] = pm.cons_dof[i];
            if (out->reaction_value) out->reaction_value[i] = react[i];
        }
        out->cond_estimate = sys.cond;
        double t3 = now_ms();
        out->assemble_ms = t1 - t0;
        out->solve_ms = t2 - t1;
        out->recover_ms = t3 - t2;
        return SVX_OK;
    });
}

}  // extern "C"
