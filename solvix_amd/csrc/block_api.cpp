// svx_solve_multi: many load cases on one assembled, constrained system.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "api_internal.hpp"
#include "block.hpp"

using namespace svx;

namespace {

// FP32-IR over the load cases of a mesh without a group plan (scalar
// systems, or the gathering SpMV selected): each case is an instance of the
// Krylov engine whose descriptor shares the system's matrix copies (CSR,
// FP64 and FP32 ELL, preconditioner) and owns its vectors, so every case runs
// the FP32-IR recurrence of a single solve (FP32 inner iterations, FP64
// residual replacement and final test, FP64 finish from the evaluation floor
// or a stall) with its own scalars, in one launch per phase. With a group
// plan the block solve runs an FP32 block phase instead (blk32_*).
int solve_multi_ir32(svx_system* sys, int32_t R, const double* F, double tol, int64_t max_iter, double* U,
                     int64_t* iterations, double* residuals, int32_t* converged, double* elapsed_ms, svx_error* err) {
    svx_mesh* m = sys->mesh;
    svx_ctx* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = m->n;
    hipEvent_t e0, e1;
    SVX_HIP(hipEventCreate(&e0));
    SVX_HIP(hipEventCreate(&e1));
    system_prepare(sys, SVX_FP32_IR);
    const size_t nr = (size_t)n * R;
    DevBuf<double> f(nr), u(nr), r(nr), p(nr), z(nr), q(nr), p_alt(nr);
    DevBuf<float> x32(nr), r32(nr), p32(nr), z32(nr), q32(nr), p32_alt(nr);
    DevBuf<InstScal> scal(R);
    std::vector<double> fh(F, F + nr);
    for (int32_t k = 0; k < R; ++k)
        for (int64_t row : sys->cons_rows) fh[(size_t)k * n + row] = 0.0;  // apply_dirichlet (solver.rs:59-61)
    const int64_t mi = max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * n);
    std::vector<InstScal> init(R, initial_scal(tol, mi, ctx->tune.ru_delta, m->max_row_len));
    std::vector<InstDesc> descs(R);
    const InstDesc base = system_desc(sys);
    for (int32_t k = 0; k < R; ++k) {
        const size_t o = (size_t)k * n;
        InstDesc d = base;
        d.f = f.p + o, d.u = u.p + o, d.r = r.p + o, d.p = p.p + o, d.z = z.p + o, d.q = q.p + o;
        d.p_alt = p_alt.p + o;
        d.x32 = x32.p + o, d.r32 = r32.p + o, d.p32 = p32.p + o, d.z32 = z32.p + o, d.q32 = q32.p + o;
        d.p32_alt = p32_alt.p + o;
        d.s = scal.p + k;
        d.mf_kind = 0;  // the FP64 finish multiplies the assembled K
        descs[k] = d;
    }
    SVX_HIP(hipEventRecord(e0, s));
    f.upload(fh.data(), nr, s);
    scal.upload(init.data(), R, s);
    Engine eng(s);
    eng.set_instances(descs, std::vector<char>(R, 1), batch_shape({sys}, ctx->tune, SVX_FP32_IR));
    EngineOptions opt;
    opt.precision = SVX_FP32_IR;
    opt.period = ctx->tune.period;
    opt.chunk = ctx->tune.chunk;
    opt.use_graph = ctx->tune.use_graph;
    eng.run(opt, ctx->h_flag, ctx->d_flag);
    if (U) u.download(U, nr, s);
    SVX_HIP(hipEventRecord(e1, s));
    std::vector<InstScal> out(R);
    scal.download(out.data(), R, s);
    SVX_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    SVX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (elapsed_ms) *elapsed_ms = ms;
    int failed = 0;
    for (int32_t k = 0; k < R; ++k) {
        if (iterations) iterations[k] = out[k].it;
        if (residuals) residuals[k] = out[k].res;
        if (converged) converged[k] = out[k].status == ST_CONVERGED;
        failed += out[k].status != ST_CONVERGED;
    }
    if (failed) {
        if (err) {
            err->code = SVX_E_NOT_CONVERGED;
            snprintf(err->message, sizeof err->message, "%d of %d load cases did not converge", failed, R);
        }
        return SVX_E_NOT_CONVERGED;
    }
    return SVX_OK;
}

// The node-group plan of the block SpMV for the group size in force (built
// once per mesh and size). Returns the group size, or 0 when the mesh takes
// the gathering SpMV.
int32_t ensure_blk_plan(svx_mesh* m, const TuneParams& t, hipStream_t s) {
    if (t.multi_group < 0) return 0;
    const int group = t.multi_group > 0 ? t.multi_group : kGroupDefault;
    if (m->blk_group != group) {
        BlkPlan B;
        build_blk_plan(m->n_nodes, m->h_coords.data(), m->hp, group, B);
        m->blk_grp.release(), m->blk_nodes.release(), m->blk_ucol.release();
        m->blk_pk_src.release(), m->blk_pk_stride.release();
        m->blk_groups = 0;
        m->blk_group = group;
        if (*B.why) return 0;
        m->blk_grp.alloc(B.grp.size());
        m->blk_nodes.alloc(B.nodes.size());
        m->blk_ucol.alloc(B.ucol.size());
        m->blk_pk_src.alloc(B.pk_src.size());
        m->blk_pk_stride.alloc(B.pk_stride.size());
        m->blk_grp.upload(B.grp.data(), B.grp.size(), s);
        m->blk_nodes.upload(B.nodes.data(), B.nodes.size(), s);
        m->blk_ucol.upload(B.ucol.data(), B.ucol.size(), s);
        m->blk_pk_src.upload(B.pk_src.data(), B.pk_src.size(), s);
        m->blk_pk_stride.upload(B.pk_stride.data(), B.pk_stride.size(), s);
        SVX_HIP(hipStreamSynchronize(s));  // the host arrays go out of scope
        m->blk_groups = (int32_t)(B.grp.size() / 4);
    }
    return m->blk_groups > 0 ? group : 0;
}

// Device state of one FP64 block solve over R load cases: packed loads,
// vectors, the launch shape (group SpMV when the mesh has a plan) and the
// per-case scalars.
struct BlockWork {
    DevBuf<double> Fin, Fp, Uv, Rv, P, P2, Q, partials, l1, pval;
    DevBuf<uint8_t> dmask;
    DevBuf<BlockScal> S;
    BlockArgs a{};
    BlockWork(svx_system* sys, int32_t R, double tol, int64_t max_iter) {
        svx_mesh* m = sys->mesh;
        svx_ctx* ctx = m->ctx;
        hipStream_t s = ctx->stream;
        const int32_t n = (int32_t)m->n;
        system_prepare(sys, SVX_FP64);
        const size_t nr = (size_t)n * R;
        for (DevBuf<double>* b : {&Fin, &Fp, &Uv, &Rv, &P, &P2, &Q}) b->alloc(nr);
        std::vector<uint8_t> mask(n, 0);
        for (int64_t r : sys->cons_rows) mask[r] = 1;
        dmask.alloc(n);
        dmask.upload(mask.data(), n, s);
        a.n = n;
        a.R = R;
        a.bs = m->ndof > 1 ? m->ndof : 1;
        a.xcd = ctx->tune.tile_order == SVX_TILES_XCD;
        // gathering SpMV (no patch plan): 48 rows = 16 / 24 nodes, 4 / 6 per
        // wavefront (measured on the 64-case 61,347-DOF C3D8 solve: 140.0 ms
        // against 147.1 ms with 96-row tiles and 164.6 ms with 24-row tiles).
        // svx_set_tuning's rows per tile override, rounded to whole nodes.
        const int rpt = ctx->tune.rows_per_tile > 0 ? ctx->tune.rows_per_tile : (a.bs > 1 ? 48 : 64);
        a.rows_per_tile = (rpt + a.bs - 1) / a.bs * a.bs;
        a.group = ensure_blk_plan(m, ctx->tune, s);
        a.n_groups = a.group > 0 ? m->blk_groups : 0;
        a.n_tiles = a.group > 0 ? (a.n_groups + 3) / 4 : (n + a.rows_per_tile - 1) / a.rows_per_tile;
        a.grp = m->blk_grp.p, a.gnodes = m->blk_nodes.p, a.ucol = m->blk_ucol.p;
        if (a.group > 0) {
            pval.alloc((size_t)m->nblk * a.bs * a.bs);
            blk_pack_values(sys->vals.p, m->blk_pk_src.p, m->blk_pk_stride.p, m->nblk, a.bs, pval.p, s);
            a.pval = pval.p;
        }
        // vector kernels (initialisation, update) on tiles of twice the
        // gathering SpMV's rows
        a.vrows_per_tile = 2 * a.rows_per_tile;
        a.n_vtiles = (n + a.vrows_per_tile - 1) / a.vrows_per_tile;
        a.row_ptr = m->row_ptr.p;
        a.brow_ptr = m->brow_ptr.p;
        a.bcol = m->bcol.p;
        a.col = m->col.p;
        a.val = sys->vals.p;
        a.dinv = sys->dinv.p;
        a.F = Fp.p;
        a.U = Uv.p, a.Rv = Rv.p, a.P = P.p, a.P2 = P2.p, a.Q = Q.p;
        partials.alloc((size_t)std::max(a.n_tiles, a.n_vtiles) * 2 * kMaxRhs);
        l1.alloc((size_t)2 * kReduceGroups * kMaxRhs);
        a.partials = partials.p;
        a.l1 = l1.p;
        S.alloc(1);
        BlockScal hs{};
        hs.tol = tol;
        hs.max_iter = max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * (int64_t)n);
        S.upload(&hs, 1, s);
        a.S = S.p;
        SVX_HIP(hipStreamSynchronize(s));  // the host arrays go out of scope
    }
};

// Runs `body` (one chunk of iterations, enqueued on s) as a captured graph
// when the context allows graphs, else kernel by kernel, until no load case
// runs; the host polls the running count once per chunk.
template <typename Body>
void run_chunks(svx_ctx* ctx, const BlockArgs& a, hipStream_t s, Body body) {
    hipGraphExec_t exec = nullptr;
    if (ctx->tune.use_graph) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            bool ok = true;
            try {
                body();
            } catch (const DeviceError&) {
                ok = false;
            }
            if (hipStreamEndCapture(s, &g) == hipSuccess && g && ok) {
                if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) exec = nullptr;
            }
            if (g) (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
    }
    try {
        for (;;) {
            if (exec)
                SVX_HIP(hipGraphLaunch(exec, s));
            else
                body();
            blk_count_running(a, ctx->d_flag, s);
            SVX_HIP(hipMemcpyAsync(ctx->h_flag, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SVX_HIP(hipStreamSynchronize(s));
            if (*ctx->h_flag == 0) break;
        }
    } catch (...) {
        if (exec) (void)hipGraphExecDestroy(exec);
        throw;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
}

// FP64 block iterations until every load case has stopped.
void run_fp64(svx_ctx* ctx, const BlockArgs& a, hipStream_t s) {
    const int chunk = (ctx->tune.chunk + 1) / 2 * 2;  // even: every chunk starts on the same direction buffer
    run_chunks(ctx, a, s, [&] {
        for (int i = 0; i < chunk; ++i) blk_iteration(a, i, s);
    });
}

// FP32-IR refresh interval of the block solve (iterations between true
// residuals; the single solve's FP32-IR checks at least this often too).
constexpr int kBlockRefresh = 32;

// Per-case results of a block solve (scalars downloaded) and the error.
int finish_multi(const BlockScal& hs, int32_t R, int64_t* iterations, double* residuals, int32_t* converged,
                 svx_error* err) {
    int failed = 0;
    for (int j = 0; j < R; ++j) {
        if (iterations) iterations[j] = hs.it[j];
        if (residuals) residuals[j] = hs.res[j];
        if (converged) converged[j] = hs.status[j] == ST_CONVERGED;
        failed += hs.status[j] != ST_CONVERGED;
    }
    if (failed) {
        if (err) {
            err->code = SVX_E_NOT_CONVERGED;
            snprintf(err->message, sizeof err->message, "%d of %d load cases did not converge", failed, R);
        }
        return SVX_E_NOT_CONVERGED;
    }
    return SVX_OK;
}

}  // namespace

extern "C" {

int svx_solve_multi(svx_system* sys, int32_t nrhs, const double* F, double tol, int64_t max_iter, int precision,
                    double* U, int64_t* iterations, double* residuals, int32_t* converged, double* elapsed_ms,
                    svx_error* err) {
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->index = -1;
    }
    if (!sys || nrhs <= 0 || nrhs > kMaxRhs || !F || !(tol > 0.0) || !sys->dirichlet) return SVX_E_ARGUMENT;
    if (precision != SVX_FP64 && precision != SVX_FP32_IR) return SVX_E_ARGUMENT;
    if (precision == SVX_FP32_IR) {
        try {
            // without a group plan each case runs the single FP32-IR solve
            if (ensure_blk_plan(sys->mesh, sys->mesh->ctx->tune, sys->mesh->ctx->stream) == 0)
                return solve_multi_ir32(sys, nrhs, F, tol, max_iter, U, iterations, residuals, converged, elapsed_ms,
                                        err);
        } catch (const std::exception& e) {
            if (err) {
                err->code = SVX_E_DEVICE;
                snprintf(err->message, sizeof err->message, "%s", e.what());
            }
            return SVX_E_DEVICE;
        }
    }
    try {
        svx_mesh* m = sys->mesh;
        svx_ctx* ctx = m->ctx;
        hipStream_t s = ctx->stream;
        const int32_t n = (int32_t)m->n, R = nrhs;
        hipEvent_t e0, e1;
        SVX_HIP(hipEventCreate(&e0));
        SVX_HIP(hipEventCreate(&e1));
        BlockWork W(sys, R, tol, max_iter);
        const size_t nr = (size_t)n * R;
        // FP32-IR (group plan in force): FP32 copies of the packed values and
        // the FP32 phase's vectors
        DevBuf<float> pval32, X32, R32, P32, P32b, Q32;
        BlockArgs& a = W.a;
        if (precision == SVX_FP32_IR) {
            pval32.alloc((size_t)m->nblk * a.bs * a.bs);
            blk_pack_values32(sys->vals.p, m->blk_pk_src.p, m->blk_pk_stride.p, m->nblk, a.bs, pval32.p, s);
            for (DevBuf<float>* b : {&X32, &R32, &P32, &P32b, &Q32}) b->alloc(nr);
            a.pval32 = pval32.p;
            a.X32 = X32.p, a.R32 = R32.p, a.P32 = P32.p, a.P32b = P32b.p, a.Q32 = Q32.p;
        }
        SVX_HIP(hipEventRecord(e0, s));
        W.Fin.upload(F, nr, s);
        blk_pack(W.Fin.p, W.Fp.p, n, R, W.dmask.p, s);
        DevBuf<double>& Fin = W.Fin;
        DevBuf<double>& Uv = W.Uv;
        DevBuf<BlockScal>& S = W.S;
        BlockScal hs{};
        if (precision == SVX_FP32_IR) {
            blk32_init(a, s);
            run_chunks(ctx, a, s, [&] {
                for (int i = 0; i < kBlockRefresh; ++i) blk32_iteration(a, i, s);
                blk32_refresh(a, s);
            });
            blk_warm_start(a, s);
        } else {
            blk_init(a, s);
        }
        run_fp64(ctx, a, s);
        blk_unpack(Uv.p, Fin.p, n, R, s);
        if (U) Fin.download(U, nr, s);
        SVX_HIP(hipEventRecord(e1, s));
        S.download(&hs, 1, s);
        SVX_HIP(hipStreamSynchronize(s));
        float ms = 0.0f;
        SVX_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (elapsed_ms) *elapsed_ms = ms;
        return finish_multi(hs, R, iterations, residuals, converged, err);
    } catch (const std::exception& e) {
        if (err) {
            err->code = SVX_E_DEVICE;
            snprintf(err->message, sizeof err->message, "%s", e.what());
        }
        return SVX_E_DEVICE;
    }
}

int svx_multi_spmv_timing(svx_system* sys, int32_t nrhs, int32_t reps, double* avg_us, double* bytes) {
    if (!sys || nrhs <= 0 || nrhs > kMaxRhs || reps <= 0 || !avg_us || !sys->dirichlet) return SVX_E_ARGUMENT;
    try {
        svx_mesh* m = sys->mesh;
        hipStream_t s = m->ctx->stream;
        const int32_t n = (int32_t)m->n;
        BlockWork W(sys, nrhs, 1e-10, 0);
        // seeded load cases: unit loads on every free row, scaled per case
        std::vector<double> F((size_t)n * nrhs);
        for (int32_t k = 0; k < nrhs; ++k)
            for (int32_t i = 0; i < n; ++i) F[(size_t)k * n + i] = 1.0 + 0.01 * k + 1e-4 * (i % 97);
        W.Fin.upload(F.data(), F.size(), s);
        blk_pack(W.Fin.p, W.Fp.p, n, nrhs, W.dmask.p, s);
        blk_init(W.a, s);
        blk_spmv_only(W.a, s);  // warm; with a group plan it also writes p
        hipEvent_t e0, e1;
        SVX_HIP(hipEventCreate(&e0));
        SVX_HIP(hipEventCreate(&e1));
        SVX_HIP(hipEventRecord(e0, s));
        for (int32_t r = 0; r < reps; ++r) blk_product_only(W.a, s);
        SVX_HIP(hipEventRecord(e1, s));
        SVX_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        SVX_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_us = 1e3 * ms / reps;
        // group SpMV: matrix values, block columns, p' in and q out; the
        // gathering SpMV also reads r, the previous direction and the
        // diagonal and writes p'
        if (bytes)
            *bytes = W.a.group > 0 ? 8.0 * (double)m->nnz + 4.0 * (double)m->nblk + 16.0 * (double)n * nrhs
                                   : 8.0 * (double)m->nnz + 4.0 * (double)m->nblk + 32.0 * (double)n * nrhs +
                                         8.0 * (double)n;
        return SVX_OK;
    } catch (const std::exception&) {
        return SVX_E_DEVICE;
    }
}

}  // extern "C"
