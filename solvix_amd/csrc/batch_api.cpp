// Batched solves: many assembled, constrained systems (uneven sizes, 2D and
// 3D together) solved by one engine — every phase of an iteration is one
// launch over the tiles of all running instances, chunks of iterations are
// captured in one hipGraph, and each instance stops on its own convergence
// test (per-instance status masks). Results are per instance and identical to
// solving the instance alone: every reduction is per instance and in a fixed
// order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>

#include "api_internal.hpp"

using namespace svx;

// One of the engines a batch solve runs side by side: its own stream (lane 0
// uses the context's), tile tables and running-count flags.
struct BatchLane {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::unique_ptr<Engine> engine;
    int32_t* h_flag = nullptr;
    int32_t* d_flag = nullptr;
    ~BatchLane() {
        engine.reset();
        if (h_flag) (void)hipHostFree(h_flag);
        if (d_flag) (void)hipFree(d_flag);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

struct svx_batch {
    svx_ctx* ctx = nullptr;
    std::vector<svx_system*> sys;
    std::vector<double> tol;
    std::vector<int64_t> max_iter;
    DevBuf<InstScal> scal;
    std::vector<InstScal> h_scal;
    std::unique_ptr<Engine> engine;
    std::vector<std::unique_ptr<BatchLane>> lanes;  // svx_batch_solve with more than one stream
    int precision = -1;
    EngineShape tuned_shape;   // from svx_batch_autotune
    int tuned_precision = -1;  // precision the tuned shape belongs to, -1 = none
    bool matrix_free = false;  // the last solve's FP64 products used the matrix-free operator
};

namespace svx {

int streaming_loads(const std::vector<svx_system*>& systems, int bs, const TuneParams& tp, int value_bytes) {
    if (tp.load_hint != SVX_LOADS_AUTO) return tp.load_hint == SVX_LOADS_STREAMING;
    int64_t bytes = 0;
    // the FP32 SpMV of FP32-IR multiplies the assembled FP32 copy whatever the operator
    for (const svx_system* s : systems)
        if (s->mesh->ndof == bs && bs > 1 && (value_bytes == 4 || !(tp.matrix_free && s->mf_ready)))
            bytes += s->mesh->ell_blocks * bs * bs * (int64_t)value_bytes;
    return bytes > kStreamBytes;
}

int streaming_vectors(const std::vector<svx_system*>& systems, int bs, const TuneParams& tp) {
    if (tp.load_hint != SVX_LOADS_AUTO) return tp.load_hint == SVX_LOADS_STREAMING;
    int64_t bytes = 0;
    for (const svx_system* s : systems)
        if (s->mesh->ndof == bs) bytes += s->mesh->n * int64_t(64);
    return bytes > kStreamBytes;
}

// Launch shape for a set of systems: the per-block-size shape every solve
// uses, so batched results equal solo ones bit for bit.
EngineShape batch_shape(const std::vector<svx_system*>& systems, const TuneParams& tp, int precision) {
    EngineShape sh;
    for (int bs = 1; bs <= 3; ++bs) {
        sh.rows_per_tile[bs] = tp.rows_per_tile;  // 0: each instance's default
        sh.lanes[bs] = tp.lanes > 0 ? tp.lanes : 16;  // CSR rows of the FP64 residual and warm start
        sh.unroll[bs] = tp.unroll > 0 ? tp.unroll : unroll_for(bs, precision);
        sh.xcd[bs] = tp.tile_order == SVX_TILES_XCD;
        sh.nt[bs] = streaming_loads(systems, bs, tp, 8);
        sh.nt32[bs] = streaming_loads(systems, bs, tp, 4);
        sh.ntv[bs] = streaming_vectors(systems, bs, tp);
    }
    if (tp.step_tiles != 0) sh.step_tiles = tp.step_tiles;
    sh.matrix_free = tp.matrix_free;  // FP32-IR: the operator of its FP64 phase
    return sh;
}

}  // namespace svx

namespace {
template <typename F>
int guarded_b(svx_error* err, F&& fn) {
    try {
        return fn();
    } catch (const DeviceError& e) {
        if (err) {
            err->code = SVX_E_DEVICE;
            snprintf(err->message, sizeof err->message, "%s", e.what());
        }
        return SVX_E_DEVICE;
    } catch (const std::exception& e) {
        if (err) {
            err->code = SVX_E_ARGUMENT;
            snprintf(err->message, sizeof err->message, "%s", e.what());
        }
        return SVX_E_ARGUMENT;
    }
}
}  // namespace

namespace {
// Splits the batch over K engines on K streams: instances in decreasing size,
// dealt to the lanes back and forth (0..K-1, K-1..0, ...) so each lane holds
// a similar share of rows; each lane keeps its instances in batch order. The
// lanes wait for the context stream's set-up work, and run_engines returns
// with every lane's stream drained.
void solve_lanes(svx_batch* b, const std::vector<InstDesc>& descs, const EngineShape& shape,
                 const EngineOptions& opt, int K) {
    svx_ctx* ctx = b->ctx;
    while ((int)b->lanes.size() < K) {
        auto l = std::make_unique<BatchLane>();
        if (b->lanes.empty()) {
            l->stream = ctx->stream;
        } else {
            SVX_HIP(hipStreamCreate(&l->stream));  // default priority
            l->own_stream = true;
        }
        l->engine = std::make_unique<Engine>(l->stream);
        SVX_HIP(hipHostMalloc((void**)&l->h_flag, sizeof(int32_t) * 4, hipHostMallocDefault));
        SVX_HIP(hipMalloc((void**)&l->d_flag, sizeof(int32_t) * 4));
        b->lanes.push_back(std::move(l));
    }
    const size_t n = descs.size();
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return descs[a].n > descs[c].n; });
    std::vector<int> lane_of(n);
    for (size_t r = 0; r < n; ++r) {
        const size_t round = r / K, pos = r % K;
        lane_of[order[r]] = (int)((round & 1) ? K - 1 - pos : pos);
    }
    SVX_HIP(hipStreamSynchronize(ctx->stream));  // batch set-up before any lane reads it
    std::vector<EngineLane> run;
    for (int k = 0; k < K; ++k) {
        std::vector<InstDesc> mine;
        for (size_t i = 0; i < n; ++i)
            if (lane_of[i] == k) mine.push_back(descs[i]);
        BatchLane& l = *b->lanes[k];
        l.engine->set_instances(mine, std::vector<char>(mine.size(), 1), shape);
        run.push_back(EngineLane{l.engine.get(), l.h_flag, l.d_flag});
    }
    run_engines(run, opt);
}
}  // namespace

extern "C" {

int svx_set_batch_streams(svx_ctx* ctx, int32_t streams) {
    if (!ctx || streams < 0 || streams > kBatchStreamsMax) return SVX_E_ARGUMENT;
    ctx->tune.batch_streams = streams;
    return SVX_OK;
}

int svx_batch_create(svx_ctx* ctx, svx_batch** out) {
    if (!ctx || !out) return SVX_E_ARGUMENT;
    auto b = std::make_unique<svx_batch>();
    b->ctx = ctx;
    *out = b.release();
    return SVX_OK;
}

void svx_batch_destroy(svx_batch* b) { delete b; }

int svx_batch_add(svx_batch* b, svx_system* sys, double tol, int64_t max_iter, int64_t* slot) {
    if (!b || !sys || !(tol > 0.0) || sys->mesh->ctx != b->ctx) return SVX_E_ARGUMENT;
    if (!sys->dirichlet) return SVX_E_ARGUMENT;
    b->sys.push_back(sys);
    b->tuned_precision = -1;  // a new member invalidates a tuned shape
    b->tol.push_back(tol);
    b->max_iter.push_back(max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * sys->mesh->n));
    if (slot) *slot = (int64_t)b->sys.size() - 1;
    return SVX_OK;
}

int64_t svx_batch_size(const svx_batch* b) { return b ? (int64_t)b->sys.size() : -1; }

}  // extern "C"

namespace {
// Device time of one iteration of the whole set: the SpMV (with its
// direction step) and the update, plus the FP32-IR period end spread over
// the default period of 8 iterations. The tile size moves work between the
// SpMV and the partial-sum reductions of every phase, so the tuner times all
// of them.
double iteration_ms(Engine& e, int precision, int reps) {
    double ms[4];
    e.time_phases(precision, reps, ms);
    return ms[0] + ms[1] + (precision == SVX_FP64 ? 0.0 : ms[2] / 8.0);
}

// Prepare every system of the batch for `precision`, reset the instance
// state and return the instance descriptors.
std::vector<InstDesc> batch_setup(svx_batch* b, int precision) {
    hipStream_t s = b->ctx->stream;
    const size_t n = b->sys.size();
    for (svx_system* sy : b->sys) system_prepare(sy, precision, b->ctx->tune.matrix_free);
    b->scal.alloc(n);
    std::vector<InstScal> init(n);
    std::vector<InstDesc> descs(n);
    for (size_t i = 0; i < n; ++i) {
        svx_system* sy = b->sys[i];
        init[i] = initial_scal(b->tol[i], b->max_iter[i], b->ctx->tune.ru_delta, sy->mesh->max_row_len);
        descs[i] = system_desc(sy);
        descs[i].s = b->scal.p + i;
    }
    b->scal.upload(init.data(), n, s);
    if (!b->engine) b->engine = std::make_unique<Engine>(s);
    return descs;
}
}  // namespace

extern "C" {

int svx_batch_autotune(svx_batch* b, int precision, int reps, int32_t* unroll, int32_t* rows_per_tile,
                       int32_t* tile_order, double* spmv_ms) {
    if (!b || reps <= 0 || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded_b(nullptr, [&]() -> int {
        const size_t n = b->sys.size();
        if (n == 0) return SVX_E_ARGUMENT;
        std::vector<InstDesc> descs = batch_setup(b, precision);
        const std::vector<char> all(n, 1);
        bool present[4] = {false, false, false, false};
        for (const InstDesc& d : descs) present[d.bs] = true;
        EngineShape best = batch_shape(b->sys, b->ctx->tune, precision);
        b->engine->set_instances(descs, all, best);
        double best_ms = iteration_ms(*b->engine, precision, reps);
        // blocks in flight per lane x ELL chunks per tile (one or two, split
        // over the wavefronts; or one per wavefront). Longer tiles time well with every
        // instance running but leave the last, largest instances of a batch
        // with few workgroups (measured: FP32-IR batch of 512 plane meshes,
        // 16-chunk tiles 6.5 s against 6.1 s with 4-chunk tiles).
        static const int kChunksPerTile[] = {1, 2, kBlock / 64};
        for (int bs = 2; bs <= 3; ++bs) {
            if (!present[bs]) continue;
            for (int x = 0; x <= 1; ++x)
                for (int u = 1; u <= kMaxUnroll; ++u)
                    for (int c : kChunksPerTile) {
                        EngineShape sh = best;
                        sh.unroll[bs] = u;
                        sh.rows_per_tile[bs] = c * kEllRows * bs;
                        sh.xcd[bs] = x;
                        b->engine->set_instances(descs, all, sh);
                        const double ms = iteration_ms(*b->engine, precision, reps);
                        if (ms < best_ms) best_ms = ms, best = sh;
                    }
        }
        b->tuned_shape = best;
        b->tuned_precision = precision;
        for (int bs = 0; bs < 4; ++bs) {
            if (unroll) unroll[bs] = best.unroll[bs];
            if (rows_per_tile) rows_per_tile[bs] = best.rows_per_tile[bs];
            if (tile_order) tile_order[bs] = best.xcd[bs] ? SVX_TILES_XCD : SVX_TILES_LINEAR;
        }
        if (spmv_ms) *spmv_ms = best_ms;
        return SVX_OK;
    });
}

int svx_batch_solve(svx_batch* b, int precision, svx_error* err) {
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->index = -1;
    }
    if (!b || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded_b(err, [&]() -> int {
        hipStream_t s = b->ctx->stream;
        const size_t n = b->sys.size();
        if (n == 0) return SVX_OK;
        std::vector<InstDesc> descs = batch_setup(b, precision);
        EngineShape shape =
            b->tuned_precision == precision ? b->tuned_shape : batch_shape(b->sys, b->ctx->tune, precision);
        shape.matrix_free = b->ctx->tune.matrix_free;  // the operator in force now
        b->matrix_free = shape.matrix_free;
        EngineOptions opt;
        opt.precision = precision;
        opt.period = b->ctx->tune.period;
        opt.chunk = b->ctx->tune.chunk;
        opt.use_graph = b->ctx->tune.use_graph;
        const int K = (int)std::min<size_t>(n, (size_t)(b->ctx->tune.batch_streams > 0 ? b->ctx->tune.batch_streams
                                                                                         : kBatchStreamsDefault));
        // the batch's own engine holds every instance either way (the timing
        // entries launch its tables)
        b->engine->set_instances(descs, std::vector<char>(n, 1), shape);
        if (K <= 1)
            b->engine->run(opt, b->ctx->h_flag, b->ctx->d_flag);
        else
            solve_lanes(b, descs, shape, opt, K);
        b->precision = precision;
        b->h_scal.resize(n);
        b->scal.download(b->h_scal.data(), n, s);
        SVX_HIP(hipStreamSynchronize(s));
        int64_t failed = 0;
        for (size_t i = 0; i < n; ++i) {
            b->sys[i]->has_u = true;
            b->sys[i]->last = b->h_scal[i];
            b->sys[i]->last_precision = precision;
            b->sys[i]->solved = b->h_scal[i].status == ST_CONVERGED;
            failed += !b->sys[i]->solved;
            system_true_residual(b->sys[i]);
        }
        if (failed && err) {
            err->code = SVX_E_NOT_CONVERGED;
            snprintf(err->message, sizeof err->message, "%lld of %zu instances did not converge", (long long)failed, n);
        }
        return failed ? SVX_E_NOT_CONVERGED : SVX_OK;
    });
}

int svx_batch_prepare(svx_batch* b, int precision) {
    if (!b || (precision != SVX_FP64 && precision != SVX_FP32_IR)) return SVX_E_ARGUMENT;
    return guarded_b(nullptr, [&]() -> int {
        const size_t n = b->sys.size();
        if (n == 0) return SVX_E_ARGUMENT;
        std::vector<InstDesc> descs = batch_setup(b, precision);
        EngineShape shape =
            b->tuned_precision == precision ? b->tuned_shape : batch_shape(b->sys, b->ctx->tune, precision);
        shape.matrix_free = b->ctx->tune.matrix_free;
        b->engine->set_instances(descs, std::vector<char>(n, 1), shape);
        b->matrix_free = shape.matrix_free;
        b->precision = precision;
        return SVX_OK;
    });
}

int svx_batch_result(const svx_batch* b, int64_t slot, int64_t* iterations, double* residual, int32_t* converged) {
    if (!b || slot < 0 || slot >= (int64_t)b->h_scal.size()) return SVX_E_ARGUMENT;
    const InstScal& sc = b->h_scal[slot];
    if (iterations) *iterations = sc.it;
    if (residual) *residual = sc.res;
    if (converged) *converged = sc.status == ST_CONVERGED;
    return SVX_OK;
}

int svx_batch_time_spmv(svx_batch* b, int reps, double* avg_ms, double* bytes_per_launch) {
    if (!b || !b->engine || reps <= 0 || b->precision < 0) return SVX_E_ARGUMENT;
    return guarded_b(nullptr, [&] {
        b->engine->tile_all();
        double ms = b->engine->time_spmv(b->precision, reps);
        double bytes = 0.0;
        const bool mf = b->matrix_free && b->precision == SVX_FP64;  // the FP32 SpMV multiplies the FP32 copy
        for (svx_system* sy : b->sys)
            bytes += mf && sy->mf_ready ? mf_spmv_bytes(sy->mesh) : spmv_bytes(sy->mesh, b->precision);
        if (avg_ms) *avg_ms = ms;
        if (bytes_per_launch) *bytes_per_launch = bytes;
        return SVX_OK;
    });
}

}  // extern "C"
