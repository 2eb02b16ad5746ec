// FP64 Jacobi-PCG on a raw CSR matrix that allocates no device memory
// (svx_csr_pcg_noalloc): for callers that cannot call hipMalloc (code under
// stream capture, tight serving loops). Its vectors, scalar state and engine
// tables live in one file-scope device array, `scratch`, that every call on a
// device shares; the kernels are those of svx_csr_pcg.
#include <mutex>

#include "api_internal.hpp"

using namespace svx;

namespace {

// Rows the scratch holds: eight FP64 vectors (f, 1/diag, u, r, p, z, q and
// the second direction), the FP32 diagonal, plus the engine tables.
constexpr int64_t kNoAllocMaxRows = 1 << 19;
constexpr size_t kVecBytes = (size_t)kNoAllocMaxRows * (8 * sizeof(double) + sizeof(float));
constexpr size_t kScratchBytes = kVecBytes + (2u << 20);

__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

namespace svx {

// Resolves this device's copy of `scratch` (with the device current); called
// by svx_ctx_create.
void noalloc_register_device(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* ptr = nullptr;
    SVX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<char*>(ptr);
}

int64_t noalloc_max_rows() { return kNoAllocMaxRows; }

}  // namespace svx

extern "C" int svx_csr_pcg_noalloc(svx_csr* csr, const double* f, double tol, int64_t max_iter, double* u,
                                   int64_t* iterations, double* residual, svx_error* err) {
    auto fail = [&](int code, const char* msg) {
        if (err) {
            std::memset(err, 0, sizeof *err);
            err->code = code;
            err->index = -1;
            snprintf(err->message, sizeof err->message, "%s", msg);
        }
        return code;
    };
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->index = -1;
    }
    if (!csr || !f || !(tol > 0.0)) return fail(SVX_E_ARGUMENT, "svx_csr_pcg_noalloc: bad argument");
    if (csr->n > kNoAllocMaxRows)
        return fail(SVX_E_CAPACITY, "svx_csr_pcg_noalloc: more rows than its scratch holds (524288)");
    svx_ctx* ctx = csr->ctx;
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
    }
    if (!base) return fail(SVX_E_DEVICE, "svx_csr_pcg_noalloc: no scratch resolved for this device");
    try {
        hipStream_t s = ctx->stream;
        SVX_HIP(hipSetDevice(ctx->device));
        const int64_t n = csr->n;
        const size_t vb = round256(sizeof(double) * (size_t)n);
        // vectors, then the scalar state, then the engine tables
        double* v[8];
        for (int k = 0; k < 8; ++k) v[k] = reinterpret_cast<double*>(base + k * vb);
        float* dinv32 = reinterpret_cast<float*>(base + 8 * vb);
        size_t off = 8 * vb + round256(sizeof(float) * (size_t)n);
        InstScal* d_scal = reinterpret_cast<InstScal*>(base + off);
        off += round256(sizeof(InstScal));
        const size_t tables = Engine::arena_bytes_for_csr(n);
        if (off + tables > kScratchBytes)
            return fail(SVX_E_CAPACITY, "svx_csr_pcg_noalloc: engine tables exceed the scratch");
        double *df = v[0], *dinv = v[1], *du = v[2], *dr = v[3], *dp = v[4], *dz = v[5], *dq = v[6], *dp_alt = v[7];
        SVX_HIP(hipMemcpyAsync(df, f, sizeof(double) * n, hipMemcpyHostToDevice, s));
        launch_diag_inv(n, csr->row_ptr.p, csr->col.p, csr->vals.p, dinv, dinv32, nullptr, nullptr, s);
        InstDesc d{};
        d.n = (int32_t)n;
        d.bs = 1;
        d.row_ptr = csr->row_ptr.p;
        d.col = csr->col.p;
        d.val = csr->vals.p;
        d.dinv = dinv;
        d.dinv32 = dinv32;
        d.f = df;
        d.u = du;
        d.r = dr, d.p = dp, d.z = dz, d.q = dq, d.p_alt = dp_alt;
        d.s = d_scal;
        const int64_t mi = max_iter > 0 ? max_iter : std::max<int64_t>(1000, 10 * n);
        InstScal sc = initial_scal(tol, mi, ctx->tune.ru_delta, csr->max_row_len);
        SVX_HIP(hipMemcpyAsync(d_scal, &sc, sizeof sc, hipMemcpyHostToDevice, s));
        // the launch shape of svx_csr_pcg, so both give the same iterates
        Engine eng(s);
        eng.use_arena(base + off, kScratchBytes - off);
        const TuneParams& tp = ctx->tune;
        EngineShape shape;
        shape.rows_per_tile[1] = tp.rows_per_tile > 0 ? tp.rows_per_tile : 256;
        shape.lanes[1] = tp.lanes > 0 ? tp.lanes : default_lanes((double)csr->nnz / n);
        eng.set_instances({d}, {1}, shape);
        EngineOptions opt;
        opt.precision = SVX_FP64;
        opt.chunk = tp.chunk;
        opt.use_graph = false;
        opt.compact = false;
        eng.run(opt, ctx->h_flag, ctx->d_flag);
        SVX_HIP(hipMemcpyAsync(&sc, d_scal, sizeof sc, hipMemcpyDeviceToHost, s));
        if (u) SVX_HIP(hipMemcpyAsync(u, du, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SVX_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = sc.it;
        if (residual) *residual = sc.res;
        if (sc.status != ST_CONVERGED) {
            char buf[256];
            snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)sc.it, sc.res);
            fail(SVX_E_NOT_CONVERGED, buf);
            if (err) err->iterations = sc.it, err->residual = sc.res;
            return SVX_E_NOT_CONVERGED;
        }
        return SVX_OK;
    } catch (const DeviceError& e) {
        return fail(SVX_E_DEVICE, e.what());
    } catch (const std::exception& e) {
        return fail(SVX_E_ARGUMENT, e.what());
    }
}
