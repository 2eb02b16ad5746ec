// Krylov engine: runs Jacobi-PCG (FP64 or FP32-IR) over a set of resident
// instances with hipGraph-captured chunks of iterations and a device-side
// convergence count polled once per chunk.
#pragma once
#include "svx_internal.hpp"

namespace svx {

struct EngineOptions {
    int precision = SVX_FP64;
    int period = 8;        // FP32-IR: iterations between residual refreshes
    int chunk = 32;        // iterations per captured graph
    bool use_graph = true;
};

class Engine {
   public:
    explicit Engine(hipStream_t s) : stream_(s) {}
    ~Engine();
    // Host descriptors (pointers already set; tile fields filled by build).
    // rows_per_tile[i] must be a multiple of the instance's block size.
    void build(std::vector<InstDesc>& descs, const std::vector<int>& rows_per_tile, int lanes, int blanes, int bs);
    // Run until every instance leaves ST_RUNNING. Scalars must be initialised
    // (status = ST_RUNNING, tol, max_iter, ru_delta) by the caller.
    void run(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag);
    const KrylovSet& set() const { return ks_; }
    int64_t launches() const { return launches_; }

   private:
    void release_graph();
    hipStream_t stream_;
    DevBuf<Tile> tiles_;
    DevBuf<InstDesc> descs_;
    DevBuf<double> partials_;
    KrylovSet ks_{};
    hipGraphExec_t exec_ = nullptr;
    int exec_precision_ = -1, exec_period_ = -1, exec_chunk_ = -1;
    int64_t launches_ = 0;
};

// Pick lanes-per-row for a CSR matrix / block-row for a BSR matrix from the
// average row length (the autotuner overrides these when it has data).
int default_lanes(double avg_row_len);
int default_rows_per_tile(int bs);

}  // namespace svx
