// Krylov engine: runs Jacobi-PCG (FP64 or FP32-IR) over a set of resident
// instances with hipGraph-captured chunks of iterations and a device-side
// convergence count polled once per chunk.
//
// Instances are grouped by node-block size (1, 2, 3); each group is a tile
// table launched as one grid per phase, and the groups' launches are captured
// into the same graph. Tiles of finished instances are dropped from the
// tables (and the graph re-captured) once they hold a quarter of all tiles;
// pending per-tile partials move with the tiles. Instances flagged FL_FRESH are initialised by the init
// phase at the start of the next chunk, so new work can join a running set.
#pragma once
#include "svx_internal.hpp"

namespace svx {

struct EngineOptions {
    int precision = SVX_FP64;
    int period = 8;        // FP32-IR: iterations between residual refreshes
    int chunk = 32;        // iterations per captured graph
    bool use_graph = true;
    bool compact = true;   // drop finished instances' tiles as they converge
    // false: enqueue one chunk (and the running count) and return -1 at once;
    // the caller synchronizes the stream (read_scalars does) before reading.
    bool wait = true;
};

constexpr int kStepKernelTiles = 1024;

struct EngineShape {
    int lanes[4] = {0, 0, 0, 0};   // lanes per row (raw CSR, block size 1)
    int unroll[4] = {0, 0, 0, 0};  // blocks in flight per lane (ELL, block sizes 2, 3)
    int rows_per_tile[4] = {0, 0, 0, 0};  // 0: per-instance default; block sizes 2, 3: whole 64-node-row chunks
    int xcd[4] = {1, 1, 1, 1};  // gathering kernels: XCD-contiguous tile order (1) or tile = workgroup (0)
    int nt[4] = {0, 0, 0, 0};   // FP64 ELL SpMV: nontemporal matrix loads (1) or cached (0)
    int nt32[4] = {0, 0, 0, 0}; // FP32 ELL SpMV (FP32-IR): the same for the FP32 copy
    int ntv[4] = {0, 0, 0, 0};  // FP64 and FP32 updates: nontemporal vector loads and stores (1) or cached (0)
    // Assembled instances with more row tiles than this take their FP64
    // pending direction step (r.r and r.z reduced over the instance's tiles,
    // convergence, budget, beta) in the step kernel, once per instance,
    // instead of in every SpMV tile, whose reduction reads every tile's
    // partials (a cost that grows with the square of the tile count) at one
    // launch more per iteration. Negative: never.
    int step_tiles = kStepKernelTiles;
    // FP64 SpMV of instances that carry a matrix-free operator (InstDesc::mf_kind)
    // on that operator instead of the assembled matrix.
    bool matrix_free = false;
};

class Engine {
   public:
    explicit Engine(hipStream_t s) : stream_(s) {}
    ~Engine();
    // Replace the instance set. Descriptors must carry valid pointers for the
    // precision that will run; tile fields are filled here. `running[i]`
    // false leaves instance i out of the tile tables.
    void set_instances(const std::vector<InstDesc>& descs, const std::vector<char>& running, const EngineShape& shape);
    // Run chunks until no instance runs (max_chunks < 0) or max_chunks
    // chunks ran. Returns the number of instances still running.
    int run(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag, int max_chunks = -1);
    // Download the per-instance scalars (status, iterations, residual).
    void read_scalars(std::vector<InstScal>& out) const;
    double time_spmv(int precision, int reps);
    // One SpMV launch (direction step and product) over the current tiles,
    // for timing several engines' products back to back on one stream;
    // spmv_prepare() forces every instance to run first.
    void spmv_prepare();
    void spmv_launch(int precision);
    // Put every instance back into the tile tables (after a solve compacted
    // them), e.g. before timing the whole set.
    void tile_all() { build_tiles(std::vector<char>(h_descs_.size(), 1)); }
    // Average device time of each phase of an iteration, launched alone on
    // the current set with every instance forced to run: [0] direction step
    // and SpMV, [1] update, [2] FP32-IR period end (0 for FP64), [3] FP32-IR
    // refresh (fold, FP64 residual, repoint; 0 for FP64). Leaves the
    // instances' solver state undefined.
    void time_phases(int precision, int reps, double ms[4]);
    int64_t launches() const { return launches_; }
    // Carve the engine's device tables (tiles, descriptors, partial sums,
    // read-back staging) from [base, base + bytes) instead of allocating
    // them: no device allocation after this call. Tile tables are then
    // rebuilt only by set_instances (no partial carry on compaction: run with
    // EngineOptions::compact = false, and without graphs, whose
    // instantiation allocates). Throws if a table does not fit.
    void use_arena(void* base, size_t bytes) {
        arena_ = static_cast<char*>(base);
        arena_bytes_ = bytes;
    }
    // Device bytes the tables of one instance of n rows (raw CSR, 256-row
    // tiles) take in an arena.
    static size_t arena_bytes_for_csr(int64_t n);
    int n_instances() const { return (int)h_descs_.size(); }

    // Run chunks of opt.precision only (no FP64 fallback phase) until no
    // instance runs or max_chunks chunks ran; returns the running count.
    int run_chunks(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag, int max_chunks) {
        return run_phase(opt, h_flag, d_flag, max_chunks);
    }
    // The steps of one chunk for a caller that drives several engines: enqueue
    // it (with the running count copied to h_flag) without waiting; after the
    // caller has synchronized the stream, settle it (re-tile as run_phase
    // does between chunks; returns the running count). begin_fallback moves
    // the FP32-IR instances that fell back into an FP64 phase (false: none).
    void enqueue_phase_chunk(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag);
    int settle_phase_chunk(const EngineOptions& opt, const int32_t* h_flag);
    bool begin_fallback();
    hipStream_t stream() const { return stream_; }

   private:
    int run_phase(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag, int max_chunks);
    void release_graph();
    // Tile tables for the instances with running[i]. carry: keep the pending
    // per-tile partials of instances that stay tiled (moved to their new
    // tile positions), so a set can be re-tiled between chunks.
    void build_tiles(const std::vector<char>& running, bool carry = false);
    void enqueue_init(const EngineOptions& opt);
    void enqueue_chunk(const EngineOptions& opt, int iters);
    hipStream_t stream_;
    std::vector<InstDesc> h_descs_;
    EngineShape shape_;
    bool uses_mf(const InstDesc& d) const;
    DevBuf<Tile> tiles_;
    DevBuf<Tile> mf_tiles_;
    DevBuf<Tile> mf_leaders_;
    DevBuf<Tile> alpha_leaders_;
    DevBuf<Tile> mf3_tiles_;
    DevBuf<InstDesc> descs_;
    DevBuf<double> partials_;
    mutable DevBuf<InstScal> scal_tmp_;  // read_scalars staging
    KrylovSet groups_[4]{};
    int tiled_instances_ = 0;
    int tiled_tiles_ = 0;
    std::vector<char> tiled_;  // instance i has tiles in the current tables
    hipGraphExec_t exec_ = nullptr;
    int exec_precision_ = -1, exec_period_ = -1, exec_chunk_ = -1;
    int64_t launches_ = 0;
    char* arena_ = nullptr;  // use_arena: tables carved here, nothing allocated
    size_t arena_bytes_ = 0, arena_off_ = 0;
    template <typename T>
    void place(DevBuf<T>& b, size_t count);
};

// One engine of a set that runs side by side, each on its own stream with
// its own running-count flags.
struct EngineLane {
    Engine* engine;
    int32_t* h_flag;  // pinned host, 2 x int32
    int32_t* d_flag;  // device, 2 x int32
};
// Run every lane's instances to the end (opt.precision, then the FP64 phase
// of FP32-IR instances that fell back), enqueueing one chunk on every lane
// before polling any, so that one lane's latency-bound products overlap
// another lane's bandwidth-bound updates. Each instance runs exactly the
// operations of a single-engine run, so its results are the same bits.
void run_engines(const std::vector<EngineLane>& lanes, const EngineOptions& opt);

int default_lanes(double avg_row_len);
// Default tile of an instance (node-block size bs, n_rows rows) and ELL
// unroll of a node-block size and precision: functions of the instance alone,
// so an instance gets the same launch shape, and the same bits, alone and
// inside any batch.
int default_rows_per_tile(int bs, int64_t n_rows, bool matrix_free = false);
int unroll_for(int bs, int precision);

}  // namespace svx
