#include "engine.hpp"

#include <algorithm>

namespace svx {

int default_lanes(double avg) {
    int l = 2;
    while (l < 64 && l * 5 < avg) l *= 2;
    return l;
}

int default_rows_per_tile(int bs) { return bs == 3 ? 288 : 256; }

Engine::~Engine() { release_graph(); }

void Engine::release_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    exec_precision_ = exec_period_ = exec_chunk_ = -1;
}

void Engine::build(std::vector<InstDesc>& descs, const std::vector<int>& rows_per_tile, int lanes, int blanes,
                   int bs) {
    release_graph();
    std::vector<Tile> tiles;
    for (size_t i = 0; i < descs.size(); ++i) {
        InstDesc& d = descs[i];
        int rpt = std::max(1, rows_per_tile[i]);
        d.tile_begin = (int32_t)tiles.size();
        for (int32_t r = 0; r < d.n; r += rpt) tiles.push_back(Tile{(int32_t)i, r, std::min(d.n, r + rpt), 0});
        d.tile_count = (int32_t)tiles.size() - d.tile_begin;
        if (d.tile_count == 0) throw std::runtime_error("instance with no rows");
    }
    tiles_.alloc(tiles.size());
    tiles_.upload(tiles.data(), tiles.size(), stream_);
    descs_.alloc(descs.size());
    descs_.upload(descs.data(), descs.size(), stream_);
    partials_.alloc(4 * tiles.size());
    ks_.tiles = tiles_.p;
    ks_.n_tiles = (int32_t)tiles.size();
    ks_.descs = descs_.p;
    ks_.n_inst = (int32_t)descs.size();
    ks_.partials = partials_.p;
    ks_.lanes = lanes;
    ks_.blanes = blanes;
    ks_.bs = bs;
    SVX_HIP(hipStreamSynchronize(stream_));
}

static void enqueue_chunk(const KrylovSet& k, const EngineOptions& opt, int iters, hipStream_t s) {
    if (opt.precision == SVX_FP64) {
        for (int i = 0; i < iters; ++i) cg64_iteration(k, s);
    } else {
        for (int i = 0; i < iters; i += opt.period) {
            for (int j = 0; j < opt.period; ++j) ir32_iteration(k, s, j == opt.period - 1);
            ir32_refresh(k, s);
        }
    }
}

void Engine::run(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag) {
    const KrylovSet& k = ks_;
    if (opt.precision == SVX_FP64)
        cg64_init(k, stream_);
    else
        ir32_init(k, stream_);
    int chunk = opt.chunk;
    if (opt.precision != SVX_FP64) chunk = std::max(opt.period, (chunk / opt.period) * opt.period);
    bool graph_ok = opt.use_graph;
    if (graph_ok && (exec_ == nullptr || exec_precision_ != opt.precision || exec_period_ != opt.period ||
                     exec_chunk_ != chunk)) {
        release_graph();
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            graph_ok = false;
        } else {
            bool enqueued = true;
            try {
                enqueue_chunk(k, opt, chunk, stream_);
            } catch (const DeviceError&) {
                enqueued = false;
            }
            if (hipStreamEndCapture(stream_, &g) != hipSuccess || g == nullptr || !enqueued) {
                if (g) (void)hipGraphDestroy(g);
                g = nullptr;
                graph_ok = false;
            } else {
                if (hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0) != hipSuccess) {
                    exec_ = nullptr;
                    graph_ok = false;
                }
                (void)hipGraphDestroy(g);
            }
        }
        (void)hipGetLastError();
        if (graph_ok) {
            exec_precision_ = opt.precision;
            exec_period_ = opt.period;
            exec_chunk_ = chunk;
        }
    }
    for (;;) {
        if (graph_ok) {
            SVX_HIP(hipGraphLaunch(exec_, stream_));
        } else {
            enqueue_chunk(k, opt, chunk, stream_);
        }
        launches_++;
        launch_count_running(k.descs, k.n_inst, d_flag, stream_);
        SVX_HIP(hipMemcpyAsync(h_flag, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
        SVX_HIP(hipStreamSynchronize(stream_));
        if (*h_flag == 0) break;
    }
}

}  // namespace svx
