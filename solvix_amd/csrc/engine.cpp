#include "engine.hpp"

#include <algorithm>

namespace svx {

int default_lanes(double avg) {
    int l = 2;
    while (l < 64 && l * 5 < avg) l *= 2;
    return l;
}

// Node-block instances: four 64-node-row ELL chunks per workgroup tile, one
// per wavefront; instances of fewer than kSmallChunks chunks (65,536 node
// rows) one chunk per tile, its blocks split over the four wavefronts, so
// that they still put four wavefronts per chunk on the device. Blocks in
// flight per lane, measured on the 265,923-DOF C3D8 block and on 2e5 /
// 4e6-DOF plane meshes: one for the FP64 3x3 blocks (nine FP64 loads per
// block already), three otherwise.
constexpr int64_t kSmallChunks = 1024;

int default_rows_per_tile(int bs, int64_t n_rows, bool matrix_free) {
    if (bs < 2) return 256;
    const int64_t chunks = (n_rows / bs + kEllRows - 1) / kEllRows;
    if (chunks < kSmallChunks) return kEllRows * bs;  // one chunk, split over the four wavefronts
    // Large assembled 3x3-block instances: two chunks per tile, two wavefronts
    // per chunk. Twice the wavefronts of one chunk per wavefront keep more
    // matrix loads in flight, which matters once the matrix streams from HBM
    // (measured on the 265,923-DOF block rotating two distinct matrices: FP64
    // SpMV 40.5 us against 46.2 us); the tile count, whose square the
    // partial-sum reductions cost, only doubles (one-chunk tiles quadruple it).
    if (bs == 3 && !matrix_free) return 2 * kEllRows * bs;
    return (kBlock / 64) * kEllRows * bs;
}

int unroll_for(int bs, int precision) {
    (void)precision;  // measured on both precisions: 3x3 blocks two in flight, smaller blocks three
    return bs == 3 ? 2 : 3;
}

Engine::~Engine() { release_graph(); }

template <typename T>
void Engine::place(DevBuf<T>& b, size_t count) {
    if (!arena_) {
        b.alloc(count);
        return;
    }
    const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
    if (arena_off_ + bytes > arena_bytes_) throw std::length_error("engine tables exceed the arena");
    b.view(reinterpret_cast<T*>(arena_ + arena_off_), count);
    arena_off_ += bytes;
}

size_t Engine::arena_bytes_for_csr(int64_t n) {
    const size_t tiles = (size_t)std::max<int64_t>(1, (n + 255) / 256);
    auto r = [](size_t b) { return (b + 255) / 256 * 256; };
    return r(tiles * sizeof(Tile)) + 4 * r(sizeof(Tile)) + r(sizeof(InstDesc)) +
           r(std::max<size_t>(kPartialSlots, kPartialSlots * tiles) * sizeof(double)) + r(sizeof(InstScal));
}

void Engine::release_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    exec_precision_ = exec_period_ = exec_chunk_ = -1;
}

void Engine::set_instances(const std::vector<InstDesc>& descs, const std::vector<char>& running,
                           const EngineShape& shape) {
    h_descs_ = descs;
    shape_ = shape;
    build_tiles(running);
}

bool Engine::uses_mf(const InstDesc& d) const {
    return shape_.matrix_free && ((d.bs == 2 && (d.mf_kind == 3 || d.mf_kind == 4)) || (d.bs == 3 && d.mf_kind == 8));
}

void Engine::build_tiles(const std::vector<char>& running, bool carry) {
    release_graph();
    // previous layout, for carrying partials
    std::vector<int32_t> old_begin(h_descs_.size(), -1);
    int old_first[4] = {0, 0, 0, 0}, old_count[4] = {0, 0, 0, 0}, old_stride[4] = {0, 0, 0, 0};
    size_t old_off[4] = {0, 0, 0, 0};
    for (int bs = 1; bs <= 3; ++bs) {
        old_first[bs] = groups_[bs].tiles ? (int)(groups_[bs].tiles - tiles_.p) : 0;
        old_count[bs] = groups_[bs].n_tiles;
        old_stride[bs] = groups_[bs].pstride;
        old_off[bs] = groups_[bs].partials ? (size_t)(groups_[bs].partials - partials_.p) : 0;
    }
    carry = carry && tiled_.size() == h_descs_.size() && partials_.p;
    if (arena_ && carry) throw std::logic_error("an engine on an arena does not re-tile mid-solve");
    arena_off_ = 0;  // an arena's tables are rebuilt in place
    if (carry)
        for (size_t i = 0; i < h_descs_.size(); ++i)
            if (tiled_[i]) old_begin[i] = h_descs_[i].tile_begin;

    // Matrix-free patch tiles (node-block size 2 only): CPS4 instances, then
    // CPS3 ones. mf_begin[i]: the instance's first patch tile.
    std::vector<Tile> mf_tiles;
    std::vector<int32_t> mf_begin(h_descs_.size(), -1);
    int32_t n_mf[2] = {0, 0}, cap_n[2] = {0, 0}, cap_e[2] = {0, 0};
    for (int kk = 0; kk < 2; ++kk) {
        const int kind = kk == 0 ? 4 : 3;
        for (size_t i = 0; i < h_descs_.size(); ++i) {
            const InstDesc& d = h_descs_[i];
            if (!running[i] || !uses_mf(d) || d.mf_kind != kind) continue;
            mf_begin[i] = (int32_t)mf_tiles.size();
            for (int32_t p = 0; p < d.mf_patches; ++p)
                mf_tiles.push_back(Tile{(int32_t)i, p, p + 1, 0, 0, P_MPQ, mf_begin[i], d.mf_patches, d.s});
            cap_n[kk] = std::max(cap_n[kk], d.mf_cap_nodes);
            cap_e[kk] = std::max(cap_e[kk], d.mf_cap_elem);
        }
        n_mf[kk] = (int32_t)mf_tiles.size() - (kk ? n_mf[0] : 0);
    }

    std::vector<Tile> tiles;
    int first[4] = {0, 0, 0, 0}, count[4] = {0, 0, 0, 0}, n_ell[4] = {0, 0, 0, 0};
    tiled_instances_ = 0;
    tiled_.assign(h_descs_.size(), 0);
    for (int bs = 1; bs <= 3; ++bs) {
        first[bs] = (int)tiles.size();
        // assembled-matrix instances first: the FP64 SpMV launch covers them
        // only, the matrix-free tiles the rest
        for (int mf = 0; mf <= 1; ++mf) {
            for (size_t i = 0; i < h_descs_.size(); ++i) {
                InstDesc& d = h_descs_[i];
                if (d.bs != bs || !running[i] || (int)uses_mf(d) != mf) continue;
                int rpt = shape_.rows_per_tile[bs] > 0 ? shape_.rows_per_tile[bs] : default_rows_per_tile(bs, d.n, mf);
                if (bs >= 2) {  // whole ELL chunks
                    const int cr = kEllRows * bs;
                    rpt = std::max(cr, (rpt + cr - 1) / cr * cr);
                }
                d.tile_begin = (int32_t)tiles.size() - first[bs];
                d.tile_count = (d.n + rpt - 1) / rpt;
                for (int32_t r = 0; r < d.n; r += rpt) {
                    Tile t{(int32_t)i, r, std::min(d.n, r + rpt), d.tile_begin, d.tile_count, P_PQ, d.tile_begin,
                           d.tile_count, d.s};
                    if (mf && bs == 2) t.pq_slot = P_MPQ, t.pq_begin = mf_begin[i], t.pq_count = d.mf_patches;
                    tiles.push_back(t);
                }
                tiled_instances_++;
                tiled_[i] = 1;
            }
            if (mf == 0) n_ell[bs] = (int)tiles.size() - first[bs];
        }
        count[bs] = (int)tiles.size() - first[bs];
    }
    // patch tiles reduce their instance's r.r / r.z over its row tiles
    for (Tile& t : mf_tiles) {
        t.tile_begin = h_descs_[t.inst].tile_begin;
        t.tile_count = h_descs_[t.inst].tile_count;
    }
    // one leader per instance whose FP64 pending step the step kernel takes:
    // group 2 (plane matrix-free instances, then large assembled 2x2-block
    // ones) first, then group 3 (C3D8 matrix-free, then large assembled
    // 3x3-block ones); C3D8 element tiles
    std::vector<Tile> mf_leaders, mf3_tiles;
    for (const Tile& t : mf_tiles)
        if (t.row_begin == 0) mf_leaders.push_back(t);
    auto large_assembled = [&](size_t i) {
        const InstDesc& d = h_descs_[i];
        return running[i] && d.bs >= 2 && !uses_mf(d) && shape_.step_tiles >= 0 && d.tile_count > shape_.step_tiles;
    };
    for (size_t i = 0; i < h_descs_.size(); ++i) {
        InstDesc& d = h_descs_[i];
        d.step_kernel = large_assembled(i) ? 1 : 0;
        if (d.step_kernel && d.bs == 2)
            mf_leaders.push_back(Tile{(int32_t)i, 0, 0, d.tile_begin, d.tile_count, P_PQ, d.tile_begin, d.tile_count, d.s});
    }
    const int32_t n_lead2 = (int32_t)mf_leaders.size();
    for (size_t i = 0; i < h_descs_.size(); ++i) {
        const InstDesc& d = h_descs_[i];
        if (!running[i] || d.bs != 3 || !uses_mf(d)) continue;
        mf_leaders.push_back(Tile{(int32_t)i, 0, 0, d.tile_begin, d.tile_count, P_PQ, d.tile_begin, d.tile_count, d.s});
        for (int32_t e = 0; e < d.mf_n_elem; e += kMf3Elems)
            mf3_tiles.push_back(Tile{(int32_t)i, e, std::min(d.mf_n_elem, e + kMf3Elems), d.tile_begin, d.tile_count,
                                     P_PQ, d.tile_begin, d.tile_count, d.s});
    }
    for (size_t i = 0; i < h_descs_.size(); ++i) {
        const InstDesc& d = h_descs_[i];
        if (d.step_kernel && d.bs == 3)
            mf_leaders.push_back(Tile{(int32_t)i, 0, 0, d.tile_begin, d.tile_count, P_PQ, d.tile_begin, d.tile_count, d.s});
    }
    // leaders whose instance has more than four times the step-kernel
    // threshold of p.q partials also take alpha once per instance
    // (cg64_alpha_kernel). Measured on plane cantilevers, update with alpha
    // against the per-tile reduction: 6,267 tiles 41.7 against 53.7 us;
    // 1,571 tiles 15.0 against 13.1 us (the extra launch costs more there).
    std::vector<Tile> alpha_leaders;
    int32_t n_alpha2 = 0;
    for (InstDesc& d : h_descs_) d.alpha_kernel = 0;
    for (size_t j = 0; j < mf_leaders.size(); ++j) {
        const Tile& t = mf_leaders[j];
        if (shape_.step_tiles < 0 || t.pq_count <= 4 * shape_.step_tiles) continue;
        h_descs_[t.inst].alpha_kernel = 1;
        alpha_leaders.push_back(t);
        if ((int32_t)j < n_lead2) n_alpha2 = (int32_t)alpha_leaders.size();
    }
    tiled_tiles_ = (int)tiles.size();
    int stride[4] = {0, 0, 0, 0};
    size_t off[4] = {0, 0, 0, 0}, total = 0;
    for (int bs = 1; bs <= 3; ++bs) {
        stride[bs] = std::max(count[bs], bs == 2 ? n_mf[0] + n_mf[1] : 0);
        off[bs] = total;
        total += (size_t)kPartialSlots * stride[bs];
    }
    // old local tile index feeding each new tile (-1: nothing to carry)
    std::vector<int32_t> src(tiles.size(), -1);
    bool any_src = false;
    if (carry)
        for (int bs = 1; bs <= 3; ++bs)
            for (int g = 0; g < count[bs]; ++g) {
                const Tile& t = tiles[first[bs] + g];
                const InstDesc& d = h_descs_[t.inst];
                if (old_begin[t.inst] < 0) continue;
                src[first[bs] + g] = old_begin[t.inst] + (g - d.tile_begin);
                any_src = true;
            }
    DevBuf<double> old_partials;
    if (any_src) old_partials = std::move(partials_);
    place(tiles_, std::max<size_t>(1, tiles.size()));
    tiles_.upload(tiles.data(), tiles.size(), stream_);
    place(mf_tiles_, std::max<size_t>(1, mf_tiles.size()));
    mf_tiles_.upload(mf_tiles.data(), mf_tiles.size(), stream_);
    place(mf_leaders_, std::max<size_t>(1, mf_leaders.size()));
    mf_leaders_.upload(mf_leaders.data(), mf_leaders.size(), stream_);
    place(alpha_leaders_, std::max<size_t>(1, alpha_leaders.size()));
    alpha_leaders_.upload(alpha_leaders.data(), alpha_leaders.size(), stream_);
    place(mf3_tiles_, std::max<size_t>(1, mf3_tiles.size()));
    mf3_tiles_.upload(mf3_tiles.data(), mf3_tiles.size(), stream_);
    place(descs_, std::max<size_t>(1, h_descs_.size()));
    descs_.upload(h_descs_.data(), h_descs_.size(), stream_);
    place(partials_, std::max<size_t>(kPartialSlots, total));
    if (arena_) place(scal_tmp_, std::max<size_t>(1, h_descs_.size()));
    if (any_src) {
        DevBuf<int32_t> d_src(src.size());
        d_src.upload(src.data(), src.size(), stream_);
        for (int bs = 1; bs <= 3; ++bs)
            launch_carry_partials(old_partials.p + old_off[bs], old_stride[bs], partials_.p + off[bs], stride[bs],
                                  d_src.p + first[bs], count[bs], stream_);
        SVX_HIP(hipStreamSynchronize(stream_));
    }
    for (int bs = 1; bs <= 3; ++bs) {
        KrylovSet& k = groups_[bs];
        k.tiles = tiles_.p + first[bs];
        k.n_tiles = count[bs];
        k.n_ell_tiles = n_ell[bs];
        k.descs = descs_.p;
        k.n_inst = (int32_t)h_descs_.size();
        k.partials = partials_.p + off[bs];
        k.pstride = stride[bs];
        k.lanes = shape_.lanes[bs] > 0 ? shape_.lanes[bs] : 16;
        k.unroll = shape_.unroll[bs] > 0 ? shape_.unroll[bs] : 2;
        k.bs = bs;
        k.xcd = shape_.xcd[bs];
        k.nt = shape_.nt[bs];
        k.nt32 = shape_.nt32[bs];
        k.ntv = shape_.ntv[bs];
        k.mf_tiles = bs == 2 ? mf_tiles_.p : nullptr;
        k.mf_leaders = bs == 2 ? mf_leaders_.p : bs == 3 ? mf_leaders_.p + n_lead2 : nullptr;
        k.n_mf_inst = bs == 2 ? n_lead2 : bs == 3 ? (int32_t)mf_leaders.size() - n_lead2 : 0;
        k.alpha_leaders = bs == 2 ? alpha_leaders_.p : bs == 3 ? alpha_leaders_.p + n_alpha2 : nullptr;
        k.n_alpha = bs == 2 ? n_alpha2 : bs == 3 ? (int32_t)alpha_leaders.size() - n_alpha2 : 0;
        k.mf3_tiles = bs == 3 ? mf3_tiles_.p : nullptr;
        k.n_mf3_tiles = bs == 3 ? (int32_t)mf3_tiles.size() : 0;
        k.n_mf4 = bs == 2 ? n_mf[0] : 0;
        k.n_mf3 = bs == 2 ? n_mf[1] : 0;
        for (int kk = 0; kk < 2; ++kk) {
            k.mf_cap_nodes[kk] = bs == 2 ? cap_n[kk] : 0;
            k.mf_cap_elem[kk] = bs == 2 ? cap_e[kk] : 0;
        }
    }
    SVX_HIP(hipStreamSynchronize(stream_));
}

void Engine::enqueue_init(const EngineOptions& opt) {
    for (int bs = 1; bs <= 3; ++bs) {
        if (opt.precision == SVX_FP64)
            cg64_init(groups_[bs], stream_);
        else
            ir32_init(groups_[bs], stream_);
    }
}

// `iters` is even (and a multiple of the even FP32-IR period): the chunk
// starts and ends with the current, complete direction in P[0] and no
// pending partial sums, so the tile tables can change between chunks.
void Engine::enqueue_chunk(const EngineOptions& opt, int iters) {
    if (opt.precision == SVX_FP64) {
        for (int i = 0; i < iters; ++i)
            for (int bs = 1; bs <= 3; ++bs) cg64_iteration(groups_[bs], stream_, i & 1);
        for (int bs = 1; bs <= 3; ++bs) cg64_chunk_end(groups_[bs], stream_);
    } else {
        for (int i = 0; i < iters; i += opt.period) {
            for (int j = 0; j < opt.period; ++j)
                for (int bs = 1; bs <= 3; ++bs) ir32_iteration(groups_[bs], stream_, j & 1);
            for (int bs = 1; bs <= 3; ++bs) ir32_period_end(groups_[bs], stream_);
            for (int bs = 1; bs <= 3; ++bs) ir32_refresh(groups_[bs], stream_);
        }
    }
}

int Engine::run(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag, int max_chunks) {
    int left = run_phase(opt, h_flag, d_flag, max_chunks);
    if (left != 0 || opt.precision == SVX_FP64) return left;
    if (!begin_fallback()) return 0;
    EngineOptions o64 = opt;
    o64.precision = SVX_FP64;
    return run_phase(o64, h_flag, d_flag, -1);
}

bool Engine::begin_fallback() {
    // FP32-IR instances that stopped improving finish as FP64 PCG from their
    // current solution (their descriptors carry the FP64 vectors too).
    std::vector<InstScal> sc;
    read_scalars(sc);
    std::vector<char> fb(sc.size());
    bool any = false;
    for (size_t i = 0; i < sc.size(); ++i) any |= (fb[i] = sc[i].status == ST_FALLBACK) != 0;
    if (!any) return false;
    launch_promote_fallback(descs_.p, (int32_t)h_descs_.size(), stream_);
    build_tiles(fb);
    return true;
}

namespace {
// Chunk length of a phase: even, and a multiple of the (even) FP32-IR period.
int phase_chunk(EngineOptions& opt) {
    opt.period = std::max(2, (opt.period + 1) / 2 * 2);  // even: the period ends with the direction in P[0]
    int chunk = std::max(2, (opt.chunk + 1) / 2 * 2);
    if (opt.precision != SVX_FP64) chunk = std::max(opt.period, (chunk / opt.period) * opt.period);
    return chunk;
}
}  // namespace

void Engine::enqueue_phase_chunk(const EngineOptions& opt_in, int32_t* h_flag, int32_t* d_flag) {
    EngineOptions opt = opt_in;
    const int chunk = phase_chunk(opt);
    enqueue_init(opt);  // initialises FL_FRESH instances only
    bool graph_ok = opt.use_graph && !arena_;  // instantiating a graph allocates
    if (graph_ok && (exec_ == nullptr || exec_precision_ != opt.precision || exec_period_ != opt.period ||
                     exec_chunk_ != chunk)) {
        release_graph();
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            graph_ok = false;
        } else {
            bool enqueued = true;
            try {
                enqueue_chunk(opt, chunk);
            } catch (const DeviceError&) {
                enqueued = false;
            }
            if (hipStreamEndCapture(stream_, &g) != hipSuccess || g == nullptr || !enqueued) {
                if (g) (void)hipGraphDestroy(g);
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
    if (graph_ok)
        SVX_HIP(hipGraphLaunch(exec_, stream_));
    else
        enqueue_chunk(opt, chunk);
    launches_++;
    launch_count_running(descs_.p, (int32_t)h_descs_.size(), d_flag, stream_);
    SVX_HIP(hipMemcpyAsync(h_flag, d_flag, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
}

int Engine::settle_phase_chunk(const EngineOptions& opt, const int32_t* h_flag) {
    const int running = h_flag[0], live_tiles = h_flag[1];
    if (running == 0) return 0;
    // Re-tile once finished instances hold a quarter of the tiles: each of
    // their workgroups still costs a launch slot and a state load per phase.
    if (opt.compact && tiled_instances_ > 1 && running < tiled_instances_ && 4 * live_tiles <= 3 * tiled_tiles_) {
        std::vector<InstScal> sc;
        read_scalars(sc);
        std::vector<char> live(h_descs_.size());
        for (size_t i = 0; i < sc.size(); ++i) live[i] = sc[i].status == ST_RUNNING;
        build_tiles(live, true);
    }
    return running;
}

int Engine::run_phase(const EngineOptions& opt, int32_t* h_flag, int32_t* d_flag, int max_chunks) {
    int done_chunks = 0;
    for (;;) {
        enqueue_phase_chunk(opt, h_flag, d_flag);
        done_chunks++;
        if (!opt.wait) return -1;
        SVX_HIP(hipStreamSynchronize(stream_));
        const int running = settle_phase_chunk(opt, h_flag);
        if (running == 0) return 0;
        if (max_chunks >= 0 && done_chunks >= max_chunks) return running;
    }
}

void run_engines(const std::vector<EngineLane>& lanes, const EngineOptions& opt) {
    struct State {
        EngineOptions opt;
        bool done = false, fallback = false;
    };
    std::vector<State> st(lanes.size());
    for (size_t k = 0; k < lanes.size(); ++k) {
        st[k].opt = opt;
        st[k].opt.wait = true;
        st[k].done = lanes[k].engine->n_instances() == 0;
    }
    for (;;) {
        bool any = false;
        // one chunk on every lane's stream before any lane is polled, so the
        // lanes' kernels run side by side
        for (size_t k = 0; k < lanes.size(); ++k)
            if (!st[k].done) {
                lanes[k].engine->enqueue_phase_chunk(st[k].opt, lanes[k].h_flag, lanes[k].d_flag);
                any = true;
            }
        if (!any) return;
        for (size_t k = 0; k < lanes.size(); ++k) {
            if (st[k].done) continue;
            SVX_HIP(hipStreamSynchronize(lanes[k].engine->stream()));
            if (lanes[k].engine->settle_phase_chunk(st[k].opt, lanes[k].h_flag) != 0) continue;
            if (st[k].opt.precision != SVX_FP64 && !st[k].fallback && lanes[k].engine->begin_fallback()) {
                st[k].fallback = true;
                st[k].opt.precision = SVX_FP64;
            } else {
                st[k].done = true;
            }
        }
    }
}

void Engine::read_scalars(std::vector<InstScal>& out) const {
    out.resize(h_descs_.size());
    if (h_descs_.empty()) return;
    if (scal_tmp_.n < h_descs_.size()) {
        if (arena_) throw std::logic_error("arena staging smaller than the instance set");
        scal_tmp_.alloc(h_descs_.size());
    }
    launch_gather_scalars(descs_.p, (int32_t)h_descs_.size(), scal_tmp_.p, stream_);
    scal_tmp_.download(out.data(), out.size(), stream_);
    SVX_HIP(hipStreamSynchronize(stream_));
}

void Engine::time_phases(int precision, int reps, double ms[4]) {
    hipEvent_t a, b;
    SVX_HIP(hipEventCreate(&a));
    SVX_HIP(hipEventCreate(&b));
    const int n_inst = (int32_t)h_descs_.size();
    for (int ph = 0; ph < 4; ++ph) {
        ms[ph] = 0.0;
        if (ph >= 2 && precision == SVX_FP64) continue;
        for (int r = 0; r <= reps; ++r) {  // r == 0 warms up
            launch_set_running(descs_.p, n_inst, stream_, ph == 3 ? FL_REFRESH : 0);
            SVX_HIP(hipEventRecord(a, stream_));
            for (int bs = 1; bs <= 3; ++bs) {
                if (ph == 0) spmv_phase(groups_[bs], precision, stream_);
                if (ph == 1) update_phase(groups_[bs], precision, stream_);
                if (ph == 2) period_end_phase(groups_[bs], precision, stream_);
                if (ph == 3) ir32_refresh(groups_[bs], stream_);
            }
            SVX_HIP(hipEventRecord(b, stream_));
            SVX_HIP(hipEventSynchronize(b));
            float t = 0.0f;
            SVX_HIP(hipEventElapsedTime(&t, a, b));
            if (r > 0) ms[ph] += t;
        }
        ms[ph] /= reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
}

void Engine::spmv_prepare() { launch_set_running(descs_.p, (int32_t)h_descs_.size(), stream_); }

void Engine::spmv_launch(int precision) {
    for (int bs = 1; bs <= 3; ++bs) spmv_phase(groups_[bs], precision, stream_);
}

double Engine::time_spmv(int precision, int reps) {
    hipEvent_t a, b;
    SVX_HIP(hipEventCreate(&a));
    SVX_HIP(hipEventCreate(&b));
    spmv_prepare();
    spmv_launch(precision);  // warm
    SVX_HIP(hipEventRecord(a, stream_));
    for (int i = 0; i < reps; ++i) spmv_launch(precision);
    SVX_HIP(hipEventRecord(b, stream_));
    SVX_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    SVX_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

}  // namespace svx
