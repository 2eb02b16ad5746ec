// Request server: continuous batching of whole-model solves on one device.
//
// Decks arrive through svx_server_submit from any host thread. Host workers
// (up to 4) validate each one and build its CSR pattern (prepare_model,
// mesh_build_host). The device thread admits prepared requests while their
// estimated device footprint fits the HBM budget and an instance slot is
// free: it uploads the mesh, assembles, applies the constraints and hands the
// system to a running Krylov engine as a fresh instance, which joins at the
// next chunk boundary. Between chunks (32 iterations) the thread retires
// finished instances — recovery on the device, results copied to the
// request, device memory freed — and admits new ones, so requests of any
// size overlap on the device as slots and memory free up.
//
// Each instance runs exactly the operations of a solo svx_solve_model with
// the same precision and launch shape (per-instance reductions, fixed lanes
// per block size), so a result does not depend on what else is in flight.
// FP32-IR instances that fall back to FP64 move to a second engine.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "api_internal.hpp"
#include "deck.hpp"

using namespace svx;

namespace {

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void put_err(svx_error* err, int code, const std::string& msg, int reason = 0, int64_t index = -1,
             int64_t iterations = 0, double residual = 0.0) {
    if (!err) return;
    err->code = code;
    err->reason = reason;
    err->index = index;
    err->iterations = iterations;
    err->residual = residual;
    snprintf(err->message, sizeof err->message, "%s", msg.c_str());
}


// Prepared requests waiting for admission at most: a prepared request holds
// its CSR pattern and incidence lists on the host (tens of MB for a
// 1e5-DOF C3D8 block), so under overload the host workers stop preparing
// once this many wait, and the backlog stays as the submitted decks.
constexpr size_t kReadyCap = 128;

// Queued: in prep_q. Preparing: a host worker holds it (nobody else may
// touch it until the worker publishes it under the lock). Prepared: in
// ready_q. Active: admitted by the device thread. Done: a waiter may take it.
enum class St { Queued, Preparing, Prepared, Active, Done };

struct Request {
    int64_t ticket = 0;
    std::unique_ptr<OwnedModel> model;  // null until a host worker reads deck_text
    std::string deck_text;              // svx_server_submit_deck
    PreparedModel pm;
    std::unique_ptr<svx_mesh> mesh;
    std::unique_ptr<svx_system> sys;
    St state = St::Queued;
    int phase_precision = SVX_FP64;  // engine the instance runs in
    int64_t bytes = 0;               // admitted device footprint (estimate)
    double t_submit = 0, t_prepared = 0, t_admit = 0, t_done = 0;
    // results
    int rc = SVX_OK;
    svx_error err{};
    std::vector<double> disp, vm, nvm, react;
    int64_t iterations = 0, refreshes = 0, fp64_iterations = 0;
    double residual = 0, cond = 0;
};

// Device bytes a request holds while admitted: mesh arrays, CSR values and
// vectors, the ELL copies of K (FP64, and FP32 for FP32-IR), and the larger
// of the transient element stiffness and nodal-stress buffers.
int64_t footprint(const svx_mesh& m, int precision) {
    const int64_t n = m.n, nn = m.n_nodes, ne = m.n_elem, nnz = m.nnz;
    const int64_t ell = (int64_t)m.hp.ell_col.size(), ell_vals = ell * m.ndof * m.ndof;
    int64_t b = 24 * nn + 12 * ne + 4 * (int64_t)m.h_conn.size() + 8 * (ne + 1);  // coords, kinds/offsets, conn
    b += 4 * (n + 1) + 4 * nnz + 4 * (nn + 1) + 4 * m.nblk;                        // CSR + node-block patterns
    b += 4 * ell + 8 * ((int64_t)m.hp.ell_nb.size() + 1);                          // ELL pattern
    b += 8 * (nn + 1) + 5 * (int64_t)m.hp.inc_elem.size() + 16 * ne;              // incidence, element offsets
    b += 8 * nnz + 8 * ell_vals + 8 * n * 10 + 4 * n;                             // CSR + ELL values, vectors, dinv32
    if (precision == SVX_FP32_IR) b += 4 * ell_vals + 4 * n * 6;                   // FP32 ELL values + vectors
    b += 8 * std::max(m.ke_total, m.ns_total + nn * 6);                            // transient assembly / recovery
    if (m.mf_state == 2 && m.mf_kind == 8) {  // matrix-free C3D8: element ranks, force slots, constraint mask
        b += 8 * ne + 24 * nn * m.mf_width + nn;
    } else if (m.mf_state == 2 && m.mf_plan) {  // matrix-free plane: patch lists, packed coordinates, mask
        const MfPlan& p = *m.mf_plan;
        b += (int64_t)(sizeof(MfPatch) * p.patches.size() + 4 * p.nodes.size() + 2 * p.econn.size() + 2 * p.adj.size());
        b += 17 * nn;
    }
    return b + (1 << 16);
}

}  // namespace

struct svx_server {
    svx_ctx* ctx = nullptr;
    // Second context on the same device: its stream and scratch carry the
    // requests' admission and retirement while Krylov chunks run on ctx's.
    svx_ctx* aux = nullptr;
    svx_server_config cfg{};
    svx::TuneParams tune;  // the context's tuning and operator when the server was created
    std::mutex mu;
    std::condition_variable cv_prep, cv_dev, cv_done, cv_waiters;
    std::deque<Request*> prep_q, ready_q;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    int64_t next_ticket = 1;
    bool stop = false;
    int waiters = 0;  // threads inside svx_server_wait
    size_t preparing = 0;  // requests the host workers hold
    std::vector<std::thread> prep_threads;
    std::thread dev_thread;
    // statistics (guarded by mu)
    int64_t submitted = 0, completed = 0, failed = 0, resident = 0, peak_resident = 0, peak_active = 0;
    int64_t dof_done = 0;
    double t_first = -1, t_last_done = 0;
    double admit_ms = 0, solve_ms = 0, retire_ms = 0, idle_ms = 0;  // device thread
    int64_t chunks = 0, retiles = 0;
    std::vector<double> latencies;
    std::string fatal;

    // systems and meshes of finished requests, freed between chunks (device
    // thread only; systems before their meshes)
    std::vector<std::unique_ptr<svx_mesh>> graveyard_mesh;
    std::vector<std::unique_ptr<svx_system>> graveyard_sys;  // destroyed first

    void prep_loop();
    void dev_loop();
    void finish(Request* r);
    // With mu held: every request not Done and not held by a host worker is
    // failed with SVX_E_DEVICE (the device thread is gone or going), the
    // queues are emptied and waiters are woken.
    void fail_pending(const std::string& why);
};

void svx_server::fail_pending(const std::string& why) {
    for (auto& kv : reqs) {
        Request* r = kv.second.get();
        if (r->state == St::Done || r->state == St::Preparing) continue;
        r->rc = SVX_E_DEVICE;
        put_err(&r->err, SVX_E_DEVICE, why);
        r->state = St::Done;
        r->t_done = wall_ms();
        failed++;
    }
    prep_q.clear();
    ready_q.clear();
    cv_done.notify_all();
}

void svx_server::prep_loop() {
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_prep.wait(lk, [&] { return stop || (!prep_q.empty() && ready_q.size() + preparing < kReadyCap); });
            if (stop) return;
            r = prep_q.front();
            prep_q.pop_front();
            r->state = St::Preparing;
            preparing++;
        }
        svx_error err{};
        err.index = -1;
        int rc = SVX_OK;
        try {
            if (!r->model) {  // deck text: read it here, off the submitting thread
                auto om = std::make_unique<OwnedModel>();
                std::string msg;
                const int64_t line = parse_inp(r->deck_text.data(), r->deck_text.size(), *om, &msg);
                std::string().swap(r->deck_text);
                if (line) {
                    rc = SVX_E_PARSE;
                    put_err(&err, SVX_E_PARSE, "line " + std::to_string(line) + ": " + msg, 0, line);
                } else {
                    std::lock_guard<std::mutex> lk(mu);  // waiters read model sizes under the lock
                    r->model = std::move(om);
                }
            }
            if (rc == SVX_OK) rc = prepare_model(&r->model->desc, r->pm, &err);
            if (rc == SVX_OK) {
                r->mesh = std::make_unique<svx_mesh>();
                r->mesh->ctx = aux;
                const svx_model_desc& md = r->model->desc;
                mesh_build_host(r->mesh.get(), r->pm.ndof, md.n_nodes, md.n_elem, md.kinds, md.conn_ptr,
                                r->pm.conn_pos.data());
                // the matrix-free plan (host work) here rather than on the
                // device thread; only systems of one material use it
                const std::vector<double>& E = r->pm.E;
                const std::vector<double>& nu = r->pm.nu;
                bool uniform = tune.matrix_free && !E.empty();
                for (size_t e = 1; e < E.size() && uniform; ++e) uniform = E[e] == E[0] && nu[e] == nu[0];
                if (uniform) mesh_mf_plan(r->mesh.get(), md.coords);
                r->bytes = footprint(*r->mesh, cfg.precision);
            }
        } catch (const std::exception& e) {
            rc = SVX_E_ARGUMENT;
            put_err(&err, rc, e.what());
        }
        std::lock_guard<std::mutex> lk(mu);
        preparing--;
        r->t_prepared = wall_ms();
        if (stop) {  // the server failed or is being destroyed: nothing will admit it
            r->rc = SVX_E_DEVICE;
            put_err(&r->err, SVX_E_DEVICE, fatal.empty() ? "server stopped" : fatal);
            r->state = St::Done;
            r->t_done = r->t_prepared;
            failed++;
            cv_done.notify_all();
        } else if (rc != SVX_OK) {
            r->rc = rc;
            r->err = err;
            r->state = St::Done;
            r->t_done = r->t_prepared;
            failed++;
            cv_done.notify_all();
        } else if (r->bytes > cfg.hbm_budget) {
            r->rc = SVX_E_CAPACITY;
            put_err(&r->err, SVX_E_CAPACITY, "deck exceeds the server's device memory budget");
            r->state = St::Done;
            r->t_done = r->t_prepared;
            failed++;
            cv_done.notify_all();
        } else {
            r->state = St::Prepared;
            ready_q.push_back(r);
            cv_dev.notify_all();
        }
    }
}

// Called on the device thread, without the lock, for a request that left
// the engines (or failed during admission).
void svx_server::finish(Request* r) {
    // Device memory is released by the device thread once no chunk is in
    // flight (a free waits for every stream of the device).
    if (r->sys) graveyard_sys.push_back(std::move(r->sys));
    if (r->mesh) graveyard_mesh.push_back(std::move(r->mesh));
    std::lock_guard<std::mutex> lk(mu);
    r->t_done = wall_ms();
    r->state = St::Done;
    resident -= r->bytes;
    if (r->rc == SVX_OK) {
        completed++;
        dof_done += (int64_t)r->model->desc.n_nodes * r->pm.ndof;
        latencies.push_back(r->t_done - r->t_submit);
        t_last_done = r->t_done;
    } else {
        failed++;
    }
    cv_done.notify_all();
}

void svx_server::dev_loop() {
    try {
        SVX_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;   // Krylov chunks
        hipStream_t sa = aux->stream;  // admission and retirement, beside the running chunk
        Engine eng32(s), eng64(s);
        std::vector<Request*> in32, in64;  // instance order of each engine
        bool dirty32 = false, dirty64 = false;
        const int main_prec = cfg.precision;
        const bool mf = tune.matrix_free;  // FP64 products on the matrix-free operator where a system has one
        EngineOptions opt;
        opt.period = tune.period;
        opt.chunk = cfg.chunk > 0 ? cfg.chunk : tune.chunk;
        // measured on the served mix at 60 requests/s: 5.34 M DOF/s kernel by
        // kernel against 5.12-5.24 M with a capture after every membership change
        opt.use_graph = tune.use_graph && cfg.use_graph != 0;
        // Admission on the aux stream: upload, assembly, constraints, solver
        // set-up. Returns false (after finishing the request) on a model error.
        auto admit_one = [&](Request* r) -> bool {
            r->t_admit = wall_ms();
            const svx_model_desc& md = r->model->desc;
            mesh_upload(r->mesh.get(), md.coords);
            r->sys = std::make_unique<svx_system>();
            r->sys->mesh = r->mesh.get();
            int64_t bad = -1;
            int why = 0;
            system_assemble(r->sys.get(), r->pm.E.data(), r->pm.nu.data(), md.thickness, &bad, &why);
            if (bad >= 0) {
                char buf[256];
                snprintf(buf, sizeof buf, "element %u: %s", md.elem_ids[bad], reason_text(why));
                r->rc = SVX_E_INVALID_MODEL;
                put_err(&r->err, SVX_E_INVALID_MODEL, buf, why, bad);
                finish(r);
                return false;
            }
            const std::string& le = !r->pm.load_error.empty() ? r->pm.load_error : r->pm.bc_error;
            if (!le.empty()) {
                r->rc = SVX_E_INVALID_MODEL;
                put_err(&r->err, SVX_E_INVALID_MODEL, le);
                finish(r);
                return false;
            }
            system_set_f(r->sys.get(), r->pm.f.data());
            system_dirichlet(r->sys.get(), (int64_t)r->pm.cons_rows.size(), r->pm.cons_rows.data());
            system_prepare(r->sys.get(), main_prec, mf);
            const int64_t n = r->mesh->n;
            InstScal sc = initial_scal(cfg.tol, std::max<int64_t>(1000, 10 * n), tune.ru_delta, r->mesh->max_row_len);
            r->sys->scal.upload(&sc, 1, sa);
            SVX_HIP(hipStreamSynchronize(sa));  // complete before a chunk on s reads it
            r->phase_precision = main_prec;
            return true;
        };
        // Retirement on the aux stream of a request that left the engines:
        // status, u, recovery, results, memory.
        auto retire = [&](Request* r, const InstScal& st) {
            r->iterations = st.it;
            r->refreshes = st.refreshes;
            r->fp64_iterations = st.fp64_it != 0 ? st.it + st.fp64_it : 0;
            r->residual = st.res;
            if (st.status != ST_CONVERGED) {
                char buf[256];
                snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)st.it,
                         st.res);
                r->rc = SVX_E_NOT_CONVERGED;
                put_err(&r->err, SVX_E_NOT_CONVERGED, buf, 0, -1, st.it, st.res);
                finish(r);
                return;
            }
            r->sys->has_u = true;
            const svx_model_desc& md = r->model->desc;
            const int nd = r->pm.ndof;
            std::vector<double> u(r->mesh->n);
            r->sys->u.download(u.data(), u.size(), sa);
            SVX_HIP(hipStreamSynchronize(sa));
            r->disp.assign(md.n_nodes * 3, 0.0);
            for (int64_t k = 0; k < md.n_nodes; ++k)
                for (int c = 0; c < nd; ++c) r->disp[k * 3 + c] = u[k * nd + c];
            r->vm.resize(md.n_elem);
            r->nvm.resize(md.n_nodes);
            r->react.resize(r->pm.cons_rows.size());
            system_recover(r->sys.get(), r->vm.data(), r->nvm.data(), (int64_t)r->pm.cons_rows.size(),
                           r->pm.cons_rows.data(), r->react.data());
            r->cond = r->sys->cond;
            r->rc = SVX_OK;
            finish(r);
        };
        // One chunk per non-empty engine, enqueued without waiting for it.
        auto step = [&](Engine& eng, std::vector<Request*>& in, bool& dirty, int prec) -> int {
            if (in.empty()) return 0;
            int retiled = 0;
            if (dirty) {
                std::vector<InstDesc> descs;
                std::vector<svx_system*> systems;
                for (Request* r : in) {
                    descs.push_back(system_desc(r->sys.get()));
                    systems.push_back(r->sys.get());
                }
                eng.set_instances(descs, std::vector<char>(in.size(), 1), batch_shape(systems, tune, prec));
                dirty = false;
                retiled = 1;
            }
            EngineOptions o = opt;
            o.precision = prec;
            o.compact = false;  // membership is rebuilt here as instances retire
            o.wait = false;
            eng.run_chunks(o, ctx->h_flag, ctx->d_flag, 1);
            return retiled;
        };
        std::vector<Request*> admitted;                       // set up on sa; join at the next chunk
        std::vector<std::pair<Request*, InstScal>> retiring;  // left the engines at the last sweep
        std::vector<Request*> moved;
        // Reads back the membership an engine stepped over (instance i is
        // ran[i]) once its chunk is done: running instances stay, finished
        // ones retire, FP32-IR instances that fall back continue in FP64.
        auto sweep = [&](Engine& eng, const std::vector<Request*>& in, std::vector<Request*>& next, bool& dirty,
                         bool fp32) {
            if (in.empty()) return;
            std::vector<InstScal> sc;
            eng.read_scalars(sc);  // waits for the chunk
            if (sc.size() != in.size()) throw std::logic_error("server: engine membership out of step");
            std::vector<Request*> keep;
            for (size_t i = 0; i < in.size(); ++i) {
                Request* r = in[i];
                const InstScal& st = sc[i];
                if (st.status == ST_RUNNING) {
                    keep.push_back(r);
                    continue;
                }
                dirty = true;
                if (fp32 && st.status == ST_FALLBACK) {  // continue as FP64 from the current u
                    InstScal moved_state = st;
                    moved_state.status = ST_RUNNING;
                    moved_state.flags = FL_FRESH | FL_WARM;
                    moved_state.fp64_it = -st.it;
                    r->sys->scal.upload(&moved_state, 1, s);
                    SVX_HIP(hipStreamSynchronize(s));
                    r->phase_precision = SVX_FP64;
                    moved.push_back(r);
                    continue;
                }
                retiring.emplace_back(r, st);
            }
            next.swap(keep);
        };
        for (;;) {
            for (Request* r : admitted) {
                (r->phase_precision == SVX_FP64 ? in64 : in32).push_back(r);
                (r->phase_precision == SVX_FP64 ? dirty64 : dirty32) = true;
            }
            admitted.clear();
            const bool any = !in32.empty() || !in64.empty();
            const double t0 = wall_ms();
            // Each engine steps over its current membership. Instances that
            // left at the last sweep are not in it: a dirty engine rebuilds
            // its tile tables before it launches, so from here on no kernel
            // on s touches a retiring request and its memory may go.
            const std::vector<Request*> ran32 = in32, ran64 = in64;
            int64_t n_retiles = step(eng32, in32, dirty32, SVX_FP32_IR);
            n_retiles += step(eng64, in64, dirty64, SVX_FP64);
            const int64_t n_chunks = (int64_t)!ran32.empty() + (int64_t)!ran64.empty();
            const double t1 = wall_ms();
            // while the chunk runs: admit what fits, then retire
            std::vector<Request*> take;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (stop) {
                    (void)hipStreamSynchronize(s);  // no chunk in flight past the loop
                    break;
                }
                while (!ready_q.empty()) {
                    Request* r = ready_q.front();
                    const int64_t n_active = (int64_t)(in32.size() + in64.size() + take.size());
                    if (n_active >= cfg.max_active || resident + r->bytes > cfg.hbm_budget) break;
                    ready_q.pop_front();
                    cv_prep.notify_one();  // room for another prepared request
                    r->state = St::Active;
                    resident += r->bytes;
                    peak_resident = std::max(peak_resident, resident);
                    take.push_back(r);
                }
                peak_active = std::max<int64_t>(peak_active, (int64_t)(in32.size() + in64.size() + take.size()));
            }
            for (Request* r : take)
                if (admit_one(r)) admitted.push_back(r);
            const double t2 = wall_ms();
            for (auto& e : retiring) retire(e.first, e.second);
            retiring.clear();
            const double t3 = wall_ms();
            sweep(eng32, ran32, in32, dirty32, true);
            sweep(eng64, ran64, in64, dirty64, false);
            graveyard_sys.clear();  // both streams idle: the frees wait for nothing
            graveyard_mesh.clear();
            if (!moved.empty()) {
                in64.insert(in64.end(), moved.begin(), moved.end());
                dirty64 = true;
                moved.clear();
            }
            const double t4 = wall_ms();
            double t5 = t4;
            if (!any && admitted.empty() && retiring.empty() && in32.empty() && in64.empty()) {
                std::unique_lock<std::mutex> lk(mu);
                cv_dev.wait(lk, [&] { return stop || !ready_q.empty(); });
                if (stop) break;
                t5 = wall_ms();
            }
            std::lock_guard<std::mutex> lk(mu);
            solve_ms += (t1 - t0) + (t4 - t3);
            admit_ms += t2 - t1;
            retire_ms += t3 - t2;
            idle_ms += t5 - t4;
            chunks += n_chunks;
            retiles += n_retiles;
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = e.what();
        stop = true;
        // Requests a host worker still prepares are failed by that worker
        // when it publishes them (it sees stop under the lock).
        fail_pending(fatal);
        cv_prep.notify_all();
    }
}

extern "C" {

int svx_server_create(svx_ctx* ctx, const svx_server_config* cfg, svx_server** out) {
    if (!ctx || !cfg || !out || (cfg->precision != SVX_FP64 && cfg->precision != SVX_FP32_IR) || !(cfg->tol > 0.0) ||
        cfg->hbm_budget <= 0 || cfg->max_active <= 0 || cfg->chunk < 0 || (cfg->use_graph != 0 && cfg->use_graph != 1))
        return SVX_E_ARGUMENT;
    try {
        auto s = std::make_unique<svx_server>();
        s->ctx = ctx;
        s->cfg = *cfg;
        s->tune = ctx->tune;
        if (svx_ctx_create(ctx->device, &s->aux) != SVX_OK) return SVX_E_DEVICE;
        s->aux->tune = ctx->tune;
        svx_server* p = s.get();
        const int workers = std::max(1, std::min(4, (int)std::thread::hardware_concurrency() / 4));
        for (int i = 0; i < workers; ++i) p->prep_threads.emplace_back([p] { p->prep_loop(); });
        p->dev_thread = std::thread([p] { p->dev_loop(); });
        *out = s.release();
        return SVX_OK;
    } catch (const std::exception&) {
        return SVX_E_DEVICE;
    }
}

void svx_server_destroy(svx_server* srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_prep.notify_all();
    srv->cv_dev.notify_all();
    for (std::thread& t : srv->prep_threads)
        if (t.joinable()) t.join();
    if (srv->dev_thread.joinable()) srv->dev_thread.join();
    {
        // No worker or device thread is left: fail what they left behind and
        // wait for every waiter to leave before the server goes.
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->fail_pending("server destroyed with the request in flight");
        srv->cv_waiters.wait(lk, [&] { return srv->waiters == 0; });
    }
    svx_ctx* aux = srv->aux;
    delete srv;  // requests (and their device memory on aux) first
    svx_ctx_destroy(aux);
}

int svx_server_submit(svx_server* srv, const svx_model_desc* model, int64_t* ticket) {
    if (!srv || !model || !ticket) return SVX_E_ARGUMENT;
    try {
        auto r = std::make_unique<Request>();
        r->model = std::make_unique<OwnedModel>(*model);
        r->err.index = -1;
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return SVX_E_DEVICE;
        r->ticket = srv->next_ticket++;
        r->t_submit = wall_ms();
        if (srv->t_first < 0) srv->t_first = r->t_submit;
        *ticket = r->ticket;
        srv->prep_q.push_back(r.get());
        srv->reqs[r->ticket] = std::move(r);
        srv->submitted++;
        srv->cv_prep.notify_one();
        return SVX_OK;
    } catch (const std::exception&) {
        return SVX_E_DEVICE;
    }
}

int svx_server_submit_deck(svx_server* srv, const char* text, size_t len, int64_t* ticket) {
    if (!srv || !ticket || (!text && len)) return SVX_E_ARGUMENT;
    try {
        auto r = std::make_unique<Request>();
        r->deck_text.assign(text ? text : "", len);
        r->err.index = -1;
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return SVX_E_DEVICE;
        r->ticket = srv->next_ticket++;
        r->t_submit = wall_ms();
        if (srv->t_first < 0) srv->t_first = r->t_submit;
        *ticket = r->ticket;
        srv->prep_q.push_back(r.get());
        srv->reqs[r->ticket] = std::move(r);
        srv->submitted++;
        srv->cv_prep.notify_one();
        return SVX_OK;
    } catch (const std::exception&) {
        return SVX_E_DEVICE;
    }
}

int svx_server_wait(svx_server* srv, int64_t ticket, double timeout_ms, svx_solution* out, svx_request_info* info,
                    svx_error* err) {
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->index = -1;
    }
    if (!srv) return SVX_E_ARGUMENT;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->waiters++;
        struct Leave {  // every return path leaves the waiter count
            svx_server* s;
            ~Leave() {
                if (--s->waiters == 0) s->cv_waiters.notify_all();
            }
        } leave{srv};
        auto it = srv->reqs.find(ticket);
        if (it == srv->reqs.end()) return SVX_E_ARGUMENT;
        Request* q = it->second.get();
        auto done = [&] { return q->state == St::Done; };
        if (timeout_ms < 0)
            srv->cv_done.wait(lk, done);
        else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done))
            return SVX_E_TIMEOUT;
        if (!out) {  // sizes and status only; the ticket stays valid
            if (info) {
                *info = svx_request_info{};
                if (q->model) {
                    info->n_nodes = q->model->desc.n_nodes;
                    info->n_elem = q->model->desc.n_elem;
                    info->n_reactions = svx_model_reaction_count(&q->model->desc);
                }
                info->n_dof = q->model ? info->n_nodes * q->pm.ndof : 0;
            }
            if (q->rc != SVX_OK && err) *err = q->err;
            return q->rc;
        }
        const bool wants_reactions = out->reaction_node || out->reaction_dof || out->reaction_value;
        if (q->rc == SVX_OK && wants_reactions && (int64_t)q->pm.cons_rows.size() > out->reaction_capacity) {
            if (err) put_err(err, SVX_E_ARGUMENT, "svx_server_wait: reaction arrays smaller than the request's reaction count");
            return SVX_E_ARGUMENT;  // the ticket stays valid
        }
        r = std::move(it->second);
        srv->reqs.erase(it);
    }
    if (info) {
        info->latency_ms = r->t_done - r->t_submit;
        info->queue_ms = r->t_admit > 0 ? r->t_admit - r->t_submit : r->t_done - r->t_submit;
        info->prepare_ms = r->t_prepared - r->t_submit;
        info->iterations = r->iterations;
        info->refreshes = r->refreshes;
        info->fp64_iterations = r->fp64_iterations;
        info->device_bytes = r->bytes;
        info->n_dof = r->model ? (int64_t)r->model->desc.n_nodes * r->pm.ndof : 0;
        info->n_nodes = r->model ? r->model->desc.n_nodes : 0;
        info->n_elem = r->model ? r->model->desc.n_elem : 0;
        info->n_reactions = r->model ? svx_model_reaction_count(&r->model->desc) : 0;
    }
    if (r->rc != SVX_OK) {
        if (err) *err = r->err;
        return r->rc;
    }
    if (out) {
        const svx_model_desc& md = r->model->desc;
        if (out->displacements) std::copy(r->disp.begin(), r->disp.end(), out->displacements);
        if (out->von_mises) std::copy(r->vm.begin(), r->vm.end(), out->von_mises);
        if (out->nodal_von_mises) std::copy(r->nvm.begin(), r->nvm.end(), out->nodal_von_mises);
        for (size_t i = 0; i < r->pm.cons_rows.size(); ++i) {
            if (out->reaction_node) out->reaction_node[i] = r->pm.cons_node[i];
            if (out->reaction_dof) out->reaction_dof[i] = r->pm.cons_dof[i];
            if (out->reaction_value) out->reaction_value[i] = r->react[i];
        }
        out->iterations = r->iterations;
        out->final_residual = r->residual;
        out->cond_estimate = r->cond;
        out->solver_method = 0;
        out->ndof = r->pm.ndof;
        out->assemble_ms = 0.0;
        out->solve_ms = r->t_done - r->t_admit;
        out->recover_ms = 0.0;
        (void)md;
    }
    return SVX_OK;
}

int svx_server_stats(svx_server* srv, svx_server_summary* out) {
    if (!srv || !out) return SVX_E_ARGUMENT;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = svx_server_summary{};
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->queued = (int64_t)(srv->prep_q.size() + srv->ready_q.size());
    out->resident_bytes = srv->resident;
    out->peak_resident_bytes = srv->peak_resident;
    out->peak_active = srv->peak_active;
    out->solved_dof = srv->dof_done;
    out->admit_ms = srv->admit_ms;
    out->solve_ms = srv->solve_ms;
    out->retire_ms = srv->retire_ms;
    out->idle_ms = srv->idle_ms;
    out->chunks = srv->chunks;
    out->retiles = srv->retiles;
    if (!srv->latencies.empty()) {
        std::vector<double> l = srv->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) { return l[std::min(l.size() - 1, (size_t)(q * (l.size() - 1) + 0.5))]; };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        out->max_ms = l.back();
        const double span = srv->t_last_done - srv->t_first;
        out->dof_per_s = span > 0 ? srv->dof_done / (span * 1e-3) : 0.0;
    }
    return srv->fatal.empty() ? SVX_OK : SVX_E_DEVICE;
}

}  // extern "C"
