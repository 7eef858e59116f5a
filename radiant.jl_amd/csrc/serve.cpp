// serve.cpp -- batched solves and the request server (continuous batching under an HBM budget).
//
// Every problem keeps its own device-resident state (fluxes, Krylov basis).  The server admits
// queued requests in arrival order whenever the bytes they need fit the budget and a slot is
// free, so a request that arrives while others run starts as soon as memory allows.  In the
// mixed-precision modes admitted 1D and 2D requests join the slots of one lockstep BatchSolver at
// the next round boundary (continuous batching: each round launches one pass per resident
// problem, and a finished problem leaves its slot at the end of the round that finished it).
// 3D requests, FP64 requests and problems the row-scan sweep cannot take are solved on their own
// streams by worker threads, concurrently on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "engine.hpp"


// Deep copy of a problem description (the caller's buffers may go away after submit).
struct DescCopy {
    rt_problem_desc d{};
    std::vector<double> dx, dy, dz, sigt, sigs, qvol, mu, eta, xi, w, M, D, Msurf[6], Dsurf[6], ssrc[6], Sb, dE, T, Mfp, omega_x, omega_y, omega_z;
    std::vector<int64_t> mat, pl, n2p[6];
    std::vector<int32_t> bc;

    explicit DescCopy(const rt_problem_desc& s) {
        d = s;
        const size_t Nc = (size_t)s.Nx * s.Ny * s.Nz;
        auto cp = [](auto& v, const auto* p, size_t n) {
            if (p) v.assign(p, p + n);
        };
        cp(dx, s.dx, s.Nx); cp(dy, s.dy, s.Ny); cp(dz, s.dz, s.Nz);
        cp(mat, s.mat, Nc);
        cp(bc, s.bc, 2 * s.ndims);
        cp(sigt, s.sigt, (size_t)s.G * s.Nmat);
        cp(sigs, s.sigs, (size_t)s.Nmat * s.G * s.G * (s.Ls + 1));
        cp(qvol, s.qvol, (size_t)s.G * Nc);
        cp(mu, s.mu, s.Nd); cp(eta, s.eta, s.Nd); cp(xi, s.xi, s.Nd); cp(w, s.w, s.Nd);
        cp(M, s.M, (size_t)s.Nd * s.Np);
        cp(D, s.D, (size_t)s.Np * s.Nd);
        cp(pl, s.pl, s.Np);
        d.dx = dx.data(); d.dy = s.dy ? dy.data() : nullptr; d.dz = s.dz ? dz.data() : nullptr;
        d.mat = mat.data(); d.bc = bc.data(); d.sigt = sigt.data(); d.sigs = sigs.data(); d.qvol = qvol.data();
        d.mu = mu.data(); d.eta = eta.data(); d.xi = xi.data(); d.w = w.data(); d.M = M.data(); d.D = D.data();
        d.pl = pl.data();
        if (s.csd) {
            cp(Sb, s.Sb, (size_t)(s.G + 1) * s.Nmat);
            cp(dE, s.dE, s.G);
            d.Sb = s.Sb ? Sb.data() : nullptr;
            d.dE = s.dE ? dE.data() : nullptr;
        }
        cp(T, s.T, (size_t)s.G * s.Nmat);
        cp(Mfp, s.Mfp, (size_t)s.Np * s.Np);
        d.T = s.T ? T.data() : nullptr;
        d.Mfp = s.Mfp ? Mfp.data() : nullptr;
        cp(omega_x, s.omega_x, (size_t)(s.order_x > 1 ? s.order_x : 1) + 1);
        d.omega_x = s.omega_x ? omega_x.data() : nullptr;
        cp(omega_y, s.omega_y, (size_t)(s.order_y > 1 ? s.order_y : 1) + 1);
        d.omega_y = s.omega_y ? omega_y.data() : nullptr;
        cp(omega_z, s.omega_z, (size_t)(s.order_z > 1 ? s.order_z : 1) + 1);
        d.omega_z = s.omega_z ? omega_z.data() : nullptr;
        for (int f = 0; f < 6; ++f) {
            d.Msurf[f] = nullptr; d.Dsurf[f] = nullptr; d.n_to_np[f] = nullptr;
            if (f < 2 * s.ndims) {
                size_t ndp = 0;
                if (s.n_to_np[f]) {
                    n2p[f].assign(s.n_to_np[f], s.n_to_np[f] + s.Nd);
                    for (int64_t v : n2p[f]) if (v != 0) ++ndp;
                    d.n_to_np[f] = n2p[f].data();
                }
                if (s.Dsurf[f]) { Dsurf[f].assign(s.Dsurf[f], s.Dsurf[f] + (size_t)s.Nps * ndp); d.Dsurf[f] = Dsurf[f].data(); }
                if (s.Msurf[f]) { Msurf[f].assign(s.Msurf[f], s.Msurf[f] + (size_t)s.Nps * ndp); d.Msurf[f] = Msurf[f].data(); }
            }
            d.ssrc[f] = nullptr;
            if (f < 2 * s.ndims && s.ssrc[f] && s.Np_source > 0) {
                const size_t npf = f < 2 ? (size_t)s.Ny * s.Nz : (f < 4 ? (size_t)s.Nx * s.Nz : (size_t)s.Nx * s.Ny);
                ssrc[f].assign(s.ssrc[f], s.ssrc[f] + (size_t)s.G * s.Np_source * npf);
                d.ssrc[f] = ssrc[f].data();
            }
        }
    }
};

// A resident batch: the problems and (for the mixed-precision modes) lockstep solvers.  The
// problems are dealt round-robin to a few sub-batches, each a BatchSolver with its own streams
// driven by its own host thread, so that one sub-batch's transport passes overlap another's
// Arnoldi work and host turnaround.  Without the lockstep solver each problem keeps its own
// stream and a host thread drives it.
struct rt_batch {
    std::vector<std::unique_ptr<rt::Problem>> probs;
    std::vector<hipStream_t> sts;                      // one per sub-batch
    std::vector<std::unique_ptr<rt::BatchSolver>> bss;  // sub-batch solvers
    std::vector<std::pair<int, int>> where;             // problem -> (sub-batch, index in it)
    std::vector<std::vector<rt_group_stats>> stats;
    ~rt_batch() {
        bss.clear();
        probs.clear();
        for (hipStream_t s : sts) (void)hipStreamDestroy(s);
    }
};

// Served mixed-precision requests: 1D and 2D problems join the lockstep engine (a 1D/2D sweep is
// one wave per direction, so a single problem fills a small part of the device and lockstep
// rounds of many problems fill it); a 3D problem's sweep is a workgroup of up to 16 waves per
// direction, which keeps the device busy on its own stream, so 3D requests are solved on worker
// threads with their own streams, concurrently with each other and with the engine's rounds
// (measured on the 3D request stream: the lockstep engine reached 3.6e9 updates/s against 8.2e9
// for one such problem solved alone).  RADIANT_HIP_SERVE_3D=lockstep sends 3D requests to the
// lockstep engine as well.
static bool lockstep_request(const rt_problem_desc& d) {
    static const bool all = [] {
        const char* e = std::getenv("RADIANT_HIP_SERVE_3D");
        return e && e[0] == 'l';
    }();
    return d.ndims < 3 || all;
}

// RADIANT_HIP_SERVE_LOG=1: one line per request solved on a worker stream (queueing, set-up and
// solve times) on stderr.
static bool serve_log() {
    static const bool on = [] {
        const char* e = std::getenv("RADIANT_HIP_SERVE_LOG");
        return e && e[0] == '1';
    }();
    return on;
}

static int sub_batches(int n) {
    const char* e = std::getenv("RADIANT_HIP_SUBBATCH");
    int k = e ? std::atoi(e) : 4;
    k = std::max(1, std::min(k, 4));
    return n >= 2 * k ? k : 1;
}

static rt_batch* make_batch(int32_t n, const rt_problem_desc* const* p) {
    std::unique_ptr<rt_batch> b(new rt_batch());
    bool lockstep = n > 0;
    for (int i = 0; i < n; ++i)
        if (p[i]->precision == RT_PREC_FP64) lockstep = false;
    if (lockstep) {
        const int nsub = sub_batches(n);
        b->sts.resize(nsub, nullptr);
        for (int s = 0; s < nsub; ++s) RT_CHECK(hipStreamCreateWithFlags(&b->sts[s], hipStreamNonBlocking));
        std::vector<std::vector<rt::Problem*>> raw(nsub);
        for (int i = 0; i < n; ++i) {
            const int s = i % nsub;
            b->probs.emplace_back(new rt::Problem(*p[i], b->sts[s]));
            b->where.push_back({s, (int)raw[s].size()});
            raw[s].push_back(b->probs.back().get());
        }
        try {
            for (int s = 0; s < nsub; ++s) b->bss.emplace_back(new rt::BatchSolver(raw[s], b->sts[s]));
        } catch (const rt::Error& e) {
            if (e.code != RT_EINVAL) throw;
            b->bss.clear();  // a problem the row-scan sweep cannot take: per-problem streams
        }
        if (b->bss.empty()) {
            b->probs.clear();
            b->where.clear();
            for (hipStream_t s : b->sts) (void)hipStreamDestroy(s);
            b->sts.clear();
        }
    }
    if (b->bss.empty())
        for (int i = 0; i < n; ++i) b->probs.emplace_back(new rt::Problem(*p[i]));
    b->stats.resize(n);
    for (int i = 0; i < n; ++i) b->stats[i].assign(b->probs[i]->G, rt_group_stats{});
    return b.release();
}

static void solve_batch(rt_batch* b, rt_solve_info* info) {
    const int n = (int)b->probs.size();
    if (!b->bss.empty()) {
        const int nsub = (int)b->bss.size();
        std::vector<rt_solve_info> infos(nsub);
        std::vector<std::string> errs(nsub);
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (int s = 0; s < nsub; ++s)
            th.emplace_back([&, s] {
                rt::set_autotune_measuring(nsub == 1);  // concurrent sub-batches are not timed
                try {
                    b->bss[s]->run(&infos[s]);
                } catch (const std::exception& e) {
                    errs[s] = e.what();
                }
            });
        for (auto& t : th) t.join();
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int s = 0; s < nsub; ++s)
            if (!errs[s].empty()) throw rt::Error(RT_EDEVICE, "sub-batch " + std::to_string(s) + ": " + errs[s]);
        for (int i = 0; i < n; ++i) b->stats[i] = b->bss[b->where[i].first]->stats(b->where[i].second);
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->ms_device = wall;
            for (auto& x : infos) {
                info->passes += x.passes;
                info->passes_lo += x.passes_lo;
                info->updates_raw += x.updates_raw;
                info->updates_problem += x.updates_problem;
                info->graph_launches += x.graph_launches;
                info->graph_instantiations += x.graph_instantiations;
            }
        }
        return;
    }
    hipEvent_t e0, e1;
    RT_CHECK(hipEventCreate(&e0));
    RT_CHECK(hipEventCreate(&e1));
    RT_CHECK(hipDeviceSynchronize());
    RT_CHECK(hipEventRecord(e0, nullptr));
    std::vector<std::thread> th;
    std::vector<std::string> errs(n);
    std::vector<rt_solve_info> infos(n);
    for (int i = 0; i < n; ++i)
        th.emplace_back([&, i] {
            rt::set_autotune_measuring(n == 1);  // concurrent problems are not timed
            try {
                b->probs[i]->solve(b->stats[i].data(), &infos[i]);
            } catch (const std::exception& e) {
                errs[i] = e.what();
            }
        });
    for (auto& t : th) t.join();
    RT_CHECK(hipDeviceSynchronize());
    RT_CHECK(hipEventRecord(e1, nullptr));
    RT_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    RT_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int i = 0; i < n; ++i)
        if (!errs[i].empty()) throw rt::Error(RT_EDEVICE, "problem " + std::to_string(i) + ": " + errs[i]);
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->ms_device = ms;
        for (auto& x : infos) {
            info->passes += x.passes;
            info->passes_lo += x.passes_lo;
            info->updates_raw += x.updates_raw;
            info->updates_problem += x.updates_problem;
            info->graph_launches += x.graph_launches;
            info->graph_instantiations += x.graph_instantiations;
        }
    }
}

template <class F>
static int guard(F f) {
    try {
        f();
        return RT_OK;
    } catch (const rt::Error& e) {
        rt::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        rt::set_last_error(e.what());
        return RT_EDEVICE;
    }
}

extern "C" {

int rt_batch_create(int32_t n, const rt_problem_desc* const* p, rt_batch** out) {
    if (n < 1 || !p || !out) {
        rt::set_last_error("invalid batch arguments");
        return RT_EINVAL;
    }
    *out = nullptr;
    return guard([&] { *out = make_batch(n, p); });
}

int rt_batch_solve(rt_batch* b, rt_group_stats* const* stats_out, rt_solve_info* info) {
    if (!b) {
        rt::set_last_error("null batch");
        return RT_EINVAL;
    }
    return guard([&] {
        solve_batch(b, info);
        if (stats_out)
            for (size_t i = 0; i < b->probs.size(); ++i)
                if (stats_out[i]) std::copy(b->stats[i].begin(), b->stats[i].end(), stats_out[i]);
    });
}

int rt_batch_get_flux(rt_batch* b, int32_t i, double* phi_out) {
    if (!b || i < 0 || i >= (int32_t)b->probs.size() || !phi_out) {
        rt::set_last_error("invalid batch flux request");
        return RT_EINVAL;
    }
    return guard([&] { b->probs[i]->get_flux(phi_out); });
}

int rt_batch_get_flux_cutoff(rt_batch* b, int32_t i, double* cutoff_out) {
    if (!b || i < 0 || i >= (int32_t)b->probs.size() || !cutoff_out) {
        rt::set_last_error("invalid batch cutoff-flux request");
        return RT_EINVAL;
    }
    return guard([&] { b->probs[i]->get_flux_cutoff(cutoff_out); });
}

int rt_batch_kernel_times(rt_batch* b, double* ms, int64_t* launches, double* updates, double* flops,
                          double* bytes) {
    if (!b || b->bss.empty()) {
        rt::set_last_error("kernel times are recorded by lockstep (mixed-precision) batches only");
        return RT_EINVAL;
    }
    for (int p = 0; p < 3; ++p)
        for (int k = 0; k < 4; ++k) {
            double t_ms = 0, t_upd = 0, t_fl = 0, t_by = 0;
            int64_t t_n = 0;
            for (auto& bs : b->bss) {
                t_ms += bs->timing.ms[p][k];
                t_n += bs->timing.n[p][k];
                t_upd += bs->timing.upd[p][k];
                t_fl += bs->timing.flop[p][k];
                t_by += bs->timing.bytes[p][k];
            }
            if (ms) ms[4 * p + k] = t_ms;
            if (launches) launches[4 * p + k] = t_n;
            if (updates) updates[4 * p + k] = t_upd;
            if (flops) flops[4 * p + k] = t_fl;
            if (bytes) bytes[4 * p + k] = t_by;
        }
    return RT_OK;
}

int rt_batch_destroy(rt_batch* b) {
    return guard([&] { delete b; });
}

int rt_solve_batch(int32_t n, const rt_problem_desc* const* p, double* const* phi_out, rt_group_stats* const* stats_out,
                   rt_solve_info* info) {
    if (n < 0 || (n > 0 && (!p || !phi_out))) {
        rt::set_last_error("invalid batch arguments");
        return RT_EINVAL;
    }
    if (n == 0) return RT_OK;
    return guard([&] {
        std::unique_ptr<rt_batch> b(make_batch(n, p));
        solve_batch(b.get(), info);
        for (int i = 0; i < n; ++i) {
            if (stats_out && stats_out[i]) std::copy(b->stats[i].begin(), b->stats[i].end(), stats_out[i]);
            b->probs[i]->get_flux(phi_out[i]);
        }
    });
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Server
// ------------------------------------------------------------------------------------------
struct Request {
    int64_t id;
    std::unique_ptr<DescCopy> desc;
    int64_t bytes = 0;
    enum State { QUEUED, RUNNING, DONE, FAILED, CANCELLED } state = QUEUED;
    std::unique_ptr<rt::Problem> result;  // a finished solve: flux moments resident until collected
    std::vector<rt_group_stats> stats;
    std::string err;
    int code = 0;
    std::chrono::steady_clock::time_point t_submit, t_done;
    std::atomic<bool> cancel{false};
};

struct rt_server {
    rt_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> queue;
    std::map<int64_t, std::shared_ptr<Request>> reqs;
    int64_t next_id = 0;
    int64_t bytes_in_use = 0;
    int active = 0;    // admitted requests (lockstep slots and workers)
    int in_batch = 0;  // of which in lockstep slots (scheduler thread only)
    bool stop = false;
    std::thread sched;
    // Results dropped with rt_server_cancel are released on this thread: freeing device memory
    // waits for the device's outstanding work, which would stall the caller.  Their bytes return
    // to the budget once released.
    std::thread reaper;
    std::deque<std::pair<std::unique_ptr<rt::Problem>, int64_t>> trash;
    std::condition_variable rcv;
    bool reaper_stop = false;
    void reap_results() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            rcv.wait(lk, [&] { return reaper_stop || !trash.empty(); });
            if (trash.empty() && reaper_stop) break;
            auto item = std::move(trash.front());
            trash.pop_front();
            lk.unlock();
            try {
                item.first.reset();
            } catch (const std::exception&) {
            }
            lk.lock();
            bytes_in_use -= item.second;
            cv.notify_all();
        }
    }
    std::map<int64_t, std::thread> workers;  // requests solved on their own stream
    std::vector<int64_t> finished;           // workers to join

    // Lockstep engine of the mixed-precision servers: one BatchSolver whose slots take admitted
    // requests at round boundaries.
    hipStream_t est = nullptr;
    std::unique_ptr<rt::BatchSolver> bs;
    struct Slot {
        std::shared_ptr<Request> r;
        std::unique_ptr<rt::Problem> P;
    };
    std::vector<Slot> slots;

    // Publishes a request's outcome (state, results and completion time together) and releases
    // its slot; called with mu held.  A finished solve keeps its device memory, and its bytes stay
    // counted against the budget, until a poll collects the flux.
    void publish(Request& r, Request::State state, std::unique_ptr<rt::Problem>& P, std::vector<rt_group_stats>& st,
                 const std::string& err, int code) {
        if (state == Request::DONE) r.result = std::move(P);
        else bytes_in_use -= r.bytes;
        r.stats.swap(st);
        r.err = err;
        r.code = code;
        r.t_done = std::chrono::steady_clock::now();
        r.state = state;
        r.desc.reset();
        --active;
    }

    void run_request(std::shared_ptr<Request> r) {
        rt::set_autotune_measuring(false);  // served requests overlap on the device
        // The outcome is published after the device memory of the solve has been released.
        Request::State state = Request::CANCELLED;
        std::unique_ptr<rt::Problem> prob;
        std::vector<rt_group_stats> st;
        std::string err;
        int code = 0;
        try {
            if (!r->cancel.load()) {
                const auto t0 = std::chrono::steady_clock::now();
                prob.reset(new rt::Problem(r->desc->d));
                const auto t1 = std::chrono::steady_clock::now();
                st.resize(r->desc->d.G);
                prob->solve(st.data(), nullptr);
                state = r->cancel.load() ? Request::CANCELLED : Request::DONE;
                if (serve_log()) {
                    const auto t2 = std::chrono::steady_clock::now();
                    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
                    std::fprintf(stderr, "[serve] request %lld: queued %.1f ms, set-up %.1f ms, solve %.1f ms\n",
                                 (long long)r->id, ms(r->t_submit, t0), ms(t0, t1), ms(t1, t2));
                }
            }
        } catch (const rt::Error& e) {
            state = Request::FAILED; err = e.what(); code = e.code;
        } catch (const std::exception& e) {
            state = Request::FAILED; err = e.what(); code = RT_EDEVICE;
        }
        if (state != Request::DONE) prob.reset();
        {
            std::lock_guard<std::mutex> lk(mu);
            publish(*r, state, prob, st, err, code);
            finished.push_back(r->id);
        }
        cv.notify_all();
    }

    // Joins the worker threads that have ended (called with mu held; they no longer take it).
    void reap(std::unique_lock<std::mutex>& lk) {
        std::vector<std::thread> done;
        for (int64_t id : finished) {
            auto it = workers.find(id);
            if (it != workers.end()) {
                done.push_back(std::move(it->second));
                workers.erase(it);
            }
        }
        finished.clear();
        lk.unlock();
        for (auto& t : done) t.join();
        lk.lock();
    }

    bool head_ready() {
        if (queue.empty()) return false;
        auto& r = reqs[queue.front()];
        return r->cancel.load() || r->bytes > cfg.hbm_budget_bytes ||
               (active < cfg.max_active && bytes_in_use + r->bytes <= cfg.hbm_budget_bytes);
    }

    // Takes queued requests in arrival order while the head request fits the budget and a slot;
    // called with mu held.  Returns the requests admitted to the lockstep engine.
    std::vector<std::shared_ptr<Request>> admit() {
        std::vector<std::shared_ptr<Request>> out;
        while (!queue.empty()) {
            auto r = reqs[queue.front()];
            if (r->cancel.load()) {
                queue.pop_front();
                r->state = Request::CANCELLED;
                r->t_done = std::chrono::steady_clock::now();
                r->desc.reset();
                continue;
            }
            if (r->bytes > cfg.hbm_budget_bytes) {  // can never fit: refused at once
                queue.pop_front();
                r->state = Request::FAILED;
                r->code = RT_ENOMEM;
                r->err = "request needs " + std::to_string(r->bytes) + " device bytes, above the server budget";
                r->t_done = std::chrono::steady_clock::now();
                r->desc.reset();
                continue;
            }
            // Otherwise it waits for a slot and for the bytes held by running requests and by
            // finished results not yet collected (or dropped with rt_server_cancel).
            if (active >= cfg.max_active || bytes_in_use + r->bytes > cfg.hbm_budget_bytes) break;
            queue.pop_front();
            r->state = Request::RUNNING;
            bytes_in_use += r->bytes;
            ++active;
            if (bs && lockstep_request(r->desc->d)) out.push_back(r);
            else workers.emplace(r->id, std::thread([this, r] { run_request(r); }));
        }
        return out;
    }

    // Puts admitted requests into free slots of the lockstep engine (mu not held).  A problem the
    // row-scan sweep cannot take is solved on its own stream instead.
    void place(std::vector<std::shared_ptr<Request>>& rs) {
        for (auto& r : rs) {
            std::unique_ptr<rt::Problem> P;
            std::string err;
            int code = 0;
            bool own = false;
            try {
                P.reset(new rt::Problem(r->desc->d, est));
                try {
                    const int k = bs->add(P.get());
                    slots[k].r = r;
                    slots[k].P = std::move(P);
                    ++in_batch;
                    continue;
                } catch (const rt::Error& e) {
                    if (e.code != RT_EINVAL) throw;
                    own = true;
                }
            } catch (const rt::Error& e) {
                err = e.what(); code = e.code;
            } catch (const std::exception& e) {
                err = e.what(); code = RT_EDEVICE;
            }
            P.reset();
            std::lock_guard<std::mutex> lk(mu);
            if (own) {
                workers.emplace(r->id, std::thread([this, r] { run_request(r); }));
            } else {
                std::unique_ptr<rt::Problem> none;
                std::vector<rt_group_stats> st;
                publish(*r, Request::FAILED, none, st, err, code);
            }
        }
    }

    // One lockstep round, then the finished and cancelled requests leave their slots (mu not
    // held).  A device error fails every request in the slots and restarts the engine.
    void step() {
        struct Out {
            std::shared_ptr<Request> r;
            Request::State state;
            std::unique_ptr<rt::Problem> P;
            std::vector<rt_group_stats> st;
            std::string err;
            int code;
        };
        std::vector<Out> outs;
        std::string err;
        int code = 0;
        try {
            bs->round();
            for (int k = 0; k < (int)slots.size(); ++k) {
                Slot& sl = slots[k];
                if (!sl.r) continue;
                const bool cancelled = sl.r->cancel.load();
                if (!cancelled && !bs->done(k)) continue;
                Out o{sl.r, Request::CANCELLED, nullptr, {}, {}, 0};
                if (!cancelled) {
                    o.state = Request::DONE;
                    o.st = bs->stats(k);
                }
                bs->remove(k);
                if (!cancelled) o.P = std::move(sl.P);  // the flux stays in device memory
                sl.P.reset();
                sl.r.reset();
                --in_batch;
                outs.push_back(std::move(o));
            }
        } catch (const rt::Error& e) {
            err = e.what(); code = e.code;
        } catch (const std::exception& e) {
            err = e.what(); code = RT_EDEVICE;
        }
        if (!err.empty()) {
            for (auto& sl : slots) {
                if (!sl.r) continue;
                outs.push_back(Out{sl.r, Request::FAILED, nullptr, {}, err, code});
                sl.r.reset();
            }
            in_batch = 0;
            restart_engine();
        }
        if (outs.empty()) return;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& o : outs) publish(*o.r, o.state, o.P, o.st, o.err, o.code);
        }
        cv.notify_all();
    }

    void restart_engine() {
        try {
            if (est) (void)hipStreamSynchronize(est);
            bs.reset();
            for (auto& sl : slots) sl.P.reset();
            bs.reset(new rt::BatchSolver(cfg.max_active, est));
        } catch (const std::exception&) {
            bs.reset();  // later requests are solved on their own streams
        }
    }

    void scheduler() {
        if (cfg.precision != RT_PREC_FP64) {
            try {
                RT_CHECK(hipStreamCreateWithFlags(&est, hipStreamNonBlocking));
                slots.resize(cfg.max_active);
                bs.reset(new rt::BatchSolver(cfg.max_active, est));
            } catch (const std::exception&) {
                bs.reset();
            }
        }
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv.wait(lk, [&] { return stop || !finished.empty() || in_batch > 0 || head_ready(); });
            if (!finished.empty()) reap(lk);
            if (stop) break;
            auto adm = admit();
            lk.unlock();
            if (!adm.empty()) place(adm);
            if (in_batch > 0) step();
            lk.lock();
        }
        // shutdown: requests still in the slots end as cancelled
        lk.unlock();
        std::vector<std::shared_ptr<Request>> left;
        for (int k = 0; k < (int)slots.size(); ++k) {
            if (!slots[k].r) continue;
            left.push_back(slots[k].r);
            try {
                if (bs) bs->remove(k);
            } catch (const std::exception&) {
            }
            slots[k].P.reset();
            slots[k].r.reset();
        }
        in_batch = 0;
        bs.reset();
        lk.lock();
        for (auto& r : left) {
            std::unique_ptr<rt::Problem> none;
            std::vector<rt_group_stats> st;
            publish(*r, Request::CANCELLED, none, st, std::string(), 0);
        }
        // Finished results of the lockstep engine that no poll collected were solved on est and
        // synchronise it when released: release them while the stream still exists.  Nothing can
        // collect them any more (the server is being destroyed).
        std::vector<std::unique_ptr<rt::Problem>> uncollected;
        for (auto& kv : reqs) {
            Request& r = *kv.second;
            if (r.state == Request::DONE && r.result && r.result->st == est) {
                bytes_in_use -= r.bytes;
                uncollected.push_back(std::move(r.result));
                r.state = Request::CANCELLED;
            }
        }
        // results dropped with rt_server_cancel are released by the reaper, which is still running:
        // wait for it, since those solved on est synchronise it when released
        rcv.notify_all();
        cv.wait(lk, [&] { return trash.empty(); });
        lk.unlock();
        uncollected.clear();
        if (est) (void)hipStreamDestroy(est);
        est = nullptr;
        lk.lock();
    }
};

extern "C" {

int rt_server_create(const rt_server_config* cfg, rt_server** out) {
    if (!cfg || !out || cfg->hbm_budget_bytes <= 0 || cfg->max_active < 1) {
        rt::set_last_error("invalid server configuration");
        return RT_EINVAL;
    }
    // a serving session starts from an empty device-memory cache: blocks cached for earlier work of
    // other sizes would only fill the cache's cap and be evicted (each eviction frees, which
    // synchronises the device) while the requests allocate their own
    rt::dev_trim();
    auto* s = new rt_server();
    s->cfg = *cfg;
    s->reaper = std::thread([s] { s->reap_results(); });
    s->sched = std::thread([s] {
        rt::set_autotune_measuring(false);  // served requests overlap on the device
        s->scheduler();
    });
    *out = s;
    return RT_OK;
}

int rt_server_submit(rt_server* s, const rt_problem_desc* p, int64_t* req) {
    if (!s || !p || !req) {
        rt::set_last_error("null argument");
        return RT_EINVAL;
    }
    auto r = std::make_shared<Request>();
    try {
        r->desc.reset(new DescCopy(*p));
        r->desc->d.precision = s->cfg.precision;
        if (r->desc->d.precision != RT_PREC_FP64) r->desc->d.accel = RT_ACCEL_GMRES;
        r->bytes = rt::Problem::device_bytes(r->desc->d);
    } catch (const std::exception& e) {
        rt::set_last_error(e.what());
        return RT_EINVAL;
    }
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->queue.push_back(r->id);
        *req = r->id;
    }
    s->cv.notify_all();
    return RT_OK;
}

int rt_server_poll(rt_server* s, int64_t req, double* phi_out, rt_group_stats* stats_out, double* latency_ms) {
    return rt_server_poll_cutoff(s, req, phi_out, nullptr, stats_out, latency_ms);
}

int rt_server_poll_cutoff(rt_server* s, int64_t req, double* phi_out, double* cutoff_out, rt_group_stats* stats_out,
                          double* latency_ms) {
    if (!s) {
        rt::set_last_error("null server");
        return RT_EINVAL;
    }
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req);
        if (it == s->reqs.end()) {
            rt::set_last_error("unknown request id");
            return RT_EINVAL;
        }
        r = it->second;
        if (r->state == Request::QUEUED || r->state == Request::RUNNING) return RT_EPENDING;
        if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        if (stats_out && r->state == Request::DONE)
            std::memcpy(stats_out, r->stats.data(), r->stats.size() * sizeof(rt_group_stats));
        if (r->state == Request::DONE && !phi_out) return RT_OK;  // status query: the flux stays
        s->reqs.erase(it);  // a final status is reported once; the id is released here
        if (r->state == Request::CANCELLED) return RT_ECANCELLED;
        if (r->state == Request::FAILED) {
            rt::set_last_error(r->err);
            return r->code ? r->code : RT_EDEVICE;
        }
    }
    // Collection: the flux is copied out of device memory on a stream of this call, outside the
    // server's lock, then the request's memory returns to the budget.
    int rc = guard([&] {
        hipStream_t cs = nullptr;
        RT_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        try {
            r->result->get_flux(phi_out, cs);
            if (cutoff_out) r->result->get_flux_cutoff(cutoff_out, cs);
        } catch (...) {
            (void)hipStreamDestroy(cs);
            throw;
        }
        RT_CHECK(hipStreamDestroy(cs));
    });
    r->result.reset();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->bytes_in_use -= r->bytes;
    }
    s->cv.notify_all();
    return rc;
}

int rt_server_cancel(rt_server* s, int64_t req) {
    if (!s) return RT_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req);
        if (it == s->reqs.end()) {
            rt::set_last_error("unknown request id");
            return RT_EINVAL;
        }
        Request& r = *it->second;
        r.cancel.store(true);
        if (r.state == Request::DONE && r.result) {  // a finished, uncollected result: released by the reaper
            s->trash.emplace_back(std::move(r.result), r.bytes);
            r.state = Request::CANCELLED;
        }
    }
    s->rcv.notify_all();
    s->cv.notify_all();
    return RT_OK;
}

int rt_server_destroy(rt_server* s) {
    if (!s) return RT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->sched.joinable()) s->sched.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->reaper_stop = true;
    }
    s->rcv.notify_all();
    if (s->reaper.joinable()) s->reaper.join();
    for (auto& kv : s->workers)
        if (kv.second.joinable()) kv.second.join();
    delete s;
    return RT_OK;
}

}  // extern "C"
