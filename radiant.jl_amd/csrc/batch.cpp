// batch.cpp -- lockstep solver for many independent problems (mixed precision, reduced state).
//
// Every round, each unfinished problem performs exactly one transport pass of its current
// phase (c = T(0), a low-precision GMRES matvec, an FP64 refinement residual or the FP64
// reconstruction); the passes of all problems share the same kernel launches (the contraction
// and sweep task tables carry the problem index).  The Arnoldi step is classical Gram-Schmidt in
// FP64: h = V^T w, then w -= V h with |w|^2 in the same pass over the basis, batched over
// problems with per-problem deterministic reductions.  The inner GMRES only has to reach 1e-5
// (3e-3 with BF16) and its result is corrected by FP64 residuals, so the loss of orthogonality
// of single-pass CGS (of order cond * eps) is immaterial; RADIANT_HIP_CGS2=1 re-orthogonalises
// once more (CGS2) for cross-checks.  The host reads every problem's scalars with one copy per round and
// advances each problem's state machine (the logic of sn_flux.jl:172-205 + sn_one_speed.jl
// with the iterative refinement of the mixed-precision mode).
#include "batch.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>

namespace rt {

namespace {
constexpr int SLOTS = 128;  // per problem: [0,40) h1, [40,80) h2, 80 |w|^2, 81 generic norm^2
constexpr int S_H1 = 0, S_H2 = 40, S_WN = 80, S_NRM = 81;
}  // namespace

BatchSolver::BatchSolver(const std::vector<Problem*>& probs, hipStream_t stream) : st(stream) {
    init_device((int)probs.size());
    for (Problem* P : probs) add(P);
    RT_CHECK(hipStreamSynchronize(st));
}

BatchSolver::BatchSolver(int capacity, hipStream_t stream) : st(stream) {
    if (capacity < 1) throw Error(RT_EINVAL, "batched solver: capacity must be positive");
    init_device(capacity);
    RT_CHECK(hipStreamSynchronize(st));
}

void BatchSolver::init_device(int n) {
    ps.resize(n);
    for (auto& s : ps) s.phase = DONE;
    auto al = [&](size_t bytes) {
        void* p = dev_alloc(std::max<size_t>(bytes, 16));
        mem.push_back(p);
        return p;
    };
    d_pd = (PassDesc*)al(sizeof(PassDesc) * n);
    d_wl = (unsigned*)al(128 * sizeof(unsigned));
    RT_CHECK(hipStreamCreateWithFlags(&st64, hipStreamNonBlocking));
    {
        const char* e = std::getenv("RADIANT_HIP_CGS2");  // re-orthogonalise every Arnoldi step
        reorth = e && e[0] == '1';
        // 1: FP64 passes on a second stream beside the low-precision ones.  Off by default: the
        // sub-batches of a batch already overlap on the device, and measured faster without it.
        const char* f = std::getenv("RADIANT_HIP_FORK");
        fork_streams = f && f[0] == '1';
        // 1: every round's launches are captured into a HIP graph and run as one graph launch
        const char* gr = std::getenv("RADIANT_HIP_GRAPHS");
        use_graphs = gr && gr[0] == '1';
    }
    RT_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    RT_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    d_slots = (double*)al(sizeof(double) * SLOTS * n);
    d_partial = (double*)al(sizeof(double) * GS_BPI_MAX * GS_KMAX * (size_t)n);
    RT_CHECK(hipHostMalloc((void**)&h_slots, sizeof(double) * SLOTS * n, hipHostMallocDefault));
    arena_cap = (sizeof(PassDesc) + 4 * sizeof(GsItem) + 16 * sizeof(VecItem) + 64 * sizeof(double) +
                 64 * sizeof(int)) * (size_t)n + (1 << 20);
    d_arena = (unsigned char*)al(arena_cap);
    RT_CHECK(hipHostMalloc((void**)&h_arena, arena_cap, hipHostMallocDefault));
}

void BatchSolver::ensure_task_capacity() {
    size_t ng = 0, nsw = 0;
    for (auto& s : ps) {
        if (!s.P) continue;
        const Problem* P = s.P;
        const size_t ntile = (P->Nc + GEMM_CELLS - 1) / GEMM_CELLS;
        ng += ntile * ((std::max(P->Nd, P->Np) + GEMM_ROWS - 1) / GEMM_ROWS);
        nsw += (size_t)P->Nd;
    }
    if (ng <= cap_gemm && nsw <= cap_sweep) return;
    // the tables are rewritten every round: after the stream drains they can be replaced
    RT_CHECK(hipStreamSynchronize(st));
    if (st64) RT_CHECK(hipStreamSynchronize(st64));
    for (void*& p : task_mem)
        if (p) {
            dev_free(p);
            p = nullptr;
        }
    cap_gemm = std::max(cap_gemm, ng + ng / 2);
    cap_sweep = std::max(cap_sweep, nsw + nsw / 2);
    auto al = [&](size_t bytes) { return dev_alloc(std::max<size_t>(bytes, 16)); };
    for (int pr = 0; pr < 3; ++pr) {
        task_mem[3 * pr] = d_mg[pr] = (GemmTask*)al(sizeof(GemmTask) * cap_gemm);
        task_mem[3 * pr + 1] = d_dg[pr] = (GemmTask*)al(sizeof(GemmTask) * cap_gemm);
        task_mem[3 * pr + 2] = d_sw[pr] = (SweepTask*)al(sizeof(SweepTask) * cap_sweep);
    }
}

int BatchSolver::add(Problem* P) {
    if (P->precision == RT_PREC_FP64)
        throw Error(RT_EINVAL, "the batched solver runs the mixed-precision modes");
    if (P->em) throw Error(RT_EINVAL, "a magnetic field runs in the single-problem solve, not in lockstep batches");
    if (P->hog) throw Error(RT_EINVAL, "closures beyond the register walkers run in the single-problem solve");
    if (!P->ho_K && (P->scan_cpl <= 0 || sweep_scan_lds_bytes(P64, P->ndims, P->Nx, P->Ny, P->Nz, P->Nmat) > 160 * 1024))
        throw Error(RT_EINVAL, "a batched problem does not fit the row-scan sweep");
    int slot = -1;
    for (int i = 0; i < (int)ps.size(); ++i)
        if (!ps[i].P) {
            slot = i;
            break;
        }
    if (slot < 0) throw Error(RT_ENOMEM, "batched solver: no free slot");
    PS& s = ps[slot];
    s = PS{};
    s.P = P;
    s.m = std::min(std::max(1, P->restart), 32);
    s.lo = (P->precision == RT_PREC_MIXED_BF16) ? PBF16 : P32;
    s.inner_tol = inner_tol_for(s.lo);
    const size_t S = P->S_r;
    for (int k = 0; k <= s.m; ++k) s.V.push_back(P->vec("bV" + std::to_string(k), S));
    s.w = P->vec("bw", S);
    s.c = P->vec("bc", S);
    s.z = P->vec("bz", S);
    s.r = P->vec("br", S);
    s.t = P->vec("bt", S);
    s.psiC = P->vec("bpsiC", (size_t)P->Nd * P->Nc);
    RT_CHECK(hipMemcpy(P->d_Vptr, s.V.data(), (s.m + 1) * sizeof(double*), hipMemcpyHostToDevice));
    s.stats.assign(P->G, rt_group_stats{});
    try {
        ensure_task_capacity();
    } catch (...) {
        s = PS{};
        s.phase = DONE;
        throw;
    }
    start(slot);
    return slot;
}

void BatchSolver::start(int slot) {
    PS& s = ps[slot];
    s.P->counters = SolveCounters{};
    // no clearing of the flux: a group's out-of-group source reads only the groups before it
    // (outsrc_kernel), and every group's moments are written whole before that: by the
    // reconstruction's full-row contraction, or by the fill of a group without source, which
    // round() issues ahead of the next group's start
    s.phase = START;
    s.g = 0;
    std::fill(s.stats.begin(), s.stats.end(), rt_group_stats{});
}

// Vector operations recorded for the next round may write into the slot's memory: they run
// before the problem is released.
void BatchSolver::remove(int slot) {
    drain_pre();
    PS& s = ps[slot];
    s = PS{};
    s.phase = DONE;
}

void BatchSolver::drain_pre() {
    if (pre.empty()) return;
    arena_off = 0;
    run_vecops(pre);
    pre.clear();
    flush();
    RT_CHECK(hipStreamSynchronize(st));
}

// Copies host data into the staging arena and returns the device address it will have after
// the next flush().  The arena is reused only after the end-of-round synchronisation.
template <class T>
T* BatchSolver::stage(const T* data, size_t count) {
    const size_t bytes = count * sizeof(T);
    const size_t off = (arena_off + 255) & ~size_t(255);
    if (off + bytes > arena_cap) throw Error(RT_ENOMEM, "batched solver: staging arena exhausted");
    std::memcpy(h_arena + off, data, bytes);
    arena_off = off + bytes;
    return (T*)(d_arena + off);
}

void BatchSolver::flush() {
    if (use_graphs && !profiling && !fork_streams) {
        flush_graph();
        return;
    }
    if (arena_off > 0) RT_CHECK(hipMemcpyAsync(d_arena, h_arena, arena_off, hipMemcpyHostToDevice, st));
    for (auto& f : plan) f();
    plan.clear();
}

// Graph form of flush(): the arena upload and the recorded launches are captured from the stream
// into a graph.  Rounds of one shape (same node kinds and kernels in the same order: the same
// phase mix of the same problems) reuse an executable graph, updated in place with the new
// launch parameters; any other shape is instantiated and cached.  The captured work has no
// memset nodes (buffers are cleared by fill kernels) and no events (timing and the autotuner's
// measurements are off while capturing; tuned or default shapes are used).
void BatchSolver::flush_graph() {
    if (arena_off == 0 && plan.empty()) return;
    const bool meas = autotune_measuring();
    set_autotune_measuring(false);
    RT_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    hipGraph_t g = nullptr;
    try {
        if (arena_off > 0) RT_CHECK(hipMemcpyAsync(d_arena, h_arena, arena_off, hipMemcpyHostToDevice, st));
        for (auto& f : plan) f();
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        set_autotune_measuring(meas);
        plan.clear();
        throw;
    }
    set_autotune_measuring(meas);
    plan.clear();
    RT_CHECK(hipStreamEndCapture(st, &g));
    // shape of the captured graph: node kinds, and the kernel of each kernel node, in order
    size_t nn = 0;
    RT_CHECK(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    if (nn) RT_CHECK(hipGraphGetNodes(g, nodes.data(), &nn));
    std::vector<uintptr_t> sig;
    sig.reserve(2 * nn);
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType ty;
        RT_CHECK(hipGraphNodeGetType(nd, &ty));
        sig.push_back((uintptr_t)ty);
        if (ty == hipGraphNodeTypeKernel) {
            hipKernelNodeParams kp{};
            RT_CHECK(hipGraphKernelNodeGetParams(nd, &kp));
            sig.push_back((uintptr_t)kp.func);
        }
    }
    hipGraphExec_t ex = nullptr;
    for (size_t c = 0; c < gcache.size(); ++c) {
        if (gcache[c].sig != sig) continue;
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(gcache[c].exec, g, &bad, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess) {
            ex = gcache[c].exec;
            ++graph_updates;
        } else {  // not updatable after all: dropped, instantiated afresh below
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(gcache[c].exec);
            gcache.erase(gcache.begin() + (long)c);
        }
        break;
    }
    if (!ex) {
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            throw Error(RT_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        if (gcache.size() >= GRAPH_CACHE) {
            (void)hipGraphExecDestroy(gcache.front().exec);
            gcache.erase(gcache.begin());
        }
        gcache.push_back({sig, ex});
        ++graph_instantiations;
    }
    (void)hipGraphDestroy(g);
    RT_CHECK(hipGraphLaunch(ex, st));
}

// The contractions of a round with FP64 and FP32 passes run as one launch of the mixed-precision
// kernel (each workgroup takes its problem's precision).  RADIANT_HIP_SPLIT_GEMM=1 runs them as
// an FP64 launch and then an FP32 launch, each kernel at its own occupancy instead; the results
// are the same bits, and on the 256-problem 2D batch the single mixed launch measured faster
// (1021 vs 1037 ms per batch): the FP64 launch alone leaves the device underfilled.
static bool mixed_gemm() {
    static const bool on = [] {
        const char* e = std::getenv("RADIANT_HIP_SPLIT_GEMM");
        return !(e && e[0] == '1');
    }();
    return on;
}

std::function<void()> BatchSolver::timed(int kernel, int prec, double upd, double flop, double bytes,
                                         hipStream_t s, std::function<void()> fn) {
    if (!profiling) return fn;
    while (ev_pool.size() < ev_next + 2) {
        hipEvent_t e;
        RT_CHECK(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    hipEvent_t a = ev_pool[ev_next], b = ev_pool[ev_next + 1];
    ev_next += 2;
    ev_used.push_back({a, b, kernel, prec, upd, flop, bytes});
    (void)s;
    return [=] {
        set_launch_events(a, b);  // recorded at the kernel's own dispatch start and end
        fn();
        set_launch_events(nullptr, nullptr);
    };
}

void BatchSolver::collect_times() {
    for (auto& r : ev_used) {
        float ms = 0.f;
        RT_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        timing.ms[r.prec][r.kernel] += ms;
        timing.n[r.prec][r.kernel] += 1;
        timing.upd[r.prec][r.kernel] += r.upd;
        timing.flop[r.prec][r.kernel] += r.flop;
        timing.bytes[r.prec][r.kernel] += r.bytes;
    }
    ev_used.clear();
    ev_next = 0;
}

BatchSolver::~BatchSolver() {
    (void)hipStreamSynchronize(st);
    if (st64) (void)hipStreamSynchronize(st64);  // no pending work uses the blocks given back below
    for (auto& c : gcache) (void)hipGraphExecDestroy(c.exec);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (void* p : task_mem)
        if (p) dev_free(p);
    if (st64) {
        (void)hipStreamSynchronize(st64);
        (void)hipStreamDestroy(st64);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    for (void* p : mem) dev_free(p);
    if (h_slots) (void)hipHostFree(h_slots);
    if (h_arena) (void)hipHostFree(h_arena);
}

// Runs host-built item lists through the device.
void BatchSolver::run_vecops(const std::vector<VecItem>& items) {
    if (items.empty()) return;
    VecItem* d = stage(items.data(), items.size());
    long nmax = 0;
    for (auto& it : items) nmax = std::max(nmax, it.n);
    const int bpi = (int)std::min<long>(256, std::max<long>(1, (nmax + 2047) / 2048));
    const int ni = (int)items.size();
    hipStream_t s = st;
    plan.push_back([=] { launch_vecops(d, ni, bpi, s); });
}

void BatchSolver::run_gs(const std::vector<GsItem>& items, bool update, bool norm) {
    if (items.empty()) return;
    int kmax = 1, bmax = 1;
    for (auto& it : items) {
        kmax = std::max(kmax, it.k);
        bmax = std::max(bmax, it.bpi);
    }
    GsItem* d = stage(items.data(), items.size());
    const int ni = (int)items.size();
    double* part = d_partial;
    hipStream_t s = st;
    plan.push_back([=] { launch_gs(d, ni, kmax, bmax, update, norm, part, s); });
}

// RADIANT_HIP_LPT=0 keeps the sweep tasks in problem order (for measuring the ordering).
static bool sweep_lpt() {
    static const bool on = [] {
        const char* e = std::getenv("RADIANT_HIP_LPT");
        return !(e && e[0] == '0');
    }();
    return on;
}

void BatchSolver::launch_passes(const std::vector<int>& who, const std::vector<PassDesc>& pds,
                                const std::vector<int>& prec, const std::vector<int>& full) {
    const int n = (int)ps.size();
    // PassDesc of every problem (index = batch index).  A pass of a problem with the angular
    // Fokker-Planck term on a non-zero state reads its source from the problem's fpq buffer,
    // qlout + T ℳ Φ, formed by fp_source_kernel at the start of the round (a reconstruction
    // runs no M contraction: its angular flux already holds the term).
    std::vector<PassDesc> all(n);
    std::vector<int> fpw;
    for (size_t i = 0; i < who.size(); ++i) {
        all[who[i]] = pds[i];
        PassDesc& a = all[who[i]];
        if (a.fpT && a.phi_in && full[i] != 2) {
            a.fp_qin = a.qlout;
            a.fpq = ps[who[i]].P->d_fpq;
            a.qlout = a.fpq;
            fpw.push_back(who[i]);
        }
    }
    const PassDesc* dpd = stage(all.data(), all.size());
    // With fork_streams the FP64 passes (refinement residuals, reconstructions, c = T(0)) of this
    // round run on st64, concurrently with the low-precision Arnoldi passes on st; st waits for
    // them before the Krylov work.  Every group has its own task tables.
    bool has64 = false, haslo = false, has32 = false;
    for (int p : prec) {
        (p == P64 ? has64 : haslo) = true;
        if (p == P32) has32 = true;
    }
    const bool fork = has64 && haslo && fork_streams;
    // Without the fork, the FP64 and FP32 passes of a round share one launch of each kernel
    // (every workgroup takes its problem's precision): a sweep costs the latency of its 128-ish
    // dependent rows whatever the number of problems in it.  Such launches are recorded in the
    // FP32 row of the kernel timing, with FP64 updates counted twice (FP32-equivalent bytes).
    const bool mix = has64 && has32 && !fork;
    auto group_of = [&](int p) { return (mix && (p == P64 || p == P32)) ? (int)PMIX : p; };
    // Non-void faces and surface sources: each such problem's incoming face fluxes are formed
    // from its state's boundary moments (and sources) before the sweeps, and after the last D
    // contraction one launch writes every problem's exit-face moments.
    for (int w : fpw) {
        const PassDesc* one = dpd + w;
        const long na = (long)ps[w].P->Nact * ps[w].P->Nc;
        hipStream_t s0 = st;
        plan.push_back([=] { launch_fp_source(one, na, s0); });
    }
    bool any_bout = false;
    int max_nb = 0;
    for (size_t i = 0; i < who.size(); ++i) {
        const Problem* P = ps[who[i]].P;
        if (pds[i].bin_x && full[i] != 2) {
            const PassDesc* one = dpd + who[i];
            const int tot = (int)P->bin_total;
            hipStream_t s0 = st;
            plan.push_back([=] { launch_bndin(one, tot, s0); });
        }
        if (pds[i].bout_x) {
            any_bout = true;
            max_nb = std::max(max_nb, (int)P->nb);
        }
    }
    if (fork) {
        hipStream_t s0 = st, s1 = st64;
        hipEvent_t ef = ev_fork;
        plan.push_back([=] {
            RT_CHECK(hipEventRecord(ef, s0));
            RT_CHECK(hipStreamWaitEvent(s1, ef, 0));
        });
    }
    for (int pr : {(int)PMIX, (int)P64, (int)P32, (int)PBF16}) {
        const int tix = pr == PMIX ? (int)P32 : pr;  // task tables and timing row of the group
        std::vector<int> mmeta, dmeta;
        // (cells per lane, waves, precision of an inflow-taking sweep or -1) -> sweep meta: the
        // sweeps that take boundary inflows run in their own launches, one per precision
        using SKey = std::array<int, 3>;
        std::map<SKey, std::vector<int>> sweeps;
        int moff = 0, doff = 0;
        std::map<SKey, int> soff;
        std::map<SKey, size_t> slds;
        std::map<SKey, double> s_upd;
        double g_upd = 0, g_mflop = 0, g_dflop = 0, d_upd = 0;
        // algorithmic bytes: M reads the Nact source moments (8 B) and writes the angular source
        // (es); the sweep reads and writes the angular flux; D reads it, writes its output rows
        // and, for an I - A pass, reads the input state's rows
        double g_mby = 0, d_by = 0;
        std::map<SKey, double> s_by;
        std::vector<size_t> ho_here;  // problems of this group with a closure above order 1
        // In a round with FP64 and FP32 passes the sweeps share one launch (latency-bound); with
        // the split the contractions run as an FP64 launch and then an FP32 launch, the FP64
        // problems' tasks first in the task tables.
        const bool split = pr == PMIX && !mixed_gemm();
        int moff64 = 0, doff64 = 0;
        for (int part = 0; part < (split ? 2 : 1); ++part) {
        if (part == 1) { moff64 = moff; doff64 = doff; }
        for (size_t i = 0; i < who.size(); ++i) {
            if (group_of(prec[i]) != pr) continue;
            if (split && (prec[i] == P64) != (part == 0)) continue;
            const Problem* P = ps[who[i]].P;
            const int ntile = (P->Nc + GEMM_CELLS - 1) / GEMM_CELLS;
            const int nrg_m = (P->Nd + GEMM_ROWS - 1) / GEMM_ROWS;
            const int rows = full[i] ? P->Np : P->Nact;
            const int nrg_d = (rows + GEMM_ROWS - 1) / GEMM_ROWS;
            dmeta.insert(dmeta.end(), {who[i], ntile, nrg_d, doff});
            doff += ntile * nrg_d;
            const double upd = (double)P->Nc * P->Nd * ((pr == PMIX && prec[i] == P64) ? 2.0 : 1.0);
            d_upd += upd;
            g_dflop += 2.0 * rows * upd;
            const double es = prec[i] == P64 ? 8.0 : 4.0, ndc = (double)P->Nd * P->Nc;
            d_by += es * ndc + 8.0 * rows * (double)P->Nc * (pds[i].zin ? 2.0 : 1.0);
            if (full[i] == 2) continue;  // contraction only: no moment->direction contraction, no sweep
            mmeta.insert(mmeta.end(), {who[i], ntile, nrg_m, moff});
            moff += ntile * nrg_m;
            if (P->ho_K) {  // a closure above order 1: its own sweep kernel, in its pass's precision
                ho_here.push_back(i);
                g_upd += upd;
                g_mflop += 2.0 * P->Nact * upd;
                g_mby += es * ndc + 8.0 * P->Nact * (double)P->Nc;
                continue;
            }
            const int bprec = pds[i].bin_x ? (prec[i] == P64 ? (int)P64 : (int)P32) : -1;
            const SKey key{P->scan_cpl, P->scan_waves, bprec};
            auto& v = sweeps[key];
            v.insert(v.end(), {who[i], P->Nd, soff[key]});
            soff[key] += P->Nd;
            g_upd += upd;
            g_mflop += 2.0 * P->Nact * upd;
            s_upd[key] += upd;
            g_mby += es * ndc + 8.0 * P->Nact * (double)P->Nc;
            s_by[key] += 2.0 * es * ndc;
            const int sp = bprec >= 0 ? bprec : (pr == PBF16 ? (int)P32 : (pr == PMIX ? (int)P64 : pr));
            slds[key] = std::max(slds[key], sweep_scan_lds_bytes(sp, P->ndims, P->Nx, P->Ny, P->Nz, P->Nmat));
        }
        }
        if (dmeta.empty()) continue;
        // Largest problems first in every sweep launch: workgroups are dispatched in task order,
        // so the directions of the problems with the most cells start first and the launch does
        // not end on a tail of long directions started last (longest-processing-time order of an
        // uneven batch).  The order changes no result: each task is one independent direction.
        if (sweep_lpt())
            for (auto& kv : sweeps) {
                std::vector<int>& v = kv.second;
                std::vector<std::array<int, 2>> e;
                for (size_t k = 0; k + 2 < v.size(); k += 3) e.push_back({v[k], v[k + 1]});
                std::stable_sort(e.begin(), e.end(), [&](const std::array<int, 2>& a, const std::array<int, 2>& b) {
                    return ps[a[0]].P->Ncell > ps[b[0]].P->Ncell;
                });
                v.clear();
                int off = 0;
                for (auto& x : e) {
                    v.insert(v.end(), {x[0], x[1], off});
                    off += x[1];
                }
            }
        // task tables expanded on the device from the small per-problem metadata
        const int nm = (int)mmeta.size() / 4, nd = (int)dmeta.size() / 4;
        const int* dmm = nm ? stage(mmeta.data(), mmeta.size()) : nullptr;
        const int* ddm = stage(dmeta.data(), dmeta.size());
        GemmTask *mg = d_mg[tix], *dg = d_dg[tix];
        SweepTask* swbase = d_sw[tix];
        hipStream_t s = (fork && pr == P64) ? st64 : st;
        const int ps_ = pr;
        plan.push_back([=] {
            if (nm) launch_expand_gemm(dmm, nm, mg, s);
            launch_expand_gemm(ddm, nd, dg, s);
        });
        // the split launches of a mixed round: FP64 tasks [0, n64), FP32 tasks [n64, n); the
        // timing events of the pair start at the first launch and stop at the last
        auto gemm2 = [split](void (*fn)(int, const PassDesc*, const GemmTask*, int, hipStream_t), int prec_,
                             const PassDesc* pd, GemmTask* t, int n64, int n, hipStream_t s_) {
            if (!split || n64 == 0 || n64 == n) {
                fn(split ? (n64 == n ? (int)P64 : (int)P32) : prec_, pd, t, n, s_);
                return;
            }
            hipEvent_t a, b;
            take_launch_events(&a, &b);
            set_launch_events(a, nullptr);
            fn(P64, pd, t, n64, s_);
            set_launch_events(nullptr, b);
            fn(P32, pd, t + n64, n - n64, s_);
            set_launch_events(nullptr, nullptr);
        };
        if (nm)
            plan.push_back(timed(0, tix, g_upd, g_mflop, g_mby, s, [=] { gemm2(launch_mgemm, ps_, dpd, mg, moff64, moff, s); }));
        // slowing down (FP64 inhomogeneous passes): the energy inflow joins the angular source
        // before the sweep, and the energy outflow 2 psi - psiE_in is formed after it
        std::vector<std::pair<const PassDesc*, long>> csd_here;
        for (size_t i = 0; i < who.size(); ++i)
            if (group_of(prec[i]) == pr && pds[i].psiE_in && full[i] != 2 && !ps[who[i]].P->bfp)
                csd_here.push_back({dpd + who[i], (long)ps[who[i]].P->Nd * ps[who[i]].P->Nc});
        if (!csd_here.empty())
            plan.push_back([=] {
                for (auto& c : csd_here) launch_csd_source(c.first, c.second, s);
            });
        for (auto& kv : sweeps) {
            const std::vector<int>& sm = kv.second;
            const int* dsm = stage(sm.data(), sm.size());
            const int nsm = (int)sm.size() / 3;
            const int ntask = soff[kv.first];
            const int cpl = kv.first[0], waves = kv.first[1], bprec = kv.first[2];
            const size_t lds = slds[kv.first];
            SweepTask* sb = swbase;
            unsigned* wl = d_wl + (s == st64 ? 64 : 0);  // work-list counter of this stream
            plan.push_back([=] { launch_expand_sweep(dsm, nsm, sb, s, wl); });
            plan.push_back(timed(1, tix, s_upd[kv.first], 0.0, s_by[kv.first], s, [=] {
                const int kind = bprec >= 0 ? bprec : (ps_ == PBF16 ? (int)P32 : ps_);
                set_sweep_worklist(wl);
                launch_sweep_scan(kind, dpd, sb, ntask, cpl, waves, lds, s, bprec >= 0);
                set_sweep_worklist(nullptr);
            }));
            swbase += ntask;
        }
        for (size_t i : ho_here) {
            const Problem* P = ps[who[i]].P;
            const PassDesc* one = dpd + who[i];
            const int sp = prec[i] == P64 ? (int)P64 : (int)P32;
            const int nd_ = P->Nd, K = P->ho_K, kx = P->ho_Kx, ky = P->ho_Ky, kz = P->ho_Kz, dims = P->ndims;
            const bool fc = P->ho_fc, bfp = P->bfp;
            const int ke = P->ho_KE;
            const size_t lds = P->ho_lds(sp);
            const double hu = (double)P->Ncell * P->Nd * ((pr == PMIX && prec[i] == P64) ? 2.0 : 1.0);
            const double hb = 2.0 * (prec[i] == P64 ? 8.0 : 4.0) * (double)P->Nd * P->Nc;
            plan.push_back(timed(1, tix, hu, 0.0, hb, s, [=] {
                if (bfp) launch_sweep_bfp(sp, one, nd_, ke, kx, ky, kz, fc, lds, s);
                else if (dims == 3) launch_sweep_3d_ho(sp, one, nd_, kx, ky, kz, fc, lds, s);
                else if (dims == 2) launch_sweep_2d_ho(sp, one, nd_, kx, ky, fc, lds, s);
                else launch_sweep_1d_ho(sp, one, nd_, K, s);
            }));
        }
        if (!csd_here.empty())
            plan.push_back([=] {
                for (auto& c : csd_here) launch_csd_out(c.first, c.second, s);
            });
        plan.push_back(timed(2, tix, d_upd, g_dflop, d_by, s, [=] { gemm2(launch_dgemm, ps_, dpd, dg, doff64, doff, s); }));
    }
    if (fork) {
        hipStream_t s0 = st, s1 = st64;
        hipEvent_t ej = ev_join;
        plan.push_back([=] {
            RT_CHECK(hipEventRecord(ej, s1));
            RT_CHECK(hipStreamWaitEvent(s0, ej, 0));
        });
    }
    // the reconstruction's angular flux is T(z)'s (the residual pass's plus c = T(0)'s, summed
    // before this round): its energy outflow is the group's
    for (size_t i = 0; i < who.size(); ++i)
        if (full[i] == 2 && pds[i].psiE_out) {
            const PassDesc* one = dpd + who[i];
            const Problem* P = ps[who[i]].P;
            const long ne = P->bfp ? (long)P->nE() : (long)P->Nd * P->Nc;
            const bool bfp = P->bfp;
            hipStream_t s0 = st;
            plan.push_back([=] {
                if (bfp) launch_bfp_out(one, ne, s0);
                else launch_csd_out(one, ne, s0);
            });
        }
    if (any_bout) {
        const int nall = n;
        hipStream_t s0 = st;
        plan.push_back([=] { launch_bndmom(dpd, nall, max_nb, s0); });
    }
    for (size_t i = 0; i < who.size(); ++i) {
        Problem* P = ps[who[i]].P;
        P->counters.passes += 1;
        if (prec[i] != P64) P->counters.passes_lo += 1;
        if (full[i] != 2) P->counters.updates_raw += (double)P->Ncell * P->Nd;  // sweeps only
    }
}

// One round: every unfinished problem does one pass of its current phase, then its Krylov work;
// the host reads the scalars and advances each problem's state machine.
bool BatchSolver::round() {
    using clk = std::chrono::steady_clock;
    auto us = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::micro>(b - a).count(); };
    const int n = (int)ps.size();
    const auto t0 = clk::now();
    // Vector operations recorded for this round run first: among them is the flux fill of a
    // group without source, whose moments the next group's out-of-group source reads.
    arena_off = 0;
    run_vecops(pre);
    pre.clear();
    // ---- group starts: out-of-group + volume source (sn_flux.jl:175-179)
    for (auto& s : ps) {
        if (s.phase != START) continue;
        Problem* P = s.P;
        const int g = s.g;
        hipStream_t s0 = st;
        plan.push_back([=] {
            launch_outsrc(P->d_qlout, P->d_phi, P->d_sigo, P->d_qvol, P->d_mat, P->d_act, g, P->G, P->Nmat, P->Nact,
                          P->Np, P->Nc, s0);
            if (P->csd) {  // the energy flux enters group g rescaled by ΔE_g/ΔE_{g-1}; zero for the first (sn_flux.jl:166,183)
                const size_t ne = P->nE();
                if (g == 0) {
                    k_fill(P->d_psiE_in, 0.0, ne, s0);
                } else {
                    k_copy(P->d_psiE_in, P->d_psiE_out, ne, s0);
                    k_scale(P->d_psiE_in, P->dE[g] / P->dE[g - 1], ne, s0);
                }
            }
        });
        s.phase = CPASS;
        s.niter = 0;
        s.refs = 0;
        s.p0 = P->counters.passes;
        s.l0 = P->counters.passes_lo;
    }
    // ---- one pass per active problem
    std::vector<int> who, prec, full;
    std::vector<PassDesc> pds;
    for (int i = 0; i < n; ++i) {
        PS& s = ps[i];
        if (s.phase == DONE) continue;
        Problem* P = s.P;
        PassDesc pd;
        switch (s.phase) {
            case CPASS:  // c = T(0); its angular flux stays in psiC for the reconstruction
                P->desc_reduced(pd, s.g, P64, MODE_T, nullptr, s.c);
                pd.psi = s.psiC;
                prec.push_back(P64); full.push_back(0);
                break;
            case ARNOLDI: P->desc_reduced(pd, s.g, s.lo, MODE_IMA, s.V[s.j], s.w); prec.push_back(s.lo); full.push_back(0); break;
            case RESID: P->desc_reduced(pd, s.g, P64, MODE_IMA, s.z, s.t); prec.push_back(P64); full.push_back(0); break;
            // T(z) = A z + T(0): the angular flux of A z is still in psi from the residual pass on the
            // same z, psiC was added to it before this round, so only the full-row D contraction runs
            case RECON: P->desc_reconstruct(pd, s.g, s.z); prec.push_back(P64); full.push_back(2); break;
            default: continue;
        }
        who.push_back(i);
        pds.push_back(pd);
    }
    if (who.empty()) {
        flush();  // vector operations recorded for problems that have just finished
        RT_CHECK(hipStreamSynchronize(st));
        collect_times();
        return false;
    }
    ++rounds;
    launch_passes(who, pds, prec, full);
    // ---- Krylov work after the pass: norms of c and r; CGS2 for the Arnoldi steps
    std::vector<VecItem> v1, v4;
    std::vector<GsItem> gn, g1, g2, g3;
    for (int i : who) {
        PS& s = ps[i];
        double* slots = d_slots + (size_t)SLOTS * i;
        const long S = (long)s.P->S_r;
        const int bpi = gs_blocks(S);
        auto gs = [&](double* y, int k, const double* coef, double* out) {
            GsItem it{};
            it.V = (const double* const*)s.P->d_Vptr; it.coef = coef; it.y = y; it.out = out;
            it.n = S; it.k = k; it.bpi = bpi;
            return it;
        };
        if (s.phase == CPASS) gn.push_back(gs(s.c, 0, nullptr, slots + S_NRM));
        if (s.phase == RESID) {
            VecItem it{};
            it.op = VOP_SUB; it.n = S; it.y = s.r; it.x = s.c; it.x2 = s.t;
            v1.push_back(it);
            gn.push_back(gs(s.r, 0, nullptr, slots + S_NRM));
        }
        if (s.phase == ARNOLDI) {
            const int k = s.j + 1;
            g1.push_back(gs(s.w, k, nullptr, slots + S_H1));        // h1 = V^T w
            if (reorth) {
                g2.push_back(gs(s.w, k, slots + S_H1, slots + S_H2));   // w -= V h1 ; h2 = V^T w
                g3.push_back(gs(s.w, k, slots + S_H2, slots + S_WN));   // w -= V h2 ; |w|^2
            } else {
                g3.push_back(gs(s.w, k, slots + S_H1, slots + S_WN));   // w -= V h1 ; |w|^2
            }
            VecItem nm{};
            nm.op = VOP_NORMALIZE; nm.n = S; nm.y = s.V[s.j + 1]; nm.x = s.w; nm.slot = slots + S_WN;
            v4.push_back(nm);
        }
    }
    run_vecops(v1);
    run_gs(gn, false, true);
    run_gs(g1, false, false);
    run_gs(g2, true, false);
    run_gs(g3, true, true);
    run_vecops(v4);
    const auto t1 = clk::now();
    flush();
    RT_CHECK(hipMemcpyAsync(h_slots, d_slots, sizeof(double) * SLOTS * n, hipMemcpyDeviceToHost, st));
    const auto t2 = clk::now();
    RT_CHECK(hipStreamSynchronize(st));
    const auto t3 = clk::now();
    collect_times();
    // ---- host state machines
    std::vector<double> coefs;
    std::vector<VecItem> adds;
    std::vector<int> add_who;
    for (int i : who) advance(i, pre, adds, add_who, coefs);
    const auto t4 = clk::now();
    host_plan += us(t0, t1);
    host_submit += us(t1, t2);
    host_wait += us(t2, t3);
    host_advance += us(t3, t4);
    // A problem that has just finished may have a flux fill pending (a group without source):
    // the pending operations run now, so that its flux is final when round() returns.
    bool fin = false;
    for (int i : who) fin = fin || ps[i].phase == DONE;
    if (!adds.empty() || (fin && !pre.empty())) {
        // z += sum y V with the coefficients in the staging arena (free again after the
        // synchronisation above)
        arena_off = 0;
        if (!adds.empty()) {
            const double* dc = stage(coefs.data(), coefs.size());
            for (size_t a = 0; a < adds.size(); ++a) adds[a].coef = dc + 64 * a;
            run_vecops(adds);
        }
        if (fin) {
            run_vecops(pre);
            pre.clear();
        }
        flush();
        RT_CHECK(hipStreamSynchronize(st));
    }
    return true;
}

void BatchSolver::run(rt_solve_info* info) {
    hipEvent_t e0, e1;
    RT_CHECK(hipEventCreate(&e0));
    RT_CHECK(hipEventCreate(&e1));
    for (int i = 0; i < (int)ps.size(); ++i)
        if (ps[i].P) start(i);
    {
        const char* e = std::getenv("RADIANT_HIP_PROFILE");
        profiling = e && e[0] == '1';
    }
    timing = Timing{};
    RT_CHECK(hipEventRecord(e0, st));
    pre.clear();
    rounds = 0;
    const int64_t graph0 = graph_updates + graph_instantiations, inst0 = graph_instantiations;
    host_plan = host_submit = host_wait = host_advance = 0;
    while (round()) {
    }
    {
        const char* e = std::getenv("RADIANT_HIP_HOSTPROF");
        if (e && e[0] == '1')
            std::fprintf(stderr, "[batch host] rounds %d plan %.0f us submit %.0f us wait %.0f us advance %.0f us\n",
                         rounds, host_plan, host_submit, host_wait, host_advance);
    }
    RT_CHECK(hipEventRecord(e1, st));
    RT_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    RT_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->ms_device = ms;
        for (auto& s : ps) {
            if (!s.P) continue;
            info->passes += s.P->counters.passes;
            info->passes_lo += s.P->counters.passes_lo;
            info->updates_raw += s.P->counters.updates_raw;
            info->updates_problem += (double)s.P->Ncell * s.P->Nd * s.P->G;
        }
        info->graph_launches = graph_updates + graph_instantiations - graph0;
        info->graph_instantiations = graph_instantiations - inst0;
    }
}

// One step of problem i's state machine after its pass of this round.
void BatchSolver::advance(int i, std::vector<VecItem>& pre, std::vector<VecItem>& adds, std::vector<int>& add_who,
                          std::vector<double>& coefs) {
    PS& s = ps[i];
    Problem* P = s.P;
    const double* slots = h_slots + (size_t)SLOTS * i;
    const long S = (long)P->S_r;
    auto item = [&](int op, double* y, const double* x, double alpha) {
        VecItem it{};
        it.op = op; it.n = S; it.y = y; it.x = x; it.alpha = alpha;
        return it;
    };
    auto finish_group = [&](bool conv) {
        rt_group_stats& g = s.stats[s.g];
        g.niter = s.niter;
        g.resid = s.resid;
        g.rho = std::numeric_limits<double>::quiet_NaN();
        g.converged = conv ? 1 : 0;
        g.passes = (int32_t)(P->counters.passes - s.p0);
        g.passes_lo = (int32_t)(P->counters.passes_lo - s.l0);
        g.refinements = s.refs;
        ++s.g;
        s.phase = (s.g >= P->G) ? DONE : START;
    };
    switch (s.phase) {
        case CPASS: {
            s.bnorm = std::sqrt(slots[S_NRM]);
            if (s.bnorm == 0.0) {  // no source anywhere in the group (sn_one_speed.jl:116-120)
                pre.push_back([&] {
                    VecItem it{};
                    it.op = VOP_FILL; it.n = (long)P->Np * P->Nc; it.y = P->d_phi + (size_t)s.g * P->Np * P->Nc;
                    return it;
                }());
                if (P->csd) {  // and no energy flux leaves the group (sn_one_speed.jl:119)
                    VecItem it{};
                    it.op = VOP_FILL; it.n = (long)P->nE(); it.y = P->d_psiE_out;
                    pre.push_back(it);
                }
                s.resid = 0.0;
                finish_group(true);
                break;
            }
            s.resid = 1.0;
            // (r is first read after the residual pass writes r = c - (I - A) z)
            pre.push_back(item(VOP_FILL, s.z, nullptr, 0.0));
            pre.push_back(item(VOP_SCALE_COPY, s.V[0], s.c, 1.0 / s.bnorm));
            s.beta = s.bnorm;
            start_arnoldi(s);
            break;
        }
        case ARNOLDI: {
            const int j = s.j;
            ++s.niter;
            for (int k = 0; k <= j; ++k) s.H[k + 33 * j] = slots[S_H1 + k] + (reorth ? slots[S_H2 + k] : 0.0);
            const double hjp1 = std::sqrt(std::max(0.0, slots[S_WN]));
            s.H[j + 1 + 33 * j] = hjp1;
            auto h = [&](int a, int b) -> double& { return s.H[a + 33 * b]; };
            for (int k = 0; k < j; ++k) {
                const double t = s.cs[k] * h(k, j) + s.sn[k] * h(k + 1, j);
                h(k + 1, j) = -s.sn[k] * h(k, j) + s.cs[k] * h(k + 1, j);
                h(k, j) = t;
            }
            const double den = std::hypot(h(j, j), h(j + 1, j));
            if (den == 0.0) { s.cs[j] = 1.0; s.sn[j] = 0.0; }
            else { s.cs[j] = h(j, j) / den; s.sn[j] = h(j + 1, j) / den; }
            h(j, j) = s.cs[j] * h(j, j) + s.sn[j] * h(j + 1, j);
            h(j + 1, j) = 0.0;
            s.gv[j + 1] = -s.sn[j] * s.gv[j];
            s.gv[j] = s.cs[j] * s.gv[j];
            const double ri = std::fabs(s.gv[j + 1]) / s.beta;
            const bool stop = ri < s.inner_tol || j + 1 == s.m || hjp1 == 0.0 || s.niter >= P->maxit;
            if (!stop) {
                ++s.j;
                break;
            }
            const int k = j + 1;
            std::vector<double> y(k, 0.0);
            for (int a = k - 1; a >= 0; --a) {
                double v = s.gv[a];
                for (int l = a + 1; l < k; ++l) v -= h(a, l) * y[l];
                y[a] = h(a, a) != 0.0 ? v / h(a, a) : 0.0;
            }
            // coefficients: 64 per add item, staged by the caller (coef set there)
            coefs.resize(64 * (adds.size() + 1), 0.0);
            for (int a = 0; a < k; ++a) coefs[64 * adds.size() + a] = y[a];
            VecItem it{};
            it.op = VOP_ADDMULTI; it.n = S; it.y = s.z; it.V = (const double* const*)P->d_Vptr; it.k = k;
            adds.push_back(it);
            add_who.push_back(i);
            s.phase = RESID;
            ++s.refs;
            break;
        }
        case RESID: {
            const double rn = std::sqrt(slots[S_NRM]);
            s.resid = rn / s.bnorm;
            if (s.resid < P->tol || s.refs >= 12 || s.niter >= P->maxit) {
                s.conv = s.resid < P->tol;
                s.phase = RECON;
                // psi (angular flux of A z from this residual pass) += psiC (of c = T(0))
                VecItem it{};
                it.op = VOP_ADD; it.n = (long)P->Nd * P->Nc; it.y = (double*)P->d_psi; it.x = (double*)P->d_psi;
                it.x2 = s.psiC;
                pre.push_back(it);
                break;
            }
            s.beta = rn;
            pre.push_back(item(VOP_SCALE_COPY, s.V[0], s.r, 1.0 / rn));
            start_arnoldi(s);
            break;
        }
        case RECON:
            finish_group(s.conv);
            break;
        default:
            break;
    }
}

void BatchSolver::start_arnoldi(PS& s) {
    s.phase = ARNOLDI;
    s.j = 0;
    std::fill(std::begin(s.H), std::end(s.H), 0.0);
    std::fill(std::begin(s.cs), std::end(s.cs), 0.0);
    std::fill(std::begin(s.sn), std::end(s.sn), 0.0);
    std::fill(std::begin(s.gv), std::end(s.gv), 0.0);
    s.gv[0] = s.beta;
}

}  // namespace rt
