// batch.hpp -- lockstep solver for many independent problems in one set of launches.
#pragma once
#include <functional>
#include <vector>

#include "engine.hpp"

namespace rt {

class BatchSolver {
public:
    // A fixed batch: slot i holds probs[i].
    BatchSolver(const std::vector<Problem*>& probs, hipStream_t stream);
    // Empty slots for continuous batching (add / remove between rounds).
    BatchSolver(int capacity, hipStream_t stream);
    ~BatchSolver();
    // Solves every problem in the slots from group 0 to the end.
    void run(rt_solve_info* info);
    // Continuous batching.  add() puts a problem into a free slot; it starts from group 0 at the
    // next round.  round() advances every unfinished problem by one pass (plus its Krylov work)
    // and returns false when none is left; the flux of a problem that is done() when round()
    // returns is final.  A finished problem keeps its slot until remove(), which may also take an
    // unfinished problem out (cancellation).  Problems in the slots are used on the solver's
    // stream and must outlive their slot.
    int add(Problem* P);
    void remove(int slot);
    bool done(int slot) const { return ps[slot].phase == DONE; }
    bool round();
    int capacity() const { return (int)ps.size(); }
    const std::vector<rt_group_stats>& stats(int i) const { return ps[i].stats; }
    int rounds = 0;
    // Per-kernel device time of the last run() when RADIANT_HIP_PROFILE=1 (HIP events on the stream
    // each launch runs on), [precision][mgemm, sweep, dgemm, -], with the work the launches did:
    // cell*direction updates, contraction flops and algorithmic HBM bytes.
    struct Timing {
        double ms[3][4] = {}, upd[3][4] = {}, flop[3][4] = {}, bytes[3][4] = {};
        int64_t n[3][4] = {};
    };
    Timing timing;

private:
    enum Phase { START, CPASS, ARNOLDI, RESID, RECON, DONE };
    struct PS {
        Problem* P = nullptr;
        int g = 0, phase = START, j = 0, m = 30, lo = P32, refs = 0;
        int64_t niter = 0, p0 = 0, l0 = 0;
        double bnorm = 0, beta = 0, resid = 1, inner_tol = 1e-5;
        bool conv = false;
        double H[33 * 32] = {}, cs[32] = {}, sn[32] = {}, gv[33] = {};
        std::vector<double*> V;
        double *w = nullptr, *c = nullptr, *z = nullptr, *r = nullptr, *t = nullptr;
        double* psiC = nullptr;  // angular flux of the group's c = T(0) pass (FP64), kept for the reconstruction
        std::vector<rt_group_stats> stats;
    };
    std::vector<PS> ps;
    hipStream_t st;
    std::vector<VecItem> pre;  // vector operations to run before the next round's passes
    void init_device(int capacity);
    void ensure_task_capacity();  // grows the task tables to the problems in the slots
    void start(int slot);         // resets a slot's problem to group 0
    void drain_pre();             // runs the vector operations recorded for the next round
    bool reorth = false;        // CGS2 instead of CGS
    bool fork_streams = false;  // FP64 passes on st64 beside the low-precision ones
    std::vector<void*> mem;
    PassDesc* d_pd = nullptr;
    // task tables per precision group (the FP64 group runs on st64, concurrently with the others)
    GemmTask *d_mg[3] = {}, *d_dg[3] = {};
    SweepTask* d_sw[3] = {};
    hipStream_t st64 = nullptr;
    unsigned* d_wl = nullptr;  // work-list counters of the line sweeps: [0] on st, [64] on st64
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    size_t cap_gemm = 0, cap_sweep = 0;
    void* task_mem[9] = {};
    double *d_slots = nullptr, *h_slots = nullptr, *d_partial = nullptr;
    // Staging arena: host (pinned) and device copies at equal offsets.  A round stages its
    // inputs and records its launches; flush() uploads the arena with one copy, then launches.
    unsigned char *h_arena = nullptr, *d_arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
    std::vector<std::function<void()>> plan;
    template <class T> T* stage(const T* data, size_t count);
    void flush();
    // RADIANT_HIP_GRAPHS=1: rounds run as captured HIP graphs, executable graphs cached by shape
    bool use_graphs = false;
    struct GraphEntry {
        std::vector<uintptr_t> sig;
        hipGraphExec_t exec;
    };
    static constexpr size_t GRAPH_CACHE = 32;
    std::vector<GraphEntry> gcache;
    void flush_graph();

public:
    int64_t graph_updates = 0, graph_instantiations = 0;  // rounds run by updating a cached graph / new ones

private:

    void launch_passes(const std::vector<int>& who, const std::vector<PassDesc>& pds, const std::vector<int>& prec,
                       const std::vector<int>& full);
    void run_vecops(const std::vector<VecItem>& items);
    void run_gs(const std::vector<GsItem>& items, bool update, bool norm);
    void advance(int i, std::vector<VecItem>& pre, std::vector<VecItem>& adds, std::vector<int>& add_who,
                 std::vector<double>& coefs);
    void start_arnoldi(PS& s);

    double host_plan = 0, host_submit = 0, host_wait = 0, host_advance = 0;  // RADIANT_HIP_HOSTPROF
    bool profiling = false;
    struct EvRec { hipEvent_t a, b; int kernel, prec; double upd, flop, bytes; };
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    std::vector<EvRec> ev_used;
    // wraps a launch closure with start/stop events when profiling
    std::function<void()> timed(int kernel, int prec, double upd, double flop, double bytes, hipStream_t s,
                                std::function<void()> fn);
    void collect_times();  // after the end-of-round synchronisation
};

}  // namespace rt
