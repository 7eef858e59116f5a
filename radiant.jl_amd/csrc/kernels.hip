// kernels.hip -- gfx950 kernels of the discrete-ordinates transport pass and its Krylov solver.
//
// One transport pass (src/particle_transport/sn_inner_pass.jl:212-332) is three kernels:
//   mgemm  : psi[n][c] = sum_a M[n][a] * Ql[a][c], with Ql = Qlout + Σs_l Φ computed while the
//            MFMA operand is loaded (scattering_source.jl:25-30 fused into the contraction);
//   sweep  : the diamond-difference wavefront of every direction, in place on psi
//            (sn_sweep_{1,2,3}D.jl with the order-1 DD cell solve of sn_{1,2,3}D_bte.jl);
//   dgemm  : Φ[r][c] = sum_n D[r][n] * psi[n][c]  (the D accumulation of the sweeps), with the
//            I - A epilogue of the GMRES matvec (sn_one_speed.jl:147) fused in;
// plus the half-range boundary moments of the exit faces (bndmom) for the faithful state.
//
// Layouts: cells c = ix + Nx*(iy + Ny*iz); moment/direction-major arrays [row][c] so that every
// contraction streams 16-cell rows (128 B in FP64) per MFMA operand.
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hip_bf16.h>
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace rt {

#define RT_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                       " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// Per-thread pair of events that the next timed launch records at its own dispatch start and end
// (hipExtLaunchKernel): the kernel's execution interval -- what a kernel trace reports -- even while
// kernels of other streams occupy the device (events recorded on the stream around the launch
// would also count the wait for those).
static thread_local hipEvent_t tl_ev_start = nullptr, tl_ev_stop = nullptr;
void set_launch_events(hipEvent_t start, hipEvent_t stop) {
    tl_ev_start = start;
    tl_ev_stop = stop;
}
void take_launch_events(hipEvent_t* start, hipEvent_t* stop) {
    *start = tl_ev_start;
    *stop = tl_ev_stop;
    tl_ev_start = tl_ev_stop = nullptr;
}
template <typename... KArgs, typename... Args>
static void launch_timed(void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shmem, hipStream_t st,
                         Args... args) {
    if (tl_ev_start || tl_ev_stop) {  // either may be null (a timed pair of launches)
        hipEvent_t a = tl_ev_start, b = tl_ev_stop;
        tl_ev_start = tl_ev_stop = nullptr;
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, a, b, 0u, static_cast<KArgs>(args)...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, static_cast<KArgs>(args)...);
    }
}

// ------------------------------------------------------------------------------------------
// Launch-shape autotuner.  A sweep or contraction launch has variants that compute bit-identical
// results and differ only in their schedule: the rows in flight of the line kernel (1D/2D), for
// the 3D scan kernel the pencils per item with their prefetch depth, the waves per workgroup and
// the split of each direction over 1, 2 or 4 workgroups,
// and the pipeline depth (k steps in flight) of the M and D contractions.  In
// mode 1 the first launches of each key (launch class, kernel, cells per lane, default waves,
// inflow flag, log2 of the task count) cycle through the candidates with dispatch start/stop
// events; once every candidate has TUNE_REPS completed timings, the one with the fastest timing
// serves the rest of the process (the fastest, as other streams' work can only slow a launch).  Mode 0 keeps the first (default) candidate; mode 2 moves to
// the next candidate on every launch (for tests of the variants' identity).
// ------------------------------------------------------------------------------------------
namespace {
constexpr int TUNE_REPS = 5;
struct TuneKey {
    int kind, kernel, cpl, waves, bin, lg;
    bool operator<(const TuneKey& o) const {
        return std::tie(kind, kernel, cpl, waves, bin, lg) < std::tie(o.kind, o.kernel, o.cpl, o.waves, o.bin, o.lg);
    }
};
struct TunePending { int cand; hipEvent_t a, b; };
// A launch chosen for timing: its events join the key's pending list only once the launch that
// records them has been issued (tune_commit), so no other thread can query them before.
struct TuneTicket { TuneKey key; int cand; hipEvent_t a = nullptr, b = nullptr; };
struct TuneState {
    int ncand = 0, best = -1, next = 0;
    std::vector<std::vector<float>> ms;
    std::vector<TunePending> pending;
    std::vector<int> issuing;  // tickets handed out and not yet committed, per candidate
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneState> g_tune;
std::atomic<int> g_tune_mode{-1};
// Launches issued from this host thread may be timed (false on threads whose launches share the
// device with other concurrent streams: their timings would measure the neighbours' work too).
thread_local bool tl_tune_measure = true;

int tune_mode() {
    int m = g_tune_mode.load();
    if (m < 0) {
        const char* e = std::getenv("RADIANT_HIP_AUTOTUNE");
        m = (e && (e[0] == '0' || e[0] == '2')) ? e[0] - '0' : 1;
        g_tune_mode.store(m);
    }
    return m;
}

int ilog2(int n) {
    int l = 0;
    while ((1 << (l + 1)) <= n) ++l;
    return l;
}

// the fastest of a candidate's timings: launches of concurrent streams only ever slow one down
float best_of(const std::vector<float>& v) { return v.empty() ? 0.f : *std::min_element(v.begin(), v.end()); }

// Completed timings of a key's issued launches move into its samples (lock held).  The queries'
// status codes are cleared from the thread's last-error state.
void tune_harvest(TuneState& s) {
    for (size_t i = 0; i < s.pending.size();) {
        TunePending& p = s.pending[i];
        const hipError_t q = hipEventQuery(p.b);
        if (q == hipErrorNotReady) { ++i; continue; }
        float ms = 0.f;
        if (q == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) s.ms[p.cand].push_back(ms);
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
        s.pending.erase(s.pending.begin() + i);
    }
    (void)hipGetLastError();
}

// The candidate for this launch; when the launch is to be timed, t.a and t.b receive fresh events
// (to be handed to launch_timed through set_launch_events, and to tune_commit after the launch).
int tune_choose(const TuneKey& key, int ncand, TuneTicket& t) {
    t.key = key;
    t.a = t.b = nullptr;
    const int mode = tune_mode();
    if (ncand <= 1 || mode == 0) return t.cand = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneState& s = g_tune[key];
    if (s.ncand == 0) { s.ncand = ncand; s.ms.assign(ncand, {}); s.issuing.assign(ncand, 0); }
    if (mode == 2) { const int c = s.next; s.next = (s.next + 1) % ncand; return t.cand = c; }
    if (!s.pending.empty()) tune_harvest(s);
    if (s.best >= 0) return t.cand = s.best;
    if (tl_ev_start || tl_ev_stop || !tl_tune_measure) return t.cand = 0;  // timed by the caller, or not timed alone
    bool done = true;
    for (auto& v : s.ms) done = done && v.size() >= (size_t)TUNE_REPS;
    if (done) {
        int best = 0;
        for (int c = 1; c < ncand; ++c)
            if (best_of(s.ms[c]) < best_of(s.ms[best])) best = c;
        s.best = best;
        return t.cand = best;
    }
    // the next candidate still short of samples (counting those in flight)
    for (int k = 0; k < ncand; ++k) {
        const int c = (s.next + k) % ncand;
        size_t inflight = 0;
        for (auto& p : s.pending) inflight += p.cand == c;
        if (s.ms[c].size() + inflight + s.issuing[c] < (size_t)TUNE_REPS) {
            s.next = (c + 1) % ncand;
            t.cand = c;
            if (hipEventCreate(&t.a) != hipSuccess) { t.a = nullptr; (void)hipGetLastError(); return c; }
            if (hipEventCreate(&t.b) != hipSuccess) {
                (void)hipEventDestroy(t.a);
                t.a = t.b = nullptr;
                (void)hipGetLastError();
                return c;
            }
            s.issuing[c] += 1;
            return c;
        }
    }
    return t.cand = 0;  // every candidate has its timings in flight
}

// After the timed launch of a ticket has been issued: its events join the pending list.
void tune_commit(const TuneTicket& t) {
    if (!t.a) return;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneState& s = g_tune[t.key];
    if (s.ncand == 0) {  // the key was forgotten meanwhile (set_autotune_mode): drop the timing
        g_tune.erase(t.key);
        (void)hipEventSynchronize(t.b);
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
        (void)hipGetLastError();
        return;
    }
    s.issuing[t.cand] -= 1;
    s.pending.push_back({t.cand, t.a, t.b});
}
}  // namespace

void set_autotune_measuring(bool on) { tl_tune_measure = on; }
bool autotune_measuring() { return tl_tune_measure; }

void set_autotune_mode(int mode) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_mode.store(mode);
    for (auto& kv : g_tune)
        for (auto& p : kv.second.pending) {
            (void)hipEventSynchronize(p.b);
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    g_tune.clear();
}

std::string autotune_report() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    static const char* kinds[] = {"fp64", "fp32", "bf16", "mixed"};
    std::string r;
    for (auto& kv : g_tune) {
        const TuneKey& k = kv.first;
        const TuneState& st = kv.second;
        char line[256];
        static const char* knames[] = {"line", "scan3d", "mgemm", "dgemm"};
        std::snprintf(line, sizeof line, "%s %s cpl=%d waves=%d inflow=%d splits=%d tasks~2^%d: ", kinds[k.kind & 3],
                      knames[k.kernel & 3], k.cpl, k.waves, k.bin & 1, k.bin >> 1, k.lg);
        r += line;
        if (st.best < 0) r += "tuning";
        else {
            std::snprintf(line, sizeof line, "variant %d of %d;", st.best, st.ncand);
            r += line;
            for (int c = 0; c < st.ncand; ++c) {
                std::snprintf(line, sizeof line, " %.1f", 1e3 * best_of(st.ms[c]));
                r += line;
            }
            r += " us";
        }
        r += "\n";
    }
    return r;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<double> { typedef v4d type; };
template <> struct Acc<float> { typedef v4f type; };

// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.  Operand layout (wave64):
//   A[16x4]: lane l holds A[l%16][l/16];  B[4x16]: lane l holds B[l/16][l%16];
//   C[16x16]: lane l, element r holds C[row][l%16] with
//     f32: row = 4*(l/16) + r      f64: row = l/16 + 4*r
// (measured on gfx950 with rt_test_contraction against a host GEMM).
__device__ inline v4d mfma4(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline v4f mfma4(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
template <typename T> __device__ inline int acc_row(int lane, int r);
template <> __device__ inline int acc_row<double>(int lane, int r) { return (lane >> 4) + 4 * r; }
template <> __device__ inline int acc_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

// Global-memory accessors.  Pointers read from a PassDesc are generic, and generic accesses
// compile to FLAT instructions, which count against lgkmcnt: every LDS or cross-lane wait would
// then also wait for the outstanding HBM loads.  Casting to the global address space gives
// global_load/global_store, tracked by vmcnt only.
#define RT_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ inline T gld(const T* p) { return *(const RT_GLOBAL T*)p; }
template <typename T> __device__ inline void gst(T* p, T v) { *(RT_GLOBAL T*)p = v; }

// Cell-solve division: IEEE division in FP64 (the parity path); in FP32 the hardware reciprocal
// (1 ulp) -- the low-precision passes only feed an FP64-refined iteration.
__device__ inline double dd_div(double q, double s) { return q / s; }
__device__ inline float dd_div(float q, float s) { return q * __builtin_amdgcn_rcpf(s); }

// bf16 round-to-nearest-even of a float, as raw bits.
__device__ inline short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    unsigned r = ((u >> 16) & 1u) + 0x7fffu;
    return (short)((u + r) >> 16);
}

// ------------------------------------------------------------------------------------------
// Contractions on the matrix cores.
//
//   mgemm: psi[n][c]  = sum_a MT[a][n] * Ql[a][c],  Ql = Qlout + Σs_l Φ computed on operand load
//          (scattering_source.jl:25-30 fused into the moment->direction contraction);
//   dgemm: out[r][c]  = sum_n DT[n][r] * psi[n][c]  (the D accumulation of the sweeps), with the
//          I - A epilogue of the GMRES matvec (sn_one_speed.jl:147): out = zin - D psi.
// The operators are stored k-major (MT[K][M]) so that each MFMA operand load is a contiguous
// 16-element row segment.  A wave computes RT x CT tiles of 16x16 (RT*16 rows, CT*16 cells):
// every A fragment is reused CT times and every B fragment RT times from registers.
//
// The k loop is software-pipelined over NS register stages: the operands of steps s+1 .. s+NS-1
// are in flight while the MFMAs of step s run.  Every operand load is unconditional (padding rows,
// cells and k read clamped in-range addresses, absent inputs read a zero line), so the wait
// counts stay exact and the loads of the next stage are never waited for early.
// ------------------------------------------------------------------------------------------
__device__ __align__(256) double g_zero[64];  // zero-initialised, never written

template <typename T, int RT, int CT, typename V>
__device__ inline void epilogue_dgemm(const PassDesc& d, const V (&acc)[RT][CT], int c0, int r0, int li, int lane,
                                      int M, int Nc) {
    const bool hasz = d.zin != nullptr;
    const double* __restrict__ zb = hasz ? d.zin : g_zero;
    const size_t zrs = hasz ? (size_t)Nc : 0;
    const double zs = hasz ? 1.0 : 0.0, sg = hasz ? -1.0 : 1.0;
    // one row tile at a time (the zin values of a whole wave tile would hold 2*RT*CT*4 VGPRs
    // and halve the waves a SIMD can hold); the loads of tile t + 1 are issued before the
    // stores of tile t
    double z[2][CT][4];
    auto zload = [&](int t, double (&zt)[CT][4]) {
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = c0 + 16 * u + li, row = r0 + 16 * t + acc_row<T>(lane, r);
                const int cc = col < Nc ? col : Nc - 1, rc = row < M ? row : M - 1;
                zt[u][r] = gld(zb + (size_t)rc * zrs + (hasz ? cc : li));
            }
    };
    zload(0, z[0]);
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        if (t + 1 < RT) zload(t + 1, z[(t + 1) & 1]);
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = c0 + 16 * u + li, row = r0 + 16 * t + acc_row<T>(lane, r);
                const double v = zs * z[t & 1][u][r] + sg * (double)acc[t][u][r];
                if (col < Nc && row < M) gst(d.out + (size_t)row * Nc + col, v);
            }
    }
}

template <typename T, int RT, int CT, bool MG, int NS, bool ADD = false>
__device__ __forceinline__ void contract_body(const PassDesc& d, const GemmTask tk, int* rows_s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Nc = d.Nc;
    const int K = MG ? d.Nact : d.Nd;
    const int M = MG ? d.Nd : d.Nrows;
    const T* __restrict__ At = (const T*)(MG ? d.Mact : d.Drows);
    const T* __restrict__ psi_in = (const T*)d.psi;
    const T* __restrict__ psi_ad = ADD ? (const T*)d.psi_add : nullptr;
    const double* __restrict__ qlout = d.qlout;
    const double* __restrict__ phi_in = d.phi_in;
    const double* __restrict__ sigw = d.sigw;
    const bool hasq = qlout != nullptr, hasp = phi_in != nullptr;
    const int c0 = tk.tile * GEMM_CELLS + wave * (CT * 16);
    const int r0 = tk.rgroup * GEMM_ROWS;
    const int kk = lane >> 4, li = lane & 15;
    int colc[CT], rowc[RT];
    bool colok[CT], rowok[RT];
#pragma unroll
    for (int u = 0; u < CT; ++u) {
        const int col = c0 + 16 * u + li;
        colok[u] = col < Nc;
        colc[u] = colok[u] ? col : Nc - 1;
    }
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = r0 + 16 * t + li;
        rowok[t] = row < M;
        rowc[t] = rowok[t] ? row : M - 1;
    }
    // Absent inputs are selected here, once, as address arithmetic: a missing Qlout or Φ is
    // read from the zero line with stride 0 (selects inside the k loop become branches).
    int qcol[CT], pcol[CT], wrow[CT];
    const double *qb = g_zero, *pb = g_zero, *wb = g_zero;
    size_t qks = 0, pks = 0;
    int wks = 0;
    if (MG) {
        const int* __restrict__ rows = d.phi_rows;
        for (int i = threadIdx.x; i < K; i += blockDim.x) rows_s[i] = rows ? rows[i] : i;
        if (hasq) { qb = qlout; qks = (size_t)Nc; }
        if (hasp) { pb = phi_in; pks = (size_t)Nc; wb = sigw; wks = 1; }
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            const int mc = gld(d.mat + colc[u]);
            qcol[u] = hasq ? colc[u] : li;
            pcol[u] = hasp ? colc[u] : li;
            wrow[u] = hasp ? mc * K : li;
        }
        __syncthreads();
    }
    typename Acc<T>::type acc[RT][CT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u) acc[t][u] = (typename Acc<T>::type)(0);
    // A stage holds the raw loads of one k step; padding (k >= K, rows >= M, cells >= Nc) reads
    // clamped in-range data and is zeroed by a 0/1 factor when the MFMA operands are formed (a
    // select would let the compiler move a load into a branch; forming the operands at load
    // time would wait for the loads right after issuing them).
    struct Stage { T a[RT]; double q[CT], w[CT], p[CT]; T b[CT], b2[ADD ? CT : 1]; };
    auto load = [&](int s, Stage& g) {
        const int k = 4 * s + kk;
        const int kc = k < K ? k : K - 1;
#pragma unroll
        for (int t = 0; t < RT; ++t) g.a[t] = gld(At + (size_t)kc * M + rowc[t]);
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            if (MG) {
                g.q[u] = gld(qb + kc * qks + qcol[u]);
                g.w[u] = gld(wb + kc * wks + wrow[u]);
                g.p[u] = gld(pb + (size_t)rows_s[kc] * pks + pcol[u]);
            } else {
                g.b[u] = gld(psi_in + (size_t)kc * Nc + colc[u]);
                if constexpr (ADD) g.b2[u] = gld(psi_ad + (size_t)kc * Nc + colc[u]);
            }
        }
    };
    auto mma = [&](int s, const Stage& g) {
        const bool kok = 4 * s + kk < K;
        T a[RT], b[CT];
#pragma unroll
        for (int t = 0; t < RT; ++t) a[t] = g.a[t] * ((kok && rowok[t]) ? (T)1 : (T)0);
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            if (MG) b[u] = (T)((g.q[u] + g.w[u] * g.p[u]) * ((kok && colok[u]) ? 1.0 : 0.0));
            else if constexpr (ADD) b[u] = (g.b[u] + g.b2[u]) * ((kok && colok[u]) ? (T)1 : (T)0);
            else b[u] = g.b[u] * ((kok && colok[u]) ? (T)1 : (T)0);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t)
            if (r0 + 16 * t < M)  // wave-uniform
#pragma unroll
                for (int u = 0; u < CT; ++u) acc[t][u] = mfma4(a[t], b[u], acc[t][u]);
    };
    // NS stages: step s + NS - 1 is loaded into the stage that step s - 1 has just freed.
    const int nsteps = (K + 3) >> 2;
    Stage g[NS];
#pragma unroll
    for (int i = 0; i < NS - 1; ++i) {
        __builtin_amdgcn_sched_barrier(0);
        load(i, g[i]);
    }
    for (int s = 0; s < nsteps; s += NS) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            __builtin_amdgcn_sched_barrier(0);
            load(s + i + NS - 1, g[(i + NS - 1) % NS]);
            __builtin_amdgcn_sched_barrier(0);
            if (s + i < nsteps) mma(s + i, g[i]);  // wave-uniform
        }
    }
    if (MG) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int u = 0; u < CT; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = c0 + 16 * u + li, row = r0 + 16 * t + acc_row<T>(lane, r);
                    if (col < Nc && row < M) gst((T*)d.psi + (size_t)row * Nc + col, (T)acc[t][u][r]);
                }
    } else {
        // out = zin - D psi (I - A matvec) or D psi: every zin value is loaded before the first
        // store, unconditionally, and combined as zs*z + sg*v (a select would branch the loads).
        epilogue_dgemm<T, RT, CT>(d, acc, c0, r0, li, lane, M, Nc);
    }
}

template <typename T, int RT, int CT, bool MG, int NS, bool ADD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && !MG) ? 3 : 2))) void contract_kernel(const PassDesc* __restrict__ pds, const GemmTask* __restrict__ tasks) {
    __shared__ int rows_s[MG ? CONTRACT_KMAX : 1];
    const GemmTask tk = tasks[blockIdx.x];
    contract_body<T, RT, CT, MG, NS, ADD>(pds[tk.prob], tk, rows_s);
}

// FP64 and FP32 passes of different problems in one launch (PassDesc::prec per workgroup).
template <int RT, int CT, bool MG, int NS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void contract_mixed_kernel(const PassDesc* __restrict__ pds,
                                                             const GemmTask* __restrict__ tasks) {
    __shared__ int rows_s[MG ? CONTRACT_KMAX : 1];
    const GemmTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    // the FP64 body keeps at most 2 stages in flight: with 4, its registers (FP64 operands and
    // accumulators) would leave room for one wave per SIMD for the FP32 passes of the launch too
    if (d.prec == P64) contract_body<double, RT, CT, MG, (NS > 2 ? 2 : NS)>(d, tk, rows_s);
    else contract_body<float, RT, CT, MG, NS>(d, tk, rows_s);
}

// BF16 operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16): A[16x16] lane l holds
// A[l%16][4*(l/16)+j], B[16x16] lane l holds B[4*(l/16)+j][l%16], j = 0..3; C as the f32 case.
// Pipelined as contract_kernel; the operands are staged in FP32 and rounded to BF16 when used.
template <int RT, int CT, bool MG>
__global__ __launch_bounds__(256) void contract_bf16_kernel(const PassDesc* __restrict__ pds, const GemmTask* __restrict__ tasks) {
    __shared__ int rows_s[MG ? CONTRACT_KMAX : 1];
    const GemmTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Nc = d.Nc;
    const int K = MG ? d.Nact : d.Nd;
    const int M = MG ? d.Nd : d.Nrows;
    const float* __restrict__ At = (const float*)(MG ? d.Mact : d.Drows);
    const float* __restrict__ psi_in = (const float*)d.psi;
    const double* __restrict__ qlout = d.qlout;
    const double* __restrict__ phi_in = d.phi_in;
    const double* __restrict__ sigw = d.sigw;
    const bool hasq = qlout != nullptr, hasp = phi_in != nullptr;
    const int c0 = tk.tile * GEMM_CELLS + wave * (CT * 16);
    const int r0 = tk.rgroup * GEMM_ROWS;
    const int kq = 4 * (lane >> 4), li = lane & 15;
    int colc[CT], rowc[RT];
    bool colok[CT], rowok[RT];
#pragma unroll
    for (int u = 0; u < CT; ++u) {
        const int col = c0 + 16 * u + li;
        colok[u] = col < Nc;
        colc[u] = colok[u] ? col : Nc - 1;
    }
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = r0 + 16 * t + li;
        rowok[t] = row < M;
        rowc[t] = rowok[t] ? row : M - 1;
    }
    // Absent inputs are selected here, once, as address arithmetic: a missing Qlout or Φ is
    // read from the zero line with stride 0 (selects inside the k loop become branches).
    int qcol[CT], pcol[CT], wrow[CT];
    const double *qb = g_zero, *pb = g_zero, *wb = g_zero;
    size_t qks = 0, pks = 0;
    int wks = 0;
    if (MG) {
        const int* __restrict__ rows = d.phi_rows;
        for (int i = threadIdx.x; i < K; i += blockDim.x) rows_s[i] = rows ? rows[i] : i;
        if (hasq) { qb = qlout; qks = (size_t)Nc; }
        if (hasp) { pb = phi_in; pks = (size_t)Nc; wb = sigw; wks = 1; }
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            const int mc = gld(d.mat + colc[u]);
            qcol[u] = hasq ? colc[u] : li;
            pcol[u] = hasp ? colc[u] : li;
            wrow[u] = hasp ? mc * K : li;
        }
        __syncthreads();
    }
    v4f acc[RT][CT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u) acc[t][u] = (v4f)(0.0f);
    struct Stage { float a[RT][4]; double q[CT][4], w[CT][4], p[CT][4]; float b[CT][4]; };
    auto load = [&](int s, Stage& g) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * s + kq + j;
            const int kc = k < K ? k : K - 1;
#pragma unroll
            for (int t = 0; t < RT; ++t) g.a[t][j] = gld(At + (size_t)kc * M + rowc[t]);
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                if (MG) {
                    g.q[u][j] = gld(qb + kc * qks + qcol[u]);
                    g.w[u][j] = gld(wb + kc * wks + wrow[u]);
                    g.p[u][j] = gld(pb + (size_t)rows_s[kc] * pks + pcol[u]);
                } else {
                    g.b[u][j] = gld(psi_in + (size_t)kc * Nc + colc[u]);
                }
            }
        }
    };
    auto mma = [&](int s, const Stage& g) {
        v4s ab[RT], bb[CT];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool kok = 16 * s + kq + j < K;
#pragma unroll
            for (int t = 0; t < RT; ++t) ab[t][j] = f2bf(g.a[t][j] * ((kok && rowok[t]) ? 1.0f : 0.0f));
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                const float v = MG ? (float)((g.q[u][j] + g.w[u][j] * g.p[u][j]) * ((kok && colok[u]) ? 1.0 : 0.0))
                                   : g.b[u][j] * ((kok && colok[u]) ? 1.0f : 0.0f);
                bb[u][j] = f2bf(v);
            }
        }
#pragma unroll
        for (int t = 0; t < RT; ++t)
            if (r0 + 16 * t < M)
#pragma unroll
                for (int u = 0; u < CT; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab[t], bb[u], acc[t][u], 0, 0, 0);
    };
    const int nsteps = (K + 15) >> 4;
    Stage g0, g1;
    load(0, g0);
    for (int s = 0; s < nsteps; s += 2) {
        __builtin_amdgcn_sched_barrier(0);
        load(s + 1, g1);
        __builtin_amdgcn_sched_barrier(0);
        mma(s, g0);
        __builtin_amdgcn_sched_barrier(0);
        load(s + 2, g0);
        __builtin_amdgcn_sched_barrier(0);
        if (s + 1 < nsteps) mma(s + 1, g1);  // wave-uniform
    }
    if (MG) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int u = 0; u < CT; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = c0 + 16 * u + li, row = r0 + 16 * t + acc_row<float>(lane, r);
                    if (col < Nc && row < M) gst((float*)d.psi + (size_t)row * Nc + col, acc[t][u][r]);
                }
    } else {
        epilogue_dgemm<float, RT, CT>(d, acc, c0, r0, li, lane, M, Nc);
    }
}

// ------------------------------------------------------------------------------------------
// Diamond-difference wavefront sweep of one direction per workgroup.
//
// Every thread owns PPT pencils (lines of cells along x, indexed in the direction's own sweep
// frame by (jy, jz)); at step s a pencil processes jx = s - jy - jz.  The incoming x-face flux
// stays in a register; the y/z face fluxes come from the neighbouring pencils through LDS,
// double-buffered by step parity (NBUF = 2) or read/written in two barrier phases (NBUF = 1).
// The angular source of the next PF steps is prefetched into a register ring.
// Cell solve (order-1 DD, sn_3D_bte.jl:42-153 with O = 1, C = 1, w = [-1, 2]):
//   psi = (Q + 2hx fx + 2hy fy + 2hz fz) / (St + 2hx + 2hy + 2hz),  f_out = 2 psi - f_in.
// ------------------------------------------------------------------------------------------
template <typename T, int PPT, int PF, int NBUF>
__global__ __launch_bounds__(1024) void sweep_kernel(const PassDesc* __restrict__ pds, const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    const int n = tk.dir;
    const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz, Nc = d.Nc, P = Ny * Nz;
    const int tid = threadIdx.x, bdim = blockDim.x;
    T* hx2 = (T*)smem;
    T* hy2 = hx2 + Nx;
    T* hz2 = hy2 + Ny;
    T* st = hz2 + Nz;
    T* Fy = st + d.Nmat;
    T* Fz = Fy + NBUF * P;
    const double mu = d.mu[n], eta = d.eta[n], xi = d.xi[n];
    const bool tang = (d.ndims == 1) && (fabs(mu) <= 1e-10);  // sn_1D_bte.jl:154-155
    const double amu = tang ? 0.0 : fabs(mu);
    for (int i = tid; i < Nx; i += bdim) hx2[i] = (T)(2.0 * (amu / d.dx[i]));
    for (int i = tid; i < Ny; i += bdim) hy2[i] = (T)(d.ndims >= 2 ? 2.0 * (fabs(eta) / d.dy[i]) : 0.0);
    for (int i = tid; i < Nz; i += bdim) hz2[i] = (T)(d.ndims >= 3 ? 2.0 * (fabs(xi) / d.dz[i]) : 0.0);
    for (int i = tid; i < d.Nmat; i += bdim) st[i] = (T)d.sigt[i];
    __syncthreads();
    const bool fxs = mu >= 0, fys = eta >= 0, fzs = xi >= 0;
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* __restrict__ bx = (T*)d.bx + (size_t)n * P;
    T* __restrict__ by = (T*)d.by + (size_t)n * Nx * Nz;
    T* __restrict__ bz = (T*)d.bz + (size_t)n * Nx * Ny;
    const int* __restrict__ mat = d.mat;

    int jy[PPT], jz[PPT], iy[PPT], iz[PPT], roff[PPT];
    bool valid[PPT];
    T fx[PPT];
    T ring[PPT][PF];
    int mring[PPT][PF];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int q = tid + k * bdim;
        valid[k] = q < P;
        jy[k] = valid[k] ? q % Ny : 0;
        jz[k] = valid[k] ? q / Ny : 0;
        iy[k] = fys ? jy[k] : Ny - 1 - jy[k];
        iz[k] = fzs ? jz[k] : Nz - 1 - jz[k];
        roff[k] = Nx * (iy[k] + Ny * iz[k]);
        fx[k] = (T)0;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int jx = i - jy[k] - jz[k];
            ring[k][i] = (T)0;
            mring[k][i] = 0;
            if (valid[k] && jx >= 0 && jx < Nx) {
                const int ix = fxs ? jx : Nx - 1 - jx;
                ring[k][i] = gld(psi + roff[k] + ix);
                mring[k][i] = gld(mat + roff[k] + ix);
            }
        }
    }
    const int nsteps = Nx + Ny + Nz - 2;
    for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int s = s0 + i;
            if (s < nsteps) {
                const int rb = (NBUF == 2) ? ((s + 1) & 1) : 0;
                const int wb = (NBUF == 2) ? (s & 1) : 0;
                T ofy[PPT], ofz[PPT];
                bool act[PPT];
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const int q = tid + k * bdim;
                    const int jx = s - jy[k] - jz[k];
                    act[k] = valid[k] && jx >= 0 && jx < Nx;
                    const T v = ring[k][i];
                    const int m = mring[k][i];
                    const int jx2 = jx + PF;
                    if (valid[k] && jx2 >= 0 && jx2 < Nx) {
                        const int ix2 = fxs ? jx2 : Nx - 1 - jx2;
                        ring[k][i] = gld(psi + roff[k] + ix2);
                        mring[k][i] = gld(mat + roff[k] + ix2);
                    }
                    ofy[k] = (T)0;
                    ofz[k] = (T)0;
                    if (act[k]) {
                        const T fy = jy[k] > 0 ? Fy[rb * P + q - 1] : (T)0;
                        const T fz = jz[k] > 0 ? Fz[rb * P + q - Ny] : (T)0;
                        const int ix = fxs ? jx : Nx - 1 - jx;
                        const T hx = hx2[ix], hy = hy2[iy[k]], hz = hz2[iz[k]];
                        const T S = ((st[m] + hx) + hy) + hz;
                        const T Q = ((v + hx * fx[k]) + hy * fy) + hz * fz;
                        const T p = dd_div(Q, S);
                        gst(psi + roff[k] + ix, p);
                        const T p2 = p + p;
                        if (!tang) fx[k] = p2 - fx[k];
                        ofy[k] = p2 - fy;
                        ofz[k] = p2 - fz;
                        if (jx == Nx - 1) gst(bx + iy[k] + Ny * iz[k], fx[k]);
                        if (d.ndims >= 2 && jy[k] == Ny - 1) gst(by + ix + Nx * iz[k], ofy[k]);
                        if (d.ndims >= 3 && jz[k] == Nz - 1) gst(bz + ix + Nx * iy[k], ofz[k]);
                        if (NBUF == 2) {
                            Fy[wb * P + q] = ofy[k];
                            Fz[wb * P + q] = ofz[k];
                        }
                    }
                }
                if (NBUF == 1) {
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < PPT; ++k) {
                        if (act[k]) {
                            const int q = tid + k * bdim;
                            Fy[q] = ofy[k];
                            Fz[q] = ofz[k];
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// 1D/2D sweep: one wavefront (64 lanes) per direction, R rows per lane (rows l, l+64, ...).
// Row jy starts at step jy and walks x; the y-face flux of row jy comes from row jy-1 computed
// one step earlier -- a cross-lane shuffle inside the wave (lane 0 takes the previous slot of
// lane 63), so the sweep needs no barrier.  The angular source of the next PF steps of every
// row is prefetched into a register ring.
// ------------------------------------------------------------------------------------------
template <typename T, int R, int PF>
__global__ __launch_bounds__(64) void sweep_rows_kernel(const PassDesc* __restrict__ pds, const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    const int n = tk.dir;
    const int Nx = d.Nx, Ny = d.Ny, Nc = d.Nc;
    const int lane = threadIdx.x;
    T* hx2 = (T*)smem;
    T* st = hx2 + Nx;
    const double mu = d.mu[n], eta = d.eta[n];
    const bool tang = (d.ndims == 1) && (fabs(mu) <= 1e-10);  // sn_1D_bte.jl:154-155
    const double amu = tang ? 0.0 : fabs(mu);
    for (int i = lane; i < Nx; i += 64) hx2[i] = (T)(2.0 * (amu / d.dx[i]));
    for (int i = lane; i < d.Nmat; i += 64) st[i] = (T)d.sigt[i];
    __syncthreads();
    const bool fxs = mu >= 0, fys = eta >= 0;
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* __restrict__ bx = (T*)d.bx + (size_t)n * Ny;
    T* __restrict__ by = (T*)d.by + (size_t)n * Nx;
    const int* __restrict__ mat = d.mat;
    int jy[R], roff[R], iyv[R];
    bool valid[R];
    T hy[R], fx[R], fyo[R];
    T ring[R][PF];
    int mring[R][PF];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        jy[r] = lane + 64 * r;
        valid[r] = jy[r] < Ny;
        iyv[r] = valid[r] ? (fys ? jy[r] : Ny - 1 - jy[r]) : 0;
        roff[r] = Nx * iyv[r];
        hy[r] = (T)((d.ndims >= 2 && valid[r]) ? 2.0 * (fabs(eta) / d.dy[iyv[r]]) : 0.0);
        fx[r] = (T)0;
        fyo[r] = (T)0;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int jx = i - jy[r];
            ring[r][i] = (T)0;
            mring[r][i] = 0;
            if (valid[r] && jx >= 0 && jx < Nx) {
                const int ix = fxs ? jx : Nx - 1 - jx;
                ring[r][i] = gld(psi + roff[r] + ix);
                mring[r][i] = gld(mat + roff[r] + ix);
            }
        }
    }
    const int nsteps = Nx + Ny - 1;
    for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int s = s0 + i;
            if (s < nsteps) {
                T fyin[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const T up = __shfl_up(fyo[r], 1);
                    const T wrap = (r > 0) ? __shfl(fyo[r > 0 ? r - 1 : 0], 63) : (T)0;
                    fyin[r] = lane > 0 ? up : wrap;
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int jx = s - jy[r];
                    const T v = ring[r][i];
                    const int m = mring[r][i];
                    const int jx2 = jx + PF;
                    if (valid[r] && jx2 >= 0 && jx2 < Nx) {
                        const int ix2 = fxs ? jx2 : Nx - 1 - jx2;
                        ring[r][i] = gld(psi + roff[r] + ix2);
                        mring[r][i] = gld(mat + roff[r] + ix2);
                    }
                    if (valid[r] && jx >= 0 && jx < Nx) {
                        const int ix = fxs ? jx : Nx - 1 - jx;
                        const T hx = hx2[ix];
                        const T S = (st[m] + hx) + hy[r];
                        const T Q = (v + hx * fx[r]) + hy[r] * fyin[r];
                        const T p = dd_div(Q, S);
                        gst(psi + roff[r] + ix, p);
                        const T p2 = p + p;
                        if (!tang) fx[r] = p2 - fx[r];
                        fyo[r] = p2 - fyin[r];
                        if (jx == Nx - 1) gst(bx + iyv[r], fx[r]);
                        if (d.ndims >= 2 && jy[r] == Ny - 1) gst(by + ix, fyo[r]);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// 3D sweep: one workgroup per direction.  Lane l owns y-rows l + 64 r (r < R); wave w owns the
// ZW z-levels w*ZW .. w*ZW+ZW-1; every (y, z) pencil walks x with jx = s - jy - jz.
//   x-face flux: register of the pencil;  y-face: shuffle from lane l-1 (lane 0 takes lane 63's
//   previous row slot);  z-face: register of the pencil one z-level below in the same lane, or
//   for the first z-level of a wave the last z-level of the previous wave through LDS
//   (double-buffered by step parity: one barrier per step).
// ------------------------------------------------------------------------------------------
template <typename T, int R, int ZW, int PF, int WMAX>
__global__ __launch_bounds__(64 * WMAX) void sweep_planes_kernel(const PassDesc* __restrict__ pds, const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int K = R * ZW;  // pencils per lane
    const SweepTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    const int n = tk.dir;
    const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz, Nc = d.Nc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, W = blockDim.x >> 6;
    T* hx2 = (T*)smem;
    T* hy2 = hx2 + Nx;
    T* hz2 = hy2 + Ny;
    T* st = hz2 + Nz;
    T* Fz = st + d.Nmat;  // [2][W][64*R]
    const double mu = d.mu[n], eta = d.eta[n], xi = d.xi[n];
    for (int i = tid; i < Nx; i += blockDim.x) hx2[i] = (T)(2.0 * (fabs(mu) / d.dx[i]));
    for (int i = tid; i < Ny; i += blockDim.x) hy2[i] = (T)(2.0 * (fabs(eta) / d.dy[i]));
    for (int i = tid; i < Nz; i += blockDim.x) hz2[i] = (T)(2.0 * (fabs(xi) / d.dz[i]));
    for (int i = tid; i < d.Nmat; i += blockDim.x) st[i] = (T)d.sigt[i];
    __syncthreads();
    const bool fxs = mu >= 0, fys = eta >= 0, fzs = xi >= 0;
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* __restrict__ bx = (T*)d.bx + (size_t)n * Ny * Nz;
    T* __restrict__ by = (T*)d.by + (size_t)n * Nx * Nz;
    T* __restrict__ bz = (T*)d.bz + (size_t)n * Nx * Ny;
    const int* __restrict__ mat = d.mat;
    int js[K], roff[K], iyv[K], izv[K];
    bool valid[K];
    T hy[K], hz[K], fx[K], fyo[K], fzo[K];
    T ring[K][PF];
    int mring[K][PF];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int r = k / ZW, kz = k % ZW;
        const int jy = lane + 64 * r, jz = wave * ZW + kz;
        valid[k] = jy < Ny && jz < Nz;
        js[k] = jy + jz;
        iyv[k] = valid[k] ? (fys ? jy : Ny - 1 - jy) : 0;
        izv[k] = valid[k] ? (fzs ? jz : Nz - 1 - jz) : 0;
        roff[k] = Nx * (iyv[k] + Ny * izv[k]);
        hy[k] = hy2[iyv[k]];
        hz[k] = hz2[izv[k]];
        fx[k] = (T)0;
        fyo[k] = (T)0;
        fzo[k] = (T)0;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int jx = i - js[k];
            ring[k][i] = (T)0;
            mring[k][i] = 0;
            if (valid[k] && jx >= 0 && jx < Nx) {
                const int ix = fxs ? jx : Nx - 1 - jx;
                ring[k][i] = gld(psi + roff[k] + ix);
                mring[k][i] = gld(mat + roff[k] + ix);
            }
        }
    }
    const int nsteps = Nx + Ny + Nz - 2;
    const int lrow = 64 * R;
    for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int s = s0 + i;
            if (s < nsteps) {
                const int rb = (s + 1) & 1, wb = s & 1;
                T fyin[K], fzin[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int r = k / ZW, kz = k % ZW;
                    const T up = __shfl_up(fyo[k], 1);
                    const T wrap = (r > 0) ? __shfl(fyo[r > 0 ? k - ZW : 0], 63) : (T)0;
                    fyin[k] = lane > 0 ? up : wrap;
                    if (kz > 0) fzin[k] = fzo[kz > 0 ? k - 1 : 0];
                    else fzin[k] = wave > 0 ? Fz[(rb * W + wave - 1) * lrow + lane + 64 * r] : (T)0;
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int jx = s - js[k];
                    const T v = ring[k][i];
                    const int m = mring[k][i];
                    const int jx2 = jx + PF;
                    if (valid[k] && jx2 >= 0 && jx2 < Nx) {
                        const int ix2 = fxs ? jx2 : Nx - 1 - jx2;
                        ring[k][i] = gld(psi + roff[k] + ix2);
                        mring[k][i] = gld(mat + roff[k] + ix2);
                    }
                    if (valid[k] && jx >= 0 && jx < Nx) {
                        const int ix = fxs ? jx : Nx - 1 - jx;
                        const T hx = hx2[ix];
                        const T S = ((st[m] + hx) + hy[k]) + hz[k];
                        const T Q = ((v + hx * fx[k]) + hy[k] * fyin[k]) + hz[k] * fzin[k];
                        const T p = dd_div(Q, S);
                        gst(psi + roff[k] + ix, p);
                        const T p2 = p + p;
                        fx[k] = p2 - fx[k];
                        fyo[k] = p2 - fyin[k];
                        fzo[k] = p2 - fzin[k];
                        const int jy = lane + 64 * (k / ZW), jz = wave * ZW + k % ZW;
                        if (jx == Nx - 1) gst(bx + iyv[k] + Ny * izv[k], fx[k]);
                        if (jy == Ny - 1) gst(by + ix + Nx * izv[k], fyo[k]);
                        if (jz == Nz - 1) gst(bz + ix + Nx * iyv[k], fzo[k]);
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r) Fz[(wb * W + wave) * lrow + lane + 64 * r] = fzo[r * ZW + ZW - 1];
                __syncthreads();
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Cross-lane affine-map scan.  Maps x -> A + B x are composed "later after earlier":
// (A2, B2) o (A1, B1) = (A2 + B2 A1, B2 B1).  The DPP version builds the inclusive scan with
// row_shr:1,2,4,8 inside each 16-lane row, then row_bcast:15 / row_bcast:31 across rows (out-of-
// range and masked lanes receive the identity (0, 1)); the shuffle version is a Hillis-Steele scan
// through ds_bpermute.
// ------------------------------------------------------------------------------------------
template <int CTRL, int RMASK>
__device__ inline float dpp_move(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, RMASK, 0xF, false));
}
template <int CTRL, int RMASK>
__device__ inline double dpp_move(double old, double src) {
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffff), (int)(v & 0xffffffff), CTRL, RMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, RMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int RMASK, typename T>
__device__ inline void dpp_compose(T& A, T& B) {
    const T As = dpp_move<CTRL, RMASK>((T)0, A);
    const T Bs = dpp_move<CTRL, RMASK>((T)1, B);
    A = A + B * As;
    B = B * Bs;
}
// Shuffle scan within segments of SEG lanes (ls: the lane's index in its segment).
template <int SEG, typename T>
__device__ inline void affine_scan_seg(T& A, T& B, int ls) {
#pragma unroll
    for (int off = 1; off < SEG; off <<= 1) {
        const T As = __shfl_up(A, off);
        const T Bs = __shfl_up(B, off);
        if (ls >= off) {
            A = A + B * As;
            B = B * Bs;
        }
    }
}
template <bool DPP, typename T>
__device__ inline void affine_scan(T& A, T& B, int lane) {
    if (DPP) {
        dpp_compose<0x111, 0xF>(A, B);  // row_shr:1
        dpp_compose<0x112, 0xF>(A, B);  // row_shr:2
        dpp_compose<0x114, 0xF>(A, B);  // row_shr:4
        dpp_compose<0x118, 0xF>(A, B);  // row_shr:8
        dpp_compose<0x142, 0xA>(A, B);  // row_bcast:15 into rows 1 and 3
        dpp_compose<0x143, 0xC>(A, B);  // row_bcast:31 into rows 2 and 3
    } else {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T As = __shfl_up(A, off);
            const T Bs = __shfl_up(B, off);
            if (lane >= off) {
                A = A + B * As;
                B = B * Bs;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Row-scan sweeps: lanes along x.
//
// Along a pencil (a row of cells in x) the DD cell solve is an affine map of the incoming
// x-face flux:  psi_i = c0_i + c1_i fx_i  and  fx_{i+1} = 2 psi_i - fx_i = a_i + b_i fx_i with
//   c0 = (Q + hy fy + hz fz)/S,  c1 = hx/S,  a = 2 c0,  b = 2 c1 - 1   (|b| < 1),
// so a whole pencil is solved at once by an exclusive scan of the maps across the lanes (each
// lane first composes its CPL consecutive cells).
//
// Every global load and store is issued unconditionally: loads of inactive lanes or items read
// a clamped in-range address and their values are discarded; stores of inactive lanes go to a
// trash line.  With no memory operation under a branch the wait counters stay exact, so the
// loads issued for later pencils remain in flight while the current pencil is solved (a
// branch around any of them makes the compiler wait for all outstanding loads instead).
// ------------------------------------------------------------------------------------------
__device__ __align__(256) double g_sweep_trash[64 * 12];

// Slot k of a lane; slots of one lane are 64 elements apart so that stores to consecutive
// slots stay separate instructions (the count of stores in flight is part of the design).
template <typename T>
__device__ inline T* trash_slot(int lane, int k) { return (T*)g_sweep_trash + k * 64 + lane; }

// 1D/2D: one wave per direction; the rows y = 0..Ny-1 are solved in order with the y-face
// fluxes in registers (each lane owns its cells).  The angular source and material of row
// y + PD are loaded while row y is solved.  BIN: the x- and y-face inflows come from
// PassDesc::bin_x / bin_y (non-void boundaries) instead of zero.
template <typename T, int CPL, int PD, bool DPP, bool BIN = false>
__device__ __forceinline__ void sweep_line_body(const PassDesc& d, const int n, unsigned char* smem) {
    const int ndims = d.ndims, Nx = d.Nx, Ny = d.Ny, Nc = d.Nc, Nmat = d.Nmat;
    const int lane = threadIdx.x;
    T* hy2 = (T*)smem;   // [Ny]
    T* st = hy2 + Ny;    // [Nmat]
    const double mu = d.mu[n], eta = ndims >= 2 ? d.eta[n] : 0.0;
    const bool tang = (ndims == 1) && (fabs(mu) <= 1e-10);  // sn_1D_bte.jl:154-155
    const double amu = tang ? 0.0 : fabs(mu);
    for (int i = lane; i < Ny; i += 64) hy2[i] = (T)(ndims >= 2 ? 2.0 * (fabs(eta) / d.dy[i]) : 0.0);
    for (int i = lane; i < Nmat; i += 64) st[i] = (T)d.sigt[i];
    const bool fxs = mu >= 0, fys = eta >= 0;
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* __restrict__ bx = (T*)d.bx + (size_t)n * Ny;
    T* __restrict__ by = (T*)d.by + (size_t)n * Nx;
    const int* __restrict__ mat = d.mat;
    int ixc[CPL];
    bool inx[CPL];
    T hx[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const int jx = lane * CPL + c;
        inx[c] = jx < Nx;
        ixc[c] = inx[c] ? (fxs ? jx : Nx - 1 - jx) : 0;
        hx[c] = inx[c] ? (T)(2.0 * (amu / d.dx[ixc[c]])) : (T)0;
    }
    const bool last_lane = lane == (Nx - 1) / CPL;
    __syncthreads();
    auto row_of = [&](int jy) {
        const int j = jy < Ny ? jy : Ny - 1;
        return fys ? j : Ny - 1 - j;
    };
    const size_t npx = (size_t)Ny * d.Nz, npy = (size_t)Nx * d.Nz;  // face positions (Nz = 1 here)
    T fy[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) fy[c] = BIN ? (T)gld(d.bin_y + n * npy + ixc[c]) : (T)0;
    T vr[PD][CPL];
    int mr[PD][CPL];
    T xr[PD];  // x-face inflow of the row in each slot (BIN)
#pragma unroll
    for (int i = 0; i < PD; ++i) {
        // issued in slot order, so that slot 0 is the oldest load when the loop starts
        __builtin_amdgcn_sched_barrier(0);
        const int ro = Nx * row_of(i);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            vr[i][c] = gld(psi + ro + ixc[c]);
            mr[i][c] = gld(mat + ro + ixc[c]);
        }
        xr[i] = BIN ? (T)gld(d.bin_x + n * npx + row_of(i)) : (T)0;
    }
    __builtin_amdgcn_sched_barrier(0);
    for (int r0 = 0; r0 < Ny; r0 += PD) {
#pragma unroll
        for (int i = 0; i < PD; ++i) {
            const int jy = r0 + i;
            const bool rv = jy < Ny;
            const int iy = row_of(jy);
            const int ro = Nx * iy;
            const T hy = hy2[iy];
            T c0[CPL], c1[CPL];
            T A = (T)0, B = (T)1;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const T S = (st[mr[i][c]] + hx[c]) + hy;
                const T rS = dd_div((T)1, S);
                c0[c] = inx[c] ? (vr[i][c] + hy * fy[c]) * rS : (T)0;
                c1[c] = hx[c] * rS;
                const T a = tang ? (T)0 : c0[c] + c0[c];
                const T b = (tang || !inx[c]) ? (T)1 : (c1[c] + c1[c]) - (T)1;
                A = a + b * A;
                B = b * B;
            }
            {
                // Row jy + PD goes into the ring slot just freed.  The loads stay here: the
                // scheduler would otherwise sink them towards their first use, and a slot whose
                // old and new values overlap is rotated through copies that wait for the loads.
                __builtin_amdgcn_sched_barrier(0);
                const int ron = Nx * row_of(jy + PD);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    vr[i][c] = gld(psi + ron + ixc[c]);
                    mr[i][c] = gld(mat + ron + ixc[c]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            affine_scan<DPP>(A, B, lane);
            T fx = __shfl_up(A, 1);  // exclusive prefix applied to the row's inflow (0 if void)
            if (lane == 0) fx = (T)0;
            if constexpr (BIN) {
                T bb = __shfl_up(B, 1);
                if (lane == 0) bb = (T)1;
                fx = fx + bb * xr[i];
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const T p = c0[c] + c1[c] * fx;
                gst((rv && inx[c]) ? psi + ro + ixc[c] : trash_slot<T>(lane, c), p);
                const T p2 = p + p;
                fx = (tang || !inx[c]) ? fx : p2 - fx;
                fy[c] = rv ? p2 - fy[c] : fy[c];
            }
            gst((rv && last_lane) ? bx + iy : trash_slot<T>(lane, 8), fx);
            // the slot's next inflow is loaded after its last use (see the ring above)
            if constexpr (BIN) xr[i] = (T)gld(d.bin_x + n * npx + row_of(jy + PD));
        }
    }
    if (ndims >= 2) {
#pragma unroll
        for (int c = 0; c < CPL; ++c)
            if (inx[c]) gst(by + ixc[c], fy[c]);  // exit y-face after the last row
    }
}

// ------------------------------------------------------------------------------------------
// Row-scan line sweep, buffer-addressed form (the DPP scan's kernels).
//
// Each row is reached through a raw buffer resource whose base is the row and whose
// num_records is its byte length, so a lane's offsets stay fixed for the whole direction and
// the row address is scalar work.  Loads beyond num_records return 0 and stores beyond it are
// dropped by the hardware: a cell past Nx gets the offset BUF_OOB, a row past Ny gets
// num_records 0, and neither needs a select, a branch or a trash line.
//
// A cell past Nx is given the identity map (a = 0, b = 1): its source loads as 0, its y-face
// flux starts and stays 0, its hx is 0 (so c0 = c1 = 0) with 1 in place of hx in the removal
// (no 0/0 in a void material), and b = 2 c1 + bias with bias = 1 instead of -1.  The exit
// x-face of a row is then the flux after a lane's last cell, whatever the lane holds.
// ------------------------------------------------------------------------------------------
constexpr unsigned BUF_OOB = 0x80000000u;  // beyond every num_records used here
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
template <typename T> __device__ inline T bld(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ inline float bld<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <> __device__ inline int bld<int>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
template <> __device__ inline double bld<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
typedef unsigned rt_u32x2 __attribute__((ext_vector_type(2)));
__device__ inline void bst(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
__device__ inline void bst(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(rt_u32x2, v), r, (int)off, 0, 0);
}
// With a scalar (wave-uniform) byte offset soff added to the address.  Every access made this
// way keeps voff + soff below num_records when it is meant to land and puts voff at BUF_OOB
// when it is not, so it is masked the same whether or not the range check includes soff.
template <typename T> __device__ inline T bld_s(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff);
template <> __device__ inline float bld_s<float>(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
}
template <> __device__ inline int bld_s<int>(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0);
}
template <> __device__ inline double bld_s<double>(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
__device__ inline void bst_s(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, (int)soff, 0);
}
__device__ inline void bst_s(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(rt_u32x2, v), r, (int)off, (int)soff, 0);
}

// Inclusive affine scan of one FP32 map per lane with the cross-lane read fused into the
// arithmetic (v_fmac_f32_dpp: A += A[src] * B; v_mul_f32_dpp: B *= B[src]).  bound_ctrl is off,
// so a lane whose source lane is out of range, or whose row the row mask leaves out, is not
// written and keeps its map.  A DPP read of a VGPR needs two wait states after a VALU write of
// it: the s_nop 1 in front covers the compiler's last writes of A and B, the s_nop 0 between
// steps the fmac -> next fmac distance, and the s_nop 1 at the end the compiler's next DPP read
// (it does not count instructions inside an asm statement as VALU writes).
__device__ inline void affine_scan_f32_fused(float& A, float& B) {
    asm volatile(
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "s_nop 1"
        : "+v"(A), "+v"(B));
}
// One row_shr step of the FP64 scan.  There is no 64-bit DPP ALU form for row_shr, so the
// halves move separately; a lane whose source is out of range reads 0 (bound_ctrl), which is
// the identity of A's term and the low word of B's identity 1.0, so only B's high word needs an
// old value (0x3ff00000).
template <int CTRL>
__device__ inline void dpp_compose_shr_f64(double& A, double& B) {
    const long long a = __double_as_longlong(A), b = __double_as_longlong(B);
    const int alo = __builtin_amdgcn_mov_dpp((int)(a & 0xffffffff), CTRL, 0xF, 0xF, true);
    const int ahi = __builtin_amdgcn_mov_dpp((int)(a >> 32), CTRL, 0xF, 0xF, true);
    const int blo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, true);
    const int bhi = __builtin_amdgcn_update_dpp(0x3ff00000, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    const double As = __longlong_as_double(((long long)ahi << 32) | (unsigned)alo);
    const double Bs = __longlong_as_double(((long long)bhi << 32) | (unsigned)blo);
    A = A + B * As;
    B = B * Bs;
}
template <typename T>
__device__ inline void affine_scan_dpp(T& A, T& B) {
    if constexpr (sizeof(T) == 4) {
        affine_scan_f32_fused(A, B);
    } else {
        dpp_compose_shr_f64<0x111>(A, B);  // row_shr:1
        dpp_compose_shr_f64<0x112>(A, B);  // row_shr:2
        dpp_compose_shr_f64<0x114>(A, B);  // row_shr:4
        dpp_compose_shr_f64<0x118>(A, B);  // row_shr:8
        dpp_compose<0x142, 0xA>(A, B);     // row_bcast:15 into rows 1 and 3
        dpp_compose<0x143, 0xC>(A, B);     // row_bcast:31 into rows 2 and 3
    }
}
// The same scan within each half of the wave (lanes 0-31 and 32-63 scan separately): the row
// steps, then row_bcast:15 into rows 1 and 3 only.
__device__ inline void affine_scan_f32_fused32(float& A, float& B) {
    asm volatile(
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 0\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_mul_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1"
        : "+v"(A), "+v"(B));
}
template <typename T>
__device__ inline void affine_scan_dpp32(T& A, T& B) {
    if constexpr (sizeof(T) == 4) {
        affine_scan_f32_fused32(A, B);
    } else {
        dpp_compose_shr_f64<0x111>(A, B);
        dpp_compose_shr_f64<0x112>(A, B);
        dpp_compose_shr_f64<0x114>(A, B);
        dpp_compose_shr_f64<0x118>(A, B);
        dpp_compose<0x142, 0xA>(A, B);  // row_bcast:15 into rows 1 and 3
    }
}
// wave_shr:1 with lane 0 reading 0: the exclusive prefix of A
template <typename T>
__device__ inline T wave_shr_zero(T v) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
    } else {
        const long long a = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_mov_dpp((int)(a & 0xffffffff), 0x138, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_mov_dpp((int)(a >> 32), 0x138, 0xF, 0xF, true);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}
// 1 / s for the FP64 line sweep: the hardware reciprocal (about 2^-23 relative) refined by two
// Newton steps, r <- r + r (1 - s r), each doubling the correct bits; within about 1 ulp of the
// correctly rounded quotient, in 5 instructions instead of the 10 of an IEEE division.
__device__ inline double rcp_nr(double s) {
    double r = __builtin_amdgcn_rcp(s);
    double e = fma(-s, r, 1.0);
    r = fma(r, e, r);
    e = fma(-s, r, 1.0);
    return fma(r, e, r);
}
__device__ inline float rcp_nr(float s) { return __builtin_amdgcn_rcpf(s); }

// 1D/2D, one wave per direction, rows in order with the y-face fluxes in registers; the source
// and material of row y + PD are loaded while row y is solved.  BIN: the x- and y-face inflows
// come from PassDesc::bin_x / bin_y.
template <typename T, int CPL, int PD, bool BIN = false>
__device__ __forceinline__ void sweep_line_fast_body(const PassDesc& d, const int n, unsigned char* smem) {
    const int ndims = d.ndims, Nx = d.Nx, Ny = d.Ny, Nc = d.Nc, Nmat = d.Nmat;
    const int lane = threadIdx.x;
    T* hy2 = (T*)smem;   // [Ny]
    T* st = hy2 + Ny;    // [Nmat]
    const double mu = d.mu[n], eta = ndims >= 2 ? d.eta[n] : 0.0;
    // direction-uniform: the row order and the tangential 1D case (sn_1D_bte.jl:154-155)
    const int fys = __builtin_amdgcn_readfirstlane(eta >= 0 ? 1 : 0);
    const int tang = __builtin_amdgcn_readfirstlane((ndims == 1 && fabs(mu) <= 1e-10) ? 1 : 0);
    const bool fxs = mu >= 0;
    const double amu = fabs(mu);
    for (int i = lane; i < Ny; i += 64) hy2[i] = (T)(ndims >= 2 ? 2.0 * (fabs(eta) / d.dy[i]) : 0.0);
    for (int i = lane; i < Nmat; i += 64) st[i] = (T)d.sigt[i];
    const T* psi = (const T*)d.psi + (size_t)n * Nc;
    const int* mat = d.mat;
    unsigned offT[CPL], off4[CPL];
    T hx[CPL], hxs[CPL], bias[CPL];
    int ixc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const int jx = lane * CPL + c;
        const bool in = jx < Nx;
        const int ix = in ? (fxs ? jx : Nx - 1 - jx) : 0;
        ixc[c] = in ? ix : -1;
        offT[c] = in ? (unsigned)ix * (unsigned)sizeof(T) : BUF_OOB;
        off4[c] = in ? (unsigned)ix * 4u : BUF_OOB;
        hx[c] = in ? (T)(2.0 * (amu / d.dx[ix])) : (T)0;
        hxs[c] = in ? hx[c] : (T)1;
        bias[c] = in ? (T)-1 : (T)1;
    }
    const unsigned rowT = (unsigned)Nx * (unsigned)sizeof(T), row4 = (unsigned)Nx * 4u;
    const unsigned bxoff = lane == (Nx - 1) / CPL ? 0u : BUF_OOB;  // the lane with the row's last cell
    T* bx = (T*)d.bx + (size_t)n * Ny;
    const size_t npx = (size_t)Ny * d.Nz, npy = (size_t)Nx * d.Nz;  // face positions (Nz = 1 here)
    __syncthreads();
    if (tang) {  // psi = Q / Σt; the x-face flux passes through
        const auto rp = buf_rsrc(psi, rowT), rm = buf_rsrc(mat, row4);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const T v = bld<T>(rp, offT[c]);
            const int m = bld<int>(rm, off4[c]);
            bst(rp, offT[c], dd_div(v, st[m]));
        }
        const T fx0 = BIN ? (T)gld(d.bin_x + n * npx) : (T)0;
        bst(buf_rsrc(bx, (unsigned)sizeof(T)), bxoff, fx0);
        return;
    }
    auto row_of = [&](int jy) {  // scalar; rows past the grid map to the last one
        const int j = jy < Ny ? jy : Ny - 1;
        return fys ? j : Ny - 1 - j;
    };
    T fy[CPL];
    if constexpr (BIN) {
        const auto rb = buf_rsrc(d.bin_y + n * npy, (unsigned)Nx * 8u);
#pragma unroll
        for (int c = 0; c < CPL; ++c) fy[c] = (T)bld<double>(rb, ixc[c] >= 0 ? (unsigned)ixc[c] * 8u : BUF_OOB);
    } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) fy[c] = (T)0;
    }
    T vr[PD][CPL];
    int mr[PD][CPL];
    T xr[PD];  // x-face inflow of the row in each slot (BIN)
    // one resource per array for the whole direction (a row is a scalar byte offset), and one
    // with num_records 0 for the rows past the grid, which load zeros and move no bytes
    const auto rpsi = buf_rsrc(psi, (unsigned)Nc * (unsigned)sizeof(T)), rpsi0 = buf_rsrc(psi, 0u);
    const auto rmat = buf_rsrc(mat, (unsigned)Nc * 4u), rmat0 = buf_rsrc(mat, 0u);
    const auto rbx = buf_rsrc(bx, (unsigned)Ny * (unsigned)sizeof(T)), rbx0 = buf_rsrc(bx, 0u);
    auto load_row = [&](int i, int jy) __attribute__((always_inline)) {
        const int iy = row_of(jy);
        const bool live = jy < Ny;
        const unsigned so = (unsigned)(Nx * iy);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            vr[i][c] = bld_s<T>(live ? rpsi : rpsi0, offT[c], so * (unsigned)sizeof(T));
            mr[i][c] = bld_s<int>(live ? rmat : rmat0, off4[c], so * 4u);
        }
    };
#pragma unroll
    for (int i = 0; i < PD; ++i) {
        __builtin_amdgcn_sched_barrier(0);  // issued in slot order: slot 0 is the oldest load
        load_row(i, i);
        if constexpr (BIN) xr[i] = (T)gld(d.bin_x + n * npx + row_of(i));
    }
    __builtin_amdgcn_sched_barrier(0);
    // one row of slot i; rv is false only for the rows past the grid in the last block
    auto solve_row = [&](int i, int jy, bool rv) __attribute__((always_inline)) {
        const int iy = row_of(jy);
        const T hy = hy2[iy];
        T c0[CPL], c1[CPL], a[CPL], b[CPL];
        T A = (T)0, B = (T)1;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const T S = (st[mr[i][c]] + hxs[c]) + hy;
            const T rS = rcp_nr(S);
            c0[c] = (vr[i][c] + hy * fy[c]) * rS;
            c1[c] = hx[c] * rS;
            a[c] = c0[c] + c0[c];
            b[c] = (c1[c] + c1[c]) + bias[c];
            if (c == 0) {
                A = a[c];
                B = b[c];
            } else {
                A = a[c] + b[c] * A;
                B = b[c] * B;
            }
        }
        {
            // row jy + PD goes into the ring slot just freed (kept here: the scheduler would
            // otherwise sink the loads towards their first use)
            __builtin_amdgcn_sched_barrier(0);
            load_row(i, jy + PD);
            __builtin_amdgcn_sched_barrier(0);
        }
        affine_scan_dpp<T>(A, B);
        T fx = wave_shr_zero(A);  // the exclusive prefix, 0 in lane 0
        if constexpr (BIN) fx = fx + dpp_move<0x138, 0xF>((T)1, B) * xr[i];
        const unsigned so = (unsigned)(Nx * iy) * (unsigned)sizeof(T);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const T p = c0[c] + c1[c] * fx;
            bst_s(rv ? rpsi : rpsi0, offT[c], so, p);
            const T fyn = (p + p) - fy[c];
            fy[c] = rv ? fyn : fy[c];
            fx = a[c] + b[c] * fx;
        }
        bst_s(rv ? rbx : rbx0, bxoff, (unsigned)iy * (unsigned)sizeof(T), fx);
        if constexpr (BIN) xr[i] = (T)gld(d.bin_x + n * npx + row_of(jy + PD));
    };
    int r0 = 0;
    for (; r0 + PD <= Ny; r0 += PD) {
#pragma unroll
        for (int i = 0; i < PD; ++i) solve_row(i, r0 + i, true);
    }
    if (r0 < Ny) {
#pragma unroll
        for (int i = 0; i < PD; ++i) solve_row(i, r0 + i, r0 + i < Ny);
    }
    if (ndims >= 2) {
        T* by = (T*)d.by + (size_t)n * Nx;
#pragma unroll
        for (int c = 0; c < CPL; ++c)
            if (ixc[c] >= 0) gst(by + ixc[c], fy[c]);  // exit y-face after the last row
    }
}

template <typename T, int CPL, int PD, bool DPP, bool BIN = false>
__global__ __launch_bounds__(64) void sweep_line_kernel(const PassDesc* __restrict__ pds,
                                                        const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    if constexpr (DPP) sweep_line_fast_body<T, CPL, PD, BIN>(pds[tk.prob], tk.dir, smem);
    else sweep_line_body<T, CPL, PD, DPP, BIN>(pds[tk.prob], tk.dir, smem);
}

// FP64 and FP32 passes of different problems in one launch: each workgroup takes the precision of
// its problem's pass (PassDesc::prec; BF16 passes sweep in FP32).
template <int CPL, int PD, bool DPP>
__global__ __launch_bounds__(64) void sweep_line_mixed_kernel(const PassDesc* __restrict__ pds,
                                                              const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    if constexpr (DPP) {
        if (d.prec == P64) sweep_line_fast_body<double, CPL, PD>(d, tk.dir, smem);
        else sweep_line_fast_body<float, CPL, PD>(d, tk.dir, smem);
    } else {
        if (d.prec == P64) sweep_line_body<double, CPL, PD, DPP>(d, tk.dir, smem);
        else sweep_line_body<float, CPL, PD, DPP>(d, tk.dir, smem);
    }
}

// The same as a persistent work list: a fixed grid of single-wave workgroups takes the tasks in
// table order (largest problems first) from a counter, zeroed by the task expansion kernel that
// precedes the launch on its stream, until the table is exhausted.  A wave's tasks run one after
// another; every wave leaves once its fetch passes the end.
template <int CPL, int PD, bool DPP>
__global__ __launch_bounds__(64) void sweep_line_worklist_kernel(const PassDesc* __restrict__ pds,
                                                                 const SweepTask* __restrict__ tasks, int ntasks,
                                                                 unsigned* ctr) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (;;) {
        unsigned k = 0;
        if (threadIdx.x == 0) k = atomicAdd(ctr, 1u);
        k = __builtin_amdgcn_readfirstlane(k);
        if (k >= (unsigned)ntasks) return;
        const SweepTask tk = tasks[k];
        const PassDesc& d = pds[tk.prob];
        if constexpr (DPP) {
            if (d.prec == P64) sweep_line_fast_body<double, CPL, PD>(d, tk.dir, smem);
            else sweep_line_fast_body<float, CPL, PD>(d, tk.dir, smem);
        } else {
            if (d.prec == P64) sweep_line_body<double, CPL, PD, DPP>(d, tk.dir, smem);
            else sweep_line_body<float, CPL, PD, DPP>(d, tk.dir, smem);
        }
    }
}

// 3D: one workgroup per direction.  Pencils (jy, jz) advance by diagonals t = jy + jz; the
// y/z face fluxes of the previous diagonal live in LDS (FY[jz][x], FZ[jy][x]), one barrier per
// diagonal.  After the last diagonal FY and FZ hold the exit y- and z-faces, which are written
// out at the end with coalesced stores.  PPW pencils per wave: with PPW = 2 each half of the wave
// (32 lanes, CPL cells per lane) solves its own pencil of the item, and the scans run per half.
//
// SPL: the direction's pencils are split by z into S slabs of Zb = ceil(Nz / S) layers, one
// workgroup each (scan_split_kernel).  Slab k's top layer hands its outgoing z-face rows to slab
// k + 1 through device memory: every value goes out as 64-bit words that carry the launch's tag in
// their high half (one word per float, two per double), so a word is valid exactly when its tag is
// the current one and no fence or flag is needed.  The slab above copies the rows into its FZ in
// LDS with one extra wave (the courier, the workgroup's last), a few rows ahead of the diagonal
// that reads them, so the solving waves' loads and waits stay as in one workgroup per direction.
// The workgroups of a direction take their slabs in the order they start (a counter per
// direction, tagged the same way), so a slab only ever waits for a workgroup that is already
// running, and the first never waits: the launch cannot deadlock whatever the dispatch order.
struct ScanSplit {
    int S;         // slabs per direction (1: no split)
    unsigned tag;  // launch tag, never 0
    int lag;       // rows the slab below is ahead when a slab starts
};
constexpr unsigned SCAN_OOB = 0x40000000u;  // buffer offset past every num_records of the 3D scan
constexpr int SCAN_SPLIT_ROWS = 6;  // rows the courier keeps in flight
constexpr int SCAN_SPLIT_LAG = 9;   // rows the slab below is ahead when a slab starts (default)
constexpr int SCAN_SPLIT_SPIN = 1 << 20;  // polls of handoff rows per workgroup before values are given up (NaN)

// The handed words are read and written with device-scope read-modify-write atomics, which the
// workgroups of every XCD see at once (plain loads and stores, even agent-scope ones, can be served
// by an XCD's own L2 for a long time).  Global (not flat) accesses: a flat access also counts on the
// LDS counter, so a wait for the solving waves' LDS reads would wait for the handoff too.
__device__ __forceinline__ unsigned long long split_ld(const unsigned long long* p) {
    return __hip_atomic_fetch_or((RT_GLOBAL unsigned long long*)p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void split_st(unsigned long long* p, unsigned long long v) {
    (void)__hip_atomic_exchange((RT_GLOBAL unsigned long long*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int CPL, int U, int PF, bool DPP, bool BIN = false, int PPW = 1, bool SPL = false>
__device__ __forceinline__ void sweep_scan_body(const PassDesc& d, const int n, unsigned char* smem,
                                                const ScanSplit sp = {1, 0u, 0}) {
    const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz, Nc = d.Nc, Nmat = d.Nmat;
    // the wave index is uniform: in a scalar register, the item cursors and barrier counts are too
    // (W counts the waves that solve pencils; with SPL the courier wave comes after them)
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int W = (blockDim.x >> 6) - (SPL ? 1 : 0);
    static_assert(PPW == 1 || PPW == 2, "one or two pencils per wave");
    const int half = PPW == 2 ? lane >> 5 : 0;    // the pencil of the item this lane works on
    const int ll = PPW == 2 ? lane & 31 : lane;    // lane within the pencil
    constexpr int UP = U * PPW;                    // pencils per item
    constexpr int WPV = sizeof(T) / 4;             // handoff words per value
    // slab of this workgroup: layers z0 .. z0 + Zc - 1 (sweep order) of Zb per slab
    int role = 0, z0 = 0, Zc = Nz, Zb = Nz;
    unsigned long long* hout = nullptr;        // rows handed to the slab above (null: top slab)
    const unsigned long long* hin = nullptr;   // rows handed from the slab below (null: first slab)
    if constexpr (SPL) {
        __shared__ int s_role;
        if (tid == 0) {
            unsigned long long* cw = d.sclaim + n;
            unsigned long long old = split_ld(cw);
            int r = 0;
            for (;;) {  // at most S rounds: every failed exchange is another workgroup's success
                const bool cur = (old >> 8) == (unsigned long long)sp.tag;
                r = cur ? (int)(old & 255) : 0;
                const unsigned long long want = cur ? old + 1 : (((unsigned long long)sp.tag << 8) | 1ull);
                if (__hip_atomic_compare_exchange_strong(cw, &old, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
            s_role = r;
        }
        __syncthreads();
        role = __builtin_amdgcn_readfirstlane(s_role);
        Zb = (Nz + sp.S - 1) / sp.S;
        z0 = role * Zb;
        Zc = Nz - z0 < Zb ? Nz - z0 : Zb;
        const size_t hrow = (size_t)Ny * Nx * 2;  // words per handed face (sized for doubles)
        unsigned long long* hb = d.shand + (size_t)n * (SCAN_SPLIT_MAX - 1) * hrow;
        if (role + 1 < sp.S) hout = hb + (size_t)role * hrow;
        if (role > 0) hin = hb + (size_t)(role - 1) * hrow;
    }
    // A face row of Nx cells is held in RS slots with cell jx at slot(jx) = (jx % CPL) La + jx / CPL
    // (La lanes hold the row's cells): lane l's cell c sits at c La + l, so the lanes of one LDS
    // access read consecutive slots, free of bank conflicts.
    const int La = (Nx + CPL - 1) / CPL, RS = CPL * La;
    auto slot = [&](int jx) { return (jx % CPL) * La + jx / CPL; };
    T* hy2 = (T*)smem;
    T* hz2 = hy2 + Ny;
    T* st = hz2 + Nz;
    T* FY = st + Nmat;                   // [Zb][RS]: y-faces of the slab's layers
    T* FZ = FY + (size_t)Zb * RS;        // [Ny][RS]
    T* FT = FZ + (size_t)Ny * RS;        // [64*CPL] trash for inactive lanes
    const double mu = d.mu[n], eta = d.eta[n], xi = d.xi[n];
    for (int i = tid; i < Ny; i += blockDim.x) hy2[i] = (T)(2.0 * (fabs(eta) / d.dy[i]));
    for (int i = tid; i < Nz; i += blockDim.x) hz2[i] = (T)(2.0 * (fabs(xi) / d.dz[i]));
    for (int i = tid; i < Nmat; i += blockDim.x) st[i] = (T)d.sigt[i];
    // the octant is uniform: scalar, so that the pencil index work below stays scalar
    const bool fxs = __builtin_amdgcn_readfirstlane((int)(mu >= 0)) != 0;
    const bool fys = __builtin_amdgcn_readfirstlane((int)(eta >= 0)) != 0;
    const bool fzs = __builtin_amdgcn_readfirstlane((int)(xi >= 0)) != 0;
    const size_t npx = (size_t)Ny * Nz, npy = (size_t)Nx * Nz, npz = (size_t)Nx * Ny;  // face positions
    // inflow on the entry y- and z-faces: zero for void boundaries; the z-faces of a slab above
    // the first come from the slab below (written by the courier)
    const int nfz = (Zb + (role == 0 ? Ny : 0)) * Nx;
    for (int i = tid; i < nfz; i += blockDim.x) {
        const int r = i / Nx, jx = i - r * Nx;  // row of FY, then of FZ, and cell
        T v = (T)0;
        if constexpr (BIN) {
            if (r < Zb) {
                const int jz = z0 + r;
                const int ix = fxs ? jx : Nx - 1 - jx, iz = fzs ? jz : Nz - 1 - jz;
                if (jz < Nz) v = (T)gld(d.bin_y + n * npy + ix + (size_t)Nx * iz);
            } else {
                const int jy = r - Zb;
                const int ix = fxs ? jx : Nx - 1 - jx, iy = fys ? jy : Ny - 1 - jy;
                v = (T)gld(d.bin_z + n * npz + ix + (size_t)Nx * iy);
            }
        }
        FY[(size_t)r * RS + slot(jx)] = v;
    }
    int ixc[CPL];
    bool inx[CPL];
    T hx[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const int jx = ll * CPL + c;
        inx[c] = jx < Nx;
        ixc[c] = inx[c] ? (fxs ? jx : Nx - 1 - jx) : 0;
        hx[c] = inx[c] ? (T)(2.0 * (fabs(mu) / d.dx[ixc[c]])) : (T)0;
    }
    const bool last_lane = ll == (Nx - 1) / CPL;
    if constexpr (SPL) {
        if (wave == W) {
            // The courier: row jy of the slab below's top z-faces goes into FZ[jy] before barrier
            // jy (the one that opens diagonal jy, whose pencil (jy, z0) reads it first), with the
            // rows up to SCAN_SPLIT_ROWS ahead already requested.  It passes the same 1 + ndiag
            // barriers as the solving waves.
            constexpr int CR = SCAN_SPLIT_ROWS, KW = 2;  // KW words per lane hold a row (Nx * WPV <= 128)
            const int ndg = Ny + Zc - 1, nw = Nx * WPV;
            const unsigned long long want = (unsigned long long)sp.tag << 32;
            auto valid = [&](unsigned long long w) { return (w >> 32) == (unsigned long long)sp.tag; };
            auto issue = [&](unsigned long long (&w)[KW], int jy) {
#pragma unroll
                for (int k = 0; k < KW; ++k) {
                    const int i = lane + 64 * k;
                    w[k] = i < nw ? split_ld(hin + (size_t)jy * nw + i) : want;
                }
            };
            // polls the words of row jy that are not valid yet, a bounded number of times in all:
            // once the bound is spent the remaining rows are not waited for
            int spins = 0;
            auto settle = [&](unsigned long long (&w)[KW], int jy) {
                bool ok = true;
#pragma unroll
                for (int k = 0; k < KW; ++k) ok = ok && valid(w[k]);
                for (; spins < SCAN_SPLIT_SPIN && !__all(ok); ++spins) {
                    __builtin_amdgcn_s_sleep(8);
                    ok = true;
#pragma unroll
                    for (int k = 0; k < KW; ++k) {
                        const int i = lane + 64 * k;
                        if (!valid(w[k])) w[k] = split_ld(hin + (size_t)jy * nw + i);
                        ok = ok && valid(w[k]);
                    }
                }
            };
            unsigned long long rw[CR][KW];
            if (role > 0) {
                const int jl = sp.lag < Ny - 1 ? sp.lag : Ny - 1;
                issue(rw[0], jl);
                settle(rw[0], jl);  // start once the slab below is sp.lag rows ahead
#pragma unroll
                for (int r = 0; r < CR; ++r)
                    if (r < Ny) issue(rw[r], r);
            }
            for (int t0 = 0; t0 <= ndg; t0 += CR) {
#pragma unroll
                for (int r = 0; r < CR; ++r) {
                    const int t = t0 + r;
                    if (t > ndg) break;
                    if (role > 0 && t < Ny) {
                        settle(rw[r], t);
                        unsigned char* dst = (unsigned char*)(FZ + (size_t)t * RS);
#pragma unroll
                        for (int k = 0; k < KW; ++k) {
                            const int i = lane + 64 * k;  // word w of cell jx (rows are handed in cell order)
                            const int jx = i / WPV, w = i - jx * WPV;
                            // a word still invalid after the bound becomes NaN: the solve fails loudly
                            const unsigned v = valid(rw[r][k]) ? (unsigned)rw[r][k] : 0xffffffffu;
                            if (i < nw) __builtin_memcpy(dst + 4 * ((size_t)slot(jx) * WPV + w), &v, 4);
                        }
                        if (t + CR < Ny) issue(rw[r], t + CR);
                    }
                    __syncthreads();
                }
            }
            return;
        }
    }
    __syncthreads();
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T* __restrict__ bx = (T*)d.bx + (size_t)n * Ny * Nz;
    T* __restrict__ by = (T*)d.by + (size_t)n * Nx * Nz;
    T* __restrict__ bz = (T*)d.bz + (size_t)n * Nx * Ny;
    // The angular flux, the material map and the x exit face go through buffer resources: offsets
    // are 32-bit, and a load or store at an offset of SCAN_OOB or more (past num_records: the host
    // keeps a direction's cells within SCAN_OOB bytes) reads 0 or is dropped, so a pencil that does
    // not exist and a cell past Nx need neither a clamped address nor a trash line.
    constexpr unsigned es = sizeof(T);
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(psi, (unsigned)((size_t)Nc * es));
    const __amdgpu_buffer_rsrc_t rm = buf_rsrc(d.mat, (unsigned)((size_t)Nc * 4));
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(bx, (unsigned)((size_t)Ny * Nz * es));
    unsigned xl[CPL], xs[CPL];  // a cell's byte offset in its row, for loads and stores (SCAN_OOB: past Nx)
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        xl[c] = (unsigned)ixc[c] * es;
        xs[c] = inx[c] ? xl[c] : SCAN_OOB;
    }
    // Pencil indices.  In sweep order pencil (jy, jz) of diagonal t = jy + jz (jz within the slab)
    // has iy = ay + sy jy, iz = az + sz jz and row iy + Ny iz = rb0 + sy t + dr jz.  The item's first
    // pencil q0 is scalar work; this lane's pencil is q0 + hq (hq: its half-wave with two pencils per
    // wave), whose indices differ from q0's by per-lane constants.
    const int sy = fys ? 1 : -1, ay = fys ? 0 : Ny - 1;
    const int sz = fzs ? 1 : -1, az = fzs ? z0 : Nz - 1 - z0;
    const int dr = Ny * sz - sy, rb0 = ay + Ny * az;
    const int hq = PPW == 2 ? half : 0;
    const int h_row = hq * dr, h_rs = hq * RS, h_nx = hq * Nx, h_iy = -sy * hq, h_iz = sz * hq;
    // byte offsets of the lane's pencil past q0's: angular flux row, material row, x exit face
    const unsigned h_rp = (unsigned)(h_row * Nx) * es, h_rm = (unsigned)(h_row * Nx) * 4u, h_rb = (unsigned)h_row * es;
    // A scalar value passed through readfirstlane stays scalar: the compiler would otherwise fold
    // the per-lane constant in before a multiply and do the multiply per lane.
    auto sc = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    const int ndiag = Ny + Zc - 1;
    // Work items of this wave, in order: (diagonal t, first pencil q) for q = wave*U, wave*U + W*U,
    // ... below the diagonal's pencil count, diagonal by diagonal (diagonals where the wave has no
    // pencil are skipped).  The angular source and material of the item PF places ahead are
    // loaded while the current one is solved, into a ring of PF register slots.
    auto diag_lo = [&](int t) { return t - (Ny - 1) > 0 ? t - (Ny - 1) : 0; };
    auto diag_cnt = [&](int t) { return (t < Zc - 1 ? t : Zc - 1) - diag_lo(t) + 1; };  // <= 0 past the end
    auto advance = [&](int& t, int& q) {
        if (q + W * UP < diag_cnt(t)) {
            q += W * UP;
            return;
        }
        ++t;
        while (t < ndiag && wave * UP >= diag_cnt(t)) ++t;
        q = wave * UP;
    };
    T vb[PF][U][CPL];
    int mb[PF][U][CPL];
    T xb[PF][U];  // x-face inflow of each pencil of the item in a slot (BIN)
    // the row (iy + Ny iz) of pencil q0 of diagonal t (scalar)
    auto row_of = [&](int t, int q0) { return rb0 + sy * t + dr * (diag_lo(t) + q0); };
    auto load_inflow = [&](T (&x)[U], int t, int q0) {
        const int tt = t < ndiag ? t : ndiag - 1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int qu = q0 + u * PPW;
            const bool okl = hq < diag_cnt(tt) - qu;
            x[u] = (T)gld(d.bin_x + n * npx + (okl ? sc(row_of(tt, qu)) + h_row : 0));
        }
    };
    auto load_item = [&](T (&v)[U][CPL], int (&m)[U][CPL], int t, int q0) {
        const int tt = t < ndiag ? t : ndiag - 1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int qu = q0 + u * PPW, row = row_of(tt, qu);
            const bool okl = hq < diag_cnt(tt) - qu;
            // byte offsets of this lane's row in the angular flux and in the material map
            const unsigned rp_o = okl ? (unsigned)sc(row * Nx * (int)es) + h_rp : SCAN_OOB;
            const unsigned rm_o = okl ? (unsigned)sc(row * Nx * 4) + h_rm : SCAN_OOB;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                v[u][c] = bld<T>(rp, rp_o + xl[c]);
                m[u][c] = bld<int>(rm, rm_o + (unsigned)ixc[c] * 4u);
            }
        }
    };
    int ct = 0, cq = wave * UP;  // item being solved
    while (ct < ndiag && wave * UP >= diag_cnt(ct)) ++ct;
    int lt = ct, lq = cq;       // next item to load
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int sl = 0; sl < PF; ++sl) {
        load_item(vb[sl], mb[sl], lt, lq);
        if constexpr (BIN) load_inflow(xb[sl], lt, lq);
        advance(lt, lq);
        // As many stores as an item issues after its prefetch, so that the loop is entered with
        // the same memory operations in flight as it repeats with (exact wait counts).
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < U * (CPL + 1); ++k) gst(trash_slot<T>(lane, k), (T)0);
        __builtin_amdgcn_sched_barrier(0);
    }
    int bt = 0;  // barriers passed = diagonals completed by every wave
    while (ct < ndiag) {
#pragma unroll
        for (int sl = 0; sl < PF; ++sl) {
            // An item of diagonal t needs every item of diagonals < t done.  Items past the end
            // (at most PF - 1 per wave) solve a clamped pencil and store only to trash.
            const int tb = ct < ndiag ? ct : ndiag;
            while (bt < tb) {
                __syncthreads();
                ++bt;
            }
            const int t = ct < ndiag ? ct : ndiag - 1;
            const int cnt = diag_cnt(t), jz_lo = diag_lo(t);
            T fy[U][CPL], fz[U][CPL], c0[U][CPL], c1[U][CPL];
            unsigned soff[U], boff[U];
            int fyo_i[U], fzo_i[U], hro_i[U];
            bool ok[U], top[U];
            T A[U], B[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q0 = cq + u * PPW, jz0 = jz_lo + q0;  // the item's first pencil (scalar)
                ok[u] = ct < ndiag && hq < cnt - q0;
                top[u] = ok[u] && jz0 + hq == Zc - 1;
                const int row = rb0 + sy * t + dr * jz0;
                soff[u] = ok[u] ? (unsigned)sc(row * Nx * (int)es) + h_rp : SCAN_OOB;
                boff[u] = (ok[u] && last_lane) ? (unsigned)sc(row * (int)es) + h_rb : SCAN_OOB;
                // LDS indices, clamped into their arrays for a pencil that does not exist (it stores
                // only to trash)
                fyo_i[u] = min(sc(jz0 * RS) + h_rs, (Zc - 1) * RS);
                fzo_i[u] = max(sc((t - jz0) * RS) - h_rs, 0);
                if constexpr (SPL) hro_i[u] = sc((t - jz0) * Nx) - h_nx;  // handed row of jy (cell order)
                const int iy = min(max(sc(ay + sy * (t - jz0)) + h_iy, 0), Ny - 1);
                const int iz = min(max(sc(az + sz * jz0) + h_iz, 0), Nz - 1);
                const T hy = hy2[iy], hz = hz2[iz];
                A[u] = (T)0;
                B[u] = (T)1;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int jx = inx[c] ? c * La + ll : 0;  // slot of the lane's cell c
                    fy[u][c] = FY[fyo_i[u] + jx];
                    fz[u][c] = FZ[fzo_i[u] + jx];
                    const T S = ((st[mb[sl][u][c]] + hx[c]) + hy) + hz;
                    const T rS = rcp_nr(S);
                    c0[u][c] = inx[c] ? ((vb[sl][u][c] + hy * fy[u][c]) + hz * fz[u][c]) * rS : (T)0;
                    c1[u][c] = hx[c] * rS;
                    // compose this lane's maps (inactive cells: identity)
                    const T a = c0[u][c] + c0[u][c];
                    const T b = inx[c] ? (c1[u][c] + c1[u][c]) - (T)1 : (T)1;
                    A[u] = a + b * A[u];
                    B[u] = b * B[u];
                }
            }
            // The composed maps depend on every use of the slot's values: pinning them here keeps
            // that arithmetic ahead of the slot's reload (instruction selection would otherwise
            // place it after the loads, and the ring would rotate through copies that wait).
#pragma unroll
            for (int u = 0; u < U; ++u) asm volatile("" ::"v"(A[u]), "v"(B[u]));
            const int xt = lt, xq = lq;  // the item whose inflow goes into this slot after its last use
            {
                // the item PF ahead goes into the slot just consumed (see sweep_line_kernel)
                __builtin_amdgcn_sched_barrier(0);
                load_item(vb[sl], mb[sl], lt, lq);
                advance(lt, lq);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (DPP && PPW == 2) affine_scan_dpp32(A[u], B[u]);
                else if constexpr (DPP) affine_scan_dpp(A[u], B[u]);
                else affine_scan_seg<PPW == 2 ? 32 : 64>(A[u], B[u], ll);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // exclusive prefix applied to the pencil's inflow (0 if void)
                T fx;
                if constexpr (DPP) {
                    fx = wave_shr_zero(A[u]);
                    if (PPW == 2 && ll == 0) fx = (T)0;
                } else {
                    fx = __shfl_up(A[u], 1);
                    if (ll == 0) fx = (T)0;
                }
                if constexpr (BIN) {
                    T bb = __shfl_up(B[u], 1);
                    if (ll == 0) bb = (T)1;
                    fx = fx + bb * xb[sl][u];
                }
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const bool act = ok[u] && inx[c];
                    const int jx = c * La + ll, jt = lane * CPL + c;  // slot of cell c; trash slot
                    const T p = c0[u][c] + c1[u][c] * fx;
                    bst(rp, soff[u] + xs[c], p);
                    const T p2 = p + p;
                    fx = inx[c] ? p2 - fx : fx;
                    *(act ? FY + fyo_i[u] + jx : FT + jt) = p2 - fy[u][c];
                    *(act ? FZ + fzo_i[u] + jx : FT + jt) = p2 - fz[u][c];
                    if constexpr (SPL) {
                        // the top layer's outgoing z-face goes to the slab above as tagged words
                        if (hout && act && top[u]) {
                            const T zo = p2 - fz[u][c];
                            unsigned wv[WPV];
                            __builtin_memcpy(wv, &zo, sizeof(T));
#pragma unroll
                            for (int w = 0; w < WPV; ++w)
                                split_st(hout + (size_t)(hro_i[u] + ll * CPL + c) * WPV + w,
                                         ((unsigned long long)sp.tag << 32) | wv[w]);
                        }
                    }
                }
                bst(rb, boff[u], fx);
            }
            if constexpr (BIN) load_inflow(xb[sl], xt, xq);
            advance(ct, cq);
        }
    }
    while (bt < ndiag) {
        __syncthreads();
        ++bt;
    }
    // exit faces: FY[jz][jx] -> by[ix + Nx iz] (the slab's layers), FZ[jy][jx] -> bz[ix + Nx iy]
    // (the top slab)
    for (int i = tid; i < Zc * Nx; i += 64 * W) {
        const int r = i / Nx, jz = z0 + r, jx = i - r * Nx;
        const int ix = fxs ? jx : Nx - 1 - jx, iz = fzs ? jz : Nz - 1 - jz;
        gst(by + ix + Nx * iz, FY[(size_t)r * RS + slot(jx)]);
    }
    if (!hout)
        for (int i = tid; i < Ny * Nx; i += 64 * W) {
            const int jy = i / Nx, jx = i - jy * Nx;
            const int ix = fxs ? jx : Nx - 1 - jx, iy = fys ? jy : Ny - 1 - jy;
            gst(bz + ix + Nx * iy, FZ[(size_t)jy * RS + slot(jx)]);
        }
}

template <typename T, int CPL, int U, int PF, int MAXT, bool DPP, bool BIN = false, int PPW = 1>
__global__ __launch_bounds__(MAXT) void sweep_scan_kernel(const PassDesc* __restrict__ pds, const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    sweep_scan_body<T, CPL, U, PF, DPP, BIN, PPW>(pds[tk.prob], tk.dir, smem);
}

// A direction's pencils split by z over S workgroups (blockIdx / S is the task); the last wave of
// each workgroup is the courier of its slab.
template <typename T, int CPL, int PF, int PPW>
__global__ __launch_bounds__(1024) void sweep_scan_split_kernel(const PassDesc* __restrict__ pds,
                                                                const SweepTask* __restrict__ tasks, int S,
                                                                unsigned tag, int lag) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x / S];
    sweep_scan_body<T, CPL, 1, PF, true, false, PPW, true>(pds[tk.prob], tk.dir, smem, ScanSplit{S, tag, lag});
}

template <int CPL, int U, int PF, int MAXT, bool DPP, int PPW = 1>
__global__ __launch_bounds__(MAXT) void sweep_scan_mixed_kernel(const PassDesc* __restrict__ pds,
                                                                const SweepTask* __restrict__ tasks) {
    extern __shared__ __align__(16) unsigned char smem[];
    const SweepTask tk = tasks[blockIdx.x];
    const PassDesc& d = pds[tk.prob];
    if (d.prec == P64) sweep_scan_body<double, CPL, U, PF, DPP, false, PPW>(d, tk.dir, smem);
    else sweep_scan_body<float, CPL, U, PF, DPP, false, PPW>(d, tk.dir, smem);
}

// ------------------------------------------------------------------------------------------
// Half-range moments of the exit faces (sn_sweep_*D.jl "Save boundary fluxes" + the
// 𝚽*12_temp accumulation of sn_inner_pass.jl:322-329), summed over directions in index order.
// One thread per (axis, face, moment, position).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bndmom_kernel(const PassDesc* __restrict__ pds, int nprob) {
    for (int pi = 0; pi < nprob; ++pi) {
        const PassDesc& d = pds[pi];
        if (!d.bout_x) continue;
        const int Nps = d.Nps, Nd = d.Nd;
        // positions times face moments per axis (general 2D closures; 1 otherwise)
        const int fm[3] = {d.fm_x > 1 ? d.fm_x : 1, d.fm_y > 1 ? d.fm_y : 1, d.fm_z > 1 ? d.fm_z : 1};
        const int npos[3] = {d.Ny * d.Nz * fm[0], d.Nx * d.Nz * fm[1], d.Nx * d.Ny * fm[2]};
        const size_t tot = (size_t)2 * Nps * (npos[0] + (d.ndims >= 2 ? npos[1] : 0) + (d.ndims >= 3 ? npos[2] : 0));
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
            size_t j = i;
            int ax = 0;
            while (j >= (size_t)2 * Nps * npos[ax]) { j -= (size_t)2 * Nps * npos[ax]; ++ax; }
            // j = p + Nps*(m + fm*(f + 2*pos)); the exit fluxes are [Nd][pos][m]
            const int p = (int)(j % Nps);
            const int m = (int)((j / Nps) % fm[ax]);
            const int f = (int)((j / ((size_t)Nps * fm[ax])) % 2);
            const size_t pos = (j / ((size_t)2 * Nps * fm[ax])) * fm[ax] + m;
            const double* Dex = ax == 0 ? d.Dexit_x : (ax == 1 ? d.Dexit_y : d.Dexit_z);
            const int* lst = (ax == 0 ? d.exlist_x : (ax == 1 ? d.exlist_y : d.exlist_z)) + f * Nd;
            const int cnt = d.excnt[ax][f];
            // exit-face fluxes in the pass's precision
            const void* bsrc = ax == 0 ? d.bx : (ax == 1 ? d.by : d.bz);
            double v = 0.0;
            if (d.prec == P64) {
                const double* b64 = (const double*)bsrc;
#pragma unroll 8
                for (int t = 0; t < cnt; ++t) {
                    const int n = lst[t];
                    v += gld(Dex + (size_t)n * Nps + p) * gld(b64 + (size_t)n * npos[ax] + pos);
                }
            } else {
                const float* b32 = (const float*)bsrc;
#pragma unroll 8
                for (int t = 0; t < cnt; ++t) {
                    const int n = lst[t];
                    v += gld(Dex + (size_t)n * Nps + p) * (double)gld(b32 + (size_t)n * npos[ax] + pos);
                }
            }
            double* out = ax == 0 ? d.bout_x : (ax == 1 ? d.bout_y : d.bout_z);
            const double* zin = ax == 0 ? d.bzin_x : (ax == 1 ? d.bzin_y : d.bzin_z);
            gst(out + j, zin ? gld(zin + j) - v : v);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Incoming face fluxes (the boundary branches of sn_sweep_{1,2,3}D.jl, e.g. sn_sweep_2D.jl:70-99):
// a direction entering through a face receives sum_p M_surf[n, p] * (the face's surface-source
// moments, in inhomogeneous passes) and then, for a non-void condition, sum_p M_surf[n, p] * (exit
// half-range moments of the input state on the same face for reflective, on the opposite face for
// periodic), each summed in moment order.
// One thread per (axis, direction, position).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bndin_kernel(const PassDesc* __restrict__ pds) {
    const PassDesc& d = pds[0];
    const int Nd = d.Nd, Nps = d.Nps;
    // positions times face moments per axis (the moments of a general 2D closure; 1 otherwise)
    const int fm[3] = {d.fm_x > 1 ? d.fm_x : 1, d.fm_y > 1 ? d.fm_y : 1, d.fm_z > 1 ? d.fm_z : 1};
    const size_t np[3] = {(size_t)d.Ny * d.Nz, (size_t)d.Nx * d.Nz, (size_t)d.Nx * d.Ny};
    size_t tot = 0;
    for (int ax = 0; ax < d.ndims; ++ax) tot += (size_t)Nd * np[ax] * fm[ax];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        size_t j = i;
        int ax = 0;
        while (j >= (size_t)Nd * np[ax] * fm[ax]) { j -= (size_t)Nd * np[ax] * fm[ax]; ++ax; }
        const int n = (int)(j / (np[ax] * fm[ax]));
        const size_t pm = j - (size_t)n * np[ax] * fm[ax];
        const size_t pos = pm / fm[ax];
        const int m = (int)(pm - pos * fm[ax]);
        const int* insrc = ax == 0 ? d.insrc_x : (ax == 1 ? d.insrc_y : d.insrc_z);
        const double* Me = ax == 0 ? d.Mentry_x : (ax == 1 ? d.Mentry_y : d.Mentry_z);
        const double* bs = ax == 0 ? d.bsin_x : (ax == 1 ? d.bsin_y : d.bsin_z);
        double* out = ax == 0 ? d.bin_x : (ax == 1 ? d.bin_y : d.bin_z);
        const int src = insrc[n];
        const double* ss = ax == 0 ? d.ssrc_x : (ax == 1 ? d.ssrc_y : d.ssrc_z);
        const int* sf = ax == 0 ? d.sface_x : (ax == 1 ? d.sface_y : d.sface_z);
        double v = 0.0;
        if (ss && sf[n] >= 0 && m == 0) {  // surface source of the entry face (first face moment)
            const double* q = ss + (size_t)sf[n] * d.ssrc_fs[ax] + pos;
            for (int p = 0; p < Nps; ++p) v += Me[(size_t)n * Nps + p] * q[(size_t)p * np[ax]];
        }
        if (src >= 0 && bs)
            for (int p = 0; p < Nps; ++p)
                v += Me[(size_t)n * Nps + p] * bs[p + (size_t)Nps * (m + (size_t)fm[ax] * (src + 2 * pos))];
        out[j] = v;
    }
}

// ------------------------------------------------------------------------------------------
// Continuous slowing down with the DD scheme of order 1 in energy (flux_{1,2,3}D_BFP of
// sn_{1,2,3}D_bfp.jl with O = 1): the cell solve gains 2 S⁺ in its removal (folded into the
// group's Σt table by the host) and (S⁺ + S⁻) psiE_in in its source, which enters the angular
// source of the sweep here; the energy face leaves with psiE_out = 2 psi - psiE_in.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void csd_source_kernel(const PassDesc* __restrict__ pds, long n) {
    const PassDesc& d = pds[0];
    double* psi = (double*)d.psi;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long c = i % d.Nc;
        psi[i] += d.csdK[d.mat[c]] * d.psiE_in[i];
    }
}

__global__ __launch_bounds__(256) void csd_out_kernel(const PassDesc* __restrict__ pds, long n) {
    const PassDesc& d = pds[0];
    const double* psi = (const double*)d.psi;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d.psiE_out[i] = 2.0 * psi[i] - d.psiE_in[i];
}

// Angular Fokker-Planck source (fokker_planck_source.jl:26-36) added to the out-of-group source
// before the moment -> direction contraction: one thread per (moment k, cell c), cells fastest.
__global__ __launch_bounds__(256) void fp_source_kernel(const PassDesc* __restrict__ pds, long n) {
    const PassDesc& d = pds[0];
    const int K = d.Nact, Nc = d.Nc;
    const int* __restrict__ rows = d.phi_rows;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / Nc), c = (int)(i - (long)k * Nc);
        const double* __restrict__ Mk = d.fpM + (size_t)k * K;
        double s = 0.0;
        for (int j = 0; j < K; ++j) s += Mk[j] * d.phi_in[(size_t)(rows ? rows[j] : j) * Nc + c];
        d.fpq[i] = (d.fp_qin ? d.fp_qin[i] : 0.0) + d.fpT[d.mat[c]] * s;
    }
}

// ------------------------------------------------------------------------------------------
// 1D sweep with K spatial Legendre moments per cell and a general closure (sn_sweep_1D.jl:51-124
// with flux_1D_BTE, sn_1D_bte.jl:27-64): one wave per direction.  A cell's K x K system
//   S_ij = St d_ij - [i > j] C_i hx sx C_j (1 - (-1)^(i-j)) + C_i hx sx^(i-1) C_j sx^(j-1) w_(j+1)
// is solved for two right-hand sides, the source moments Q and r_j = C_j hx (sx^(j-1) w_0 -
// (-sx)^(j-1)), so that psi = S^-1 Q - f_in S^-1 r and the outgoing face flux
// f_out = w_0 f_in + sum_j C_j sx^(j-1) w_(j+1) psi_j is an affine map a f_in + b of the incoming
// one (a = 1, b = 0 for a tangential direction, which passes the face flux through).  The 64 lanes
// take CPL consecutive cells each, compose their maps, and scan them across the wave; one step
// covers 64 * CPL cells.  Source and flux moments: psi[n][j + K * cell] in the pass's precision.
// ------------------------------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(64) void sweep_1d_ho_kernel(const PassDesc* __restrict__ pds) {
    constexpr int CPL = 4;
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, lane = threadIdx.x;
    const int Nx = d.Nx, Nc = d.Nc;
    double mu = d.mu[n];
    const bool fwd = mu >= 0.0;
    const bool tang = fabs(mu) <= 1e-10;
    const T sx = (tang || mu > 0.0) ? T(1) : T(-1);
    if (tang) mu = 0.0;
    T om[K + 1], Cc[K], sp[K], sn[K];   // w, C_j, sx^(j-1), (-sx)^(j-1)
#pragma unroll
    for (int i = 0; i <= K; ++i) om[i] = (T)d.ho_om[i];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        Cc[i] = (T)sqrt(2.0 * i + 1.0);
        sp[i] = i == 0 ? T(1) : sp[i - 1] * sx;
        sn[i] = i == 0 ? T(1) : -sn[i - 1] * sx;
    }
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    T f = d.bin_x ? (T)d.bin_x[n] : T(0);  // incoming face flux of the current step
    for (int base = 0; base < Nx; base += 64 * CPL) {
        T a[CPL], b[CPL], xq[CPL][K], xr[CPL][K];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int k = base + lane * CPL + c;
            a[c] = T(1);
            b[c] = T(0);
#pragma unroll
            for (int j = 0; j < K; ++j) { xq[c][j] = T(0); xr[c][j] = T(0); }
            if (k < Nx) {
                const int ix = fwd ? k : Nx - 1 - k;
                const T hx = (T)(fabs(mu) / d.dx[ix]);
                const T St = (T)d.sigt[d.mat[(size_t)K * ix]];
                T S[K][K];
#pragma unroll
                for (int i = 0; i < K; ++i) {
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        T e = i == j ? St : T(0);
                        if (i > j && ((i - j) & 1)) e -= T(2) * Cc[i] * hx * sx * Cc[j];
                        e += Cc[i] * hx * sp[i] * Cc[j] * sp[j] * om[j + 1];
                        S[i][j] = e;
                    }
                    xq[c][i] = psi[i + (size_t)K * ix];
                    xr[c][i] = Cc[i] * hx * (sp[i] * om[0] - sn[i]);
                }
                // Gaussian elimination with partial pivoting (row swaps as selects)
#pragma unroll
                for (int p = 0; p < K; ++p) {
#pragma unroll
                    for (int i = p + 1; i < K; ++i) {
                        const bool sw = fabs(S[i][p]) > fabs(S[p][p]);
#pragma unroll
                        for (int j = 0; j < K; ++j) {
                            const T t = S[p][j];
                            S[p][j] = sw ? S[i][j] : t;
                            S[i][j] = sw ? t : S[i][j];
                        }
                        const T tq = xq[c][p], tr = xr[c][p];
                        xq[c][p] = sw ? xq[c][i] : tq;
                        xq[c][i] = sw ? tq : xq[c][i];
                        xr[c][p] = sw ? xr[c][i] : tr;
                        xr[c][i] = sw ? tr : xr[c][i];
                    }
#pragma unroll
                    for (int i = p + 1; i < K; ++i) {
                        const T fct = S[i][p] / S[p][p];
#pragma unroll
                        for (int j = p + 1; j < K; ++j) S[i][j] -= fct * S[p][j];
                        xq[c][i] -= fct * xq[c][p];
                        xr[c][i] -= fct * xr[c][p];
                    }
                }
#pragma unroll
                for (int i = K - 1; i >= 0; --i) {
                    T vq = xq[c][i], vr = xr[c][i];
#pragma unroll
                    for (int j = i + 1; j < K; ++j) { vq -= S[i][j] * xq[c][j]; vr -= S[i][j] * xr[c][j]; }
                    xq[c][i] = vq / S[i][i];
                    xr[c][i] = vr / S[i][i];
                }
                if (!tang) {
                    T ca = om[0], cb = T(0);
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const T cj = Cc[j] * sp[j] * om[j + 1];
                        ca -= cj * xr[c][j];
                        cb += cj * xq[c][j];
                    }
                    a[c] = ca;
                    b[c] = cb;
                }
            }
        }
        // the lane's map over its cells, then an inclusive scan of the maps across the wave
        T A = a[0], B = b[0];
#pragma unroll
        for (int c = 1; c < CPL; ++c) { B = a[c] * B + b[c]; A = a[c] * A; }
#pragma unroll
        for (int dl = 1; dl < 64; dl <<= 1) {
            const T Ap = __shfl_up(A, dl), Bp = __shfl_up(B, dl);
            if (lane >= dl) { B = A * Bp + B; A = A * Ap; }
        }
        T Ae = __shfl_up(A, 1), Be = __shfl_up(B, 1);
        if (lane == 0) { Ae = T(1); Be = T(0); }
        const T fnext = __shfl(A, 63) * f + __shfl(B, 63);
        T fin = Ae * f + Be;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int k = base + lane * CPL + c;
            if (k < Nx) {
                const int ix = fwd ? k : Nx - 1 - k;
#pragma unroll
                for (int j = 0; j < K; ++j) psi[j + (size_t)K * ix] = xq[c][j] - xr[c][j] * fin;
                fin = a[c] * fin + b[c];
            }
        }
        f = fnext;
    }
    if (lane == 0 && d.bx) ((T*)d.bx)[n] = f;  // exit face flux (one position in 1D)
}

// ------------------------------------------------------------------------------------------
// 2D sweep with a general closure of orders KX x KY (sn_sweep_2D.jl:54-200 with flux_2D_BTE,
// sn_2D_bte.jl:27-116): one workgroup per direction walks the anti-diagonals of the cell grid in
// sweep order; the cells of a diagonal are independent, one per thread.  A cell reads the x-face
// moments of its row (FX[iy][KY]) and the y-face moments of its column (FY[ix][KX]) from LDS,
// solves its NM x NM system (NM = KX*KY with full coupling, KX + KY - 1 without; basis (ix, iy)
// at KX*iy + ix, or ix for iy = 0 and KX - 1 + iy for ix = 0)
//   S_ij = St d_ij - [iy=jy, ix>jx] C_ix hx sx C_jx (1 - (-1)^(ix-jx)) + [iy=jy] C_ix hx sx^ix C_jx sx^jx wx_(jx+1)
//          - [ix=jx, iy>jy] C_iy hy sy C_jy (1 - (-1)^(iy-jy)) + [ix=jx] C_iy hy sy^iy C_jy sy^jy wy_(jy+1)
//   Q_j = Qn_j - C_jx hx (sx^jx wx_0 - (-sx)^jx) fx_jy - C_jy hy (sy^jy wy_0 - (-sy)^jy) fy_jx
// (0-based orders, C_k = sqrt(2k+1)), and writes back its outgoing face moments
//   fx_jy <- wx_0 fx_jy + sum_jx C_jx sx^jx wx_(jx+1) psi_(jx,jy),  fy_jx likewise.
// Face inflows come from bin_x / bin_y ([Nd][pos][moment]); the exit faces leave in bx / by.
// ------------------------------------------------------------------------------------------
template <int KX, int KY, bool FC>
__device__ __forceinline__ int ho2_idx(int ix, int iy) {
    if constexpr (FC) return KX * iy + ix;
    return iy > 0 ? KX - 1 + iy : ix;
}

template <typename T, int KX, int KY, bool FC>
__global__ __launch_bounds__(256) void sweep_2d_ho_kernel(const PassDesc* __restrict__ pds) {
    constexpr int NM = FC ? KX * KY : KX + KY - 1;
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny, Nc = d.Nc;
    T* FX = (T*)smem;                  // [Ny][KY]
    T* FY = FX + (size_t)Ny * KY;      // [Nx][KX]
    const double mu = d.mu[n], eta = d.eta[n];
    const bool fxs = mu >= 0, fys = eta >= 0;
    const T sx = mu > 0 ? T(1) : (mu < 0 ? T(-1) : T(0)), sy = eta > 0 ? T(1) : (eta < 0 ? T(-1) : T(0));
    T wx[KX + 1], wy[KY + 1], Cx[KX], Cy[KY], px[KX], nx[KX], py[KY], ny[KY];
#pragma unroll
    for (int k = 0; k <= KX; ++k) wx[k] = (T)d.ho_om[k];
#pragma unroll
    for (int k = 0; k <= KY; ++k) wy[k] = (T)d.ho_omy[k];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        Cx[k] = (T)sqrt(2.0 * k + 1.0);
        px[k] = k == 0 ? T(1) : px[k - 1] * sx;
        nx[k] = k == 0 ? T(1) : -nx[k - 1] * sx;
    }
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        Cy[k] = (T)sqrt(2.0 * k + 1.0);
        py[k] = k == 0 ? T(1) : py[k - 1] * sy;
        ny[k] = k == 0 ? T(1) : -ny[k - 1] * sy;
    }
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    for (int i = tid; i < Ny * KY; i += nt) FX[i] = d.bin_x ? (T)d.bin_x[(size_t)n * Ny * KY + i] : T(0);
    for (int i = tid; i < Nx * KX; i += nt) FY[i] = d.bin_y ? (T)d.bin_y[(size_t)n * Nx * KX + i] : T(0);
    __syncthreads();
    for (int dg = 0; dg < Nx + Ny - 1; ++dg) {
        const int k0 = dg - (Ny - 1) > 0 ? dg - (Ny - 1) : 0, k1 = dg < Nx - 1 ? dg : Nx - 1;
        for (int kx = k0 + tid; kx <= k1; kx += nt) {
            const int ky = dg - kx;
            const int ix = fxs ? kx : Nx - 1 - kx, iy = fys ? ky : Ny - 1 - ky;
            const size_t c = (size_t)ix + (size_t)Nx * iy;
            const T hx = (T)(fabs(mu) / d.dx[ix]), hy = (T)(fabs(eta) / d.dy[iy]);
            const T St = (T)d.sigt[d.mat[c * NM]];
            T* fx = FX + (size_t)iy * KY;
            T* fy = FY + (size_t)ix * KX;
            T S[NM][NM], q[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
#pragma unroll
                for (int j = 0; j < NM; ++j) S[i][j] = i == j ? St : T(0);
#pragma unroll
            for (int ax = 0; ax < KX; ++ax)
#pragma unroll
                for (int ay = 0; ay < KY; ++ay) {
                    if (!FC && ax > 0 && ay > 0) continue;
                    const int i = ho2_idx<KX, KY, FC>(ax, ay);
#pragma unroll
                    for (int bx = 0; bx < KX; ++bx)
#pragma unroll
                        for (int by = 0; by < KY; ++by) {
                            if (!FC && bx > 0 && by > 0) continue;
                            const int j = ho2_idx<KX, KY, FC>(bx, by);
                            if (ay == by) {
                                if (ax > bx && ((ax - bx) & 1)) S[i][j] -= T(2) * Cx[ax] * hx * sx * Cx[bx];
                                S[i][j] += Cx[ax] * hx * px[ax] * Cx[bx] * px[bx] * wx[bx + 1];
                            }
                            if (ax == bx) {
                                if (ay > by && ((ay - by) & 1)) S[i][j] -= T(2) * Cy[ay] * hy * sy * Cy[by];
                                S[i][j] += Cy[ay] * hy * py[ay] * Cy[by] * py[by] * wy[by + 1];
                            }
                        }
                    q[i] = psi[(size_t)i + NM * c] - Cx[ax] * hx * (px[ax] * wx[0] - nx[ax]) * fx[ay] -
                           Cy[ay] * hy * (py[ay] * wy[0] - ny[ay]) * fy[ax];
                }
            // Gaussian elimination with partial pivoting (row swaps as selects)
#pragma unroll
            for (int p = 0; p < NM; ++p) {
#pragma unroll
                for (int i = p + 1; i < NM; ++i) {
                    const bool sw = fabs(S[i][p]) > fabs(S[p][p]);
#pragma unroll
                    for (int j = 0; j < NM; ++j) {
                        const T t = S[p][j];
                        S[p][j] = sw ? S[i][j] : t;
                        S[i][j] = sw ? t : S[i][j];
                    }
                    const T tq = q[p];
                    q[p] = sw ? q[i] : tq;
                    q[i] = sw ? tq : q[i];
                }
#pragma unroll
                for (int i = p + 1; i < NM; ++i) {
                    const T f = S[i][p] / S[p][p];
#pragma unroll
                    for (int j = p + 1; j < NM; ++j) S[i][j] -= f * S[p][j];
                    q[i] -= f * q[p];
                }
            }
#pragma unroll
            for (int i = NM - 1; i >= 0; --i) {
                T v = q[i];
#pragma unroll
                for (int j = i + 1; j < NM; ++j) v -= S[i][j] * q[j];
                q[i] = v / S[i][i];
            }
#pragma unroll
            for (int i = 0; i < NM; ++i) psi[(size_t)i + NM * c] = q[i];
            // outgoing face moments
            T ox[KY], oy[KX];
#pragma unroll
            for (int ay = 0; ay < KY; ++ay) ox[ay] = wx[0] * fx[ay];
#pragma unroll
            for (int ax = 0; ax < KX; ++ax) oy[ax] = wy[0] * fy[ax];
#pragma unroll
            for (int ax = 0; ax < KX; ++ax)
#pragma unroll
                for (int ay = 0; ay < KY; ++ay) {
                    if (!FC && ax > 0 && ay > 0) continue;
                    const T v = q[ho2_idx<KX, KY, FC>(ax, ay)];
                    ox[ay] += Cx[ax] * px[ax] * wx[ax + 1] * v;
                    oy[ax] += Cy[ay] * py[ay] * wy[ay + 1] * v;
                }
#pragma unroll
            for (int ay = 0; ay < KY; ++ay) fx[ay] = ox[ay];
#pragma unroll
            for (int ax = 0; ax < KX; ++ax) fy[ax] = oy[ax];
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * KY; i += nt) ((T*)d.bx)[(size_t)n * Ny * KY + i] = FX[i];
    if (d.by)
        for (int i = tid; i < Nx * KX; i += nt) ((T*)d.by)[(size_t)n * Nx * KX + i] = FY[i];
}

// ------------------------------------------------------------------------------------------
// 3D sweep with a general closure of orders KX x KY x KZ (sn_sweep_3D.jl:57-240 with flux_3D_BTE,
// sn_3D_bte.jl:34-152): one workgroup per direction walks the diagonal planes kx + ky + kz = d in
// sweep order, one cell per thread.  Cell moments (a, b, c): KX*KY*c + KX*b + a with full coupling;
// without it only one of a, b, c is non-zero, at a + b + c (+ KX-1 when b > 0, + KX-1+KY-1 when
// c > 0).  The x, y, z faces carry the moments over (b, c), (a, c), (a, b) with the same rule, held
// in LDS per pencil: FX[Ny*Nz][F0], FY[Nx*Nz][F1], FZ[Nx*Ny][F2].  The cell system and the face
// updates are those of the 2D kernel with a third axis; an axis's terms couple moments that agree on
// the other two axes.
// ------------------------------------------------------------------------------------------
template <int KX, int KY, int KZ, bool FC>
struct Ho3 {
    static constexpr int NM = FC ? KX * KY * KZ : KX + KY + KZ - 2;
    static constexpr int F0 = FC ? KY * KZ : KY + KZ - 1, F1 = FC ? KX * KZ : KX + KZ - 1, F2 = FC ? KX * KY : KX + KY - 1;
    __device__ static constexpr bool valid(int a, int b, int c) { return FC || (a > 0) + (b > 0) + (c > 0) <= 1; }
    __device__ static constexpr int idx(int a, int b, int c) {
        return FC ? KX * KY * c + KX * b + a : a + b + c + (b > 0 ? KX - 1 : 0) + (c > 0 ? KX - 1 + KY - 1 : 0);
    }
    __device__ static constexpr int fx(int b, int c) { return FC ? KY * c + b : b + c + (c > 0 ? KY - 1 : 0); }
    __device__ static constexpr int fy(int a, int c) { return FC ? KX * c + a : a + c + (c > 0 ? KX - 1 : 0); }
    __device__ static constexpr int fz(int a, int b) { return FC ? KX * b + a : a + b + (b > 0 ? KX - 1 : 0); }
};

template <typename T, int KX, int KY, int KZ, bool FC>
__global__ __launch_bounds__(256) void sweep_3d_ho_kernel(const PassDesc* __restrict__ pds) {
    using H = Ho3<KX, KY, KZ, FC>;
    constexpr int NM = H::NM, F0 = H::F0, F1 = H::F1, F2 = H::F2;
    constexpr int KM = KX > KY ? (KX > KZ ? KX : KZ) : (KY > KZ ? KY : KZ);
    extern __shared__ __align__(16) unsigned char smem[];
    const PassDesc& d = pds[0];
    const int n = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz, Nc = d.Nc;
    // face moments in LDS, or in this direction's slice of hoface (the same walk: the waves of the
    // workgroup share one CU, so __syncthreads orders their device-memory writes and reads too)
    const size_t nface = (size_t)Ny * Nz * F0 + (size_t)Nx * Nz * F1 + (size_t)Nx * Ny * F2;
    T* FX = d.hoface ? (T*)d.hoface + (size_t)n * nface : (T*)smem;  // [Ny*Nz][F0]
    T* FY = FX + (size_t)Ny * Nz * F0;       // [Nx*Nz][F1]
    T* FZ = FY + (size_t)Nx * Nz * F1;       // [Nx*Ny][F2]
    const double om[3] = {d.mu[n], d.eta[n], d.xi[n]};
    const int K3[3] = {KX, KY, KZ};
    T s[3], w[3][KM + 1], Cc[KM], sp[3][KM], sn[3][KM];
    const double* ws[3] = {d.ho_om, d.ho_omy, d.ho_omz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a] = om[a] > 0 ? T(1) : (om[a] < 0 ? T(-1) : T(0));
#pragma unroll
        for (int k = 0; k <= KM; ++k) w[a][k] = k <= K3[a] ? (T)ws[a][k] : T(0);
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            sp[a][k] = k == 0 ? T(1) : sp[a][k - 1] * s[a];
            sn[a][k] = k == 0 ? T(1) : -sn[a][k - 1] * s[a];
        }
    }
#pragma unroll
    for (int k = 0; k < KM; ++k) Cc[k] = (T)sqrt(2.0 * k + 1.0);
    T* __restrict__ psi = (T*)d.psi + (size_t)n * Nc;
    for (int i = tid; i < Ny * Nz * F0; i += nt) FX[i] = d.bin_x ? (T)d.bin_x[(size_t)n * Ny * Nz * F0 + i] : T(0);
    for (int i = tid; i < Nx * Nz * F1; i += nt) FY[i] = d.bin_y ? (T)d.bin_y[(size_t)n * Nx * Nz * F1 + i] : T(0);
    for (int i = tid; i < Nx * Ny * F2; i += nt) FZ[i] = d.bin_z ? (T)d.bin_z[(size_t)n * Nx * Ny * F2 + i] : T(0);
    __syncthreads();
    const bool fw[3] = {om[0] >= 0, om[1] >= 0, om[2] >= 0};
    for (int dg = 0; dg < Nx + Ny + Nz - 2; ++dg) {
        for (int i = tid; i < Nx * Ny; i += nt) {
            const int kx = i % Nx, ky = i / Nx, kz = dg - kx - ky;
            if (kz < 0 || kz >= Nz) continue;
            const int ix = fw[0] ? kx : Nx - 1 - kx, iy = fw[1] ? ky : Ny - 1 - ky, iz = fw[2] ? kz : Nz - 1 - kz;
            const size_t c = (size_t)ix + (size_t)Nx * (iy + (size_t)Ny * iz);
            const T h[3] = {(T)(fabs(om[0]) / d.dx[ix]), (T)(fabs(om[1]) / d.dy[iy]), (T)(fabs(om[2]) / d.dz[iz])};
            const T St = (T)d.sigt[d.mat[c * NM]];
            T* fx = FX + ((size_t)iy + (size_t)Ny * iz) * F0;
            T* fy = FY + ((size_t)ix + (size_t)Nx * iz) * F1;
            T* fz = FZ + ((size_t)ix + (size_t)Nx * iy) * F2;
            T S[NM][NM], q[NM];
#pragma unroll
            for (int i2 = 0; i2 < NM; ++i2)
#pragma unroll
                for (int j = 0; j < NM; ++j) S[i2][j] = i2 == j ? St : T(0);
#pragma unroll
            for (int a = 0; a < KX; ++a)
#pragma unroll
                for (int b = 0; b < KY; ++b)
#pragma unroll
                    for (int cz = 0; cz < KZ; ++cz) {
                        if (!H::valid(a, b, cz)) continue;
                        const int ii = H::idx(a, b, cz);
                        const int ia[3] = {a, b, cz};
#pragma unroll
                        for (int a2 = 0; a2 < KX; ++a2)
#pragma unroll
                            for (int b2 = 0; b2 < KY; ++b2)
#pragma unroll
                                for (int c2 = 0; c2 < KZ; ++c2) {
                                    if (!H::valid(a2, b2, c2)) continue;
                                    const int jj = H::idx(a2, b2, c2);
                                    const int jb[3] = {a2, b2, c2};
#pragma unroll
                                    for (int ax = 0; ax < 3; ++ax) {
                                        const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                                        if (ia[o1] != jb[o1] || ia[o2] != jb[o2]) continue;
                                        const int u = ia[ax], v = jb[ax];
                                        if (u > v && ((u - v) & 1)) S[ii][jj] -= T(2) * Cc[u] * h[ax] * s[ax] * Cc[v];
                                        S[ii][jj] += Cc[u] * h[ax] * sp[ax][u] * Cc[v] * sp[ax][v] * w[ax][v + 1];
                                    }
                                }
                        q[ii] = psi[(size_t)ii + NM * c] - Cc[a] * h[0] * (sp[0][a] * w[0][0] - sn[0][a]) * fx[H::fx(b, cz)] -
                                Cc[b] * h[1] * (sp[1][b] * w[1][0] - sn[1][b]) * fy[H::fy(a, cz)] -
                                Cc[cz] * h[2] * (sp[2][cz] * w[2][0] - sn[2][cz]) * fz[H::fz(a, b)];
                    }
#pragma unroll
            for (int p = 0; p < NM; ++p) {
#pragma unroll
                for (int i2 = p + 1; i2 < NM; ++i2) {
                    const bool sw = fabs(S[i2][p]) > fabs(S[p][p]);
#pragma unroll
                    for (int j = 0; j < NM; ++j) {
                        const T t = S[p][j];
                        S[p][j] = sw ? S[i2][j] : t;
                        S[i2][j] = sw ? t : S[i2][j];
                    }
                    const T tq = q[p];
                    q[p] = sw ? q[i2] : tq;
                    q[i2] = sw ? tq : q[i2];
                }
#pragma unroll
                for (int i2 = p + 1; i2 < NM; ++i2) {
                    const T f = S[i2][p] / S[p][p];
#pragma unroll
                    for (int j = p + 1; j < NM; ++j) S[i2][j] -= f * S[p][j];
                    q[i2] -= f * q[p];
                }
            }
#pragma unroll
            for (int i2 = NM - 1; i2 >= 0; --i2) {
                T v = q[i2];
#pragma unroll
                for (int j = i2 + 1; j < NM; ++j) v -= S[i2][j] * q[j];
                q[i2] = v / S[i2][i2];
            }
#pragma unroll
            for (int i2 = 0; i2 < NM; ++i2) psi[(size_t)i2 + NM * c] = q[i2];
            T ox[F0], oy[F1], oz[F2];
#pragma unroll
            for (int k = 0; k < F0; ++k) ox[k] = w[0][0] * fx[k];
#pragma unroll
            for (int k = 0; k < F1; ++k) oy[k] = w[1][0] * fy[k];
#pragma unroll
            for (int k = 0; k < F2; ++k) oz[k] = w[2][0] * fz[k];
#pragma unroll
            for (int a = 0; a < KX; ++a)
#pragma unroll
                for (int b = 0; b < KY; ++b)
#pragma unroll
                    for (int cz = 0; cz < KZ; ++cz) {
                        if (!H::valid(a, b, cz)) continue;
                        const T v = q[H::idx(a, b, cz)];
                        ox[H::fx(b, cz)] += Cc[a] * sp[0][a] * w[0][a + 1] * v;
                        oy[H::fy(a, cz)] += Cc[b] * sp[1][b] * w[1][b + 1] * v;
                        oz[H::fz(a, b)] += Cc[cz] * sp[2][cz] * w[2][cz + 1] * v;
                    }
#pragma unroll
            for (int k = 0; k < F0; ++k) fx[k] = ox[k];
#pragma unroll
            for (int k = 0; k < F1; ++k) fy[k] = oy[k];
#pragma unroll
            for (int k = 0; k < F2; ++k) fz[k] = oz[k];
        }
        __syncthreads();
    }
    if (d.bx)
        for (int i = tid; i < Ny * Nz * F0; i += nt) ((T*)d.bx)[(size_t)n * Ny * Nz * F0 + i] = FX[i];
    if (d.by)
        for (int i = tid; i < Nx * Nz * F1; i += nt) ((T*)d.by)[(size_t)n * Nx * Nz * F1 + i] = FY[i];
    if (d.bz)
        for (int i = tid; i < Nx * Ny * F2; i += nt) ((T*)d.bz)[(size_t)n * Nx * Ny * F2 + i] = FZ[i];
}

// ------------------------------------------------------------------------------------------
// Out-of-group + volume source (scattering_source.jl:59-68, sn_flux.jl:179).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void outsrc_kernel(double* __restrict__ qlout, const double* __restrict__ phi_all,
                                                     const double* __restrict__ sigo, const double* __restrict__ qvol,
                                                     const int* __restrict__ mat, const int* __restrict__ act, int g,
                                                     int G, int Nmat, int Nact, int Np, int Nc) {
    const size_t tot = (size_t)Nact * Nc;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(i / Nc);
        const int c = (int)(i % Nc);
        const int m = mat[c];
        const int p = act[a];
        double q = 0.0;
        for (int gi = 0; gi < g; ++gi)
            q += sigo[(((size_t)g * G + gi) * Nmat + m) * Nact + a] * phi_all[((size_t)gi * Np + p) * Nc + c];
        if (p == 0) q += qvol[(size_t)g * Nc + c];
        qlout[i] = q;
    }
}

// ------------------------------------------------------------------------------------------
// BLAS-1 with deterministic two-level reductions: a fixed grid writes one partial per block
// (fixed tree inside the block), then a one-block kernel sums the partials in a fixed order.
// A last-block ticket instead would need a device-scope release fence in every block, which on
// this GPU writes back the XCD's L2 each time -- far costlier than the second launch.
// ------------------------------------------------------------------------------------------
__device__ inline double block_sum(double v) {
    __shared__ double sh[RED_THREADS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(RED_THREADS) void reduce_finish_kernel(const double* __restrict__ partial, int np,
                                                                    double* out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += partial[i];
    v = block_sum(v);
    if (threadIdx.x == 0) *out = v;
}

__global__ __launch_bounds__(RED_THREADS) void dot_kernel(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                                                          double* partial) {
    double v = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v += gld(a + i) * gld(b + i);
    v = block_sum(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// w -= h v, then the partial of w . vnext (NEXT) or w . w
template <bool NEXT>
__global__ __launch_bounds__(RED_THREADS) void mgs_kernel(double* __restrict__ w, const double* __restrict__ v,
                                                          const double* __restrict__ vnext, size_t n, const double* hptr,
                                                          double* partial) {
    const double h = *hptr;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double x = gld(w + i) - h * gld(v + i);
        gst(w + i, x);
        acc += x * (NEXT ? gld(vnext + i) : x);
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void fill_kernel(double* x, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void copy_kernel(double* __restrict__ d, const double* __restrict__ s, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}
__global__ void scale_kernel(double* x, double a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] *= a;
}
__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += a * x[i];
}
__global__ void neg_add_kernel(double* __restrict__ r, const double* __restrict__ c, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) r[i] = -r[i] + c[i];
}
__global__ void sub_kernel(double* out, const double* a, const double* b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) out[i] = a[i] - b[i];
}
__global__ void scale_inv_slot_kernel(double* __restrict__ d, const double* __restrict__ s, const double* scal, int slot, size_t n) {
    const double h = sqrt(scal[slot]);  // the slot holds a squared norm
    if (!(h > 0.0)) return;
    const double a = 1.0 / h;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i] * a;
}
__global__ void lincomb_kernel(double* __restrict__ z, const double* const* __restrict__ V, const double* __restrict__ y, int k,
                               size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = z[i];
        for (int t = 0; t < k; ++t) v += y[t] * V[t][i];
        z[i] = v;
    }
}
__global__ void cast_kernel(float* __restrict__ d, const double* __restrict__ s, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = (float)s[i];
}
__global__ void scatter_rows_kernel(double* __restrict__ dst, const double* __restrict__ src, int rows, const int* __restrict__ map,
                                    int Nc) {
    const size_t tot = (size_t)rows * Nc;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / Nc), c = (int)(i % Nc);
        dst[(size_t)map[r] * Nc + c] = src[i];
    }
}
__global__ void set_pd_kernel(PassDesc v, PassDesc* dst) { *dst = v; }
__global__ void set_vals_kernel(Vals32 v, double* dst, int k) {
    if ((int)threadIdx.x < k) dst[threadIdx.x] = v.v[threadIdx.x];
}
// phi_all [g][p][c] -> Radiant 𝚽l[g + G*(p + Np*c)]
// dst[g + G (p + Np (c - c0))] = src[(g Np + p) Nc + c] for the cells c0 <= c < c0 + nc
__global__ void phi_to_radiant_kernel(double* __restrict__ dst, const double* __restrict__ src, int G, int Np, int Nc,
                                      int c0, int nc) {
    const size_t tot = (size_t)G * Np * nc;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t g = i % G, r = i / G;
        const size_t p = r % Np, c = c0 + r / Np;
        dst[i] = src[(g * Np + p) * Nc + c];
    }
}
// Flux at the cutoff energy (sn_flux.jl:213-218): out[p + Np*c] = sum_n D[p, n] psiE[n][c], the
// directions summed in increasing order; DT is the k-major operator DT[n][p] = D[p, n].
__global__ void cutoff_kernel(double* __restrict__ out, const double* __restrict__ DT, const double* __restrict__ psiE,
                              int Np, int Nd, int Ncell, int K, int FE) {
    // out[p + Np*(is + K*cell)] = sum_n D[p, n] psiE[n][cell][is] for is < FE, zero above
    const size_t tot = (size_t)Np * Ncell * K;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t p = i % Np, col = i / Np, cell = col / K, is = col % K;
        double acc = 0.0;
        if (is < (size_t)FE)
            for (int n = 0; n < Nd; ++n)
                acc = fma(DT[(size_t)n * Np + p], psiE[((size_t)n * Ncell + cell) * FE + is], acc);
        out[i] = acc;
    }
}
__global__ void transpose_kernel(double* __restrict__ dst, const double* __restrict__ src, int R, int C) {
    // dst[c*R + r] = src[r*C + c]
    const size_t tot = (size_t)R * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i % R, c = i / R;
        dst[i] = src[r * C + c];
    }
}

// ------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------
static inline int grid_for(size_t n, int block = 256, int cap = 2048) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

template <bool MG, int NS>
static void contract_launch(int prec, const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st) {
    if (prec == PMIX) launch_timed(contract_mixed_kernel<GEMM_RT, GEMM_CT, MG, NS>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    else if (prec == P64) launch_timed(contract_kernel<double, GEMM_RT, GEMM_CT, MG, NS>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    else launch_timed(contract_kernel<float, GEMM_RT, GEMM_CT, MG, NS>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
}

// Pipeline depth (k steps in flight) of the contractions: the default first, then the autotuner's
// alternative (same accumulation order, so the same bits); BF16 operands keep their one shape.
void launch_mgemm(int prec, const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st) {
    if (ntasks <= 0) return;
    if (prec == PBF16) {
        launch_timed(contract_bf16_kernel<GEMM_RT, GEMM_CT, true>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    } else {
        TuneTicket t;
        const int v = tune_choose({prec, 2, 0, 0, 0, ilog2(ntasks)}, 2, t);
        if (t.a) set_launch_events(t.a, t.b);
        if (v == 0) contract_launch<true, 2>(prec, d_pd, tasks, ntasks, st);
        else contract_launch<true, 3>(prec, d_pd, tasks, ntasks, st);
        tune_commit(t);
    }
    RT_HIP(hipGetLastError());
}

void launch_dgemm_add(const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st) {
    if (ntasks <= 0) return;
    TuneTicket t;
    const int v = tune_choose({P64, 4, 0, 0, 0, ilog2(ntasks)}, 2, t);
    if (t.a) set_launch_events(t.a, t.b);
    if (v == 0) launch_timed(contract_kernel<double, GEMM_RT, GEMM_CT, false, 4, true>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    else launch_timed(contract_kernel<double, GEMM_RT, GEMM_CT, false, 2, true>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    tune_commit(t);
    RT_HIP(hipGetLastError());
}

void launch_dgemm(int prec, const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st) {
    if (ntasks <= 0) return;
    if (prec == PBF16) {
        launch_timed(contract_bf16_kernel<GEMM_RT, GEMM_CT, false>, dim3(ntasks), dim3(256), 0, st, d_pd, tasks);
    } else {
        TuneTicket t;
        const int v = tune_choose({prec, 3, 0, 0, 0, ilog2(ntasks)}, 2, t);
        if (t.a) set_launch_events(t.a, t.b);
        if (v == 0) contract_launch<false, 4>(prec, d_pd, tasks, ntasks, st);
        else contract_launch<false, 2>(prec, d_pd, tasks, ntasks, st);
        tune_commit(t);
    }
    RT_HIP(hipGetLastError());
}

int sweep_pencils_per_thread(int P) {
    for (int ppt : {1, 2, 4, 8, 16})
        if (P <= 1024 * ppt) return ppt;
    return -1;
}

size_t sweep_lds_bytes(int prec, int Nx, int Ny, int Nz, int Nmat, int nbuf) {
    const size_t es = (prec == P64) ? 8 : 4;
    return es * ((size_t)Nx + Ny + Nz + Nmat + 2ull * nbuf * Ny * Nz);
}

template <typename T, int PPT, int NBUF>
static void sweep_inst(const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int block, size_t lds, hipStream_t st) {
    constexpr int PF = (PPT <= 4) ? 2 : 1;
    auto kfn = sweep_kernel<T, PPT, PF, NBUF>;
    static bool attr_set = false;
    if (!attr_set) {
        RT_HIP(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    launch_timed(kfn, dim3(ntasks), dim3(block), lds, st, d_pd, tasks);
}

template <typename T, int NBUF>
static void sweep_ppt(int ppt, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int block, size_t lds,
                      hipStream_t st) {
    switch (ppt) {
        case 1: sweep_inst<T, 1, NBUF>(d_pd, tasks, ntasks, block, lds, st); break;
        case 2: sweep_inst<T, 2, NBUF>(d_pd, tasks, ntasks, block, lds, st); break;
        case 4: sweep_inst<T, 4, NBUF>(d_pd, tasks, ntasks, block, lds, st); break;
        case 8: sweep_inst<T, 8, NBUF>(d_pd, tasks, ntasks, block, lds, st); break;
        case 16: sweep_inst<T, 16, NBUF>(d_pd, tasks, ntasks, block, lds, st); break;
        default: throw std::runtime_error("sweep: unsupported pencils per thread");
    }
}

template <typename T, int R>
static void sweep_rows_inst(const PassDesc* d_pd, const SweepTask* tasks, int ntasks, size_t lds, hipStream_t st) {
    constexpr int PF = (sizeof(T) == 8 && R != 2) ? 8 : 16;
    auto kfn = sweep_rows_kernel<T, R, PF>;
    static bool attr_set = false;
    if (!attr_set) {
        RT_HIP(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    launch_timed(kfn, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
}

int sweep_rows_per_lane(int ndims, int Ny) {
    if (ndims > 2) return -1;
    for (int r : {1, 2, 4})
        if (Ny <= 64 * r) return r;
    return -1;
}

size_t sweep_rows_lds_bytes(int prec, int Nx, int Nmat) { return (prec == P64 ? 8 : 4) * ((size_t)Nx + Nmat); }

void launch_sweep_rows(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int rows_per_lane,
                       size_t lds, hipStream_t st) {
    if (ntasks <= 0) return;
    if (prec == P64) {
        if (rows_per_lane == 1) sweep_rows_inst<double, 1>(d_pd, tasks, ntasks, lds, st);
        else if (rows_per_lane == 2) sweep_rows_inst<double, 2>(d_pd, tasks, ntasks, lds, st);
        else sweep_rows_inst<double, 4>(d_pd, tasks, ntasks, lds, st);
    } else {
        if (rows_per_lane == 1) sweep_rows_inst<float, 1>(d_pd, tasks, ntasks, lds, st);
        else if (rows_per_lane == 2) sweep_rows_inst<float, 2>(d_pd, tasks, ntasks, lds, st);
        else sweep_rows_inst<float, 4>(d_pd, tasks, ntasks, lds, st);
    }
    RT_HIP(hipGetLastError());
}

bool sweep_planes_shape(int prec, int Ny, int Nz, int& R, int& ZW, int& waves) {
    // pencils per lane K = R*ZW <= 8 with at most 8 waves (256 VGPRs per lane), or, in FP32,
    // K = 16 with up to 16 waves; otherwise the pencil kernel handles the grid.
    R = Ny <= 64 ? 1 : (Ny <= 128 ? 2 : -1);
    if (R < 0) return false;
    for (int zw : {1, 2, 4, 8}) {
        const int w = (Nz + zw - 1) / zw;
        if (w <= 8 && R * zw <= 8) { ZW = zw; waves = w; return true; }
    }
    if (prec != P64) {
        const int w = (Nz + 7) / 8;
        if (w <= 16) { ZW = 8; waves = w; return true; }
    }
    return false;
}

size_t sweep_planes_lds_bytes(int prec, int Nx, int Ny, int Nz, int Nmat, int R, int waves) {
    return (prec == P64 ? 8 : 4) * ((size_t)Nx + Ny + Nz + Nmat + 2ull * waves * 64 * R);
}

template <typename T, int R, int ZW, int WMAX>
static void planes_inst(const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds, hipStream_t st) {
    constexpr int K = R * ZW;
    constexpr bool dbl = sizeof(T) == 8;
    constexpr int PF = K <= 2 ? 8 : (K <= 4 ? 4 : (K <= 8 && !dbl ? 2 : 1));
    auto kfn = sweep_planes_kernel<T, R, ZW, PF, WMAX>;
    static bool attr_set = false;
    if (!attr_set) {
        RT_HIP(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    launch_timed(kfn, dim3(ntasks), dim3(64 * waves), lds, st, d_pd, tasks);
}

template <typename T>
static void planes_dispatch(int R, int ZW, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds,
                            hipStream_t st) {
    if (R == 1) {
        if (ZW == 1) planes_inst<T, 1, 1, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else if (ZW == 2) planes_inst<T, 1, 2, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else if (ZW == 4) planes_inst<T, 1, 4, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else if (waves <= 8) planes_inst<T, 1, 8, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else planes_inst<T, 1, 8, 16>(d_pd, tasks, ntasks, waves, lds, st);
    } else {
        if (ZW == 1) planes_inst<T, 2, 1, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else if (ZW == 2) planes_inst<T, 2, 2, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else if (ZW == 4) planes_inst<T, 2, 4, 8>(d_pd, tasks, ntasks, waves, lds, st);
        else planes_inst<T, 2, 8, 16>(d_pd, tasks, ntasks, waves, lds, st);
    }
}

void launch_sweep_planes(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int R, int ZW, int waves,
                         size_t lds, hipStream_t st) {
    if (ntasks <= 0) return;
    if (prec == P64) planes_dispatch<double>(R, ZW, d_pd, tasks, ntasks, waves, lds, st);
    else planes_dispatch<float>(R, ZW, d_pd, tasks, ntasks, waves, lds, st);
    RT_HIP(hipGetLastError());
}

int sweep_scan_cpl(int Nx) {
    for (int c : {1, 2, 4, 8})
        if (Nx <= 64 * c) return c;
    return -1;
}

// 3D: cells per lane, plus SCAN_PAIR when each half-wave solves its own pencil (32 lanes per
// pencil).  RADIANT_HIP_SCAN_PPW=1 or 2 forces one pencil or two pencils per wave; by default
// grids with Nx <= SCAN_PAIR_NX take two.
static int scan_ppw_env() {
    const char* e = std::getenv("RADIANT_HIP_SCAN_PPW");
    return e ? std::atoi(e) : 0;
}
int sweep_scan_cpl3d(int Nx) {
    const int env = scan_ppw_env();
    const bool pair = env == 2 || (env != 1 && Nx <= SCAN_PAIR_NX);
    if (pair)
        for (int c : {1, 2, 4})
            if (Nx <= 32 * c) return c | SCAN_PAIR;
    return sweep_scan_cpl(Nx);
}

size_t sweep_scan_lds_bytes(int prec, int ndims, int Nx, int Ny, int Nz, int Nmat) {
    const size_t es = prec == P64 ? 8 : 4;
    if (ndims < 3) return es * ((size_t)Ny + Nmat);
    int cpl = sweep_scan_cpl3d(Nx) & (SCAN_PAIR - 1);
    if (cpl <= 0) cpl = 8;
    const size_t rs = (size_t)cpl * ((Nx + cpl - 1) / cpl);  // slots of a face row (sweep_scan_body)
    return es * ((size_t)Ny + Nz + Nmat + ((size_t)Nz + Ny) * rs + 64 * (size_t)cpl);
}

int sweep_scan_waves(int ndims, int Ny, int Nz, int Nx) {
    if (ndims < 3) return 0;              // 1D/2D: the line kernel (one wave per direction)
    const int diag = Ny < Nz ? Ny : Nz;  // longest diagonal of the (y, z) pencil grid
    const int per = (sweep_scan_cpl3d(Nx) & SCAN_PAIR) ? 4 : 2;
    int w = (diag + per - 1) / per;      // two items per wave per diagonal
    return w < 1 ? 1 : (w > 16 ? 16 : w);
}

static bool scan_use_dpp() {
    // RADIANT_HIP_SCAN=shfl selects the ds_bpermute scan (kept for cross-checking the DPP one)
    const char* e = std::getenv("RADIANT_HIP_SCAN");
    return !(e && e[0] == 's');
}

// 3D shapes (pencils per item U, prefetch depth PF, threads per workgroup at most MAXT): shape 0
// (U 1, PF 4) is the default; RADIANT_HIP_SCAN3D=1 selects U 2 with a one-item prefetch, for
// cross-checking.  Measured on 48^3..96^3 grids the two are within 10% of each other (the kernel
// is bound by its instruction issue, not by the latency of its loads).
static int scan3d_shape() {
    const char* e = std::getenv("RADIANT_HIP_SCAN3D");
    return e && e[0] == '1' ? 1 : 0;
}

template <int CPL, int U, int PF, int MAXT, bool DPP, bool BIN = false, int PPW = 1>
static void scan3d_launch(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds,
                          hipStream_t st) {
    waves = waves < MAXT / 64 ? waves : MAXT / 64;
    if constexpr (BIN) {  // FP64 and FP32 passes of one problem
        auto f64 = sweep_scan_kernel<double, CPL, U, PF, MAXT, DPP, true, PPW>;
        auto f32 = sweep_scan_kernel<float, CPL, U, PF, MAXT, DPP, true, PPW>;
        static bool battr[2] = {false, false};
        const int bi = kind == P64 ? 0 : 1;
        if (!battr[bi]) {
            RT_HIP(hipFuncSetAttribute(bi == 0 ? (const void*)f64 : (const void*)f32,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            battr[bi] = true;
        }
        if (bi == 0) launch_timed(f64, dim3(ntasks), dim3(64 * waves), lds, st, d_pd, tasks);
        else launch_timed(f32, dim3(ntasks), dim3(64 * waves), lds, st, d_pd, tasks);
        return;
    }
    const void* fn = kind == PMIX ? (const void*)sweep_scan_mixed_kernel<CPL, U, PF, MAXT, DPP, PPW>
                     : kind == P64 ? (const void*)sweep_scan_kernel<double, CPL, U, PF, MAXT, DPP, false, PPW>
                                   : (const void*)sweep_scan_kernel<float, CPL, U, PF, MAXT, DPP, false, PPW>;
    static bool attr_set[3] = {};
    const int ai = kind == PMIX ? 0 : (kind == P64 ? 1 : 2);
    if (!attr_set[ai]) {
        RT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set[ai] = true;
    }
    if (kind == PMIX)
        launch_timed(sweep_scan_mixed_kernel<CPL, U, PF, MAXT, DPP, PPW>, dim3(ntasks), dim3(64 * waves), lds, st, d_pd,
                     tasks);
    else if (kind == P64)
        launch_timed(sweep_scan_kernel<double, CPL, U, PF, MAXT, DPP, false, PPW>, dim3(ntasks), dim3(64 * waves), lds, st,
                     d_pd, tasks);
    else
        launch_timed(sweep_scan_kernel<float, CPL, U, PF, MAXT, DPP, false, PPW>, dim3(ntasks), dim3(64 * waves), lds, st,
                     d_pd, tasks);
}

// z-split launches: RADIANT_HIP_SCAN_SPLIT=1 never splits, =2 or =4 always splits into that many
// slabs where the grid allows it; unset, the split shapes are autotuner candidates.
static int scan_split_env() {
    const char* e = std::getenv("RADIANT_HIP_SCAN_SPLIT");
    return e ? std::atoi(e) : 0;
}

// Solving waves of a slab workgroup (its courier comes on top), 0 when S slabs do not fit the grid:
// every slab needs a layer, and a handed row must fit the courier's two words per lane.  items:
// items per wave on the slab's longest diagonal (two pencils per item).
int scan_split_waves(const ScanSplitGeom& g, int S, int prec, int items) {
    if (S < 2 || S > SCAN_SPLIT_MAX || g.Nz < 2 * S) return 0;
    if (g.Nx * (prec == P64 ? 2 : 1) > 128) return 0;
    const int Zb = (g.Nz + S - 1) / S;
    if ((S - 1) * Zb >= g.Nz) return 0;
    const int diag = g.Ny < Zb ? g.Ny : Zb;
    const int w = (diag + 2 * items - 1) / (2 * items);
    return w < 1 ? 1 : (w > 15 ? 15 : w);
}

static std::atomic<unsigned> g_split_tag{0};

template <int CPL>
static void scan3d_split_launch(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, const ScanSplitGeom& g,
                                int S, int items, size_t lds, hipStream_t st) {
    const int ws = scan_split_waves(g, S, kind, items);
    const size_t es = kind == P64 ? 8 : 4;
    const int Zb = (g.Nz + S - 1) / S;
    const size_t rs = (size_t)CPL * ((g.Nx + CPL - 1) / CPL);  // slots of a face row
    const size_t lds_s = lds - es * (size_t)(g.Nz - Zb) * rs;    // FY holds the slab's layers only
    unsigned tag = ++g_split_tag;
    if (tag == 0) tag = ++g_split_tag;  // 0 is the value of words never written
    auto f64 = sweep_scan_split_kernel<double, CPL, 4, 2>;
    auto f32 = sweep_scan_split_kernel<float, CPL, 4, 2>;
    static bool attr[2] = {false, false};
    const int ai = kind == P64 ? 0 : 1;
    if (!attr[ai]) {
        // 160 KiB less the kernel's static LDS (the slab role)
        RT_HIP(hipFuncSetAttribute(ai == 0 ? (const void*)f64 : (const void*)f32,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256));
        attr[ai] = true;
    }
    static const int lag = [] {
        const char* e = std::getenv("RADIANT_HIP_SCAN_SPLIT_LAG");
        return e ? std::atoi(e) : SCAN_SPLIT_LAG;
    }();
    if (ai == 0) launch_timed(f64, dim3(ntasks * S), dim3(64 * (ws + 1)), lds_s, st, d_pd, tasks, S, tag, lag);
    else launch_timed(f32, dim3(ntasks * S), dim3(64 * (ws + 1)), lds_s, st, d_pd, tasks, S, tag, lag);
}

// Two pencils per wave (32 lanes each): one item shape, the default waves and half of them, and
// for single-problem passes whose PassDesc carries the split buffers (sg) the z-split into 2 and 4
// slabs per direction.
template <int CPL, bool BIN>
static void scan3d_pair_inst(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds,
                             hipStream_t st, const ScanSplitGeom* sg) {
    const int NW = waves >= 2 ? 2 : 1;
    // split candidates: (slabs, items per wave on the longest diagonal)
    int splits[4], items[4], ns = 0;
    if constexpr (!BIN && CPL <= 2) {
        if (sg && kind != PMIX && scan_split_env() != 1)
            for (int S : {2, 4})
                if (scan_split_waves(*sg, S, kind, 1) > 0)
                    for (int it : {1, 2}) { splits[ns] = S; items[ns] = it; ++ns; }
        const int forced = scan_split_env();
        const char* fi = std::getenv("RADIANT_HIP_SCAN_SPLIT_ITEMS");
        const int fit = fi ? std::atoi(fi) : 1;
        for (int k = 0; k < ns; ++k)
            if (splits[k] == forced && items[k] == fit) {
                scan3d_split_launch<CPL>(kind, d_pd, tasks, ntasks, *sg, forced, fit, lds, st);
                return;
            }
    }
    TuneTicket t;
    const int v = tune_choose({kind, 1, CPL | SCAN_PAIR, waves, (BIN ? 1 : 0) + 2 * (ns / 2), ilog2(ntasks)}, NW + ns, t);
    if (t.a) set_launch_events(t.a, t.b);
    if (v >= NW) {
        if constexpr (!BIN && CPL <= 2)
            scan3d_split_launch<CPL>(kind, d_pd, tasks, ntasks, *sg, splits[v - NW], items[v - NW], lds, st);
    } else {
        const int w = v ? (waves + 1) / 2 : waves;
        constexpr int PF = CPL >= 4 ? 2 : 4;
        scan3d_launch<CPL, 1, PF, 1024, true, BIN, 2>(kind, d_pd, tasks, ntasks, w, lds, st);
    }
    tune_commit(t);
}

template <int CPL, bool DPP, bool BIN>
static void scan3d_inst(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds,
                        hipStream_t st) {
    constexpr int U2 = CPL >= 4 ? 1 : 2;
    // candidates: item shapes (one unless DPP, fewer than 8 cells per lane and no inflow) times
    // waves per workgroup (the default and half of it)
    constexpr int NS = (CPL >= 8 || !DPP || BIN) ? 1 : 2;
    const int NW = waves >= 2 ? 2 : 1;
    TuneTicket t;
    int v = tune_choose({kind, 1, CPL, waves, BIN ? 1 : 0, ilog2(ntasks)}, NS * NW, t);
    if (scan3d_shape() == 1 && NS == 2 && v % 2 == 0) v += 1;  // RADIANT_HIP_SCAN3D=1 forces shape 1
    if (t.a) set_launch_events(t.a, t.b);
    const int shape = v % NS, w = (v / NS) ? (waves + 1) / 2 : waves;
    if constexpr (CPL >= 8 || !DPP) {  // 8 cells per lane: 512 threads, so that the registers do not spill
        scan3d_launch<CPL, 1, 2, 512, DPP, BIN>(kind, d_pd, tasks, ntasks, w, lds, st);
    } else if constexpr (BIN) {
        scan3d_launch<CPL, 1, 2, 1024, DPP, BIN>(kind, d_pd, tasks, ntasks, w, lds, st);
    } else {
        constexpr int PF = CPL >= 4 ? 2 : 4;  // 4 cells per lane: a deeper ring would spill at 1024 threads
        if (shape == 1) scan3d_launch<CPL, U2, 1, 1024, DPP>(kind, d_pd, tasks, ntasks, w, lds, st);
        else scan3d_launch<CPL, 1, PF, 1024, DPP>(kind, d_pd, tasks, ntasks, w, lds, st);
    }
    (void)shape;
    tune_commit(t);
}

// Persistent work list for the lockstep line sweeps (set_sweep_worklist): off unless
// RADIANT_HIP_WORKLIST=1; RADIANT_HIP_WORKLIST_WG single-wave workgroups per CU (default 8), or
// RADIANT_HIP_WORKLIST_GRID workgroups in all.
thread_local unsigned* tl_worklist = nullptr;
void set_sweep_worklist(unsigned* ctr) { tl_worklist = ctr; }
static bool worklist_on() {
    const char* e = std::getenv("RADIANT_HIP_WORKLIST");
    return e && e[0] == '1';
}
static int worklist_grid() {
    static const int cus = [] {
        int dev = 0, n = 256;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            n = 256;
        }
        return n;
    }();
    const char* g = std::getenv("RADIANT_HIP_WORKLIST_GRID");
    if (g && std::atoi(g) > 0) return std::atoi(g);
    const char* e = std::getenv("RADIANT_HIP_WORKLIST_WG");
    return cus * (e && std::atoi(e) > 0 ? std::atoi(e) : 8);
}

template <int CPL, int PD, bool DPP, bool BIN>
static void line_launch(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, size_t lds, hipStream_t st) {
    if constexpr (BIN) {  // FP64 and FP32 passes of one problem
        auto f64 = sweep_line_kernel<double, CPL, PD, DPP, true>;
        auto f32 = sweep_line_kernel<float, CPL, PD, DPP, true>;
        static bool battr[2] = {false, false};
        const int bi = kind == P64 ? 0 : 1;
        if (!battr[bi]) {
            RT_HIP(hipFuncSetAttribute(bi == 0 ? (const void*)f64 : (const void*)f32,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            battr[bi] = true;
        }
        if (bi == 0) launch_timed(f64, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
        else launch_timed(f32, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
        return;
    }
    const void* fn = kind == PMIX ? (const void*)sweep_line_mixed_kernel<CPL, PD, DPP>
                     : kind == P64 ? (const void*)sweep_line_kernel<double, CPL, PD, DPP>
                                   : (const void*)sweep_line_kernel<float, CPL, PD, DPP>;
    static bool attr_set[3] = {};
    const int ai = kind == PMIX ? 0 : (kind == P64 ? 1 : 2);
    if (!attr_set[ai]) {
        RT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set[ai] = true;
    }
    if (kind == PMIX && tl_worklist && worklist_on()) {
        auto wl = sweep_line_worklist_kernel<CPL, PD, DPP>;
        static bool wattr = false;
        if (!wattr) {
            RT_HIP(hipFuncSetAttribute((const void*)wl, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            wattr = true;
        }
        const int grid = ntasks < worklist_grid() ? ntasks : worklist_grid();
        unsigned* ctr = tl_worklist;
        tl_worklist = nullptr;
        launch_timed(wl, dim3(grid), dim3(64), lds, st, d_pd, tasks, ntasks, ctr);
    } else if (kind == PMIX)
        launch_timed(sweep_line_mixed_kernel<CPL, PD, DPP>, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
    else if (kind == P64)
        launch_timed(sweep_line_kernel<double, CPL, PD, DPP>, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
    else
        launch_timed(sweep_line_kernel<float, CPL, PD, DPP>, dim3(ntasks), dim3(64), lds, st, d_pd, tasks);
}

template <int CPL, bool DPP, bool BIN>
static void line_inst(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, size_t lds, hipStream_t st) {
    // rows in flight: at most PD0 keeps the outstanding loads and stores of PD rows within vmcnt's
    // 63; fewer rows in flight are the autotuner's alternatives (same arithmetic)
    constexpr int PD0 = CPL <= 2 ? 8 : (CPL == 4 ? 4 : 2);
    constexpr int NC = CPL <= 2 ? 3 : (CPL == 4 ? 2 : 1);
    TuneTicket t;
    const int v = tune_choose({kind, 0, CPL, 0, BIN ? 1 : 0, ilog2(ntasks)}, NC, t);
    if (t.a) set_launch_events(t.a, t.b);
    const int pd = PD0 >> v;
    if constexpr (CPL <= 2) {
        if (pd == 8) line_launch<CPL, 8, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
        else if (pd == 4) line_launch<CPL, 4, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
        else line_launch<CPL, 2, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
    } else if constexpr (CPL == 4) {
        if (pd == 4) line_launch<CPL, 4, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
        else line_launch<CPL, 2, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
    } else {
        line_launch<CPL, 2, DPP, BIN>(kind, d_pd, tasks, ntasks, lds, st);
    }
    tune_commit(t);
}

template <int CPL, bool BIN>
static void scan_any(int kind, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int waves, size_t lds,
                     hipStream_t st) {
    if (scan_use_dpp()) {
        if (waves == 0) line_inst<CPL, true, BIN>(kind, d_pd, tasks, ntasks, lds, st);
        else scan3d_inst<CPL, true, BIN>(kind, d_pd, tasks, ntasks, waves, lds, st);
    } else {
        if (waves == 0) line_inst<CPL, false, BIN>(kind, d_pd, tasks, ntasks, lds, st);
        else scan3d_inst<CPL, false, BIN>(kind, d_pd, tasks, ntasks, waves, lds, st);
    }
}

void launch_sweep_scan(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int cpl, int waves, size_t lds,
                       hipStream_t st, bool bin, const ScanSplitGeom* sg) {
    if (ntasks <= 0) return;
    const int kind = prec == PMIX ? PMIX : (prec == P64 ? P64 : P32);
    if (bin && kind == PMIX) throw std::runtime_error("boundary inflow is supported by single-problem passes only");
    if (cpl & SCAN_PAIR) {
        cpl &= SCAN_PAIR - 1;
        if (waves > 0 && scan_use_dpp()) {  // the shuffle cross-check runs one pencil per wave
            switch (cpl) {
                case 1: bin ? scan3d_pair_inst<1, true>(kind, d_pd, tasks, ntasks, waves, lds, st, nullptr)
                            : scan3d_pair_inst<1, false>(kind, d_pd, tasks, ntasks, waves, lds, st, sg); break;
                case 2: bin ? scan3d_pair_inst<2, true>(kind, d_pd, tasks, ntasks, waves, lds, st, nullptr)
                            : scan3d_pair_inst<2, false>(kind, d_pd, tasks, ntasks, waves, lds, st, sg); break;
                default: bin ? scan3d_pair_inst<4, true>(kind, d_pd, tasks, ntasks, waves, lds, st, nullptr)
                             : scan3d_pair_inst<4, false>(kind, d_pd, tasks, ntasks, waves, lds, st, sg); break;
            }
            RT_HIP(hipGetLastError());
            return;
        }
    }
    if (bin) {
        switch (cpl) {
            case 1: scan_any<1, true>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            case 2: scan_any<2, true>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            case 4: scan_any<4, true>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            default: scan_any<8, true>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
        }
    } else {
        switch (cpl) {
            case 1: scan_any<1, false>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            case 2: scan_any<2, false>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            case 4: scan_any<4, false>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
            default: scan_any<8, false>(kind, d_pd, tasks, ntasks, waves, lds, st); break;
        }
    }
    RT_HIP(hipGetLastError());
}

void launch_csd_source(const PassDesc* d_pd, long n, hipStream_t st) {
    hipLaunchKernelGGL(csd_source_kernel, dim3(grid_for((size_t)n)), dim3(256), 0, st, d_pd, n);
    RT_HIP(hipGetLastError());
}
void launch_csd_out(const PassDesc* d_pd, long n, hipStream_t st) {
    hipLaunchKernelGGL(csd_out_kernel, dim3(grid_for((size_t)n)), dim3(256), 0, st, d_pd, n);
    RT_HIP(hipGetLastError());
}
template <typename T>
static void ho_launch(const PassDesc* d_pd, int Nd, int K, hipStream_t st) {
    switch (K) {
        case 1: launch_timed(sweep_1d_ho_kernel<T, 1>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 2: launch_timed(sweep_1d_ho_kernel<T, 2>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 3: launch_timed(sweep_1d_ho_kernel<T, 3>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 4: launch_timed(sweep_1d_ho_kernel<T, 4>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 5: launch_timed(sweep_1d_ho_kernel<T, 5>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 6: launch_timed(sweep_1d_ho_kernel<T, 6>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 7: launch_timed(sweep_1d_ho_kernel<T, 7>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        case 8: launch_timed(sweep_1d_ho_kernel<T, 8>, dim3(Nd), dim3(64), 0, st, d_pd); break;
        default: throw std::runtime_error("1D closure order out of range");
    }
}
size_t sweep_2d_ho_lds_bytes(int prec, int Nx, int Ny, int Kx, int Ky) {
    return (prec == P64 ? 8 : 4) * ((size_t)Ny * Ky + (size_t)Nx * Kx);
}

template <typename T, int KX, int KY, bool FC>
static void ho2_launch(const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    auto fn = sweep_2d_ho_kernel<T, KX, KY, FC>;
    static bool attr = false;
    if (!attr) {
        RT_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    launch_timed(fn, dim3(Nd), dim3(256), (uint32_t)lds, st, d_pd);  // with the caller's timing events
}

bool sweep_2d_ho_supported(int Kx, int Ky, bool fc) {
    if (Kx < 1 || Ky < 1 || Kx > HO2_KMAX || Ky > HO2_KMAX) return false;
    return (fc ? Kx * Ky : Kx + Ky - 1) <= HO2_NM_MAX;
}

template <typename T, int KX, int KY>
static void ho2_fc(bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    // partial coupling differs from full coupling only when both orders exceed 1
    if constexpr (KX > 1 && KY > 1) {
        if (!fc) return ho2_launch<T, KX, KY, false>(d_pd, Nd, lds, st);
    }
    // the cell system lives in registers: full coupling up to HO2_NM_MAX moments
    if constexpr (KX * KY <= HO2_NM_MAX) ho2_launch<T, KX, KY, true>(d_pd, Nd, lds, st);
    else throw std::runtime_error("2D closure with more than HO2_NM_MAX moments per cell");
}

template <typename T, int KX>
static void ho2_y(int Ky, bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    switch (Ky) {
        case 1: return ho2_fc<T, KX, 1>(fc, d_pd, Nd, lds, st);
        case 2: return ho2_fc<T, KX, 2>(fc, d_pd, Nd, lds, st);
        case 3: return ho2_fc<T, KX, 3>(fc, d_pd, Nd, lds, st);
        case 4: return ho2_fc<T, KX, 4>(fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("2D closure orders out of range");
    }
}

template <typename T>
static void ho2_dispatch(const PassDesc* d_pd, int Nd, int Kx, int Ky, bool fc, size_t lds, hipStream_t st) {
    if (!sweep_2d_ho_supported(Kx, Ky, fc)) throw std::runtime_error("2D closure orders out of range");
    switch (Kx) {
        case 1: return ho2_y<T, 1>(Ky, fc, d_pd, Nd, lds, st);
        case 2: return ho2_y<T, 2>(Ky, fc, d_pd, Nd, lds, st);
        case 3: return ho2_y<T, 3>(Ky, fc, d_pd, Nd, lds, st);
        case 4: return ho2_y<T, 4>(Ky, fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("2D closure orders out of range");
    }
}

void launch_sweep_2d_ho(int prec, const PassDesc* d_pd, int Nd, int Kx, int Ky, bool fc, size_t lds, hipStream_t st) {
    if (Nd <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("2D closure face moments exceed the LDS");
    if (prec == P64) ho2_dispatch<double>(d_pd, Nd, Kx, Ky, fc, lds, st);
    else ho2_dispatch<float>(d_pd, Nd, Kx, Ky, fc, lds, st);
    RT_HIP(hipGetLastError());
}

size_t sweep_3d_ho_lds_bytes(int prec, int Nx, int Ny, int Nz, int Kx, int Ky, int Kz, bool fc) {
    const size_t f0 = fc ? Ky * Kz : Ky + Kz - 1, f1 = fc ? Kx * Kz : Kx + Kz - 1, f2 = fc ? Kx * Ky : Kx + Ky - 1;
    return (prec == P64 ? 8 : 4) * ((size_t)Ny * Nz * f0 + (size_t)Nx * Nz * f1 + (size_t)Nx * Ny * f2);
}

template <typename T, int KX, int KY, int KZ, bool FC>
static void ho3_launch(const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    auto fn = sweep_3d_ho_kernel<T, KX, KY, KZ, FC>;
    static bool attr = false;
    if (!attr) {
        RT_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    launch_timed(fn, dim3(Nd), dim3(256), (uint32_t)lds, st, d_pd);  // with the caller's timing events
}

bool sweep_3d_ho_supported(int Kx, int Ky, int Kz, bool fc) {
    if (Kx < 1 || Ky < 1 || Kz < 1 || Kx > HO3_KMAX || Ky > HO3_KMAX || Kz > HO3_KMAX) return false;
    return (fc ? Kx * Ky * Kz : Kx + Ky + Kz - 2) <= HO3_NM_MAX;
}

template <typename T, int KX, int KY, int KZ>
static void ho3_fc(bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    // partial coupling differs from full coupling only when two orders or more exceed 1
    if constexpr ((KX > 1) + (KY > 1) + (KZ > 1) >= 2) {
        if (!fc) return ho3_launch<T, KX, KY, KZ, false>(d_pd, Nd, lds, st);
    }
    // the cell system lives in registers: full coupling up to HO3_NM_MAX moments
    if constexpr (KX * KY * KZ <= HO3_NM_MAX) ho3_launch<T, KX, KY, KZ, true>(d_pd, Nd, lds, st);
    else throw std::runtime_error("3D closure with more than HO3_NM_MAX moments per cell");
}

template <typename T, int KX, int KY>
static void ho3_z(int Kz, bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    switch (Kz) {
        case 1: return ho3_fc<T, KX, KY, 1>(fc, d_pd, Nd, lds, st);
        case 2: return ho3_fc<T, KX, KY, 2>(fc, d_pd, Nd, lds, st);
        case 3: return ho3_fc<T, KX, KY, 3>(fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("3D closure orders out of range");
    }
}

template <typename T, int KX>
static void ho3_yz(int Ky, int Kz, bool fc, const PassDesc* d_pd, int Nd, size_t lds, hipStream_t st) {
    switch (Ky) {
        case 1: return ho3_z<T, KX, 1>(Kz, fc, d_pd, Nd, lds, st);
        case 2: return ho3_z<T, KX, 2>(Kz, fc, d_pd, Nd, lds, st);
        case 3: return ho3_z<T, KX, 3>(Kz, fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("3D closure orders out of range");
    }
}

template <typename T>
static void ho3_dispatch(const PassDesc* d_pd, int Nd, int Kx, int Ky, int Kz, bool fc, size_t lds, hipStream_t st) {
    if (!sweep_3d_ho_supported(Kx, Ky, Kz, fc)) throw std::runtime_error("3D closure orders out of range");
    switch (Kx) {
        case 1: return ho3_yz<T, 1>(Ky, Kz, fc, d_pd, Nd, lds, st);
        case 2: return ho3_yz<T, 2>(Ky, Kz, fc, d_pd, Nd, lds, st);
        case 3: return ho3_yz<T, 3>(Ky, Kz, fc, d_pd, Nd, lds, st);
        default: throw std::runtime_error("3D closure orders out of range");
    }
}

void launch_sweep_3d_ho(int prec, const PassDesc* d_pd, int Nd, int Kx, int Ky, int Kz, bool fc, size_t lds,
                        hipStream_t st) {
    if (Nd <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("3D closure face moments exceed the LDS");
    if (prec == P64) ho3_dispatch<double>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
    else ho3_dispatch<float>(d_pd, Nd, Kx, Ky, Kz, fc, lds, st);
    RT_HIP(hipGetLastError());
}

void launch_sweep_1d_ho(int prec, const PassDesc* d_pd, int Nd, int K, hipStream_t st) {
    if (Nd <= 0) return;
    if (prec == P64) ho_launch<double>(d_pd, Nd, K, st);
    else ho_launch<float>(d_pd, Nd, K, st);
    RT_HIP(hipGetLastError());
}

void launch_fp_source(const PassDesc* d_pd, long n, hipStream_t st) {
    hipLaunchKernelGGL(fp_source_kernel, dim3(grid_for((size_t)n)), dim3(256), 0, st, d_pd, n);
    RT_HIP(hipGetLastError());
}

void launch_bndin(const PassDesc* d_pd, int total, hipStream_t st) {
    if (total <= 0) return;
    hipLaunchKernelGGL(bndin_kernel, dim3(grid_for((size_t)total)), dim3(256), 0, st, d_pd);
    RT_HIP(hipGetLastError());
}

void launch_sweep(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int ppt, int nbuf, int block,
                  size_t lds, hipStream_t st) {
    if (ntasks <= 0) return;
    if (prec == P64) {
        if (nbuf == 2) sweep_ppt<double, 2>(ppt, d_pd, tasks, ntasks, block, lds, st);
        else sweep_ppt<double, 1>(ppt, d_pd, tasks, ntasks, block, lds, st);
    } else {
        if (nbuf == 2) sweep_ppt<float, 2>(ppt, d_pd, tasks, ntasks, block, lds, st);
        else sweep_ppt<float, 1>(ppt, d_pd, tasks, ntasks, block, lds, st);
    }
    RT_HIP(hipGetLastError());
}

void launch_bndmom(const PassDesc* d_pd, int nprob, int max_items, hipStream_t st) {
    launch_timed(bndmom_kernel, dim3(grid_for((size_t)max_items)), dim3(256), 0, st, d_pd, nprob);
    RT_HIP(hipGetLastError());
}

void launch_outsrc(double* qlout, const double* phi_all, const double* sigo, const double* qvol, const int* mat,
                   const int* act, int g, int G, int Nmat, int Nact, int Np, int Nc, hipStream_t st) {
    hipLaunchKernelGGL(outsrc_kernel, dim3(grid_for((size_t)Nact * Nc)), dim3(256), 0, st, qlout, phi_all, sigo, qvol, mat,
                       act, g, G, Nmat, Nact, Np, Nc);
    RT_HIP(hipGetLastError());
}

void k_dot(const RedWork& rw, int slot, const double* a, const double* b, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(dot_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, a, b, n, rw.partial);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(RED_THREADS), 0, st, rw.partial, RED_BLOCKS, rw.scal + slot);
    RT_HIP(hipGetLastError());
}
void k_mgs(const RedWork& rw, int hslot, int nslot, double* w, const double* v, const double* vnext, size_t n, hipStream_t st) {
    if (vnext)
        hipLaunchKernelGGL(mgs_kernel<true>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, w, v, vnext, n, rw.scal + hslot,
                           rw.partial);
    else
        hipLaunchKernelGGL(mgs_kernel<false>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, w, v, vnext, n,
                           rw.scal + hslot, rw.partial);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(RED_THREADS), 0, st, rw.partial, RED_BLOCKS, rw.scal + nslot);
    RT_HIP(hipGetLastError());
}
void k_fill(double* x, double v, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, v, n);
    RT_HIP(hipGetLastError());
}
void k_copy(double* dst, const double* src, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, st, dst, src, n);
    RT_HIP(hipGetLastError());
}
void k_scale(double* x, double a, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, a, n);
    RT_HIP(hipGetLastError());
}
void k_axpy(double a, const double* x, double* y, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, x, y, n);
    RT_HIP(hipGetLastError());
}
void k_neg_add(double* r, const double* c, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(neg_add_kernel, dim3(grid_for(n)), dim3(256), 0, st, r, c, n);
    RT_HIP(hipGetLastError());
}
void k_sub(double* out, const double* a, const double* b, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(sub_kernel, dim3(grid_for(n)), dim3(256), 0, st, out, a, b, n);
    RT_HIP(hipGetLastError());
}
void k_scale_by_inv_slot(double* dst, const double* src, const double* scal, int slot, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(scale_inv_slot_kernel, dim3(grid_for(n)), dim3(256), 0, st, dst, src, scal, slot, n);
    RT_HIP(hipGetLastError());
}
void k_lincomb(double* z, const double* const* V, const double* y, int k, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n)), dim3(256), 0, st, z, V, y, k, n);
    RT_HIP(hipGetLastError());
}
void k_cast_f64_f32(float* dst, const double* src, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(cast_kernel, dim3(grid_for(n)), dim3(256), 0, st, dst, src, n);
    RT_HIP(hipGetLastError());
}
void k_scatter_state(double* dst, const double* src, int rows, const int* rowmap, int Nc, hipStream_t st) {
    hipLaunchKernelGGL(scatter_rows_kernel, dim3(grid_for((size_t)rows * Nc)), dim3(256), 0, st, dst, src, rows, rowmap, Nc);
    RT_HIP(hipGetLastError());
}
void k_set_pd(const PassDesc& v, PassDesc* dst, hipStream_t st) {
    hipLaunchKernelGGL(set_pd_kernel, dim3(1), dim3(1), 0, st, v, dst);
    RT_HIP(hipGetLastError());
}
void k_set_vals(const Vals32& v, double* dst, int k, hipStream_t st) {
    hipLaunchKernelGGL(set_vals_kernel, dim3(1), dim3(32), 0, st, v, dst, k);
    RT_HIP(hipGetLastError());
}
void k_phi_to_radiant(double* dst, const double* phi_all, int G, int Np, int Nc, int c0, int nc, hipStream_t st) {
    hipLaunchKernelGGL(phi_to_radiant_kernel, dim3(grid_for((size_t)G * Np * nc)), dim3(256), 0, st, dst, phi_all, G, Np,
                       Nc, c0, nc);
    RT_HIP(hipGetLastError());
}
void k_cutoff(double* out, const double* DT, const double* psiE, int Np, int Nd, int Ncell, int K, int FE,
              hipStream_t st) {
    hipLaunchKernelGGL(cutoff_kernel, dim3(grid_for((size_t)Np * Ncell * K)), dim3(256), 0, st, out, DT, psiE, Np, Nd,
                       Ncell, K, FE);
    RT_HIP(hipGetLastError());
}
void k_state_moments_in(double* dst, const double* src, int Np, int Nc, hipStream_t st) {
    // src [c][p] (p fastest) -> dst [p][c]
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for((size_t)Np * Nc)), dim3(256), 0, st, dst, src, Nc, Np);
    RT_HIP(hipGetLastError());
}
void k_state_moments_out(double* dst, const double* src, int Np, int Nc, hipStream_t st) {
    // src [p][c] -> dst [c][p]
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for((size_t)Np * Nc)), dim3(256), 0, st, dst, src, Np, Nc);
    RT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------
// Batched helpers: task-table expansion, segmented dot products, vector-op lists.
// ------------------------------------------------------------------------------------------
__global__ void expand_gemm_kernel(const int* __restrict__ meta, int n, GemmTask* __restrict__ out) {
    // meta: [prob, ntile, nrg, offset] per entry
    const int e = blockIdx.x;
    if (e >= n) return;
    const int prob = meta[4 * e], ntile = meta[4 * e + 1], nrg = meta[4 * e + 2], off = meta[4 * e + 3];
    for (int i = threadIdx.x; i < ntile * nrg; i += blockDim.x) {
        int tile, rg;
        gemm_task_order(i, ntile, nrg, tile, rg);
        out[off + i] = GemmTask{prob, tile, rg, 0};
    }
}

__global__ void expand_sweep_kernel(const int* __restrict__ meta, int n, SweepTask* __restrict__ out, unsigned* ctr) {
    // meta: [prob, ndir, offset] per entry
    const int e = blockIdx.x;
    if (e == 0 && threadIdx.x == 0 && ctr) *ctr = 0u;  // work-list counter of the sweep that follows
    if (e >= n) return;
    const int prob = meta[3 * e], nd = meta[3 * e + 1], off = meta[3 * e + 2];
    for (int i = threadIdx.x; i < nd; i += blockDim.x) out[off + i] = SweepTask{prob, i};
}

// UPD: y -= sum_i coef_i V_i first (in place); NRM: the output is y . y, else V_i . y.  A block
// streams its contiguous chunk in batches of UE elements per thread, two batches in flight: the
// loads of batch b+1 are issued before batch b is worked, and every load and store is
// unconditional (padding lanes read a clamped element and store to the trash line), so the
// waits never cover the next batch.  Slots i >= k alias y with coefficient 0 (x - 0*v == x).
template <int KM, bool UPD, bool NRM>
__global__ __launch_bounds__(256) void gs_kernel(const GsItem* __restrict__ items, double* partial) {
    constexpr int UE = KM <= 8 ? 2 : 1;
    const GsItem it = items[blockIdx.y];
    if ((int)blockIdx.x >= it.bpi) return;
    const int k = it.k;
    const double* __restrict__ Vp[KM];
    double c[KM], acc[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        Vp[i] = i < k ? it.V[i] : it.y;
        c[i] = (UPD && i < k) ? it.coef[i] : 0.0;
        acc[i] = 0.0;
    }
    double* __restrict__ y = it.y;
    const long chunk = (it.n + it.bpi - 1) / it.bpi;
    const long lo = (long)blockIdx.x * chunk;
    const long hi = lo + chunk < it.n ? lo + chunk : it.n;
    const long nb = hi > lo ? (hi - lo + 256 * UE - 1) / (256 * UE) : 0;
    const int tid = threadIdx.x;
    struct Batch { double v[UE][KM]; double x[UE]; };
    auto load = [&](long b, Batch& B) {
#pragma unroll
        for (int u = 0; u < UE; ++u) {
            const long e = lo + (b * UE + u) * 256 + tid;
            const long ec = e < hi ? e : hi - 1;
#pragma unroll
            for (int i = 0; i < KM; ++i) B.v[u][i] = gld(Vp[i] + ec);
            B.x[u] = gld(y + ec);
        }
    };
    auto work = [&](long b, const Batch& B) {
#pragma unroll
        for (int u = 0; u < UE; ++u) {
            const long e = lo + (b * UE + u) * 256 + tid;
            const double m = e < hi ? 1.0 : 0.0;
            double x = B.x[u];
            if (UPD) {
#pragma unroll
                for (int i = 0; i < KM; ++i) x -= c[i] * B.v[u][i];
                gst(e < hi ? y + e : (double*)trash_slot<double>(tid & 63, u), x);
            }
            const double xm = x * m;
            if (NRM) {
                acc[0] += xm * x;
            } else {
#pragma unroll
                for (int i = 0; i < KM; ++i) acc[i] += B.v[u][i] * xm;
            }
        }
    };
    if (nb > 0) {
        Batch A, B;
        load(0, A);
        for (long b = 0; b < nb; b += 2) {
            __builtin_amdgcn_sched_barrier(0);
            load(b + 1, B);
            __builtin_amdgcn_sched_barrier(0);
            work(b, A);
            __builtin_amdgcn_sched_barrier(0);
            load(b + 2, A);
            __builtin_amdgcn_sched_barrier(0);
            work(b + 1, B);  // past the chunk when b + 1 == nb: masked, stores to the trash line
        }
    }
    // block partials: butterfly sums within each wave, then the 4 waves in order
    const int nout = NRM ? 1 : k;
    __shared__ double red[4][KM];
    const int lane = tid & 63, wv = tid >> 6;
    auto block_reduce = [&](double (&val)[KM]) {
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            double s = val[i];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) red[wv][i] = s;
        }
        __syncthreads();
    };
    block_reduce(acc);
    double* part = partial + ((size_t)blockIdx.y * GS_BPI_MAX + blockIdx.x) * KM;
    for (int i = tid; i < nout; i += blockDim.x) part[i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

// One block per item: thread b loads the partials of block b (all at once), then a fixed-order
// block sum of each output.
template <int KM, bool NRM>
__global__ __launch_bounds__(256) void gs_finish_kernel(const GsItem* __restrict__ items, const double* __restrict__ partial) {
    const GsItem it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nout = NRM ? 1 : it.k;
    const int b = tid < it.bpi ? tid : it.bpi - 1;
    const double keep = tid < it.bpi ? 1.0 : 0.0;
    __shared__ double red[4][KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        double s = partial[((size_t)blockIdx.x * GS_BPI_MAX + b) * KM + i] * keep;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[wv][i] = s;
    }
    __syncthreads();
    for (int i = tid; i < nout; i += blockDim.x) it.out[i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

template <int KM>
static void gs_launch(dim3 grid, const GsItem* items, bool update, bool norm, double* partial, hipStream_t st) {
    if (update && norm) hipLaunchKernelGGL((gs_kernel<KM, true, true>), grid, dim3(256), 0, st, items, partial);
    else if (update) hipLaunchKernelGGL((gs_kernel<KM, true, false>), grid, dim3(256), 0, st, items, partial);
    else if (norm) hipLaunchKernelGGL((gs_kernel<KM, false, true>), grid, dim3(256), 0, st, items, partial);
    else hipLaunchKernelGGL((gs_kernel<KM, false, false>), grid, dim3(256), 0, st, items, partial);
    if (norm) hipLaunchKernelGGL((gs_finish_kernel<KM, true>), dim3(grid.y), dim3(256), 0, st, items, partial);
    else hipLaunchKernelGGL((gs_finish_kernel<KM, false>), dim3(grid.y), dim3(256), 0, st, items, partial);
}

int gs_blocks(long n) {
    const long b = (n + 4095) / 4096;
    return (int)(b < 8 ? 8 : (b > GS_BPI_MAX ? GS_BPI_MAX : b));
}

void launch_gs(const GsItem* items, int nitems, int kmax, int bpi_max, bool update, bool norm, double* partial,
               hipStream_t st) {
    if (nitems <= 0) return;
    if (kmax > GS_KMAX || bpi_max > GS_BPI_MAX) throw std::runtime_error("launch_gs: item too large");
    const dim3 grid(bpi_max, nitems);
    if (kmax <= 2) gs_launch<2>(grid, items, update, norm, partial, st);
    else if (kmax <= 4) gs_launch<4>(grid, items, update, norm, partial, st);
    else if (kmax <= 8) gs_launch<8>(grid, items, update, norm, partial, st);
    else if (kmax <= 16) gs_launch<16>(grid, items, update, norm, partial, st);
    else gs_launch<GS_KMAX>(grid, items, update, norm, partial, st);
    RT_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void vecop_kernel(const VecItem* __restrict__ items) {
    const VecItem it = items[blockIdx.y];
    double a = it.alpha;
    if (it.op == VOP_NORMALIZE) {
        const double h = *it.slot;
        if (!(h > 0.0)) return;
        a = 1.0 / sqrt(h);
    }
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < it.n; i += (long)gridDim.x * blockDim.x) {
        switch (it.op) {
            case VOP_FILL: it.y[i] = a; break;
            case VOP_COPY: it.y[i] = it.x[i]; break;
            case VOP_SCALE_COPY:
            case VOP_NORMALIZE: it.y[i] = it.x[i] * a; break;
            case VOP_SUB: it.y[i] = it.x[i] - it.x2[i]; break;
            case VOP_ADD: it.y[i] = it.x[i] + it.x2[i]; break;
            case VOP_SUBMULTI: {
                double v = it.y[i];
                for (int t = 0; t < it.k; ++t) v -= it.coef[t] * it.V[t][i];
                it.y[i] = v;
                break;
            }
            case VOP_ADDMULTI: {
                double v = it.y[i];
                for (int t = 0; t < it.k; ++t) v += it.coef[t] * it.V[t][i];
                it.y[i] = v;
                break;
            }
            default: break;
        }
    }
}

void launch_expand_gemm(const int* meta, int n, GemmTask* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(expand_gemm_kernel, dim3(n), dim3(256), 0, st, meta, n, out);
    RT_HIP(hipGetLastError());
}
void launch_expand_sweep(const int* meta, int n, SweepTask* out, hipStream_t st, unsigned* ctr) {
    if (n <= 0) return;
    hipLaunchKernelGGL(expand_sweep_kernel, dim3(n), dim3(256), 0, st, meta, n, out, ctr);
    RT_HIP(hipGetLastError());
}
void launch_vecops(const VecItem* items, int nitems, int blocks_per_item, hipStream_t st) {
    if (nitems <= 0) return;
    hipLaunchKernelGGL(vecop_kernel, dim3(blocks_per_item, nitems), dim3(256), 0, st, items);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
