// engine.hpp -- device-resident SN problem and its solver (host orchestration of the kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/radiant_hip.h"
#include "kernels.hpp"

namespace rt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define RT_CHECK(x)                                                                                        \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess)                                                                              \
            throw ::rt::Error(RT_EDEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ") at " + \
                                              __FILE__ + ":" + std::to_string(__LINE__));                 \
    } while (0)

// Device allocations through the cache of pool.cpp: dev_free takes a block back once no pending
// device work uses it (it is kept mapped for the next allocation of a fitting size).
void* dev_alloc(size_t bytes);
void dev_free(void* p);
size_t dev_cached_bytes();
void dev_trim();

// Message returned by rt_last_error() on the calling thread.
void set_last_error(const std::string& m);

// Relative residual at which the low-precision inner GMRES of an iterative-refinement step stops
// (FP32: 1e-5, BF16 operators: 3e-3); RADIANT_HIP_INNER_TOL overrides the FP32 value.
double inner_tol_for(int prec);

// Runs one contraction kernel in isolation (test hook behind rt_test_contraction).
void test_contraction(int which, int prec, int M, int K, int N, const double* A, const double* B, double* C);

// Pass kinds of sn_inner_pass! / the GMRES matvec.
enum PassMode { MODE_T = 0, MODE_A = 1, MODE_IMA = 2 };

struct SolveCounters {
    int64_t passes = 0, passes_lo = 0;
    double updates_raw = 0;
};

// Per-kernel device time, recorded with HIP events on the problem's stream when profiling is on
// (environment RADIANT_HIP_PROFILE=1).  Index: 0 mgemm, 1 sweep, 2 dgemm, 3 bndmom.
struct KernelTimes {
    double ms[4] = {0, 0, 0, 0};
    int64_t launches[4] = {0, 0, 0, 0};
    // per-precision split of the sweep and contraction kernels [prec][kernel]
    double ms_p[3][4] = {};
    int64_t n_p[3][4] = {};
};

class BatchSolver;

class Problem {
    friend class BatchSolver;

public:
    explicit Problem(const rt_problem_desc& d, hipStream_t stream = nullptr);
    ~Problem();
    Problem(const Problem&) = delete;
    Problem& operator=(const Problem&) = delete;

    static int64_t device_bytes(const rt_problem_desc& d);

    // Solve all groups (sn_flux.jl:150-224); fluxes stay on the device.
    void solve(rt_group_stats* stats, rt_solve_info* info);
    // Copy 𝚽l to the host in Radiant's layout.
    // Flux moments in Radiant's layout into host memory, on stream s (default: the problem's).
    void get_flux(double* phi_out, hipStream_t s = nullptr);
    static constexpr size_t FLUX_STAGE = (size_t)8 << 20;  // doubles staged per copy-out step
    // 𝚽cutoff[Np, 1, Nx, Ny, Nz] (sn_flux.jl:213-218): the energy-face flux leaving the last group
    // of the last solve, projected with Dn; zeros unless continuous slowing down is on.
    void get_flux_cutoff(double* out, hipStream_t s = nullptr);
    // Kernel-level hook: one faithful-state pass on host vectors (see rt_problem_pass).
    void host_pass(int g, int mode, const double* qlout_host, const double* zin, double* zout);
    // The same pass on device vectors, with every temporary in the caller-given scratch of cap
    // doubles (rt_problem_pass_device): no device memory is allocated.
    void device_pass(int g, int mode, const double* qlout_dev, const double* zin_dev, double* zout_dev,
                     double* scratch, size_t cap);
    int64_t faithful_state_length() const { return (int64_t)S_f; }

    // ---- pieces used by the batched and served drivers ----
    int ndims, Nx, Ny, Nz, Nc, G, Nmat, Nd, Np, Nps, Nact;
    int Ncell = 0;   // cells; Nc counts moment columns (Ncell times the spatial moments per cell)
    int ho_K = 0;    // spatial moments per cell of a general closure (0: diamond difference, order 1)
    int ho_Kx = 1, ho_Ky = 1, ho_Kz = 1;  // 2D / 3D closure orders
    bool ho_fc = true;         // full coupling of the 2D moments
    int fm[3] = {1, 1, 1};     // moments of the face fluxes per axis (Ky, Kx for a 2D closure)
    // slowing-down cells of a general closure (csd with any order or closure other than diamond
    // difference of order 1): energy order, moments of the energy face; the cell moments ho_K
    // then span (E, x, y, z)
    bool bfp = false;
    bool adapt = false;  // adaptive weighted diamond closure (sweep_awd.hip), FP64 passes only
    bool hog = false;       // a 2D/3D closure beyond the register walkers (sweep_hog.hip)
    double* d_hog = nullptr;  // its cell-system scratch
    double awd_inv_dE = 0.0;  // 1/ΔE of the group of the pass being set up (adaptive slowing-down cells)
    int ho_KE = 1, fm_e = 1;
    // the 3D and slowing-down walkers keep their face moments in device memory when they exceed
    // the LDS (d_hoface, [Nd][faces] doubles); dynamic LDS of their launch in precision sp
    double* d_hoface = nullptr;
    // z-split 3D scan sweeps (PassDesc::sclaim / shand; null: the split shapes are not offered)
    unsigned long long* d_sclaim = nullptr;
    unsigned long long* d_shand = nullptr;
    size_t ho_lds(int sp) const;
    size_t nE() const { return (size_t)Nd * Ncell * fm_e; }  // energy-face angular flux values
    int accel, restart, depth, maxit, precision, state_mode, verbose;
    double tol;
    hipStream_t st;
    bool own_stream = false;

    // per-group solve (used by solve() and by the batch/server schedulers)
    void solve_group(int g, rt_group_stats& s);
    SolveCounters counters;
    KernelTimes ktimes;
    bool profiling = false;

    // device data
    double* d_phi = nullptr;  // [G][Np][Nc]

private:
    std::vector<int> act;
    size_t npos[3] = {0, 0, 0};
    size_t nb = 0, S_f = 0, S_r = 0;
    std::vector<void*> allocs;
    template <class T> T* dalloc(size_t n);

    int* d_mat = nullptr;
    double *d_dx = nullptr, *d_dy = nullptr, *d_dz = nullptr, *d_mu = nullptr, *d_eta = nullptr, *d_xi = nullptr;
    double *d_sigt = nullptr, *d_sigw = nullptr, *d_sigo = nullptr, *d_qvol = nullptr;
    double *d_Mact64 = nullptr, *d_D64 = nullptr, *d_Dact64 = nullptr;
    float *d_Mact32 = nullptr, *d_D32 = nullptr, *d_Dact32 = nullptr;
    int* d_act = nullptr;
    double* d_Dexit[3] = {nullptr, nullptr, nullptr};
    int* d_exitf[3] = {nullptr, nullptr, nullptr};
    int* d_exlist[3] = {nullptr, nullptr, nullptr};
    int excnt[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    // reflective / periodic faces (faithful FP64): entry M rows, inflow source faces, inflows
    bool has_bc = false, has_ss = false;  // non-void faces; surface sources
    double* d_ssrc[3] = {nullptr, nullptr, nullptr};   // per axis [2][G][Nps][npos]
    int* d_sface[3] = {nullptr, nullptr, nullptr};     // per axis [Nd] entry face with a source, or -1
    std::vector<char> grp_ss;                          // per group: some surface source is nonzero
    // continuous slowing down: (S⁻ + S⁺) per group and material, energy-face angular fluxes, ΔE
    bool csd = false;
    double* d_csdK = nullptr;
    double *d_psiE_in = nullptr, *d_psiE_out = nullptr;
    std::vector<double> dE;
    void set_csd(PassDesc& pd, int g);
    // angular Fokker-Planck term (solver types BFP, FP): ℳ on the active moments, T per group and
    // material, and the source qlout + T ℳ Φ the contraction reads
    bool fp = false;
    double *d_fpM = nullptr, *d_fpT = nullptr, *d_fpq = nullptr;
    bool em = false;                                 // magnetic-field source (em_source.hip)
    double *d_emM = nullptr, *d_emq = nullptr;       // ℳ_EM per group [G][Nact][Nact], its source rows
    int em_g = 0;                                    // group of the pass being set up
    double *d_omx = nullptr, *d_omy = nullptr, *d_omz = nullptr;  // closure weights of the general cells per axis
    double* d_ome = nullptr;   // energy closure weights (slowing-down cells)
    double* d_csdS = nullptr;  // S⁻/ΔE, S⁺/ΔE per group and material [G][Nmat][2] (slowing-down cells)
    int* d_emap = nullptr;     // cell moment of (energy moment, energy-face moment), -1 when dropped
    double* d_Mentry[3] = {nullptr, nullptr, nullptr};
    int* d_insrc[3] = {nullptr, nullptr, nullptr};
    double* d_bin[3] = {nullptr, nullptr, nullptr};
    int64_t bin_total = 0;
    double* d_qlout = nullptr;
    void* d_psi = nullptr;
    void* d_bex[3] = {nullptr, nullptr, nullptr};
    double* d_bscratch = nullptr;   // boundary part of a reconstruction pass
    PassDesc* d_pd = nullptr;
    GemmTask *d_mg = nullptr, *d_dgf = nullptr, *d_dga = nullptr;
    int n_mg = 0, n_dgf = 0, n_dga = 0;
    SweepTask* d_sw = nullptr;
    int sw_ppt = 1, sw_block = 64, sw_nbuf[3] = {2, 2, 2};
    int rows_per_lane = -1;   // > 0: 1D/2D wave-per-direction sweep
    int sweep_kind = 0;       // 0 row-scan, 1 rows/planes, 2 pencil
    int scan_cpl = -1, scan_waves = 1;
    bool planes_ok[2] = {false, false};   // 3D workgroup-per-direction sweep, per precision
    int pl_R[2] = {1, 1}, pl_ZW[2] = {1, 1}, pl_waves[2] = {1, 1};
    size_t sw_lds[3] = {0, 0, 0};

    // Krylov work vectors (allocated on first use, keyed by name and length)
    std::map<std::string, double*> pool;
    double* vec(const std::string& name, size_t n);
    RedWork rw{};
    double* h_scal = nullptr;   // pinned mirror of the scalar slots
    GsItem* d_gsi = nullptr;    // the two Gram-Schmidt items of a classical Arnoldi step
    double* d_gspart = nullptr; // their block partials
    double** d_Vptr = nullptr;  // device table of Krylov basis pointers
    double* d_y = nullptr;

    // passes
    void pass_faithful(int g, int prec, int mode, const double* zin, double* zout);
    void pass_reduced(int g, int prec, int mode, const double* zin, double* zout, void* psi = nullptr);
    void pass_reconstruct(int g, const double* zin_reduced);
    void desc_reduced(PassDesc& pd, int g, int prec, int mode, const double* zin, double* zout);
    void desc_reconstruct(PassDesc& pd, int g, const double* zin_reduced);
    void set_boundary(PassDesc& pd, int g, int mode, const double* zin, double* zout, size_t off);
    void launch_pass(PassDesc& pd, int prec, bool full_rows);
    struct EvPair { hipEvent_t a, b; int kernel, prec; };
    std::vector<hipEvent_t> ev_pool;
    std::vector<EvPair> ev_used;
    size_t ev_next = 0;
    hipEvent_t next_event();
    void collect_kernel_times();
    void fill_geometry(PassDesc& pd, int g);

    // BLAS-1 helpers
    double dot(const double* a, const double* b, size_t n);
    double norm(const double* a, size_t n) { return std::sqrt(dot(a, a, n)); }

    // solvers over a state of length n; pass(out, in) applies the operator
    template <class MV>
    void gmres(size_t n, double* z, MV matvec, const double* c, int m, int64_t maxit_, double tol_, int64_t& iter,
               double& resid, bool& conv, double& rho, const std::string& tag, bool zero_guess = false,
               bool cgs = false);
    // T(z) = A z + T(0) from the angular fluxes of the last A z pass (in d_psi) and of c = T(0)
    // (psi_c, or psi_c alone when no A z pass ran): one full-row D contraction, no sweep.
    void reconstruct_sweep_free(int g, const double* zin_reduced, const double* psi_c, bool have_az);
    template <class FP>
    void anderson(size_t n, double* z, FP fixedpoint, int64_t& iter, double& resid, bool& conv, double& rho);
    template <class FP>
    void livolant(size_t n, double* z, FP fixedpoint, int period, int64_t& iter, double& resid, bool& conv,
                  double& rho);
    template <class MV>
    void bicgstab(size_t n, double* z, MV matvec, const double* c, int64_t& iter, double& resid, bool& conv,
                  double& rho);

    void solve_group_plain(int g, rt_group_stats& s);
    void solve_group_mixed(int g, rt_group_stats& s);
};

// Address of this device's scratch array of the allocation-free pass (taken once per device) and
// its capacity in doubles.
double* noalloc_scratch(size_t* cap);

}  // namespace rt
