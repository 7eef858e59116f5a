// engine.cpp -- host orchestration of the SN solve on one device.
//
// Follows src/particle_transport/sn_flux.jl:150-224 (group loop), sn_one_speed.jl:81-196
// (in-group solve) and tools/{gmres,anderson,livolant,bicgstab}.jl, with every state vector
// resident in HBM and the transport pass of sn_inner_pass.jl:212-332 done by the kernels of
// kernels.hip.
#include "engine.hpp"
#include "sweep_awd.hpp"
#include "em_source.hpp"
#include "sweep_hog.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>

namespace rt {

// out[g][Kc * c] = in[c][g] for the G x Ncell volume source values (out zero elsewhere)
__global__ void qvol_layout_kernel(double* __restrict__ out, const double* __restrict__ in, int G, long Ncell, int Kc,
                                   long Nc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // output order: coalesced writes
    if (i >= (long)G * Ncell) return;
    const int g = (int)(i / Ncell);
    const long c = i - (long)g * Ncell;
    out[(size_t)g * Nc + (size_t)Kc * c] = in[(size_t)G * c + g];
}

static inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

double inner_tol_for(int prec) {
    if (prec == PBF16) return 3e-3;
    static const double t = [] {
        const char* e = std::getenv("RADIANT_HIP_INNER_TOL");
        const double v = e ? std::atof(e) : 0.0;
        return (v > 0.0 && v < 1.0) ? v : 1e-5;
    }();
    return t;
}

template <class T>
T* Problem::dalloc(size_t n) {
    if (n == 0) n = 1;
    void* p = dev_alloc(n * sizeof(T));
    allocs.push_back(p);
    return (T*)p;
}

// Synchronous upload: the host vectors are temporaries (setup cost only).
template <class T>
static void upload(T* d, const std::vector<T>& h, hipStream_t st) {
    RT_CHECK(hipStreamSynchronize(st));
    if (!h.empty()) RT_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// SN solver type (SN.jl get_solver_type): 0 in the description means BTE, or BCSD with csd.
static int solver_type_of(const rt_problem_desc& d) { return d.solver_type ? d.solver_type : (d.csd ? 3 : 1); }
// Spatial moments per cell of a general closure (SN.set_scheme, set_is_full_coupling): 0 for
// diamond difference of order 1 on every axis, which the order-1 kernels take; in 1D the x order,
// in 2D Kx*Ky with full coupling and Kx + Ky - 1 without.
static int ho_order(const rt_problem_desc& d) {
    const int Kx = d.order_x > 1 ? d.order_x : 1, Ky = d.ndims >= 2 && d.order_y > 1 ? d.order_y : 1;
    const int Kz = d.ndims >= 3 && d.order_z > 1 ? d.order_z : 1;
    const int Ke = d.csd && d.order_e > 1 ? d.order_e : 1;
    auto dd1 = [](const double* w) { return !w || (w[0] == -1.0 && w[1] == 2.0); };
    // an adaptive closure walks its cells in order, at order 1 too (sweep_awd.hip)
    if (!d.adaptive && Kx == 1 && Ky == 1 && Kz == 1 && Ke == 1 && dd1(d.omega_x) && (d.ndims < 2 || dd1(d.omega_y)) &&
        (d.ndims < 3 || dd1(d.omega_z)) && (!d.csd || dd1(d.omega_e)))
        return 0;
    if (d.csd) return bfp_cell_moments(Ke, Kx, Ky, Kz, d.partial_coupling == 0);  // moments over (E, x, y, z)
    if (d.ndims == 1) return Kx;
    if (d.ndims == 2) return d.partial_coupling ? Kx + Ky - 1 : Kx * Ky;
    return d.partial_coupling ? Kx + Ky + Kz - 2 : Kx * Ky * Kz;
}

// The 3D order-1 scan sweep with two pencils per wave can split each direction's pencils by z
// over several workgroups (kernels.hip, sweep_scan_body): offered when the handed rows fit its
// courier (Nx <= 64) and the grid has layers for two slabs at least.
static bool split_scan_offered(const rt_problem_desc& d) {
    if (d.ndims != 3 || ho_order(d) || d.Nx > 64 || d.Nz < 4) return false;
    if (!(sweep_scan_cpl3d(d.Nx) & SCAN_PAIR)) return false;
    const char* e = std::getenv("RADIANT_HIP_SCAN_SPLIT");
    return !(e && e[0] == '1' && e[1] == 0);
}
// words of the handed rows: [Nd][SCAN_SPLIT_MAX - 1][Ny][Nx][2]
static size_t split_scan_words(const rt_problem_desc& d) {
    return (size_t)d.Nd * (SCAN_SPLIT_MAX - 1) * d.Ny * d.Nx * 2;
}

static bool has_fp(const rt_problem_desc& d) {
    const int t = solver_type_of(d);
    return (t == 2 || t == 4) && d.T && d.Mfp;
}

// Moments that carry a source: the Legendre orders with Σs_l != 0 for some material/group pair,
// and with the Fokker-Planck term or a magnetic field every moment that ℳ or ℳ_EM reads or writes.
static std::vector<int> active_moments(const rt_problem_desc& d) {
    std::vector<bool> lnz(d.Ls + 1, false);
    const size_t per_l = (size_t)d.Nmat * d.G * d.G;
    for (int l = 0; l <= d.Ls; ++l)
        for (size_t i = 0; i < per_l; ++i)
            if (d.sigs[i + per_l * l] != 0.0) { lnz[l] = true; break; }
    std::vector<bool> fp(d.Np, false);
    if (has_fp(d)) {
        bool tnz = false;
        for (size_t i = 0; i < (size_t)d.G * d.Nmat; ++i) tnz = tnz || d.T[i] != 0.0;
        for (int p = 0; p < d.Np && tnz; ++p)
            for (int q = 0; q < d.Np; ++q)
                if (d.Mfp[p + (size_t)d.Np * q] != 0.0) fp[p] = fp[q] = true;
    }
    if (d.Mem)  // the magnetic-field operator: every moment it reads or writes
        for (int p = 0; p < d.Np; ++p)
            for (int q = 0; q < d.Np; ++q)
                for (int g = 0; g < d.G; ++g)
                    if (d.Mem[g + (size_t)d.G * (p + (size_t)d.Np * q)] != 0.0) { fp[p] = fp[q] = true; break; }
    std::vector<int> act;
    for (int p = 0; p < d.Np; ++p) {
        const int64_t l = d.pl[p];
        if (l == 0 || (l <= d.Ls && lnz[l]) || fp[p]) act.push_back(p);
    }
    return act;
}

static void validate(const rt_problem_desc& d) {
    auto bad = [](const std::string& m) { throw Error(RT_EINVAL, m); };
    if (d.ndims < 1 || d.ndims > 3) bad("Dimension is not 1, 2 or 3.");
    if (d.Nx < 1 || d.Ny < 1 || d.Nz < 1) bad("number of voxels must be positive");
    if (d.ndims < 2 && d.Ny != 1) bad("Ny must be 1 below 2D");
    if (d.ndims < 3 && d.Nz != 1) bad("Nz must be 1 below 3D");
    if (!d.dx || (d.ndims >= 2 && !d.dy) || (d.ndims >= 3 && !d.dz)) bad("missing voxel widths");
    if (!d.mat || !d.sigt || !d.sigs || !d.qvol || !d.mu || !d.eta || !d.xi || !d.w || !d.M || !d.D || !d.pl)
        bad("missing input array");
    if (d.Nmat < 1 || d.G < 1 || d.Ls < 0) bad("invalid cross-section extents");
    if (d.Nd < 1 || d.Np < 1 || d.Nps < 0) bad("invalid angular extents");
    if (d.accel < RT_ACCEL_NONE || d.accel > RT_ACCEL_BICGSTAB) bad("Unknown acceleration method.");
    if (d.precision < RT_PREC_FP64 || d.precision > RT_PREC_MIXED_BF16) bad("unknown precision mode");
    if (d.state_mode != RT_STATE_FAITHFUL && d.state_mode != RT_STATE_REDUCED) bad("unknown state mode");
    if (d.precision != RT_PREC_FP64 && d.accel != RT_ACCEL_GMRES)
        bad("mixed precision requires the gmres acceleration (FP64 Arnoldi, iterative refinement)");
    if (d.restart < 1 && d.accel == RT_ACCEL_GMRES) bad("gmres restart must be >= 1");
    if (d.maxit < 1) bad("maximum number of iterations must be >= 1");
    if (!(d.tol > 0)) bad("convergence criterion must be positive");
    if (!d.bc) bad("missing boundary conditions");
    bool nonvoid = false;
    for (int f = 0; f < 2 * d.ndims; ++f) {
        if (d.bc[f] < 0 || d.bc[f] > 2) bad("Unknown boundary condition type.");
        nonvoid = nonvoid || d.bc[f] != 0;
    }
    bool surf = false;
    for (int f = 0; f < 2 * d.ndims; ++f) surf = surf || d.ssrc[f] != nullptr;
    if (nonvoid || surf) {
        for (int f = 0; f < 2 * d.ndims; ++f)
            if (!d.Msurf[f] || !d.Dsurf[f] || !d.n_to_np[f])
                bad("reflective/periodic boundaries and surface sources need the half-range operators of every face");
    }
    if (surf && (d.Np_source < 1 || d.Np_source > d.Nps))
        bad("Np_source must lie in 1..Nps when a surface source is given");
    const int kmax = d.ndims == 1 ? HO_KMAX : HOG_KMAX;
    if (d.order_x < 0 || d.order_x > kmax || d.order_y < 0 || d.order_y > kmax || d.order_z < 0 || d.order_z > kmax)
        bad("spatial closure orders must lie in 1.." + std::to_string(kmax) + " in " + std::to_string(d.ndims) + "D");
    if (d.order_x > 1 && !d.omega_x) bad("closures of order above 1 need their weights omega_x");
    if (d.ndims >= 2 && d.order_y > 1 && !d.omega_y) bad("closures of order above 1 need their weights omega_y");
    if (d.ndims >= 3 && d.order_z > 1 && !d.omega_z) bad("closures of order above 1 need their weights omega_z");
    // 2D/3D closures: the register walkers up to 12 moments per cell, the general walker beyond
    if (d.ndims >= 2 && !d.csd &&
        !sweep_hog_supported(d.ndims, std::max(1, d.order_x), std::max(1, d.order_y),
                             d.ndims >= 3 ? std::max(1, d.order_z) : 1, d.partial_coupling == 0))
        bad(std::to_string(d.ndims) + "D closures: orders up to " + std::to_string(HOG_KMAX) + " per axis and at most " +
            std::to_string(HOG_NM_MAX) + " moments per cell");
    if (!d.csd && d.order_e > 1) bad("an energy closure needs the slowing-down term (csd)");
    if (d.adaptive) {  // scheme "AWD" (adaptive.jl)
        if (d.adaptive != 1) bad("adaptive must be 0 or 1");
        const bool sd11 = d.order_x <= 1 && d.order_y <= 1 && d.order_z <= 1 && d.order_e <= 1;
        const bool sd22 = d.ndims == 1 && d.order_x == 2 && d.order_e == 2 && d.partial_coupling == 0;
        if (d.csd && !sd11 && !sd22)
            bad("the adaptive closure (AWD) with slowing down: order 1 on every axis, space and energy, or (2, 2) "
                "in (x, E) with full coupling in 1D");
        if (d.precision != RT_PREC_FP64)
            bad("the adaptive closure (AWD) is not linear in the flux: it runs in the FP64 solve, whose "
                "reconstruction is a full pass");
        if (d.state_mode != RT_STATE_FAITHFUL)
            bad("the adaptive closure (AWD) runs in the faithful state: for a pass that is not linear the "
                "Krylov iterates depend on the state they run over");
        if (!sweep_awd_supported(d.ndims, std::max(1, d.order_x), d.ndims >= 2 ? std::max(1, d.order_y) : 1,
                                 d.ndims >= 3 ? std::max(1, d.order_z) : 1))
            bad("the adaptive closure (AWD): orders 1 and 2 in 1D, order 1 on every axis in 2D and 3D, or (2, 2) "
                "with full coupling in 2D");
        if (d.ndims == 2 && d.order_x == 2 && d.partial_coupling)
            bad("the adaptive closure (AWD) of orders (2, 2) needs full coupling");
    }
    if (d.csd && ho_order(d)) {
        const int Ke = d.order_e > 1 ? d.order_e : 1;
        if (Ke > 1 && !d.omega_e) bad("energy closures of order above 1 need their weights omega_e");
        const int kx = std::max(1, d.order_x), ky = d.ndims >= 2 ? std::max(1, d.order_y) : 1,
                  kz = d.ndims >= 3 ? std::max(1, d.order_z) : 1;
        if (!sweep_bfp_supported(d.ndims, Ke, kx, ky, kz, d.partial_coupling == 0) &&
            !sweep_hog_bfp_supported(d.ndims, Ke, kx, ky, kz, d.partial_coupling == 0))
            bad("slowing-down closures: energy order at most 2, spatial orders up to " + std::to_string(HOG_KMAX) +
                " per axis, at most " + std::to_string(HOG_NM_MAX) + " moments per cell");
    }
    if (d.solver_type < 0 || d.solver_type > 6) bad("Unknown type of solver.");
    const int stype = solver_type_of(d);
    if ((stype >= 2) != (d.csd != 0))
        bad("solver types BFP, BCSD, FP, CSD and BFP-EF carry the slowing-down term (csd), BTE does not");
    if ((stype == 2 || stype == 4) && (!d.T || !d.Mfp))
        bad("solver types BFP and FP need the momentum transfer T and the Fokker-Planck matrix Mfp");
    if (d.csd) {
        if (!d.Sb || !d.dE) bad("continuous slowing down needs the boundary stopping powers and group widths");
        for (int g = 0; g < d.G; ++g)
            if (!(d.dE[g] > 0)) bad("energy group widths must be positive");
    }
    const size_t nc = (size_t)d.Nx * d.Ny * d.Nz;
    for (size_t c = 0; c < nc; ++c)
        if (d.mat[c] < 1 || d.mat[c] > d.Nmat) bad("material index out of range at voxel " + std::to_string(c));
    for (int p = 0; p < d.Np; ++p)
        if (d.pl[p] < 0) bad("negative Legendre order in pl");
    if (d.state_mode == RT_STATE_FAITHFUL && d.precision == RT_PREC_FP64) {
        for (int f = 0; f < 2 * d.ndims; ++f)
            if (!d.Dsurf[f] || !d.n_to_np[f]) bad("faithful state needs the half-range operators of every face");
    }
    if (sweep_pencils_per_thread(d.Ny * d.Nz) < 0) bad("transverse grid too large for the sweep kernel (Ny*Nz > 16384)");
}

Problem::Problem(const rt_problem_desc& d, hipStream_t stream) {
    validate(d);
    ndims = d.ndims; Nx = d.Nx; Ny = d.Ny; Nz = d.Nz; Ncell = Nx * Ny * Nz;
    ho_K = ho_order(d);
    bfp = d.csd && ho_K;
    adapt = d.adaptive != 0;
    if (bfp) {
        ho_KE = d.order_e > 1 ? d.order_e : 1;
        ho_Kx = d.order_x > 1 ? d.order_x : 1;
        ho_Ky = ndims >= 2 && d.order_y > 1 ? d.order_y : 1;
        ho_Kz = ndims >= 3 && d.order_z > 1 ? d.order_z : 1;
        ho_fc = d.partial_coupling == 0;
        for (int a = 0; a < 3; ++a) fm[a] = bfp_face_moments(a + 1, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc);
        fm_e = bfp_face_moments(0, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc);
    } else if (ho_K && ndims >= 2) {
        ho_Kx = d.order_x > 1 ? d.order_x : 1;
        ho_Ky = d.order_y > 1 ? d.order_y : 1;
        ho_Kz = ndims == 3 && d.order_z > 1 ? d.order_z : 1;
        ho_fc = d.partial_coupling == 0;
        // a face carries the moments over the other axes (SN.get_schemes: Nm[1..3])
        if (ndims == 2) {
            fm[0] = ho_Ky;
            fm[1] = ho_Kx;
        } else {
            fm[0] = ho_fc ? ho_Ky * ho_Kz : ho_Ky + ho_Kz - 1;
            fm[1] = ho_fc ? ho_Kx * ho_Kz : ho_Kx + ho_Kz - 1;
            fm[2] = ho_fc ? ho_Kx * ho_Ky : ho_Kx + ho_Ky - 1;
        }
    }
    Nc = Ncell * (ho_K ? ho_K : 1);  // moment columns: spatial moment + K * cell
    G = d.G; Nmat = d.Nmat; Nd = d.Nd; Np = d.Np; Nps = d.Nps;
    accel = d.accel; restart = d.restart; depth = d.depth; maxit = d.maxit; tol = d.tol;
    precision = d.precision;
    state_mode = (d.precision != RT_PREC_FP64) ? RT_STATE_REDUCED : d.state_mode;
    verbose = d.verbose;
    act = active_moments(d);
    Nact = (int)act.size();
    if (Nact > CONTRACT_KMAX)
        throw Error(RT_EINVAL, "more than " + std::to_string(CONTRACT_KMAX) + " moments carry a scattering source");
    // boundary positions times face moments per axis
    npos[0] = (size_t)Ny * Nz * fm[0];
    npos[1] = ndims >= 2 ? (size_t)Nx * Nz * fm[1] : 0;
    npos[2] = ndims >= 3 ? (size_t)Nx * Ny * fm[2] : 0;
    nb = (size_t)2 * Nps * (npos[0] + npos[1] + npos[2]);
    for (int f = 0; f < 2 * ndims; ++f) {
        has_bc = has_bc || d.bc[f] != 0;
        has_ss = has_ss || d.ssrc[f] != nullptr;
    }
    S_f = (size_t)Np * Nc + nb;
    // the reduced state carries the boundary moments when they feed back (reflective, periodic)
    S_r = (size_t)Nact * Nc + (has_bc ? nb : 0);
    if (stream) {
        st = stream;
    } else {
        RT_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        own_stream = true;
    }
    try {
        // ---- geometry
        const int Kc = Nc / Ncell;
        std::vector<int> mat(Nc);  // per moment column
        for (int c = 0; c < Nc; ++c) mat[c] = (int)(d.mat[c / Kc] - 1);
        d_mat = dalloc<int>(Nc); upload(d_mat, mat, st);
        if (bfp) {  // closure weights of every axis and of energy, the energy-face moment map
            auto weights = [&](const double* w, int K) {
                std::vector<double> v(w ? w : nullptr, w ? w + K + 1 : nullptr);
                if (v.empty()) v = {-1.0, 2.0};
                double* p = dalloc<double>(v.size());
                upload(p, v, st);
                return p;
            };
            d_omx = weights(d.omega_x, ho_Kx);
            d_omy = weights(ndims >= 2 ? d.omega_y : nullptr, ho_Ky);
            d_omz = weights(ndims >= 3 ? d.omega_z : nullptr, ho_Kz);
            d_ome = weights(d.omega_e, ho_KE);
            // cell index of (e, a, b, c): energy fastest; without full coupling one index is
            // non-zero, after the (order - 1) of the axes before it
            const int O[4] = {ho_KE, ho_Kx, ho_Ky, ho_Kz};
            auto index = [&](const int* m, int skip) {
                int r = 0, stride = 1, off = 0;
                for (int k = 0; k < 4; ++k) {
                    if (k == skip) continue;
                    if (ho_fc) { r += m[k] * stride; stride *= O[k]; }
                    else { if (m[k] > 0) r = m[k] + off; off += O[k] - 1; }
                }
                return r;
            };
            std::vector<int> emap((size_t)ho_KE * fm_e, -1);
            for (int a = 0; a < ho_Kx; ++a)
                for (int b = 0; b < ho_Ky; ++b)
                    for (int c = 0; c < ho_Kz; ++c) {
                        if (!ho_fc && (a > 0) + (b > 0) + (c > 0) > 1) continue;
                        for (int e = 0; e < ho_KE; ++e) {
                            const int m[4] = {e, a, b, c};
                            if (!ho_fc && (e > 0) + (a > 0) + (b > 0) + (c > 0) > 1) continue;
                            emap[(size_t)e * fm_e + index(m, 0)] = index(m, -1);
                        }
                    }
            d_emap = dalloc<int>(emap.size()); upload(d_emap, emap, st);
            const size_t fb = sweep_bfp_lds_bytes(P64, ndims, Nx, Ny, Nz, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc);
            if (fb > 160 * 1024) d_hoface = dalloc<double>((size_t)Nd * fb / sizeof(double));
            if (!adapt && !sweep_bfp_supported(ndims, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc)) {
                hog = true;  // beyond the register walker: the general walker and its cell-system scratch
                d_hog = dalloc<double>(sweep_hog_bfp_scratch(Nd, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc));
            }
        } else if (ho_K) {  // closure weights per axis (diamond difference of order 1 when absent)
            if (ndims >= 2 && !adapt &&
                !(ndims == 2 ? sweep_2d_ho_supported(ho_Kx, ho_Ky, ho_fc) : sweep_3d_ho_supported(ho_Kx, ho_Ky, ho_Kz, ho_fc))) {
                hog = true;  // beyond the register walkers: the general walker and its cell-system scratch
                d_hog = dalloc<double>(sweep_hog_scratch(Nd, ho_Kx, ho_Ky, ndims == 3 ? ho_Kz : 1, ho_fc));
            }
            const int kx = ndims >= 2 ? ho_Kx : ho_K;
            std::vector<double> om(d.omega_x ? d.omega_x : nullptr, d.omega_x ? d.omega_x + kx + 1 : nullptr);
            if (om.empty()) om = {-1.0, 2.0};
            d_omx = dalloc<double>(om.size()); upload(d_omx, om, st);
            if (ndims == 2) {
                std::vector<double> oy(d.omega_y ? d.omega_y : nullptr, d.omega_y ? d.omega_y + ho_Ky + 1 : nullptr);
                if (oy.empty()) oy = {-1.0, 2.0};
                d_omy = dalloc<double>(oy.size()); upload(d_omy, oy, st);
                if (sweep_2d_ho_lds_bytes(P64, Nx, Ny, ho_Kx, ho_Ky) > 160 * 1024)
                    throw Error(RT_EINVAL, "the face moments of the 2D closure exceed the 160 KiB LDS of a workgroup");
            }
            if (ndims == 3) {
                std::vector<double> oy(d.omega_y ? d.omega_y : nullptr, d.omega_y ? d.omega_y + ho_Ky + 1 : nullptr);
                std::vector<double> oz(d.omega_z ? d.omega_z : nullptr, d.omega_z ? d.omega_z + ho_Kz + 1 : nullptr);
                if (oy.empty()) oy = {-1.0, 2.0};
                if (oz.empty()) oz = {-1.0, 2.0};
                d_omy = dalloc<double>(oy.size()); upload(d_omy, oy, st);
                d_omz = dalloc<double>(oz.size()); upload(d_omz, oz, st);
                const size_t fb = sweep_3d_ho_lds_bytes(P64, Nx, Ny, Nz, ho_Kx, ho_Ky, ho_Kz, ho_fc);
                if (fb > 160 * 1024) d_hoface = dalloc<double>((size_t)Nd * fb / sizeof(double));
            }
        }
        auto up = [&](const double* src, int n) {
            double* p = dalloc<double>(n);
            RT_CHECK(hipMemcpy(p, src, n * sizeof(double), hipMemcpyHostToDevice));
            return p;
        };
        std::vector<double> one(1, 1.0);
        d_dx = up(d.dx, Nx);
        d_dy = ndims >= 2 ? up(d.dy, Ny) : up(one.data(), 1);
        d_dz = ndims >= 3 ? up(d.dz, Nz) : up(one.data(), 1);
        d_mu = up(d.mu, Nd); d_eta = up(d.eta, Nd); d_xi = up(d.xi, Nd);
        // ---- cross sections (sn_flux.jl:69-83)
        std::vector<double> sigt((size_t)G * Nmat);
        for (int g = 0; g < G; ++g)
            for (int m = 0; m < Nmat; ++m) sigt[(size_t)g * Nmat + m] = d.sigt[g + (size_t)G * m];
        csd = d.csd != 0;
        if (csd) {
            // continuous slowing down (sn_flux.jl:86-97,183-188): S⁻ = Sb(g)/ΔE, S⁺ = Sb(g+1)/ΔE.  In
            // the order-1 cell the removal gains 2 S⁺ and the energy inflow enters with (S⁺ + S⁻);
            // the general cell takes S⁻ and S⁺ themselves.
            std::vector<double> K((size_t)G * Nmat), SS((size_t)G * Nmat * 2);
            for (int g = 0; g < G; ++g)
                for (int m = 0; m < Nmat; ++m) {
                    const double Sm = d.Sb[g + (size_t)(G + 1) * m] / d.dE[g];
                    const double Sp = d.Sb[(g + 1) + (size_t)(G + 1) * m] / d.dE[g];
                    if (!bfp) sigt[(size_t)g * Nmat + m] += 2.0 * Sp;
                    K[(size_t)g * Nmat + m] = Sp + Sm;
                    SS[((size_t)g * Nmat + m) * 2] = Sm;
                    SS[((size_t)g * Nmat + m) * 2 + 1] = Sp;
                }
            if (bfp) { d_csdS = dalloc<double>(SS.size()); upload(d_csdS, SS, st); }
            else { d_csdK = dalloc<double>(K.size()); upload(d_csdK, K, st); }
            dE.assign(d.dE, d.dE + G);
        }
        d_sigt = dalloc<double>(sigt.size()); upload(d_sigt, sigt, st);
        auto S = [&](int m, int gi, int gf, int64_t l) -> double {
            if (l > d.Ls) return 0.0;
            return d.sigs[m + (size_t)Nmat * (gi + (size_t)G * (gf + (size_t)G * l))];
        };
        // solver types 4-6 have no in-group scattering source, 4-5 none from other groups
        // (sn_inner_pass.jl:22, sn_flux.jl:176)
        const int stype = solver_type_of(d);
        const bool within = stype < 4, across = stype != 4 && stype != 5;
        std::vector<double> sigw((size_t)G * Nmat * Nact), sigo((size_t)G * G * Nmat * Nact);
        for (int g = 0; g < G; ++g)
            for (int m = 0; m < Nmat; ++m)
                for (int a = 0; a < Nact; ++a) {
                    sigw[((size_t)g * Nmat + m) * Nact + a] = within ? S(m, g, g, d.pl[act[a]]) : 0.0;
                    for (int gi = 0; gi < G; ++gi)
                        sigo[(((size_t)g * G + gi) * Nmat + m) * Nact + a] = across ? S(m, gi, g, d.pl[act[a]]) : 0.0;
                }
        if (has_fp(d)) {
            // angular Fokker-Planck source T[g, mat] ℳ Φ (fokker_planck_source.jl:26-36) on the
            // active moments: ℳ row-major [Nact][Nact], T [G][Nmat]
            fp = true;
            std::vector<double> Ma((size_t)Nact * Nact), Tg((size_t)G * Nmat);
            for (int k = 0; k < Nact; ++k)
                for (int j = 0; j < Nact; ++j) Ma[(size_t)k * Nact + j] = d.Mfp[act[k] + (size_t)Np * act[j]];
            for (int g = 0; g < G; ++g)
                for (int m = 0; m < Nmat; ++m) Tg[(size_t)g * Nmat + m] = d.T[g + (size_t)G * m];
            d_fpM = dalloc<double>(Ma.size()); upload(d_fpM, Ma, st);
            d_fpT = dalloc<double>(Tg.size()); upload(d_fpT, Tg, st);
            d_fpq = dalloc<double>((size_t)Nact * Nc);
        }
        if (d.Mem) {
            // magnetic-field source ℳ_EM[g] Φ (sn_inner_pass.jl:26-32) on the active moments:
            // [G][Nact][Nact] row-major
            em = true;
            std::vector<double> Me((size_t)G * Nact * Nact);
            for (int g = 0; g < G; ++g)
                for (int k = 0; k < Nact; ++k)
                    for (int j = 0; j < Nact; ++j)
                        Me[((size_t)g * Nact + k) * Nact + j] = d.Mem[g + (size_t)G * (act[k] + (size_t)Np * act[j])];
            d_emM = dalloc<double>(Me.size()); upload(d_emM, Me, st);
            d_emq = dalloc<double>((size_t)Nact * Nc);
        }
        d_sigw = dalloc<double>(sigw.size()); upload(d_sigw, sigw, st);
        d_sigo = dalloc<double>(sigo.size()); upload(d_sigo, sigo, st);
        // volume source: spatial moment 1 of each cell, [g][Kc * cell], rearranged on the device from
        // the description's [cell][g] (a host transpose of G x Ncell values dominated the set-up of
        // large requests)
        d_qvol = dalloc<double>((size_t)G * Nc);
        {
            void* raw = dev_alloc((size_t)G * Ncell * sizeof(double));
            try {
                RT_CHECK(hipStreamSynchronize(st));
                RT_CHECK(hipMemcpy(raw, d.qvol, (size_t)G * Ncell * sizeof(double), hipMemcpyHostToDevice));
                if (Kc > 1) RT_CHECK(hipMemsetAsync(d_qvol, 0, (size_t)G * Nc * sizeof(double), st));
                const size_t n = (size_t)G * Ncell;
                hipLaunchKernelGGL(qvol_layout_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_qvol,
                                   (const double*)raw, G, (long)Ncell, Kc, (long)Nc);
                RT_CHECK(hipGetLastError());
                RT_CHECK(hipStreamSynchronize(st));
            } catch (...) {
                dev_free(raw);
                throw;
            }
            dev_free(raw);
        }
        // ---- angular operators (row-major device copies)
        // k-major device copies: MT[a][n] = M[n, act a], DT[n][p] = D[p, n], DTact[n][a] = D[act a, n]
        std::vector<double> Mact((size_t)Nd * Nact), D((size_t)Np * Nd), Dact((size_t)Nact * Nd);
        for (int a = 0; a < Nact; ++a)
            for (int n = 0; n < Nd; ++n) Mact[(size_t)a * Nd + n] = d.M[n + (size_t)Nd * act[a]];
        for (int n = 0; n < Nd; ++n)
            for (int p = 0; p < Np; ++p) D[(size_t)n * Np + p] = d.D[p + (size_t)Np * n];
        for (int n = 0; n < Nd; ++n)
            for (int a = 0; a < Nact; ++a) Dact[(size_t)n * Nact + a] = d.D[act[a] + (size_t)Np * n];
        d_Mact64 = dalloc<double>(Mact.size()); upload(d_Mact64, Mact, st);
        d_D64 = dalloc<double>(D.size()); upload(d_D64, D, st);
        d_Dact64 = dalloc<double>(Dact.size()); upload(d_Dact64, Dact, st);
        auto tof = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
        d_Mact32 = dalloc<float>(Mact.size()); upload(d_Mact32, tof(Mact), st);
        d_D32 = dalloc<float>(D.size()); upload(d_D32, tof(D), st);
        d_Dact32 = dalloc<float>(Dact.size()); upload(d_Dact32, tof(Dact), st);
        d_act = dalloc<int>(Nact); upload(d_act, act, st);
        // ---- exit-face half-range D columns (sn_inner_pass.jl:256-317 selection)
        if (state_mode == RT_STATE_FAITHFUL || has_bc || has_ss) {
            for (int ax = 0; ax < ndims; ++ax) {
                std::vector<double> Dex((size_t)Nd * Nps, 0.0);
                std::vector<int> ef(Nd, 0);
                const double* cos_ = ax == 0 ? d.mu : (ax == 1 ? d.eta : d.xi);
                const int fm = 2 * ax, fp = 2 * ax + 1;
                for (int n = 0; n < Nd; ++n) {
                    ef[n] = cos_[n] >= 0 ? 1 : 0;
                    int f = -1, col = -1;
                    if (d.n_to_np[fm][n] != 0) { f = fm; col = (int)d.n_to_np[fm][n] - 1; }
                    else if (d.n_to_np[fp][n] != 0) { f = fp; col = (int)d.n_to_np[fp][n] - 1; }
                    if (f >= 0)
                        for (int p = 0; p < Nps; ++p) Dex[(size_t)n * Nps + p] = d.Dsurf[f][p + (size_t)Nps * col];
                }
                d_Dexit[ax] = dalloc<double>(Dex.size()); upload(d_Dexit[ax], Dex, st);
                d_exitf[ax] = dalloc<int>(Nd); upload(d_exitf[ax], ef, st);
                // directions leaving through each face of the axis, in index order (the order in
                // which sn_inner_pass.jl:322-329 accumulates them)
                std::vector<int> lst(2 * (size_t)Nd, 0);
                for (int f = 0; f < 2; ++f) {
                    int c = 0;
                    for (int n = 0; n < Nd; ++n)
                        if (ef[n] == f) lst[(size_t)f * Nd + c++] = n;
                    excnt[ax][f] = c;
                }
                d_exlist[ax] = dalloc<int>(lst.size()); upload(d_exlist[ax], lst, st);
            }
            // ---- reflective / periodic faces (sn_sweep_2D.jl:70-99 and the x/z analogues): the
            // entry half-range M row of every direction and the face whose exit moments feed it
            if (has_bc || has_ss) {
                const size_t npf[3] = {(size_t)Ny * Nz, (size_t)Nx * Nz, (size_t)Nx * Ny};
                for (int ax = 0; ax < 3; ++ax) {
                    std::vector<double> Me((size_t)Nd * std::max(Nps, 1), 0.0);
                    std::vector<int> src(Nd, -1), sface(Nd, -1);
                    if (ax < ndims) {
                        const double* cos_ = ax == 0 ? d.mu : (ax == 1 ? d.eta : d.xi);
                        const int fm = 2 * ax, fp = 2 * ax + 1;
                        int ndp[2] = {0, 0};
                        for (int k = 0; k < 2; ++k)
                            for (int n = 0; n < Nd; ++n) ndp[k] += d.n_to_np[2 * ax + k][n] != 0;
                        for (int n = 0; n < Nd; ++n) {
                            const bool fwd = cos_[n] >= 0;  // enters through the minus face
                            const int code = d.bc[fwd ? fm : fp];
                            int f = -1, row = -1;
                            if (d.n_to_np[fm][n] != 0) { f = fm; row = (int)d.n_to_np[fm][n] - 1; }
                            else if (d.n_to_np[fp][n] != 0) { f = fp; row = (int)d.n_to_np[fp][n] - 1; }
                            if (f < 0) continue;
                            if (code != 0) src[n] = fwd ? (code == 1 ? 0 : 1) : (code == 1 ? 1 : 0);
                            if (d.ssrc[fwd ? fm : fp]) sface[n] = fwd ? 0 : 1;
                            const int nd_f = ndp[f - fm];
                            for (int p = 0; p < Nps; ++p) Me[(size_t)n * Nps + p] = d.Msurf[f][row + (size_t)nd_f * p];
                        }
                        // surface sources of the axis: [2 faces][G][Nps][npos], moments beyond
                        // Np_source zero
                        if (d.ssrc[fm] || d.ssrc[fp]) {
                            const size_t npa = npf[ax];
                            grp_ss.resize(G, 0);
                            for (int k = 0; k < 2; ++k)
                                if (d.ssrc[fm + k])
                                    for (int g = 0; g < G; ++g)
                                        for (size_t i = 0; i < (size_t)d.Np_source * npa && !grp_ss[g]; ++i)
                                            grp_ss[g] = d.ssrc[fm + k][(size_t)g * d.Np_source * npa + i] != 0.0;
                            std::vector<double> q((size_t)2 * G * Nps * npa, 0.0);
                            for (int k = 0; k < 2; ++k) {
                                const double* in = d.ssrc[fm + k];
                                if (!in) continue;
                                for (int g = 0; g < G; ++g)
                                    for (int p = 0; p < d.Np_source; ++p)
                                        for (size_t ps = 0; ps < npa; ++ps)
                                            q[(((size_t)k * G + g) * Nps + p) * npa + ps] =
                                                in[((size_t)g * d.Np_source + p) * npa + ps];
                            }
                            d_ssrc[ax] = dalloc<double>(q.size()); upload(d_ssrc[ax], q, st);
                        }
                    }
                    d_sface[ax] = dalloc<int>(Nd); upload(d_sface[ax], sface, st);
                    d_Mentry[ax] = dalloc<double>(Me.size()); upload(d_Mentry[ax], Me, st);
                    d_insrc[ax] = dalloc<int>(Nd); upload(d_insrc[ax], src, st);
                    d_bin[ax] = dalloc<double>((size_t)Nd * npf[ax] * fm[ax]);
                    k_fill(d_bin[ax], 0.0, (size_t)Nd * npf[ax] * fm[ax], st);  // axes beyond ndims stay zero
                    bin_total += ax < ndims ? (int64_t)Nd * (int64_t)npf[ax] * fm[ax] : 0;
                }
            }
        }
        // ---- fluxes and work buffers
        d_phi = dalloc<double>((size_t)G * Np * Nc);
        k_fill(d_phi, 0.0, (size_t)G * Np * Nc, st);
        d_qlout = dalloc<double>((size_t)Nact * Nc);
        d_psi = dalloc<double>((size_t)Nd * Nc);
        if (csd) {
            d_psiE_in = dalloc<double>(nE());
            d_psiE_out = dalloc<double>(nE());
            k_fill(d_psiE_in, 0.0, nE(), st);
            k_fill(d_psiE_out, 0.0, nE(), st);
        }
        for (int ax = 0; ax < 3; ++ax) d_bex[ax] = dalloc<double>(std::max<size_t>(1, (size_t)Nd * (ax < ndims ? npos[ax] : 1)));
        d_bscratch = dalloc<double>(std::max<size_t>(1, nb));
        d_pd = dalloc<PassDesc>(1);
        // ---- work tables
        auto gemm_tasks = [&](int rows, int& ntask) {
            std::vector<GemmTask> t;
            const int tiles = ceil_div(Nc, GEMM_CELLS), rg = ceil_div(rows, GEMM_ROWS);
            for (int i = 0; i < tiles * rg; ++i) {
                int tile, r;
                gemm_task_order(i, tiles, rg, tile, r);
                t.push_back({0, tile, r, 0});
            }
            ntask = (int)t.size();
            GemmTask* p = dalloc<GemmTask>(t.size());
            upload(p, t, st);
            return p;
        };
        d_mg = gemm_tasks(Nd, n_mg);
        d_dgf = gemm_tasks(Np, n_dgf);
        d_dga = gemm_tasks(Nact, n_dga);
        std::vector<SweepTask> sw(Nd);
        for (int n = 0; n < Nd; ++n) sw[n] = {0, n};
        d_sw = dalloc<SweepTask>(Nd); upload(d_sw, sw, st);
        const int P = Ny * Nz;
        rows_per_lane = sweep_rows_per_lane(ndims, Ny);
        for (int pr = 0; pr < 2; ++pr)
            planes_ok[pr] = ndims == 3 && sweep_planes_shape(pr, Ny, Nz, pl_R[pr], pl_ZW[pr], pl_waves[pr]);
        scan_cpl = ndims == 3 ? sweep_scan_cpl3d(Nx) : sweep_scan_cpl(Nx);
        // the 1D/2D row-scan kernel addresses a direction's angular flux through one buffer
        // resource, whose byte count is 32-bit
        if (ndims <= 2 && (uint64_t)Nc * 8u > 0xffffffffull) scan_cpl = 0;
        // the 3D one's buffer offsets of a direction stay below 2^30 bytes
        if (ndims == 3 && (uint64_t)Nc * 8u > 0x40000000ull) scan_cpl = 0;
        scan_waves = sweep_scan_waves(ndims, Ny, Nz, Nx);
        if (split_scan_offered(d)) {
            d_sclaim = dalloc<unsigned long long>(Nd);
            d_shand = dalloc<unsigned long long>(split_scan_words(d));
            RT_CHECK(hipMemsetAsync(d_sclaim, 0, (size_t)Nd * 8, st));
            RT_CHECK(hipMemsetAsync(d_shand, 0, split_scan_words(d) * 8, st));
        }
        // RADIANT_HIP_SWEEP selects another sweep kernel for testing: s = scan (default), r = wave
        // rows / planes, g = pencil kernel
        sweep_kind = 0;
        if (const char* ev = std::getenv("RADIANT_HIP_SWEEP")) {
            if (ev[0] == 'r') sweep_kind = 1;
            if (ev[0] == 'g') { sweep_kind = 2; rows_per_lane = -1; planes_ok[0] = planes_ok[1] = false; }
        }
        if ((has_bc || has_ss) && !ho_K && !(sweep_kind == 0 && scan_cpl > 0 &&
                        sweep_scan_lds_bytes(P64, ndims, Nx, Ny, Nz, Nmat) <= 160 * 1024))
            throw Error(RT_EINVAL, "reflective/periodic boundaries and surface sources need the row-scan sweep "
                                   "(Nx <= 512 and its LDS face buffers within 160 KiB)");
        sw_ppt = sweep_pencils_per_thread(P);
        sw_block = std::min(1024, ((ceil_div(P, sw_ppt) + 63) / 64) * 64);
        for (int pr = 0; pr < 3; ++pr) {
            // FP64 passes are always used (residuals, reconstruction); low precision only in mixed modes
            const bool needed = (pr == P64) || (precision != RT_PREC_FP64);
            int nbuf = 2;
            size_t lds = sweep_lds_bytes(pr, Nx, Ny, Nz, Nmat, nbuf);
            if (lds > 160 * 1024) { nbuf = 1; lds = sweep_lds_bytes(pr, Nx, Ny, Nz, Nmat, nbuf); }
            if (lds > 160 * 1024 && needed)
                throw Error(RT_EINVAL, "sweep face buffers exceed the 160 KiB of LDS for this grid");
            sw_nbuf[pr] = nbuf;
            sw_lds[pr] = lds;
        }
        // ---- reductions
        rw.partial = dalloc<double>(RED_BLOCKS);
        rw.scal = dalloc<double>(128);
        // pinned: 128 scalar slots, then the two Gram-Schmidt items of a classical Arnoldi step
        RT_CHECK(hipHostMalloc((void**)&h_scal, 128 * sizeof(double) + 2 * sizeof(GsItem), hipHostMallocDefault));
        d_gsi = dalloc<GsItem>(2);
        d_gspart = dalloc<double>((size_t)GS_BPI_MAX * GS_KMAX);
        d_Vptr = dalloc<double*>(64);
        d_y = dalloc<double>(64);
        RT_CHECK(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        for (void* p : allocs) dev_free(p);
        allocs.clear();
        if (h_scal) (void)hipHostFree(h_scal);
        if (own_stream) (void)hipStreamDestroy(st);
        throw;
    }
}

Problem::~Problem() {
    (void)hipStreamSynchronize(st);
    for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
    for (auto& kv : pool) dev_free(kv.second);
    for (void* p : allocs) dev_free(p);
    if (h_scal) (void)hipHostFree(h_scal);
    if (own_stream) (void)hipStreamDestroy(st);
}

int64_t Problem::device_bytes(const rt_problem_desc& d) {
    const int64_t Nc = (int64_t)d.Nx * d.Ny * d.Nz * (ho_order(d) ? ho_order(d) : 1);
    const int Nact_est = (d.sigs && d.pl && d.Ls >= 0) ? (int)active_moments(d).size() : d.Np;
    int64_t b = 0;
    b += Nc * 4 + (int64_t)d.G * Nc * 8;                        // mat, qvol
    b += (int64_t)d.G * d.Np * Nc * 8;                          // fluxes
    b += (int64_t)Nact_est * Nc * 8;                            // qlout
    b += (int64_t)d.Nd * Nc * 8;                                // psi
    b += (int64_t)3 * d.Nd * d.Np * 12;                         // operators
    // face moments: at most the product of the other two orders
    const int64_t ke = d.csd ? std::max(d.order_e, 1) : 1;
    const int64_t fmax = ho_order(d) && (d.ndims >= 2 || ke > 1)
                             ? std::max<int64_t>(1, ke * std::max(d.order_x, 1) * std::max(d.order_y, 1) * std::max(d.order_z, 1))
                             : 1;
    const int64_t nposs = fmax * ((int64_t)d.Ny * d.Nz + (d.ndims >= 2 ? (int64_t)d.Nx * d.Nz : 0) +
                                  (d.ndims >= 3 ? (int64_t)d.Nx * d.Ny : 0));
    b += (int64_t)d.Nd * nposs * 8 + 2LL * d.Nps * nposs * 8;
    if (ho_order(d) && d.ndims >= 2 && nposs * 8 > 160 * 1024) b += (int64_t)d.Nd * nposs * 8;  // walker faces in HBM
    const bool faithful = d.precision == RT_PREC_FP64 && d.state_mode == RT_STATE_FAITHFUL;
    bool nonvoid = false;
    for (int f = 0; f < 2 * d.ndims && d.bc; ++f) nonvoid = nonvoid || d.bc[f] != 0;
    const int64_t S = faithful ? (int64_t)d.Np * Nc + 2LL * d.Nps * nposs
                               : (int64_t)Nact_est * Nc + (nonvoid ? 2LL * d.Nps * nposs : 0);
    int nvec = 8;
    if (d.accel == RT_ACCEL_GMRES) nvec = d.restart + 6;
    if (d.accel == RT_ACCEL_ANDERSON) nvec = 2 * d.depth + 8;
    if (d.precision != RT_PREC_FP64) nvec += 6;
    b += (int64_t)nvec * S * 8;
    if (d.precision != RT_PREC_FP64) b += (int64_t)d.Nd * Nc * 8;  // angular flux of c = T(0) kept by batches
    if (d.csd) b += 2LL * d.Nd * Nc * 8;                              // energy-face angular fluxes (at most Nc per direction)
    if (d.Mem) b += (int64_t)Nact_est * Nc * 8 + (int64_t)d.G * Nact_est * Nact_est * 8;  // magnetic-field source
    if (ho_order(d) && !d.adaptive) {  // the general walkers' cell-system scratch
        const int kx = std::max(1, d.order_x), ky = d.ndims >= 2 ? std::max(1, d.order_y) : 1,
                  kz = d.ndims >= 3 ? std::max(1, d.order_z) : 1, ke = d.csd && d.order_e > 1 ? d.order_e : 1;
        const bool fc = d.partial_coupling == 0;
        if (d.csd && !sweep_bfp_supported(d.ndims, ke, kx, ky, kz, fc))
            b += (int64_t)sweep_hog_bfp_scratch(d.Nd, ke, kx, ky, kz, fc) * 8;
        else if (!d.csd && d.ndims >= 2 &&
                 !(d.ndims == 2 ? sweep_2d_ho_supported(kx, ky, fc) : sweep_3d_ho_supported(kx, ky, kz, fc)))
            b += (int64_t)sweep_hog_scratch(d.Nd, kx, ky, kz, fc) * 8;
    }
    if (has_fp(d)) b += (int64_t)Nact_est * (Nc + Nact_est) * 8;     // Fokker-Planck source, ℳ
    b += (int64_t)std::min<int64_t>((int64_t)d.G * d.Np * Nc, std::max<int64_t>(FLUX_STAGE, (int64_t)d.G * d.Np)) * 8;
    if (split_scan_offered(d)) b += ((int64_t)d.Nd + (int64_t)split_scan_words(d)) * 8;
    return b;
}

double* Problem::vec(const std::string& name, size_t n) {
    const std::string key = name + "#" + std::to_string(n);
    auto it = pool.find(key);
    if (it != pool.end()) return it->second;
    void* p = dev_alloc(n * sizeof(double));
    pool[key] = (double*)p;
    return (double*)p;
}

double Problem::dot(const double* a, const double* b, size_t n) {
    k_dot(rw, 0, a, b, n, st);
    RT_CHECK(hipMemcpyAsync(h_scal, rw.scal, sizeof(double), hipMemcpyDeviceToHost, st));
    RT_CHECK(hipStreamSynchronize(st));
    return h_scal[0];
}

// ------------------------------------------------------------------------------------------
// Transport passes
// ------------------------------------------------------------------------------------------
void Problem::fill_geometry(PassDesc& pd, int g) {
    std::memset(&pd, 0, sizeof(pd));
    pd.ndims = ndims; pd.Nx = Nx; pd.Ny = Ny; pd.Nz = Nz; pd.Nc = Nc; pd.Nd = Nd; pd.Nmat = Nmat;
    pd.mat = d_mat; pd.dx = d_dx; pd.dy = d_dy; pd.dz = d_dz; pd.mu = d_mu; pd.eta = d_eta; pd.xi = d_xi;
    pd.sigt = d_sigt + (size_t)g * Nmat;
    pd.Nact = Nact;
    pd.sigw = d_sigw + (size_t)g * Nmat * Nact;
    pd.psi = d_psi;
    pd.bx = d_bex[0]; pd.by = d_bex[1]; pd.bz = d_bex[2];
    pd.Nps = Nps;
    if (fp) { pd.fpT = d_fpT + (size_t)g * Nmat; pd.fpM = d_fpM; }
    em_g = g;
    pd.ho_K = ho_K;
    pd.ho_om = d_omx;
    pd.ho_Kx = ho_Kx; pd.ho_Ky = ho_Ky; pd.ho_fc = ho_fc ? 1 : 0; pd.ho_omy = d_omy;
    pd.fm_x = fm[0]; pd.fm_y = fm[1];
    pd.ho_Kz = ho_Kz; pd.ho_omz = d_omz; pd.fm_z = fm[2];
    pd.hoface = d_hoface;
    pd.sclaim = d_sclaim; pd.shand = d_shand;
    if (bfp) {  // the general slowing-down cell needs S⁻, S⁺ in every pass (homogeneous ones too)
        pd.ho_KE = ho_KE; pd.ho_ome = d_ome; pd.fm_e = fm_e; pd.emap = d_emap;
        pd.csdS = d_csdS + (size_t)g * Nmat * 2;
        awd_inv_dE = 1.0 / dE[g];  // the adaptive estimate's energy h (sn_1D_bfp.jl:52)
    }
}

size_t Problem::ho_lds(int sp) const {
    if (d_hoface) return 0;  // faces in device memory
    if (bfp) return sweep_bfp_lds_bytes(sp, ndims, Nx, Ny, Nz, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc);
    if (ho_K && ndims == 3) return sweep_3d_ho_lds_bytes(sp, Nx, Ny, Nz, ho_Kx, ho_Ky, ho_Kz, ho_fc);
    if (ho_K && ndims == 2) return sweep_2d_ho_lds_bytes(sp, Nx, Ny, ho_Kx, ho_Ky);
    return 0;
}

hipEvent_t Problem::next_event() {
    if (ev_next == ev_pool.size()) {
        hipEvent_t e;
        RT_CHECK(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    return ev_pool[ev_next++];
}

void Problem::collect_kernel_times() {
    if (ev_used.empty()) return;
    RT_CHECK(hipStreamSynchronize(st));
    for (auto& p : ev_used) {
        float ms = 0.f;
        RT_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        ktimes.ms[p.kernel] += ms;
        ktimes.launches[p.kernel] += 1;
        ktimes.ms_p[p.prec][p.kernel] += ms;
        ktimes.n_p[p.prec][p.kernel] += 1;
    }
    ev_used.clear();
    ev_next = 0;
}

void Problem::launch_pass(PassDesc& pd, int prec, bool full_rows) {
    // Fokker-Planck source of a non-zero state: the contraction reads qlout + T ℳ Φ
    const bool fpsrc = pd.fpT && pd.phi_in;
    if (fpsrc) { pd.fp_qin = pd.qlout; pd.fpq = d_fpq; pd.qlout = d_fpq; }
    // magnetic-field source of a non-zero state: then the contraction reads qlout (+ T ℳ Φ) + ℳ_EM Φ
    const bool emsrc = em && pd.phi_in;
    const double* em_qin = pd.qlout;
    if (emsrc) pd.qlout = d_emq;
    k_set_pd(pd, d_pd, st);
    if (pd.bin_x) launch_bndin(d_pd, (int)bin_total, st);  // incoming fluxes of non-void faces
    if (fpsrc) launch_fp_source(d_pd, (long)Nact * Nc, st);
    if (emsrc) launch_em_source(pd.phi_in, pd.phi_rows, d_emM + (size_t)em_g * Nact * Nact, em_qin, d_emq, Nact, Nc, st);
    auto timed = [&](int kernel, auto&& fn) {
        if (!profiling) {
            fn();
            return;
        }
        if (ev_next + 2 > 4096) collect_kernel_times();
        hipEvent_t a = next_event(), b = next_event();
        set_launch_events(a, b);  // recorded at the kernel's own dispatch start and end
        fn();
        set_launch_events(nullptr, nullptr);
        ev_used.push_back({a, b, kernel, prec});
    };
    timed(0, [&] { launch_mgemm(prec, d_pd, d_mg, n_mg, st); });
    if (pd.psiE_in && !bfp) launch_csd_source(d_pd, (long)Nd * Nc, st);  // energy inflow (FP64 passes)
    timed(1, [&] {
        const int sp = prec == PBF16 ? P32 : prec;
        if (adapt) {
            if (sp != P64) throw Error(RT_EINVAL, "the adaptive closure runs FP64 passes only");
            if (bfp && ndims == 1 && ho_K == 4) launch_sweep_1d_bfp_awd1(d_pd, Nd, awd_inv_dE, st);
            else if (bfp) launch_sweep_bfp_awd(d_pd, Nd, awd_inv_dE, ho_lds(P64), st);
            else launch_sweep_awd(d_pd, Nd, ndims, ho_K, ho_lds(P64), st);
        } else if (bfp && hog)
            launch_sweep_hog_bfp(sp, d_pd, Nd, ndims, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc, ho_lds(sp), d_hog, st);
        else if (bfp)
            launch_sweep_bfp(sp, d_pd, Nd, ho_KE, ho_Kx, ho_Ky, ho_Kz, ho_fc, ho_lds(sp), st);
        else if (hog)
            launch_sweep_hog(sp, d_pd, Nd, ndims, ho_Kx, ho_Ky, ndims == 3 ? ho_Kz : 1, ho_fc, ho_lds(sp), d_hog, st);
        else if (ho_K && ndims == 3)
            launch_sweep_3d_ho(sp, d_pd, Nd, ho_Kx, ho_Ky, ho_Kz, ho_fc, ho_lds(sp), st);
        else if (ho_K && ndims == 2)
            launch_sweep_2d_ho(sp, d_pd, Nd, ho_Kx, ho_Ky, ho_fc, sweep_2d_ho_lds_bytes(sp, Nx, Ny, ho_Kx, ho_Ky), st);
        else if (ho_K)
            launch_sweep_1d_ho(sp, d_pd, Nd, ho_K, st);
        else if (sweep_kind == 0 && scan_cpl > 0 && sweep_scan_lds_bytes(sp, ndims, Nx, Ny, Nz, Nmat) <= 160 * 1024)
        {
            const ScanSplitGeom sg{Nx, Ny, Nz};
            launch_sweep_scan(sp, d_pd, d_sw, Nd, scan_cpl, scan_waves, sweep_scan_lds_bytes(sp, ndims, Nx, Ny, Nz, Nmat),
                              st, pd.bin_x != nullptr, d_sclaim ? &sg : nullptr);
        }
        else if (rows_per_lane > 0 && sweep_rows_lds_bytes(sp, Nx, Nmat) <= 160 * 1024)
            launch_sweep_rows(sp, d_pd, d_sw, Nd, rows_per_lane, sweep_rows_lds_bytes(sp, Nx, Nmat), st);
        else if (planes_ok[sp] && sweep_planes_lds_bytes(sp, Nx, Ny, Nz, Nmat, pl_R[sp], pl_waves[sp]) <= 160 * 1024)
            launch_sweep_planes(sp, d_pd, d_sw, Nd, pl_R[sp], pl_ZW[sp], pl_waves[sp],
                                sweep_planes_lds_bytes(sp, Nx, Ny, Nz, Nmat, pl_R[sp], pl_waves[sp]), st);
        else
            launch_sweep(sp, d_pd, d_sw, Nd, sw_ppt, sw_nbuf[prec], sw_block, sw_lds[prec], st);
    });
    if (pd.psiE_out && !bfp) launch_csd_out(d_pd, (long)Nd * Nc, st);
    timed(2, [&] { launch_dgemm(prec, d_pd, full_rows ? d_dgf : d_dga, full_rows ? n_dgf : n_dga, st); });
    if (pd.bout_x) timed(3, [&] { launch_bndmom(d_pd, 1, (int)nb, st); });
    counters.passes += 1;
    if (prec != P64) counters.passes_lo += 1;
    counters.updates_raw += (double)Ncell * Nd;
}

// Faithful state [Φ (Np x Nc) | boundary moments] (sn_one_speed.jl:126-134), FP64.
void Problem::pass_faithful(int g, int prec, int mode, const double* zin, double* zout) {
    PassDesc pd;
    fill_geometry(pd, g);
    pd.Mact = d_Mact64;
    pd.qlout = (mode == MODE_T) ? d_qlout : nullptr;
    pd.phi_in = zin;  // null = zero state
    pd.phi_rows = d_act;
    pd.Nrows = Np;
    pd.Drows = d_D64;
    pd.out = zout;
    pd.zin = (mode == MODE_IMA) ? zin : nullptr;
    if (mode == MODE_T) set_csd(pd, g);
    set_boundary(pd, g, mode, zin, zout, (size_t)Np * Nc);
    (void)prec;
    launch_pass(pd, P64, true);
}

// Boundary part of a pass: the exit-face moments written into zout's boundary part at offset off
// (I - A epilogue from zin's in MODE_IMA; zout null: none), and for non-void faces or surface
// sources the incoming face fluxes, formed from zin's boundary part (zin null: zero) and, in
// inhomogeneous passes, the group's surface sources.
void Problem::set_boundary(PassDesc& pd, int g, int mode, const double* zin, double* zout, size_t off) {
    auto split = [&](const double* b, const double** x, const double** y, const double** z) {
        *x = b;
        *y = ndims >= 2 ? b + 2 * (size_t)Nps * npos[0] : nullptr;
        *z = ndims >= 3 ? b + 2 * (size_t)Nps * (npos[0] + npos[1]) : nullptr;
    };
    if (zout && nb) {
        double* b = zout + off;
        pd.bout_x = b;
        pd.bout_y = ndims >= 2 ? b + 2 * (size_t)Nps * npos[0] : nullptr;
        pd.bout_z = ndims >= 3 ? b + 2 * (size_t)Nps * (npos[0] + npos[1]) : nullptr;
        if (mode == MODE_IMA) split(zin + off, &pd.bzin_x, &pd.bzin_y, &pd.bzin_z);
        pd.Dexit_x = d_Dexit[0]; pd.Dexit_y = d_Dexit[1]; pd.Dexit_z = d_Dexit[2];
        pd.exitf_x = d_exitf[0]; pd.exitf_y = d_exitf[1]; pd.exitf_z = d_exitf[2];
        pd.exlist_x = d_exlist[0]; pd.exlist_y = d_exlist[1]; pd.exlist_z = d_exlist[2];
        for (int ax = 0; ax < 3; ++ax)
            for (int f = 0; f < 2; ++f) pd.excnt[ax][f] = excnt[ax][f];
    }
    if (has_bc || has_ss) {
        pd.Mentry_x = d_Mentry[0]; pd.Mentry_y = d_Mentry[1]; pd.Mentry_z = d_Mentry[2];
        pd.insrc_x = d_insrc[0]; pd.insrc_y = d_insrc[1]; pd.insrc_z = d_insrc[2];
        pd.bin_x = d_bin[0]; pd.bin_y = d_bin[1]; pd.bin_z = d_bin[2];
        if (zin) split(zin + off, &pd.bsin_x, &pd.bsin_y, &pd.bsin_z);  // none for T(0)
        if (has_ss && mode == MODE_T) {  // surface sources enter the inhomogeneous passes only
            const size_t npf[3] = {(size_t)Ny * Nz, (size_t)Nx * Nz, (size_t)Nx * Ny};
            const double** ps[3] = {&pd.ssrc_x, &pd.ssrc_y, &pd.ssrc_z};
            for (int ax = 0; ax < 3; ++ax) {
                *ps[ax] = d_ssrc[ax] ? d_ssrc[ax] + (size_t)g * Nps * npf[ax] : nullptr;
                pd.ssrc_fs[ax] = (long)((size_t)G * Nps * npf[ax]);
            }
            pd.sface_x = d_sface[0]; pd.sface_y = d_sface[1]; pd.sface_z = d_sface[2];
        }
    }
}

// Reduced state [Φ_act (Nact x Nc)]: only the moments that feed back into the next pass.
void Problem::desc_reduced(PassDesc& pd, int g, int prec, int mode, const double* zin, double* zout) {
    fill_geometry(pd, g);
    pd.prec = prec;
    pd.Mact = (prec == P64) ? (const void*)d_Mact64 : (const void*)d_Mact32;
    pd.qlout = (mode == MODE_T) ? d_qlout : nullptr;
    pd.phi_in = zin;
    pd.phi_rows = nullptr;
    pd.Nrows = Nact;
    pd.Drows = (prec == P64) ? (const void*)d_Dact64 : (const void*)d_Dact32;
    pd.out = zout;
    pd.zin = (mode == MODE_IMA) ? zin : nullptr;
    if (mode == MODE_T && prec == P64) set_csd(pd, g);
    // non-void faces: the state's boundary moments follow its Nact x Nc moments
    if (has_bc || has_ss)
        set_boundary(pd, g, mode, has_bc ? zin : nullptr, has_bc ? zout : nullptr, (size_t)Nact * Nc);
}

void Problem::pass_reduced(int g, int prec, int mode, const double* zin, double* zout, void* psi) {
    PassDesc pd;
    desc_reduced(pd, g, prec, mode, zin, zout);
    if (psi) pd.psi = psi;  // the pass's angular flux goes there instead of d_psi
    launch_pass(pd, prec, false);
}

// Final reconstruction pass 𝚽l = T(z) from a reduced state, all Np moments in FP64.
void Problem::desc_reconstruct(PassDesc& pd, int g, const double* zin_reduced) {
    fill_geometry(pd, g);
    pd.Mact = d_Mact64;
    pd.qlout = d_qlout;
    pd.phi_in = zin_reduced;
    pd.phi_rows = nullptr;
    pd.Nrows = Np;
    pd.Drows = d_D64;
    pd.out = d_phi + (size_t)g * Np * Nc;
    pd.zin = nullptr;
    set_csd(pd, g);
    if (has_bc || has_ss) set_boundary(pd, g, MODE_T, has_bc ? zin_reduced : nullptr, nullptr, (size_t)Nact * Nc);
}

// Continuous slowing down in an inhomogeneous pass: the energy inflow psiE_in is a fixed source
// (zero in homogeneous passes, sn_inner_pass.jl:38) and the outgoing energy flux is kept.
void Problem::set_csd(PassDesc& pd, int g) {
    if (!csd) return;
    pd.psiE_in = d_psiE_in;
    pd.psiE_out = d_psiE_out;
    if (bfp) pd.csdS = d_csdS + (size_t)g * Nmat * 2;
    else pd.csdK = d_csdK + (size_t)g * Nmat;
}

void Problem::pass_reconstruct(int g, const double* zin_reduced) {
    PassDesc pd;
    desc_reconstruct(pd, g, zin_reduced);
    launch_pass(pd, P64, true);
}

// RADIANT_HIP_RECON_FOLD=0: the reconstruction sums the angular fluxes with an axpy before its D
// contraction instead of inside it (for comparisons; the results agree to rounding).
static bool fold_off() {
    static const bool off = [] {
        const char* e = std::getenv("RADIANT_HIP_RECON_FOLD");
        return e && e[0] == '0';
    }();
    return off;
}

// The sweep is linear in its angular source and inflows, so the angular flux of T(z) is that of
// A z (the FP64 residual pass on the same z: its inflows come from z's boundary moments) plus
// that of c = T(0) (volume, surface and energy-inflow sources).  𝚽l = D ψ on all rows, and the
// energy outflow of the slowing-down types is 2 ψ - ψE_in.
void Problem::reconstruct_sweep_free(int g, const double* zin_reduced, const double* psi_c, bool have_az) {
    PassDesc pd;
    desc_reconstruct(pd, g, zin_reduced);
    const size_t n = (size_t)Nd * Nc;
    // Without an energy outflow to form, the D contraction sums the two angular fluxes in its
    // operand load (one pass over psi_c instead of an axpy over both); with one, the summed
    // angular flux is formed first, as the outflow needs it.
    const bool fold = have_az && !pd.psiE_out && !fold_off();
    if (fold) {
        pd.psi = d_psi;
        pd.psi_add = psi_c;
    } else {
        if (have_az) k_axpy(1.0, psi_c, (double*)d_psi, n, st);
        else k_copy((double*)d_psi, psi_c, n, st);
        pd.psi = d_psi;
    }
    k_set_pd(pd, d_pd, st);
    if (pd.psiE_out) {
        if (bfp) launch_bfp_out(d_pd, (long)nE(), st);
        else launch_csd_out(d_pd, (long)n, st);
    }
    counters.passes += 1;  // a pass without sweep
    auto dg = [&] {
        if (fold) launch_dgemm_add(d_pd, d_dgf, n_dgf, st);
        else launch_dgemm(P64, d_pd, d_dgf, n_dgf, st);
    };
    if (profiling) {
        if (ev_next + 2 > 4096) collect_kernel_times();
        hipEvent_t a = next_event(), b = next_event();
        set_launch_events(a, b);
        dg();
        set_launch_events(nullptr, nullptr);
        ev_used.push_back({a, b, 2, P64});
    } else {
        dg();
    }
}

// ------------------------------------------------------------------------------------------
// Krylov solvers (tools/gmres.jl, anderson.jl, livolant.jl, bicgstab.jl)
// ------------------------------------------------------------------------------------------
static double rho_estimate(double r0, double r, int64_t it) {  // krylov_state.jl:139-144
    if (it > 1 && r0 > 0.0 && r > 0.0) return std::pow(r / r0, 1.0 / (double)(it - 1));
    return std::numeric_limits<double>::quiet_NaN();
}

// tools/gmres.jl:35-133 -- restarted GMRES(m), modified Gram-Schmidt, Givens rotations.
// The MGS chain runs on the device (each kernel reads its coefficient from a scalar slot and
// reduces the next dot product); the host reads one H column per Arnoldi step.  Two options
// serve the inner solves of the mixed-precision refinement, which need no parity with
// gmres.jl's iteration counts: zero_guess (z = 0, so the first residual is c without a pass) and
// cgs (classical Gram-Schmidt in two passes over the basis instead of the MGS chain).
template <class MV>
void Problem::gmres(size_t n, double* z, MV matvec, const double* c, int m, int64_t maxit_, double tol_,
                    int64_t& iter, double& resid, bool& conv, double& rho, const std::string& tag, bool zero_guess,
                    bool cgs) {
    m = std::max(1, m);
    if (m + 1 > 64) throw Error(RT_EINVAL, "gmres restart too large (max 63)");
    const double bnorm = norm(c, n);
    if (bnorm == 0.0) {
        k_fill(z, 0.0, n, st);
        iter = 0; resid = 0.0; conv = true; rho = std::numeric_limits<double>::quiet_NaN();
        return;
    }
    std::vector<double*> V(m + 1);
    for (int i = 0; i <= m; ++i) V[i] = vec(tag + "V" + std::to_string(i), n);
    double* w = vec(tag + "w", n);
    double* r = vec(tag + "r", n);
    RT_CHECK(hipMemcpyAsync(d_Vptr, V.data(), (m + 1) * sizeof(double*), hipMemcpyHostToDevice, st));
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), gv(m + 1, 0.0);
    auto h = [&](int i, int j) -> double& { return H[i + (size_t)(m + 1) * j]; };
    iter = 0; resid = 1.0; conv = false;
    double resid0 = std::numeric_limits<double>::quiet_NaN();
    bool first = true;
    while (iter < maxit_ && !conv) {
        const bool skip = first && zero_guess;  // z = 0: the initial residual is c, no pass needed
        first = false;
        if (skip) {
            k_copy(r, c, n, st);
        } else {
            matvec(r, z); ++iter;
            k_neg_add(r, c, n, st);  // r = c - r  (state_scale!(-1) ; state_axpy!(1, c))
        }
        const double beta = skip ? bnorm : norm(r, n);
        resid = beta / bnorm;
        if (std::isnan(resid0)) resid0 = resid;
        if (resid < tol_) { conv = true; break; }
        k_copy(V[0], r, n, st);
        k_scale(V[0], 1.0 / beta, n, st);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int k = 0;
        for (int j = 0; j < m; ++j) {
            if (iter >= maxit_) break;
            matvec(w, V[j]); ++iter;
            if (cgs) {
                // classical Gram-Schmidt in two passes over the basis: h = V^T w, then
                // w -= V h with |w|^2 (the item is reused after the step's synchronisation)
                GsItem* h_gsi = (GsItem*)(h_scal + 128);
                const int bpi = gs_blocks((long)n);
                h_gsi[0] = GsItem{(const double* const*)d_Vptr, nullptr, w, rw.scal, (long)n, j + 1, bpi};
                h_gsi[1] = GsItem{(const double* const*)d_Vptr, rw.scal, w, rw.scal + j + 1, (long)n, j + 1, bpi};
                RT_CHECK(hipMemcpyAsync(d_gsi, h_gsi, 2 * sizeof(GsItem), hipMemcpyHostToDevice, st));
                launch_gs(d_gsi, 1, j + 1, bpi, false, false, d_gspart, st);
                launch_gs(d_gsi + 1, 1, j + 1, bpi, true, true, d_gspart, st);
            } else {
                k_dot(rw, 0, w, V[0], n, st);
                for (int i = 0; i <= j; ++i) k_mgs(rw, i, i + 1, w, V[i], i < j ? V[i + 1] : nullptr, n, st);
            }
            k_scale_by_inv_slot(V[j + 1], w, rw.scal, j + 1, n, st);
            RT_CHECK(hipMemcpyAsync(h_scal, rw.scal, (j + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
            RT_CHECK(hipStreamSynchronize(st));
            for (int i = 0; i <= j; ++i) h(i, j) = h_scal[i];
            const double hjp1 = std::sqrt(h_scal[j + 1]);
            h(j + 1, j) = hjp1;
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                h(i, j) = t;
            }
            const double den = std::hypot(h(j, j), h(j + 1, j));
            if (den == 0.0) { cs[j] = 1.0; sn[j] = 0.0; }
            else { cs[j] = h(j, j) / den; sn[j] = h(j + 1, j) / den; }
            h(j, j) = cs[j] * h(j, j) + sn[j] * h(j + 1, j);
            h(j + 1, j) = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            k = j + 1;
            resid = std::fabs(gv[j + 1]) / bnorm;
            if (resid < tol_) { conv = true; break; }
            if (hjp1 == 0.0) break;
        }
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
            double s = gv[i];
            for (int l = i + 1; l < k; ++l) s -= h(i, l) * y[l];
            y[i] = h(i, i) != 0.0 ? s / h(i, i) : 0.0;
        }
        for (int i0 = 0; i0 < k; i0 += 32) {
            const int kk = std::min(32, k - i0);
            Vals32 v{};
            for (int t = 0; t < kk; ++t) v.v[t] = y[i0 + t];
            k_set_vals(v, d_y, kk, st);
            k_lincomb(z, (const double* const*)(d_Vptr + i0), d_y, kk, n, st);
        }
        if (conv) break;
    }
    rho = rho_estimate(resid0, resid, iter);
}

// tools/anderson.jl:38-140 (beta = 1)
template <class FP>
void Problem::anderson(size_t n, double* z, FP fixedpoint, int64_t& iter, double& resid, bool& conv, double& rho) {
    const int m = std::max(1, depth);
    double* g = vec("ag", n);
    double* f = vec("af", n);
    double* fo = vec("afo", n);
    double* go = vec("ago", n);
    std::vector<double*> dF, dG;
    iter = 0; resid = 1.0; conv = false;
    double resid0 = std::numeric_limits<double>::quiet_NaN();
    for (int64_t it = 1; it <= maxit; ++it) {
        fixedpoint(g, z); ++iter;
        k_copy(f, g, n, st);
        k_axpy(-1.0, z, f, n, st);
        const double fnorm = norm(f, n);
        const double sc = std::max(std::max(norm(z, n), norm(g, n)), 1e-16);
        resid = fnorm / sc;
        if (std::isnan(resid0)) resid0 = resid;
        if (resid < tol) { conv = true; break; }
        if (it == 1) {
            k_axpy(1.0, f, z, n, st);
        } else {
            // history of depth m (push!, then popfirst! when longer than m): reuse the oldest buffers
            double *a, *b;
            if ((int)dF.size() < m) {
                a = vec("adF" + std::to_string(dF.size()), n);
                b = vec("adG" + std::to_string(dG.size()), n);
            } else {
                a = dF.front(); b = dG.front();
                dF.erase(dF.begin()); dG.erase(dG.begin());
            }
            dF.push_back(a); dG.push_back(b);
            k_copy(a, f, n, st); k_axpy(-1.0, fo, a, n, st);
            k_copy(b, g, n, st); k_axpy(-1.0, go, b, n, st);
            const int mk = (int)dF.size();
            std::vector<double> Nm((size_t)mk * mk), bb(mk);
            for (int i = 0; i < mk; ++i) {
                bb[i] = dot(dF[i], f, n);
                for (int j = 0; j < mk; ++j) Nm[i + (size_t)mk * j] = dot(dF[i], dF[j], n);
            }
            double tr = 0.0;
            for (int i = 0; i < mk; ++i) tr += Nm[i + (size_t)mk * i];
            const double lam = 1e-12 * (tr / mk + 1e-300);
            for (int i = 0; i < mk; ++i) Nm[i + (size_t)mk * i] += lam;
            // anderson_solve: Gaussian elimination with partial pivoting (anderson.jl:109-140)
            std::vector<double> M = Nm, x = bb;
            for (int kq = 0; kq < mk; ++kq) {
                int p = kq;
                double mx = std::fabs(M[kq + (size_t)mk * kq]);
                for (int i = kq + 1; i < mk; ++i)
                    if (std::fabs(M[i + (size_t)mk * kq]) > mx) { mx = std::fabs(M[i + (size_t)mk * kq]); p = i; }
                if (p != kq) {
                    for (int j = kq; j < mk; ++j) std::swap(M[kq + (size_t)mk * j], M[p + (size_t)mk * j]);
                    std::swap(x[kq], x[p]);
                }
                const double piv = M[kq + (size_t)mk * kq];
                if (piv == 0.0) continue;
                for (int i = kq + 1; i < mk; ++i) {
                    const double fct = M[i + (size_t)mk * kq] / piv;
                    for (int j = kq; j < mk; ++j) M[i + (size_t)mk * j] -= fct * M[kq + (size_t)mk * j];
                    x[i] -= fct * x[kq];
                }
            }
            for (int kq = mk - 1; kq >= 0; --kq) {
                double s = x[kq];
                for (int j = kq + 1; j < mk; ++j) s -= M[kq + (size_t)mk * j] * x[j];
                x[kq] = M[kq + (size_t)mk * kq] != 0.0 ? s / M[kq + (size_t)mk * kq] : 0.0;
            }
            k_axpy(1.0, f, z, n, st);
            // z <- z + f - sum gamma_i dG_i  (the (beta-1) dF term vanishes for beta = 1)
            for (int i = 0; i < mk; ++i) k_axpy(-x[i], dG[i], z, n, st);
        }
        k_copy(fo, f, n, st);
        k_copy(go, g, n, st);
    }
    rho = rho_estimate(resid0, resid, iter);
}

// tools/livolant.jl:115-155 (period < 0: plain source iteration, the "none" option)
template <class FP>
void Problem::livolant(size_t n, double* z, FP fixedpoint, int period, int64_t& iter, double& resid, bool& conv,
                       double& rho) {
    double* g = vec("lg", n);
    double* f = vec("lf", n);
    double* z1 = vec("lz1", n);
    double* z0 = vec("lz0", n);
    k_fill(z1, 0.0, n, st);
    k_fill(z0, 0.0, n, st);
    int nhist = 0;
    iter = 0; resid = 1.0; conv = false;
    double resid0 = std::numeric_limits<double>::quiet_NaN();
    for (int64_t it = 1; it <= maxit; ++it) {
        fixedpoint(g, z); ++iter;
        k_copy(f, g, n, st);
        k_axpy(-1.0, z, f, n, st);
        const double fnorm = norm(f, n);
        const double sc = std::max(norm(g, n), 1e-16);
        resid = fnorm / sc;
        if (std::isnan(resid0)) resid0 = resid;
        if (resid < tol) { k_copy(z, g, n, st); conv = true; break; }
        if (period > 0 && it % period == 0 && nhist >= 2) {
            // livolant_factor (livolant.jl:69-81): e0 = z1 - z0, e1 = g - z1, Δe = e1 - e0
            double* e0 = vec("le0", n);
            double* de = vec("lde", n);
            k_sub(e0, z1, z0, n, st);
            k_sub(de, g, z1, n, st);
            k_axpy(-1.0, e0, de, n, st);
            const double se = dot(e0, de, n), sd = dot(de, de, n);
            double mu = 1.0;
            if (sd != 0.0) { mu = -se / sd; if (mu <= 0) mu = 1.0; }
            k_copy(z, g, n, st);
            k_scale(z, mu, n, st);
            k_axpy(1.0 - mu, z1, z, n, st);
        } else {
            k_copy(z, g, n, st);
        }
        k_copy(z0, z1, n, st);
        k_copy(z1, g, n, st);
        ++nhist;
    }
    rho = rho_estimate(resid0, resid, iter);
}

// tools/bicgstab.jl:30-101
template <class MV>
void Problem::bicgstab(size_t n, double* z, MV matvec, const double* c, int64_t& iter, double& resid, bool& conv,
                       double& rho) {
    const double bnorm = norm(c, n);
    if (bnorm == 0.0) {
        k_fill(z, 0.0, n, st);
        iter = 0; resid = 0.0; conv = true; rho = std::numeric_limits<double>::quiet_NaN();
        return;
    }
    double *r = vec("br", n), *r0 = vec("br0", n), *p = vec("bp", n), *v = vec("bv", n), *s = vec("bs", n),
           *t = vec("bt", n);
    matvec(r, z);
    k_neg_add(r, c, n, st);
    iter = 1;
    resid = norm(r, n) / bnorm;
    const double resid0 = resid;
    conv = false;
    if (resid < tol) { conv = true; rho = std::numeric_limits<double>::quiet_NaN(); return; }
    k_copy(r0, r, n, st);
    double rho_old = 1.0, al = 1.0, om = 1.0;
    const double tiny = 1e-300;
    k_fill(p, 0.0, n, st);
    k_fill(v, 0.0, n, st);
    while (iter < maxit) {
        const double rn = dot(r0, r, n);
        if (std::fabs(rn) < tiny) break;
        const double be = (rn / rho_old) * (al / om);
        k_axpy(-om, v, p, n, st);
        k_scale(p, be, n, st);
        k_axpy(1.0, r, p, n, st);
        matvec(v, p); ++iter;
        const double den = dot(r0, v, n);
        if (std::fabs(den) < tiny) break;
        al = rn / den;
        k_copy(s, r, n, st);
        k_axpy(-al, v, s, n, st);
        if (norm(s, n) / bnorm < tol) {
            k_axpy(al, p, z, n, st);
            resid = norm(s, n) / bnorm;
            conv = true;
            break;
        }
        matvec(t, s); ++iter;
        const double tt = dot(t, t, n);
        if (tt < tiny) break;
        om = dot(t, s, n) / tt;
        k_axpy(al, p, z, n, st);
        k_axpy(om, s, z, n, st);
        k_copy(r, s, n, st);
        k_axpy(-om, t, r, n, st);
        resid = norm(r, n) / bnorm;
        if (resid < tol) { conv = true; break; }
        if (std::fabs(om) < tiny) break;
        rho_old = rn;
    }
    rho = rho_estimate(resid0, resid, iter);
}

// ------------------------------------------------------------------------------------------
// In-group solves
// ------------------------------------------------------------------------------------------
static void print_group(int g, const rt_group_stats& s) {
    std::printf(">>>Group %d has %sconverged ( ϵ = %.4E , N = %lld , ρ = %.4f )\n", g + 1, s.converged ? "" : "not ",
                s.resid, (long long)s.niter, s.rho);
    std::fflush(stdout);
}

// sn_one_speed.jl:81-196 with the state of sn_one_speed.jl:126-134 (faithful) or the active
// moments only (reduced); FP64 passes.
void Problem::solve_group_plain(int g, rt_group_stats& s) {
    const bool faithful = state_mode == RT_STATE_FAITHFUL;
    const size_t n = faithful ? S_f : S_r;
    const int64_t p0 = counters.passes;
    auto T = [&](double* out, const double* in) {
        if (faithful) pass_faithful(g, P64, MODE_T, in, out);
        else pass_reduced(g, P64, MODE_T, in, out);
    };
    auto MV = [&](double* out, const double* in) {
        if (faithful) pass_faithful(g, P64, MODE_IMA, in, out);
        else pass_reduced(g, P64, MODE_IMA, in, out);
    };
    double* z = vec("z", n);
    int64_t niter = 0;
    double resid = 1.0, rho = std::numeric_limits<double>::quiet_NaN();
    bool conv = false;
    if (accel == RT_ACCEL_NONE || accel == RT_ACCEL_LIVOLANT || accel == RT_ACCEL_ANDERSON) {
        // z = state_clone(work): the group flux enters as the initial guess (zero in the single
        // Gauss-Seidel pass), boundary moments zero.
        k_fill(z, 0.0, n, st);
        if (accel == RT_ACCEL_ANDERSON) anderson(n, z, T, niter, resid, conv, rho);
        else livolant(n, z, T, accel == RT_ACCEL_NONE ? -1 : 3, niter, resid, conv, rho);
    } else {
        double* c = vec("c", n);
        T(c, nullptr);  // c = T(0)
        k_fill(z, 0.0, n, st);
        if (accel == RT_ACCEL_GMRES) gmres(n, z, MV, c, restart, maxit, tol, niter, resid, conv, rho, "g");
        else bicgstab(n, z, MV, c, niter, resid, conv, rho);
    }
    // reconstruction pass 𝚽l = T(z) (sn_one_speed.jl:186)
    if (faithful) {
        double* out = vec("recon", n);
        pass_faithful(g, P64, MODE_T, z, out);
        k_copy(d_phi + (size_t)g * Np * Nc, out, (size_t)Np * Nc, st);
    } else {
        pass_reconstruct(g, z);
    }
    s.niter = niter; s.resid = resid; s.rho = rho; s.converged = conv ? 1 : 0;
    s.passes = (int32_t)(counters.passes - p0);
    s.passes_lo = 0; s.refinements = 0;
}

// Mixed precision: GMRES with FP64 Arnoldi on low-precision passes, wrapped in FP64 iterative
// refinement of (I - A) z = c; final 𝚽l = T(z) in FP64.
void Problem::solve_group_mixed(int g, rt_group_stats& s) {
    const size_t n = S_r;
    const int lo = (precision == RT_PREC_MIXED_BF16) ? PBF16 : P32;
    const double inner_tol = inner_tol_for(lo);
    const int max_ref = 12;
    const int64_t p0 = counters.passes, l0 = counters.passes_lo;
    double* c = vec("c", n);
    double* z = vec("z", n);
    double* r = vec("r", n);
    double* dz = vec("dz", n);
    double* t = vec("t", n);
    double* psi_c = vec("psi_c", (size_t)Nd * Nc);
    pass_reduced(g, P64, MODE_T, nullptr, c, psi_c);  // c = T(0) in FP64, its angular flux kept
    const double bnorm = norm(c, n);
    int64_t niter = 0;
    double resid = 0.0;
    bool conv = true;
    int refs = 0;
    bool have_az = false;  // d_psi holds the angular flux of A z at the current z
    k_fill(z, 0.0, n, st);
    if (bnorm > 0.0) {
        k_copy(r, c, n, st);
        resid = 1.0;
        conv = false;
        auto MVlo = [&](double* out, const double* in) { pass_reduced(g, lo, MODE_IMA, in, out); };
        while (true) {
            if (resid < tol) { conv = true; break; }
            if (refs >= max_ref || niter >= maxit) break;
            k_fill(dz, 0.0, n, st);
            int64_t it = 0;
            double rr, rh;
            bool cv;
            gmres(n, dz, MVlo, r, restart, maxit - niter, inner_tol, it, rr, cv, rh, "m", true, true);
            niter += it;
            k_axpy(1.0, dz, z, n, st);
            pass_reduced(g, P64, MODE_IMA, z, t);  // t = z - A z in FP64
            have_az = true;
            k_sub(r, c, t, n, st);                  // r = c - (I - A) z
            resid = norm(r, n) / bnorm;
            ++refs;
        }
    }
    reconstruct_sweep_free(g, z, psi_c, have_az);
    s.niter = niter; s.resid = resid; s.rho = std::numeric_limits<double>::quiet_NaN();
    s.converged = conv ? 1 : 0;
    s.passes = (int32_t)(counters.passes - p0);
    s.passes_lo = (int32_t)(counters.passes_lo - l0);
    s.refinements = refs;
}

void Problem::solve_group(int g, rt_group_stats& s) {
    std::memset(&s, 0, sizeof(s));
    // out-of-group + volume source (sn_flux.jl:175-179)
    launch_outsrc(d_qlout, d_phi, d_sigo, d_qvol, d_mat, d_act, g, G, Nmat, Nact, Np, Nc, st);
    if (csd) {  // the energy flux enters group g rescaled by ΔE_g/ΔE_{g-1}; zero for the first (sn_flux.jl:166,183)
        if (g == 0) k_fill(d_psiE_in, 0.0, nE(), st);
        else {
            k_copy(d_psiE_in, d_psiE_out, nE(), st);
            k_scale(d_psiE_in, dE[g] / dE[g - 1], nE(), st);
        }
    }
    // zero-source shortcut (sn_one_speed.jl:116-120): no volume or scattering source, no surface
    // source and no incoming energy flux in the group
    if (!(has_ss && grp_ss[g]) && !(csd && dot(d_psiE_in, d_psiE_in, nE()) != 0.0) &&
        dot(d_qlout, d_qlout, (size_t)Nact * Nc) == 0.0) {
        if (csd) k_fill(d_psiE_out, 0.0, nE(), st);
        k_fill(d_phi + (size_t)g * Np * Nc, 0.0, (size_t)Np * Nc, st);
        s.niter = 0; s.resid = 0.0; s.rho = std::numeric_limits<double>::quiet_NaN(); s.converged = 1; s.passes = 0;
        if (verbose) print_group(g, s);
        return;
    }
    if (precision == RT_PREC_FP64) solve_group_plain(g, s);
    else solve_group_mixed(g, s);
    if (verbose) print_group(g, s);
}

void Problem::solve(rt_group_stats* stats, rt_solve_info* info) {
    const char* pe = std::getenv("RADIANT_HIP_PROFILE");
    profiling = pe && pe[0] == '1';
    ktimes = KernelTimes{};
    ev_used.clear();
    ev_next = 0;
    hipEvent_t e0, e1;
    RT_CHECK(hipEventCreate(&e0));
    RT_CHECK(hipEventCreate(&e1));
    counters = SolveCounters{};
    // no clearing of the flux: every group's moments are written whole (reconstruction, or the
    // fill of a group without source) before a later group's out-of-group source reads them
    RT_CHECK(hipEventRecord(e0, st));
    for (int g = 0; g < G; ++g) {
        rt_group_stats s;
        solve_group(g, s);
        if (stats) stats[g] = s;
    }
    RT_CHECK(hipEventRecord(e1, st));
    RT_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    RT_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    collect_kernel_times();
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->ms_device = ms;
        info->passes = counters.passes;
        info->passes_lo = counters.passes_lo;
        info->updates_raw = counters.updates_raw;
        info->updates_problem = (double)Ncell * Nd * G;
        if (profiling) {
            info->ms_mgemm = ktimes.ms[0];
            info->ms_sweep = ktimes.ms[1];
            info->ms_dgemm = ktimes.ms[2] + ktimes.ms[3];
            info->ms_krylov = ms - ktimes.ms[0] - ktimes.ms[1] - ktimes.ms[2] - ktimes.ms[3];
        }
    }
}

// Radiant's layout [G, Np, Nc] has the cells outermost, so a range of cells is one contiguous
// piece of the output: the flux is reordered and copied out one range at a time through a
// staging buffer of at most FLUX_STAGE doubles.
// Copy-out through two pinned host buffers of FLUX_STAGE doubles, shared by the process (one
// copy-out at a time): the device reorders chunk k + 1 into Radiant's layout and copies it into
// one pinned buffer while host threads copy chunk k from the other into the caller's memory.  A
// device-to-host copy into pageable memory runs at a fraction of the link's speed.
namespace {
struct PinnedStage {
    std::mutex mu;
    double* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool tried = false;
    bool ready() {
        if (tried) return buf[1] != nullptr;
        tried = true;
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc((void**)&buf[i], Problem::FLUX_STAGE * sizeof(double), hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                if (buf[0]) (void)hipHostFree(buf[0]);
                if (buf[1]) (void)hipHostFree(buf[1]);
                buf[0] = buf[1] = nullptr;
                return false;
            }
        }
        return true;
    }
};
PinnedStage g_flux_stage;

void parallel_copy(double* dst, const double* src, size_t n) {
    const size_t nt = std::max<size_t>(1, std::min<size_t>(8, n / (size_t(1) << 18)));
    if (nt == 1) {
        std::memcpy(dst, src, n * sizeof(double));
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 1; t < nt; ++t) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a < b) th.emplace_back([=] { std::memcpy(dst + a, src + a, (b - a) * sizeof(double)); });
    }
    std::memcpy(dst, src, std::min(per, n) * sizeof(double));
    for (auto& x : th) x.join();
}
}  // namespace

void Problem::get_flux(double* phi_out, hipStream_t s) {
    if (!s) s = st;
    const size_t per_cell = (size_t)G * Np;
    const size_t cells = std::max<size_t>(1, std::min<size_t>(Nc, FLUX_STAGE / per_cell));
    double* tmp = vec("flux_stage", cells * per_cell);
    const size_t nchunks = ((size_t)Nc + cells - 1) / cells;
    std::unique_lock<std::mutex> lk(g_flux_stage.mu);
    if (!g_flux_stage.ready()) {  // no pinned memory: copy into the caller's memory directly
        lk.unlock();
        for (size_t c0 = 0; c0 < (size_t)Nc; c0 += cells) {
            const size_t nc = std::min(cells, (size_t)Nc - c0);
            k_phi_to_radiant(tmp, d_phi, G, Np, Nc, (int)c0, (int)nc, s);
            RT_CHECK(hipMemcpyAsync(phi_out + c0 * per_cell, tmp, nc * per_cell * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        RT_CHECK(hipStreamSynchronize(s));
        return;
    }
    auto issue = [&](size_t k) {
        const size_t c0 = k * cells, nc = std::min(cells, (size_t)Nc - c0);
        k_phi_to_radiant(tmp, d_phi, G, Np, Nc, (int)c0, (int)nc, s);
        RT_CHECK(hipMemcpyAsync(g_flux_stage.buf[k & 1], tmp, nc * per_cell * sizeof(double), hipMemcpyDeviceToHost, s));
        RT_CHECK(hipEventRecord(g_flux_stage.ev[k & 1], s));
    };
    issue(0);
    for (size_t k = 0; k < nchunks; ++k) {
        if (k + 1 < nchunks) issue(k + 1);  // its pinned buffer held chunk k - 1, copied out already
        RT_CHECK(hipEventSynchronize(g_flux_stage.ev[k & 1]));
        const size_t c0 = k * cells, nc = std::min(cells, (size_t)Nc - c0);
        parallel_copy(phi_out + c0 * per_cell, g_flux_stage.buf[k & 1], nc * per_cell);
    }
}

void Problem::get_flux_cutoff(double* out, hipStream_t s) {
    if (!s) s = st;
    const size_t n = (size_t)Np * Nc;
    if (!csd) {
        std::fill(out, out + n, 0.0);
        return;
    }
    double* tmp = vec("cutoff", n);
    k_cutoff(tmp, d_D64, d_psiE_out, Np, Nd, Ncell, Nc / Ncell, fm_e, s);
    RT_CHECK(hipMemcpyAsync(out, tmp, n * sizeof(double), hipMemcpyDeviceToHost, s));
    RT_CHECK(hipStreamSynchronize(s));
}

void Problem::host_pass(int g, int mode, const double* qlout_host, const double* zin, double* zout) {
    if (g < 0 || g >= G) throw Error(RT_EINVAL, "group out of range");
    if (mode < 0 || mode > 2) throw Error(RT_EINVAL, "pass mode must be 0 (T), 1 (A) or 2 (I - A)");
    if (state_mode != RT_STATE_FAITHFUL) throw Error(RT_EINVAL, "rt_problem_pass needs a faithful FP64 handle");
    const size_t n = S_f;
    double* a = vec("hp_in", n);
    double* b = vec("hp_out", n);
    double* tmp = vec("hp_tmp", n);
    RT_CHECK(hipMemcpyAsync(tmp, zin, n * sizeof(double), hipMemcpyHostToDevice, st));
    k_state_moments_in(a, tmp, Np, Nc, st);
    k_copy(a + (size_t)Np * Nc, tmp + (size_t)Np * Nc, nb, st);
    // qlout [Np][Nc] host (Np fastest) -> active rows; the other rows must be zero, as the group
    // loop produces them (their Σs_l vanish and the volume source only feeds p = 1)
    {
        std::vector<bool> is_act(Np, false);
        for (int a2 : act) is_act[a2] = true;
        for (size_t c = 0; c < (size_t)Nc; ++c)
            for (int p = 0; p < Np; ++p)
                if (!is_act[p] && qlout_host[p + (size_t)Np * c] != 0.0)
                    throw Error(RT_EINVAL, "qlout is nonzero on a moment that carries no scattering source");
    }
    double* q = vec("hp_q", (size_t)Np * Nc);
    RT_CHECK(hipMemcpyAsync(tmp, qlout_host, (size_t)Np * Nc * sizeof(double), hipMemcpyHostToDevice, st));
    k_state_moments_in(q, tmp, Np, Nc, st);
    for (int a2 = 0; a2 < Nact; ++a2)
        RT_CHECK(hipMemcpyAsync(d_qlout + (size_t)a2 * Nc, q + (size_t)act[a2] * Nc, Nc * sizeof(double),
                                hipMemcpyDeviceToDevice, st));
    pass_faithful(g, P64, mode == 2 ? MODE_IMA : (mode == 1 ? MODE_A : MODE_T), a, b);
    k_state_moments_out(tmp, b, Np, Nc, st);
    k_copy(tmp + (size_t)Np * Nc, b + (size_t)Np * Nc, nb, st);
    RT_CHECK(hipMemcpyAsync(zout, tmp, n * sizeof(double), hipMemcpyDeviceToHost, st));
    RT_CHECK(hipStreamSynchronize(st));
}

// ---- allocation-free pass (rt_problem_pass_device) ----
// Scratch of the allocation-free pass: the input and output states in the engine's moment-major
// layout and the out-of-group source, 2 S + Np Nc doubles at most NOALLOC_DOUBLES.  Every device
// holds its own copy until the process ends; every call on a device shares it, so calls that
// overlap on one device race.  Host code reaches it only through the address hipGetSymbolAddress
// gives for the device (the array's name is a host-side stand-in, not a device address).
static constexpr size_t NOALLOC_DOUBLES = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

double* noalloc_scratch(size_t* cap) {
    int dev = 0;
    RT_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw Error(RT_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error(RT_EDEVICE, std::string("hipGetSymbolAddress of the pass scratch failed: ") +
                                        hipGetErrorString(e));
        }
        g_noalloc_ptr[dev] = (double*)p;
    }
    *cap = NOALLOC_DOUBLES;
    return g_noalloc_ptr[dev];
}

void Problem::device_pass(int g, int mode, const double* qlout_dev, const double* zin_dev, double* zout_dev,
                          double* scr, size_t cap) {
    if (g < 0 || g >= G) throw Error(RT_EINVAL, "group out of range");
    if (mode < 0 || mode > 2) throw Error(RT_EINVAL, "pass mode must be 0 (T), 1 (A) or 2 (I - A)");
    if (state_mode != RT_STATE_FAITHFUL) throw Error(RT_EINVAL, "rt_problem_pass_device needs a faithful FP64 handle");
    const size_t n = S_f, nphi = (size_t)Np * Nc, need = 2 * n + nphi;
    if (need > cap)
        throw Error(RT_EINVAL, "rt_problem_pass_device: the pass needs " + std::to_string(need) +
                                   " doubles of scratch (2 x state + Np x cells), above its maximum of " +
                                   std::to_string(cap) + "; use rt_problem_pass");
    double* a = scr;
    double* b = scr + n;
    double* q = scr + 2 * n;
    const bool meas = autotune_measuring();
    set_autotune_measuring(false);  // no timing events: tuned or default launch shapes
    try {
        k_state_moments_in(a, zin_dev, Np, Nc, st);
        k_copy(a + nphi, zin_dev + nphi, nb, st);
        // active rows of the out-of-group source; the other rows are not read
        k_state_moments_in(q, qlout_dev, Np, Nc, st);
        for (int a2 = 0; a2 < Nact; ++a2)
            RT_CHECK(hipMemcpyAsync(d_qlout + (size_t)a2 * Nc, q + (size_t)act[a2] * Nc, Nc * sizeof(double),
                                    hipMemcpyDeviceToDevice, st));
        pass_faithful(g, P64, mode == 2 ? MODE_IMA : (mode == 1 ? MODE_A : MODE_T), a, b);
        k_state_moments_out(zout_dev, b, Np, Nc, st);
        k_copy(zout_dev + nphi, b + nphi, nb, st);
        RT_CHECK(hipStreamSynchronize(st));
    } catch (...) {
        set_autotune_measuring(meas);
        throw;
    }
    set_autotune_measuring(meas);
}

void test_contraction(int which, int prec, int M, int K, int N, const double* A, const double* B, double* C) {
    // which 0: psi = A . B through the moment->direction kernel (B enters as the source operand)
    // which 1: C = A . B through the direction->moment kernel
    std::vector<void*> mem;
    auto al = [&](size_t bytes) {
        void* p = nullptr;
        RT_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
        mem.push_back(p);
        return p;
    };
    try {
        const size_t es = prec == P64 ? 8 : 4;
        void* dA = al((size_t)M * K * es);
        std::vector<double> At((size_t)M * K);  // operators are stored k-major on the device
        for (int i = 0; i < M; ++i)
            for (int k = 0; k < K; ++k) At[(size_t)k * M + i] = A[(size_t)i * K + k];
        if (prec == P64) {
            RT_CHECK(hipMemcpy(dA, At.data(), (size_t)M * K * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(At.begin(), At.end());
            RT_CHECK(hipMemcpy(dA, f.data(), f.size() * 4, hipMemcpyHostToDevice));
        }
        double* dC = (double*)al((size_t)M * N * 8);
        void* dpsi = al((size_t)std::max(M, K) * N * 8);
        int* dmat = (int*)al((size_t)N * 4);
        RT_CHECK(hipMemset(dmat, 0, (size_t)N * 4));
        double* dsig = (double*)al((size_t)K * 8);
        RT_CHECK(hipMemset(dsig, 0, (size_t)K * 8));
        PassDesc pd;
        std::memset(&pd, 0, sizeof(pd));
        pd.Nc = N; pd.Nmat = 1; pd.mat = dmat; pd.sigw = dsig;
        std::vector<GemmTask> t;
        const int rows = which == 0 ? M : M;
        for (int r = 0; r < (rows + GEMM_ROWS - 1) / GEMM_ROWS; ++r)
            for (int i = 0; i < (N + GEMM_CELLS - 1) / GEMM_CELLS; ++i) t.push_back({0, i, r, 0});
        GemmTask* dt = (GemmTask*)al(t.size() * sizeof(GemmTask));
        RT_CHECK(hipMemcpy(dt, t.data(), t.size() * sizeof(GemmTask), hipMemcpyHostToDevice));
        PassDesc* dpd = (PassDesc*)al(sizeof(PassDesc));
        if (which == 0) {
            double* dB = (double*)al((size_t)K * N * 8);
            RT_CHECK(hipMemcpy(dB, B, (size_t)K * N * 8, hipMemcpyHostToDevice));
            pd.Nd = M; pd.Nact = K; pd.Mact = dA; pd.qlout = dB; pd.phi_in = nullptr; pd.psi = dpsi;
            RT_CHECK(hipMemcpy(dpd, &pd, sizeof(pd), hipMemcpyHostToDevice));
            launch_mgemm(prec, dpd, dt, (int)t.size(), nullptr);
            RT_CHECK(hipDeviceSynchronize());
            if (prec == P64) {
                RT_CHECK(hipMemcpy(C, dpsi, (size_t)M * N * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> f((size_t)M * N);
                RT_CHECK(hipMemcpy(f.data(), dpsi, f.size() * 4, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < f.size(); ++i) C[i] = f[i];
            }
        } else {
            if (prec == P64) {
                RT_CHECK(hipMemcpy(dpsi, B, (size_t)K * N * 8, hipMemcpyHostToDevice));
            } else {
                std::vector<float> f(B, B + (size_t)K * N);
                RT_CHECK(hipMemcpy(dpsi, f.data(), f.size() * 4, hipMemcpyHostToDevice));
            }
            pd.Nd = K; pd.Nrows = M; pd.Drows = dA; pd.psi = dpsi; pd.out = dC; pd.zin = nullptr;
            RT_CHECK(hipMemcpy(dpd, &pd, sizeof(pd), hipMemcpyHostToDevice));
            launch_dgemm(prec, dpd, dt, (int)t.size(), nullptr);
            RT_CHECK(hipDeviceSynchronize());
            RT_CHECK(hipMemcpy(C, dC, (size_t)M * N * 8, hipMemcpyDeviceToHost));
        }
    } catch (...) {
        for (void* p : mem) (void)hipFree(p);
        throw;
    }
    for (void* p : mem) (void)hipFree(p);
}

}  // namespace rt
