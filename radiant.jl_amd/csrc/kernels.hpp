// kernels.hpp -- device kernels of the SN transport backend (gfx950) and their launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace rt {

// Per-problem arguments of one transport pass.  Lives in device memory (one entry per problem
// of a batch); the kernels read it through a task table.
struct PassDesc {
    // geometry
    int ndims, Nx, Ny, Nz, Nc, Nd, Nmat;
    const int* mat;               // [Nc] 0-based material
    const double *dx, *dy, *dz;   // cell widths
    const double *mu, *eta, *xi;  // direction cosines [Nd]
    const double* sigt;           // Σt of the current group [Nmat]
    // moment -> direction contraction (with the scattering source fused in the operand load)
    int Nact;                     // moments that can carry a source (Legendre order <= XS order)
    const void* Mact;             // M restricted to the active moments, k-major: MT[Nact][Nd], compute type
    const double* sigw;           // within-group Σs_l of the current group [Nmat][Nact]
    const double* qlout;          // out-of-group + volume source [Nact][Nc]; null = homogeneous pass
    const double* phi_in;         // moments of the input state, row r at phi_in + r*Nc (null = zero state)
    const int* phi_rows;          // [Nact] row of each active moment in phi_in (null = identity)
    // angular flux (source in, flux out, in place) and exit-face angular fluxes
    void* psi;                    // [Nd][Nc] compute type
    void *bx, *by, *bz;           // [Nd][Ny*Nz], [Nd][Nx*Nz], [Nd][Nx*Ny] compute type
    // direction -> moment contraction
    int Nrows;                    // rows of Drows
    const void* Drows;            // D rows, k-major: DT[Nd][Nrows], compute type
    double* out;                  // [Nrows][Nc]
    const double* zin;            // null: out = D psi ; else out = zin - D psi (the I - A matvec)
    // boundary moments (faithful state): [Nps][2][pos] per axis, null = skip
    int Nps;
    double *bout_x, *bout_y, *bout_z;
    const double *bzin_x, *bzin_y, *bzin_z;  // I - A epilogue inputs (null = plain)
    const double *Dexit_x, *Dexit_y, *Dexit_z;  // [Nd][Nps] half-range D column of the exit face
    const int *exitf_x, *exitf_y, *exitf_z;      // [Nd] exit face (0 = minus, 1 = plus)
    const int *exlist_x, *exlist_y, *exlist_z;   // [2][Nd] directions leaving through face f, in index order
    int excnt[3][2];                             // lengths of those lists
    int prec;                                    // Prec of this pass (operators, psi, bx/by/bz)
    // Non-void boundaries (FP64 faithful passes; null = every face void).  The incoming face flux
    // of direction n at position pos of axis a is bin_a[n*npos_a + pos] = sum_p Mentry_a[n][p] *
    // bsin_a[p + Nps*(insrc_a[n] + 2*pos)] (0 when insrc_a[n] < 0), with npos = Ny*Nz, Nx*Nz, Nx*Ny.
    const double *Mentry_x, *Mentry_y, *Mentry_z;  // [Nd][Nps] half-range M row of each direction
    const int *insrc_x, *insrc_y, *insrc_z;        // [Nd] face whose exit moments come back in (-1: none)
    const double *bsin_x, *bsin_y, *bsin_z;        // boundary moments of the input state (null: zero)
    double *bin_x, *bin_y, *bin_z;                 // incoming face fluxes [Nd][npos]
    // surface sources of the current group (inhomogeneous passes; null = none): per axis the
    // moments [2 faces][Nps][npos] at face stride ssrc_fs, and the face each direction enters
    // through when that face carries a source (-1: none).  Their term comes before the boundary one.
    const double *ssrc_x, *ssrc_y, *ssrc_z;
    const int *sface_x, *sface_y, *sface_z;
    long ssrc_fs[3];
    // continuous slowing down (FP64 inhomogeneous passes; null = off): the incoming and outgoing
    // energy-face angular fluxes [Nd][Nc] and the group's (S⁻ + S⁺) per material
    const double* psiE_in;
    double* psiE_out;
    const double* csdK;
    // angular Fokker-Planck term (null = none): T of the group per material, ℳ on the active
    // moments [Nact][Nact], and for a pass with a non-zero state the source it builds,
    // fpq = fp_qin + T[mat] ℳ Φ [Nact][Nc] (fp_qin: the out-of-group source, null = zero)
    const double* fpT;
    const double* fpM;
    const double* fp_qin;
    double* fpq;
    // 1D cells with ho_K spatial Legendre moments and the closure weights ho_om[ho_K + 1] (0 = the
    // diamond-difference kernels of order 1); moment columns are then spatial moment + ho_K * cell
    int ho_K;
    const double* ho_om;
    // 2D cells of orders ho_Kx x ho_Ky (ho_K = Kx*Ky with full coupling, Kx + Ky - 1 without) and the
    // y-axis closure weights ho_omy[ho_Ky + 1]; the x faces then carry fm_x = Ky moments per position
    // and the y faces fm_y = Kx (face arrays [Nd][npos][fm], boundary moments p + Nps*(m + fm*(f +
    // 2*pos))).  fm 0 stands for 1.
    int ho_Kx, ho_Ky, ho_fc;
    const double* ho_omy;
    int fm_x, fm_y;
    // 3D: the z order and weights; faces carry the moments over the other two axes (fm_z: x, y)
    int ho_Kz;
    const double* ho_omz;
    int fm_z;
    // Slowing-down cells of a general closure (ho_KE > 0; csdK then unused): energy order ho_KE and
    // its weights ho_ome[ho_KE + 1]; a cell holds the moments over (E, x, y, z), energy fastest, and
    // the x, y, z faces carry the energy moments too.  The energy face carries fm_e moments over
    // the spatial axes; psiE_in / psiE_out are then [Nd][cells][fm_e].  csdS: the group's S⁻/ΔE and
    // S⁺/ΔE per material [Nmat][2].  emap[e * fm_e + m]: the cell moment of energy moment e and
    // energy-face moment m (-1 when the partial coupling drops it).
    int ho_KE;
    const double* ho_ome;
    int fm_e;
    const double* csdS;
    const int* emap;
    // D contraction of launch_dgemm_add: the angular flux is psi + psi_add (FP64), so a
    // reconstruction T(z) = A z + T(0) sums the two passes' angular fluxes in its operand load
    const double* psi_add;
    // Face moments of the 3D and slowing-down walkers in device memory instead of LDS (grids whose
    // faces exceed the 160 KiB of a workgroup): [Nd][faces] in the pass's compute type, launched
    // with no dynamic LDS.  Null: LDS.
    void* hoface;
    // 3D scan sweeps split by z over several workgroups per direction (null: not offered): a
    // claim counter per direction [Nd] and the handed z-face rows [Nd][SCAN_SPLIT_MAX - 1][Ny][Nx][2]
    // of tagged 64-bit words, both zero before the first launch
    unsigned long long* sclaim;
    unsigned long long* shand;
};

enum Prec { P64 = 0, P32 = 1, PBF16 = 2,
            PMIX = 3 };  // launch level: FP64 and FP32 passes of different problems in one launch

// The next launch of a pass kernel on this host thread (mgemm, sweep, dgemm, bndmom) records start
// and stop at its own dispatch (kernel execution interval); null, null: untimed.
void set_launch_events(hipEvent_t start, hipEvent_t stop);

// Autotuner of the sweep and contraction launch shapes (variants with bit-identical results): mode 0 default
// variants, 1 online tuning (the default; environment RADIANT_HIP_AUTOTUNE), 2 a new variant on
// every launch.  Setting the mode forgets every tuned key.  The report lists each key's choice
// and the fastest time of every candidate.
void set_autotune_mode(int mode);
// Whether the calling host thread's launches may be timed for tuning (default true); threads whose
// launches run beside other streams' work (concurrent sub-batches, served requests) turn it off
// and use the shapes tuned elsewhere or the defaults.
void set_autotune_measuring(bool on);
bool autotune_measuring();
std::string autotune_report();

// Work tiles of the contraction kernels: x = problem, y = 64-cell tile, z = row group.
struct GemmTask { int prob, tile, rgroup, pad; };
struct SweepTask { int prob, dir; };

// Sweep geometry needed by the host to size a launch.
int sweep_pencils_per_thread(int P);
size_t sweep_lds_bytes(int prec, int Nx, int Ny, int Nz, int Nmat, int nbuf);

void launch_mgemm(int prec, const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st);
// FP64 D contraction of (psi + psi_add), one problem (PassDesc::psi_add non-null)
void launch_dgemm_add(const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st);
void launch_sweep(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int ppt, int nbuf,
                  int block, size_t lds, hipStream_t st);
// 1D/2D wave-per-direction sweep (rows per lane 1, 2 or 4; -1 when the grid does not fit).
int sweep_rows_per_lane(int ndims, int Ny);
size_t sweep_rows_lds_bytes(int prec, int Nx, int Nmat);
void launch_sweep_rows(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int rows_per_lane,
                       size_t lds, hipStream_t st);
// 3D workgroup-per-direction sweep: lanes = y rows (R per lane), waves = z slabs of ZW levels.
bool sweep_planes_shape(int prec, int Ny, int Nz, int& R, int& ZW, int& waves);
size_t sweep_planes_lds_bytes(int prec, int Nx, int Ny, int Nz, int Nmat, int R, int waves);
void launch_sweep_planes(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int R, int ZW, int waves,
                         size_t lds, hipStream_t st);
// Row-scan sweep (lanes along x, affine scan of the x-recurrence); cells per lane 1..8.
int sweep_scan_cpl(int Nx);
// 3D scan: cells per lane, with SCAN_PAIR set when each half-wave solves its own pencil
constexpr int SCAN_PAIR = 16;
constexpr int SCAN_PAIR_NX = 64;
int sweep_scan_cpl3d(int Nx);
size_t sweep_scan_lds_bytes(int prec, int ndims, int Nx, int Ny, int Nz, int Nmat);
int sweep_scan_waves(int ndims, int Ny, int Nz, int Nx);
// 3D scan sweeps split by z over several workgroups per direction (PassDesc::sclaim / shand)
constexpr int SCAN_SPLIT_MAX = 4;  // slabs per direction at most: shand holds SCAN_SPLIT_MAX - 1 faces
struct ScanSplitGeom { int Nx, Ny, Nz; };
int scan_split_waves(const ScanSplitGeom& g, int S, int prec, int items);
// bin: the passes take their inflow from PassDesc::bin_* (FP64 only).  sg: the grid of a
// single-problem pass whose PassDesc carries the split buffers (null: no split shapes).
void launch_sweep_scan(int prec, const PassDesc* d_pd, const SweepTask* tasks, int ntasks, int cpl, int waves, size_t lds,
                       hipStream_t st, bool bin = false, const ScanSplitGeom* sg = nullptr);
// incoming face fluxes of non-void boundaries (PassDesc::bin_*), one problem
void launch_bndin(const PassDesc* d_pd, int total, hipStream_t st);
// continuous slowing down, one problem: psi += K[mat] * psiE_in before the sweep; psiE_out = 2 psi -
// psiE_in after it (FP64 angular fluxes [Nd][Nc])
void launch_csd_source(const PassDesc* d_pd, long n, hipStream_t st);
void launch_csd_out(const PassDesc* d_pd, long n, hipStream_t st);
// angular Fokker-Planck source, one problem: fpq[k][c] = fp_qin[k][c] + T[mat c] sum_j ℳ[k][j]
// phi_in[row j][c], n = Nact * Nc
void launch_fp_source(const PassDesc* d_pd, long n, hipStream_t st);
// 1D sweep of a general closure with K = PassDesc::ho_K (1..HO_KMAX) spatial moments per cell, one
// problem, one wave per direction
constexpr int HO_KMAX = 8;
void launch_sweep_1d_ho(int prec, const PassDesc* d_pd, int Nd, int K, hipStream_t st);
// 2D sweep of a general closure of orders Kx, Ky (1..HO2_KMAX each, at most HO2_NM_MAX moments per
// cell), one problem, one workgroup per direction walking the anti-diagonals; lds = face moments of
// every row and column
constexpr int HO2_KMAX = 4;
constexpr int HO2_NM_MAX = 12;
bool sweep_2d_ho_supported(int Kx, int Ky, bool fc);
size_t sweep_2d_ho_lds_bytes(int prec, int Nx, int Ny, int Kx, int Ky);
void launch_sweep_2d_ho(int prec, const PassDesc* d_pd, int Nd, int Kx, int Ky, bool fc, size_t lds, hipStream_t st);
// 3D sweep of a general closure of orders up to HO3_KMAX per axis and at most HO3_NM_MAX moments per
// cell (full coupling: the product of the orders; without: their sum - 2): one workgroup per
// direction walking the diagonal planes; lds = face moments of every pencil of the three axes
constexpr int HO3_KMAX = 3;
constexpr int HO3_NM_MAX = 12;
bool sweep_3d_ho_supported(int Kx, int Ky, int Kz, bool fc);
size_t sweep_3d_ho_lds_bytes(int prec, int Nx, int Ny, int Nz, int Kx, int Ky, int Kz, bool fc);
// Slowing-down sweep with a general closure in space and energy (sweep_bfp.hip): one workgroup per
// direction walks the diagonal planes; orders Ke 1..2, Kx 1..4 in 1D and 1..2 per axis in 2D/3D,
// at most SWEEP_BFP_NM moments per cell.
constexpr int SWEEP_BFP_NM = 8;
int bfp_cell_moments(int Ke, int Kx, int Ky, int Kz, bool fc);
int bfp_face_moments(int axis, int Ke, int Kx, int Ky, int Kz, bool fc);  // axis 0 = E, 1..3 = x, y, z
bool sweep_bfp_supported(int ndims, int Ke, int Kx, int Ky, int Kz, bool fc);
size_t sweep_bfp_lds_bytes(int prec, int ndims, int Nx, int Ny, int Nz, int Ke, int Kx, int Ky, int Kz, bool fc);
void launch_sweep_bfp(int prec, const PassDesc* d_pd, int Nd, int Ke, int Kx, int Ky, int Kz, bool fc, size_t lds,
                      hipStream_t st);
// energy outflow of the slowing-down closure from a pass's angular flux (sweep-free reconstruction):
// psiE_out = w_0 psiE_in + sum_e C_e (-1)^e w_(e+1) psi_(e, m), n = Nd * cells * fm_e entries
void launch_bfp_out(const PassDesc* d_pd, long n, hipStream_t st);
// the timing events set for the next launch (set_launch_events), taken and cleared
void take_launch_events(hipEvent_t* start, hipEvent_t* stop);
void launch_sweep_3d_ho(int prec, const PassDesc* d_pd, int Nd, int Kx, int Ky, int Kz, bool fc, size_t lds,
                        hipStream_t st);
void launch_dgemm(int prec, const PassDesc* d_pd, const GemmTask* tasks, int ntasks, hipStream_t st);
void launch_bndmom(const PassDesc* d_pd, int nprob, int max_items, hipStream_t st);

// Order of the work tiles of one contraction: task i -> (cell tile, row group).  The row groups of
// a cell tile sit XCD_WAYS task indices apart, within a group of XCD_WAYS consecutive tiles, so that
// the workgroups reading the same cells' operands run close in time on the same XCD (workgroups are
// dealt to the XCDs round-robin by index) and share its L2.
constexpr int XCD_WAYS = 8;
__host__ __device__ inline void gemm_task_order(int i, int ntile, int nrg, int& tile, int& rg) {
    const int full = (ntile / XCD_WAYS) * XCD_WAYS;
    if (i < full * nrg) {
        const int grp = i / (XCD_WAYS * nrg), k = i - grp * XCD_WAYS * nrg;
        rg = k / XCD_WAYS;
        tile = grp * XCD_WAYS + (k - rg * XCD_WAYS);
    } else {
        const int j = i - full * nrg, w = ntile - full;
        rg = j / w;
        tile = full + (j - rg * w);
    }
}

// Row groups of the contraction kernels (rows per work tile).
constexpr int GEMM_RT = 4;                      // 16-row MFMA tiles per wave
constexpr int GEMM_CT = 4;                      // 16-cell MFMA tiles per wave
constexpr int GEMM_ROWS = 16 * GEMM_RT;         // rows per work tile
constexpr int GEMM_CELLS = 4 * 16 * GEMM_CT;    // cells per work tile (4 waves)
constexpr int CONTRACT_KMAX = 1024;             // active moments at most (row map held in LDS)

// Out-of-group + volume source of group g (sn_flux.jl:175-179):
//   qlout[a][c] = sum_{gi<g} sigo[g][gi][mat c][a] * phi[gi][act a][c] + (p_a == 0) * qvol[g][c]
void launch_outsrc(double* qlout, const double* phi_all, const double* sigo, const double* qvol, const int* mat,
                   const int* act, int g, int G, int Nmat, int Nact, int Np, int Nc, hipStream_t st);

// Krylov BLAS-1 with deterministic reductions into device scalar slots.
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
struct RedWork { double* partial; double* scal; };   // partial: RED_BLOCKS doubles
void k_dot(const RedWork& rw, int slot, const double* a, const double* b, size_t n, hipStream_t st);
// w -= scal[hslot]*v ; then scal[nslot] = dot(w, vnext) (vnext null: dot(w, w))
void k_mgs(const RedWork& rw, int hslot, int nslot, double* w, const double* v, const double* vnext, size_t n,
           hipStream_t st);
void k_fill(double* x, double v, size_t n, hipStream_t st);
void k_copy(double* dst, const double* src, size_t n, hipStream_t st);
void k_scale(double* x, double a, size_t n, hipStream_t st);
void k_axpy(double a, const double* x, double* y, size_t n, hipStream_t st);          // y += a x
void k_neg_add(double* r, const double* c, size_t n, hipStream_t st);                 // r = -r + c
// dst = src * (1/sqrt(scal[slot])) when scal[slot] > 0 (the slot holds a squared norm)
void k_scale_by_inv_slot(double* dst, const double* src, const double* scal, int slot, size_t n, hipStream_t st);
struct Vals32 { double v[32]; };
void k_set_vals(const Vals32& v, double* dst, int k, hipStream_t st);   // stream-ordered small upload
void k_set_pd(const PassDesc& v, PassDesc* dst, hipStream_t st);       // stream-ordered PassDesc upload
void k_lincomb(double* z, const double* const* V, const double* y, int k, size_t n, hipStream_t st);
void k_sub(double* out, const double* a, const double* b, size_t n, hipStream_t st);   // out = a - b
void k_cast_f64_f32(float* dst, const double* src, size_t n, hipStream_t st);
void k_scatter_state(double* dst, const double* src, int rows, const int* rowmap, int Nc, hipStream_t st);

// Batched helpers (one launch serves every problem of a batch).
enum VecOpCode { VOP_FILL = 0, VOP_COPY = 1, VOP_SCALE_COPY = 2, VOP_NORMALIZE = 3, VOP_SUB = 4, VOP_SUBMULTI = 5,
                 VOP_ADDMULTI = 6, VOP_ADD = 7 };  // VOP_ADD: y = x + x2
struct VecItem {
    int op, k;
    long n;
    double* y;
    const double* x;
    const double* x2;
    const double* const* V;  // device table of k vectors
    const double* coef;      // device array of k coefficients
    const double* slot;      // squared norm for VOP_NORMALIZE
    double alpha;
};
void launch_expand_gemm(const int* meta, int n, GemmTask* out, hipStream_t st);   // meta [prob, ntile, nrg, off]
void launch_expand_sweep(const int* meta, int n, SweepTask* out, hipStream_t st,
                         unsigned* ctr = nullptr);  // meta [prob, ndir, off]; zeroes *ctr when given
// Lockstep line sweeps as a persistent work list: the next mixed-precision line sweep launched by
// launch_sweep_scan from this host thread runs a fixed grid of workgroups that take their tasks in
// order from the counter ctr (zeroed by launch_expand_sweep on the same stream), when
// RADIANT_HIP_WORKLIST=1; null clears it.
void set_sweep_worklist(unsigned* ctr);
// Batched Gram-Schmidt step, one item per problem: optionally (update) y -= sum_i coef_i V_i in
// place, then out_i = V_i . y for i < k, or out_0 = y . y (norm).  Every item is reduced over its
// own bpi blocks in a fixed order, so its result does not depend on the other items.
constexpr int GS_KMAX = 33;      // vectors per item at most
constexpr int GS_BPI_MAX = 256;  // blocks per item at most
struct GsItem {
    const double* const* V;  // device table of k vectors
    const double* coef;      // k coefficients (update only)
    double* y;
    double* out;
    long n;
    int k, bpi;
};
int gs_blocks(long n);   // the block count of an item of length n
void launch_gs(const GsItem* items, int nitems, int kmax, int bpi_max, bool update, bool norm, double* partial,
               hipStream_t st);
void launch_vecops(const VecItem* items, int nitems, int blocks_per_item, hipStream_t st);

// Layout converters of the host boundary (Radiant order <-> device order).
void k_phi_to_radiant(double* dst, const double* phi_all, int G, int Np, int Nc, int c0, int nc, hipStream_t st);
// out[p + Np*c] = sum_n DT[n][p] psiE[n][c] (the cutoff-energy flux moments, Radiant's layout)
// 𝚽cutoff [Np][K*cells] from the energy-face angular flux psiE [Nd][cells][FE]: the first FE of each
// cell's K moments, the others zero
void k_cutoff(double* out, const double* DT, const double* psiE, int Np, int Nd, int Ncell, int K, int FE,
              hipStream_t st);
void k_state_moments_in(double* dst, const double* src, int Np, int Nc, hipStream_t st);   // [c][p] -> [p][c]
void k_state_moments_out(double* dst, const double* src, int Np, int Nc, hipStream_t st);  // [p][c] -> [c][p]

}  // namespace rt
