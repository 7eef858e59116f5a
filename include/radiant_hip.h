/*
 * radiant_hip.h -- C ABI of the MI355X (gfx950) backend for Radiant.jl's discrete-ordinates
 * transport solve.
 *
 * The boundary replaces the body of
 *     compute_flux(cross_sections, geometry, solver::SN, source, electromagnetic_field)
 * (src/particle_transport/sn_flux.jl:22-234): Radiant's host code keeps building the
 * quadrature, the Galerkin operators, the closure weights and the cross-section arrays
 * (sn_flux.jl:27-148); the group loop and the in-group solves (sn_flux.jl:150-224, which call
 * sn_one_speed.jl:81-196 and sn_inner_pass.jl:212-332) run on the GPU behind one call.
 *
 * Conventions
 *   - Plain pointers and extents only; arrays are Julia column-major as Radiant builds them.
 *   - The caller owns every host buffer for the duration of a call; the library copies in/out.
 *   - Device memory is owned by the library (a problem handle or a server handle).
 *   - Return value 0 on success, a negative RT_E* status on failure; rt_last_error() gives the
 *     message of the calling thread's last failure (the Julia shim maps it to error()).
 *   - rt_solve is blocking and may be called from several host threads.
 */
#ifndef RADIANT_HIP_H
#define RADIANT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RT_OK 0
#define RT_EINVAL -1      /* invalid argument / unsupported option */
#define RT_EDEVICE -2     /* HIP runtime error */
#define RT_ENOMEM -3      /* device memory budget exceeded */
#define RT_ECANCELLED -4  /* request cancelled (server) */
#define RT_EPENDING -5    /* request not finished yet (server poll) */

/* acceleration of the in-group iteration (structures/SN.jl:44-63, sn_one_speed.jl:151-181) */
#define RT_ACCEL_NONE 0
#define RT_ACCEL_LIVOLANT 1
#define RT_ACCEL_ANDERSON 2
#define RT_ACCEL_GMRES 3
#define RT_ACCEL_BICGSTAB 4

/* precision of the transport passes */
#define RT_PREC_FP64 0          /* plain FP64 (parity mode) */
#define RT_PREC_MIXED_FP32 1    /* FP32 sweep + FP32 M/D, FP64 Arnoldi and iterative refinement */
#define RT_PREC_MIXED_BF16 2    /* as above with BF16 M/D operands (FP32 accumulation) */

/* Krylov state */
#define RT_STATE_FAITHFUL 0     /* all Np moments + boundary moments, as sn_one_speed.jl:126-134 */
#define RT_STATE_REDUCED 1      /* the moments that feed back (Legendre order <= XS order) only */

typedef struct rt_problem_desc {
    /* geometry (Geometry.get_*: dimension, number_of_voxels, voxels_width, material_per_voxel,
       boundary_conditions) */
    int32_t ndims, Nx, Ny, Nz;          /* Ny = Nz = 1 below their dimension */
    const double *dx, *dy, *dz;         /* cell widths per axis */
    const int64_t *mat;                 /* [Nx*Ny*Nz] 1-based material index, x fastest */
    const int32_t *bc;                  /* [2*ndims] X-,X+,Y-,Y+,Z-,Z+ : 0 void, 1 reflective, 2 periodic */
    /* cross sections (sn_flux.jl:69-83) */
    int32_t Nmat, G;
    const double *sigt;                 /* Σtot[G, Nmat] */
    const double *sigs;                 /* Σs[Nmat, G(from), G(to), Ls+1] */
    int32_t Ls;                         /* Legendre slots - 1 of sigs */
    /* volume source: moment p=1 of Source.volume_sources (Source.jl:132-137), [G, Nx*Ny*Nz] */
    const double *qvol;
    /* angular discretisation (sn_flux.jl:44-48) */
    int32_t Nd, Np, Nps;
    const double *mu, *eta, *xi, *w;    /* [Nd] */
    const double *M;                    /* Mn[Nd, Np]  (moment -> direction) */
    const double *D;                    /* Dn[Np, Nd]  (direction -> moment) */
    const int64_t *pl;                  /* [Np] Legendre order of each moment */
    const double *Msurf[6];             /* Mn_surf[f][Nd+, Nps] per face (X-,X+,Y-,Y+,Z-,Z+) */
    const double *Dsurf[6];             /* Dn_surf[f][Nps, Nd+] */
    const int64_t *n_to_np[6];          /* n_to_n⁺[f][Nd], 1-based, 0 = direction not entering f */
    /* solver (SN.get_acceleration, get_gmres_restart, get_anderson_depth,
       get_maximum_iteration, get_convergence_criterion) */
    int32_t accel, restart, depth, maxit;
    double tol;
    /* backend options */
    int32_t precision;                  /* RT_PREC_* */
    int32_t state_mode;                 /* RT_STATE_* (RT_PREC_MIXED_* implies REDUCED) */
    int32_t verbose;                    /* print Radiant's per-group convergence lines */
    /* Surface sources (surface_source.jl:22-408, accumulated as Source.jl:152-190 does and applied
     * in the sweeps' boundary blocks, e.g. sn_sweep_2D.jl:70-99): per face f (x-, x+, y-, y+, z-,
     * z+), NULL or the half-range moments [G][Np_source][npos_f], row-major with the position
     * fastest (npos = Ny*Nz for x faces, Nx*Nz for y faces, Nx*Ny for z faces; positions ordered
     * like the boundary moments: iy + Ny*iz, ix + Nx*iz, ix + Nx*iy).  The first Np_source
     * half-range moments of each face carry the source.  Every precision and state, single
     * problems, lockstep batches and served requests alike.  Callers that do not use surface
     * sources leave these fields zero. */
    const double *ssrc[6];
    int32_t Np_source;
    /* Continuous slowing down (SN solver type "BCSD" with the DD scheme of order 1 in energy;
     * sn_{1,2,3}D_bfp.jl, sn_flux.jl:86-97,166,183-188): 0 = off.  Sb: stopping powers at the group
     * boundaries [G+1][Nmat] column-major (cross_sections.get_boundary_stopping_powers), dE: group
     * widths [G] (get_energy_width).  Single problems, lockstep batches and served requests. */
    int32_t csd;
    const double *Sb;
    const double *dE;
    /* SN solver type (SN.jl get_solver_type; sn_flux.jl:54): 1 BTE, 2 BFP, 3 BCSD, 4 FP, 5 CSD,
     * 6 BFP-EF; 0 = BTE, or BCSD when csd is set.  Types 2-6 need csd = 1.  The caller passes Σtot
     * and Σs as sn_flux.jl:69-112 prepares them (absorption and zero Σs for FP/CSD, + T λ₀ for
     * BFP/FP, minus the l = 0 self-scattering for BFP-EF); the solve leaves the in-group scattering
     * out for types 4-6 and the out-of-group scattering for types 4-5 (sn_inner_pass.jl:22,
     * sn_flux.jl:176).  Types 2 and 4 add the angular Fokker-Planck source T[g, mat] ℳ 𝚽l
     * (fokker_planck_source.jl:26-36): T the momentum transfer [G, Nmat] column-major, Mfp = ℳ
     * [Np, Np] column-major (fokker_planck_scattering_matrix.jl). */
    int32_t solver_type;
    const double *T;
    const double *Mfp;
    /* Spatial closure along x (SN.set_scheme("x", type, order); scheme_weights.jl:25-90,
     * sn_1D_bte.jl:27-64): order_x = 𝒪x spatial Legendre moments per cell (0 or 1: first order, at
     * most 8 in 1D) and omega_x[𝒪x + 1] the closure weights ω of the x axis (NULL: diamond difference of
     * order 1).  Any order or closure other than diamond difference of order 1 works in single
     * problems, lockstep batches and served requests, every precision and state; with the
     * slowing-down term (csd = 1) see order_e below.  The flux moments then carry Nm5 = 𝒪x
     * spatial moments per cell: phi_out is 𝚽l[G, Np, Nm5, Nx] and the qlout of rt_problem_pass is
     * [Np, Nm5 * Nx]; qvol keeps one value per cell (spatial moment 1, Source.jl:135). */
    int32_t order_x;
    const double *omega_x;
    /* The y axis of 2D problems likewise (order_y at most 8, omega_y[𝒪y + 1]; with order_x at most 8
     * and at most 64 moments per cell: up to 12 moments and orders 4 the cell system is held in
     * registers, beyond that it is solved in device scratch of Nd * 64 * (Nm5^2 + Nm5) doubles),
     * and the coupling of the high-order moments (SN.set_is_full_coupling; 0 = full, the default:
     * Nm5 = 𝒪x*𝒪y moments per cell; 1 = 𝒪x + 𝒪y - 1).  The x-face boundary moments then carry 𝒪y
     * moments per position and the y faces 𝒪x (sn_one_speed.jl:92-106: 𝚽x12[Nps, Nm1, 2, Ny], ...),
     * and the faithful state and rt_problem_pass follow that layout. */
    int32_t order_y;
    const double *omega_y;
    int32_t partial_coupling;
    /* The z axis of 3D problems likewise (orders at most 8 per axis in 3D and at most 64 moments per
     * cell, in registers up to orders 3 and 12 moments; Nm5 = 𝒪x*𝒪y*𝒪z moments, or 𝒪x + 𝒪y + 𝒪z - 2 without full coupling; a face carries the
     * moments over the other two axes). */
    int32_t order_z;
    const double *omega_z;
    /* The energy axis of the slowing-down types (csd = 1; SN.set_scheme("E", type, order),
     * scheme_weights.jl with isCSD, flux_{1,2,3}D_BFP of sn_{1,2,3}D_bfp.jl): order_e = 𝒪E energy
     * Legendre moments per cell (0 or 1: first order, at most 2) and omega_e[𝒪E + 1] its closure
     * weights (NULL: diamond difference of order 1).  With any closure other than diamond
     * difference of order 1 on every axis (space or energy), the slowing-down types take the
     * general cell: spatial orders up to 8 per axis, DD or DG on each axis, at most 64 moments per
     * cell (in registers up to 8 moments and orders 4 in 1D / 2 per axis in 2D/3D; beyond that in
     * device scratch of Nd * 64 * (Nm5^2 + Nm5) doubles).  A cell then holds Nm5 moments over (E, x[, y[, z]]), energy
     * fastest (𝒪E*𝒪x*... with full coupling, 𝒪E + 𝒪x + ... - (ndims) without); the faces carry the
     * energy moments too (Nm1 = 𝒪E*𝒪y*𝒪z, ...), and the cutoff flux fills the first
     * Nm4 = 𝒪x*𝒪y*𝒪z moments of each cell (sn_flux.jl:216). */
    int32_t order_e;
    const double *omega_e;
    /* The adaptive weighted diamond closure (SN scheme "AWD"; scheme_weights.jl:32-37, adaptive.jl):
     * 1 makes every spatial axis adaptive, the closure weights of each cell then following from its
     * incoming face fluxes and source moments in every sweep (omega_* are not used).  Orders 1 and 2
     * in 1D (AWD0, AWD1), order 1 on every axis in 2D and 3D (AWD0), orders (2, 2) with full
     * coupling in 2D (the coupled linear AWD1 of constant_linear.jl); with slowing down (csd = 1)
     * order 1 on every axis and in energy, the energy axis adapted too (1/ΔE its h), or orders
     * (2, 2) in (x, E) with full coupling in 1D,
     * precision RT_PREC_FP64 and state_mode RT_STATE_FAITHFUL only: the closure is not linear in the
     * flux, so the mixed-precision solve's sweep-free reconstruction T(z) = A z + T(0) does not hold
     * for it, and the Krylov iterates depend on the state they run over (the faithful state is the
     * reference's).  0: off. */
    int32_t adaptive;
    /* An external magnetic field (Electromagnetic_Field; sn_flux.jl:116-120): the Lorentz-force
     * operator ℳ_EM of electromagnetic_scattering_matrix.jl, [G][Np][Np] laid out as Radiant's
     * ℳ_EM[ig, p, q] array (column-major: g + G*(p + Np*q)); every pass adds ℳ_EM[g] Φ to the
     * source of each cell and spatial moment (sn_inner_pass.jl:26-32).  Its λ₀_EM is zero, so Σtot
     * is unchanged.  Every precision mode and state; the lockstep batch solver does not take such
     * a problem, so rt_solve_batch and the server solve it on its own stream.  NULL: no field. */
    const double *Mem;
} rt_problem_desc;

/* Per-group result of the in-group solve (the values sn_one_speed.jl:189-193 prints). */
typedef struct rt_group_stats {
    int64_t niter;      /* matvecs / fixed-point applications of the solver */
    double resid;       /* final relative residual */
    double rho;         /* spectral-radius estimate (krylov_state.jl:139-144), NaN if undefined */
    int32_t converged;
    int32_t passes;     /* transport passes of the group (all precisions) */
    int32_t passes_lo;  /* of which low-precision passes (mixed modes) */
    int32_t refinements;/* iterative-refinement steps (mixed modes) */
} rt_group_stats;

/* Timing of the last solve on a handle (device events). */
typedef struct rt_solve_info {
    double ms_device;           /* wall time of the solve on the device stream */
    int64_t passes;             /* transport passes, all groups */
    int64_t passes_lo;          /* low-precision passes */
    double updates_raw;         /* sum over passes of Nc*Nd */
    double updates_problem;     /* Nc*Nd*G */
    double ms_sweep, ms_mgemm, ms_dgemm, ms_krylov;  /* per-phase device time (profiling mode) */
    int64_t graph_launches;       /* lockstep rounds run as HIP graph launches (RADIANT_HIP_GRAPHS=1) */
    int64_t graph_instantiations; /* of which needed a new executable graph (the rest updated one) */
} rt_solve_info;

/* ------------------------------------------------------------------------------------------
 * One-shot solve: replaces sn_flux.jl:150-224 and returns 𝚽l[G, Np, 1, Nx, Ny, Nz]
 * (phi_out, caller-allocated, G*Np*Nx*Ny*Nz doubles) and stats_out[G] (may be NULL).
 * ------------------------------------------------------------------------------------------ */
int rt_solve(const rt_problem_desc *p, double *phi_out, rt_group_stats *stats_out);

/* One-shot solve that also returns the flux at the cutoff energy of the slowing-down solver
 * types (sn_flux.jl:151,213-218, stored by flux.add_flux_cutoff at sn_flux.jl:229):
 * 𝚽cutoff[p, 1, Nx, Ny, Nz] = sum_n Dn[p, n] 𝚽E12[n, 1, Nx, Ny, Nz], the energy-face angular flux
 * leaving the last group.  cutoff_out (caller-allocated, Np*Nx*Ny*Nz doubles, may be NULL) is
 * zero-filled when continuous slowing down is off (csd = 0). */
int rt_solve_cutoff(const rt_problem_desc *p, double *phi_out, double *cutoff_out, rt_group_stats *stats_out);

/* Resident problem: upload once, solve many times (inputs stay in HBM). */
typedef struct rt_problem rt_problem;
int rt_problem_create(const rt_problem_desc *p, rt_problem **out);
int rt_problem_solve(rt_problem *h, rt_group_stats *stats_out, rt_solve_info *info);
int rt_problem_get_flux(rt_problem *h, double *phi_out);
int rt_problem_destroy(rt_problem *h);
/* 𝚽cutoff of the last solve on a handle, as rt_solve_cutoff returns it. */
int rt_problem_get_flux_cutoff(rt_problem *h, double *cutoff_out);
/* Device bytes a problem needs (admission control). */
int rt_problem_device_bytes(const rt_problem_desc *p, int64_t *bytes);

/* Kernel-level hook: one transport pass of group g (0-based) on a host state vector laid out
 * as sn_one_speed.jl:126-134 builds it: [𝚽l_g (Np fastest, then cells) | 𝚽x12 | 𝚽y12 | 𝚽z12]
 * with 𝚽x12[Nps, 2, Ny, Nz], 𝚽y12[Nps, 2, Nx, Nz], 𝚽z12[Nps, 2, Nx, Ny].
 * mode 0: T(z) (sn_inner_pass! homogeneous=false), 1: A z (homogeneous=true),
 * 2: z - A z (the GMRES matvec, sn_one_speed.jl:147).  qlout: out-of-group source
 * [Np, Nx*Ny*Nz] (Np fastest).  z_out may alias z_in.  Precision follows the handle. */
int rt_problem_pass(rt_problem *h, int32_t g, int32_t mode, const double *qlout, const double *z_in,
                    double *z_out);
/* Allocation-free variant of rt_problem_pass.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc), e.g. an external Krylov solver applying the transport operator to vectors
 * it keeps on the device.  qlout, z_in and z_out are DEVICE pointers (caller-allocated) in the
 * layouts of rt_problem_pass; the rows of qlout for moments without a scattering source are not
 * read.  Its temporaries live in a fixed device array of 64 MiB per GPU, held until the process
 * ends: 2*S + Np*Nx*Ny*Nz doubles (S = rt_problem_state_length) must fit in 8,388,608, else
 * RT_EINVAL.  Every call on a GPU shares that array, so calls that overlap on one GPU race.
 * Blocking (the handle's stream is synchronised); needs a faithful FP64 handle. */
int rt_problem_pass_device(rt_problem *h, int32_t g, int32_t mode, const double *qlout, const double *z_in,
                           double *z_out);
/* Per-kernel device time of the last solve on a handle, recorded with HIP events on the
 * handle's stream when the environment has RADIANT_HIP_PROFILE=1 (profiling hook, no
 * counterpart in the reference).  ms[12] and launches[12] are indexed [precision][kernel] with
 * precision RT_PREC_* and kernel 0 moment->direction contraction, 1 sweep,
 * 2 direction->moment contraction, 3 boundary moments. */
int rt_problem_kernel_times(rt_problem *h, double *ms, int64_t *launches);
/* State length of the faithful Krylov state of a handle. */
int rt_problem_state_length(rt_problem *h, int64_t *S);

/* Kernel-level test hook (no counterpart in the reference): C[M,N] = A[M,K] . B[K,N], all
 * row-major FP64 on the host, through the moment->direction (which = 0) or the
 * direction->moment (which = 1) contraction kernel at precision prec (RT_PREC_*). */
int rt_test_contraction(int32_t which, int32_t prec, int32_t M, int32_t K, int32_t N, const double *A,
                        const double *B, double *C);

/* ------------------------------------------------------------------------------------------
 * Batched solve: n independent problems advanced together (one sweep launch covers every
 * problem's directions, per-problem Krylov state). phi_out[i] as rt_solve; stats_out[i][G_i].
 * ------------------------------------------------------------------------------------------ */
int rt_solve_batch(int32_t n, const rt_problem_desc *const *p, double *const *phi_out,
                   rt_group_stats *const *stats_out, rt_solve_info *info);

/* Resident batch: n problems uploaded once and solved together (all RT_PREC_MIXED_*: one
 * lockstep launch stream; otherwise one stream per problem). */
typedef struct rt_batch rt_batch;
int rt_batch_create(int32_t n, const rt_problem_desc *const *p, rt_batch **out);
int rt_batch_solve(rt_batch *b, rt_group_stats *const *stats_out, rt_solve_info *info);
int rt_batch_get_flux(rt_batch *b, int32_t i, double *phi_out);
/* 𝚽cutoff of problem i after the last rt_batch_solve, as rt_solve_cutoff returns it. */
int rt_batch_get_flux_cutoff(rt_batch *b, int32_t i, double *cutoff_out);
int rt_batch_destroy(rt_batch *b);
/* Profiling hook (no counterpart in the reference): per-kernel device time of the last
 * rt_batch_solve run with RADIANT_HIP_PROFILE=1, lockstep batches only.  Arrays of 12 indexed
 * [precision][kernel] as rt_problem_kernel_times; updates[] holds the cell*direction updates the
 * launches processed, flops[] the contraction flops and bytes[] their algorithmic HBM bytes
 * (angular flux/source es = 8 or 4 B per value, moments 8 B: M es*Nd*Nc + 8*Nact*Nc, sweep
 * 2*es*Nd*Nc, D es*Nd*Nc + 8*rows*Nc, twice that last term for an I - A pass, per problem pass). */
int rt_batch_kernel_times(rt_batch *b, double *ms, int64_t *launches, double *updates, double *flops,
                          double *bytes);

/* ------------------------------------------------------------------------------------------
 * Serving: continuous batching of fixed-source solves admitted by an HBM budget.
 * ------------------------------------------------------------------------------------------ */
typedef struct rt_server rt_server;
typedef struct rt_server_config {
    int64_t hbm_budget_bytes;   /* device memory the server may hold for requests */
    int32_t max_active;         /* maximum number of requests resident at once */
    int32_t precision;          /* RT_PREC_* applied to every request */
} rt_server_config;
int rt_server_create(const rt_server_config *cfg, rt_server **out);
/* Copies the problem; returns a request id (>= 0) in *req. */
int rt_server_submit(rt_server *s, const rt_problem_desc *p, int64_t *req);
/* RT_EPENDING while queued/running; otherwise the final status: RT_OK with the results,
 * RT_ECANCELLED, or the error of a failed request.  The poll that reports a final status
 * releases the request id, except a finished request polled with phi_out == NULL (a status
 * query that fills stats_out and latency_ms and keeps the flux for a later poll).
 * A finished request keeps its flux moments in device memory, counted against the server's
 * budget, until a poll with phi_out collects them (copied out on a stream of that call) or
 * rt_server_cancel drops them.  Queued requests are admitted in arrival order when a slot is
 * free and their bytes fit beside the running requests and the uncollected results; a request
 * larger than the whole budget fails with RT_ENOMEM.
 * latency_ms is submit -> flux final in device memory. */
int rt_server_poll(rt_server *s, int64_t req, double *phi_out, rt_group_stats *stats_out,
                   double *latency_ms);
/* rt_server_poll that also collects 𝚽cutoff (as rt_solve_cutoff) when cutoff_out is not NULL. */
int rt_server_poll_cutoff(rt_server *s, int64_t req, double *phi_out, double *cutoff_out,
                          rt_group_stats *stats_out, double *latency_ms);
/* Cancels a queued or running request; a finished request that no poll has collected releases
 * its flux moments (their bytes return to the budget at once).  The next poll of the id reports
 * RT_ECANCELLED and releases it. */
int rt_server_cancel(rt_server *s, int64_t req);
int rt_server_destroy(rt_server *s);

const char *rt_last_error(void);
/* Library/device identification string (e.g. "radiant_hip 0.1 gfx950 ..."). */
const char *rt_version(void);

/* Launch-shape autotuner of the sweep and contraction kernels (no counterpart in the reference; a
 * backend tuning hook).  The candidates of a launch compute bit-identical results and differ only
 * in their schedule (rows in flight of the 1D/2D line kernel; pencils per item and waves per
 * workgroup of the 3D scan kernel; k steps in flight of the M and D contractions).  mode 1 (the default, or environment RADIANT_HIP_AUTOTUNE): the first
 * launches of each shape time every candidate and the fastest serves the rest of the process;
 * 0: default candidates only; 2: a new candidate on every launch (identity tests).  Setting the
 * mode forgets what was tuned.  The report writes one line per launch shape (its choice and the
 * fastest time of each candidate), NUL-terminated and truncated to cap bytes; returns the full
 * length. */
int rt_autotune(int32_t mode);
int64_t rt_autotune_report(char *buf, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* RADIANT_HIP_H */
