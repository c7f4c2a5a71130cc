/*
 * aderdg.h -- C-ABI of the MI355X-native ADER-DG element step.
 *
 * The reference solver (dasikasunder/ADER-DG) is a Fortran program whose hot
 * path is a sequence of external subroutines with implicit interfaces called
 * by its driver (Euler/main.f90:24-100).  This library replaces those
 * subroutines with hand-written HIP kernels for gfx950.  Every entry point
 * below names the reference routine(s) it stands in for; the Fortran binding
 * module (ader-dg_amd/fortran/aderdg_hip.f90) declares them with
 * iso_c_binding so that a driver with main.f90's call order runs on the GPU.
 *
 * Conventions
 *  - Host arrays cross the boundary in the reference's Fortran layouts
 *    (Euler/ader_dg.f90:85-93): uh(nVar,N+1,N+1,IMAX,JMAX) etc., float64.
 *  - Device state is owned by the context and stays resident between calls.
 *  - Return 0 on success or a negative ADERDG_E* code; aderdg_last_error()
 *    gives the message (the reference aborts on errors instead, e.g.
 *    Euler/la.f90:39-44; the Fortran binding turns a non-zero code into `stop`).
 *  - One host thread per context.  Calls are asynchronous on the context's
 *    stream except those that return a value computed on the device.
 */
#ifndef ADERDG_H
#define ADERDG_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ADERDG_OK 0
#define ADERDG_EINVAL -1   /* invalid argument or descriptor */
#define ADERDG_EHIP -2     /* HIP runtime error */
#define ADERDG_ESINGULAR -3/* singular operator matrix (Euler/la.f90:39-44) */
#define ADERDG_ESTATE -4   /* call out of order */
#define ADERDG_ECAPACITY -5/* a fixed-size buffer is too small */
#define ADERDG_EFULL -6    /* no room for another solve in a batch arena (admission refused) */

/* 2-D Euler (nVar 4, Euler/), 2-D Baer-Nunziato two-phase flow (nVar 9,
 * Baer-Nunziato/, path-conservative fluxes, transmissive or periodic
 * boundaries) and 2-D linear advection-diffusion (nVar 1,
 * Advection-Diffusion/, viscous LLF fluxes, no limiter) */
enum aderdg_pde { ADERDG_PDE_EULER = 0, ADERDG_PDE_BN = 1, ADERDG_PDE_AD = 2 };
/* Boundary modes.  DMR reproduces the reference at HEAD exactly
 * (Euler/riemann_problem.f90:10-116 and the ghost overrides of
 * Euler/limiter.f90:491-522).  PERIODIC wraps both the face coupling and the
 * limiter neighbour table (Euler/initialize.f90:241-255, boundary code 3).
 * TRANSMISSIVE copies the interior state into the ghost face state
 * (Euler/riemann_problem.f90:50-53); the limiter neighbour table clamps to the
 * cell itself (Euler/initialize.f90:231-237). */
/* Boundaries of the Euler and Baer-Nunziato systems.  DMR: the Double Mach
 * Reflection set of Euler/riemann_problem.f90:10-116 (Euler only).  PERIODIC /
 * TRANSMISSIVE: every side.  SIDES: a code per side in aderdg_desc.bc, the
 * reference's params.ini convention (Euler/params.ini:8-11): */
enum aderdg_bc { ADERDG_BC_DMR = 0, ADERDG_BC_PERIODIC = 1, ADERDG_BC_TRANSMISSIVE = 2, ADERDG_BC_SIDES = 3 };
/*   0 inflow: the ghost state is aderdg_desc.bc_inflow, its flux the state's flux
 *     (the DMR inflow, Euler/riemann_problem.f90:186-193);
 *   1 transmissive: the ghost is the inside face state and flux
 *     (Baer-Nunziato/riemann_problem.f90:26-34);
 *   2 reflective: the inside face state with the normal momentum of every phase
 *     negated, and every flux component but the normal momentum negated (the DMR
 *     wall, Euler/riemann_problem.f90:225-228);
 *   3 periodic: the opposite side (left with right, bottom with top, in pairs).
 * The subcell recompute's ghost subcells follow the same rule: the neighbour
 * table's clamp (or wrap) gives them, then inflow replaces and reflective
 * mirrors them (Euler/limiter.f90:491-522), x sides before y sides. */
enum aderdg_bc_code { ADERDG_BC_INFLOW = 0, ADERDG_BC_TRANSMIT = 1, ADERDG_BC_REFLECT = 2, ADERDG_BC_WRAP = 3 };

typedef struct aderdg_ctx aderdg_ctx;

/* Operator matrices of Euler/initialize.f90:80-213, column-major as in the
 * reference module.  Pointers may be NULL in aderdg_desc.ops, in which case
 * the library builds them with the same algorithm (quadrature.f90, basis.f90,
 * la.f90). */
typedef struct {
    const double *wGPN;    /* (N+1)            Gauss-Legendre weights on [0,1] */
    const double *Kxi;     /* (N+1,N+1)        stiffness                       */
    const double *iK1;     /* (N+1,N+1)        (F1 - Kxi)^-1                   */
    const double *F0;      /* (N+1)            basis at xi = 0                 */
    const double *FLcoeff; /* (N+1)                                            */
    const double *FRcoeff; /* (N+1)                                            */
    const double *uh2lim;  /* (2N+1,N+1)       DG -> subcell averages          */
    const double *lim2uh;  /* (N+1,2N+1)       subcell averages -> DG          */
    const double *uh2lob;  /* (N+1,N+1)        DG -> Gauss-Lobatto values      */
} aderdg_ops;

typedef struct {
    int pde;        /* enum aderdg_pde                                   */
    int N;          /* polynomial degree, 1..4 (Euler/ader_dg.f90:17)    */
    int imax, jmax; /* cells per direction                               */
    double xl[2], xr[2];
    int bc_mode;    /* enum aderdg_bc                                    */
    double gamma;   /* EQN%GAMMA (Euler)                                 */
    double cfl;     /* CFL (Euler/ader_dg.f90:19), 0.9 in the reference  */
    int exact;      /* reserved, must be 0                               */
    int device;     /* HIP device ordinal                                */
    aderdg_ops ops; /* optional host operator matrices                   */
    double bn_eos[4]; /* Baer-Nunziato EQN: GAMMA_S, GAMMA_G, PI_S, PI_G
                         (Baer-Nunziato/initialize.f90:23-26)            */
    double ad_c[2];   /* advection-diffusion EQN%cx, EQN%cy              */
    double ad_a;      /* EQN%a, the diffusivity (Advection-Diffusion/ader_dg.f90:74) */
    int ad_bc[4];     /* bC(1..4) left, right, bottom, top: 1 Neumann, 2 Dirichlet,
                         3 periodic (Advection-Diffusion/riemann_problem.f90:8-113);
                         bc_mode is not used for this system             */
    int ad_fixed_predictor; /* 0: the reference predictor, which evaluates only
                         the viscous flux (Advection-Diffusion/space_time_predictor.f90:78);
                         1: PDEFusedViscConvFlux there                    */
    int ad_pad;
    int bc[4];        /* bc_mode SIDES: left, right, bottom, top codes
                         (enum aderdg_bc_code)                            */
    double bc_inflow[9]; /* conserved ghost state of inflow sides (nVar values) */
} aderdg_desc;

/* Setup: replaces the allocation and operator part of `initialize`
 * (Euler/initialize.f90:80-145).  The initial condition comes through
 * aderdg_upload_uh. */
int aderdg_create(const aderdg_desc *desc, aderdg_ctx **out);
int aderdg_destroy(aderdg_ctx *ctx);
const char *aderdg_last_error(void);
/* 1 if the library was compiled with floating-point contraction off
 * (bit-identical mode), 0 for the fused-multiply-add build */
int aderdg_build_exact(void);

/* Host-only construction of the operator matrices of Euler/initialize.f90:80-213
 * for degree N (no device needed): xi, wGPN (N+1); Kxi, dudx, iK1 (N+1,N+1);
 * F0, FLcoeff, FRcoeff (N+1); uh2lim (2N+1,N+1); lim2uh (N+1,2N+1); uh2lob (N+1,N+1). */
int aderdg_operators(int N, double *xi, double *wGPN, double *Kxi, double *dudx, double *iK1, double *F0,
                     double *FLcoeff, double *FRcoeff, double *uh2lim, double *lim2uh, double *uh2lob);

/* Host-only initial condition at the Gauss nodes, uh(nVar,N+1,N+1,IMAX,JMAX), the
 * point-value projection of Euler/initialize.f90:262-273.  ictype follows the
 * reference catalogue (Euler/params.ini:17-85): 4 isentropic vortex (strength 5
 * on (1,1,1,1), centred in the domain), 5 Lax-Liu 2-D Riemann problem 3 split at
 * the domain centre, 11 Double Mach Reflection (Euler/initialize.f90:311-341). */
int aderdg_initial_condition(int ictype, int N, int imax, int jmax, const double *xl, const double *xr, double gamma,
                             double *uh);

/* Operators the context uses (host copies, same layout as aderdg_ops; any pointer may be NULL) */
int aderdg_get_ops(aderdg_ctx *ctx, double *wGPN, double *Kxi, double *iK1, double *F0, double *FLcoeff,
                   double *FRcoeff, double *uh2lim, double *lim2uh, double *uh2lob);

/* Upload the solution (nVar = 4 Euler, 9 Baer-Nunziato) and reset the limiter to "no troubled cell"
 * (Euler/initialize.f90:160-165).  Seeds the per-cell min/max and time-step
 * candidates that the next aderdg_calc_time_step reduces. */
int aderdg_upload_uh(aderdg_ctx *ctx, const double *uh);
int aderdg_download_uh(aderdg_ctx *ctx, double *uh);
/* Limiter state in host layout: status(IMAX,JMAX), Lh(nVar,2N+1,2N+1,IMAX,JMAX).
 * set: cells with status /= 0 become limited with the given Lh.
 * get: status and recompute maps of the last step, Lh of limited cells (0
 * elsewhere), lmin/lmax(nVar,IMAX,JMAX) of the last step's get_min_max. Any
 * output pointer may be NULL. */
int aderdg_set_limiter(aderdg_ctx *ctx, const int32_t *status, const double *Lh);

int aderdg_get_limiter(aderdg_ctx *ctx, int32_t *status, int32_t *recompute, double *Lh, double *lmin,
                       double *lmax);

/* calc_time_step (Euler/calc_time_step.f90:9-52): dt = min over all nodes of
 * CFL*PNPM(N)/sum_d max|lambda_d|/dx_d, clipped to tend - time.  Synchronous:
 * returns dt in *dt and keeps it on the device for the following calls. */
int aderdg_calc_time_step(aderdg_ctx *ctx, double time, double tend, double *dt);
/* get_min_max + `recompute = 0` + save_olduh (Euler/main.f90:36-40,
 * Euler/limiter.f90:104-171,251-265).  The per-cell pass of get_min_max is
 * produced by the previous step's update, the 3x3 neighbourhood pass runs
 * inside the predictor launch and olduh is a ping-pong buffer, so this call
 * only opens the step. */
int aderdg_limiter_begin(aderdg_ctx *ctx);
/* ader_space_time_predictor + volume_integral over all cells
 * (Euler/space_time_predictor.f90:8-123, Euler/volume_integral.f90:8-36,
 * loops of Euler/main.f90:44-58). */
int aderdg_space_time_predictor(aderdg_ctx *ctx);
/* solve_riemann_problem + surface_integral + element_update + DMP over all
 * cells (Euler/riemann_problem.f90:10-116, Euler/surface_integral.f90:8-39,
 * Euler/element_update.f90:8-30, Euler/limiter.f90:179-245, loops of
 * Euler/main.f90:62-88).  If nrecompute is not NULL the call waits and
 * returns the number of troubled cells (`nRecompute` of Euler/main.f90). */
int aderdg_riemann_surface_update(aderdg_ctx *ctx, int32_t *nrecompute);
/* spread_recompute + allocate_limiter + subcell_recompute + update_limiter
 * (Euler/limiter.f90:329-354,302-322,411-647,272-296).  Troubled cells are
 * compacted into a work list on the device. */
int aderdg_limiter_end(aderdg_ctx *ctx);
/* One whole time step in the order of Euler/main.f90:24-100. */
int aderdg_step(aderdg_ctx *ctx, double time, double tend, double *dt, int32_t *nrecompute);
/* nsteps device-resident steps replayed from one captured hipGraph; time is
 * advanced on the device (time = time + dt, Euler/main.f90:97).  Waits and
 * returns the final time and the last step's dt / nRecompute. */
int aderdg_run(aderdg_ctx *ctx, int nsteps, double time, double tend, double *time_out, double *dt_out,
               int32_t *nrecompute_out);

/* Batch form of ader_space_time_predictor with the reference argument list
 * (Euler/space_time_predictor.f90:8-17) for ncells independent cells, host
 * arrays in reference layout, dt given.  Used to check the kernel against the
 * reference routine cell by cell. */
int aderdg_predictor_cells(aderdg_ctx *ctx, int ncells, double dt, const double *luh, double *lqhi, double *lFhi,
                           double *lqbnd, double *lFbnd);

/* Baer-Nunziato form of aderdg_predictor_cells with the argument list of
 * Baer-Nunziato/space_time_predictor.f90:8-18 (lShi = time-averaged NCP source). */
int aderdg_predictor_cells_bn(aderdg_ctx *ctx, int ncells, double dt, const double *luh, double *lqhi, double *lFhi,
                              double *lShi, double *lqbnd, double *lFbnd);

/* Per-cell forms of the other element routines, with the reference argument lists, for
 * callers that keep the reference's loops over cells (the symbol-compatible Fortran shim
 * ader-dg_amd/fortran/aderdg_percell.f90).  ncells independent cells, host arrays in the
 * reference's per-cell layout repeated cell after cell; synchronous.  Euler and
 * Baer-Nunziato contexts; the context supplies the operators, mesh spacing and equation of
 * state.  In the build without contraction every result is bit-identical to the reference
 * routine's. */

/* volume_integral(lduh, lFhi) (Euler/volume_integral.f90:8-36) and its Baer-Nunziato form
 * volume_integral(lduh, lFhi, lShi) (Baer-Nunziato/volume_integral.f90): lduh (V,n,n) per
 * cell out; lFhi (V,2,n,n) per cell in; lShi (V,n,n) per cell, required for Baer-Nunziato,
 * ignored (may be NULL) for Euler. */
int aderdg_volume_integral_cells(aderdg_ctx *ctx, int ncells, double *lduh, const double *lFhi, const double *lShi);

/* surface_integral(lduh, lFl, lFr, lGb, lGt) (Euler/surface_integral.f90:8-39): lduh (V,n,n)
 * per cell in and out; the left, right, bottom and top face fluxes (V,n) per cell in. */
int aderdg_surface_integral_cells(aderdg_ctx *ctx, int ncells, double *lduh, const double *lFl, const double *lFr,
                                  const double *lGb, const double *lGt);

/* element_update(luh, lduh) (Euler/element_update.f90:8-30) with the module's dt given:
 * lduh is divided by the Gauss weights in place and luh advanced by dt * lduh. */
int aderdg_element_update_cells(aderdg_ctx *ctx, int ncells, double dt, double *luh, double *lduh);

/* DMP(dmpresult, arguh, argLimiter, argmax) (Euler/limiter.f90:179-245; Baer-Nunziato/limiter.f90
 * with its PDEAssurePositivity): dmpresult[c] = 1 where the cell passes (the reference's
 * .true.), 0 where it is troubled; arguh (V,n,n) per cell; lmin, lmax (V) per cell are the
 * argLimiter%lmin / %lmax the reference reads (tLimiter, Euler/ader_dg.f90:73-79, is not
 * C-interoperable: the shim passes those two members). */
int aderdg_dmp_cells(aderdg_ctx *ctx, int ncells, int32_t *dmpresult, const double *arguh, const double *lmin,
                     const double *lmax, double argmax);

/* aderdg_predictor_cells / aderdg_predictor_cells_bn without any device allocation: no
 * counterpart in the reference.  For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc): the device staging of
 * the cells goes through a file-scope array of 16 MiB held on every GPU the library runs on,
 * whose address is resolved when a context is created on that GPU.  Same arguments and
 * results (bit for bit) as the allocating forms; lShi is required for a Baer-Nunziato context
 * and ignored (may be NULL) for Euler.  ncells is limited to
 * aderdg_predictor_cells_noalloc_max(pde, N) (e.g. 3,177 Euler or 1,412 Baer-Nunziato cells at
 * N = 4); more returns ADERDG_ECAPACITY.  All no-allocation calls on one GPU share its array:
 * calls that overlap in time on one GPU (other threads, other contexts) race. */
int aderdg_predictor_cells_noalloc(aderdg_ctx *ctx, int ncells, double dt, const double *luh, double *lqhi,
                                   double *lFhi, double *lShi, double *lqbnd, double *lFbnd);
/* Largest ncells aderdg_predictor_cells_noalloc accepts for (pde, N); 0 for other systems. */
int aderdg_predictor_cells_noalloc_max(int pde, int N);

/* Kernel timing: when enabled, the kernels of each step are bracketed with
 * hipEvents on the context stream.  Returns accumulated milliseconds and
 * launch counts per kernel: [0] dt reduction, [1] predictor, [2] surface
 * integral + update + DMP, [3] spread/compact, [4] subcell recompute,
 * [5] face Riemann fluxes, [6] the 3x3 neighbourhood ranges of get_min_max
 * (the DMP bounds of the step). */
#define ADERDG_TIMING_SLOTS 7
int aderdg_timing_enable(aderdg_ctx *ctx, int on);
int aderdg_timing_get(aderdg_ctx *ctx, double ms[ADERDG_TIMING_SLOTS], int64_t launches[ADERDG_TIMING_SLOTS]);

/* Predictor kernel autotuning.  No counterpart in the reference (a serial CPU
 * program).  The predictor has ADERDG_PREDICTOR_VARIANTS kernel variants with
 * identical results: 0 time-line threads (one thread per spatial node holds its
 * time line), 1/2/3 space-time-node threads in workgroups of 64/128/256 (1 is
 * absent for N=4, whose space-time cell has 125 nodes), 4 time-line threads
 * with a smaller register footprint (right-hand sides kept in the time line,
 * state and operators in LDS), 5 (N<=3, libaderdg_amd.so only) time-line threads
 * with the per-slice contractions, the volume integral and the boundary
 * extrapolation on the FP64 matrix cores, equal to the others within rounding.
 * Advection-diffusion has variant 0 only.  The first context or batch arena of
 * a (device, pde, N, launch mode) times every variant on a synthetic 256x256
 * solve and keeps the fastest for the process (variant 5 competes at N=3 only;
 * below it pads each cell to 4x4 nodes).
 * aderdg_tune_force(pde, N, v) makes contexts and arenas created afterwards use
 * variant v (-1: back to the tuner); the environment variable
 * ADERDG_PREDICTOR_VARIANT=<v> does the same for every (pde, N), and
 * ADERDG_PREDICTOR_VARIANT=<N>:<v>,<N>:<v>... per degree.
 * aderdg_tune_query returns the tuner's choice and the milliseconds per
 * launch it measured (-1 for variants it did not run); persistent selects
 * batch arenas (persistent grids) rather than single contexts. */
#define ADERDG_PREDICTOR_VARIANTS 6
int aderdg_tune_force(int pde, int N, int variant);
int aderdg_tune_query(int device, int pde, int N, int persistent, int *variant,
                      double ms[ADERDG_PREDICTOR_VARIANTS]);
/* The same for either tuned kernel: kernel 0 the predictor (variants as above),
 * kernel 1 the update (ADERDG_UPDATE_VARIANTS variants with identical results:
 * 0 each thread forms the first passes of the subcell and Lobatto maps it needs
 * and LDS atomics reduce the ranges, 1 those passes shared in LDS and the ranges
 * reduced by the cell's threads).  ms holds the variant count of the kernel. */
#define ADERDG_UPDATE_VARIANTS 2
int aderdg_tune_force_kernel(int kernel, int pde, int N, int variant);
int aderdg_tune_query_kernel(int kernel, int device, int pde, int N, int persistent, int *variant, double *ms);

/* Algorithmic floating-point operations of one cell-step of the predictor
 * launch (predictor + volume integral) and of the update launch, counted as
 * + - * / sqrt = 1 (the convention of the oracle's op counter).  For
 * advection-diffusion the predictor count is that of the reference's
 * viscous-only predictor flux; aderdg_flops_predictor_ad(N, 1) counts the
 * fused convective + viscous flux, and the update count is 0 (not modelled). */
double aderdg_flops_predictor(int pde, int N);
double aderdg_flops_predictor_ad(int N, int fused);
double aderdg_flops_update(int pde, int N);

/* ------------------------------------------------------------------------
 * Batched engine.  No counterpart in the reference, which runs one solve per
 * process (Euler/main.f90:1-107); each solve of a batch follows the same
 * per-step routines with its own mesh, degree, boundaries, time and dt
 * (the state of Euler/ader_dg.f90:30-101, one copy per solve).
 *
 * Solves of one (pde, N) share an arena (a group): flat device arrays sized
 * at aderdg_batch_reserve, in which a solve takes contiguous ranges of cells
 * and faces.  One captured graph steps every active solve of every group;
 * the kernels loop over device-resident work tables, so adding or removing
 * solves never changes the graph.  A solve stops after max_steps steps or,
 * with stop_at_tend, once its time reaches tend; it keeps its state until
 * it is removed.  All calls come from one host thread.
 * ------------------------------------------------------------------------ */
typedef struct aderdg_batch aderdg_batch;

/* Initial condition given by a formula, evaluated where the solve lives (on
 * the device for batched solves) at the Gauss-Legendre nodes with the node
 * arithmetic of Euler/initialize.f90:151,262-273.  Parameters p[]:
 *   VORTEX    (Euler)  beta, xc, yc, u0, v0: isentropic vortex on rho = p = 1
 *   QUADRANT  (Euler)  x0, y0, then (rho,u,v,p) of NE, NW, SW, SE
 *   DMR       (Euler)  x_split0, tan(pi/3), (rho,u,v,p) left, (rho,u,v,p) right
 *                      (InitialField, Euler/initialize.f90:311-341)
 *   BN_CIRCLE (Baer-Nunziato) xc, yc, radius, p_in, p_out, rho_s, rho_g,
 *                      phi_s inside, phi_s outside; both phases at rest
 *   BN_DISC   (Baer-Nunziato) xc, yc, radius, then the primitive states
 *                      (rho_s,u_s,v_s,p_s,rho_g,u_g,v_g,p_g,phi_s) inside and
 *                      outside (e.g. the two sides of a Riemann problem of
 *                      Baer-Nunziato/initialize.f90:314-390)
 *   SINE      (advection-diffusion) amplitude, kx, ky: A sin(2 pi kx x) sin(2 pi ky y) */
enum aderdg_ic_kind { ADERDG_IC_VORTEX = 1, ADERDG_IC_QUADRANT = 2, ADERDG_IC_DMR = 3, ADERDG_IC_BN_CIRCLE = 4,
                      ADERDG_IC_BN_DISC = 5, ADERDG_IC_SINE = 6, ADERDG_IC_BN_RIEMANN = 7 };
typedef struct {
    int kind;
    int pad;
    double p[24];
} aderdg_ic;

/* Host evaluation of an initial condition into uh(nVar,N+1,N+1,IMAX,JMAX). */
int aderdg_ic_eval(const aderdg_desc *solve, const aderdg_ic *ic, double *uh);

/* The formula of a reference ICType: Euler 4, 5, 11 (Euler/params.ini:17-85), or the
 * Baer-Nunziato one-dimensional Riemann problems 1, 2, 3, 7
 * (Baer-Nunziato/initialize.f90:314-403; ADERDG_IC_BN_RIEMANN: split at x = p[0],
 * primitive states p[1..9] left, p[10..18] right). */
int aderdg_ic_of_ictype(int pde, int ictype, const double xl[2], const double xr[2], aderdg_ic *out);

/* The troubled-cell seeding at the end of initialize (Euler/initialize.f90:274-303,
 * Baer-Nunziato/initialize.f90:272-301): after the formula initial condition ic has
 * been projected and uploaded, every cell whose projection fails the DMP against its
 * own nodal range with margin 0.1, or the positivity check of its subcell values,
 * starts limited with its subcell averages taken from ic at the subcell barycentres
 * (the reference's uninitialised x0 read as the cell corner).  Euler and
 * Baer-Nunziato contexts. */
int aderdg_seed_limiter(aderdg_ctx *ctx, const aderdg_ic *ic);

typedef struct {
    int device;          /* HIP device ordinal                                   */
    int persist_blocks;  /* workgroups of each table-driven kernel per group;
                            0 = eight per compute unit                           */
} aderdg_batch_desc;

int aderdg_batch_create(const aderdg_batch_desc *desc, aderdg_batch **out);
int aderdg_batch_destroy(aderdg_batch *b);
/* Arena for up to max_cells cells in up to max_solves solves of one (pde, N).
 * Face capacity is sized for solves of at least 8 cells per side. */
int aderdg_batch_reserve(aderdg_batch *b, int pde, int N, int64_t max_cells, int max_solves);
/* The same with a smaller limited-cell pool.  The reference allocates the subcell
 * averages Lh(nVar,2N+1,2N+1) per limited cell (Euler/limiter.f90:302-322,
 * Baer-Nunziato/limiter.f90:301-321, :575) and frees them when the cell is no
 * longer limited (Baer-Nunziato/limiter.f90:271-295); an arena keeps them in a pool of slots, and a
 * solve of C cells owns ceil(limited_fraction * C) of them (aderdg_batch_reserve:
 * limited_fraction = 1, one per cell).  A step that limits more cells than its
 * solve owns ends that solve: its solution is that of the step, and
 * aderdg_batch_error reports ADERDG_ECAPACITY.  0 < limited_fraction <= 1. */
int aderdg_batch_reserve_limited(aderdg_batch *b, int pde, int N, int64_t max_cells, int max_solves,
                                 double limited_fraction);
/* Device bytes a solve of this shape occupies in an arena (admission by HBM footprint). */
int64_t aderdg_solve_footprint(int pde, int N, int imax, int jmax);
/* ... in an arena reserved with aderdg_batch_reserve_limited. */
int64_t aderdg_solve_footprint_limited(int pde, int N, int imax, int jmax, double limited_fraction);
/* Device bytes reserved by all arenas of the batch. */
int64_t aderdg_batch_reserved_bytes(aderdg_batch *b);
/* Add a solve: descriptor (ops must be NULL: built-in operators), initial
 * uh(nVar,N+1,N+1,IMAX,JMAX), start time, end time, step limit (> 0) and
 * whether reaching tend ends it.  *id receives a handle.  ADERDG_EFULL: no
 * free slot or range in the (pde, N) arena (nothing was added). */
int aderdg_batch_add(aderdg_batch *b, const aderdg_desc *solve, const double *uh, double time, double tend,
                     int max_steps, int stop_at_tend, int *id);
/* aderdg_batch_add with the initial state computed on the device from ic
 * (no host array, no transfer). */
int aderdg_batch_add_ic(aderdg_batch *b, const aderdg_desc *solve, const aderdg_ic *ic, double time, double tend,
                        int max_steps, int stop_at_tend, int *id);
/* Enqueue nsteps replays of the step graph (asynchronous). */
int aderdg_batch_step(aderdg_batch *b, int nsteps);
int aderdg_batch_sync(aderdg_batch *b);
/* State of one solve (synchronous): time, last dt, steps taken, still active. */
int aderdg_batch_query(aderdg_batch *b, int id, double *time, double *dt, int *steps, int *active);
/* Handles of solves that have stopped and are not removed yet (synchronous);
 * *count receives how many were written (at most max_ids). */
int aderdg_batch_finished(aderdg_batch *b, int *ids, int max_ids, int *count);
/* Current solution of a solve in the reference layout (synchronous). */
int aderdg_batch_download(aderdg_batch *b, int id, double *uh);
/* Limiter status and recompute flags of a solve's last step (synchronous). */
int aderdg_batch_limiter(aderdg_batch *b, int id, int32_t *status, int32_t *recompute);
/* The whole limiter state of a solve, in the layouts of aderdg_get_limiter: status,
 * recompute, Lh(nVar,2N+1,2N+1,IMAX,JMAX) of limited cells (0 elsewhere) and
 * lmin/lmax(nVar,IMAX,JMAX) of the last step (synchronous; any output may be NULL).
 * Replaces reading Limiter(i,j)%status/%Lh/%lmin/%lmax of module ader_dg
 * (Euler/ader_dg.f90:73-79) for one solve of a batch. */
int aderdg_batch_limiter_state(aderdg_batch *b, int id, int32_t *status, int32_t *recompute, double *Lh,
                               double *lmin, double *lmax);
/* Health of a solve (synchronous, counted on the device): *nonfinite = DOFs of the current
 * solution that are NaN or infinite, *troubled = cells with recompute 1 in the last step
 * (the reference's nRecompute, Euler/main.f90:86-92).  No counterpart in the reference:
 * for monitors of many resident solves that cannot download every solution. */
int aderdg_batch_health(aderdg_batch *b, int id, int64_t *nonfinite, int64_t *troubled);
/* *error = 0, or ADERDG_ECAPACITY when the solve stopped because it limited more
 * cells than its share of the limited-cell pool (synchronous). */
int aderdg_batch_error(aderdg_batch *b, int id, int *error);
/* Free a solve's slot and ranges. */
int aderdg_batch_remove(aderdg_batch *b, int id);
/* Solves placed and cells in use over all arenas. */
int aderdg_batch_usage(aderdg_batch *b, int *solves, int64_t *cells);
/* Kernel timing of batch steps: when enabled, aderdg_batch_step launches the
 * step's kernels one by one between hipEvents instead of replaying the graph.
 * aderdg_batch_timing_get returns the totals of the (pde, N) arena in the
 * slot order of aderdg_timing_get. */
int aderdg_batch_timing_enable(aderdg_batch *b, int on);
int aderdg_batch_timing_get(aderdg_batch *b, int pde, int N, double ms[ADERDG_TIMING_SLOTS],
                            int64_t launches[ADERDG_TIMING_SLOTS]);

/* ------------------------------------------------------------------------
 * Serving loop.  No counterpart in the reference.  A worker thread owns a
 * batch: requests (a solve, a formula initial condition, an end time) wait
 * in arrival order until the (pde, N) arena has room and the admitted
 * footprint stays within hbm_budget, are then added to the running batch
 * (continuous batching: the step graph keeps running while solves join and
 * leave), and are retired when they reach their end time.  Submission and
 * polling are thread-safe; everything on the device is done by the worker.
 * ------------------------------------------------------------------------ */
typedef struct aderdg_server aderdg_server;

typedef struct {
    int device;
    int steps_per_round;  /* step-graph replays between completion checks (0: 2)  */
    int persist_blocks;   /* as in aderdg_batch_desc                              */
    int pad;
    int64_t hbm_budget;   /* bytes of admitted solve footprint (0: arenas only)   */
} aderdg_server_desc;

typedef struct {
    aderdg_desc solve;    /* ops must be NULL                                     */
    aderdg_ic ic;
    double tend;          /* the solve ends once its time reaches tend            */
    int max_steps;        /* and at the latest after this many steps (> 0)        */
    int keep_result;      /* 1: keep the final uh until aderdg_server_result      */
} aderdg_request;

typedef struct {
    int64_t submitted, admitted, completed, failed, queued, running;
    int64_t dof_updates;  /* nVar (N+1)^2 cells x steps of finished requests      */
    int64_t bytes_in_use, peak_bytes;
    double busy_seconds;  /* worker time with solves on the device                */
} aderdg_server_stats;

int aderdg_server_create(const aderdg_server_desc *desc, aderdg_server **out);
/* Arenas, before aderdg_server_start (aderdg_batch_reserve, aderdg_batch_reserve_limited).
 * A request whose solve runs out of limited-cell pool slots finishes with status
 * ADERDG_ECAPACITY. */
int aderdg_server_reserve(aderdg_server *s, int pde, int N, int64_t max_cells, int max_solves);
int aderdg_server_reserve_limited(aderdg_server *s, int pde, int N, int64_t max_cells, int max_solves,
                                  double limited_fraction);
int aderdg_server_start(aderdg_server *s);
/* Queue a request; *ticket identifies it in poll and result. */
int aderdg_server_submit(aderdg_server *s, const aderdg_request *req, int64_t *ticket);
/* Requests finished since the last poll: ticket, seconds from submit to
 * completion, status (0 done, else the error code), steps taken and whether
 * the solve reached tend (1) or stopped at max_steps (0); *count written.
 * Any of the output arrays except tickets may be NULL. */
int aderdg_server_poll(aderdg_server *s, int64_t *tickets, double *latency_s, int *status, int *steps,
                       int *reached_tend, int max, int *count);
/* Final state of a finished request kept with keep_result (released here). */
int aderdg_server_result(aderdg_server *s, int64_t ticket, double *uh, double *time, int *steps);
/* Wait until no request is queued or running (timeout_s <= 0: no limit);
 * ADERDG_ESTATE on timeout. */
int aderdg_server_drain(aderdg_server *s, double timeout_s);
int aderdg_server_stats_get(aderdg_server *s, aderdg_server_stats *st);
/* Stop the worker (queued requests are dropped) and free everything. */
int aderdg_server_destroy(aderdg_server *s);

/* Measured peaks of a device (no counterpart in the reference; the denominators of the
 * benchmark's roofline fractions beside the datasheet's 78.6 TFLOP/s and 8 TB/s):
 * FP64 FMA rate of eight independent chains per thread (TFLOP/s), and HBM bandwidth of a
 * 1 GiB stream read and of a 1 GiB copy (GB/s, read plus write bytes).  Allocates 2 GiB
 * for the duration of the call. */
int aderdg_probe_peaks(int device, double *fp64_tflops, double *hbm_read_gbs, double *hbm_copy_gbs);

#ifdef __cplusplus
}
#endif
#endif
