/*
 * aderdg_oracle.h -- CPU restatement of the reference ADER-DG element step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker against which the HIP path is
 * compared; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links it.
 *
 * All arrays use the reference's Fortran column-major layouts
 * (Euler/ader_dg.f90:85-93): uh(V,n,n,IMAX,JMAX), qbnd/Fbnd(V,4,n,IMAX,JMAX),
 * Fhi(V,2,n,n,...), Fu(V,n,IMAX+1,JMAX), Gu(V,n,IMAX,JMAX+1), with n = N+1.
 * Operator matrices are column-major n x n etc. as in Euler/ader_dg.f90:30-38.
 *
 * Parity pinning: every routine is checked bit-for-bit against the reference
 * routines compiled from their own sources (oracle/_ref) by the CPU tests, and
 * against the committed golden vectors in tests/golden.
 */
#ifndef ADERDG_ORACLE_H
#define ADERDG_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PDE_EULER = 0, ORC_PDE_BN = 1, ORC_PDE_AD = 2 };
enum { ORC_BC_DMR = 0, ORC_BC_PERIODIC = 1, ORC_BC_TRANSMISSIVE = 2, ORC_BC_SIDES = 3 };

typedef struct orc_ctx orc_ctx;

/* create: mesh, degree, boundary mode, EOS; builds the operators of
 * Euler/initialize.f90:80-213 and allocates the state of Euler/initialize.f90:135-145 */
orc_ctx *orc_create(int pde, int N, int imax, int jmax, const double xl[2], const double xr[2],
                    int bc_mode, const double *eos);
void orc_destroy(orc_ctx *c);

/* operators: xi, w, MM, Kxi, dudx, iK1, F0, FLcoeff, FRcoeff, uh2lim, lim2uh, uh2lob, xiLob, wLob, K1 */
void orc_get_ops(const orc_ctx *c, double *xi, double *w, double *mm, double *kxi, double *dudx,
                 double *ik1, double *f0, double *flc, double *frc, double *u2l, double *l2u,
                 double *u2b, double *xlob, double *wlob, double *k1);

void orc_set_scalars(orc_ctx *c, double dt, double time, double tend);
void orc_get_scalars(const orc_ctx *c, double *dt, double *time, double *tend);
void orc_set_uh(orc_ctx *c, const double *uh);
void orc_get_uh(const orc_ctx *c, double *uh);
void orc_set_limiter(orc_ctx *c, const int *status, const double *lh);
void orc_get_limiter(const orc_ctx *c, int *status, int *recompute, double *lh, double *lmin, double *lmax);

/* per-cell routines with the reference argument lists */
void orc_predictor(orc_ctx *c, double *lqhi, double *lFhi, double *lqbnd, double *lFbnd, const double *luh);
void orc_volume(orc_ctx *c, double *lduh, const double *lFhi);
void orc_surface(orc_ctx *c, double *lduh, const double *fl, const double *fr, const double *gb, const double *gt);
void orc_update(orc_ctx *c, double *luh, double *lduh);
void orc_flux(orc_ctx *c, const double *q, double *f);
void orc_eigen(orc_ctx *c, const double *q, const double *nv, double *lam);
void orc_rusanov(orc_ctx *c, const double *ql, const double *fl, const double *qr, const double *fr,
                 const double *nv, double *flux);
void orc_subcell_data(orc_ctx *c, double *lim, const double *luh);
void orc_lobatto_data(orc_ctx *c, double *lob, const double *luh);
void orc_put_subcell_data(orc_ctx *c, double *luh, const double *lim);
int orc_dmp(orc_ctx *c, const double *luh, const double *lmin, const double *lmax, double argmax);
/* signed DMP margin (negative = troubled), used for the near-threshold rule */
double orc_dmp_margin(orc_ctx *c, const double *luh, const double *lmin, const double *lmax, double argmax);
void orc_record_margins(orc_ctx *c, int on);
void orc_get_margins(const orc_ctx *c, double *out);

/* Baer-Nunziato (pde = ORC_PDE_BN, eos = GAMMA_S, GAMMA_G, PI_S, PI_G) */
void orc_predictor_bn(orc_ctx *c, double *lqhi, double *lFhi, double *lShi, double *lqbnd, double *lFbnd,
                      const double *luh);
void orc_volume_bn(orc_ctx *c, double *lduh, const double *lFhi, const double *lShi);
void orc_rusanov_bn(orc_ctx *c, const double *ql, const double *fl, const double *qr, const double *fr,
                    const double *nv, double *dm, double *dp);
void orc_ncp(orc_ctx *c, const double *q, const double *gradq, double *src);  /* PDEFusedSrcNCP */
void orc_roe(orc_ctx *c, const double *qa, const double *qb, const double *nv, double *b);
void orc_prim2cons(orc_ctx *c, const double *v, double *q);

/* whole-mesh phases */
double orc_calc_dt(orc_ctx *c);
int orc_step(orc_ctx *c);                       /* returns nRecompute */
void orc_get_face_fluxes(const orc_ctx *c, double *fu, double *gu);

/* algorithmic floating-point operation counter (+ - * / sqrt count 1 each) */
/* Advection-diffusion (pde = ORC_PDE_AD, nVar = 1): advection speeds, diffusivity,
 * boundary codes bC(1..4) of Advection-Diffusion/params.ini (1 Neumann, 2 Dirichlet,
 * 3 periodic) and the predictor flux: fixed = 0 the reference's PDEViscFlux,
 * fixed = 1 PDEFusedViscConvFlux.  Call right after orc_create. */
void orc_set_ad(orc_ctx *c, double cx, double cy, double a, const int bc[4], int fixed);

/* Euler / Baer-Nunziato boundary code per side (left, right, bottom, top), the
 * reference's params.ini convention (Euler/params.ini:8-11): 0 inflow with the
 * conserved state qin, 1 transmissive, 2 reflective, 3 periodic (in pairs).  The
 * reference computes only the DMR set (Euler/riemann_problem.f90:10-116) and
 * transmissive faces (Baer-Nunziato/riemann_problem.f90:9-80); the other codes
 * restate those patterns per side: inflow as the DMR inflow (:186-193), reflective
 * as the DMR wall (:225-228), and the subcell ghosts as the DMR overrides of
 * Euler/limiter.f90:491-522 on the neighbour table's clamp.  Parity of the codes
 * the reference lacks is pinned by invariants (tests/test_oracle_sides.py): all 1
 * and all 3 equal the transmissive and periodic modes bit for bit, a reflective
 * wall equals the mirror image of the doubled domain, a uniform state with inflow
 * of the same state stays uniform, and a supersonic inflow state fills the mesh.
 * Call right after orc_create. */
void orc_set_sides(orc_ctx *c, const int codes[4], const double *qin);
void orc_predictor_ad(orc_ctx *c, double *lqhi, double *lFhi, double *lqbnd, double *lgradqbnd, const double *luh);

void orc_reset_opcount(void);
double orc_get_opcount(void);

#ifdef __cplusplus
}
#endif
#endif
