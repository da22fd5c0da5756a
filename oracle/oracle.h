/*
 * oracle.h -- CPU restatement of code_mge's multigrid Jameson-RK solve step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests/, the smoke
 * check and the cpu_baseline leg of bench.py may load it, and only as the
 * checker.  The product path (code_mge_amd/) never links, loads or calls it.
 *
 * Every routine follows the reference Fortran line by line, including the
 * evaluation order of each expression (left to right, as gfortran evaluates
 * it without -ffast-math), so that the HIP path can be compared at the
 * last-bit level on single stages and at 1e-10 after many cycles.
 * Build with -O2 -ffp-contract=off (no fused multiply-add).
 *
 * Layout: every field is a Fortran array (0:ib,0:jb,0:kb[,5]) stored with
 * i fastest; component n of a 5-vector field lives at offset n*E.
 * Level 1 is the coarsest level (the reference's mgl = 1).
 */
#ifndef MGE_ORACLE_H
#define MGE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error / warning codes shared with the product's C ABI (include/mge.h) */
#define ORC_OK               0
#define ORC_ERR_NEG_C2     -10  /* timestep: c^2 < 0 or NaN   (srcf90/timestep.f90:45-49)  */
#define ORC_ERR_WALL_RHO   -11  /* efluxb: zero wall density   (srcf90/efluxb.f90:61-64)  */
#define ORC_ERR_WSOL_NAN   -12  /* bc_cutcell: NaN mirror state (srcf90/bc_cutcell.f90:147-157) */
#define ORC_ERR_RES_NAN    -13  /* solver: NaN residual        (srcf90/solver.f90:180)      */
#define ORC_ERR_PERIODIC   -14  /* bcond: unpaired periodic    (srcf90/bcond.f90:26-31)     */
#define ORC_ERR_SMOO_DIM   -16  /* psmoo_old with < 2 cells    (srcf90/psmoo.f90:132,166)   */
#define ORC_ERR_ARG        -20
#define ORC_WARN_ZERO_CLOUD  1  /* bc_cutcell: empty cloud, early return (bc_cutcell.f90:65-70) */
#define ORC_WARN_BCTYPE      2  /* bcond: unknown bctype (bcond.f90:35) */

typedef struct {
  int    mstage;
  double cstp[8], cdis[8];
  double cfl;          /* |cfl| (input_param.f90:37-39) */
  int    vt;           /* 0: local time step, 1: global */
  double vis2, vis4;   /* vis4 already divided by 32 (input_param.f90:40) */
  double smoo;
  int    mgtype;       /* 1 V-cycle, 2 W-cycle */
  double gamma;
  double rho0, p0, c0, ei0, u0, v0, w0, h0;
  int    eqntype;      /* 0 Euler, 1 laminar Navier-Stokes */
  int    dissip;       /* 0 scalar, 1 matrix */
  double viscl, prlmean, gas_const;
  double drop;         /* cycle exit when ini_res/res > drop (1e6 in solver.f90:238) */
  int    distype;      /* 0 central (JST), 1 Roe first order, 2 Roe + MUSCL-minmod */
  double epsentr;      /* Roe entropy-fix coefficient */
  int    execmode;     /* 0 steady, 1 unsteady dual time stepping (BDF2 physical time) */
  double dtstep;       /* nondimensional physical time step (input_param.f90:81-96) */
  int    smoother;     /* 0 psmoo_old (constant coefficient, what euler.f90:376 calls);
                          1 psmoo + progon3 (per-cell coefficients, psmoo.f90:1-95) */
  double ffreeze;      /* 0: bc_far's inflow/outflow choice at every call (bc_far.f90:58-60);
                          > 0: after the finest-level cycle whose residual has dropped by this
                          factor from the first, each far-field ghost point keeps its last choice */
  double lfreeze;      /* Roe r2 only.  0: minmod at every call; > 0: after the finest-level cycle
                          whose residual has dropped by this factor, every cell keeps the slope
                          choices (2-bit codes per direction and component) its limiter made last */
} orc_params;

typedef struct {
  int    faceid, zz;
  int    ccid[3], scid[3];
  double psurf[3], nvec[3], radcur[2], rr;
  int    npoint;
  int    ijk[8][3];
  double coef[8];
  double psol, wsol[5];
} orc_cut;

typedef struct orc_case orc_case;

orc_case *orc_case_new(const orc_params *prm, int nlevel);
void      orc_case_free(orc_case *c);

/* Level lvl (1..nlevel, 1 = coarsest).  xv/yv/zv: vertex coordinates
 * x(1..ib) (index 0 unused, kept 0 like init_metric.f90:59).
 * blank_int: ni*nj*nk bytes, i fastest, 0 = solid.  ratio: coarsening
 * ratio (1 or 2 per direction) from this level to lvl-1. */
int orc_set_level(orc_case *c, int lvl, int ni, int nj, int nk,
                  const double *xv, const double *yv, const double *zv,
                  double cellvolume, const int bctype[6],
                  const uint8_t *blank_int, int nmir, const orc_cut *cuts,
                  const int ratio[3], int mxcyc);

int  orc_init_flow(orc_case *c);                               /* init_flow.f90 */
int  orc_timestep(orc_case *c, int lvl);                       /* timestep.f90  */
int  orc_euler(orc_case *c, int lvl, int itime, int ncyc, int mgl, int mlvl,
               int lres, int lfir);                            /* euler.f90     */
int  orc_bcond(orc_case *c, int lvl);                          /* bcond.f90     */
int  orc_bc_cutcell(orc_case *c, int lvl);                     /* bc_cutcell.f90*/
int  orc_movco(orc_case *c, int mgl);                          /* movco.f90     */
int  orc_movfin(orc_case *c, int mgl, int mlvl);               /* movfin.f90    */
/* solver.f90: runs the FMG ramp; writes (level, ncyc, res) triples into
 * reslog (capacity cap entries) and returns the count in *nlog. */
int  orc_solve(orc_case *c, const int *mcyc, double *reslog, int cap, int *nlog,
               long long *cell_stages);
/* one physical time step of solver.f90's alop loop at time level itime (the
 * FMG ramp of orc_solve with euler's transient terms), then, for execmode 1,
 * the time-level shift w2 = w1, w1 = w on every level (solver.f90:263-276) */
int  orc_solve_step(orc_case *c, const int *mcyc, int itime, double *reslog, int cap, int *nlog,
                    long long *cell_stages);
int  orc_shift_time_levels(orc_case *c);                       /* solver.f90:263-276 */
/* upload_solut.f90:89-241: 7 (Euler) or 11 (NS) fields per interior cell,
 * field-major, i fastest; p_tot <= 0: freestream total pressure */
int  orc_upload_solut(orc_case *c, int lvl, double p_tot, double *out);

/* unit-level entry points for the tests: psmoo_old on the level's dw
 * (psmoo.f90:97-242) and dflux_matrix (dflux.f90:283-339) */
int  orc_psmoo(orc_case *c, int lvl);
/* psmoo (psmoo.f90:1-95, progon3.f90:1-62) on the level's dw with the
 * coefficients epsx/epsy/epsz of the last orc_timestep */
int  orc_psmoo_var(orc_case *c, int lvl);
/* progon3 itself on 0-based arrays of n (n >= 2) */
void orc_progon3(const double *a, const double *b, const double *c, const double *d, double *x, int n);
/* one Roe face (own design, no reference source): primitive states qL, qR
 * (rho,u,v,w,p), face vector S; e = central part 0.5*(F(UL)+F(UR)),
 * d = cdis*0.5*|A_roe|(UR-UL) with the Harten entropy fix eps.  With
 * signed != 0 the eigenvalues keep their sign (no fix): d = cdis*0.5*A(UR-UL). */
void orc_roe_face(double gamma, double eps, double cdis, const double *qL, const double *qR,
                  const double *S, int signed_lambda, double *e, double *d);
/* MUSCL-minmod face states from four cell primitives (q(a-2)..q(a+1)) and
 * their fluid flags; order 1 or 2 */
void orc_muscl(int order, const double *qm2, const double *qm1, const double *qp0, const double *qp1,
               const int *fl, double *qL, double *qR);
void orc_dflux_matrix(double gamma, const double *u1, const double *u2, const double *dw,
                      const double *ss, double *ff);

/* field access.  name: "w","p","dw","fw","ws","wn","wr","qf","radi","radj",
 * "radk","vdtim","si","sj","sk","w1","w2","epsx","epsy","epsz".  Copies the full (0:ib,...) array. */
int    orc_get(orc_case *c, int lvl, const char *name, double *out);
void   orc_set_frozen(orc_case *c, int far, int lim);   /* freeze flags, set directly (tests) */
int    orc_cycles(orc_case *c, int mlvl, int ncyc0, int count, double *res, long long *cell_stages);
/* embedded multi-block hierarchy over one-level cases (solver.f90:119-261
 * with bc_embed, bc_overlap and partial-coverage collc/addw) */
int    orc_comp_solve(orc_case **cases, int nb, const int *level, const int *parent, const int *g0,
                      const int *mcyc, double *reslog, int cap, int *nlog, long long *cell_stages);
int    orc_comp_op(orc_case **cases, int nb, const int *level, const int *parent, const int *g0, int op,
                   int L, int mlvl);
int    orc_set(orc_case *c, int lvl, const char *name, const double *in);
int    orc_get_cuts(orc_case *c, int lvl, orc_cut *out);
double orc_get_rtrms(orc_case *c, int lvl, int ncyc);
double orc_get_dtim(orc_case *c, int lvl);
/* build-defined FP64 wall force (Fx,Fy,Fz) from the wall-face flux that the
 * scheme applies (efluxb.f90:66-68), summed over records in order. */
int    orc_wall_force(orc_case *c, int lvl, double out[3]);

/* viscous-flux test hooks (own laminar flux; see oracle.c) */
void orc_visc_face(int dir, const double *pL, const double *pR, const double *t1L, const double *t1R,
                   const double *t2L, const double *t2R, double hn, double S, double mu, double kap, double *g);
int  orc_visc(orc_case *c, int lvl, double cdis);

#ifdef __cplusplus
}
#endif
#endif
