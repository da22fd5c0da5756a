/*
 * mge.h -- C ABI of the MI355X-native multigrid Jameson-RK solve step.
 *
 * The entry points keep the granularity of code_mge's Fortran call
 * surface, so a Fortran driver binds them 1:1 with iso_c_binding:
 *
 *   mge_timestep   <- timestep(b, lwork)                       srcf90/timestep.f90:1
 *   mge_euler      <- euler(b,itime,ncyc,mgl,mlvl,lres,lfir,…) srcf90/euler.f90:1
 *   mge_bcond      <- bcond(b)                                 srcf90/bcond.f90:1
 *   mge_bc_cutcell <- bc_cutcell(b)                            srcf90/bc_cutcell.f90:1
 *   mge_bc_overlap <- bc_overlap(blk, lv)                      srcf90/bc_overlap.f90:1
 *   mge_movco      <- movco(blk, mgl, nmesh, mblks)            srcf90/movco.f90:1
 *   mge_movfin     <- movfin(blk, mgl, mlvl, nmesh, mblks)     srcf90/movfin.f90:1
 *   mge_init_flow  <- init_flow(blk, lrflow, imodel)           srcf90/init_flow.f90:1
 *   mge_solve      <- solver(blk, project, job, model)         srcf90/solver.f90:1
 *   mge_solve_step, mge_shift_time_levels
 *                  <- solver's physical time loop              srcf90/solver.f90:110-290
 *   mge_upload_solut <- upload_solut(blk, p_tot, atcellcenter) srcf90/upload_solut.f90:89
 *   mge_write_solut/mge_read_solut <- write_solut/read_solut   srcf90/wsolut.f90:47, rsolut.f90:1
 *   mge_level_define/mge_case_create  <- make_flow + init_metric
 *                                       (srcf90/make_flow.f90:1, init_metric.f90:1)
 *   mge_geom_level <- the cut-cell part of make_fsijob + solid_ipolcoef
 *                     (srcf90/main.f90:44,57; their library is not in the
 *                     reference tree, so the record layout follows the
 *                     fields that bc_cutcell/efluxb/dfluxb read)
 *
 * Conventions
 *  - Every function returns int: 0 = OK, < 0 = error (one code per
 *    reference `stop` site), > 0 = warning (reference printed and went on).
 *  - The case handle owns all device memory.  Host arrays are read or
 *    written only during the call that receives them; none is retained.
 *  - Calls on one handle are ordered on the handle's HIP stream and are not
 *    re-entrant.  Different handles are independent.
 *  - Levels are numbered 1..nlevel with 1 the coarsest (the reference's
 *    mgl); each level is one Cartesian block (blk = 1).
 *  - Field arrays crossing the ABI are Fortran-ordered (0:ib,0:jb,0:kb[,5])
 *    with ib = ni+3: i fastest, component slowest, FP64 (an FP32 case
 *    converts on download/upload).
 */
#ifndef MGE_H
#define MGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MGE_OK               0
#define MGE_ERR_NEG_C2     -10  /* c^2 < 0 or NaN in timestep   (srcf90/timestep.f90:45-49)      */
#define MGE_ERR_WALL_RHO   -11  /* zero density at a wall face  (srcf90/efluxb.f90:61-64)      */
#define MGE_ERR_WSOL_NAN   -12  /* NaN mirror state             (srcf90/bc_cutcell.f90:147-157)  */
#define MGE_ERR_RES_NAN    -13  /* NaN density residual         (srcf90/solver.f90:180)          */
#define MGE_ERR_PERIODIC   -14  /* unpaired periodic face       (srcf90/bcond.f90:26-31)         */
#define MGE_ERR_DISTYPE    -15  /* bad distype / dissip         (srcf90/euler.f90:212-214)       */
#define MGE_ERR_SMOO_DIM   -16  /* smoothing needs >= 2 cells   (srcf90/psmoo.f90:132,166)       */
#define MGE_ERR_ARG        -20  /* bad argument                                                  */
#define MGE_ERR_HIP        -21  /* HIP runtime failure                                           */
#define MGE_ERR_NOMEM      -22  /* device allocation failed                                      */
#define MGE_ERR_STATE      -23  /* call out of order (e.g. level not defined)                    */
#define MGE_ERR_RESTART    -24  /* restart file does not match the case (srcf90/rsolut.f90:28-32) */
#define MGE_WARN_ZERO_CLOUD  1  /* empty reflected-node cloud: early return (bc_cutcell.f90:65-70) */
#define MGE_WARN_BCTYPE      2  /* unknown bctype, face skipped (bcond.f90:35)                    */

/* scheme and freestream parameters (typepara, filled as input_param.f90:13-115) */
typedef struct {
  int    mstage;              /* RK stages (<= 8)                                  */
  double cstp[8], cdis[8];    /* stage and dissipation-blending coefficients       */
  double cfl;                 /* |cfl|                                             */
  int    vt;                  /* 0 local time step (jobsheet cfl < 0), 1 global    */
  double vis2, vis4;          /* vis4 already divided by 32                        */
  double smoo;                /* constant residual-smoothing coefficient           */
  int    mgtype;              /* 1 V-cycle, 2 W-cycle                              */
  double gamma;
  double rho0, p0, c0, ei0, u0, v0, w0, h0;
  int    eqntype;             /* 0 Euler, 1 laminar Navier-Stokes                  */
  int    dissip;              /* 0 scalar, 1 matrix                                */
  double viscl, prlmean, gas_const;
  double drop;                /* solve exits when ini_res/res > drop (1e6 in the reference) */
  int    distype;             /* 0 central (JST); 1 Roe first order (r1); 2 Roe with
                                 MUSCL-minmod (r2)                                 */
  int    precision;           /* 64 (or 0): FP64 throughout.  32: FP32 bulk state,
                                 fluxes, smoothing and updates; the wall mirror
                                 states, residual norms, global time step and
                                 force integral stay FP64                          */
  double epsentr;             /* Roe entropy-fix coefficient (jobsheet r2_epsentr,
                                 input_param.f90:102)                              */
  int    execmode;            /* 0 steady; 1 unsteady dual time stepping: BDF2
                                 physical time derivative in euler, pseudo step
                                 capped by the physical one (jobsheet runtype)     */
  double dtstep;              /* nondimensional physical time step (> 0 when
                                 execmode = 1; input_param.f90:81-96)              */
  int    smoother;            /* implicit residual smoothing: 0 psmoo_old, the
                                 constant coefficient smo that euler.f90:376 calls;
                                 1 psmoo, per-cell coefficients from timestep
                                 (timestep.f90:286-309) solved line by line as
                                 progon3 does (psmoo.f90:1-95, progon3.f90:1-62) */
  double ffreeze;             /* 0: the far-field rule of bc_far.f90:58-112 as is
                                 (inflow or outflow by the sign of the normal
                                 velocity at every call).  > 0: once the finest
                                 level's residual has dropped by this factor from
                                 its first cycle, each far-field ghost point keeps
                                 the branch it last took, so points where the
                                 normal velocity is near zero (a wake crossing a
                                 face parallel to the stream) stop switching
                                 between the two branches cycle after cycle       */
  double lfreeze;             /* Roe + MUSCL-minmod only (distype 2).  0: the
                                 minmod limiter at every stage.  > 0: once the
                                 finest level's residual has dropped by this
                                 factor from its first cycle, every cell keeps
                                 the slope choice (zero, backward or forward
                                 difference, per direction and component) its
                                 limiter made last, so the face states stop
                                 switching between limiter branches and the
                                 iteration converges to the solution of the
                                 frozen scheme (no counterpart in the reference,
                                 whose roepak source is absent)                    */
} mge_params;

/* one wall (cut-cell, solid-cell) face pair: the fields of a txcut record
 * that bc_cutcell.f90, efluxb.f90 and dfluxb.f90 read.  Indices follow the
 * Fortran block numbering (interior cells 2..ncell+1). */
typedef struct {
  int    faceid;              /* 1..6: side of the cut cell holding the solid cell */
  int    zz;                  /* direction 1..3                                    */
  int    ccid[3], scid[3];    /* cut cell, solid cell                              */
  double psurf[3], nvec[3], radcur[2], rr;
  int    npoint;              /* reflected-node cloud size (<= 8)                  */
  int    ijk[8][3];
  double coef[8];
  double psol, wsol[5];       /* mirror state (output of bc_cutcell)               */
} mge_cutrec;

/* analytic immersed bodies for the geometry preprocessor */
#define MGE_BODY_NONE      0
#define MGE_BODY_CYLINDER  1  /* circular cylinder, radius semi[0], axis `axis` */
#define MGE_BODY_SPHERE    2  /* radius semi[0]                                  */
#define MGE_BODY_ELLIPSOID 3  /* semi-axes semi[0..2] in body frame, rot = body->world (row-major) */
typedef struct {
  int    kind;
  int    axis;
  double center[3];
  double semi[3];
  double rot[9];
} mge_body;

typedef struct mge_case mge_case;

/* ---- input files (the role of make_fsijob's readers, srcf90/main.f90:44) */
/* one grid of a .pin file: vertex counts, origin, extent */
typedef struct {
  int    name;
  int    points[3];
  double origin[3], length[3];
} mge_pin_grid;

/* what a jobsheet (with its .pin and .p3d) defines for the solve step */
typedef struct {
  char   job[64], model[64];
  char   mesh[512], surface[512];   /* paths of the .pin and .p3d files       */
  int    nlevel;                    /* entries of mcyc                        */
  int    mcyc[16];                  /* cycles per level of the FMG ramp       */
  double mach, reynolds, gamma, prlmean, alphad, betad;
  double cflf, vis2, vis4, smoo, epsentr;
  int    mgtype, mstage;
  double cstp[8], cdis[8];
  int    eqntype, distype, dissip;
  int    n[3];                      /* finest-level cells                     */
  double lo[3], hi[3];              /* block extent                           */
  double dom_lo[3], dom_hi[3];      /* DOMAIN bounds                          */
  int    bctype[6];                 /* faces i-lo, j-lo, k-lo, i-hi, j-hi, k-hi */
  int    iflow;                     /* CONTROL iflow: 1 = start from <job>_restart.dat */
  int    execmode;                  /* CONTROL runtype: 0 steady, 1 unsteady    */
  int    nperiod, nstep;            /* CONTROL period, step: physical steps = period*step */
  double dt;                        /* CONTROL dt: physical time step [s]      */
  double pmean, tmean, rgasmean;    /* FLOW: freestream p [Pa], T [K], R [J/(kg K)] */
} mge_job;

/* Read a jobsheet and the single-block .pin it names (make_fsijob's job
 * and grid input; the key mapping is this library's, ctrlib being absent
 * from the reference tree).  Paths in the jobsheet are relative to it. */
int mge_read_job(const char *jobsheet, mge_job *job);
/* the scheme parameters the job defines, derived as input_param.f90:13-115 */
int mge_job_params(const mge_job *job, int precision, mge_params *prm);
/* cells and coarsening ratio (to lvl-1) of level lvl, 1 = coarsest */
int mge_job_level(const mge_job *job, int lvl, int n[3], int ratio[3]);
/* grids of a .pin file (cap entries written, *ngrid = count) */
int mge_read_pin(const char *path, int cap, mge_pin_grid *grids, int *ngrid);
/* triangles (9 values each) of the surface patches of a PLOT3D .p3d file
 * (cap triangles written, *ntri = count) */
int mge_read_p3d(const char *path, int cap, double *tri, int *ntri);
/* message of the last failed input call */
const char *mge_input_error(void);

/* ---- geometry preprocessor (host only, no GPU needed) ------------------ */
/* Classifies the interior cells of one block (cell centre inside the body =
 * solid) and builds one record per (fluid, solid) face neighbour with the
 * surface point, normal, mirror distance and a trilinear reflected-node
 * cloud over fluid cells.  xv/yv/zv: vertex coordinates x(0..ib).
 * blank: ni*nj*nk bytes (i fastest, 1 fluid, 0 solid).  If recs is NULL or
 * cap is too small only *nmir is set. */
int mge_geom_level(const mge_body *body, int ni, int nj, int nk,
                   const double *xv, const double *yv, const double *zv,
                   uint8_t *blank, mge_cutrec *recs, int cap, int *nmir);
/* The same for a closed triangulated surface (tri: 9 values per triangle,
 * the three vertices; e.g. the quads of the PLOT3D surface patches that
 * make_fsijob reads, srcf90/main.f90:44, split in two): a cell centre is
 * inside when the grid line through it crosses the surface an odd number
 * of times before it; the record normal is the crossed triangle's. */
int mge_geom_level_tri(const double *tri, int ntri, int ni, int nj, int nk,
                       const double *xv, const double *yv, const double *zv,
                       uint8_t *blank, mge_cutrec *recs, int cap, int *nmir);

/* ---- case lifetime (make_flow.f90 / init_metric.f90) -------------------- */
int mge_case_create(const mge_params *prm, int nlevel, mge_case **out);
int mge_case_destroy(mge_case *h);
/* ratio: coarsening factor (1 or 2 per direction) from this level to lvl-1 */
int mge_level_define(mge_case *h, int lvl, int ni, int nj, int nk,
                     const double *xv, const double *yv, const double *zv,
                     double cellvolume, const int bctype[6], const int ratio[3],
                     const uint8_t *blank, int nmir, const mge_cutrec *recs,
                     int mxcyc);

/* ---- the solve step (reference call surface) ---------------------------- */
int mge_init_flow(mge_case *h);
int mge_timestep(mge_case *h, int lvl, int blk);
/* rtrms: density-residual RMS of this call when lres == 0 (may be NULL) */
int mge_euler(mge_case *h, int lvl, int blk, int itime, int ncyc, int mgl,
              int mlvl, int lres, int lfir, double *rtrms);
int mge_bcond(mge_case *h, int lvl, int blk);
int mge_bc_cutcell(mge_case *h, int lvl, int blk);
int mge_bc_overlap(mge_case *h, int lvl);
int mge_movco(mge_case *h, int mgl);
int mge_movfin(mge_case *h, int mgl, int mlvl);
/* full FMG/FAS solve (solver.f90); reslog receives (level, ncyc, res)
 * triples like <job>_res.dat; cell_stages counts residual+update passes. */
int mge_solve(mge_case *h, const int *mcyc, double *reslog, int cap, int *nlog,
              long long *cell_stages);

/* one physical time step of solver.f90's time loop (the alop/rstep body,
 * solver.f90:110-290): the FMG ramp of mge_solve with euler at time level
 * itime (0 for the first step; euler.f90:249-275 adds the BDF2 term from
 * itime 2 on), then, when execmode = 1, the time-level shift of
 * mge_shift_time_levels.  With execmode = 0 and itime = 0 this is mge_solve. */
int mge_solve_step(mge_case *h, const int *mcyc, int itime, double *reslog, int cap, int *nlog,
                   long long *cell_stages);
/* w2 = w1, w1 = w on every level (solver.f90:263-276); execmode = 1 only */
int mge_shift_time_levels(mge_case *h);

/* `count` multigrid cycles with finest active level mlvl and no drop test
 * (the lop0 body of solver.f90:124-254 for ncyc = ncyc0..ncyc0+count-1);
 * res receives the last cycle's residual (may be NULL). */
int mge_cycles(mge_case *h, int mlvl, int ncyc0, int count, double *res, long long *cell_stages);

/* ---- data movement and results ------------------------------------------ */
/* field: "w","p","dw","fw","ws","wr","qf","vdtim","radsum" (radi+radj+radk);
 * unsteady cases also "w1","w2" (physical time levels n and n-1) */
int mge_download(mge_case *h, int lvl, int blk, const char *field, double *out);
int mge_upload(mge_case *h, int lvl, int blk, const char *field, const double *in);
int mge_get_cuts(mge_case *h, int lvl, int blk, mge_cutrec *out);
double mge_get_dtim(mge_case *h, int lvl, int blk);
/* upload_solut (srcf90/upload_solut.f90:89-241): the output fields of every
 * interior cell in FP64, field-major with i fastest over ni*nj*nk cells --
 * rho, u, v, w, cp = (p/p0 - 1)*2/(gamma*M^2), Mach, total-pressure loss
 * p*(1 + 0.2*Mach^2)^3.5/p_tot - 1, and for laminar flow rho*E/rho and three
 * zero gradient slots.  *nval receives 7 (Euler) or 11 (Navier-Stokes);
 * solut may be NULL to query it.  p_tot <= 0 takes the freestream total
 * pressure p0*(1 + 0.2*M^2)^3.5 (the reference passes gas_ptot). */
int mge_upload_solut(mge_case *h, int lvl, int blk, double p_tot, double *solut, int *nval);
/* mge_upload_solut without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops or harnesses that forbid it).  The
 * fields pass through the library's fixed device scratch array of
 * MGE_SCRATCH_BYTES, which every GPU the library runs on holds until the
 * process ends, in chunks of whole k-planes; a level whose one k-plane of
 * fields (ni*nj*nval*8 bytes) exceeds it returns MGE_ERR_ARG.  Calls that
 * overlap on one GPU share that array and race: serialise them. */
#define MGE_SCRATCH_BYTES (32ull << 20)
int mge_upload_solut_noalloc(mge_case *h, int lvl, int blk, double p_tot, double *solut, int *nval);

/* FP64 wall force (Fx,Fy,Fz) on the body at level lvl: the momentum flux
 * the scheme applies at each wall face (efluxb.f90:66-68), in record order */
int mge_wall_force(mge_case *h, int lvl, double out[3]);
/* restart files of every level: format 0 the reference's text file
 * (write_solut, srcf90/wsolut.f90:47-93: per block a "(5I5)" header code,
 * low(4), upp(4), then w(0:ib,0:jb,0:kb,1:5) in "(5f16.7)"; block code
 * 100*lvl + 1); reading it (read_solut, srcf90/rsolut.f90:1-41) checks code
 * and dimensions, initialises the flow, takes w from the file and rebuilds
 * p and the wall mirror states from it.  Format 1, the full-precision
 * variant, also holds p and the mirror states, so a run restarted from it
 * continues exactly as the run that wrote it. */
int mge_write_solut(mge_case *h, const char *path, int format);
int mge_read_solut(mge_case *h, const char *path, int format);
int mge_synchronize(mge_case *h);
/* HIP-event timing of the library's launches on the case stream: mean
 * duration (ms), launch count and mean algorithmic bytes per launch of the
 * named kernel since the last reset; enable turns recording on or off;
 * name NULL resets the statistics.  (No counterpart in the reference: it is
 * the measurement hook of the benchmark.) */
int mge_kernel_timing(mge_case *h, int enable, const char *name, double *ms, long long *count,
                      double *bytes);

/* Autotuner of the LDS-tiled stage kernels (no counterpart in the
 * reference): times stage 1 of level lvl for the tile rows tj = 4 and 8
 * (256 or 512 threads per workgroup) and every k-chunk length kc = 1, 2,
 * 4, ... nk, and keeps the fastest pair for the level; results are cached
 * per (shape, precision, scheme) for the process.  The launch shape changes
 * no arithmetic.  Overwrites only scratch state (a non-current state
 * buffer, dw, fw, the time-step arrays); call it between solves, not while
 * a batch is recording.  *kc (0 for the matrix-dissipation kernel, which has
 * no chunks) and *us (best launch time, microseconds) may be NULL; the
 * chosen tile rows are read with mge_stage_tile. */
int mge_autotune(mge_case *h, int lvl, int *kc, double *us);
/* force the k-chunk length of level lvl (0: the built-in fill rule) */
int mge_set_kc(mge_case *h, int lvl, int kc);
/* set the case's freeze flags directly: far-field branches (ffreeze) and
 * limiter choices (lfreeze); each flag is otherwise set by the case's own
 * freeze test and cleared by mge_init_flow (no counterpart in the
 * reference; for tests and callers that decide the freeze themselves) */
int mge_set_freeze(mge_case *h, int far, int lim);
/* the tile rows and k-chunk length the stage kernel of level lvl launches
 * with (no counterpart in the reference) */
int mge_stage_tile(mge_case *h, int lvl, int *tj, int *kc);
/* force the tile rows (4 or 8; 0: the precision's default) and k-chunk
 * length (0: the fill rule) of level lvl (no counterpart in the reference) */
int mge_set_tile(mge_case *h, int lvl, int tj, int kc);
/* ---- embedded multi-block hierarchies (composite grids) ----------------
 * Several blocks per multigrid level, each finer level covering part of the
 * coarser one (the reference's embedded grids: bc_embed.f90:50-143,
 * bc_overlap.f90:32-90, partial-coverage collc.f90:58-146 and
 * addw.f90:84-188 inside movco.f90:21-49 / movfin.f90:20-40, with the block
 * loops of solver.f90:119-261).  Each block is a one-level case (its own
 * mge_case_create + mge_level_define, one level, the same parameters);
 * block i has composite level level[i] (1 = coarsest), parent[i] (the
 * block of level - 1 holding it, -1 on level 1) and g0[3i..3i+2], the index
 * of its cell 2 on its level's lattice (the level-1 lattice origin at 0;
 * each level refines the one below by 2 in every direction, so a child's
 * g0 and extents are even and lie within twice its parent's).  A face of a
 * child strictly inside its parent has bctype 0 (filled by bc_embed and
 * bc_overlap); a face on its parent's boundary keeps the domain condition.
 * While in a composite the blocks run on its stream and must not be used
 * on their own. */
typedef struct mge_comp mge_comp;
int mge_comp_create(int nblock, mge_case *const *blocks, const int *level, const int *parent, const int *g0,
                    mge_comp **out);
/* solver.f90:119-261 over the composite (steady): mcyc[l] cycles with the
 * finest active level l + 1, the FMG transfer after each; the residual of
 * a cycle is the smallest over the active level's blocks (solver.f90:164-179);
 * reslog triples (level, cycle, residual) as mge_solve */
int mge_comp_solve(mge_comp *c, const int *mcyc, double *reslog, int cap, int *nlog, long long *cell_stages);
/* release the composite; its blocks return to their own streams */
int mge_comp_destroy(mge_comp *c);
const char *mge_last_error(void);

/* ---- batches of cases --------------------------------------------------
 * No counterpart in the reference (it advances one case per process); for
 * callers that solve many independent cases: throughput runs and the
 * serving loop.  A batch advances its cases together, two multigrid cycles
 * (finest level, no full-multigrid ramp) per period, with one launch per
 * operation for all cases, replayed as a hipGraph when use_graph != 0.
 * Its cases must share the number of levels, scheme, stages, cycle type,
 * smoothing on/off and precision; they may differ in grid, body and flow
 * parameters.  Each case stops at the first cycle whose finest-level
 * residual has dropped by more than `drop` from its first cycle (the drop
 * test of srcf90/solver.f90:238), at maxcyc cycles, or on an error; it then
 * holds exactly the state its own single-case run of those cycles holds.
 * The cases stay owned by the caller: remove a case before destroying it. */
/* device bytes held by a case (every level's arrays); no counterpart in
 * the reference -- for callers that admit cases by a memory budget */
long long mge_case_bytes(mge_case *h);

typedef struct mge_batch mge_batch;
int mge_batch_create(int use_graph, int log_cap, mge_batch **out);
int mge_batch_destroy(mge_batch *b);
/* returns the case's slot (>= 0); the case must have been initialised */
int mge_batch_add(mge_batch *b, mge_case *h, double drop, int maxcyc);
int mge_batch_remove(mge_batch *b, int slot);
/* run up to `periods` periods (2 cycles each) or until no case is active */
int mge_batch_run(mge_batch *b, int periods, int *nactive);
/* status: 0 running, 1 converged, 2 cycle limit, 3 removed, < 0 error code */
int mge_batch_status(mge_batch *b, int slot, int *status, int *ncyc, double *ini, double *last);
/* residual history of a slot (one value per cycle); returns the count */
int mge_batch_log(mge_batch *b, int slot, double *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
