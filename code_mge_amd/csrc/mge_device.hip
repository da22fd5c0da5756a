// mge_device.hip -- HIP kernels for gfx950 and the C ABI of include/mge.h.
//
// One Cartesian block per multigrid level lives in HBM for the whole solve:
//   S[3]     three state buffers (w: 5 components, then p), rotated across
//            RK stages so W^(0) (the reference's wn) is never copied
//   dw, fw   residual and blended dissipation (5 components each)
//   ws,wr,qf multigrid saved state, restricted residual, forcing
//   radsum   radi+radj+radk, vdtim: local time step (interior cells)
//   flags    bit0 fluid (blank != 0), bit1 touched by a wall record
// All fields use the reference numbering (0:ib,0:jb,0:kb), i fastest,
// component-major, so host<->device copies need no reshuffling.
//
// Stage pipeline (euler.f90:118-423), per RK stage m:
//   k_stage     fluxes + dissipation + blending + forcing + dt scaling
//               (+ update when the stage is not smoothed) for every cell
//               not touched by a wall record
//   k_records   the wall corrections of dfluxb.f90/efluxb.f90 in record
//               order for the touched cells, then the same finish
//   k_smooth_x/y/z  psmoo_old.f90 line recurrences (z fuses the update)
//   k_ghost_copy, k_bc_cut, k_bc_face x6, k_bc_corner  (bc_cutcell, bcond)
//   k_rtrms     residual norm after the last stage (euler.f90:431-447)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mge.h"
#include "mge_kernels.h"

using namespace mge;

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int set_err(int code, const char *msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      char b_[256];                                                        \
      snprintf(b_, sizeof b_, "%s: %s", #x, hipGetErrorString(e_));        \
      return set_err(MGE_ERR_HIP, b_);                                     \
    }                                                                      \
  } while (0)

// device error flag: the first code raised stays (the reference stops at
// its first failing check)
__device__ inline void raise_err(int *flag, int code) { atomicCAS(flag, 0, code); }

// ------------------------------------------------------------------ records

struct DevRec {
  int faceid, zz;
  int c[3], s[3];
  int npoint, pad;
  long cloud[8];
  double coef[8];
  double nvec[3];
};

// op kinds applied to one touched cell, in the reference order
enum OpKind : int {
  OP_HI_C_UNDO = 0,  // fw(s)   -= dfs(s) - dfs(s+1)        dfluxb.f90:47
  OP_HI_N_UNDO,      // fw(s-1) -= dfs(s-1) - dfs(s)        dfluxb.f90:48
  OP_HI_C_ADD,       // fw(s)   += fs1 - fs2                dfluxb.f90:104
  OP_HI_N_ADD,       // fw(s-1) += dfs(s-1) - fs1           dfluxb.f90:105
  OP_LO_C_UNDO,      // fw(s)   -= dfs(s) - dfs(s+1)        dfluxb.f90:109
  OP_LO_N_UNDO,      // fw(s+1) -= dfs(s+1) - dfs(s+2)      dfluxb.f90:110
  OP_LO_C_ADD,       // fw(s)   += fs1 - fs2                dfluxb.f90:170
  OP_LO_N_ADD,       // fw(s+1) += fs2 - dfs(s+2)           dfluxb.f90:171
  OP_E_UNDO,         // dw(s)   -= efs(s) - efs(s+1)        efluxb.f90:257
  OP_E_ADD           // dw(s)   += fa - fs  or  fs - fa     efluxb.f90:262,266
};

// ------------------------------------------------------------------ kernel args

struct StageArgs {
  FluxCtx fx;
  double *wdst, *pdst;      // output state (update)
  const double *wn;         // W^(0)
  double *dw, *fw;
  double *qf;
  const double *wr;
  const uint8_t *marker;
  const uint8_t *flags;
  const double *vdtim;
  const double *dtim;       // device scalar (global time step)
  double cellvolume;
  double blend;             // 1 - cdis(m)
  double fn;                // cstp(m)*cfl
  int m;                    // stage number (1-based)
  int has_diss;             // cdis(m) > 0
  int coarse;               // mgl < mlvl
  int lfir1;                // lfir and m == 1
  int lres;
  int smoothed;             // psmoo_old follows (odd stage, smoo > 0)
  int store_fw;             // a later stage reads fw
  int vt;
};

struct RecArgs {
  StageArgs st;
  const DevRec *recs;
  const double *wsol;       // 5*nmir, component-major
  const double *psol;
  int nmir;
  const long *tcell;        // touched cells
  const int *beg;           // op range of each touched cell (ntouch+1)
  const int2 *ops;          // (record, kind), fw ops then dw ops per cell
  double *opval;            // 10 doubles per op
  int ntouch, nops;
  int *err;
};

// ------------------------------------------------------------------ finish of a cell

// dw = dw + fw, forcing, blanking, time-step scaling and (unsmoothed) the
// update with the EOS (euler.f90:217-223, 294-342, 363-399)
__device__ inline void finish_cell(const StageArgs &A, long q, int i, int j, int k,
                                   const double *f, const double *e) {
  const long E = A.fx.d.E;
  double r[5];
#pragma unroll
  for (int n = 0; n < 5; n++) r[n] = e[n] + f[n];
  if (A.coarse) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      double qv;
      if (A.lfir1) {
        qv = 0.0;
        if (A.marker[q] == 1) qv = 1.0 * A.wr[n * E + q] - r[n];
        A.qf[n * E + q] = qv;
      } else {
        qv = A.qf[n * E + q];
      }
      r[n] = r[n] + qv;
    }
  }
  double bl = (A.flags[q] & F_FLUID) ? 1.0 : 0.0;
#pragma unroll
  for (int n = 0; n < 5; n++) r[n] = r[n] * bl;
  if (A.store_fw) {
#pragma unroll
    for (int n = 0; n < 5; n++) A.fw[n * E + q] = f[n];
  }
  if (A.lres) {
#pragma unroll
    for (int n = 0; n < 5; n++) A.dw[n * E + q] = r[n];
    return;
  }
  double de = A.fn * A.vdtim[q] / A.cellvolume;
  if (A.vt == 1) de = A.fn * A.dtim[0] / A.cellvolume;
#pragma unroll
  for (int n = 0; n < 5; n++) r[n] = de * r[n];
#pragma unroll
  for (int n = 0; n < 5; n++) A.dw[n * E + q] = r[n];
  if (!A.smoothed) {
    double v[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      v[n] = A.wn[n * E + q] - r[n] / 1.0;
      A.wdst[n * E + q] = v[n];
    }
    A.pdst[q] = eos(A.fx.gamma, v);
  }
}

// ------------------------------------------------------------------ k_stage

__global__ void __launch_bounds__(256) k_stage(StageArgs A) {
  const Dims &d = A.fx.d;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 2 + blockIdx.z;
  if (i > d.il || j > d.jl || k > d.kl) return;
  const long E = d.E;
  long q = ix(d, i, j, k);
  double f[5];
  if (A.m > 1) {
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = A.fw[n * E + q] * A.blend;
  } else {
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = 0.0;
  }
  if (A.has_diss) {
    double a[5], b[5];
    long bx = ix(d, 0, j, k), by = ix(d, i, 0, k), bz = ix(d, i, j, 0);
    dface<0>(A.fx, bx, i, i, j, k, a);
    dface<0>(A.fx, bx, i + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = f[n] + (a[n] - b[n]);
    dface<1>(A.fx, by, j, i, j, k, a);
    dface<1>(A.fx, by, j + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = f[n] + (a[n] - b[n]);
    dface<2>(A.fx, bz, k, i, j, k, a);
    dface<2>(A.fx, bz, k + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = f[n] + (a[n] - b[n]);
  }
  double e[5];
  {
    double a[5], b[5];
    long bx = ix(d, 0, j, k), by = ix(d, i, 0, k), bz = ix(d, i, j, 0);
    eface<0>(A.fx, bx, i, i, j, k, a);
    eface<0>(A.fx, bx, i + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) e[n] = a[n] - b[n];
    eface<1>(A.fx, by, j, i, j, k, a);
    eface<1>(A.fx, by, j + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) e[n] = e[n] + a[n] - b[n];
    eface<2>(A.fx, bz, k, i, j, k, a);
    eface<2>(A.fx, bz, k + 1, i, j, k, b);
#pragma unroll
    for (int n = 0; n < 5; n++) e[n] = e[n] + a[n] - b[n];
  }
  if (A.flags[q] & F_TOUCHED) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      A.fw[n * E + q] = f[n];
      A.dw[n * E + q] = e[n];
    }
    return;
  }
  finish_cell(A, q, i, j, k, f, e);
}

// ------------------------------------------------------------------ k_stage2
//
// LDS-tiled form of k_stage.  A workgroup owns a TI x TJ column tile of
// (i,j) and marches through a chunk of k planes.  Per plane the i- and
// j-direction quantities every face needs -- the convective cell fluxes
// F (ceflux.f90), the pressure switches fp and the enthalpy differences diw
// (dflux.f90:70-86) -- are computed once per cell into LDS (tile + halo);
// each thread then forms its faces from LDS.  The k-direction face values
// are carried from plane to plane in registers, so every k face is
// evaluated once.  Each face value is computed with the same expression
// tree as the reference, so results equal k_stage's bit for bit.
constexpr int TI = 64, TJ = 4;

struct StageSmem {
  double fx[5][TJ][TI + 2];   // F_x at i = i0-1 .. i0+TI
  double fy[5][TJ + 2][TI];   // F_y at j = j0-1 .. j0+TJ
  double px[TJ][TI + 4];      // fp_x at i = i0-2 .. i0+TI+1
  double py[TJ + 4][TI];      // fp_y at j = j0-2 .. j0+TJ+1
  double dx[5][TJ][TI + 3];   // diw_x at i = i0-1 .. i0+TI+1
  double dy[5][TJ + 3][TI];   // diw_y at j = j0-1 .. j0+TJ+1
};

// JST face coefficients from the switches fp(a-2..a+1) and the radii sums
// on both sides (dflux.f90:88-96); the face value of component n is then
// (dis2*diw(a) - dis4*(diw(a+1) - 2 diw(a) + diw(a-1))) * lam
__device__ inline void face_coef(const FluxCtx &c, double f0, double f1, double f2, double f3, double lm, double lp,
                                 double *dis2o, double *dis4o, double *lamo) {
  double ep = maxd(maxd(maxd(f0, f1), f2), f3);
  double lam = 0.5 * (lm + lp);
  double dis2 = c.vis2 * ep;
  double dis4 = c.vis4;
  dis4 = dimd(dis4, dis2);
  dis2 = dis2 * c.cdis;
  dis4 = dis4 * c.cdis;
  *dis2o = dis2;
  *dis4o = dis4;
  *lamo = lam;
}

template <int DIR>
__device__ inline void cell_flux(const FluxCtx &c, int i, int j, int k, double *o) {
  long q = ix(c.d, i, j, k);
  double v[5];
  load5(c.s.w, c.d.E, q, v);
  cflux<DIR>(v, c.s.p[q], face_s<DIR>(c.g, i, j, k), o);
}

__global__ void __launch_bounds__(256, 2) k_stage2(StageArgs A, int kc) {
  __shared__ StageSmem sm;
  const FluxCtx &c = A.fx;
  const Dims &d = c.d;
  const long E = d.E;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TI * ty;
  const int i0 = 2 + blockIdx.x * TI, j0 = 2 + blockIdx.y * TJ;
  const int kb0 = 2 + blockIdx.z * kc;
  const int kb1 = min(kb0 + kc - 1, d.kl);
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = i <= d.il && j <= d.jl;
  const bool diss = A.has_diss;
  const long bz = ix(d, i, j, 0);
  // k-direction carry (face k) and rings for face k+1
  double Dk[5], Ek[5], fz[4], dzr[3][5], rz[2], Fz[2][5];
  if (valid) {
    int k = kb0;
    eface<2>(c, bz, k, i, j, k, Ek);
    cell_flux<2>(c, i, j, k, Fz[0]);
    cell_flux<2>(c, i, j, k + 1, Fz[1]);
    if (diss) {
      dface<2>(c, bz, k, i, j, k, Dk);
      for (int t = 0; t < 4; t++) fz[t] = fp_at<2>(c, bz, k - 1 + t);
      for (int t = 0; t < 3; t++) diw_at<2>(c, bz, k + t, dzr[t]);
      rz[0] = rad_at<2>(c, bz, k);
      rz[1] = rad_at<2>(c, bz, k + 1);
    }
  }
  for (int k = kb0; k <= kb1; k++) {
    // ---- fill the plane's i/j quantities (cells within the arrays)
    {
      // own cell
      if (i <= d.ib && j <= d.jb) {
        long brow = ix(d, 0, j, k), bcol = ix(d, i, 0, k);
        double o[5];
        cell_flux<0>(c, i, j, k, o);
        for (int n = 0; n < 5; n++) sm.fx[n][ty][tx + 1] = o[n];
        cell_flux<1>(c, i, j, k, o);
        for (int n = 0; n < 5; n++) sm.fy[n][ty + 1][tx] = o[n];
        if (diss) {
          sm.px[ty][tx + 2] = fp_at<0>(c, brow, i);
          sm.py[ty + 2][tx] = fp_at<1>(c, bcol, j);
          diw_at<0>(c, brow, i, o);
          for (int n = 0; n < 5; n++) sm.dx[n][ty][tx + 1] = o[n];
          diw_at<1>(c, bcol, j, o);
          for (int n = 0; n < 5; n++) sm.dy[n][ty + 1][tx] = o[n];
        }
      }
      // halo items
      const int nF = 2 * TJ + 2 * TI;
      const int nP = diss ? 4 * TJ + 4 * TI : 0;
      const int nD = diss ? 3 * TJ + 3 * TI : 0;
      for (int t = tid; t < nF + nP + nD; t += TI * TJ) {
        double o[5];
        if (t < 2 * TJ) {  // F_x at i0-1, i0+TI
          int r = t % TJ, side = t / TJ;
          int ii = side ? i0 + TI : i0 - 1, jj = j0 + r, col = side ? TI + 1 : 0;
          if (ii <= d.ib && jj <= d.jb) {
            cell_flux<0>(c, ii, jj, k, o);
            for (int n = 0; n < 5; n++) sm.fx[n][r][col] = o[n];
          }
        } else if (t < nF) {  // F_y at j0-1, j0+TJ
          int u = t - 2 * TJ;
          int col = u % TI, side = u / TI;
          int ii = i0 + col, jj = side ? j0 + TJ : j0 - 1, row = side ? TJ + 1 : 0;
          if (ii <= d.ib && jj <= d.jb) {
            cell_flux<1>(c, ii, jj, k, o);
            for (int n = 0; n < 5; n++) sm.fy[n][row][col] = o[n];
          }
        } else if (t < nF + 4 * TJ) {  // fp_x at i0-2, i0-1, i0+TI, i0+TI+1
          int u = t - nF;
          int r = u % TJ, w4 = u / TJ;
          int off = w4 < 2 ? w4 - 2 : TI + (w4 - 2);
          int ii = i0 + off, jj = j0 + r, col = off + 2;
          if (jj <= d.jb) sm.px[r][col] = (ii >= 0 && ii <= d.ib) ? fp_at<0>(c, ix(d, 0, jj, k), ii) : 0.0;
        } else if (t < nF + nP) {  // fp_y at j0-2, j0-1, j0+TJ, j0+TJ+1
          int u = t - nF - 4 * TJ;
          int col = u % TI, w4 = u / TI;
          int off = w4 < 2 ? w4 - 2 : TJ + (w4 - 2);
          int ii = i0 + col, jj = j0 + off, row = off + 2;
          if (ii <= d.ib) sm.py[row][col] = (jj >= 0 && jj <= d.jb) ? fp_at<1>(c, ix(d, ii, 0, k), jj) : 0.0;
        } else if (t < nF + nP + 3 * TJ) {  // diw_x at i0-1, i0+TI, i0+TI+1
          int u = t - nF - nP;
          int r = u % TJ, w3 = u / TJ;
          int off = w3 == 0 ? -1 : TI + (w3 - 1);
          int ii = i0 + off, jj = j0 + r, col = off + 1;
          if (jj <= d.jb && ii <= d.ib) {
            diw_at<0>(c, ix(d, 0, jj, k), ii, o);
            for (int n = 0; n < 5; n++) sm.dx[n][r][col] = o[n];
          }
        } else {  // diw_y at j0-1, j0+TJ, j0+TJ+1
          int u = t - nF - nP - 3 * TJ;
          int col = u % TI, w3 = u / TI;
          int off = w3 == 0 ? -1 : TJ + (w3 - 1);
          int ii = i0 + col, jj = j0 + off, row = off + 1;
          if (ii <= d.ib && jj <= d.jb) {
            diw_at<1>(c, ix(d, ii, 0, k), jj, o);
            for (int n = 0; n < 5; n++) sm.dy[n][row][col] = o[n];
          }
        }
      }
    }
    __syncthreads();
    if (valid) {
      const long q = ix(d, i, j, k);
      double f[5];
      if (A.m > 1) {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = A.fw[n * E + q] * A.blend;
      } else {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = 0.0;
      }
      double Dk1[5], Ek1[5];
      if (diss) {
        // face coefficients (dflux.f90:88-96), then per component the two
        // faces of each direction straight from LDS
        double da2, da4, dal, db2, db4, dbl;
        const long brow = ix(d, 0, j, k), bcol = ix(d, i, 0, k);
        {
          double r_m = rad_at<0>(c, brow, i - 1), r_0 = rad_at<0>(c, brow, i), r_p = rad_at<0>(c, brow, i + 1);
          face_coef(c, sm.px[ty][tx], sm.px[ty][tx + 1], sm.px[ty][tx + 2], sm.px[ty][tx + 3], r_m, r_0, &da2, &da4,
                    &dal);
          face_coef(c, sm.px[ty][tx + 1], sm.px[ty][tx + 2], sm.px[ty][tx + 3], sm.px[ty][tx + 4], r_0, r_p, &db2,
                    &db4, &dbl);
#pragma unroll
          for (int n = 0; n < 5; n++) {
            double d0 = sm.dx[n][ty][tx], d1 = sm.dx[n][ty][tx + 1], d2 = sm.dx[n][ty][tx + 2],
                   d3 = sm.dx[n][ty][tx + 3];
            double fa = (da2 * d1 - da4 * (d2 - 2.0 * d1 + d0)) * dal;
            double fb = (db2 * d2 - db4 * (d3 - 2.0 * d2 + d1)) * dbl;
            f[n] = f[n] + (fa - fb);
          }
        }
        {
          double r_m = rad_at<1>(c, bcol, j - 1), r_0 = rad_at<1>(c, bcol, j), r_p = rad_at<1>(c, bcol, j + 1);
          face_coef(c, sm.py[ty][tx], sm.py[ty + 1][tx], sm.py[ty + 2][tx], sm.py[ty + 3][tx], r_m, r_0, &da2, &da4,
                    &dal);
          face_coef(c, sm.py[ty + 1][tx], sm.py[ty + 2][tx], sm.py[ty + 3][tx], sm.py[ty + 4][tx], r_0, r_p, &db2,
                    &db4, &dbl);
#pragma unroll
          for (int n = 0; n < 5; n++) {
            double d0 = sm.dy[n][ty][tx], d1 = sm.dy[n][ty + 1][tx], d2 = sm.dy[n][ty + 2][tx],
                   d3 = sm.dy[n][ty + 3][tx];
            double fa = (da2 * d1 - da4 * (d2 - 2.0 * d1 + d0)) * dal;
            double fb = (db2 * d2 - db4 * (d3 - 2.0 * d2 + d1)) * dbl;
            f[n] = f[n] + (fa - fb);
          }
        }
        face_coef(c, fz[0], fz[1], fz[2], fz[3], rz[0], rz[1], &da2, &da4, &dal);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          Dk1[n] = (da2 * dzr[1][n] - da4 * (dzr[2][n] - 2.0 * dzr[1][n] + dzr[0][n])) * dal;
          f[n] = f[n] + (Dk[n] - Dk1[n]);
        }
      }
      double e[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        double ea = 0.5 * (sm.fx[n][ty][tx + 1] + sm.fx[n][ty][tx]);
        double eb = 0.5 * (sm.fx[n][ty][tx + 2] + sm.fx[n][ty][tx + 1]);
        e[n] = ea - eb;
      }
#pragma unroll
      for (int n = 0; n < 5; n++) {
        double ea = 0.5 * (sm.fy[n][ty + 1][tx] + sm.fy[n][ty][tx]);
        double eb = 0.5 * (sm.fy[n][ty + 2][tx] + sm.fy[n][ty + 1][tx]);
        e[n] = e[n] + ea - eb;
      }
#pragma unroll
      for (int n = 0; n < 5; n++) {
        Ek1[n] = 0.5 * (Fz[1][n] + Fz[0][n]);
        e[n] = e[n] + Ek[n] - Ek1[n];
      }
      if (A.flags[q] & F_TOUCHED) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          A.fw[n * E + q] = f[n];
          A.dw[n * E + q] = e[n];
        }
      } else {
        finish_cell(A, q, i, j, k, f, e);
      }
      // advance the k-direction carry to face k+2
      if (k < kb1) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          Ek[n] = Ek1[n];
          Fz[0][n] = Fz[1][n];
        }
        cell_flux<2>(c, i, j, k + 2, Fz[1]);
        if (diss) {
#pragma unroll
          for (int n = 0; n < 5; n++) {
            Dk[n] = Dk1[n];
            dzr[0][n] = dzr[1][n];
            dzr[1][n] = dzr[2][n];
          }
          fz[0] = fz[1];
          fz[1] = fz[2];
          fz[2] = fz[3];
          fz[3] = fp_at<2>(c, bz, k + 3);
          diw_at<2>(c, bz, k + 3, dzr[2]);
          rz[0] = rz[1];
          rz[1] = rad_at<2>(c, bz, k + 2);
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ wall records

template <int DIR>
__device__ inline void rec_dfs(const FluxCtx &c, int i, int j, int k, int a, double *o) {
  long base = DIR == 0 ? ix(c.d, 0, j, k) : (DIR == 1 ? ix(c.d, i, 0, k) : ix(c.d, i, j, 0));
  dface<DIR>(c, base, a, i, j, k, o);
}
__device__ inline void rec_dfs_any(const FluxCtx &c, int zz, int i, int j, int k, int a, double *o) {
  if (zz == 1) rec_dfs<0>(c, i, j, k, a, o);
  else if (zz == 2) rec_dfs<1>(c, i, j, k, a, o);
  else rec_dfs<2>(c, i, j, k, a, o);
}
template <int DIR>
__device__ inline void rec_efs(const FluxCtx &c, int i, int j, int k, int a, double *o) {
  long base = DIR == 0 ? ix(c.d, 0, j, k) : (DIR == 1 ? ix(c.d, i, 0, k) : ix(c.d, i, j, 0));
  eface<DIR>(c, base, a, i, j, k, o);
}
__device__ inline void rec_efs_any(const FluxCtx &c, int zz, int i, int j, int k, int a, double *o) {
  if (zz == 1) rec_efs<0>(c, i, j, k, a, o);
  else if (zz == 2) rec_efs<1>(c, i, j, k, a, o);
  else rec_efs<2>(c, i, j, k, a, o);
}

__device__ inline double face_s_any(const Geo &g, int zz, int i, int j, int k) {
  if (zz == 1) return face_s<0>(g, i, j, k);
  if (zz == 2) return face_s<1>(g, i, j, k);
  return face_s<2>(g, i, j, k);
}

// the two recomputed faces of dfluxb.f90 for record r: fs1 (inner face of
// the pair) and fs2 (the wall face or the face behind it)
__device__ void dfluxb_rec(const FluxCtx &c, const DevRec &r, const double *wsolv, double psol,
                           double *fs1, double *fs2) {
  const Dims &d = c.d;
  const long E = d.E;
  int zz = r.zz, id = r.faceid;
  int i = r.c[0], j = r.c[1], k = r.c[2];
  int s0 = zz == 1 ? i : (zz == 2 ? j : k);
  long st = zz == 1 ? 1L : (zz == 2 ? d.sy : d.sz);
  long base = ix(d, i, j, k) - (long)s0 * st;
  int nbmax = zz == 1 ? d.ib : (zz == 2 ? d.jb : d.kb);
  double wsol[5];
  for (int n = 0; n < 5; n++) wsol[n] = wsolv[n];
  wsol[4] = wsol[4] + psol;
  auto P = [&](int s) { return c.s.p[base + (long)s * st]; };
  auto H = [&](int s, int n) {
    long q = base + (long)s * st;
    return n < 4 ? c.s.w[n * E + q] : c.s.w[4 * E + q] + c.s.p[q];
  };
  auto RS = [&](int s) { return c.radsum[base + (long)s * st]; };
  auto SS = [&](int s, double *ss) {
    ss[0] = ss[1] = ss[2] = 0.0;
    int fi = zz == 1 ? s : i, fj = zz == 2 ? s : j, fk = zz == 3 ? s : k;
    ss[zz - 1] = face_s_any(c.g, zz, fi, fj, fk);
  };
  double fp[3], ep, lm, lp, lam, dis2, dis4, diw[3][5];
  int hi = (id >= 4);
  int s = s0;
  if (hi) {
    if (zz == 1) {
      fp[0] = fabs(P(s - 2) - 2.0 * P(s - 1) + P(s)) / fabs(P(s - 2) + 2.0 * P(s - 1) + P(s));
      fp[1] = fabs(P(s - 1) - 2.0 * P(s) + psol) / fabs(P(s - 1) + 2.0 * P(s) + psol);
    } else {
      fp[0] = fabs(P(s) - 2.0 * P(s - 1) + P(s - 2)) / fabs(P(s) + 2.0 * P(s - 1) + P(s - 2));
      fp[1] = fabs(psol - 2.0 * P(s) + P(s - 1)) / fabs(psol + 2.0 * P(s) + P(s - 1));
    }
    ep = maxd(fp[0], fp[1]);
    lm = RS(s);
    lam = lm;
    dis2 = c.vis2 * ep; dis4 = c.vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * c.cdis; dis4 = dis4 * c.cdis;
    for (int n = 0; n < 5; n++) {
      diw[0][n] = H(s, n) - H(s - 1, n);
      diw[1][n] = wsol[n] - H(s, n);
      diw[2][n] = diw[1][n];
      fs2[n] = diw[2][n] - 2.0 * diw[1][n] + diw[0][n];
      fs2[n] = dis2 * diw[1][n] - dis4 * fs2[n];
    }
    if (c.dissip == 0) {
      for (int n = 0; n < 5; n++) fs2[n] = fs2[n] * lam;
    } else {
      double uc[5], ss[3], o[5];
      for (int n = 0; n < 5; n++) uc[n] = H(s, n);
      SS(s + 1, ss);
      dmatrix(c.gamma, uc, wsol, fs2, ss, o);
      for (int n = 0; n < 5; n++) fs2[n] = o[n];
    }
    if ((s - 3) >= 0)
      fp[2] = fabs(P(s - 1) - 2.0 * P(s - 2) + P(s - 3)) / fabs(P(s - 1) + 2.0 * P(s - 2) + P(s - 3));
    else
      fp[2] = 0.0;
    ep = maxd(maxd(fp[0], fp[1]), fp[2]);
    lm = RS(s - 1);
    lp = RS(s);
    lam = 0.5 * (lm + lp);
    dis2 = c.vis2 * ep; dis4 = c.vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * c.cdis; dis4 = dis4 * c.cdis;
    for (int n = 0; n < 5; n++) {
      diw[0][n] = H(s - 1, n) - H(s - 2, n);
      diw[1][n] = H(s, n) - H(s - 1, n);
      diw[2][n] = wsol[n] - H(s, n);
      fs1[n] = diw[2][n] - 2.0 * diw[1][n] + diw[0][n];
      fs1[n] = dis2 * diw[1][n] - dis4 * fs1[n];
    }
    if (c.dissip == 0) {
      for (int n = 0; n < 5; n++) fs1[n] = fs1[n] * lam;
    } else {
      double um[5], uc[5], ss[3], o[5];
      for (int n = 0; n < 5; n++) { um[n] = H(s - 1, n); uc[n] = H(s, n); }
      SS(s, ss);
      dmatrix(c.gamma, um, uc, fs1, ss, o);
      for (int n = 0; n < 5; n++) fs1[n] = o[n];
    }
  } else {
    fp[0] = fabs(P(s + 1) - 2.0 * P(s) + psol) / fabs(P(s + 1) + 2.0 * P(s) + psol);
    fp[1] = fabs(P(s + 2) - 2.0 * P(s + 1) + P(s)) / fabs(P(s + 2) + 2.0 * P(s + 1) + P(s));
    if ((s + 3) <= nbmax)
      fp[2] = fabs(P(s + 3) - 2.0 * P(s + 2) + P(s + 1)) / fabs(P(s + 3) + 2.0 * P(s + 2) + P(s + 1));
    else
      fp[2] = 0.0;
    ep = maxd(maxd(fp[0], fp[1]), fp[2]);
    lm = RS(s);
    lp = RS(s + 1);
    lam = 0.5 * (lm + lp);
    dis2 = c.vis2 * ep; dis4 = c.vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * c.cdis; dis4 = dis4 * c.cdis;
    for (int n = 0; n < 5; n++) {
      diw[0][n] = H(s, n) - wsol[n];
      diw[1][n] = H(s + 1, n) - H(s, n);
      diw[2][n] = H(s + 2, n) - H(s + 1, n);
      fs2[n] = diw[2][n] - 2.0 * diw[1][n] + diw[0][n];
      fs2[n] = dis2 * diw[1][n] - dis4 * fs2[n];
    }
    if (c.dissip == 0) {
      for (int n = 0; n < 5; n++) fs2[n] = fs2[n] * lam;
    } else {
      double uc[5], up[5], ss[3], o[5];
      for (int n = 0; n < 5; n++) { uc[n] = H(s, n); up[n] = H(s + 1, n); }
      SS(s + 1, ss);
      dmatrix(c.gamma, uc, up, fs2, ss, o);
      for (int n = 0; n < 5; n++) fs2[n] = o[n];
    }
    ep = maxd(fp[0], fp[1]);
    lp = RS(s);
    lam = lp;
    dis2 = c.vis2 * ep; dis4 = c.vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * c.cdis; dis4 = dis4 * c.cdis;
    for (int n = 0; n < 5; n++) {
      diw[1][n] = H(s, n) - wsol[n];
      diw[0][n] = diw[1][n];
      diw[2][n] = H(s + 1, n) - H(s, n);
      fs1[n] = diw[2][n] - 2.0 * diw[1][n] + diw[0][n];
      fs1[n] = dis2 * diw[1][n] - dis4 * fs1[n];
    }
    if (c.dissip == 0) {
      for (int n = 0; n < 5; n++) fs1[n] = fs1[n] * lam;
    } else {
      double uc[5], ss[3], o[5];
      for (int n = 0; n < 5; n++) uc[n] = H(s, n);
      SS(s, ss);
      dmatrix(c.gamma, wsol, uc, fs1, ss, o);
      for (int n = 0; n < 5; n++) fs1[n] = o[n];
    }
  }
}

// the wall flux fs and the opposite-face flux fa of efluxb.f90:235-253
__device__ void efluxb_rec(const FluxCtx &c, const DevRec &r, const double *wsol, double psol,
                           double *fs, double *fa, int *err) {
  const Dims &d = c.d;
  int zz = r.zz, id = r.faceid;
  int i = r.c[0], j = r.c[1], k = r.c[2];
  long qc = ix(d, i, j, k);
  double sc[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
  sc[zz - 1] = face_s_any(c.g, zz, i, j, k);
  ss[zz - 1] = face_s_any(c.g, zz, r.s[0], r.s[1], r.s[2]);
  double wc[5];
  load5(c.s.w, d.E, qc, wc);
  if (fabs(wc[0]) < (double)1.0e-8f) raise_err(err, MGE_ERR_WALL_RHO);
  double fs1[5], fs2[5], fs3[5];
  cflux3(wc, c.s.p[qc], sc, fs1);
  cflux3(wsol, psol, ss, fs2);
  for (int n = 0; n < 5; n++) fs[n] = 0.5 * (fs1[n] + fs2[n]);
  int hi = (id >= 4);
  int oi = i, oj = j, ok = k;
  int dlt = hi ? -1 : 1;
  if (zz == 1) oi += dlt;
  else if (zz == 2) oj += dlt;
  else ok += dlt;
  long qo = ix(d, oi, oj, ok);
  double wo[5], so[3] = {0, 0, 0};
  load5(c.s.w, d.E, qo, wo);
  so[zz - 1] = face_s_any(c.g, zz, oi, oj, ok);
  cflux3(wo, c.s.p[qo], so, fs3);
  for (int n = 0; n < 5; n++) fa[n] = 0.5 * (fs1[n] + fs3[n]);
}

// the value each op contributes (dfluxb.f90 / efluxb.f90 terms), one
// thread per op; k_rec_apply adds them in the reference order
__global__ void __launch_bounds__(64) k_rec_ops(RecArgs R) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= R.nops) return;
  const StageArgs &A = R.st;
  int2 op = R.ops[o];
  double *out = R.opval + (long)o * 10;
  const DevRec &r = R.recs[op.x];
  if (op.y < OP_E_UNDO && !A.has_diss) return;
  int zz = r.zz;
  int s0 = zz == 1 ? r.c[0] : (zz == 2 ? r.c[1] : r.c[2]);
  double ws[5];
  for (int n = 0; n < 5; n++) ws[n] = R.wsol[n * R.nmir + op.x];
  double ps = R.psol[op.x];
  double a[5], b[5], fs1[5], fs2[5];
  int ci = r.c[0], cj = r.c[1], ck = r.c[2];
  switch (op.y) {
    case OP_HI_C_UNDO:
    case OP_LO_C_UNDO:
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0, a);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 + 1, b);
      for (int n = 0; n < 5; n++) out[n] = a[n] - b[n];
      break;
    case OP_HI_N_UNDO:
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 - 1, a);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0, b);
      for (int n = 0; n < 5; n++) out[n] = a[n] - b[n];
      break;
    case OP_LO_N_UNDO:
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 + 1, a);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 + 2, b);
      for (int n = 0; n < 5; n++) out[n] = a[n] - b[n];
      break;
    case OP_HI_C_ADD:
    case OP_LO_C_ADD:
      dfluxb_rec(A.fx, r, ws, ps, fs1, fs2);
      for (int n = 0; n < 5; n++) out[n] = fs1[n] - fs2[n];
      break;
    case OP_HI_N_ADD:
      dfluxb_rec(A.fx, r, ws, ps, fs1, fs2);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 - 1, a);
      for (int n = 0; n < 5; n++) out[n] = a[n] - fs1[n];
      break;
    case OP_LO_N_ADD:
      dfluxb_rec(A.fx, r, ws, ps, fs1, fs2);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 + 2, a);
      for (int n = 0; n < 5; n++) out[n] = fs2[n] - a[n];
      break;
    case OP_E_UNDO:
      rec_efs_any(A.fx, zz, ci, cj, ck, s0, a);
      rec_efs_any(A.fx, zz, ci, cj, ck, s0 + 1, b);
      for (int n = 0; n < 5; n++) out[n] = a[n] - b[n];
      break;
    case OP_E_ADD: {
      double fs[5], fa[5];
      efluxb_rec(A.fx, r, ws, ps, fs, fa, R.err);
      bool hi = r.faceid >= 4;
      for (int n = 0; n < 5; n++) {
        out[n] = hi ? fa[n] : fs[n];
        out[5 + n] = hi ? fs[n] : fa[n];
      }
      break;
    }
    default:
      break;
  }
}

// per touched cell: apply the op values in record order, then finish
__global__ void __launch_bounds__(64) k_rec_apply(RecArgs R) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.ntouch) return;
  const StageArgs &A = R.st;
  const Dims &d = A.fx.d;
  const long E = d.E;
  long q = R.tcell[t];
  int k = (int)(q / d.sz);
  int j = (int)((q - (long)k * d.sz) / d.sy);
  int i = (int)(q - (long)k * d.sz - (long)j * d.sy);
  double f[5], e[5];
  for (int n = 0; n < 5; n++) {
    f[n] = A.fw[n * E + q];
    e[n] = A.dw[n * E + q];
  }
  for (int o = R.beg[t]; o < R.beg[t + 1]; o++) {
    int kind = R.ops[o].y;
    const double *v = R.opval + (long)o * 10;
    switch (kind) {
      case OP_HI_C_UNDO:
      case OP_HI_N_UNDO:
      case OP_LO_C_UNDO:
      case OP_LO_N_UNDO:
        if (A.has_diss)
          for (int n = 0; n < 5; n++) f[n] = f[n] - v[n];
        break;
      case OP_HI_C_ADD:
      case OP_HI_N_ADD:
      case OP_LO_C_ADD:
      case OP_LO_N_ADD:
        if (A.has_diss)
          for (int n = 0; n < 5; n++) f[n] = f[n] + v[n];
        break;
      case OP_E_UNDO:
        for (int n = 0; n < 5; n++) e[n] = e[n] - v[n];
        break;
      case OP_E_ADD:
        for (int n = 0; n < 5; n++) e[n] = e[n] + v[n] - v[5 + n];
        break;
      default:
        break;
    }
  }
  finish_cell(A, q, i, j, k, f, e);
}

// ------------------------------------------------------------------ smoothing

struct SmoothArgs {
  Dims d;
  double *dw;
  const uint8_t *flags;
  const double *e1, *e2;  // pivots of this direction (psmoo.f90:134-166)
  const double *fa;       // -bb/e1 (x lines: affine-map form of the forward step)
  double bb;
  // z pass fuses the update
  const double *wn;
  double *wdst, *pdst;
  double gamma;
};

// psmoo_old.f90:171-188 (x lines), one wavefront per (j,k) line and
// component.  Lane l owns CPL consecutive cells, so loads and stores are
// coalesced.  Both recurrences are first composed as affine maps
// (forward y_i = A_i y_(i-1) + B_i, backward z_i = C_i z_(i+1) + y_i),
// scanned across the wave, and then every lane re-evaluates its own cells
// with the reference formula from the scanned boundary value.  Only the
// lane-boundary values differ from the sequential evaluation (by a few
// ulp); blank cells keep their value as in the reference.
template <int CPL>
__global__ void __launch_bounds__(256) k_smooth_x_scan(SmoothArgs S) {
  const Dims &d = S.d;
  const int lane = threadIdx.x & 63;
  const long wave = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6;
  const long nl = (long)d.nj * d.nk;
  if (wave >= nl * 5) return;  // uniform per wave
  const int n = (int)(wave / nl);
  const long l = wave - (long)n * nl;
  const int j = 2 + (int)(l % d.nj), k = 2 + (int)(l / d.nj);
  double *D = S.dw + n * d.E;
  const long b0 = ix(d, 0, j, k);
  const int il = d.il, ni = d.ni;
  const int i0 = 2 + lane * CPL;
  double r[CPL];
  bool fl[CPL], ok[CPL];
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    int i = i0 + c;
    ok[c] = i <= il;
    r[c] = ok[c] ? D[b0 + i] : 0.0;
    fl[c] = ok[c] && (S.flags[b0 + i] & F_FLUID);
  }
  // forward: lane composite of y_i = A_i y_(i-1) + B_i
  double Al = 1.0, Bl = 0.0;
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    int i = i0 + c;
    double A, B;
    if (!ok[c]) { A = 1.0; B = 0.0; }
    else if (!fl[c]) { A = 0.0; B = r[c]; }
    else if (i == 2) { A = 0.0; B = r[c] / S.e1[2]; }
    else { A = S.fa[i]; B = r[c] / S.e1[i]; }
    Bl = A * Bl + B;
    Al = A * Al;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    double Ap = __shfl_up(Al, off, 64), Bp = __shfl_up(Bl, off, 64);
    if (lane >= off) {
      Bl = Al * Bp + Bl;
      Al = Al * Ap;
    }
  }
  double yin = __shfl_up(Bl, 1, 64);
  if (lane == 0) yin = 0.0;
  double y[CPL];
  double yp = yin;
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    int i = i0 + c;
    double v = r[c];
    if (ok[c] && fl[c]) v = (i == 2) ? r[c] / S.e1[2] : (r[c] - S.bb * yp) / S.e1[i];
    y[c] = v;
    yp = v;
  }
  // backward: z_i = C_i z_(i+1) + y_i, C_i = -e2_i on fluid cells 2..ni
  double Cl = 1.0, Dl = 0.0;
#pragma unroll
  for (int c = CPL - 1; c >= 0; c--) {
    int i = i0 + c;
    double C, Dv;
    if (!ok[c]) { C = 1.0; Dv = 0.0; }
    else if (fl[c] && i <= ni) { C = -S.e2[i]; Dv = y[c]; }
    else { C = 0.0; Dv = y[c]; }
    Dl = C * Dl + Dv;
    Cl = C * Cl;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    double Cn = __shfl_down(Cl, off, 64), Dn = __shfl_down(Dl, off, 64);
    if (lane + off < 64) {
      Dl = Cl * Dn + Dl;
      Cl = Cl * Cn;
    }
  }
  double zin = __shfl_down(Dl, 1, 64);
  if (lane == 63) zin = 0.0;
  double zn = zin;
#pragma unroll
  for (int c = CPL - 1; c >= 0; c--) {
    int i = i0 + c;
    if (!ok[c]) continue;
    double v = y[c];
    if (fl[c] && i <= ni) v = y[c] - S.e2[i] * zn;
    zn = v;
    if (fl[c]) D[b0 + i] = v;
  }
}

// sequential x lines for blocks wider than 512 cells
__global__ void __launch_bounds__(256) k_smooth_x_seq(SmoothArgs S) {
  const Dims &d = S.d;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long nl = (long)d.nj * d.nk;
  if (t >= nl * 5) return;
  int n = (int)(t / nl);
  long l = t - (long)n * nl;
  int j = 2 + (int)(l % d.nj), k = 2 + (int)(l / d.nj);
  double *D = S.dw + n * d.E;
  long b0 = ix(d, 0, j, k);
  if (S.flags[b0 + 2] & F_FLUID) D[b0 + 2] = D[b0 + 2] / S.e1[2];
  for (int i = 3; i <= d.il; i++)
    if (S.flags[b0 + i] & F_FLUID) D[b0 + i] = (D[b0 + i] - S.bb * D[b0 + i - 1]) / S.e1[i];
  for (int i = 2; i <= d.ni; i++) {
    int ir = d.il + 1 - i;
    if (S.flags[b0 + ir] & F_FLUID) D[b0 + ir] = D[b0 + ir] - S.e2[ir] * D[b0 + ir + 1];
  }
}

// One line of the exact sequential recurrence of psmoo_old
// (psmoo.f90:190-237): cells s = 2..nl at stride st.  Full chunks of CH
// cells are read into registers double-buffered with unconditional loads,
// so the compiler can keep the next chunk's loads in flight (counted waits)
// while the current chunk runs its dependent divisions.  The updates are
// branch-free: blank cells hold dw = +-0 (blanked before smoothing), so
// y = m*((r - bb*y)/e1) with m = 0 on blank cells gives the reference's
// values on fluid cells and keeps the zeros.  The backward sweep starts
// at s = nl with z(nl+1) = 0 and e2(nl) = 0, which returns y(nl) there.
template <int CH>
__device__ __attribute__((always_inline)) inline void line_sweep(double *__restrict__ D,
                                                                 const uint8_t *__restrict__ F, long b0, long st,
                                                                 int nl, const double *__restrict__ e1,
                                                                 const double *__restrict__ e2, double bb) {
  const int ncell = nl - 1;
  const int nfull = ncell / CH;
  double A[CH], B[CH], MA[CH], MB[CH];
  double y = 0.0;
#define LD_F(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    long q_ = b0 + (long)(2 + (c) * CH + t) * st;                     \
    BUF[t] = D[q_];                                                   \
    MK[t] = (F[q_] & F_FLUID) ? 1.0 : 0.0;                            \
  }
#define PR_F(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    int s_ = 2 + (c) * CH + t;                                        \
    y = MK[t] * ((BUF[t] - bb * y) / e1[s_]);                         \
    D[b0 + (long)s_ * st] = y;                                        \
  }
#define LD_B(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    long q_ = b0 + (long)(nl - (c) * CH - t) * st;                    \
    BUF[t] = D[q_];                                                   \
    MK[t] = (F[q_] & F_FLUID) ? 1.0 : 0.0;                            \
  }
#define PR_B(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    int s_ = nl - (c) * CH - t;                                       \
    y = BUF[t] - MK[t] * (e2[s_] * y);                                \
    D[b0 + (long)s_ * st] = y;                                        \
  }
  // forward
  if (nfull > 0) {
    LD_F(A, MA, 0);
    for (int c = 0; c < nfull; c += 2) {
      int c1 = c + 1 < nfull ? c + 1 : c;
      LD_F(B, MB, c1);
      PR_F(A, MA, c);
      if (c + 1 >= nfull) break;
      int c2 = c + 2 < nfull ? c + 2 : c + 1;
      LD_F(A, MA, c2);
      PR_F(B, MB, c + 1);
    }
  }
  for (int s_ = 2 + nfull * CH; s_ <= nl; s_++) {
    long q_ = b0 + (long)s_ * st;
    double m_ = (F[q_] & F_FLUID) ? 1.0 : 0.0;
    y = m_ * ((D[q_] - bb * y) / e1[s_]);
    D[q_] = y;
  }
  // backward
  y = 0.0;
  if (nfull > 0) {
    LD_B(A, MA, 0);
    for (int c = 0; c < nfull; c += 2) {
      int c1 = c + 1 < nfull ? c + 1 : c;
      LD_B(B, MB, c1);
      PR_B(A, MA, c);
      if (c + 1 >= nfull) break;
      int c2 = c + 2 < nfull ? c + 2 : c + 1;
      LD_B(A, MA, c2);
      PR_B(B, MB, c + 1);
    }
  }
  for (int s_ = nl - nfull * CH; s_ >= 2; s_--) {
    long q_ = b0 + (long)s_ * st;
    double m_ = (F[q_] & F_FLUID) ? 1.0 : 0.0;
    y = D[q_] - m_ * (e2[s_] * y);
    D[q_] = y;
  }
#undef LD_F
#undef PR_F
#undef LD_B
#undef PR_B
}

// psmoo_old.f90:190-206 (y lines); thread per (i,k) and component
__global__ void __launch_bounds__(64) k_smooth_y(SmoothArgs S) {
  const Dims &d = S.d;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int k = 2 + blockIdx.y;
  int n = blockIdx.z;
  if (i > d.il) return;
  line_sweep<16>(S.dw + n * d.E, S.flags, ix(d, i, 0, k), d.sy, d.jl, S.e1, S.e2, S.bb);
}

// psmoo_old.f90:211-237 (z lines); thread per (i,j) and component
__global__ void __launch_bounds__(64) k_smooth_z(SmoothArgs S) {
  const Dims &d = S.d;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y;
  int n = blockIdx.z;
  if (i > d.il) return;
  line_sweep<16>(S.dw + n * d.E, S.flags, ix(d, i, j, 0), d.sz, d.kl, S.e1, S.e2, S.bb);
}

// after smoothing: w = wn - dw and the EOS on interior cells
// (euler.f90:381-399); the ghost shell of the previous state is carried
// over, as k_ghost_copy does on unsmoothed stages
__global__ void __launch_bounds__(256) k_update_ghost(Dims d, const double *wn, const double *dw, StateC src,
                                                      State dst, double gamma) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  int k = (int)(q / d.sz);
  int j = (int)((q - (long)k * d.sz) / d.sy);
  int i = (int)(q - (long)k * d.sz - (long)j * d.sy);
  bool inner = i >= 2 && i <= d.il && j >= 2 && j <= d.jl && k >= 2 && k <= d.kl;
  if (inner) {
    double v[5];
    for (int n = 0; n < 5; n++) {
      v[n] = wn[n * d.E + q] - dw[n * d.E + q] / 1.0;
      dst.w[n * d.E + q] = v[n];
    }
    dst.p[q] = eos(gamma, v);
  } else {
    for (int n = 0; n < 5; n++) dst.w[n * d.E + q] = src.w[n * d.E + q];
    dst.p[q] = src.p[q];
  }
}

// ------------------------------------------------------------------ timestep

struct TsArgs {
  Dims d;
  StateC s;
  Geo g;
  const uint8_t *flags;
  double *radsum, *vdtim;
  unsigned long long *dtim_bits;
  double gamma, cfl, cellvolume;
  int eqntype;
  double viscl, prlmean;
  int *err;
};

// timestep.f90:36-115 (Euler) and :117-209 (NS): spectral radii, local
// time step with the adaptive factor, global minimum; blank cells zeroed
// (:212-223).  The ghost copies of the radii (:227-282) are realised by
// index clamping where the radii are read.
__global__ void __launch_bounds__(256) k_timestep(TsArgs T) {
  const Dims &d = T.d;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 2 + blockIdx.z;
  double mine = 10.0;
  if (i <= d.il && j <= d.jl && k <= d.kl) {
    const long E = d.E;
    long q = ix(d, i, j, k);
    double rho = T.s.w[q];
    double cc = T.gamma * T.s.p[q] / rho;
    if (cc < 0.0 || cc != cc) raise_err(T.err, MGE_ERR_NEG_C2);
    cc = sqrt(maxd(cc, 0.0));
    double uu = T.s.w[E + q] / rho;
    double uv = T.s.w[2 * E + q] / rho;
    double uw = T.s.w[3 * E + q] / rho;
    double dtv = 0.0;
    if (T.eqntype == 1) {
      double fmue = T.viscl / T.prlmean;
      double f1 = 4.0 / rho / 3.0;
      double f2 = T.gamma / rho;
      double fac = maxd(f1, f2);
      dtv = fac * fmue / T.cellvolume;
    }
    double sx = face_s<0>(T.g, i, j, k);
    double sy = face_s<1>(T.g, i, j, k);
    double sz = face_s<2>(T.g, i, j, k);
    double qs, ds, cs;
    qs = fabs(uu * sx);
    ds = sx * sx;
    cs = cc * sqrt(ds);
    double ri = qs + cs, rvi = dtv * ds;
    qs = fabs(uv * sy);
    ds = sy * sy;
    cs = cc * sqrt(ds);
    double rj = qs + cs, rvj = dtv * ds;
    qs = fabs(uw * sz);
    ds = sz * sz;
    cs = cc * sqrt(ds);
    double rk = qs + cs, rvk = dtv * ds;
    double fict;
    if (T.eqntype == 1) fict = T.cellvolume / (ri + rj + rk + 4.0 * (rvi + rvj + rvk));
    else fict = T.cellvolume / (ri + rj + rk);
    const double plim = 0.001;
    const double *P = T.s.p;
    double pc = P[q];
    double pip = P[q + 1], pim = P[q - 1];
    double pjp = P[q + d.sy], pjm = P[q - d.sy];
    double pkp = P[q + d.sz], pkm = P[q - d.sz];
    double dpi = fabs((pip - 2.0 * pc + pim) / (pip + 2.0 * pc + pim + plim));
    double dpj = fabs((pjp - 2.0 * pc + pjm) / (pjp + 2.0 * pc + pjm + plim));
    double dpk = fabs((pkp - 2.0 * pc + pkm) / (pkp + 2.0 * pc + pkm + plim));
    double cm1 = dimd(T.cfl, 1.0);
    double sw = 2.0 * (dpi + dpj + dpk * 1.0);
    double r = 1.0 / (1.0 + (cm1 < sw ? cm1 : sw));
    double vd = fict * r;
    mine = vd;
    if (T.flags[q] & F_FLUID) {
      T.radsum[q] = ri + rj + rk;
      T.vdtim[q] = vd;
    } else {
      T.radsum[q] = 0.0 + 0.0 + 0.0;
      T.vdtim[q] = 0.0;
    }
  }
  // block minimum of the time step -> one atomic per block
  __shared__ double sm[256];
  int tid = threadIdx.x + threadIdx.y * blockDim.x;
  sm[tid] = mine;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s && sm[tid + s] < sm[tid]) sm[tid] = sm[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    double v = sm[0];
    unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    atomicMin(T.dtim_bits, bits);
  }
}

// ------------------------------------------------------------------ residual norm

struct RtArgs {
  Dims d;
  const double *dw;
  const double *vdtim;
  const double *dtim;
  const uint8_t *flags;
  double cfl;
  int vt;
  double *partial;
};

// euler.f90:431-447: sum of (dw_1/(cfl*dt))^2 over fluid cells; per-block
// partials in a fixed tree order, then k_rtrms_final adds them in order
__global__ void __launch_bounds__(256) k_rtrms(RtArgs R) {
  const Dims &d = R.d;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 2 + blockIdx.z;
  double v = 0.0;
  if (i <= d.il && j <= d.jl && k <= d.kl) {
    long q = ix(d, i, j, k);
    double de = R.cfl * R.vdtim[q];
    if (R.vt == 1) de = R.cfl * R.dtim[0];
    double rt = 0.0;
    if (R.flags[q] & F_FLUID) rt = R.dw[q] / de;
    v = rt * rt;
  }
  __shared__ double sm[256];
  int tid = threadIdx.x + threadIdx.y * blockDim.x;
  sm[tid] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) sm[tid] = sm[tid] + sm[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    R.partial[b] = sm[0];
  }
}

__global__ void __launch_bounds__(256) k_rtrms_final(const double *partial, long nb, double ncell,
                                                     double *out, int *err) {
  __shared__ double sm[256];
  double acc = 0.0;
  for (long b = threadIdx.x; b < nb; b += 256) acc = acc + partial[b];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = sm[threadIdx.x] + sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double r = sqrt(sm[0] / ncell);
    out[0] = r;
    if (r != r) raise_err(err, MGE_ERR_RES_NAN);
  }
}

__global__ void k_dtim_decode(const unsigned long long *bits, double *dtim) {
  dtim[0] = __longlong_as_double((long long)bits[0]);
}

// ------------------------------------------------------------------ bc_cutcell

struct CutArgs {
  Dims d;
  StateC s;
  const DevRec *recs;
  double *wsol, *psol;
  int nmir, nupd, zero_at;  // records < nupd are updated; zero_at gets zeros
  double gamma;
  int eqntype;
  int *err;
};

// bc_cutcell.f90:200-428 for steady bodies (qb = 0)
__global__ void __launch_bounds__(64) k_bc_cut(CutArgs C) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == C.zero_at) {
    for (int n = 0; n < 5; n++) C.wsol[n * C.nmir + m] = 0.0;
    C.psol[m] = 0.0;
    return;
  }
  if (m >= C.nupd) return;
  const DevRec &r = C.recs[m];
  const long E = C.d.E;
  double nv[3] = {r.nvec[0], r.nvec[1], r.nvec[2]};
  double qb[3] = {0.0, 0.0, 0.0};
  double s = 0.0;
  s = s + nv[0] * nv[0];
  s = s + nv[1] * nv[1];
  s = s + nv[2] * nv[2];
  double nn = sqrt(s);
  for (int dd = 0; dd < 3; dd++) nv[dd] = nv[dd] / nn;
  double vr[5] = {0, 0, 0, 0, 0}, pr = 0.0;
  for (int n = 0; n < 5; n++)
    for (int c = 0; c < r.npoint; c++) vr[n] = vr[n] + r.coef[c] * C.s.w[n * E + r.cloud[c]];
  for (int c = 0; c < r.npoint; c++) pr = pr + r.coef[c] * C.s.p[r.cloud[c]];
  double psol = pr, ws[5];
  for (int n = 0; n < 5; n++) ws[n] = vr[n];
  if (C.eqntype == 0) {
    double qn = 2.0 * (nv[0] * (ws[1] - ws[0] * qb[0]) + nv[1] * (ws[2] - ws[0] * qb[1]) +
                       nv[2] * (ws[3] - ws[0] * qb[2]));
    for (int dd = 0; dd < 3; dd++) ws[1 + dd] = ws[1 + dd] - qn * nv[dd];
  } else {
    for (int dd = 0; dd < 3; dd++) ws[1 + dd] = 2.0 * qb[dd] * vr[0] - vr[1 + dd];
  }
  double s2 = 0.0;
  s2 = s2 + ws[1] * ws[1];
  s2 = s2 + ws[2] * ws[2];
  s2 = s2 + ws[3] * ws[3];
  ws[4] = psol / (C.gamma - 1.0) + 0.5 * s2 / ws[0];
  if (C.eqntype == 0)
    for (int n = 0; n < 5; n++)
      if (ws[n] != ws[n]) raise_err(C.err, MGE_ERR_WSOL_NAN);
  for (int n = 0; n < 5; n++) C.wsol[n * C.nmir + m] = ws[n];
  C.psol[m] = psol;
}

// ------------------------------------------------------------------ bcond

struct BcArgs {
  Dims d;
  State s;
  int face;      // 1..6
  int kind;      // -20 far, -50 symmetry, -60 periodic, -80 wall
  double gamma, rho0, p0, c0, u0, v0, w0;
};

// one face of bcond.f90:16-39 -> bc_far.f90 / bc_symm.f90 / bc_periodic.f90 /
// bc_wall.f90, over the whole (n+4)^2 face like the reference loops
__global__ void __launch_bounds__(256) k_bc_face(BcArgs B) {
  const Dims &d = B.d;
  int dir = (B.face - 1) % 3;
  int amax = dir == 0 ? d.jb : d.ib;
  int cmax = dir == 2 ? d.jb : d.kb;
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  int c2 = blockIdx.y;
  if (a > amax || c2 > cmax) return;
  int nl = dir == 0 ? d.il : (dir == 1 ? d.jl : d.kl);
  int ne = nl + 1, nb = nl + 2;
  auto L = [&](int s) -> long {
    return dir == 0 ? ix(d, s, a, c2) : (dir == 1 ? ix(d, a, s, c2) : ix(d, a, c2, s));
  };
  const long E = d.E;
  double *W = B.s.w, *P = B.s.p;
  if (B.kind == -20) {
    int phy1, dum1, dum2;
    if (B.face <= 3) { phy1 = 2; dum1 = 1; dum2 = 0; }
    else { phy1 = nl; dum1 = ne; dum2 = nb; }
    double fn[3] = {0.0, 0.0, 0.0};
    fn[dir] = B.face <= 3 ? -1.0 : 1.0;
    double vinf[3] = {B.u0, B.v0, B.w0};
    const double gamma = B.gamma, gm = 1.0 / (gamma - 1.0);
    long q1 = L(phy1), qd1 = L(dum1), qd2 = L(dum2);
    double s = 0.0;
    s = s + W[E + q1] * fn[0];
    s = s + W[2 * E + q1] * fn[1];
    s = s + W[3 * E + q1] * fn[2];
    double vel_norm = s / W[q1];
    double r, vel[3], pr;
    if (vel_norm <= 0.0) {
      double rhoe = W[q1];
      double ue = W[(1 + dir) * E + q1] / rhoe;
      double pe = P[q1];
      double ce = sqrt(gamma * pe / rhoe);
      double rpo = ue * fn[dir] + 2.0 * ce / (gamma - 1.0);
      double rmi = vinf[dir] * fn[dir] - 2.0 * B.c0 / (gamma - 1.0);
      double un = 0.5 * (rpo + rmi) * fn[dir];
      double ci = 0.25 * (gamma - 1.0) * (rpo - rmi);
      vel[0] = B.u0; vel[1] = B.v0; vel[2] = B.w0;
      vel[dir] = un;
      double si = pow(B.rho0, gamma) / B.p0;
      double cc = ci * ci / gamma;
      r = pow(si * cc, gm);
      pr = r * cc;
    } else {
      double rhoi = W[q1];
      double vv[3] = {W[E + q1] / rhoi, W[2 * E + q1] / rhoi, W[3 * E + q1] / rhoi};
      double prei = P[q1];
      double si = pow(rhoi, gamma) / prei;
      double ci = sqrt(gamma * prei / rhoi);
      double rpo = vv[dir] * fn[dir] + 2.0 * ci / (gamma - 1.0);
      double rmi = vinf[dir] * fn[dir] - 2.0 * B.c0 / (gamma - 1.0);
      double un = 0.5 * (rpo + rmi) * fn[dir];
      double ce = 0.25 * (gamma - 1.0) * (rpo - rmi);
      vel[0] = vv[0]; vel[1] = vv[1]; vel[2] = vv[2];
      vel[dir] = un;
      double cc = ce * ce / gamma;
      r = pow(si * cc, gm);
      pr = r * cc;
    }
    double o[5];
    o[0] = r;
    o[1] = r * vel[0];
    o[2] = r * vel[1];
    o[3] = r * vel[2];
    o[4] = pr * gm + 0.5 * r * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    for (int n = 0; n < 5; n++) {
      W[n * E + qd1] = o[n];
      W[n * E + qd2] = o[n];
    }
    P[qd1] = pr;
    P[qd2] = pr;
    return;
  }
  int phy1, phy2, dum1, dum2;
  if (B.kind == -60) {
    if (B.face <= 3) { dum1 = 1; dum2 = 0; phy1 = nl; phy2 = nl - 1; }
    else { phy1 = 2; phy2 = 3; dum1 = ne; dum2 = nb; }
  } else {
    if (B.face <= 3) { phy1 = 2; phy2 = 3; dum1 = 1; dum2 = 0; }
    else { phy1 = nl; phy2 = nl - 1; dum1 = ne; dum2 = nb; }
  }
  long q1 = L(phy1), q2 = L(phy2), qd1 = L(dum1), qd2 = L(dum2);
  if (B.kind == -60) {
    for (int n = 0; n < 5; n++) W[n * E + qd2] = W[n * E + q2];
    for (int n = 0; n < 5; n++) W[n * E + qd1] = W[n * E + q1];
    P[qd2] = P[q2];
    P[qd1] = P[q1];
  } else if (B.kind == -50) {
    for (int n = 0; n < 5; n++) W[n * E + qd1] = W[n * E + q1];
    W[(1 + dir) * E + qd1] = -W[(1 + dir) * E + q1];
    P[qd1] = P[q1];
    for (int n = 0; n < 5; n++) W[n * E + qd2] = W[n * E + q2];
    W[(1 + dir) * E + qd2] = -W[(1 + dir) * E + q2];
    P[qd2] = P[q2];
  } else if (B.kind == -80) {
    W[qd1] = W[q1];
    W[4 * E + qd1] = W[4 * E + q1];
    for (int n = 1; n < 4; n++) W[n * E + qd1] = -W[n * E + q1];
    W[qd2] = W[q2];
    W[4 * E + qd2] = W[4 * E + q2];
    for (int n = 1; n < 4; n++) W[n * E + qd2] = -W[n * E + q2];
    P[qd1] = P[q1];
    P[qd2] = P[q1];
  }
}

// bc_corner.f90:23-157 in its four ordered phases (one workgroup)
__device__ inline void corner_eos(const Dims &d, State s, double gamma, long q) {
  double v[5];
  for (int n = 0; n < 5; n++) v[n] = s.w[n * d.E + q];
  s.p[q] = eos(gamma, v);
}

// the four ordered phases of bc_corner.f90; within a phase every (edge,
// index) item is independent, so one 1024-thread workgroup runs each phase
// in parallel and synchronises between phases
__device__ inline void edge_fill(const Dims &d, State s, double gamma, int ti, int tj, int tk, int ai, int aj,
                                 int ak, int bi, int bj, int bk, int ci, int cj, int ck) {
  const long E = d.E;
  long qt = ix(d, ti, tj, tk), qa = ix(d, ai, aj, ak), qb = ix(d, bi, bj, bk), qc = ix(d, ci, cj, ck);
  double v[5];
  for (int n = 0; n < 5; n++) {
    v[n] = s.w[n * E + qa] + s.w[n * E + qb] - s.w[n * E + qc];
    s.w[n * E + qt] = v[n];
  }
  s.p[qt] = eos(gamma, v);
}

__global__ void __launch_bounds__(1024) k_bc_corner(Dims d, State s, double gamma) {
  const long E = d.E;
  double *W = s.w;
  const int il = d.il, jl = d.jl, kl = d.kl, ie = d.ie, je = d.je, ke = d.ke, ib = d.ib;
  const int tid = threadIdx.x, nt = blockDim.x;
  // phase 1: edges along k (bc_corner.f90:23-61)
  for (int t = tid; t < 4 * (kl - 1); t += nt) {
    int e = t & 3, k = 2 + (t >> 2);
    if (e == 0) edge_fill(d, s, gamma, 1, 1, k, 2, 1, k, 1, 2, k, 2, 2, k);
    else if (e == 1) edge_fill(d, s, gamma, ie, 1, k, il, 1, k, ie, 2, k, il, 2, k);
    else if (e == 2) edge_fill(d, s, gamma, 1, je, k, 2, je, k, 1, jl, k, 2, jl, k);
    else edge_fill(d, s, gamma, ie, je, k, il, je, k, ie, jl, k, il, jl, k);
  }
  __syncthreads();
  // phase 2: edges along j (:63-99)
  for (int t = tid; t < 4 * (jl - 1); t += nt) {
    int e = t & 3, j = 2 + (t >> 2);
    if (e == 0) edge_fill(d, s, gamma, 1, j, 1, 2, j, 1, 1, j, 2, 2, j, 2);
    else if (e == 1) edge_fill(d, s, gamma, ie, j, 1, il, j, 1, ie, j, 2, il, j, 2);
    else if (e == 2) edge_fill(d, s, gamma, 1, j, ke, 2, j, ke, 1, j, kl, 2, j, kl);
    else edge_fill(d, s, gamma, ie, j, ke, il, j, ke, ie, j, kl, il, j, kl);
  }
  __syncthreads();
  // phase 3: edges along i including the ghost i (:101-137); the third
  // line adds and subtracts the same value (:120)
  for (int t = tid; t < 4 * (ib + 1); t += nt) {
    int e = t & 3, i = t >> 2;
    if (e == 0) edge_fill(d, s, gamma, i, 1, 1, i, 1, 2, i, 2, 1, i, 2, 2);
    else if (e == 1) edge_fill(d, s, gamma, i, 1, ke, i, 2, ke, i, 1, kl, i, 2, kl);
    else if (e == 2) edge_fill(d, s, gamma, i, je, 1, i, je, 1, i, jl, 2, i, jl, 2);
    else edge_fill(d, s, gamma, i, je, ke, i, jl, ke, i, je, kl, i, jl, kl);
  }
  __syncthreads();
#define WQ(n, i, j, k) W[(n) * E + ix(d, i, j, k)]
  // phase 4: the 8 corners (:141-157)
  if (tid < 40) {
    int n = tid % 5, c = tid / 5;
    if (c == 0) WQ(n, 1, 1, 1) = (WQ(n, 2, 1, 1) + WQ(n, 1, 2, 1) + WQ(n, 1, 1, 2)) * 2.0 / 3.0 - WQ(n, 2, 2, 2);
    if (c == 1) WQ(n, 1, 1, ke) = (WQ(n, 2, 1, ke) + WQ(n, 1, 2, ke) + WQ(n, 1, 1, kl)) * 2.0 / 3.0 - WQ(n, 2, 2, kl);
    if (c == 2) WQ(n, 1, je, 1) = (WQ(n, 2, je, 1) + WQ(n, 1, jl, 1) + WQ(n, 1, je, 2)) * 2.0 / 3.0 - WQ(n, 2, jl, 2);
    if (c == 3) WQ(n, ie, 1, 1) = (WQ(n, il, 1, 1) + WQ(n, ie, 2, 1) + WQ(n, ie, 1, 2)) * 2.0 / 3.0 - WQ(n, il, 2, 2);
    if (c == 4) WQ(n, ie, 1, ke) = (WQ(n, il, 1, ke) + WQ(n, ie, 2, ke) + WQ(n, ie, 1, kl)) * 2.0 / 3.0 - WQ(n, il, 2, kl);
    if (c == 5) WQ(n, ie, je, 1) = (WQ(n, il, je, 1) + WQ(n, ie, jl, 1) + WQ(n, ie, je, 2)) * 2.0 / 3.0 - WQ(n, il, jl, 2);
    if (c == 6) WQ(n, 1, je, ke) = (WQ(n, 2, je, ke) + WQ(n, 1, jl, ke) + WQ(n, 1, je, kl)) * 2.0 / 3.0 - WQ(n, 2, jl, kl);
    if (c == 7) WQ(n, ie, je, ke) = (WQ(n, il, je, ke) + WQ(n, ie, jl, ke) + WQ(n, ie, je, kl)) * 2.0 / 3.0 - WQ(n, il, jl, kl);
  }
  __syncthreads();
#undef WQ
  if (tid < 8) {
    int c = tid;
    int ci = (c & 1) ? ie : 1, cj = (c & 2) ? je : 1, ck = (c & 4) ? ke : 1;
    corner_eos(d, s, gamma, ix(d, ci, cj, ck));
  }
}

// ------------------------------------------------------------------ misc kernels

// ghost shell of src -> dst (w and p), so bcond sees the same ghost values
// that the reference's single in-place array holds at that point
__global__ void __launch_bounds__(256) k_ghost_copy(Dims d, StateC src, State dst) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  int k = (int)(q / d.sz);
  int j = (int)((q - (long)k * d.sz) / d.sy);
  int i = (int)(q - (long)k * d.sz - (long)j * d.sy);
  bool inner = i >= 2 && i <= d.il && j >= 2 && j <= d.jl && k >= 2 && k <= d.kl;
  if (inner) return;
  for (int n = 0; n < 5; n++) dst.w[n * d.E + q] = src.w[n * d.E + q];
  dst.p[q] = src.p[q];
}

__global__ void __launch_bounds__(256) k_eos_all(Dims d, State s, double gamma) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  double v[5];
  for (int n = 0; n < 5; n++) v[n] = s.w[n * d.E + q];
  s.p[q] = eos(gamma, v);
}

__global__ void __launch_bounds__(256) k_fill_freestream(Dims d, State s, double r0, double ru,
                                                         double rv, double rw, double re) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  s.w[q] = r0;
  s.w[d.E + q] = ru;
  s.w[2 * d.E + q] = rv;
  s.w[3 * d.E + q] = rw;
  s.w[4 * d.E + q] = re;
}

// ------------------------------------------------------------------ transfers

struct XferArgs {
  Dims dc, df;           // coarse, fine
  int rx, ry, rz;        // fine/coarse ratio per direction
  const double *wf, *dwf;
  double *wc;
  const double *wsc;
  double *wr;
  uint8_t *marker;
  double volc, volf;
  int fmg;               // mgl-1 == mlvl: interpolate the solution itself
  const uint8_t *flagsf;
  double *wfo;           // fine state (prolongation, in place)
};

__device__ inline void rstencil(int I, int r, int *s, double *f) {
  if (r == 2) { *s = 2 * I - 2; f[0] = 0.5; f[1] = 0.5; }
  else { *s = I; f[0] = 1.0; f[1] = 0.0; }
}
__device__ inline void pstencil(int i, int r, int *s, double *f) {
  if (r == 2) {
    if ((i & 1) == 0) { *s = i / 2; f[0] = 0.25; f[1] = 0.75; }
    else { *s = (i - 1) / 2 + 1; f[0] = 0.75; f[1] = 0.25; }
  } else { *s = i; f[0] = 1.0; f[1] = 0.0; }
}

// collc.f90:110-198 for aligned ratio-1/2 levels: w and the residual
// restricted with the cell-centre trilinear weights; marker = 1
__global__ void __launch_bounds__(256) k_restrict(XferArgs X) {
  const Dims &c = X.dc, &f = X.df;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 2 + blockIdx.z;
  if (i > c.il || j > c.jl || k > c.kl) return;
  int s[3];
  double fa[2], fb[2], fc[2], fac[8];
  rstencil(i, X.rx, &s[0], fa);
  rstencil(j, X.ry, &s[1], fb);
  rstencil(k, X.rz, &s[2], fc);
  for (int c8 = 0; c8 < 8; c8++) fac[c8] = fa[c8 & 1] * fb[(c8 >> 1) & 1] * fc[(c8 >> 2) & 1];
  long qc = ix(c, i, j, k);
  long qf[8];
  for (int c8 = 0; c8 < 8; c8++) qf[c8] = ix(f, s[0] + (c8 & 1), s[1] + ((c8 >> 1) & 1), s[2] + ((c8 >> 2) & 1));
  for (int n = 0; n < 5; n++) {
    double acc = 0.0;
    for (int c8 = 0; c8 < 8; c8++) {
      double val = X.dwf[n * f.E + qf[c8]] / X.volf;
      acc = acc + val * fac[c8];
    }
    X.wr[n * c.E + qc] = acc * X.volc;
  }
  for (int n = 0; n < 5; n++) {
    double acc = 0.0;
    for (int c8 = 0; c8 < 8; c8++) acc = acc + X.wf[n * f.E + qf[c8]] * fac[c8];
    X.wc[n * c.E + qc] = acc;
  }
  X.marker[qc] = 1;
}

// addw.f90:369-438: trilinear prolongation of the correction w-ws (or of
// the solution on the full-multigrid step) to fluid cells
__global__ void __launch_bounds__(256) k_prolong(XferArgs X) {
  const Dims &c = X.dc, &f = X.df;
  int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  int k = 2 + blockIdx.z;
  if (i > f.il || j > f.jl || k > f.kl) return;
  int s[3];
  double fa[2], fb[2], fc[2], fac[8];
  pstencil(i, X.rx, &s[0], fa);
  pstencil(j, X.ry, &s[1], fb);
  pstencil(k, X.rz, &s[2], fc);
  for (int c8 = 0; c8 < 8; c8++) fac[c8] = fa[c8 & 1] * fb[(c8 >> 1) & 1] * fc[(c8 >> 2) & 1];
  long qf = ix(f, i, j, k);
  if (!(X.flagsf[qf] & F_FLUID)) return;
  long qc[8];
  for (int c8 = 0; c8 < 8; c8++) qc[c8] = ix(c, s[0] + (c8 & 1), s[1] + ((c8 >> 1) & 1), s[2] + ((c8 >> 2) & 1));
  for (int n = 0; n < 5; n++) {
    double acc = 0.0;
    for (int c8 = 0; c8 < 8; c8++) {
      double val = X.wc[n * c.E + qc[c8]];
      if (!X.fmg) val = val - X.wsc[n * c.E + qc[c8]];
      acc = acc + val * fac[c8];
    }
    if (X.fmg) X.wfo[n * f.E + qf] = acc;
    else X.wfo[n * f.E + qf] = X.wfo[n * f.E + qf] + acc;
  }
}

// wall force: momentum flux of the scheme's wall faces (efluxb.f90:251-253);
// one thread per record writes its term, one thread adds them in record order
__global__ void k_wall_terms(Dims d, StateC s, Geo g, const DevRec *recs, const double *wsol, const double *psol,
                             int nmir, double *terms) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmir) return;
  const DevRec &r = recs[m];
  int zz = r.zz;
  double sc[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
  sc[zz - 1] = face_s_any(g, zz, r.c[0], r.c[1], r.c[2]);
  ss[zz - 1] = face_s_any(g, zz, r.s[0], r.s[1], r.s[2]);
  long qc = ix(d, r.c[0], r.c[1], r.c[2]);
  double wc[5], ws[5], f1[5], f2[5];
  load5(s.w, d.E, qc, wc);
  for (int n = 0; n < 5; n++) ws[n] = wsol[n * nmir + m];
  cflux3(wc, s.p[qc], sc, f1);
  cflux3(ws, psol[m], ss, f2);
  double sgn = (r.faceid >= 4) ? -1.0 : 1.0;
  for (int dd = 0; dd < 3; dd++) terms[3 * m + dd] = sgn * (0.5 * (f1[1 + dd] + f2[1 + dd]));
}

__global__ void k_wall_sum(const double *terms, int nmir, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double F[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < nmir; m++)
    for (int dd = 0; dd < 3; dd++) F[dd] = F[dd] + terms[3 * m + dd];
  for (int dd = 0; dd < 3; dd++) out[dd] = F[dd];
}

// ================================================================== host side

namespace {

struct Level {
  bool defined = false;
  Dims d{};
  double cellvolume = 0.0;
  int bctype[6] = {0, 0, 0, 0, 0, 0};
  int ratio[3] = {2, 2, 2};
  double *S[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  double *ws = nullptr, *wr = nullptr, *qf = nullptr, *dw = nullptr, *fw = nullptr;
  double *radsum = nullptr, *vdtim = nullptr;
  uint8_t *flags = nullptr, *marker = nullptr;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr;
  double *piv = nullptr;  // e1x,e2x,e1y,e2y,e1z,e2z each (nmax+2)
  double *fax = nullptr;  // -bb/e1x
  int npiv = 0;
  int nmir = 0, nupd = 0, zero_at = -1;
  DevRec *recs = nullptr;
  double *wsol = nullptr, *psol = nullptr;
  std::vector<mge_cutrec> hrecs;
  int ntouch = 0, nops = 0;
  long *tcell = nullptr;
  int *beg = nullptr;
  int2 *ops = nullptr;
  double *opval = nullptr;
  double *partial = nullptr;
  long npartial = 0;
  unsigned long long *dtim_bits = nullptr;
  double *dtim = nullptr;
  double *rtrms = nullptr;
  double *force = nullptr;
  int mxcyc = 1;
  std::vector<void *> allocs;
};

struct Timer {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};

struct TStat {
  double ms = 0.0;
  long long count = 0;
  double bytes = 0.0;
};

}  // namespace

struct mge_case {
  mge_params prm{};
  int nlevel = 0;
  std::vector<Level> lv;
  hipStream_t stream = nullptr;
  int *err = nullptr;
  bool timing = false;
  std::vector<Timer> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, TStat> tstat;
  double next_bytes = 0.0;  // algorithmic bytes of the next timed launch
  double smoo_bb = 0.0;
};

namespace {

template <class T>
int dalloc(Level &L, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, n * sizeof(T) + 64) != hipSuccess) return set_err(MGE_ERR_NOMEM, "hipMalloc failed");
  hipMemset(q, 0, n * sizeof(T) + 64);
  L.allocs.push_back(q);
  *p = (T *)q;
  return MGE_OK;
}

void level_free(Level &L) {
  for (void *p : L.allocs) hipFree(p);
  L.allocs.clear();
  L = Level();
}

hipEvent_t ev_get(mge_case *h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

void tic(mge_case *h, const char *name, Timer *t) {
  t->bytes = h->next_bytes;
  h->next_bytes = 0.0;
  if (!h->timing) return;
  t->name = name;
  t->a = ev_get(h);
  t->b = ev_get(h);
  hipEventRecord(t->a, h->stream);
}
void toc(mge_case *h, Timer *t) {
  if (!h->timing) return;
  hipEventRecord(t->b, h->stream);
  h->pending.push_back(*t);
}
void resolve_timers(mge_case *h) {
  for (auto &t : h->pending) {
    hipEventSynchronize(t.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, t.a, t.b);
    auto &s = h->tstat[t.name];
    s.ms += ms;
    s.count += 1;
    s.bytes += t.bytes;
    h->pool.push_back(t.a);
    h->pool.push_back(t.b);
  }
  h->pending.clear();
}

#define LAUNCH(h, name, kern, grid, block, ...)                    \
  do {                                                             \
    Timer t_;                                                      \
    tic(h, name, &t_);                                             \
    hipLaunchKernelGGL(kern, grid, block, 0, (h)->stream, __VA_ARGS__); \
    toc(h, &t_);                                                   \
  } while (0)

Geo geo(const Level &L) { return Geo{L.dx, L.dy, L.dz}; }
StateC stc(const Level &L, int b) { return StateC{L.S[b], L.S[b] + 5 * L.d.E}; }
State stm(const Level &L, int b) { return State{L.S[b], L.S[b] + 5 * L.d.E}; }

dim3 cell_grid(const Dims &d) { return dim3((d.ni + 63) / 64, (d.nj + 3) / 4, d.nk); }
dim3 cell_block() { return dim3(64, 4, 1); }

Level *get_level(mge_case *h, int lvl) {
  if (!h || lvl < 1 || lvl > h->nlevel) return nullptr;
  Level *L = &h->lv[lvl];
  return L->defined ? L : nullptr;
}

int check_err(mge_case *h) {
  int e = 0;
  if (hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return set_err(MGE_ERR_HIP, "error-flag copy failed");
  if (hipStreamSynchronize(h->stream) != hipSuccess) return set_err(MGE_ERR_HIP, "stream sync failed");
  if (h->timing) resolve_timers(h);
  if (e < 0) {
    const char *m = "device error";
    if (e == MGE_ERR_NEG_C2) m = "timestep: negative or NaN speed of sound squared";
    else if (e == MGE_ERR_WALL_RHO) m = "efluxb: zero density at a wall face";
    else if (e == MGE_ERR_WSOL_NAN) m = "bc_cutcell: mirror state is NaN";
    else if (e == MGE_ERR_RES_NAN) m = "solver: residual is NaN";
    int z = 0;
    hipMemcpy(h->err, &z, sizeof(int), hipMemcpyHostToDevice);
    return set_err(e, m);
  }
  return MGE_OK;
}

// ---- boundary conditions ------------------------------------------------

int do_bcond(mge_case *h, Level &L, int buf) {
  const mge_params &P = h->prm;
  int warn = MGE_OK;
  for (int f = 1; f <= 6; f++) {
    int t = L.bctype[f - 1];
    if (t == 0) continue;
    if (t != -20 && t != -50 && t != -60 && t != -80) {
      warn = MGE_WARN_BCTYPE;
      continue;
    }
    if (t == -60 && f <= 3 && L.bctype[f + 2] != -60)
      return set_err(MGE_ERR_PERIODIC, "Periodic boundary must appear in pairs");
    BcArgs B;
    B.d = L.d;
    B.s = stm(L, buf);
    B.face = f;
    B.kind = t;
    B.gamma = P.gamma; B.rho0 = P.rho0; B.p0 = P.p0; B.c0 = P.c0;
    B.u0 = P.u0; B.v0 = P.v0; B.w0 = P.w0;
    int dir = (f - 1) % 3;
    int amax = dir == 0 ? L.d.jb : L.d.ib;
    int cmax = dir == 2 ? L.d.jb : L.d.kb;
    dim3 g((amax + 1 + 255) / 256, cmax + 1);
    LAUNCH(h, "k_bc_face", k_bc_face, g, dim3(256), B);
  }
  LAUNCH(h, "k_bc_corner", k_bc_corner, dim3(1), dim3(1024), L.d, stm(L, buf), P.gamma);
  return warn;
}

int do_bc_cutcell(mge_case *h, Level &L, int buf) {
  if (L.nmir == 0) return MGE_OK;
  CutArgs C;
  C.d = L.d;
  C.s = stc(L, buf);
  C.recs = L.recs;
  C.wsol = L.wsol;
  C.psol = L.psol;
  C.nmir = L.nmir;
  C.nupd = L.nupd;
  C.zero_at = L.zero_at;
  C.gamma = h->prm.gamma;
  C.eqntype = h->prm.eqntype;
  C.err = h->err;
  LAUNCH(h, "k_bc_cut", k_bc_cut, dim3((L.nmir + 63) / 64), dim3(64), C);
  return L.zero_at >= 0 ? MGE_WARN_ZERO_CLOUD : MGE_OK;
}

// ---- timestep --------------------------------------------------------------

int do_timestep(mge_case *h, Level &L) {
  const mge_params &P = h->prm;
  unsigned long long init = (unsigned long long)0x4024000000000000ULL;  // 10.0
  HIPCHK(hipMemcpyAsync(L.dtim_bits, &init, sizeof(init), hipMemcpyHostToDevice, h->stream));
  TsArgs T;
  T.d = L.d;
  T.s = stc(L, L.cur);
  T.g = geo(L);
  T.flags = L.flags;
  T.radsum = L.radsum;
  T.vdtim = L.vdtim;
  T.dtim_bits = L.dtim_bits;
  T.gamma = P.gamma;
  T.cfl = P.cfl;
  T.cellvolume = L.cellvolume;
  T.eqntype = P.eqntype;
  T.viscl = P.viscl;
  T.prlmean = P.prlmean;
  T.err = h->err;
  LAUNCH(h, "k_timestep", k_timestep, cell_grid(L.d), cell_block(), T);
  LAUNCH(h, "k_dtim_decode", k_dtim_decode, dim3(1), dim3(1), L.dtim_bits, L.dtim);
  return MGE_OK;
}

// ---- euler -------------------------------------------------------------------

int do_euler(mge_case *h, Level &L, int ncyc, int mgl, int mlvl, int lres, int lfir) {
  const mge_params &P = h->prm;
  const Dims &d = L.d;
  const long E = d.E;
  if (P.smoo > 0.0 && (d.ni < 2 || d.nj < 2 || d.nk < 2))
    return set_err(MGE_ERR_SMOO_DIM, "residual smoothing needs at least 2 cells per direction");
  if (lfir) HIPCHK(hipMemcpyAsync(L.ws, L.S[L.cur], sizeof(double) * 5 * E, hipMemcpyDeviceToDevice, h->stream));
  const int w0 = L.cur;
  int src = w0;
  int warn = MGE_OK;
  for (int m = 1; m <= P.mstage; m++) {
    int dst = 0;
    while (dst == w0 || dst == src) dst++;
    StageArgs A;
    A.fx.d = d;
    A.fx.s = stc(L, src);
    A.fx.g = geo(L);
    A.fx.radsum = L.radsum;
    A.fx.gamma = P.gamma;
    A.fx.vis2 = P.vis2;
    A.fx.vis4 = P.vis4;
    A.fx.cdis = P.cdis[m - 1];
    A.fx.dissip = P.dissip;
    A.wdst = L.S[dst];
    A.pdst = L.S[dst] + 5 * E;
    A.wn = L.S[w0];
    A.dw = L.dw;
    A.fw = L.fw;
    A.qf = L.qf;
    A.wr = L.wr;
    A.marker = L.marker;
    A.flags = L.flags;
    A.vdtim = L.vdtim;
    A.dtim = L.dtim;
    A.cellvolume = L.cellvolume;
    A.blend = 1.0 - P.cdis[m - 1];
    A.fn = P.cstp[m - 1] * P.cfl;
    A.m = m;
    A.has_diss = P.cdis[m - 1] > 0.0;
    A.coarse = mgl < mlvl;
    A.lfir1 = lfir && m == 1;
    A.lres = lres;
    A.smoothed = (m % 2 != 0) && P.smoo > 0.0;
    A.store_fw = (m < P.mstage) && (A.has_diss || m == 1 || A.blend != 1.0);
    A.vt = P.vt;
    {
      // algorithmic bytes per cell: read W^(m-1); write W^(m) and read W^(0)
      // when this kernel updates, else write dw; forcing on coarse levels;
      // one flag byte
      double s8 = 8.0, per = 5 * s8 + 1;
      bool upd = !lres && !A.smoothed;
      per += upd ? 10 * s8 : 5 * s8;
      if (A.coarse) per += A.lfir1 ? 10 * s8 : 5 * s8;
      h->next_bytes = per * (double)d.ni * d.nj * d.nk;
    }
    {
      int kc = d.nk;
      long tiles = (long)((d.ni + TI - 1) / TI) * ((d.nj + TJ - 1) / TJ);
      // enough workgroups to fill 256 CUs: split k into chunks
      while (kc > 8 && tiles * ((d.nk + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
      dim3 g((d.ni + TI - 1) / TI, (d.nj + TJ - 1) / TJ, (d.nk + kc - 1) / kc);
      if (P.dissip == 0) LAUNCH(h, "k_stage", k_stage2, g, dim3(TI, TJ), A, kc);
      else LAUNCH(h, "k_stage", k_stage, cell_grid(d), cell_block(), A);  // matrix dissipation
    }
    if (L.ntouch > 0) {
      RecArgs R;
      R.st = A;
      R.st.store_fw = 1;
      R.recs = L.recs;
      R.wsol = L.wsol;
      R.psol = L.psol;
      R.nmir = L.nmir;
      R.tcell = L.tcell;
      R.beg = L.beg;
      R.ops = L.ops;
      R.opval = L.opval;
      R.ntouch = L.ntouch;
      R.nops = L.nops;
      R.err = h->err;
      LAUNCH(h, "k_rec_ops", k_rec_ops, dim3((L.nops + 63) / 64), dim3(64), R);
      LAUNCH(h, "k_rec_apply", k_rec_apply, dim3((L.ntouch + 63) / 64), dim3(64), R);
    }
    if (lres) return MGE_OK;
    if (A.smoothed) {
      SmoothArgs S;
      S.d = d;
      S.dw = L.dw;
      S.flags = L.flags;
      S.bb = h->smoo_bb;
      S.wn = L.S[w0];
      S.wdst = L.S[dst];
      S.pdst = L.S[dst] + 5 * E;
      S.gamma = P.gamma;
      S.e1 = L.piv;
      S.e2 = L.piv + L.npiv;
      S.fa = L.fax;
      long nwave = (long)d.nj * d.nk * 5;
      dim3 gx((unsigned)((nwave * 64 + 255) / 256));
      if (d.ni <= 64) LAUNCH(h, "k_smooth_x", k_smooth_x_scan<1>, gx, dim3(256), S);
      else if (d.ni <= 128) LAUNCH(h, "k_smooth_x", k_smooth_x_scan<2>, gx, dim3(256), S);
      else if (d.ni <= 256) LAUNCH(h, "k_smooth_x", k_smooth_x_scan<4>, gx, dim3(256), S);
      else if (d.ni <= 512) LAUNCH(h, "k_smooth_x", k_smooth_x_scan<8>, gx, dim3(256), S);
      else LAUNCH(h, "k_smooth_x", k_smooth_x_seq, dim3((unsigned)((nwave + 255) / 256)), dim3(256), S);
      S.e1 = L.piv + 2 * L.npiv;
      S.e2 = L.piv + 3 * L.npiv;
      LAUNCH(h, "k_smooth_y", k_smooth_y, dim3((d.ni + 63) / 64, d.nk, 5), dim3(64), S);
      S.e1 = L.piv + 4 * L.npiv;
      S.e2 = L.piv + 5 * L.npiv;
      LAUNCH(h, "k_smooth_z", k_smooth_z, dim3((d.ni + 63) / 64, d.nj, 5), dim3(64), S);
      LAUNCH(h, "k_update_ghost", k_update_ghost, dim3((unsigned)((E + 255) / 256)), dim3(256), d,
             (const double *)L.S[w0], (const double *)L.dw, stc(L, src), stm(L, dst), P.gamma);
    } else {
      LAUNCH(h, "k_ghost_copy", k_ghost_copy, dim3((unsigned)((E + 255) / 256)), dim3(256), d, stc(L, src),
             stm(L, dst));
    }
    int rc = do_bc_cutcell(h, L, dst);
    if (rc > 0) warn = rc;
    rc = do_bcond(h, L, dst);
    if (rc < 0) return rc;
    if (rc > 0) warn = rc;
    src = dst;
  }
  L.cur = src;
  RtArgs R;
  R.d = d;
  R.dw = L.dw;
  R.vdtim = L.vdtim;
  R.dtim = L.dtim;
  R.flags = L.flags;
  R.cfl = P.cfl;
  R.vt = P.vt;
  R.partial = L.partial;
  dim3 g = cell_grid(d);
  LAUNCH(h, "k_rtrms", k_rtrms, g, cell_block(), R);
  long nb = (long)g.x * g.y * g.z;
  int c = ncyc >= 1 && ncyc <= L.mxcyc ? ncyc : 0;
  LAUNCH(h, "k_rtrms_final", k_rtrms_final, dim3(1), dim3(256), (const double *)L.partial, nb,
         (double)((long)d.ni * d.nj * d.nk), L.rtrms + c, h->err);
  return warn;
}

// ---- transfers ------------------------------------------------------------------

// fine level = ratio x coarse level in every direction (aligned levels)
bool levels_nest(const Level &c, const Level &f) {
  int nc[3] = {c.d.ni, c.d.nj, c.d.nk}, nf[3] = {f.d.ni, f.d.nj, f.d.nk};
  for (int a = 0; a < 3; a++)
    if (nf[a] != f.ratio[a] * nc[a]) return false;
  return true;
}

int do_movco(mge_case *h, int mgl) {
  Level &c = h->lv[mgl], &f = h->lv[mgl + 1];
  if (!levels_nest(c, f)) return set_err(MGE_ERR_ARG, "levels do not nest with the given ratio");
  XferArgs X;
  X.dc = c.d;
  X.df = f.d;
  X.rx = f.ratio[0]; X.ry = f.ratio[1]; X.rz = f.ratio[2];
  X.wf = f.S[f.cur];
  X.dwf = f.dw;
  X.wc = c.S[c.cur];
  X.wsc = c.ws;
  X.wr = c.wr;
  X.marker = c.marker;
  X.volc = c.cellvolume;
  X.volf = f.cellvolume;
  X.fmg = 0;
  X.flagsf = f.flags;
  X.wfo = nullptr;
  LAUNCH(h, "k_restrict", k_restrict, cell_grid(c.d), cell_block(), X);
  LAUNCH(h, "k_eos_all", k_eos_all, dim3((unsigned)((c.d.E + 255) / 256)), dim3(256), c.d, stm(c, c.cur),
         h->prm.gamma);
  return do_bcond(h, c, c.cur);
}

int do_movfin(mge_case *h, int mgl, int mlvl) {
  Level &f = h->lv[mgl], &c = h->lv[mgl - 1];
  if (!levels_nest(c, f)) return set_err(MGE_ERR_ARG, "levels do not nest with the given ratio");
  XferArgs X;
  X.dc = c.d;
  X.df = f.d;
  X.rx = f.ratio[0]; X.ry = f.ratio[1]; X.rz = f.ratio[2];
  X.wf = nullptr;
  X.dwf = nullptr;
  X.wc = c.S[c.cur];
  X.wsc = c.ws;
  X.wr = nullptr;
  X.marker = nullptr;
  X.volc = c.cellvolume;
  X.volf = f.cellvolume;
  X.fmg = (mgl - 1 == mlvl);
  X.flagsf = f.flags;
  X.wfo = f.S[f.cur];
  LAUNCH(h, "k_prolong", k_prolong, cell_grid(f.d), cell_block(), X);
  LAUNCH(h, "k_eos_all", k_eos_all, dim3((unsigned)((f.d.E + 255) / 256)), dim3(256), f.d, stm(f, f.cur),
         h->prm.gamma);
  return do_bcond(h, f, f.cur);
}

int do_init_flow(mge_case *h) {
  const mge_params &P = h->prm;
  int warn = MGE_OK;
  for (int l = 1; l <= h->nlevel; l++) {
    Level &L = h->lv[l];
    if (!L.defined) return set_err(MGE_ERR_STATE, "level not defined");
    LAUNCH(h, "k_fill", k_fill_freestream, dim3((unsigned)((L.d.E + 255) / 256)), dim3(256), L.d,
           stm(L, L.cur), P.rho0, P.rho0 * P.u0, P.rho0 * P.v0, P.rho0 * P.w0, P.rho0 * P.h0 - P.p0);
  }
  for (int l = 1; l <= h->nlevel; l++) {
    Level &L = h->lv[l];
    LAUNCH(h, "k_eos_all", k_eos_all, dim3((unsigned)((L.d.E + 255) / 256)), dim3(256), L.d, stm(L, L.cur),
           P.gamma);
    int rc = do_bc_cutcell(h, L, L.cur);
    if (rc > 0) warn = rc;
    rc = do_bcond(h, L, L.cur);
    if (rc < 0) return rc;
    if (rc > 0) warn = rc;
  }
  return warn;
}

#define CHKRC(x)            \
  do {                      \
    int r_ = (x);           \
    if (r_ < 0) return r_;  \
  } while (0)

// one multigrid cycle with finest active level mlvl: the body of lopmg in
// solver.f90:127-234.  The residual of the mlvl relax is left in
// L.rtrms[ncyc] (read by the caller after the cycle: the finest level is
// relaxed once per cycle, so the value is final when the cycle ends).
int do_cycle(mge_case *h, int mlvl, int ncyc, long long *cs) {
  const int nlevel = h->nlevel;
  const mge_params &P = h->prm;
  std::vector<int> icyc(nlevel + 1, 0);
  int mgl = mlvl;
  for (;;) {
    for (;;) {
      icyc[mgl] += 1;
      int lfir = (mgl == mlvl) ? 0 : 1;
      Level &B = h->lv[mgl];
      CHKRC(do_timestep(h, B));
      CHKRC(do_euler(h, B, ncyc, mgl, mlvl, 0, lfir));
      *cs += (long long)P.mstage * B.d.ni * B.d.nj * B.d.nk;
      if (mlvl == 1) return MGE_OK;
      CHKRC(do_euler(h, B, ncyc, mgl, mlvl, 1, 0));
      *cs += (long long)B.d.ni * B.d.nj * B.d.nk;
      mgl -= 1;
      CHKRC(do_movco(h, mgl));
      if (mgl == 1) break;
    }
    int lfir = 1;
    for (;;) {
      Level &B = h->lv[mgl];
      CHKRC(do_timestep(h, B));
      CHKRC(do_euler(h, B, ncyc, mgl, mlvl, 0, lfir));
      *cs += (long long)P.mstage * B.d.ni * B.d.nj * B.d.nk;
      lfir = 0;
      mgl += 1;
      CHKRC(do_movfin(h, mgl, mlvl));
      if (mgl == mlvl) return MGE_OK;
      if (icyc[mgl] != P.mgtype) break;
      icyc[mgl] = 0;
    }
  }
}

int read_rtrms(mge_case *h, int mlvl, int ncyc, double *r) {
  Level &B = h->lv[mlvl];
  int c = ncyc >= 1 && ncyc <= B.mxcyc ? ncyc : 0;
  HIPCHK(hipMemcpyAsync(r, B.rtrms + c, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return check_err(h);
}

// solver.f90:119-261 for steady flow (nperiod = nstep = 1): FMG ramp over
// the levels, cycles with the residual log and the drop test
int do_solve(mge_case *h, const int *mcyc, double *reslog, int cap, int *nlog, long long *cs_out) {
  const int nlevel = h->nlevel;
  const mge_params &P = h->prm;
  int nl = 0;
  double ini_res = 0.0, max_res = 0.0;
  long long cs = 0;
  for (int mlvl = 1; mlvl <= nlevel; mlvl++) {
    for (int ncyc = 1; ncyc <= mcyc[mlvl - 1]; ncyc++) {
      CHKRC(do_cycle(h, mlvl, ncyc, &cs));
      double r = 0.0;
      CHKRC(read_rtrms(h, mlvl, ncyc, &r));
      max_res = 1000.0;
      if (ncyc == 1) ini_res = r;
      if (r < max_res) max_res = r;
      if (r != r) return set_err(MGE_ERR_RES_NAN, "solver: residual is NaN");
      if (reslog && nl < cap) {
        reslog[3 * nl] = mlvl;
        reslog[3 * nl + 1] = ncyc;
        reslog[3 * nl + 2] = max_res;
      }
      nl++;
      if (mlvl > 1 && ini_res / max_res > P.drop) break;
    }
    if (mlvl < nlevel) CHKRC(do_movfin(h, mlvl + 1, mlvl));
  }
  CHKRC(check_err(h));
  if (nlog) *nlog = nl;
  if (cs_out) *cs_out = cs;
  return MGE_OK;
}

// host-side record bookkeeping: touched cells and their ordered op lists
int build_records(mge_case *h, Level &L, int nmir, const mge_cutrec *recs) {
  const Dims &d = L.d;
  L.nmir = nmir;
  L.hrecs.assign(recs, recs + nmir);
  L.zero_at = -1;
  L.nupd = nmir;
  for (int m = 0; m < nmir; m++)
    if (recs[m].npoint == 0) {
      L.zero_at = m;
      L.nupd = m;
      break;
    }
  if (nmir == 0) return MGE_OK;
  std::vector<DevRec> dr(nmir);
  for (int m = 0; m < nmir; m++) {
    const mge_cutrec &r = recs[m];
    DevRec &o = dr[m];
    std::memset(&o, 0, sizeof(o));
    if (r.faceid < 1 || r.faceid > 6 || r.zz != (r.faceid - 1) % 3 + 1 || r.npoint < 0 || r.npoint > 8)
      return set_err(MGE_ERR_ARG, "bad wall record");
    o.faceid = r.faceid;
    o.zz = r.zz;
    for (int a = 0; a < 3; a++) {
      o.c[a] = r.ccid[a];
      o.s[a] = r.scid[a];
      o.nvec[a] = r.nvec[a];
    }
    int cmax[3] = {d.il, d.jl, d.kl};
    for (int a = 0; a < 3; a++) {
      // the cut cell and its stencil along zz must stay inside the arrays
      if (r.ccid[a] < 2 || r.ccid[a] > cmax[a]) return set_err(MGE_ERR_ARG, "record cell outside the block");
      if (r.scid[a] < 0 || r.scid[a] > cmax[a] + 2) return set_err(MGE_ERR_ARG, "record solid cell outside");
    }
    int s0 = r.ccid[r.zz - 1];
    int nl = cmax[r.zz - 1];
    if (s0 < 3 || s0 > nl - 1) return set_err(MGE_ERR_ARG, "wall record too close to the block edge");
    o.npoint = r.npoint;
    for (int c = 0; c < r.npoint; c++) {
      for (int a = 0; a < 3; a++)
        if (r.ijk[c][a] < 0 || r.ijk[c][a] > cmax[a] + 2) return set_err(MGE_ERR_ARG, "cloud point outside");
      o.cloud[c] = ix(d, r.ijk[c][0], r.ijk[c][1], r.ijk[c][2]);
      o.coef[c] = r.coef[c];
    }
  }
  // ordered op streams per touched cell
  std::map<long, std::pair<std::vector<int2>, std::vector<int2>>> per;
  for (int m = 0; m < nmir; m++) {
    const mge_cutrec &r = recs[m];
    long st = r.zz == 1 ? 1L : (r.zz == 2 ? d.sy : d.sz);
    long qc = ix(d, r.ccid[0], r.ccid[1], r.ccid[2]);
    bool hi = r.faceid >= 4;
    long qn = hi ? qc - st : qc + st;
    auto &pc = per[qc];
    auto &pn = per[qn];
    if (hi) {
      pc.first.push_back(make_int2(m, OP_HI_C_UNDO));
      pn.first.push_back(make_int2(m, OP_HI_N_UNDO));
      pc.first.push_back(make_int2(m, OP_HI_C_ADD));
      pn.first.push_back(make_int2(m, OP_HI_N_ADD));
    } else {
      pc.first.push_back(make_int2(m, OP_LO_C_UNDO));
      pn.first.push_back(make_int2(m, OP_LO_N_UNDO));
      pc.first.push_back(make_int2(m, OP_LO_C_ADD));
      pn.first.push_back(make_int2(m, OP_LO_N_ADD));
    }
    pc.second.push_back(make_int2(m, OP_E_UNDO));
    pc.second.push_back(make_int2(m, OP_E_ADD));
  }
  std::vector<long> tcell;
  std::vector<int> begv(1, 0);
  std::vector<int2> opv;
  for (auto &kv : per) {
    tcell.push_back(kv.first);
    for (auto &o : kv.second.first) opv.push_back(o);
    for (auto &o : kv.second.second) opv.push_back(o);
    begv.push_back((int)opv.size());
  }
  L.ntouch = (int)tcell.size();
  L.nops = (int)opv.size();
  CHKRC(dalloc(L, &L.recs, nmir));
  CHKRC(dalloc(L, &L.wsol, (size_t)5 * nmir));
  CHKRC(dalloc(L, &L.psol, nmir));
  CHKRC(dalloc(L, &L.tcell, tcell.size()));
  CHKRC(dalloc(L, &L.beg, begv.size()));
  CHKRC(dalloc(L, &L.ops, opv.size() + 1));
  CHKRC(dalloc(L, &L.opval, (opv.size() + 1) * 10));
  HIPCHK(hipMemcpy(L.recs, dr.data(), sizeof(DevRec) * nmir, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.tcell, tcell.data(), sizeof(long) * tcell.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.beg, begv.data(), sizeof(int) * begv.size(), hipMemcpyHostToDevice));
  if (!opv.empty()) HIPCHK(hipMemcpy(L.ops, opv.data(), sizeof(int2) * opv.size(), hipMemcpyHostToDevice));
  std::vector<uint8_t> fl(d.E);
  HIPCHK(hipMemcpy(fl.data(), L.flags, d.E, hipMemcpyDeviceToHost));
  for (long q : tcell) fl[q] |= F_TOUCHED;
  HIPCHK(hipMemcpy(L.flags, fl.data(), d.E, hipMemcpyHostToDevice));
  std::vector<double> w5(5 * (size_t)nmir), p1(nmir);
  for (int m = 0; m < nmir; m++) {
    for (int n = 0; n < 5; n++) w5[(size_t)n * nmir + m] = recs[m].wsol[n];
    p1[m] = recs[m].psol;
  }
  HIPCHK(hipMemcpy(L.wsol, w5.data(), sizeof(double) * w5.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.psol, p1.data(), sizeof(double) * p1.size(), hipMemcpyHostToDevice));
  return MGE_OK;
}

}  // namespace

// ================================================================== C ABI

extern "C" {

const char *mge_last_error(void) { return g_err.c_str(); }

int mge_case_create(const mge_params *prm, int nlevel, mge_case **out) {
  if (!prm || !out || nlevel < 1 || nlevel > 16) return set_err(MGE_ERR_ARG, "bad case arguments");
  if (prm->mstage < 1 || prm->mstage > 8) return set_err(MGE_ERR_ARG, "mstage must be 1..8");
  if (prm->distype != 0) return set_err(MGE_ERR_DISTYPE, "only the central (JST) scheme is available");
  if (prm->dissip != 0 && prm->dissip != 1) return set_err(MGE_ERR_DISTYPE, "dissip must be scalar or matrix");
  if (prm->precision != 64 && prm->precision != 0) return set_err(MGE_ERR_ARG, "precision must be 64");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(MGE_ERR_HIP, "no HIP device");
  mge_case *h = new mge_case();
  h->prm = *prm;
  h->nlevel = nlevel;
  h->lv.resize(nlevel + 1);
  h->smoo_bb = -prm->smoo;
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return set_err(MGE_ERR_HIP, "hipStreamCreate failed");
  }
  if (hipMalloc(&h->err, sizeof(int)) != hipSuccess) {
    hipStreamDestroy(h->stream);
    delete h;
    return set_err(MGE_ERR_NOMEM, "hipMalloc failed");
  }
  hipMemset(h->err, 0, sizeof(int));
  *out = h;
  return MGE_OK;
}

int mge_case_destroy(mge_case *h) {
  if (!h) return MGE_OK;
  hipStreamSynchronize(h->stream);
  for (auto &L : h->lv) level_free(L);
  for (auto &t : h->pending) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  for (auto e : h->pool) hipEventDestroy(e);
  hipFree(h->err);
  hipStreamDestroy(h->stream);
  delete h;
  return MGE_OK;
}

int mge_level_define(mge_case *h, int lvl, int ni, int nj, int nk, const double *xv, const double *yv,
                     const double *zv, double cellvolume, const int bctype[6], const int ratio[3],
                     const uint8_t *blank, int nmir, const mge_cutrec *recs, int mxcyc) {
  if (!h || lvl < 1 || lvl > h->nlevel) return set_err(MGE_ERR_ARG, "bad level");
  if (ni < 1 || nj < 1 || nk < 1 || !xv || !yv || !zv || !bctype || !blank || nmir < 0 ||
      (nmir > 0 && !recs))
    return set_err(MGE_ERR_ARG, "bad level arguments");
  Level &L = h->lv[lvl];
  level_free(L);
  Dims &d = L.d;
  d.ni = ni; d.nj = nj; d.nk = nk;
  d.il = ni + 1; d.jl = nj + 1; d.kl = nk + 1;
  d.ie = ni + 2; d.je = nj + 2; d.ke = nk + 2;
  d.ib = ni + 3; d.jb = nj + 3; d.kb = nk + 3;
  d.sy = d.ib + 1;
  d.sz = (long)(d.ib + 1) * (d.jb + 1);
  d.E = d.sz * (d.kb + 1);
  const long E = d.E;
  L.cellvolume = cellvolume;
  for (int f = 0; f < 6; f++) L.bctype[f] = bctype[f];
  for (int a = 0; a < 3; a++) L.ratio[a] = ratio ? ratio[a] : 2;
  for (int a = 0; a < 3; a++)
    if (L.ratio[a] != 1 && L.ratio[a] != 2) return set_err(MGE_ERR_ARG, "ratio must be 1 or 2");
  L.mxcyc = mxcyc > 0 ? mxcyc : 1;
  for (int b = 0; b < 3; b++) CHKRC(dalloc(L, &L.S[b], (size_t)6 * E));
  CHKRC(dalloc(L, &L.ws, (size_t)5 * E));
  CHKRC(dalloc(L, &L.wr, (size_t)5 * E));
  CHKRC(dalloc(L, &L.qf, (size_t)5 * E));
  CHKRC(dalloc(L, &L.dw, (size_t)5 * E));
  CHKRC(dalloc(L, &L.fw, (size_t)5 * E));
  CHKRC(dalloc(L, &L.radsum, E));
  CHKRC(dalloc(L, &L.vdtim, E));
  CHKRC(dalloc(L, &L.flags, E));
  CHKRC(dalloc(L, &L.marker, E));
  int nmax = std::max(d.ib, std::max(d.jb, d.kb)) + 2;
  CHKRC(dalloc(L, &L.dx, nmax));
  CHKRC(dalloc(L, &L.dy, nmax));
  CHKRC(dalloc(L, &L.dz, nmax));
  L.npiv = nmax;
  CHKRC(dalloc(L, &L.piv, (size_t)6 * nmax));
  dim3 g = cell_grid(d);
  L.npartial = (long)g.x * g.y * g.z;
  CHKRC(dalloc(L, &L.partial, L.npartial));
  CHKRC(dalloc(L, &L.dtim_bits, 1));
  CHKRC(dalloc(L, &L.dtim, 1));
  CHKRC(dalloc(L, &L.rtrms, L.mxcyc + 1));
  CHKRC(dalloc(L, &L.force, 3));
  // vertex differences (ghost entries of index 0 stay unused)
  std::vector<double> hx(nmax, 0.0), hy(nmax, 0.0), hz(nmax, 0.0);
  for (int i = 1; i < d.ib; i++) hx[i] = xv[i + 1] - xv[i];
  for (int j = 1; j < d.jb; j++) hy[j] = yv[j + 1] - yv[j];
  for (int k = 1; k < d.kb; k++) hz[k] = zv[k + 1] - zv[k];
  HIPCHK(hipMemcpy(L.dx, hx.data(), sizeof(double) * nmax, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.dy, hy.data(), sizeof(double) * nmax, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(L.dz, hz.data(), sizeof(double) * nmax, hipMemcpyHostToDevice));
  // smoothing pivots (psmoo.f90:134-166), computed once on the host
  std::vector<double> piv((size_t)6 * nmax, 0.0);
  if (h->prm.smoo > 0.0) {
    double aa = 1.0 + 2.0 * h->prm.smoo, bb = -h->prm.smoo;
    int nn[3] = {ni, nj, nk};
    for (int a = 0; a < 3; a++) {
      double *e1 = &piv[(size_t)(2 * a) * nmax], *e2 = &piv[(size_t)(2 * a + 1) * nmax];
      int lm = nn[a], l = nn[a] + 1;
      if (lm < 2) continue;
      e1[2] = aa;
      e2[2] = bb / aa;
      for (int s = 3; s <= lm; s++) {
        e1[s] = aa - bb * e2[s - 1];
        e2[s] = bb / e1[s];
      }
      e1[l] = aa - bb * e2[lm];
    }
  }
  HIPCHK(hipMemcpy(L.piv, piv.data(), sizeof(double) * piv.size(), hipMemcpyHostToDevice));
  std::vector<double> fa(nmax, 0.0);
  if (h->prm.smoo > 0.0)
    for (int i = 2; i <= d.il && i < nmax; i++) fa[i] = h->prm.smoo / piv[i];  // -bb/e1, bb = -smoo
  CHKRC(dalloc(L, &L.fax, nmax));
  HIPCHK(hipMemcpy(L.fax, fa.data(), sizeof(double) * nmax, hipMemcpyHostToDevice));
  // flags: fluid bit with the ghost copies of init_metric.f90:304-315
  std::vector<uint8_t> bl(E, 1);
  auto X = [&](int i, int j, int k) { return ix(d, i, j, k); };
  for (int k = 2; k <= d.kl; k++)
    for (int j = 2; j <= d.jl; j++)
      for (int i = 2; i <= d.il; i++)
        bl[X(i, j, k)] = blank[(long)(i - 2) + (long)ni * ((long)(j - 2) + (long)nj * (k - 2))] ? 1 : 0;
  for (int k = 0; k <= d.kb; k++)
    for (int j = 0; j <= d.jb; j++) {
      bl[X(0, j, k)] = bl[X(2, j, k)];
      bl[X(1, j, k)] = bl[X(2, j, k)];
      bl[X(d.ie, j, k)] = bl[X(d.il, j, k)];
      bl[X(d.ib, j, k)] = bl[X(d.il, j, k)];
    }
  for (int k = 0; k <= d.kb; k++)
    for (int i = 0; i <= d.ib; i++) {
      bl[X(i, 0, k)] = bl[X(i, 2, k)];
      bl[X(i, 1, k)] = bl[X(i, 2, k)];
      bl[X(i, d.je, k)] = bl[X(i, d.jl, k)];
      bl[X(i, d.jb, k)] = bl[X(i, d.jl, k)];
    }
  for (int j = 0; j <= d.jb; j++)
    for (int i = 0; i <= d.ib; i++) {
      bl[X(i, j, 0)] = bl[X(i, j, 2)];
      bl[X(i, j, 1)] = bl[X(i, j, 2)];
      bl[X(i, j, d.ke)] = bl[X(i, j, d.kl)];
      bl[X(i, j, d.kb)] = bl[X(i, j, d.kl)];
    }
  HIPCHK(hipMemcpy(L.flags, bl.data(), E, hipMemcpyHostToDevice));
  L.cur = 0;
  L.defined = true;
  int rc = build_records(h, L, nmir, recs);
  if (rc < 0) {
    level_free(L);
    return rc;
  }
  return MGE_OK;
}

int mge_init_flow(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  int rc = do_init_flow(h);
  if (rc < 0) return rc;
  int e = check_err(h);
  return e < 0 ? e : rc;
}

int mge_timestep(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return do_timestep(h, *L);
}

int mge_euler(mge_case *h, int lvl, int blk, int itime, int ncyc, int mgl, int mlvl, int lres, int lfir,
              double *rtrms) {
  (void)itime;
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  int rc = do_euler(h, *L, ncyc, mgl, mlvl, lres, lfir);
  if (rc < 0) return rc;
  if (rtrms && !lres) {
    int c = ncyc >= 1 && ncyc <= L->mxcyc ? ncyc : 0;
    HIPCHK(hipMemcpyAsync(rtrms, L->rtrms + c, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    int e = check_err(h);
    if (e < 0) return e;
  }
  return rc;
}

int mge_bcond(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return do_bcond(h, *L, L->cur);
}

int mge_bc_cutcell(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return do_bc_cutcell(h, *L, L->cur);
}

int mge_bc_overlap(mge_case *h, int lvl) {
  // one block per level: no same-level overlaps (bc_overlap.f90:296-354 loops over none)
  if (!get_level(h, lvl)) return set_err(MGE_ERR_ARG, "bad level");
  return MGE_OK;
}

int mge_movco(mge_case *h, int mgl) {
  if (!get_level(h, mgl) || !get_level(h, mgl + 1)) return set_err(MGE_ERR_ARG, "bad level");
  return do_movco(h, mgl);
}

int mge_movfin(mge_case *h, int mgl, int mlvl) {
  if (!get_level(h, mgl) || !get_level(h, mgl - 1)) return set_err(MGE_ERR_ARG, "bad level");
  return do_movfin(h, mgl, mlvl);
}

int mge_solve(mge_case *h, const int *mcyc, double *reslog, int cap, int *nlog, long long *cell_stages) {
  if (!h || !mcyc) return set_err(MGE_ERR_ARG, "bad solve arguments");
  for (int l = 1; l <= h->nlevel; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  return do_solve(h, mcyc, reslog, cap, nlog, cell_stages);
}

int mge_cycles(mge_case *h, int mlvl, int ncyc0, int count, double *res, long long *cell_stages) {
  if (!h || mlvl < 1 || mlvl > h->nlevel || count < 0) return set_err(MGE_ERR_ARG, "bad cycle arguments");
  for (int l = 1; l <= mlvl; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  long long cs = 0;
  for (int c = 0; c < count; c++) CHKRC(do_cycle(h, mlvl, ncyc0 + c, &cs));
  if (res && count > 0) CHKRC(read_rtrms(h, mlvl, ncyc0 + count - 1, res));
  else CHKRC(check_err(h));
  if (cell_stages) *cell_stages = cs;
  return MGE_OK;
}

static double *field_ptr(Level &L, const char *name, long *n) {
  const long E = L.d.E;
  *n = 5 * E;
  if (!strcmp(name, "w")) return L.S[L.cur];
  if (!strcmp(name, "dw")) return L.dw;
  if (!strcmp(name, "fw")) return L.fw;
  if (!strcmp(name, "ws")) return L.ws;
  if (!strcmp(name, "wr")) return L.wr;
  if (!strcmp(name, "qf")) return L.qf;
  *n = E;
  if (!strcmp(name, "p")) return L.S[L.cur] + 5 * E;
  if (!strcmp(name, "vdtim")) return L.vdtim;
  if (!strcmp(name, "radsum")) return L.radsum;
  return nullptr;
}

int mge_download(mge_case *h, int lvl, int blk, const char *field, double *out) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !field || !out) return set_err(MGE_ERR_ARG, "bad download arguments");
  long n;
  double *p = field_ptr(*L, field, &n);
  if (!p) return set_err(MGE_ERR_ARG, "unknown field");
  HIPCHK(hipMemcpyAsync(out, p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  return check_err(h);
}

int mge_upload(mge_case *h, int lvl, int blk, const char *field, const double *in) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !field || !in) return set_err(MGE_ERR_ARG, "bad upload arguments");
  long n;
  double *p = field_ptr(*L, field, &n);
  if (!p) return set_err(MGE_ERR_ARG, "unknown field");
  HIPCHK(hipMemcpyAsync(p, in, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return MGE_OK;
}

int mge_get_cuts(mge_case *h, int lvl, int blk, mge_cutrec *out) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !out) return set_err(MGE_ERR_ARG, "bad arguments");
  int n = L->nmir;
  if (n == 0) return 0;
  std::vector<double> w5((size_t)5 * n), p1(n);
  HIPCHK(hipMemcpyAsync(w5.data(), L->wsol, sizeof(double) * w5.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(p1.data(), L->psol, sizeof(double) * p1.size(), hipMemcpyDeviceToHost, h->stream));
  int e = check_err(h);
  if (e < 0) return e;
  for (int m = 0; m < n; m++) {
    out[m] = L->hrecs[m];
    for (int c = 0; c < 5; c++) out[m].wsol[c] = w5[(size_t)c * n + m];
    out[m].psol = p1[m];
  }
  return n;
}

double mge_get_dtim(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return NAN;
  double v = NAN;
  hipMemcpyAsync(&v, L->dtim, sizeof(double), hipMemcpyDeviceToHost, h->stream);
  hipStreamSynchronize(h->stream);
  return v;
}

int mge_wall_force(mge_case *h, int lvl, double out[3]) {
  Level *L = get_level(h, lvl);
  if (!L || !out) return set_err(MGE_ERR_ARG, "bad arguments");
  if (L->nmir == 0) {
    out[0] = out[1] = out[2] = 0.0;
    return MGE_OK;
  }
  LAUNCH(h, "k_wall_terms", k_wall_terms, dim3((L->nmir + 63) / 64), dim3(64), L->d, stc(*L, L->cur), geo(*L),
         L->recs, L->wsol, L->psol, L->nmir, L->opval);
  LAUNCH(h, "k_wall_sum", k_wall_sum, dim3(1), dim3(64), (const double *)L->opval, L->nmir, L->force);
  HIPCHK(hipMemcpyAsync(out, L->force, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
  return check_err(h);
}

int mge_synchronize(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  return check_err(h);
}

int mge_kernel_timing(mge_case *h, int enable, const char *name, double *ms, long long *count,
                      double *bytes) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  if (h->timing) {
    hipStreamSynchronize(h->stream);
    resolve_timers(h);
  }
  h->timing = enable != 0;
  if (!name) {
    h->tstat.clear();
    return MGE_OK;
  }
  auto it = h->tstat.find(name);
  if (it == h->tstat.end()) {
    if (ms) *ms = 0.0;
    if (count) *count = 0;
    if (bytes) *bytes = 0.0;
    return MGE_OK;
  }
  const TStat &t = it->second;
  if (ms) *ms = t.count ? t.ms / t.count : 0.0;
  if (count) *count = t.count;
  if (bytes) *bytes = t.count ? t.bytes / t.count : 0.0;
  return MGE_OK;
}

}  // extern "C"
