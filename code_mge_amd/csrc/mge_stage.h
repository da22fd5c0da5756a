// mge_stage.h -- the flux/update kernels of one Jameson-RK stage and the
// wall-record corrections, templated on the bulk type T.
//
// Stage pipeline (euler.f90:118-423), per RK stage m:
//   k_stage2 / k_stage  fluxes + dissipation + blending + forcing + dt
//               scaling (+ the update when the stage is not smoothed) for
//               every cell not touched by a wall record
//   k_rec       the wall corrections of dfluxb.f90/efluxb.f90 in record
//               order for the touched cells, then the same finish
// The wall terms that involve the mirror state (wsol/psol, always FP64) are
// evaluated in FP64 whatever T is.
#pragma once
#include "mge_kernels.h"

namespace mge {

// device error flag: the first code raised stays (the reference stops at
// its first failing check)
__device__ inline void raise_err(int *flag, int code) { atomicCAS(flag, 0, code); }

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
  OP_E_UNDO,         // dw(s)   -= efs(s) - efs(s+1)        efluxb.f90:72,86,100
  OP_E_ADD,          // dw(s)   += fa - fs  or  fs - fa     efluxb.f90:77,81
  OP_ROE_WALL        // Roe schemes: fw(s) += upwind part, dw(s) += central
                     // part of the wall face between s and its mirror state
};

template <typename T>
struct StageArgs {
  FluxCtx<T> fx;
  T *wdst, *pdst;           // output state (update)
  const T *wn;              // W^(0)
  T *dw, *fw;
  T *qf;
  const T *wr;
  const uint8_t *marker;
  const uint8_t *flags;
  const T *vdtim;
  const double *dtim;       // device scalar (global time step)
  T cellvolume;
  const T *vol;             // stretched level: per-cell volumes (else nullptr)
  T blend;                  // 1 - cdis(m)
  T fn;                     // cstp(m)*cfl
  int m;                    // stage number (1-based)
  int has_diss;             // cdis(m) > 0
  int coarse;               // mgl < mlvl
  int lfir1;                // lfir and m == 1
  int lres;
  int smoothed;             // psmoo_old follows (odd stage, smoo > 0)
  int store_fw;             // a later stage reads fw
  int vt;
  int order;                // Roe schemes: 1 first order, 2 MUSCL-minmod
  T eps;                    // Roe entropy-fix coefficient
  unsigned long long *lim;  // Roe second order with limiter freezing: limiter words per direction and cell
                            // (lim[dir*E + q], byte n = component n; else nullptr)
  const int *lfrozen;       // the case's limiter-freeze flag
  T g1;                     // gamma/(gamma-1)
  int kc;                   // k planes per workgroup (LDS-tiled kernels)
  int fw_in;                // fw already holds this stage's blend and viscous part
  // unsteady dual time stepping (euler.f90:249-275, 320-331, 383-393)
  int tsrc;                 // transient term: 0 none, 1 first step (w - w1), 2 BDF2
  int ulam;                 // 1: coarse level divides dw by lam; 2: the update does
  const T *w1, *w2;         // physical time levels n and n-1
  T tcoef;                  // 0.5*cellvolume/dtstep
  T dtstep;
};


template <typename T>
struct RecArgs {
  StageArgs<T> st;
  const DevRec *recs;
  const double *wsol;       // 5*nmir, component-major
  const double *psol;
  int nmir;
  const long *tcell;        // touched cells
  const int *beg;           // op range of each touched cell (ntouch+1)
  const int2 *ops;          // (record, kind), fw ops then dw ops per cell
  const int2 *chunks;       // touched cells [x, y) of each workgroup, at most kRecOps ops
  int ntouch, nops;
  int *err;
};

// ops of one workgroup of k_rec (its threads); no touched cell has more
// than 36 (6 records as cut cell x 4, 6 as neighbour x 2)
constexpr int kRecOps = 64;

// the Roe wall-face term of a record (defined with the Roe scheme)
template <typename T>
__device__ __attribute__((always_inline)) inline void roe_wall_rec(const FluxCtx<T> &c, const DevRec &r, const double *wsol, double psol, double gamma,
                             double eps, double cdis, double *ve, double *vd, int *err);

// ------------------------------------------------------------------ finish of a cell

// dw = dw + fw, forcing, blanking, time-step scaling and (unsmoothed) the
// update with the EOS (euler.f90:217-223, 294-342, 363-399).  With PRE the
// caller has already loaded the cell's W(0) (wnp), local time step (vdt) and
// flags (fl), so that their latency overlapped its face evaluation.  GEN
// false is the form for a launch that is not on a coarse level, not
// unsteady, not on a stretched level and uses the local time step (the
// host picks it only then): those branches and their pointers drop out, which
// frees the scalar registers they held in the stage kernels' plane loops.
template <typename T, bool PRE = false, bool GEN = true>
__device__ inline void finish_cell(const StageArgs<T> &A, long q, const T *f, const T *e, const T *wnp = nullptr,
                                   T vdt = T(0), uint8_t fl = 0) {
  const long E = A.fx.d.E;
  T r[5];
#pragma unroll
  for (int n = 0; n < 5; n++) r[n] = e[n] + f[n];
  if (GEN && A.tsrc) {
    // physical time derivative of the stage state (euler.f90:249-275)
    const T *w = A.fx.s.w;
    const T tc = A.vol ? T(0.5) * A.vol[q] / A.dtstep : A.tcoef;
    if (A.tsrc == 2) {
#pragma unroll
      for (int n = 0; n < 5; n++)
        r[n] = r[n] + tc * (T(3) * w[n * E + q] - T(4) * A.w1[n * E + q] + A.w2[n * E + q]);
    } else {
#pragma unroll
      for (int n = 0; n < 5; n++) r[n] = r[n] + tc * (w[n * E + q] - A.w1[n * E + q]);
    }
  }
  if (GEN && A.coarse) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      T qv;
      if (A.lfir1) {
        qv = T(0);
        if (A.marker[q] == 1) qv = T(1) * A.wr[n * E + q] - r[n];
        A.qf[n * E + q] = qv;
      } else {
        qv = A.qf[n * E + q];
      }
      r[n] = r[n] + qv;
    }
    if (A.ulam == 1) {
      const T lam = unsteady_lam(A.fn, A.vdtim, A.dtim, A.vt, A.dtstep, q);
#pragma unroll
      for (int n = 0; n < 5; n++) r[n] = r[n] / lam;
    }
  }
  T bl = ((PRE ? fl : A.flags[q]) & F_FLUID) ? T(1) : T(0);
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
  // de = fn*vdtim/V, or fn*dtim/V with the global step: one division
  const T dt = GEN && A.vt == 1 ? T(A.dtim[0]) : (PRE ? vdt : A.vdtim[q]);
  const T de = A.fn * dt / (GEN && A.vol ? A.vol[q] : A.cellvolume);
#pragma unroll
  for (int n = 0; n < 5; n++) r[n] = de * r[n];
#pragma unroll
  for (int n = 0; n < 5; n++) A.dw[n * E + q] = r[n];
  if (!A.smoothed) {
    T v[5];
    if (GEN && A.ulam == 2) {
      const T lam = unsteady_lam(A.fn, A.vdtim, A.dtim, A.vt, A.dtstep, q);
#pragma unroll
      for (int n = 0; n < 5; n++) r[n] = r[n] / lam;
    }
#pragma unroll
    for (int n = 0; n < 5; n++) {
      v[n] = (PRE ? wnp[n] : A.wn[n * E + q]) - r[n];
      A.wdst[n * E + q] = v[n];
    }
    A.pdst[q] = eos_st(A.fx.gamma, A.fx.inf, v);
  }
}

// ------------------------------------------------------------------ k_stage
//
// One thread per interior cell, every face evaluated from global memory.
// Used for matrix dissipation (whose per-face state is too large for the
// LDS-tiled kernel's register budget).
template <typename T>
__device__ __attribute__((always_inline)) inline void k_stage_body(const StageArgs<T> &A, int bx, int by, int bz) {
  const Dims &d = A.fx.d;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int j = 2 + by * blockDim.y + threadIdx.y;
  int k = 2 + bz;
  if (i > d.il || j > d.jl || k > d.kl) return;
  const long E = d.E;
  long q = ix(d, i, j, k);
  T f[5];
  if (A.fw_in) {
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = A.fw[n * E + q];
  } else if (A.m > 1) {
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = A.fw[n * E + q] * A.blend;
  } else {
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = T(0);
  }
  if (A.has_diss) {
    T a[5], b[5];
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
  T e[5];
  {
    T a[5], b[5];
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
  finish_cell(A, q, f, e);
}
MGE_ENTRY_T(k_stage, StageArgs, 256, 1)

// ------------------------------------------------------------------ k_stage2
//
// LDS-tiled form of k_stage.  A workgroup owns a TI x TJ column tile of
// (i,j) and marches through a chunk of k planes.  Per plane the raw state of
// the tile and its halo -- w with slot 4 replaced by w4 + p (the enthalpy
// term of the flux and of the dissipation differences), and p -- is staged
// in LDS once: a cross of two-cell arms for w and three-cell arms for p (the
// pressure switch at the tile's second halo cell reads one cell further).
// From it the i- and j-direction quantities every face needs -- the
// convective cell fluxes F (ceflux.f90) and the pressure switches fp
// (dflux.f90:70-81) -- are computed once per cell into LDS, and each thread
// forms the enthalpy differences diw (dflux.f90:82-86) of its faces from the
// staged state.  In k the raw state of the cell's own column k, k+1 (and p
// at k+2, k+3) is carried in registers, so each plane loads one new column
// value per cell;
// the k face values are carried as well, so every k face is evaluated once.
// Each face value is computed with the same expression tree as the
// reference, so results equal k_stage's bit for bit.
constexpr int TI = 64, TJ = 4;

// rows of the JST tile by default: FP64 uses 8 (512 threads, about 100 KB
// LDS, one workgroup per CU -- the two waves per SIMD its registers allow
// anyway, with half the y-halo work per cell); FP32 keeps 4 rows.  Both row
// counts (4 and 8) are built for both precisions; the autotuner may pick the
// other one per level shape.
template <typename T>
struct Stage2Tile {
  static constexpr int TJ = kDelta<T> ? 4 : 8;
  static constexpr int MINW = 2;   // waves per SIMD the launch bounds ask for
};

template <typename T, int TJ_>
struct StageSmem {
  T rw[5][TJ_ + 4][TI + 4];   // w, slot 4 = w4 + p, at i0-2+col, j0-2+row
  T rp[TJ_ + 6][TI + 6];      // p at i0-3+col, j0-3+row
  T fx[5][TJ_][TI + 2];       // F_x at i = i0-1 .. i0+TI
  T fy[5][TJ_ + 2][TI];       // F_y at j = j0-1 .. j0+TJ
  T px[TJ_][TI + 4];          // fp_x at i = i0-2 .. i0+TI+1
  T py[TJ_ + 4][TI];          // fp_y at j = j0-2 .. j0+TJ+1
  T rd[TJ_ + 2][TI + 2];      // radius sums at i0-1+col, j0-1+row (line-clamped)
  T gx[TI + 2];               // dx at i0-1 .. i0+TI
  T gy[TJ_ + 2];              // dy at j0-1 .. j0+TJ
};

// JST face coefficients from the switches fp(a-2..a+1) and the radii sums
// on both sides (dflux.f90:88-96); the face value of component n is then
// (dis2*diw(a) - dis4*(diw(a+1) - 2 diw(a) + diw(a-1))) * lam
template <typename T>
__device__ inline void face_coef(const FluxCtx<T> &c, T f0, T f1, T f2, T f3, T lm, T lp, T *dis2o, T *dis4o,
                                 T *lamo) {
  T ep = maxd(maxd(maxd(f0, f1), f2), f3);
  T lam = T(0.5) * (lm + lp);
  T dis2 = c.vis2 * ep;
  T dis4 = c.vis4;
  dis4 = dimd(dis4, dis2);
  dis2 = dis2 * c.cdis;
  dis4 = dis4 * c.cdis;
  *dis2o = dis2;
  *dis4o = dis4;
  *lamo = lam;
}

template <int DIR, typename T>
__device__ inline void cell_flux(const FluxCtx<T> &c, int i, int j, int k, T *o) {
  long q = ix(c.d, i, j, k);
  T v[5];
  load5(c.s.w, c.d.E, q, v);
  cflux_st<DIR>(c.inf, v, c.s.p[q], face_s<DIR>(c.g, i, j, k), o);
}

// raw state of one cell: w0..w3, w4 + p, p
template <typename T>
__device__ inline void raw_load(const FluxCtx<T> &c, long q, T *o) {
  const long E = c.d.E;
#pragma unroll
  for (int n = 0; n < 5; n++) o[n] = c.s.w[n * E + q];
  o[5] = c.s.p[q];
  o[4] = o[4] + o[5];
}

// convective flux of a staged cell (LDS row r, column col of rw / rp)
template <int DIR, typename T, int TJ_>
__device__ inline void lds_flux(const FluxCtx<T> &c, const StageSmem<T, TJ_> &sm, int r, int col, T S, T *o) {
  T v[5];
#pragma unroll
  for (int n = 0; n < 5; n++) v[n] = sm.rw[n][r][col];
  cflux_st_h<DIR>(c.inf, v, v[4], sm.rp[r + 1][col + 1], S, o);
}

// pressure switch fp(s) of the tile's i line (DIR 0, rp row r) or j line
// (DIR 1, rp column r) from the staged pressures, with fp_at's end
// convention; s0 is the line index of rp's first entry
template <int DIR, typename T, int TJ_>
__device__ inline T lds_fp(const FluxCtx<T> &c, const StageSmem<T, TJ_> &sm, int r, int s, int s0) {
  const int nb = dne<DIR>(c.d) + 1;
  const int sc = (s < 1 ? 1 : (s > nb - 1 ? nb - 1 : s)) - s0;
  T pm, pc, pp;
  if (DIR == 0) {
    pm = sm.rp[r][sc - 1]; pc = sm.rp[r][sc]; pp = sm.rp[r][sc + 1];
  } else {
    pm = sm.rp[sc - 1][r]; pc = sm.rp[sc][r]; pp = sm.rp[sc + 1][r];
  }
  T v = fp_val<DIR>(c.inf, pm, pc, pp);
  return (s <= 0 || s >= nb) ? T(0) : v;
}

// one halo item of the raw staging (t < 6 TJ + 6 TI): x arms (6 columns of
// the tile's rows), then y arms (6 rows of the tile's columns), at rp
// positions (row r, column col); the outermost column / row of each arm is
// needed for p only (w = false) and, like the rest, only where it lies
// inside the arrays
struct HaloItem {
  int r, col;
  long q;   // cell index in plane 0
  bool ok, w;
};
template <int TJ_>
__device__ inline HaloItem halo_item(const Dims &d, int t, int i0, int j0, bool diss) {
  HaloItem h;
  if (t < 6 * TJ_) {
    const int a = t / TJ_;
    h.r = 3 + t % TJ_;
    h.col = a < 3 ? a : TI + a;
    h.w = h.col != 0 && h.col != TI + 5;
  } else {
    const int u = t - 6 * TJ_, a = u / TI;
    h.col = 3 + u % TI;
    h.r = a < 3 ? a : TJ_ + a;
    h.w = h.r != 0 && h.r != TJ_ + 5;
  }
  const int ii = i0 - 3 + h.col, jj = j0 - 3 + h.r;
  h.ok = t < 6 * TJ_ + 6 * TI && ii >= 0 && jj >= 0 && ii <= d.ib && jj <= d.jb && (h.w || diss);
  h.q = h.ok ? ix(d, ii, jj, 0) : 0;
  return h;
}
// one arm item of the radius-sum staging (t < 2 TJ + 2 TI): x arms at
// i0-1 and i0+TI (clamped into the interior along i), y arms at j0-1 and
// j0+TJ (clamped along j), at rd position (r, col)
struct RadItem {
  int r, col;
  long q;
  bool ok;
};
template <int TJ_>
__device__ inline RadItem rad_item(const Dims &d, int t, int i0, int j0, bool diss) {
  RadItem h;
  int ii, jj;
  if (t < 2 * TJ_) {
    h.r = 1 + t % TJ_;
    h.col = t / TJ_ ? TI + 1 : 0;
    ii = max(2, min(i0 - 1 + h.col, d.il));
    jj = j0 - 1 + h.r;
    h.ok = jj <= d.jl;
  } else {
    const int u = t - 2 * TJ_;
    h.col = 1 + u % TI;
    h.r = u / TI ? TJ_ + 1 : 0;
    ii = i0 - 1 + h.col;
    jj = max(2, min(j0 - 1 + h.r, d.jl));
    h.ok = ii <= d.il;
  }
  h.ok = h.ok && diss && t < 2 * TJ_ + 2 * TI;
  h.q = h.ok ? ix(d, ii, jj, 0) : 0;
  return h;
}

template <typename T, int TJs, bool GEN>
__device__ __attribute__((always_inline)) inline void k_stage2_body(const StageArgs<T> &A, int gbx, int gby, int gbz) {
  constexpr int NT = TI * TJs;
  // FP64 loads the next plane's halo items during the current plane's
  // faces (two waves per SIMD hide little latency); FP32 loads them in place
  constexpr bool PRE = !kDelta<T>;
  constexpr bool HPRE = PRE;
  __shared__ StageSmem<T, TJs> sm;
  const FluxCtx<T> &c = A.fx;
  const Dims &d = c.d;
  const long E = d.E;
  const int kc = A.kc;
  const T two = T(2), half = T(0.5);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TI * ty;
  const int i0 = 2 + gbx * TI, j0 = 2 + gby * TJs;
  const int kb0 = 2 + gbz * kc;
  const int kb1 = min(kb0 + kc - 1, d.kl);
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = i <= d.il && j <= d.jl;
  const bool inarr = i <= d.ib && j <= d.jb;
  const bool diss = A.has_diss;
  const long bz = inarr ? ix(d, i, j, 0) : 0;
  const long sz = d.sz;

  // ---- geometry of the tile's lines
  for (int t = tid; t < TI + 2 + TJs + 2; t += NT) {
    if (t < TI + 2) {
      const int ii = i0 - 1 + t;
      if (ii <= d.ib) sm.gx[t] = c.g.dx[ii];
    } else {
      const int jj = j0 - 1 + (t - TI - 2);
      if (jj <= d.jb) sm.gy[t - TI - 2] = c.g.dy[jj];
    }
  }
  const T Sz = valid ? face_s<2>(c.g, i, j, 0) : T(0);

  // ---- staging items of this thread.  Their indices are rederived from
  // an opaque copy of the thread index in every plane: held across the
  // plane loop, the item addresses alone would take dozens of registers
  constexpr int NH = 6 * TJs + 6 * TI;
  constexpr int NS = (NH + NT - 1) / NT;
  const bool rown = diss && i <= d.ie && j <= d.je && !(i > d.il && j > d.jl);
  const long rq_own = rown ? ix(d, min(i, d.il), min(j, d.jl), 0) : 0;
  T hraw[NS][6], rdo = T(0), rdh = T(0), dzk = T(0);
  auto load_items = [&](int k) {
    int t0 = tid;
    asm volatile("" : "+v"(t0));
#pragma unroll
    for (int h = 0; h < NS; h++) {
      const HaloItem it = halo_item<TJs>(d, t0 + h * NT, i0, j0, diss);
      if (it.ok) {
        const long q = it.q + (long)k * sz;
        if (it.w) {
          raw_load(c, q, hraw[h]);
        } else {
          hraw[h][5] = c.s.p[q];
        }
      }
    }
    const RadItem ri = rad_item<TJs>(d, t0, i0, j0, diss);
    if (rown) rdo = c.radsum[rq_own + (long)k * sz];
    if (ri.ok) rdh = c.radsum[ri.q + (long)k * sz];
    dzk = c.g.dz[k];
  };
  if constexpr (HPRE) load_items(kb0);

  // ---- k-direction: raw state of the own column at k and k+1, p at k+2
  // and k+3 (cells inside the arrays); face k values, q(k-1) and the rings
  // for face k+1 (valid cells)
  T R0[6] = {}, R1[6] = {}, P2 = T(0), P3 = T(0);
  T Dk[5], Ek[5], fz[4], rz[2], Fz0[5], Qm[5];
  if (inarr) {
    raw_load(c, bz + (long)kb0 * sz, R0);
    raw_load(c, bz + (long)(kb0 + 1) * sz, R1);
    P2 = c.s.p[bz + (long)(kb0 + 2) * sz];
    P3 = c.s.p[bz + (long)min(kb0 + 3, d.kb) * sz];
  }
  if (valid) {
    const int k = kb0;
    eface<2>(c, bz, k, i, j, k, Ek);
    cflux_st_h<2>(c.inf, R0, R0[4], R0[5], Sz, Fz0);
    if (diss) {
      T t6[6];
      raw_load(c, bz + (long)(k - 1) * sz, t6);
#pragma unroll
      for (int n = 0; n < 5; n++) Qm[n] = t6[n];
      dface<2>(c, bz, k, i, j, k, Dk);
      for (int t = 0; t < 4; t++) fz[t] = fp_at<2>(c, bz, k - 1 + t);
      rz[0] = rad_at<2>(c, bz, k);
      rz[1] = rad_at<2>(c, bz, k + 1);
    }
  }
  for (int k = kb0; k <= kb1; k++) {
    const long qk = bz + (long)k * sz;
    const bool more = k < kb1;
    // ---- the finish inputs of this plane's cell and the next column
    // values, loaded first so that their latency overlaps the plane's
    // staging and faces
    T pfw[5] = {}, pwn[5] = {}, pvd = T(0), W2[5] = {}, P4 = T(0), rz2 = T(0);
    uint8_t pfl = 0;
    if (valid) {
      if (A.fw_in || A.m > 1) {
#pragma unroll
        for (int n = 0; n < 5; n++) pfw[n] = A.fw[n * E + qk];
      }
      if (PRE && !A.lres && !A.smoothed) {
#pragma unroll
        for (int n = 0; n < 5; n++) pwn[n] = A.wn[n * E + qk];
      }
      if (PRE) pvd = A.vdtim[qk];
      pfl = A.flags[qk];
    }
    if (inarr) {
      const long q2 = bz + (long)(k + 2) * sz;   // k + 2 <= kl + 2 = kb
#pragma unroll
      for (int n = 0; n < 5; n++) W2[n] = c.s.w[n * E + q2];
      if (more) P4 = c.s.p[bz + (long)min(k + 4, d.kb) * sz];
      if (valid && diss && more) rz2 = rad_at<2>(c, bz, k + 2);
    }
    const T w24 = W2[4] + P2;   // slot 4 of q(k+2)
    // ---- stage the plane: own cell from the column carry, halo items
    if (!HPRE) load_items(k);
    if (inarr) {
#pragma unroll
      for (int n = 0; n < 5; n++) sm.rw[n][ty + 2][tx + 2] = R0[n];
      sm.rp[ty + 3][tx + 3] = R0[5];
    }
    int tk = tid;
    asm volatile("" : "+v"(tk));
#pragma unroll
    for (int h = 0; h < NS; h++) {
      const HaloItem it = halo_item<TJs>(d, tk + h * NT, i0, j0, diss);
      if (it.ok) {
        if (it.w) {
#pragma unroll
          for (int n = 0; n < 5; n++) sm.rw[n][it.r - 1][it.col - 1] = hraw[h][n];
        }
        sm.rp[it.r][it.col] = hraw[h][5];
      }
    }
    if (rown) sm.rd[ty + 1][tx + 1] = rdo;
    {
      const RadItem ri = rad_item<TJs>(d, tk, i0, j0, diss);
      if (ri.ok) sm.rd[ri.r][ri.col] = rdh;
    }
    const T dz = dzk;
    if constexpr (HPRE) if (more) load_items(k + 1);
    __syncthreads();
    // ---- fill the plane's F and fp (cells within the arrays)
    if (inarr) {
      T o[5];
      lds_flux<0>(c, sm, ty + 2, tx + 2, -(sm.gy[ty + 1] * dz), o);
#pragma unroll
      for (int n = 0; n < 5; n++) sm.fx[n][ty][tx + 1] = o[n];
      lds_flux<1>(c, sm, ty + 2, tx + 2, -(sm.gx[tx + 1] * dz), o);
#pragma unroll
      for (int n = 0; n < 5; n++) sm.fy[n][ty + 1][tx] = o[n];
      if (diss) {
        sm.px[ty][tx + 2] = lds_fp<0>(c, sm, ty + 3, i, i0 - 3);
        sm.py[ty + 2][tx] = lds_fp<1>(c, sm, tx + 3, j, j0 - 3);
      }
    }
    {
      const int nF = 2 * TJs + 2 * TI;
      const int nP = diss ? 4 * TJs + 4 * TI : 0;
      for (int t = tk; t < nF + nP; t += NT) {
        T o[5];
        if (t < 2 * TJs) {  // F_x at i0-1, i0+TI
          const int r = t % TJs, side = t / TJs;
          const int ii = side ? i0 + TI : i0 - 1, jj = j0 + r, col = side ? TI + 1 : 0;
          if (ii <= d.ib && jj <= d.jb) {
            lds_flux<0>(c, sm, r + 2, ii - i0 + 2, -(sm.gy[r + 1] * dz), o);
#pragma unroll
            for (int n = 0; n < 5; n++) sm.fx[n][r][col] = o[n];
          }
        } else if (t < nF) {  // F_y at j0-1, j0+TJs
          const int u = t - 2 * TJs;
          const int cl = u % TI, side = u / TI;
          const int ii = i0 + cl, jj = side ? j0 + TJs : j0 - 1, row = side ? TJs + 1 : 0;
          if (ii <= d.ib && jj <= d.jb) {
            lds_flux<1>(c, sm, jj - j0 + 2, cl + 2, -(sm.gx[cl + 1] * dz), o);
#pragma unroll
            for (int n = 0; n < 5; n++) sm.fy[n][row][cl] = o[n];
          }
        } else if (t < nF + 4 * TJs) {  // fp_x at i0-2, i0-1, i0+TI, i0+TI+1
          const int u = t - nF;
          const int r = u % TJs, w4 = u / TJs;
          const int off = w4 < 2 ? w4 - 2 : TI + (w4 - 2);
          if (j0 + r <= d.jb) sm.px[r][off + 2] = lds_fp<0>(c, sm, r + 3, i0 + off, i0 - 3);
        } else {  // fp_y at j0-2, j0-1, j0+TJs, j0+TJs+1
          const int u = t - nF - 4 * TJs;
          const int cl = u % TI, w4 = u / TI;
          const int off = w4 < 2 ? w4 - 2 : TJs + (w4 - 2);
          if (i0 + cl <= d.ib) sm.py[off + 2][cl] = lds_fp<1>(c, sm, cl + 3, j0 + off, j0 - 3);
        }
      }
    }
    __syncthreads();
    if (valid) {
      T f[5];
      if (A.fw_in) {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = pfw[n];
      } else if (A.m > 1) {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = pfw[n] * A.blend;
      } else {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = T(0);
      }
      T Dk1[5], Ek1[5], Fz1[5];
      if (diss) {
        // face coefficients (dflux.f90:88-96), then per component the two
        // faces of each direction; diw(s) = q(s) - q(s-1) from the staged
        // state at s = a-1, a, a+1, a+2 with the line-end copies diw(1) =
        // diw(2), diw(ie+1) = diw(ie)
        T da2, da4, dal, db2, db4, dbl;
        {
          face_coef(c, sm.px[ty][tx], sm.px[ty][tx + 1], sm.px[ty][tx + 2], sm.px[ty][tx + 3], sm.rd[ty + 1][tx],
                    sm.rd[ty + 1][tx + 1], &da2, &da4, &dal);
          face_coef(c, sm.px[ty][tx + 1], sm.px[ty][tx + 2], sm.px[ty][tx + 3], sm.px[ty][tx + 4],
                    sm.rd[ty + 1][tx + 1], sm.rd[ty + 1][tx + 2], &db2, &db4, &dbl);
          const bool lo = i == 2, hi = i == d.il;
#pragma unroll
          for (int n = 0; n < 5; n++) {
            T q0 = sm.rw[n][ty + 2][tx], q1 = sm.rw[n][ty + 2][tx + 1], q2 = sm.rw[n][ty + 2][tx + 2],
              q3 = sm.rw[n][ty + 2][tx + 3], q4 = sm.rw[n][ty + 2][tx + 4];
            T d1 = q2 - q1, d2 = q3 - q2;
            T d0 = lo ? d1 : q1 - q0, d3 = hi ? d2 : q4 - q3;
            T fa = (da2 * d1 - da4 * (d2 - two * d1 + d0)) * dal;
            T fb = (db2 * d2 - db4 * (d3 - two * d2 + d1)) * dbl;
            f[n] = f[n] + (fa - fb);
          }
        }
        {
          face_coef(c, sm.py[ty][tx], sm.py[ty + 1][tx], sm.py[ty + 2][tx], sm.py[ty + 3][tx], sm.rd[ty][tx + 1],
                    sm.rd[ty + 1][tx + 1], &da2, &da4, &dal);
          face_coef(c, sm.py[ty + 1][tx], sm.py[ty + 2][tx], sm.py[ty + 3][tx], sm.py[ty + 4][tx],
                    sm.rd[ty + 1][tx + 1], sm.rd[ty + 2][tx + 1], &db2, &db4, &dbl);
          const bool lo = j == 2, hi = j == d.jl;
#pragma unroll
          for (int n = 0; n < 5; n++) {
            T q0 = sm.rw[n][ty][tx + 2], q1 = sm.rw[n][ty + 1][tx + 2], q2 = sm.rw[n][ty + 2][tx + 2],
              q3 = sm.rw[n][ty + 3][tx + 2], q4 = sm.rw[n][ty + 4][tx + 2];
            T d1 = q2 - q1, d2 = q3 - q2;
            T d0 = lo ? d1 : q1 - q0, d3 = hi ? d2 : q4 - q3;
            T fa = (da2 * d1 - da4 * (d2 - two * d1 + d0)) * dal;
            T fb = (db2 * d2 - db4 * (d3 - two * d2 + d1)) * dbl;
            f[n] = f[n] + (fa - fb);
          }
        }
        // k face k+1 from diw(k) .. diw(k+2) of the column carry; diw(k+2)
        // is diw(ke) where k+2 = ke+1
        face_coef(c, fz[0], fz[1], fz[2], fz[3], rz[0], rz[1], &da2, &da4, &dal);
        const bool top = k + 2 > d.ke;
#pragma unroll
        for (int n = 0; n < 5; n++) {
          const T q2 = n < 4 ? W2[n] : w24;
          const T z0 = R0[n] - Qm[n], z1 = R1[n] - R0[n], z2 = top ? z1 : q2 - R1[n];
          Dk1[n] = (da2 * z1 - da4 * (z2 - two * z1 + z0)) * dal;
          f[n] = f[n] + (Dk[n] - Dk1[n]);
        }
      }
      T e[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        T ea = half * (sm.fx[n][ty][tx + 1] + sm.fx[n][ty][tx]);
        T eb = half * (sm.fx[n][ty][tx + 2] + sm.fx[n][ty][tx + 1]);
        e[n] = ea - eb;
      }
#pragma unroll
      for (int n = 0; n < 5; n++) {
        T ea = half * (sm.fy[n][ty + 1][tx] + sm.fy[n][ty][tx]);
        T eb = half * (sm.fy[n][ty + 2][tx] + sm.fy[n][ty + 1][tx]);
        e[n] = e[n] + ea - eb;
      }
      cflux_st_h<2>(c.inf, R1, R1[4], R1[5], Sz, Fz1);
#pragma unroll
      for (int n = 0; n < 5; n++) {
        Ek1[n] = half * (Fz1[n] + Fz0[n]);
        e[n] = e[n] + Ek[n] - Ek1[n];
      }
      if (pfl & F_TOUCHED) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          A.fw[n * E + qk] = f[n];
          A.dw[n * E + qk] = e[n];
        }
      } else {
        finish_cell<T, PRE, GEN>(A, qk, f, e, pwn, pvd, pfl);
      }
      // advance the k-direction carry to face k+2
      if (more) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          Ek[n] = Ek1[n];
          Fz0[n] = Fz1[n];
        }
        if (diss) {
          // fp(k+3) from p(k+2..k+4) (0 at and beyond the line end)
          const int s = k + 3, nb = d.ke + 1;
          const T fpn = fp_val<2>(c.inf, P2, P3, P4);
#pragma unroll
          for (int n = 0; n < 5; n++) {
            Dk[n] = Dk1[n];
            Qm[n] = R0[n];
          }
          fz[0] = fz[1];
          fz[1] = fz[2];
          fz[2] = fz[3];
          fz[3] = s >= nb ? T(0) : fpn;
          rz[0] = rz[1];
          rz[1] = rz2;
        }
      }
    }
    // ---- shift the column carry: k+1, k+2 become k, k+1
    if (inarr && more) {
#pragma unroll
      for (int n = 0; n < 6; n++) R0[n] = R1[n];
#pragma unroll
      for (int n = 0; n < 4; n++) R1[n] = W2[n];
      R1[4] = w24;
      R1[5] = P2;
      P2 = P3;
      P3 = P4;
    }
    __syncthreads();
  }
}
// entry points as MGE_ENTRY_T_XCD makes them (XCD-contiguous block order in
// the single-case entry), with launch bounds that follow the tile's rows
template <typename T, int TJ_, bool GEN>
__global__ void __launch_bounds__(TI * TJ_, Stage2Tile<T>::MINW) k_stage2(StageArgs<T> A) {
  int bx, by, bz;
  xcd_block(bx, by, bz);
  k_stage2_body<T, TJ_, GEN>(A, bx, by, bz);
}
template <typename T, int TJ_, bool GEN>
__global__ void __launch_bounds__(TI * TJ_, Stage2Tile<T>::MINW)
    k_stage2_b(const StageArgs<T> *As, const int4 *cg, int nc, const int *act) {
  const int b = (int)blockIdx.x;
  const int c = batch_case(cg, nc, b);
  const int4 g = cg[c];
  if (!act[g.w]) return;
  const int l = b - g.x;
  k_stage2_body<T, TJ_, GEN>(As[c], l % g.y, (l / g.y) % g.z, l / (g.y * g.z));
}

// ------------------------------------------------------------------ wall records

template <int DIR, typename T>
__device__ inline void rec_dfs(const FluxCtx<T> &c, int i, int j, int k, int a, T *o) {
  long base = DIR == 0 ? ix(c.d, 0, j, k) : (DIR == 1 ? ix(c.d, i, 0, k) : ix(c.d, i, j, 0));
  dface<DIR>(c, base, a, i, j, k, o);
}
template <typename T>
__device__ inline void rec_dfs_any(const FluxCtx<T> &c, int zz, int i, int j, int k, int a, T *o) {
  if (zz == 1) rec_dfs<0>(c, i, j, k, a, o);
  else if (zz == 2) rec_dfs<1>(c, i, j, k, a, o);
  else rec_dfs<2>(c, i, j, k, a, o);
}
template <int DIR, typename T>
__device__ inline void rec_efs(const FluxCtx<T> &c, int i, int j, int k, int a, T *o) {
  long base = DIR == 0 ? ix(c.d, 0, j, k) : (DIR == 1 ? ix(c.d, i, 0, k) : ix(c.d, i, j, 0));
  eface<DIR>(c, base, a, i, j, k, o);
}
template <typename T>
__device__ inline void rec_efs_any(const FluxCtx<T> &c, int zz, int i, int j, int k, int a, T *o) {
  if (zz == 1) rec_efs<0>(c, i, j, k, a, o);
  else if (zz == 2) rec_efs<1>(c, i, j, k, a, o);
  else rec_efs<2>(c, i, j, k, a, o);
}

// face_s<zz-1> for a runtime zz: the three vertex differences are loaded
// and the product selected (branches on zz that each load through a
// different pointer of g are merged into one load through a selected field
// address, which keeps the kernel's argument struct in scratch memory)
template <typename T>
__device__ inline T face_s_any(const Geo<T> &g, int zz, int i, int j, int k) {
  const T ax = g.dx[i], ay = g.dy[j], az = g.dz[k];
  const T s0 = -(ay * az), s1 = -(ax * az), s2 = -(ax * ay);
  return zz == 1 ? s0 : (zz == 2 ? s1 : s2);
}

// the face vector of a zz-face (zz = 1, 2, 3): component zz-1 is s, the
// others 0, set by selects (a runtime index into a local array would put
// the array in scratch memory)
__device__ inline void axis_vec(double *v, int zz, double s) {
  v[0] = zz == 1 ? s : 0.0;
  v[1] = zz == 2 ? s : 0.0;
  v[2] = zz == 3 ? s : 0.0;
}

// the two recomputed faces of dfluxb.f90 for record r: fs1 (inner face of
// the pair) and fs2 (the wall face or the face behind it).  They involve the
// FP64 mirror state, so they are evaluated in FP64 from the bulk values.
template <typename T>
__device__ __attribute__((always_inline)) inline void dfluxb_rec(const FluxCtx<T> &c, const DevRec &r, const double *wsolv, double psol, double *fs1,
                           double *fs2) {
  const Dims &d = c.d;
  const long E = d.E;
  int zz = r.zz, id = r.faceid;
  int i = r.c[0], j = r.c[1], k = r.c[2];
  int s0 = zz == 1 ? i : (zz == 2 ? j : k);
  long st = zz == 1 ? 1L : (zz == 2 ? d.sy : d.sz);
  long base = ix(d, i, j, k) - (long)s0 * st;
  // a sum of the three values rather than a select between the fields (see
  // face_s_any)
  const int nbmax = (zz == 1) * d.ib + (zz == 2) * d.jb + (zz == 3) * d.kb;
  const double vis2 = c.vis2, vis4 = c.vis4, cdis = c.cdis;
  double wsol[5];
  for (int n = 0; n < 5; n++) wsol[n] = wsolv[n];
  wsol[4] = wsol[4] + psol;
  auto P = [&](int s) { return full(c.inf, 5, c.s.p[base + (long)s * st]); };
  auto H = [&](int s, int n) {
    long q = base + (long)s * st;
    return n < 4 ? full(c.inf, n, c.s.w[n * E + q]) : full(c.inf, 4, c.s.w[4 * E + q]) + full(c.inf, 5, c.s.p[q]);
  };
  auto RS = [&](int s) { return (double)c.radsum[base + (long)s * st]; };
  auto SS = [&](int s, double *ss) {
    int fi = zz == 1 ? s : i, fj = zz == 2 ? s : j, fk = zz == 3 ? s : k;
    axis_vec(ss, zz, (double)face_s_any(c.g, zz, fi, fj, fk));
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
    dis2 = vis2 * ep; dis4 = vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * cdis; dis4 = dis4 * cdis;
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
      dmatrix((double)c.gamma, uc, wsol, fs2, ss, o);
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
    dis2 = vis2 * ep; dis4 = vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * cdis; dis4 = dis4 * cdis;
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
      dmatrix((double)c.gamma, um, uc, fs1, ss, o);
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
    dis2 = vis2 * ep; dis4 = vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * cdis; dis4 = dis4 * cdis;
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
      dmatrix((double)c.gamma, uc, up, fs2, ss, o);
      for (int n = 0; n < 5; n++) fs2[n] = o[n];
    }
    ep = maxd(fp[0], fp[1]);
    lp = RS(s);
    lam = lp;
    dis2 = vis2 * ep; dis4 = vis4; dis4 = dimd(dis4, dis2);
    dis2 = dis2 * cdis; dis4 = dis4 * cdis;
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
      dmatrix((double)c.gamma, wsol, uc, fs1, ss, o);
      for (int n = 0; n < 5; n++) fs1[n] = o[n];
    }
  }
}

// the wall flux fs and the opposite-face flux fa of efluxb.f90:50-80,
// in FP64 like dfluxb_rec
template <typename T>
__device__ __attribute__((always_inline)) inline void efluxb_rec(const FluxCtx<T> &c, const DevRec &r, const double *wsol, double psol, double *fs,
                           double *fa, int *err) {
  const Dims &d = c.d;
  int zz = r.zz, id = r.faceid;
  int i = r.c[0], j = r.c[1], k = r.c[2];
  long qc = ix(d, i, j, k);
  double sc[3], ss[3];
  axis_vec(sc, zz, (double)face_s_any(c.g, zz, i, j, k));
  axis_vec(ss, zz, (double)face_s_any(c.g, zz, r.s[0], r.s[1], r.s[2]));
  double wc[5];
  for (int n = 0; n < 5; n++) wc[n] = full(c.inf, n, c.s.w[n * d.E + qc]);
  if (fabs(wc[0]) < (double)1.0e-8f) raise_err(err, MGE_ERR_WALL_RHO);
  double fs1[5], fs2[5], fs3[5];
  cflux3(wc, full(c.inf, 5, c.s.p[qc]), sc, fs1);
  cflux3(wsol, psol, ss, fs2);
  for (int n = 0; n < 5; n++) fs[n] = 0.5 * (fs1[n] + fs2[n]);
  int hi = (id >= 4);
  int oi = i, oj = j, ok = k;
  int dlt = hi ? -1 : 1;
  if (zz == 1) oi += dlt;
  else if (zz == 2) oj += dlt;
  else ok += dlt;
  long qo = ix(d, oi, oj, ok);
  double wo[5], so[3];
  for (int n = 0; n < 5; n++) wo[n] = full(c.inf, n, c.s.w[n * d.E + qo]);
  axis_vec(so, zz, (double)face_s_any(c.g, zz, oi, oj, ok));
  cflux3(wo, full(c.inf, 5, c.s.p[qo]), so, fs3);
  for (int n = 0; n < 5; n++) fa[n] = 0.5 * (fs1[n] + fs3[n]);
}

// the value op o contributes (dfluxb.f90 / efluxb.f90 terms); the undo
// terms repeat the bulk kernel's face values in T
template <typename T>
__device__ __attribute__((always_inline)) inline void rec_op_value(const RecArgs<T> &R, int o, double *out) {
  const StageArgs<T> &A = R.st;
  int2 op = R.ops[o];
  const DevRec &r = R.recs[op.x];
  if (op.y < OP_E_UNDO && !A.has_diss) return;
  int zz = r.zz;
  int s0 = zz == 1 ? r.c[0] : (zz == 2 ? r.c[1] : r.c[2]);
  double ws[5];
  for (int n = 0; n < 5; n++) ws[n] = R.wsol[n * R.nmir + op.x];
  double ps = R.psol[op.x];
  T a[5], b[5];
  double fs1[5], fs2[5];
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
      for (int n = 0; n < 5; n++) out[n] = (double)a[n] - fs1[n];
      break;
    case OP_LO_N_ADD:
      dfluxb_rec(A.fx, r, ws, ps, fs1, fs2);
      rec_dfs_any(A.fx, zz, ci, cj, ck, s0 + 2, a);
      for (int n = 0; n < 5; n++) out[n] = fs2[n] - (double)a[n];
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
    case OP_ROE_WALL: {
      double ve[5], vd[5];
      roe_wall_rec(A.fx, r, ws, ps, (double)A.fx.gamma, (double)A.eps, (double)A.fx.cdis, ve, vd, R.err);
      for (int n = 0; n < 5; n++) {
        out[n] = vd[n];
        out[5 + n] = ve[n];
      }
      break;
    }
    default:
      break;
  }
}

// The wall records of a stage: each workgroup takes whole touched cells
// with their ops (at most kRecOps), one thread per op computes its value
// into LDS, then one thread per cell adds its ops' values in the reference
// order (in FP64) and finishes the cell in T.
template <typename T>
__device__ __attribute__((always_inline)) inline void k_rec_body(const RecArgs<T> &R, int bx, int, int) {
  __shared__ double val[kRecOps * 10];
  const int2 ch = R.chunks[bx];
  const int o0 = R.beg[ch.x], o1 = R.beg[ch.y];
  const int tid = threadIdx.x;
  const int t = ch.x + tid;
  const StageArgs<T> &A = R.st;
  const Dims &d = A.fx.d;
  const long E = d.E;
  // the cell's fw and dw, and the finish's inputs (W(0) on an unsmoothed
  // stage, the local time step, the flags), are loaded before the op values
  // are formed, so their round trips overlap the ops' own
  const bool cell = t < ch.y;
  const long q = cell ? R.tcell[t] : 0;
  double f[5], e[5];
  T wnp[5] = {T(0), T(0), T(0), T(0), T(0)};
  T vdt = T(0);
  uint8_t fl = 0;
  if (cell) {
    for (int n = 0; n < 5; n++) {
      f[n] = A.fw[n * E + q];
      e[n] = A.dw[n * E + q];
    }
    fl = A.flags[q];
    vdt = A.vdtim[q];
    if (!A.smoothed && !A.lres)
      for (int n = 0; n < 5; n++) wnp[n] = A.wn[n * E + q];
  }
  if (o0 + tid < o1) rec_op_value(R, o0 + tid, val + tid * 10);
  __syncthreads();
  if (!cell) return;
  for (int o = R.beg[t]; o < R.beg[t + 1]; o++) {
    int kind = R.ops[o].y;
    const double *v = val + (o - o0) * 10;
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
      case OP_ROE_WALL:
        for (int n = 0; n < 5; n++) {
          if (A.has_diss) f[n] = f[n] + v[n];
          e[n] = e[n] + v[5 + n];
        }
        break;
      default:
        break;
    }
  }
  T ft[5], et[5];
  for (int n = 0; n < 5; n++) {
    ft[n] = (T)f[n];
    et[n] = (T)e[n];
  }
  finish_cell<T, true>(A, q, ft, et, wnp, vdt, fl);
}
MGE_ENTRY_T(k_rec, RecArgs, 64, 1)

}  // namespace mge
