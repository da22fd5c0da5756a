// mge_visc.h -- laminar viscous fluxes (the role of viscpak's fgradi/j/k
// and vflux, called from euler.f90:137-181 when cdis(m) > 0).
//
// viscpak has no source in the reference tree; the flux is this build's
// own and the CPU restatement evaluates the same expressions:
//  - velocity and temperature per cell (euler.f90:140-150: vel = m/rho,
//    tem = p/(rho R)); the FP32 path keeps their deviations from the
//    freestream so that differences of neighbouring cells keep their digits;
//  - face gradient: normal derivative (phi_a - phi_(a-1))/hn with hn the
//    centre distance; tangential derivatives the mean of the two cells'
//    centred differences, one-sided next to a blank cell, zero between two;
//  - tau = mu (G + G^T) - 2/3 mu div(u) I, q = -kappa dT/dn,
//    kappa = mu gamma R / ((gamma-1) Pr); flux (0, tau_n, u_f . tau_n - q_n)
//    times the face vector S, u_f the mean of the two cells;
//  - no bulk flux through a face with a blank side; the wall records add the
//    flux of the face between each cut cell and its no-slip mirror state
//    (normal derivatives only), in FP64 and in record order;
//  - per cell fw = blend fw + cdis ((gx_hi - gx_lo) + (gy_hi - gy_lo) +
//    (gz_hi - gz_lo)) before the dissipation of the stage is added.
#pragma once
#include "mge_stage.h"

namespace mge {

template <typename T>
struct ViscArgs {
  Dims d;
  StateC<T> s;
  const uint8_t *flags;
  Geo<T> g;
  T *prim;       // u, v, w, T per cell (4E); FP32: deviations from the freestream
  T *fw;
  Inf<T> inf;
  T uinf[3];     // freestream velocity (FP32 work term)
  T tinf;        // freestream temperature
  T rgas, mu, kap, cdis, blend;
  int m;         // stage; fw is blended in only for m > 1
};

// euler.f90:140-150 on every cell of the block
template <typename T>
__device__ __attribute__((always_inline)) inline void k_visc_prim_body(const ViscArgs<T> &A, int bx, int, int) {
  const Dims &d = A.d;
  const long E = d.E;
  long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= E) return;
  T o[4];
  if (!(A.flags[q] & F_FLUID)) {
    o[0] = o[1] = o[2] = o[3] = T(0);
  } else if constexpr (kDelta<T>) {
    const T r1 = A.s.w[q];
    const T rho = A.inf.w[0] + r1;
#pragma unroll
    for (int a = 0; a < 3; a++) o[a] = (A.s.w[(a + 1) * E + q] - A.uinf[a] * r1) / rho;
    o[3] = (A.s.p[q] - (A.tinf * A.rgas) * r1) / (rho * A.rgas);
  } else {
    const T r = A.s.w[q];
#pragma unroll
    for (int a = 0; a < 3; a++) o[a] = A.s.w[(a + 1) * E + q] / r;
    o[3] = A.s.p[q] / (r * A.rgas);
  }
#pragma unroll
  for (int a = 0; a < 4; a++) A.prim[a * E + q] = o[a];
}
MGE_ENTRY_T(k_visc_prim, ViscArgs, 256, 1)

// one face: pL/pR velocity + temperature of the two cells (FP32: deviations;
// u0 is then the freestream velocity, zero in FP64), t1*/t2* the tangential
// velocity derivatives of the two cells
template <typename U>
__device__ inline void visc_face(int dir, const U *pL, const U *pR, const U *t1L, const U *t1R, const U *t2L,
                                 const U *t2R, U hn, U S, U mu, U kap, const U *u0, U *g) {
  const int t1 = dir == 0 ? 1 : 0, t2 = dir == 2 ? 1 : 2;
  U G[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    G[a][dir] = (pR[a] - pL[a]) / hn;
    G[a][t1] = U(0.5) * (t1L[a] + t1R[a]);
    G[a][t2] = U(0.5) * (t2L[a] + t2R[a]);
  }
  U dT = (pR[3] - pL[3]) / hn;
  U div = (G[0][0] + G[1][1]) + G[2][2];
  U tau[3];
#pragma unroll
  for (int a = 0; a < 3; a++) tau[a] = mu * (G[a][dir] + G[dir][a]);
  tau[dir] = mu * (U(2) * G[dir][dir] - U(2.0 / 3.0) * div);
  U uf[3];
#pragma unroll
  for (int a = 0; a < 3; a++) uf[a] = u0 ? u0[a] + U(0.5) * (pL[a] + pR[a]) : U(0.5) * (pL[a] + pR[a]);
  U qn = -kap * dT;
  g[0] = U(0) * S;
  g[1] = tau[0] * S;
  g[2] = tau[1] * S;
  g[3] = tau[2] * S;
  g[4] = (((uf[0] * tau[0] + uf[1] * tau[1]) + uf[2] * tau[2]) - qn) * S;
}

// centred (one-sided next to a blank) velocity difference of cell q along
// an axis with stride st; c the cell's index on it, h the axis' widths
template <typename T>
__device__ inline void visc_dt(const ViscArgs<T> &A, long q, long st, int c, const T *h, T *o) {
  const long E = A.d.E;
  const bool fp = (A.flags[q + st] & F_FLUID) != 0, fm = (A.flags[q - st] & F_FLUID) != 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const T *v = A.prim + a * E;
    if (fp && fm)
      o[a] = (v[q + st] - v[q - st]) / (T(0.5) * (h[c - 1] + h[c]) + T(0.5) * (h[c] + h[c + 1]));
    else if (fp)
      o[a] = (v[q + st] - v[q]) / (T(0.5) * (h[c] + h[c + 1]));
    else if (fm)
      o[a] = (v[q] - v[q - st]) / (T(0.5) * (h[c - 1] + h[c]));
    else
      o[a] = T(0);
  }
}

// the bulk face in direction DIR whose higher cell is q (index ijk)
template <int DIR, typename T>
__device__ inline void visc_bulk_face(const ViscArgs<T> &A, long q, int i, int j, int k, T *g) {
  const Dims &d = A.d;
  const long E = d.E;
  const long st[3] = {1L, (long)d.sy, (long)d.sz};
  const int ijk[3] = {i, j, k};
  const T *hh[3] = {A.g.dx, A.g.dy, A.g.dz};
  constexpr int t1 = DIR == 0 ? 1 : 0, t2 = DIR == 2 ? 1 : 2;
  const long sd = st[DIR];
  if (!(A.flags[q] & F_FLUID) || !(A.flags[q - sd] & F_FLUID)) {
#pragma unroll
    for (int n = 0; n < 5; n++) g[n] = T(0);
    return;
  }
  T pL[4], pR[4], t1L[3], t1R[3], t2L[3], t2R[3];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    pL[a] = A.prim[a * E + q - sd];
    pR[a] = A.prim[a * E + q];
  }
  visc_dt(A, q - sd, st[t1], ijk[t1], hh[t1], t1L);
  visc_dt(A, q, st[t1], ijk[t1], hh[t1], t1R);
  visc_dt(A, q - sd, st[t2], ijk[t2], hh[t2], t2L);
  visc_dt(A, q, st[t2], ijk[t2], hh[t2], t2R);
  const T *hd = hh[DIR];
  T hn = T(0.5) * (hd[ijk[DIR] - 1] + hd[ijk[DIR]]);
  T S = face_s<DIR>(A.g, i, j, k);
  visc_face<T>(DIR, pL, pR, t1L, t1R, t2L, t2R, hn, S, A.mu, A.kap, kDelta<T> ? A.uinf : nullptr, g);
}

// fw = blend fw + cdis * (viscous balance), every interior cell
template <typename T>
__device__ __attribute__((always_inline)) inline void k_visc_body(const ViscArgs<T> &A, int bx, int by, int bz) {
  const Dims &d = A.d;
  const long E = d.E;
  const int i = 2 + bx * blockDim.x + threadIdx.x;
  const int j = 2 + by * blockDim.y + threadIdx.y;
  const int k = 2 + bz;
  if (i > d.il || j > d.jl || k > d.kl) return;
  const long q = ix(d, i, j, k);
  T acc[5], glo[5], ghi[5];
#pragma unroll
  for (int n = 0; n < 5; n++) acc[n] = T(0);
  visc_bulk_face<0>(A, q, i, j, k, glo);
  visc_bulk_face<0>(A, q + 1, i + 1, j, k, ghi);
#pragma unroll
  for (int n = 0; n < 5; n++) acc[n] = acc[n] + (ghi[n] - glo[n]);
  visc_bulk_face<1>(A, q, i, j, k, glo);
  visc_bulk_face<1>(A, q + d.sy, i, j + 1, k, ghi);
#pragma unroll
  for (int n = 0; n < 5; n++) acc[n] = acc[n] + (ghi[n] - glo[n]);
  visc_bulk_face<2>(A, q, i, j, k, glo);
  visc_bulk_face<2>(A, q + d.sz, i, j, k + 1, ghi);
#pragma unroll
  for (int n = 0; n < 5; n++) acc[n] = acc[n] + (ghi[n] - glo[n]);
#pragma unroll
  for (int n = 0; n < 5; n++) {
    T f = A.m > 1 ? A.fw[n * E + q] * A.blend : T(0);
    A.fw[n * E + q] = f + A.cdis * acc[n];
  }
}
MGE_ENTRY_T(k_visc, ViscArgs, 256, 1)

template <typename T>
struct ViscWallArgs {
  ViscArgs<T> v;
  const DevRec *recs;
  const double *wsol, *psol;
  int nmir;
  const long *vcell;  // cut cells with wall records, in order of first record
  const int *vbeg;    // records of vcell[t]: vrec[vbeg[t] .. vbeg[t+1])
  const int *vrec;
  int nvcell;
};

// the wall-face viscous flux of record r (FP64); returns whether the face
// is the cut cell's high face
template <typename T>
__device__ inline bool visc_wall_flux(const ViscArgs<T> &A, const DevRec &r, const double *ws, double ps,
                                      double *g) {
  const Dims &d = A.d;
  const long E = d.E;
  const int dir = r.zz - 1;
  const bool hi = r.faceid >= 4;
  const long qc = ix(d, r.c[0], r.c[1], r.c[2]);
  double pc[4], pg[4];
  double rc = full(A.inf, 0, A.s.w[qc]);
  for (int a = 0; a < 3; a++) pc[a] = full(A.inf, a + 1, A.s.w[(a + 1) * E + qc]) / rc;
  pc[3] = full(A.inf, 5, A.s.p[qc]) / (rc * (double)A.rgas);
  pg[0] = ws[1] / ws[0];
  pg[1] = ws[2] / ws[0];
  pg[2] = ws[3] / ws[0];
  pg[3] = ps / (ws[0] * (double)A.rgas);
  const T *hd = dir == 0 ? A.g.dx : (dir == 1 ? A.g.dy : A.g.dz);
  const int c = r.c[dir];
  const double zero[3] = {0.0, 0.0, 0.0};
  const double mu = (double)A.mu, kap = (double)A.kap;
  if (hi) {
    double S = face_s_any(A.g, r.zz, r.s[0], r.s[1], r.s[2]);
    visc_face<double>(dir, pc, pg, zero, zero, zero, zero, 0.5 * ((double)hd[c] + (double)hd[c + 1]), S, mu, kap,
                      nullptr, g);
  } else {
    double S = face_s_any(A.g, r.zz, r.c[0], r.c[1], r.c[2]);
    visc_face<double>(dir, pg, pc, zero, zero, zero, zero, 0.5 * ((double)hd[c - 1] + (double)hd[c]), S, mu, kap,
                      nullptr, g);
  }
  return hi;
}

// per cut cell: fw += cdis * (+-g) over its wall records in record order
template <typename T>
__device__ __attribute__((always_inline)) inline void k_visc_wall_body(const ViscWallArgs<T> &W, int bx, int, int) {
  const int t = bx * blockDim.x + threadIdx.x;
  if (t >= W.nvcell) return;
  const ViscArgs<T> &A = W.v;
  const long E = A.d.E;
  const long q = W.vcell[t];
  double f[5];
  for (int n = 0; n < 5; n++) f[n] = (double)A.fw[n * E + q];
  for (int o = W.vbeg[t]; o < W.vbeg[t + 1]; o++) {
    const int m = W.vrec[o];
    double ws[5], g[5];
    for (int n = 0; n < 5; n++) ws[n] = W.wsol[n * W.nmir + m];
    bool hi = visc_wall_flux(A, W.recs[m], ws, W.psol[m], g);
    for (int n = 0; n < 5; n++) f[n] = f[n] + (double)A.cdis * (hi ? g[n] : -g[n]);
  }
  for (int n = 0; n < 5; n++) A.fw[n * E + q] = (T)f[n];
}
MGE_ENTRY_T(k_visc_wall, ViscWallArgs, 64, 1)

// wall force: adds each record's viscous part (minus the viscous momentum
// flux leaving the fluid through its wall face) to the record's term
template <typename T>
__global__ void k_wall_visc_terms(ViscArgs<T> A, const DevRec *recs, const double *wsol, const double *psol,
                                  int nmir, double *terms) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmir) return;
  double ws[5], g[5];
  for (int n = 0; n < 5; n++) ws[n] = wsol[n * nmir + m];
  const bool hi = visc_wall_flux(A, recs[m], ws, psol[m], g);
  for (int dd = 0; dd < 3; dd++) terms[3 * m + dd] = terms[3 * m + dd] + (hi ? g[1 + dd] : -g[1 + dd]);
}

}  // namespace mge
