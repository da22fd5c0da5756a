// mge_roe.h -- Roe flux with MUSCL-minmod reconstruction (first or second
// order), the stage kernel that applies it and its wall-record term.
//
// The reference calls var_diffs / dflux_roe2 / eflux_roe2 and dflux_roe1 /
// eflux_roe1 from euler.f90:201-210; the roepak module has no source in the
// reference tree, so the scheme is written from the mathematics:
//  - face states from primitive variables q = (rho,u,v,w,p); second order
//    q_L = q(a-1) + minmod(dq(a-1), dq(a))/2, q_R = q(a) - minmod(dq(a),
//    dq(a+1))/2, with a slope that would reach a solid cell set to zero;
//  - central part 0.5*(F_S(q_L) + F_S(q_R)), the physical fluxes written from
//    the primitive face states (accumulated like eflux into dw);
//  - upwind part cdis*0.5*|A_roe(S)|(U_R - U_L) from Roe's averages and
//    three wave families with Harten's entropy fix (accumulated like dflux
//    into fw, so it is blended across stages as the JST dissipation is);
//  - no bulk flux through a face between a fluid and a solid cell: each
//    wall record adds the Roe flux between its cut cell (first order) and
//    its FP64 mirror state.
// The face vector S of the reference points from cell a to cell a-1, so
// e + d is minus the outward flux through the face times |S|.
// The bulk-face arithmetic is compiled with FMA contraction (the FP64 wall
// term and the CPU restatement are not), so the GPU Roe path agrees with
// the restatement to rounding rather than bit for bit.
#pragma once
#include <type_traits>

#include "mge_stage.h"

namespace mge {

// minmod(a, b): 0 when a and b differ in sign or one is zero, else the
// argument of smaller magnitude.  Written with min/max only (no selects, no
// product): max(min(a,b),0) + min(max(a,b),0) is the smaller of two positive
// arguments, the larger (nearer zero) of two negative ones and 0 otherwise,
// and the sum adds an exact zero, so the value is that argument itself.
template <typename T>
__device__ inline T minmod(T a, T b) {
  return fmax(fmin(a, b), T(0)) + fmin(fmax(a, b), T(0));
}

// primitive variables of a cell from the stored state (FP32: the full
// values rebuilt from the deviations)
template <typename T>
__device__ inline void prim_raw(const T *raw, const Inf<T> &f, T *o);

template <typename T>
__device__ inline void prim_of(const StateC<T> &s, const Inf<T> &f, long E, long q, T *o) {
  const T raw[5] = {s.w[q], s.w[E + q], s.w[2 * E + q], s.w[3 * E + q], s.p[q]};
  prim_raw(raw, f, o);
}

// the same from the loaded values rho, m1, m2, m3, p of a cell
template <typename T>
__device__ inline void prim_raw(const T *raw, const Inf<T> &f, T *o) {
  T r = raw[0], m1 = raw[1], m2 = raw[2], m3 = raw[3], p = raw[4];
  if constexpr (kDelta<T>) {
    r = r + f.w[0];
    m1 = m1 + f.w[1];
    m2 = m2 + f.w[2];
    m3 = m3 + f.w[3];
    p = p + f.p;
  }
  T rr = rcp_pos(r);
  o[0] = r;
  o[1] = m1 * rr;
  o[2] = m2 * rr;
  o[3] = m3 * rr;
  o[4] = p;
}

// physical flux of a primitive face state through the face vector whose
// single non-zero component S lies in direction DIR; returns the total
// enthalpy H = gamma/(gamma-1) p/rho + |u|^2/2 (g1 = gamma/(gamma-1))
template <int DIR, typename T>
__device__ inline T roe_state_flux(T g1, const T *q, T S, T *F, T &rr) {
#pragma clang fp contract(fast)
  rr = rcp_pos(q[0]);
  T q2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  T H = g1 * (q[4] * rr) + T(0.5) * q2;
  T m = q[0] * (q[1 + DIR] * S);
  F[0] = m;
#pragma unroll
  for (int c = 0; c < 3; c++) F[1 + c] = m * q[1 + c];
  F[1 + DIR] = m * q[1 + DIR] + q[4] * S;
  F[4] = m * H;
  return H;
}

// the same for a general face vector (FP64 wall faces)
__device__ inline double roe_state_flux_vec(double g1, const double *q, const double *S, double *F) {
  double rr = 1.0 / q[0];
  double q2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double H = g1 * (q[4] * rr) + 0.5 * q2;
  double unS = q[1] * S[0] + q[2] * S[1] + q[3] * S[2];
  double m = q[0] * unS;
  F[0] = m;
  F[1] = m * q[1] + q[4] * S[0];
  F[2] = m * q[2] + q[4] * S[1];
  F[3] = m * q[3] + q[4] * S[2];
  F[4] = m * H;
  return H;
}

// Harten's fix: |l| below del becomes (l^2 + del^2)/(2 del); hd = 0.5/del
template <typename T>
__device__ inline T fix_abs(T l, T del, T hd) {
  T a = fabs(l);
  T f = (l * l + del * del) * hd;
  return a < del ? f : a;
}

// face states of one face from the primitives of cells a-2..a+1 and the
// fluid flags of the outer two
template <typename T>
__device__ inline void muscl(int order, const T *qm2, const T *qm1, const T *qp0, const T *qp1, bool fm2, bool fp1,
                             T *qL, T *qR) {
#pragma clang fp contract(fast)
  if (order < 2) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      qL[n] = qm1[n];
      qR[n] = qp0[n];
    }
    return;
  }
#pragma unroll
  for (int n = 0; n < 5; n++) {
    T sl = fm2 ? minmod(qm1[n] - qm2[n], qp0[n] - qm1[n]) : T(0);
    T sr = fp1 ? minmod(qp0[n] - qm1[n], qp1[n] - qp0[n]) : T(0);
    qL[n] = qm1[n] + T(0.5) * sl;
    qR[n] = qp0[n] - T(0.5) * sr;
  }
}

// Limiter freezing (lfreeze > 0, second order).  For cell c, direction dir
// and component n the limiter value phi = minmod(a,b) / (0.5 (a+b)) of its
// backward and forward differences a, b (0 when they differ in sign or one
// is zero, else in (0,1]) is kept as the byte round(254 phi) at byte n of
// the cell's word of that direction (lim[dir*E + q]); 255 means not
// recorded (ghost cells, which keep the minmod itself).  Until the freeze
// each stage records the values of its cells; after it a cell's slope is
// (byte/254) * 0.5 (a+b): a fixed blend of first order and Fromm's central
// slope, linear in the state and always dissipative, so the iteration
// converges instead of cycling through limiter switches.  The wall rule (no
// slope reaching a solid cell) applies in both modes.
template <typename T>
__device__ inline T lim_frozen(T back, T fwd, unsigned byte) {
  const T k = T(0.5 / 254.0);
  return byte == 255u ? minmod(back, fwd) : (T(byte) * k) * (back + fwd);
}
template <typename T>
__device__ inline unsigned lim_byte(T m, T back, T fwd) {
  const T phi = back * fwd > T(0) ? m / (T(0.5) * (back + fwd)) : T(0);
  return (unsigned)(T(254) * phi + T(0.5));
}

// muscl with the limiter mode lm (1 record: *rec gets the word of cell a;
// 2 frozen: the slopes of cells a-1 and a from their words wl, wr)
template <typename T>
__device__ inline void muscl_lim(int lm, const T *qm2, const T *qm1, const T *qp0, const T *qp1, bool fm2, bool fp1,
                                 unsigned long long wl, unsigned long long wr, T *qL, T *qR,
                                 unsigned long long *rec) {
  unsigned long long word = ~0ull;
#pragma unroll
  for (int n = 0; n < 5; n++) {
    const T a1 = qm1[n] - qm2[n], b1 = qp0[n] - qm1[n], b2 = qp1[n] - qp0[n];
    T sl, sr;
    if (lm == 2) {
      sl = lim_frozen(a1, b1, (unsigned)(wl >> (8 * n)) & 255u);
      sr = lim_frozen(b1, b2, (unsigned)(wr >> (8 * n)) & 255u);
    } else {
      sl = minmod(a1, b1);
      sr = minmod(b1, b2);
      word &= ~(0xffull << (8 * n));
      word |= (unsigned long long)lim_byte(sr, b1, b2) << (8 * n);
    }
    sl = fm2 ? sl : T(0);
    sr = fp1 ? sr : T(0);
    {
#pragma clang fp contract(fast)
      qL[n] = qm1[n] + T(0.5) * sl;
      qR[n] = qp0[n] - T(0.5) * sr;
    }
  }
  *rec = word;
}

// one Roe face whose face vector has the single non-zero component S in
// direction DIR (axis-aligned blocks); e central part, d upwind part (only
// when diss).  |S| and S/|S| are formed as fabs(S) and sign(S), which equal
// sqrt(S*S) and S/sqrt(S*S) exactly in IEEE arithmetic.
template <int DIR, typename T>
__device__ inline void roe_face(T gamma, T g1, T eps, T cdis, bool diss, const T *qL, const T *qR, T S, T *e,
                                T *d) {
#pragma clang fp contract(fast)
  T FL[5], FR[5], rrL, rrR;
  T HL = roe_state_flux<DIR>(g1, qL, S, FL, rrL);
  T HR = roe_state_flux<DIR>(g1, qR, S, FR, rrR);
#pragma unroll
  for (int n = 0; n < 5; n++) e[n] = T(0.5) * (FL[n] + FR[n]);
  if (!diss) return;
  // Roe weights from t = sqrt(rho_R/rho_L): w_L = 1/(1+t), w_R = t w_L and
  // the averaged density rho_L t (one square root per face)
  T t = sqrt_pos(qR[0] * rrL);
  T wL = rcp_pos(T(1) + t), wR = t * wL;
  T rh = qL[0] * t;
  T uh[3];
#pragma unroll
  for (int m = 0; m < 3; m++) uh[m] = wL * qL[1 + m] + wR * qR[1 + m];
  T Hh = wL * HL + wR * HR;
  T q2h = uh[0] * uh[0] + uh[1] * uh[1] + uh[2] * uh[2];
  T c2 = (gamma - T(1)) * (Hh - T(0.5) * q2h);
  if (c2 < T(1.0e-12)) c2 = T(1.0e-12);
  T c = sqrt_pos(c2);
  T rc2 = rcp_pos(c2);
  T A = fabs(S);
  T nd = S < T(0) ? T(-1) : T(1);
  T un = uh[DIR] * nd;
  T l1 = (un - c) * A, l2 = un * A, l3 = (un + c) * A;
  T del = eps * (fabs(un) + c) * A;
  T hd = T(0.5) * rcp_pos(del);
  T a1 = fix_abs(l1, del, hd), a2 = fix_abs(l2, del, hd), a3 = fix_abs(l3, del, hd);
  T dr = qR[0] - qL[0], dp = qR[4] - qL[4];
  T du[3];
#pragma unroll
  for (int m = 0; m < 3; m++) du[m] = qR[1 + m] - qL[1 + m];
  T dun = du[DIR] * nd;
  T hc = T(0.5) * rc2;
  T al1 = (dp - rh * c * dun) * hc;
  T al3 = (dp + rh * c * dun) * hc;
  T al2 = dr - dp * rc2;
  T dut[3];
#pragma unroll
  for (int m = 0; m < 3; m++) dut[m] = du[m];
  dut[DIR] = du[DIR] - dun * nd;
  T udt = uh[0] * dut[0] + uh[1] * dut[1] + uh[2] * dut[2];
  // |A| dU = sum of the three wave families, grouped: with b_k = a_k al_k,
  // s = b1 + b2 + b3, the acoustic difference c (b3 - b1) and g = a2 rh
  const T b1 = a1 * al1, b2 = a2 * al2, b3 = a3 * al3;
  const T s13 = b1 + b3, sall = s13 + b2, dl = c * (b3 - b1), g = a2 * rh;
  T dd[5];
  dd[0] = sall;
#pragma unroll
  for (int m = 0; m < 3; m++) dd[1 + m] = sall * uh[m] + g * dut[m];
  dd[1 + DIR] = dd[1 + DIR] + dl * nd;
  dd[4] = s13 * Hh + dl * un + b2 * (T(0.5) * q2h) + g * udt;
#pragma unroll
  for (int n = 0; n < 5; n++) d[n] = cdis * (T(0.5) * dd[n]);
}

// the same face for a general face vector S (FP64 wall faces)
__device__ inline void roe_face_vec(double gamma, double eps, double cdis, const double *qL, const double *qR,
                                    const double *S, double *e, double *d) {
  const double g1 = gamma / (gamma - 1.0);
  double FL[5], FR[5];
  double HL = roe_state_flux_vec(g1, qL, S, FL);
  double HR = roe_state_flux_vec(g1, qR, S, FR);
  for (int n = 0; n < 5; n++) e[n] = 0.5 * (FL[n] + FR[n]);
  double sL = sqrt(qL[0]), sR = sqrt(qR[0]);
  double rd = 1.0 / (sL + sR);
  double wL = sL * rd, wR = sR * rd;
  double rh = sL * sR;
  double uh[3];
  for (int m = 0; m < 3; m++) uh[m] = wL * qL[1 + m] + wR * qR[1 + m];
  double Hh = wL * HL + wR * HR;
  double q2h = uh[0] * uh[0] + uh[1] * uh[1] + uh[2] * uh[2];
  double c2 = (gamma - 1.0) * (Hh - 0.5 * q2h);
  if (c2 < 1.0e-12) c2 = 1.0e-12;
  double c = sqrt(c2);
  double rc2 = 1.0 / c2;
  double A = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  double nv[3] = {S[0] / A, S[1] / A, S[2] / A};
  double un = uh[0] * nv[0] + uh[1] * nv[1] + uh[2] * nv[2];
  double l1 = (un - c) * A, l2 = un * A, l3 = (un + c) * A;
  double del = eps * (fabs(un) + c) * A;
  double hd = 0.5 / del;
  double a1 = fix_abs(l1, del, hd), a2 = fix_abs(l2, del, hd), a3 = fix_abs(l3, del, hd);
  double dr = qR[0] - qL[0], dp = qR[4] - qL[4];
  double du[3] = {qR[1] - qL[1], qR[2] - qL[2], qR[3] - qL[3]};
  double dun = du[0] * nv[0] + du[1] * nv[1] + du[2] * nv[2];
  double hc = 0.5 * rc2;
  double al1 = (dp - rh * c * dun) * hc;
  double al3 = (dp + rh * c * dun) * hc;
  double al2 = dr - dp * rc2;
  double dut[3];
  for (int m = 0; m < 3; m++) dut[m] = du[m] - dun * nv[m];
  double udt = uh[0] * dut[0] + uh[1] * dut[1] + uh[2] * dut[2];
  double dd[5];
  dd[0] = a1 * al1 + a2 * al2 + a3 * al3;
  for (int m = 0; m < 3; m++)
    dd[1 + m] = a1 * al1 * (uh[m] - c * nv[m]) + a2 * (al2 * uh[m] + rh * dut[m]) + a3 * al3 * (uh[m] + c * nv[m]);
  dd[4] = a1 * al1 * (Hh - c * un) + a2 * (al2 * 0.5 * q2h + rh * udt) + a3 * al3 * (Hh + c * un);
  for (int n = 0; n < 5; n++) d[n] = cdis * (0.5 * dd[n]);
}

// wall face of record r: the signed contributions (ve, vd) to the cut cell,
// in FP64 from the bulk state and the mirror state
template <typename T>
__device__ __attribute__((always_inline)) inline void roe_wall_rec(const FluxCtx<T> &c, const DevRec &r, const double *wsol, double psol, double gamma,
                             double eps, double cdis, double *ve, double *vd, int *err) {
  const Dims &d = c.d;
  int zz = r.zz;
  bool hi = r.faceid >= 4;
  long qc = ix(d, r.c[0], r.c[1], r.c[2]);
  if (fabs(full(c.inf, 0, c.s.w[qc])) < (double)1.0e-8f) raise_err(err, MGE_ERR_WALL_RHO);
  int fi = r.c[0], fj = r.c[1], fk = r.c[2];
  if (hi) {
    if (zz == 1) fi++;
    else if (zz == 2) fj++;
    else fk++;
  }
  double qcv[5], qm[5], S[3], e[5], dv[5];
  double rc = full(c.inf, 0, c.s.w[qc]);
  double rr = 1.0 / rc;
  qcv[0] = rc;
  qcv[1] = full(c.inf, 1, c.s.w[d.E + qc]) * rr;
  qcv[2] = full(c.inf, 2, c.s.w[2 * d.E + qc]) * rr;
  qcv[3] = full(c.inf, 3, c.s.w[3 * d.E + qc]) * rr;
  qcv[4] = full(c.inf, 5, c.s.p[qc]);
  double rm = 1.0 / wsol[0];
  qm[0] = wsol[0];
  qm[1] = wsol[1] * rm;
  qm[2] = wsol[2] * rm;
  qm[3] = wsol[3] * rm;
  qm[4] = psol;
  axis_vec(S, zz, (double)face_s_any(c.g, zz, fi, fj, fk));
  if (hi) roe_face_vec(gamma, eps, cdis, qcv, qm, S, e, dv);
  else roe_face_vec(gamma, eps, cdis, qm, qcv, S, e, dv);
  double sgn = hi ? -1.0 : 1.0;
  for (int n = 0; n < 5; n++) {
    ve[n] = sgn * e[n];
    vd[n] = sgn * dv[n];
  }
}

// ------------------------------------------------------------------ k_stage_roe
//
// Per k plane, the workgroup's TI x TJ tile plus a two-cell halo is
// converted to primitive variables in LDS; every i and j face of the tile
// is evaluated once (each cell its low face, plus the faces at the tile's
// high edges) into LDS; k faces are carried from plane to plane in
// registers with the primitives of the cell column k-1..k+2.  By default
// FP64 uses an 8-row tile (512 threads, 121 KB LDS, one workgroup per CU:
// the two waves per SIMD its registers allow anyway, with half the halo per
// cell) and FP32 4 rows (four waves per SIMD fit its registers only at 256
// threads per workgroup); both row counts are built for both precisions and
// the autotuner may pick the other one per level shape.
template <typename T>
struct RoeTile {
  static constexpr int TJ = kDelta<T> ? 4 : 8;
  static constexpr int MINW = 2;   // waves per SIMD the launch bounds ask for
};

template <typename T, int TJ_, bool LIM>
struct RoeSmem {
  T q[5][TJ_ + 4][TI + 4];
  uint8_t fl[TJ_ + 4][TI + 4];
  // limiter words of i and j (frozen limiter; one placeholder word without)
  unsigned long long lw[LIM ? 2 : 1][LIM ? TJ_ + 4 : 1][LIM ? TI + 4 : 1];
  T xe[5][TJ_][TI + 1], xd[5][TJ_][TI + 1];
  T ye[5][TJ_ + 1][TI], yd[5][TJ_ + 1][TI];
};

template <typename T>
__device__ inline void roe_load(const FluxCtx<T> &c, const uint8_t *flags, long q, T *o, bool *f) {
  prim_of(c.s, c.inf, c.d.E, q, o);
  *f = (flags[q] & F_FLUID) != 0;
}

// the face between two cells of the LDS tile at (row, col-1) | (row, col)
// for DIR 0, or (row-1, col) | (row, col) for DIR 1
// lm: limiter mode (0 none, 1 record into *rec, 2 frozen); a face that is
// not evaluated records "not recorded" (its slopes are never used: a cell
// next to a solid one takes no slope toward it)
template <int DIR, typename T, int TJ_, bool LIM>
__device__ __attribute__((always_inline)) inline void roe_lds_face(const RoeSmem<T, TJ_, LIM> &sm, const StageArgs<T> &A, int row, int col, T S, T *e,
                                    T *dd, int lm = 0, unsigned long long *rec = nullptr) {
  const bool diss = A.has_diss;
  bool f[4];
  T v[4][5];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    int r = DIR == 1 ? row + t - 2 : row, cc = DIR == 0 ? col + t - 2 : col;
    f[t] = sm.fl[r][cc] != 0;
#pragma unroll
    for (int n = 0; n < 5; n++) v[t][n] = sm.q[n][r][cc];
  }
  if (rec) *rec = ~0ull;
  if (!(f[1] && f[2])) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      e[n] = T(0);
      dd[n] = T(0);
    }
    return;
  }
  T qL[5], qR[5];
  if (!LIM || lm == 0) {
    muscl(A.order, v[0], v[1], v[2], v[3], f[0], f[3], qL, qR);
  } else {
    const unsigned long long wl = lm == 2 ? sm.lw[DIR][DIR == 1 ? row - 1 : row][DIR == 0 ? col - 1 : col] : 0ull;
    const unsigned long long wr = lm == 2 ? sm.lw[DIR][row][col] : 0ull;
    unsigned long long word;
    muscl_lim(lm, v[0], v[1], v[2], v[3], f[0], f[3], wl, wr, qL, qR, &word);
    if (rec) *rec = word;
  }
  roe_face<DIR>(A.fx.gamma, A.g1, A.eps, A.fx.cdis, diss, qL, qR, S, e, dd);
  if (!diss)
#pragma unroll
    for (int n = 0; n < 5; n++) dd[n] = T(0);
}

template <typename T>
__device__ __attribute__((always_inline)) inline void roe_z_face(const StageArgs<T> &A, T (*zq)[5], const bool *zf, T S, T *e, T *dd,
                                                               int lm = 0, const unsigned long long *zw = nullptr,
                                                               unsigned long long *rec = nullptr) {
  const bool diss = A.has_diss;
  if (rec) *rec = ~0ull;
  if (!(zf[1] && zf[2])) {
#pragma unroll
    for (int n = 0; n < 5; n++) {
      e[n] = T(0);
      dd[n] = T(0);
    }
    return;
  }
  T qL[5], qR[5];
  if (lm == 0) {
    muscl(A.order, zq[0], zq[1], zq[2], zq[3], zf[0], zf[3], qL, qR);
  } else {
    unsigned long long word;
    muscl_lim(lm, zq[0], zq[1], zq[2], zq[3], zf[0], zf[3], lm == 2 ? zw[1] : 0ull, lm == 2 ? zw[2] : 0ull, qL, qR,
              &word);
    if (rec) *rec = word;
  }
  roe_face<2>(A.fx.gamma, A.g1, A.eps, A.fx.cdis, diss, qL, qR, S, e, dd);
  if (!diss)
#pragma unroll
    for (int n = 0; n < 5; n++) dd[n] = T(0);
}

// LIM: built with the limiter-freezing paths (A.lim set), else without;
// GEN: finish_cell's general form (see there)
template <typename T, int TJr, bool LIM, bool GEN>
__device__ __attribute__((always_inline)) inline void k_stage_roe_body(const StageArgs<T> &A, int bx, int by, int bz) {
  __shared__ RoeSmem<T, TJr, LIM> sm;
  const FluxCtx<T> &c = A.fx;
  const Dims &d = c.d;
  const long E = d.E;
  const int kc = A.kc;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TI * ty;
  const int i0 = 2 + bx * TI, j0 = 2 + by * TJr;
  const int kb0 = 2 + bz * kc;
  const int kb1 = min(kb0 + kc - 1, d.kl);
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = i <= d.il && j <= d.jl;
  const bool inarr = i <= d.ib && j <= d.jb;
  static_assert(TJr >= 3, "the high-edge faces are spread over waves 1 and 2");
  const bool diss = A.has_diss;
  const int lm = LIM ? (*A.lfrozen ? 2 : 1) : 0;   // limiter: none, record, frozen
  const T Sz = face_s<2>(c.g, min(i, d.ib), min(j, d.jb), 0);  // clamped: read only for valid cells
  // halo ring items of this thread (t = tid, tid + TI*TJr, ...): their LDS
  // slot and cell do not change with k.  FP64 loads the next plane's items
  // during the current plane's faces (two waves per SIMD hide little
  // latency); FP32 loads them in place
  constexpr bool PRE = !kDelta<T>;
  constexpr int NH = 4 * (TI + 4) + 4 * TJr;
  constexpr int NS = (NH + TI * TJr - 1) / (TI * TJr);   // halo items per thread (2 for 4 rows, 1 for 8)
  int hr[NS] = {}, hc[NS] = {};
  long hq[NS] = {};
  bool hok[NS] = {};
  T hraw[NS][5] = {};
  uint8_t hfl[NS] = {};
  unsigned long long hlw[NS][2] = {};
  if constexpr (PRE) {
#pragma unroll
    for (int h = 0; h < NS; h++) {
      const int t = tid + h * TI * TJr;
      int r, col;
      if (t < 4 * (TI + 4)) {
        int rr = t / (TI + 4);
        col = t % (TI + 4);
        r = rr < 2 ? rr : TJr + rr;
      } else {
        int u = t - 4 * (TI + 4);
        int cc = u / TJr;
        r = 2 + u % TJr;
        col = cc < 2 ? cc : TI + cc;
      }
      const int ii = i0 - 2 + col, jj = j0 - 2 + r;
      hr[h] = r;
      hc[h] = col;
      hok[h] = t < NH && ii <= d.ib && jj <= d.jb;
      hq[h] = hok[h] ? ix(d, ii, jj, 0) : 0;
    }
#pragma unroll
    for (int h = 0; h < NS; h++)
      if (hok[h]) {
        const long q = hq[h] + (long)kb0 * d.sz;
#pragma unroll
        for (int n = 0; n < 4; n++) hraw[h][n] = c.s.w[n * E + q];
        hraw[h][4] = c.s.p[q];
        hfl[h] = A.flags[q];
        if (lm == 2) {
          hlw[h][0] = A.lim[q];
          hlw[h][1] = A.lim[E + q];
        }
      }
  }
  // column carry: primitives of k-1..k+2 and the bottom face of plane k;
  // with the limiter frozen also the cells' limiter words, and while
  // recording the k codes of cell k (from its bottom face)
  T zq[4][5] = {}, Ez[5], Dz[5];
  bool zf[4] = {false, false, false, false};
  unsigned long long zw[4] = {~0ull, ~0ull, ~0ull, ~0ull}, zbits = ~0ull;
  if (valid) {
    const long bz = ix(d, i, j, 0);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const long q = bz + (long)(kb0 - 2 + t) * d.sz;
      roe_load(c, A.flags, q, zq[t], &zf[t]);
      if (lm == 2) zw[t] = A.lim[2 * E + q];
    }
    roe_z_face(A, zq, zf, Sz, Ez, Dz, lm, zw, &zbits);
#pragma unroll
    for (int t = 0; t < 3; t++) {
#pragma unroll
      for (int n = 0; n < 5; n++) zq[t][n] = zq[t + 1][n];
      zf[t] = zf[t + 1];
      zw[t] = zw[t + 1];
    }
    roe_load(c, A.flags, bz + (long)(kb0 + 2) * d.sz, zq[3], &zf[3]);
    if (lm == 2) zw[3] = A.lim[2 * E + bz + (long)(kb0 + 2) * d.sz];
  }
  for (int k = kb0; k <= kb1; k++) {
    // ---- FP64: the finish inputs of this plane's cell, loaded first so
    // that their latency overlaps the face evaluation (two waves per SIMD
    // hide little); FP32 runs four waves per SIMD and keeps its registers
    constexpr bool PRE = !kDelta<T>;
    const long qk = ix(d, i, j, k);
    T pfw[5] = {}, pwn[5] = {}, pvd = T(0);
    uint8_t pfl = 0;
    if (PRE && valid) {
      if (A.fw_in || A.m > 1) {
#pragma unroll
        for (int n = 0; n < 5; n++) pfw[n] = A.fw[n * E + qk];
      } else {
#pragma unroll
        for (int n = 0; n < 5; n++) pfw[n] = T(0);
      }
      if (!A.lres && !A.smoothed) {
#pragma unroll
        for (int n = 0; n < 5; n++) pwn[n] = A.wn[n * E + qk];
      } else {
#pragma unroll
        for (int n = 0; n < 5; n++) pwn[n] = T(0);
      }
      pvd = A.vdtim[qk];
      pfl = A.flags[qk];
    }
    // ---- primitives of the plane: own cell, then the halo ring
    if (inarr) {
      T o[5];
      bool f;
      if (valid) {
#pragma unroll
        for (int n = 0; n < 5; n++) o[n] = zq[1][n];
        f = zf[1];
      } else {
        roe_load(c, A.flags, ix(d, i, j, k), o, &f);
      }
#pragma unroll
      for (int n = 0; n < 5; n++) sm.q[n][ty + 2][tx + 2] = o[n];
      sm.fl[ty + 2][tx + 2] = f;
      if constexpr (LIM) if (lm == 2) {
        const long q = ix(d, i, j, k);
        sm.lw[0][ty + 2][tx + 2] = A.lim[q];
        sm.lw[1][ty + 2][tx + 2] = A.lim[E + q];
      }
    }
    if constexpr (PRE) {
#pragma unroll
      for (int h = 0; h < NS; h++)
        if (hok[h]) {
          T o[5];
          prim_raw(hraw[h], c.inf, o);
#pragma unroll
          for (int n = 0; n < 5; n++) sm.q[n][hr[h]][hc[h]] = o[n];
          sm.fl[hr[h]][hc[h]] = (hfl[h] & F_FLUID) != 0;
          if constexpr (LIM) if (lm == 2) {
            sm.lw[0][hr[h]][hc[h]] = hlw[h][0];
            sm.lw[1][hr[h]][hc[h]] = hlw[h][1];
          }
        }
    } else {
      for (int t = tid; t < NH; t += TI * TJr) {
        int r, col;
        if (t < 4 * (TI + 4)) {
          int rr = t / (TI + 4);
          col = t % (TI + 4);
          r = rr < 2 ? rr : TJr + rr;
        } else {
          int u = t - 4 * (TI + 4);
          int cc = u / TJr;
          r = 2 + u % TJr;
          col = cc < 2 ? cc : TI + cc;
        }
        int ii = i0 - 2 + col, jj = j0 - 2 + r;
        if (ii <= d.ib && jj <= d.jb) {
          T o[5];
          bool f;
          roe_load(c, A.flags, ix(d, ii, jj, k), o, &f);
#pragma unroll
          for (int n = 0; n < 5; n++) sm.q[n][r][col] = o[n];
          sm.fl[r][col] = f;
          if constexpr (LIM) if (lm == 2) {
            const long q = ix(d, ii, jj, k);
            sm.lw[0][r][col] = A.lim[q];
            sm.lw[1][r][col] = A.lim[E + q];
          }
        }
      }
    }
    __syncthreads();
    // FP64: the next plane's halo items, in flight during this plane's faces
    if constexpr (PRE) if (k < kb1) {
#pragma unroll
      for (int h = 0; h < NS; h++)
        if (hok[h]) {
          const long q = hq[h] + (long)(k + 1) * d.sz;
#pragma unroll
          for (int n = 0; n < 4; n++) hraw[h][n] = c.s.w[n * E + q];
          hraw[h][4] = c.s.p[q];
          hfl[h] = A.flags[q];
          if (lm == 2) {
            hlw[h][0] = A.lim[q];
            hlw[h][1] = A.lim[E + q];
          }
        }
    }
    // ---- faces: own i face and j face, then the tile's high-edge faces
    unsigned long long rx = ~0ull, ry = ~0ull;   // the i and j words of the own cell (recording)
    {
      T e[5], dd[5];
      if (j <= d.jl && i <= d.ie) {
        roe_lds_face<0>(sm, A, ty + 2, tx + 2, face_s<0>(c.g, i, j, k), e, dd, lm, &rx);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          sm.xe[n][ty][tx] = e[n];
          sm.xd[n][ty][tx] = dd[n];
        }
      }
      if (i <= d.il && j <= d.je) {
        roe_lds_face<1>(sm, A, ty + 2, tx + 2, face_s<1>(c.g, i, j, k), e, dd, lm, &ry);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          sm.ye[n][ty][tx] = e[n];
          sm.yd[n][ty][tx] = dd[n];
        }
      }
      // the tile's high-edge faces go to two different waves (x edge:
      // wave 2, y edge: wave 1) so no wave evaluates more than one extra face
      if (tid >= 2 * TI && tid < 2 * TI + TJr) {
        int jj = j0 + tid - 2 * TI, a = i0 + TI;
        if (jj <= d.jl && a <= d.ie) {
          roe_lds_face<0>(sm, A, tid - 2 * TI + 2, TI + 2, face_s<0>(c.g, a, jj, k), e, dd, lm);
#pragma unroll
          for (int n = 0; n < 5; n++) {
            sm.xe[n][tid - 2 * TI][TI] = e[n];
            sm.xd[n][tid - 2 * TI][TI] = dd[n];
          }
        }
      } else if (tid >= TI && tid < 2 * TI) {
        int cidx = tid - TI, ii = i0 + cidx, a = j0 + TJr;
        if (ii <= d.il && a <= d.je) {
          roe_lds_face<1>(sm, A, TJr + 2, cidx + 2, face_s<1>(c.g, ii, a, k), e, dd, lm);
#pragma unroll
          for (int n = 0; n < 5; n++) {
            sm.ye[n][TJr][cidx] = e[n];
            sm.yd[n][TJr][cidx] = dd[n];
          }
        }
      }
    }
    T Ez1[5], Dz1[5];
    unsigned long long zn = ~0ull;   // the k word of cell k+1 (recording)
    roe_z_face(A, zq, zf, Sz, Ez1, Dz1, lm, zw, &zn);  // a lane outside the block computes from zeros and stores nothing
    if (lm == 1 && valid) {
      A.lim[qk] = rx;
      A.lim[E + qk] = ry;
      A.lim[2 * E + qk] = zbits;
    }
    zbits = zn;
    __syncthreads();
    // ---- accumulate and finish
    if (valid) {
      const long q = qk;
      T f[5], e[5];
      if (A.fw_in) {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = PRE ? pfw[n] : A.fw[n * E + q];
      } else if (A.m > 1) {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = (PRE ? pfw[n] : A.fw[n * E + q]) * A.blend;
      } else {
#pragma unroll
        for (int n = 0; n < 5; n++) f[n] = T(0);
      }
      if (diss) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          f[n] = f[n] + (sm.xd[n][ty][tx] - sm.xd[n][ty][tx + 1]);
          f[n] = f[n] + (sm.yd[n][ty][tx] - sm.yd[n][ty + 1][tx]);
          f[n] = f[n] + (Dz[n] - Dz1[n]);
        }
      }
#pragma unroll
      for (int n = 0; n < 5; n++) {
        e[n] = sm.xe[n][ty][tx] - sm.xe[n][ty][tx + 1];
        e[n] = e[n] + (sm.ye[n][ty][tx] - sm.ye[n][ty + 1][tx]);
        e[n] = e[n] + (Ez[n] - Ez1[n]);
      }
      if ((PRE ? pfl : A.flags[q]) & F_TOUCHED) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          A.fw[n * E + q] = f[n];
          A.dw[n * E + q] = e[n];
        }
      } else {
        finish_cell<T, PRE, GEN>(A, q, f, e, pwn, pvd, pfl);
      }
      // advance the column carry
#pragma unroll
      for (int n = 0; n < 5; n++) {
        Ez[n] = Ez1[n];
        Dz[n] = Dz1[n];
      }
      if (k < kb1) {
#pragma unroll
        for (int t = 0; t < 3; t++) {
#pragma unroll
          for (int n = 0; n < 5; n++) zq[t][n] = zq[t + 1][n];
          zf[t] = zf[t + 1];
          zw[t] = zw[t + 1];
        }
        roe_load(c, A.flags, ix(d, i, j, k + 3), zq[3], &zf[3]);
        if (lm == 2) zw[3] = A.lim[2 * E + ix(d, i, j, k + 3)];
      }
    }
  }
}
// entry points as MGE_ENTRY_T_XCD makes them (XCD-contiguous block order in
// the single-case entry), with launch bounds that follow the tile's rows
template <typename T, int TJ_, bool LIM, bool GEN>
__global__ void __launch_bounds__(TI * TJ_, RoeTile<T>::MINW) k_stage_roe(StageArgs<T> A) {
  int bx, by, bz;
  xcd_block(bx, by, bz);
  k_stage_roe_body<T, TJ_, LIM, GEN>(A, bx, by, bz);
}
template <typename T, int TJ_, bool LIM, bool GEN>
__global__ void __launch_bounds__(TI * TJ_, RoeTile<T>::MINW)
    k_stage_roe_b(const StageArgs<T> *As, const int4 *cg, int nc, const int *act) {
  const int b = (int)blockIdx.x;
  const int c = batch_case(cg, nc, b);
  const int4 g = cg[c];
  if (!act[g.w]) return;
  const int l = b - g.x;
  k_stage_roe_body<T, TJ_, LIM, GEN>(As[c], l % g.y, (l / g.y) % g.z, l / (g.y * g.z));
}

}  // namespace mge
