// mge_bc.h -- local time step, residual norm, boundary conditions, level
// transfers and the wall force, templated on the bulk type T.  The
// residual norm, the global time step, the cut-cell mirror state and the
// force integral are evaluated in FP64 whatever T is.
#pragma once
#include "mge_stage.h"

namespace mge {

// ------------------------------------------------------------------ timestep

template <typename T>
struct TsArgs {
  Dims d;
  StateC<T> s;
  Geo<T> g;
  const uint8_t *flags;
  T *radsum, *vdtim;
  unsigned long long *dtim_bits;
  T gamma, cfl, cellvolume;
  const T *vol;  // stretched level: per-cell volumes (else nullptr)
  int eqntype;
  T viscl, prlmean;
  T cfac;  // viscous factor of the scheme (timestep.f90:119-124)
  int *err;
  Inf<T> inf;
  int execmode;  // 1: the pseudo step is capped by dtime
  T dtime;       // 2*dtstep/(3*cfl)
  T *eps[3];     // psmoo's per-cell coefficients (nullptr: not formed)
  T smoo, cflrat;  // smo and sqrt(1 + 4 smo)
  int want_dt;     // form the global minimum (dtim_bits)
};

// one coefficient of timestep.f90:286-309: the ratio r of the other two
// radii to this one, e = max(0, min(smo, ((sqrt(1+4 smo)/(1+r/16))^2 - 1)/4))
template <typename T>
__device__ __attribute__((always_inline)) inline T psmoo_coef(T r, T smoo, T cflrat) {
  T t = cflrat / (T(1) + T(0.0625) * r);
  T ex = T(0.25) * (t * t - T(1));
  ex = smoo < ex ? smoo : ex;
  return T(0) > ex ? T(0) : ex;
}

// timestep.f90:36-115 (Euler) and :117-209 (NS): spectral radii, local
// time step with the adaptive factor, global minimum; blank cells zeroed
// (:212-223).  The ghost copies of the radii (:227-282) are realised by
// index clamping where the radii are read.
template <typename T>
__device__ __attribute__((always_inline)) inline void k_timestep_body(const TsArgs<T> &A, int bx, int by, int bz) {
  const Dims &d = A.d;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int j = 2 + by * blockDim.y + threadIdx.y;
  int k = 2 + bz;
  const T one = T(1), two = T(2);
  T mine = T(10);
  if (i <= d.il && j <= d.jl && k <= d.kl) {
    const long E = d.E;
    long q = ix(d, i, j, k);
    T rho = A.s.w[q], pq = A.s.p[q], m1 = A.s.w[E + q], m2 = A.s.w[2 * E + q], m3 = A.s.w[3 * E + q];
    if constexpr (kDelta<T>) {
      rho = rho + A.inf.w[0];
      pq = pq + A.inf.p;
      m1 = m1 + A.inf.w[1];
      m2 = m2 + A.inf.w[2];
      m3 = m3 + A.inf.w[3];
    }
    T cc = A.gamma * pq / rho;
    if (cc < T(0) || cc != cc) raise_err(A.err, MGE_ERR_NEG_C2);
    cc = sqrt(maxd(cc, T(0)));
    T uu = m1 / rho;
    T uv = m2 / rho;
    T uw = m3 / rho;
    T dtv = T(0);
    const T vol = A.vol ? A.vol[q] : A.cellvolume;
    if (A.eqntype == 1) {
      T fmue = A.viscl / A.prlmean;
      T f1 = T(4) / rho / T(3);
      T f2 = A.gamma / rho;
      T fac = maxd(f1, f2);
      dtv = fac * fmue / vol;
    }
    T sx = face_s<0>(A.g, i, j, k);
    T sy = face_s<1>(A.g, i, j, k);
    T sz = face_s<2>(A.g, i, j, k);
    T qs, ds, cs;
    qs = fabs(uu * sx);
    ds = sx * sx;
    cs = cc * sqrt(ds);
    T ri = qs + cs, rvi = dtv * ds;
    qs = fabs(uv * sy);
    ds = sy * sy;
    cs = cc * sqrt(ds);
    T rj = qs + cs, rvj = dtv * ds;
    qs = fabs(uw * sz);
    ds = sz * sz;
    cs = cc * sqrt(ds);
    T rk = qs + cs, rvk = dtv * ds;
    T fict;
    if (A.eqntype == 1) fict = vol / (ri + rj + rk + A.cfac * (rvi + rvj + rvk));
    else fict = vol / (ri + rj + rk);
    // unsteady: capped by the physical step (timestep.f90:86-99, 180-192)
    if (A.execmode == 1) fict = fict < A.dtime ? fict : A.dtime;
    const T plim = T(0.001);
    const T *P = A.s.p;
    T pc = P[q];
    T pip = P[q + 1], pim = P[q - 1];
    T pjp = P[q + d.sy], pjm = P[q - d.sy];
    T pkp = P[q + d.sz], pkm = P[q - d.sz];
    T four_p = T(0);
    if constexpr (kDelta<T>) four_p = T(4) * A.inf.p;
    T dpi = fabs((pip - two * pc + pim) / (pip + two * pc + pim + four_p + plim));
    T dpj = fabs((pjp - two * pc + pjm) / (pjp + two * pc + pjm + four_p + plim));
    T dpk = fabs((pkp - two * pc + pkm) / (pkp + two * pc + pkm + four_p + plim));
    T cm1 = dimd(A.cfl, one);
    T sw = two * (dpi + dpj + dpk * one);
    T r = one / (one + (cm1 < sw ? cm1 : sw));
    T vd = fict * r;
    mine = vd;
    if (A.flags[q] & F_FLUID) {
      A.radsum[q] = ri + rj + rk;
      A.vdtim[q] = vd;
    } else {
      A.radsum[q] = T(0) + T(0) + T(0);
      A.vdtim[q] = T(0);
    }
    if (A.eps[0]) {
      // a blank cell has zero radii and takes coefficient 0: its row is the
      // identity and its zero residual stays out of its lines
      const bool fl = A.flags[q] & F_FLUID;
      A.eps[0][q] = fl ? psmoo_coef(rj / ri + rk / ri, A.smoo, A.cflrat) : T(0);
      A.eps[1][q] = fl ? psmoo_coef(rk / rj + ri / rj, A.smoo, A.cflrat) : T(0);
      A.eps[2][q] = fl ? psmoo_coef(ri / rk + rj / rk, A.smoo, A.cflrat) : T(0);
    }
  }
  // block minimum of the time step (wavefront shuffles, then the block's
  // waves) -> one atomic per block on the FP64 bit pattern (positive
  // doubles order like their bits), skipped when the block cannot lower the
  // value already there, so the single address sees few atomics
  if (!A.want_dt) return;  // uniform over the launch
  T m = mine;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(m, off, 64);
    m = o < m ? o : m;
  }
  __shared__ T wmin[16];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  if ((tid & 63) == 0) wmin[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    T b = wmin[0];
    for (int w = 1; w < nw; w++) b = wmin[w] < b ? wmin[w] : b;
    unsigned long long bits = (unsigned long long)__double_as_longlong((double)b);
    if (bits < __atomic_load_n(A.dtim_bits, __ATOMIC_RELAXED)) atomicMin(A.dtim_bits, bits);
  }
}
MGE_ENTRY_T(k_timestep, TsArgs, 256, 1)

// global time step: bits of 10.0 before k_timestep, decoded after it
struct DtimArgs {
  unsigned long long *bits;
  double *dtim;
};
__device__ __attribute__((always_inline)) inline void k_dtim_init_body(const DtimArgs &A, int, int, int) {
  if (threadIdx.x == 0) A.bits[0] = 0x4024000000000000ULL;
}
MGE_ENTRY(k_dtim_init, DtimArgs, 64, 1)
__device__ __attribute__((always_inline)) inline void k_dtim_decode_body(const DtimArgs &A, int, int, int) {
  if (threadIdx.x == 0) A.dtim[0] = __longlong_as_double((long long)A.bits[0]);
}
MGE_ENTRY(k_dtim_decode, DtimArgs, 64, 1)

// ------------------------------------------------------------------ residual norm

constexpr int kRtrmsPlanes = 8;  // k planes per thread of k_rtrms

template <typename T>
struct RtArgs {
  Dims d;
  const T *dw;
  const T *vdtim;
  const double *dtim;
  const uint8_t *flags;
  double cfl;
  int vt;
  double *partial;
  int gx, gy;  // grid of the launch (partial index)
};

// euler.f90:431-447: sum of (dw_1/(cfl*dt))^2 over fluid cells in FP64;
// per-block partials in a fixed tree order, then k_rtrms_final adds them
// in order, so the norm is deterministic
template <typename T>
__device__ __attribute__((always_inline)) inline void k_rtrms_body(const RtArgs<T> &R, int bx, int by, int bz) {
  const Dims &d = R.d;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int j = 2 + by * blockDim.y + threadIdx.y;
  double v = 0.0;
  // kRtrmsPlanes planes per thread, summed in k order, so that the final
  // pass reads few block partials
  for (int k = 2 + bz * kRtrmsPlanes; k < 2 + (bz + 1) * kRtrmsPlanes; k++) {
    if (i <= d.il && j <= d.jl && k <= d.kl) {
      long q = ix(d, i, j, k);
      double de = R.cfl * (double)R.vdtim[q];
      if (R.vt == 1) de = R.cfl * R.dtim[0];
      double rt = 0.0;
      if (R.flags[q] & F_FLUID) rt = (double)R.dw[q] / de;
      v = v + rt * rt;
    }
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
    long b = bx + (long)R.gx * (by + (long)R.gy * bz);
    R.partial[b] = sm[0];
  }
}
MGE_ENTRY_T(k_rtrms, RtArgs, 256, 1)

struct RfArgs {
  const double *partial;
  long nb;
  double ncell;
  double *out;
  int *err;
};

__device__ __attribute__((always_inline)) inline void k_rtrms_final_body(const RfArgs &F, int, int, int) {
  const double *partial = F.partial;
  const long nb = F.nb;
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
    double r = sqrt(sm[0] / F.ncell);
    F.out[0] = r;
    if (r != r) raise_err(F.err, MGE_ERR_RES_NAN);
  }
}
MGE_ENTRY(k_rtrms_final, RfArgs, 256, 1)

// ------------------------------------------------------------------ bc_cutcell

template <typename T>
struct CutArgs {
  Dims d;
  StateC<T> s;
  Inf<T> inf;
  const DevRec *recs;
  double *wsol, *psol;
  int nmir, nupd, zero_at;  // records < nupd are updated; zero_at gets zeros
  double gamma;
  int eqntype;
  int *err;
};

// bc_cutcell.f90:40-219 for steady bodies (qb = 0), in FP64
template <typename T>
__device__ __attribute__((always_inline)) inline void k_bc_cut_body(const CutArgs<T> &C, int bx, int, int) {
  int m = bx * blockDim.x + threadIdx.x;
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
  // every cloud value is loaded before the sums use any (the loads are in
  // flight together rather than one round trip each); the sums keep the
  // reference order
  const int np = r.npoint;
  T cw[8][6];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    if (c < np) {
      const long qc = r.cloud[c];
#pragma unroll
      for (int n = 0; n < 5; n++) cw[c][n] = C.s.w[n * E + qc];
      cw[c][5] = C.s.p[qc];
    }
  }
  double vr[5] = {0, 0, 0, 0, 0}, pr = 0.0;
#pragma unroll
  for (int n = 0; n < 5; n++)
#pragma unroll
    for (int c = 0; c < 8; c++)
      if (c < np) vr[n] = vr[n] + r.coef[c] * full(C.inf, n, cw[c][n]);
#pragma unroll
  for (int c = 0; c < 8; c++)
    if (c < np) pr = pr + r.coef[c] * full(C.inf, 5, cw[c][5]);
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
MGE_ENTRY_T(k_bc_cut, CutArgs, 64, 1)

// ------------------------------------------------------------------ bcond

// ---- the six faces of bcond in one launch
//
// bcond.f90:16-39 runs the faces one after another in place, each over its
// whole (n+4)^2 face, so a ghost cell beyond two or three faces (an edge or
// corner of the ghost shell) is written by each of them and keeps the value
// of the last, the highest face number f.  Face f read that cell's sources
// (phy1/phy2 along its own direction) when every other face writing them had
// run, since their writers are the cell's other faces, all numbered below f.
// So the value bcond leaves in a ghost cell is a chain: start from cells no
// face writes, apply the cell's faces in increasing order, each to the
// source cells of the previous.  k_bc_faces gives every ghost cell of the
// shell a thread that evaluates its chain (at most three faces deep),
// reading only cells that no face writes, so the threads are independent.
// The chain argument needs every source to be a cell with fewer writers,
// which fails where a direction has one cell (phy2 = nl - 1 is then a ghost
// cell); such a level runs the six faces one after another inside a single
// workgroup instead.  Far-field branch bytes are recorded where the face is
// the cell's last writer: those are the only positions a frozen solve reads.
template <typename T>
struct FacesArgs {
  Dims d;
  State<T> s;
  int kind[6];        // faces 1..6: -20 far, -50 symmetry, -60 periodic, -80 wall, 0 not written
  int seq;            // 1: faces in order inside one workgroup (a direction with one cell)
  double gamma, rho0, p0, c0, u0, v0, w0;
  double s0;          // freestream entropy rho0^gamma / p0
  Inf<T> inf;
  uint8_t *ffdec[6];  // far field: each face's branch per ghost point, or nullptr
  const int *frozen;
  // bc_cutcell's mirror states in the same launch (blocks nfb..): it runs
  // before bcond, reads only interior cells (the level's clouds are checked
  // on the host) and writes only wsol/psol, so the two are independent
  long nfb;           // blocks of the faces
  CutArgs<T> cut;     // cut.nmir == 0: no records in this launch
};

template <typename T>
struct CellV {
  T w[5];
  T p;
};

// the highest-numbered face that writes cell (x, y, z) (0: none)
template <typename T>
__device__ inline int last_face(const FacesArgs<T> &A, int x, int y, int z) {
  const Dims &d = A.d;
  int f = 0;
  if (x <= 1 && A.kind[0]) f = 1;
  if (y <= 1 && A.kind[1]) f = 2;
  if (z <= 1 && A.kind[2]) f = 3;
  if (x >= d.ie && A.kind[3]) f = 4;
  if (y >= d.je && A.kind[4]) f = 5;
  if (z >= d.ke && A.kind[5]) f = 6;
  return f;
}

__device__ inline double pick3(int dir, double a, double b, double c) { return dir == 0 ? a : (dir == 1 ? b : c); }

// bc_far.f90:26-284: the ghost state of a far-field face from the state at
// phy1, in FP64 on full state values (FP32 stores deviations); components
// are picked by selects rather than indexed by the face direction
template <typename T>
__device__ inline CellV<T> far_state(const FacesArgs<T> &A, int f, CellV<T> v, uint8_t *dec,
                                                        bool rec) {
  CellV<T> o;
  const Inf<T> &fi = A.inf;
  const int dir = (f - 1) % 3;
  const double sg = f <= 3 ? -1.0 : 1.0;
  const double fn0 = dir == 0 ? sg : 0.0, fn1 = dir == 1 ? sg : 0.0, fn2 = dir == 2 ? sg : 0.0;
  const double gamma = A.gamma, gm = 1.0 / (gamma - 1.0);
  const double w0 = full(fi, 0, v.w[0]), w1 = full(fi, 1, v.w[1]), w2 = full(fi, 2, v.w[2]),
               w3 = full(fi, 3, v.w[3]);
  double s = 0.0;
  s = s + w1 * fn0;
  s = s + w2 * fn1;
  s = s + w3 * fn2;
  const double vel_norm = s / w0;
  bool inflow = vel_norm <= 0.0;
  if (dec) {
    if (*A.frozen) inflow = *dec == 0;
    else if (rec) *dec = inflow ? 0 : 1;
  }
  const double fnd = sg, vinfd = pick3(dir, A.u0, A.v0, A.w0);
  double r, v0, v1, v2, pr;
  if (inflow) {
    double rhoe = w0;
    double ue = pick3(dir, w1, w2, w3) / rhoe;
    double pe = full(fi, 5, v.p);
    double ce = sqrt(gamma * pe / rhoe);
    double rpo = ue * fnd + 2.0 * ce / (gamma - 1.0);
    double rmi = vinfd * fnd - 2.0 * A.c0 / (gamma - 1.0);
    double un = 0.5 * (rpo + rmi) * fnd;
    double ci = 0.25 * (gamma - 1.0) * (rpo - rmi);
    v0 = dir == 0 ? un : A.u0;
    v1 = dir == 1 ? un : A.v0;
    v2 = dir == 2 ? un : A.w0;
    double si = A.s0;
    double cc = ci * ci / gamma;
    r = pow(si * cc, gm);
    pr = r * cc;
  } else {
    double rhoi = w0;
    double vv0 = w1 / rhoi, vv1 = w2 / rhoi, vv2 = w3 / rhoi;
    double prei = full(fi, 5, v.p);
    double si = pow(rhoi, gamma) / prei;
    double ci = sqrt(gamma * prei / rhoi);
    double rpo = pick3(dir, vv0, vv1, vv2) * fnd + 2.0 * ci / (gamma - 1.0);
    double rmi = vinfd * fnd - 2.0 * A.c0 / (gamma - 1.0);
    double un = 0.5 * (rpo + rmi) * fnd;
    double ce = 0.25 * (gamma - 1.0) * (rpo - rmi);
    v0 = dir == 0 ? un : vv0;
    v1 = dir == 1 ? un : vv1;
    v2 = dir == 2 ? un : vv2;
    double cc = ce * ce / gamma;
    r = pow(si * cc, gm);
    pr = r * cc;
  }
  o.w[0] = stored(fi, 0, r);
  o.w[1] = stored(fi, 1, r * v0);
  o.w[2] = stored(fi, 2, r * v1);
  o.w[3] = stored(fi, 3, r * v2);
  o.w[4] = stored(fi, 4, pr * gm + 0.5 * r * (v0 * v0 + v1 * v1 + v2 * v2));
  o.p = stored(fi, 5, pr);
  return o;
}

template <typename T, int D>
__device__ __attribute__((always_inline)) inline void face_chain(const FacesArgs<T> &A, int x, int y, int z,
                                                                 bool top, CellV<T> &o);

// face f's value in its ghost cell (x, y, z) (bc_far.f90, bc_symm.f90,
// bc_periodic.f90, bc_wall.f90), its sources taken as face_chain<D> of
// them; `top`: the thread's own cell (it records the far-field branch)
template <typename T, int D>
__device__ __attribute__((always_inline)) inline void face_step(const FacesArgs<T> &A, int f, int x, int y, int z,
                                                                bool top, CellV<T> &o) {
  const Dims &d = A.d;
  const int dir = (f - 1) % 3;
  const bool lo = f <= 3;
  const int nl = dir == 0 ? d.il : (dir == 1 ? d.jl : d.kl);
  const int cd = dir == 0 ? x : (dir == 1 ? y : z);
  const bool l1 = lo ? cd == 1 : cd == nl + 1;  // first ghost layer
  const int kind = A.kind[f - 1];
  // source index along dir: phy1 (far field, the first layer, the wall's
  // pressure) or phy2 (the second layer)
  const int p1 = kind == -60 ? (lo ? nl : 2) : (lo ? 2 : nl);
  const int p2 = kind == -60 ? (lo ? nl - 1 : 3) : (lo ? 3 : nl - 1);
  const int sd = (kind == -20 || l1) ? p1 : p2;
  const int sx = dir == 0 ? sd : x, sy = dir == 1 ? sd : y, sz = dir == 2 ? sd : z;
  face_chain<T, D>(A, sx, sy, sz, false, o);
  if (kind == -20) {
    uint8_t *dec = nullptr;
    if (A.ffdec[f - 1]) {
      const int a = dir == 0 ? y : x, c2 = dir == 2 ? y : z;
      const int amax = dir == 0 ? d.jb : d.ib;
      dec = A.ffdec[f - 1] + a + (long)(amax + 1) * c2;
    }
    o = far_state(A, f, o, dec, top && l1);
    return;
  }
  const Inf<T> &fi = A.inf;
  if (kind == -50 || kind == -80) {
#pragma unroll
    for (int n = 1; n < 4; n++)
      if (kind == -80 || n == 1 + dir) o.w[n] = stored(fi, n, -full(fi, n, o.w[n]));
  }
  if (kind == -80 && !l1) {  // bc_wall: the second ghost layer takes the pressure of phy1
    CellV<T> v;
    face_chain<T, D>(A, dir == 0 ? p1 : x, dir == 1 ? p1 : y, dir == 2 ? p1 : z, false, v);
    o.p = v.p;
  }
}

// the value bcond leaves in cell (x, y, z), with at most D faces writing it
template <typename T, int D>
__device__ __attribute__((always_inline)) inline void face_chain(const FacesArgs<T> &A, int x, int y, int z,
                                                                 bool top, CellV<T> &o) {
  if constexpr (D > 0) {
    const int f = last_face(A, x, y, z);
    if (f != 0) {
      face_step<T, D - 1>(A, f, x, y, z, top, o);
      return;
    }
  }
  const Dims &d = A.d;
  const long q = ix(d, x, y, z), E = d.E;
#pragma unroll
  for (int n = 0; n < 5; n++) o.w[n] = A.s.w[n * E + q];
  o.p = A.s.p[q];
}

template <typename T>
__device__ __attribute__((always_inline)) inline void k_bc_faces_body(const FacesArgs<T> &A, int bx, int, int) {
  const Dims &d = A.d;
  if (A.seq) {  // one workgroup: the faces in order, as bcond.f90
    for (int f = 1; f <= 6; f++) {
      if (A.kind[f - 1] == 0) continue;
      const int dir = (f - 1) % 3;
      const int nl = dir == 0 ? d.il : (dir == 1 ? d.jl : d.kl);
      const int g1 = f <= 3 ? 1 : nl + 1, g2 = f <= 3 ? 0 : nl + 2;  // the two ghost layers
      const int amax = dir == 0 ? d.jb : d.ib, cmax = dir == 2 ? d.jb : d.kb;
      const long npos = (long)(amax + 1) * (cmax + 1);
      for (long m = threadIdx.x; m < npos; m += blockDim.x) {
        const int a = (int)(m % (amax + 1)), c2 = (int)(m / (amax + 1));
        // both layers from the cells as they stand, then both written (a
        // source may be the other layer when the direction has one cell)
        CellV<T> v1, v2;
        int x1, y1, z1, x2, y2, z2;
        if (dir == 0) { x1 = g1; x2 = g2; y1 = y2 = a; z1 = z2 = c2; }
        else if (dir == 1) { y1 = g1; y2 = g2; x1 = x2 = a; z1 = z2 = c2; }
        else { z1 = g1; z2 = g2; x1 = x2 = a; y1 = y2 = c2; }
        face_step<T, 0>(A, f, x1, y1, z1, true, v1);
        face_step<T, 0>(A, f, x2, y2, z2, false, v2);
        const long E = d.E, q1 = ix(d, x1, y1, z1), q2 = ix(d, x2, y2, z2);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          A.s.w[n * E + q2] = v2.w[n];
          A.s.w[n * E + q1] = v1.w[n];
        }
        A.s.p[q2] = v2.p;
        A.s.p[q1] = v1.p;
      }
      __syncthreads();
    }
    return;
  }
  if (bx >= A.nfb) {
    k_bc_cut_body(A.cut, (int)(bx - A.nfb), 0, 0);
    return;
  }
  const long t = (long)bx * blockDim.x + threadIdx.x;
  int x, y, z;
  if (!shell_cell_by_depth(d, t, x, y, z)) return;
  if (last_face(A, x, y, z) == 0) return;
  CellV<T> o;
  face_chain<T, 3>(A, x, y, z, true, o);
  const long q = ix(d, x, y, z), E = d.E;
  for (int n = 0; n < 5; n++) A.s.w[n * E + q] = o.w[n];
  A.s.p[q] = o.p;
}
MGE_ENTRY_T(k_bc_faces, FacesArgs, 256, 1)

// the ordered edge extrapolation of bc_corner.f90 (one target cell)
template <typename T>
__device__ inline void edge_fill(const Dims &d, State<T> s, T gamma, const Inf<T> &f, int ti, int tj, int tk, int ai,
                                 int aj, int ak, int bi, int bj, int bk, int ci, int cj, int ck) {
  const long E = d.E;
  long qt = ix(d, ti, tj, tk), qa = ix(d, ai, aj, ak), qb = ix(d, bi, bj, bk), qc = ix(d, ci, cj, ck);
  T v[5];
  for (int n = 0; n < 5; n++) {
    v[n] = s.w[n * E + qa] + s.w[n * E + qb] - s.w[n * E + qc];
    s.w[n * E + qt] = v[n];
  }
  s.p[qt] = eos_st(gamma, f, v);
}

// bc_corner.f90:23-157.  Every edge cell of the first ghost layer is
// t = a + b - c of three cells that no edge writes (face ghosts and interior
// cells; the reference's third i-edge line reads its own target and adds and
// subtracts the same value, :120): edges along k (:23-61), along j (:63-99)
// and along i (:101-137).  The i-edge cells at i = 1 and i = ie are the
// eight corners, which :141-157 overwrite without reading them; each corner
// reads three edge cells, which its thread evaluates itself with the same
// expression.  So every item is independent: one launch over all of them.
__device__ __attribute__((always_inline)) inline void edge_cells(const Dims &d, int ph, int e, int x, long *q) {
  const int il = d.il, jl = d.jl, kl = d.kl, ie = d.ie, je = d.je, ke = d.ke;
  // the row e of each table as a sum over the unrolled entries (a runtime
  // row index, or a select chain the compiler turns into one, would put the
  // tables in scratch memory)
  auto row = [e](const int (&T)[4][12], int m) {
    return (e == 0) * T[0][m] + (e == 1) * T[1][m] + (e == 2) * T[2][m] + (e == 3) * T[3][m];
  };
  int c[12];
  if (ph == 1) {  // edges along k, x = k
    const int T[4][12] = {{1, 1, 0, 2, 1, 0, 1, 2, 0, 2, 2, 0}, {ie, 1, 0, il, 1, 0, ie, 2, 0, il, 2, 0},
                          {1, je, 0, 2, je, 0, 1, jl, 0, 2, jl, 0}, {ie, je, 0, il, je, 0, ie, jl, 0, il, jl, 0}};
#pragma unroll
    for (int m = 0; m < 12; m++) c[m] = (m % 3 == 2) ? x : row(T, m);
  } else if (ph == 2) {  // edges along j, x = j
    const int T[4][12] = {{1, 0, 1, 2, 0, 1, 1, 0, 2, 2, 0, 2}, {ie, 0, 1, il, 0, 1, ie, 0, 2, il, 0, 2},
                          {1, 0, ke, 2, 0, ke, 1, 0, kl, 2, 0, kl}, {ie, 0, ke, il, 0, ke, ie, 0, kl, il, 0, kl}};
#pragma unroll
    for (int m = 0; m < 12; m++) c[m] = (m % 3 == 1) ? x : row(T, m);
  } else {  // edges along i, x = i
    const int T[4][12] = {{0, 1, 1, 0, 1, 2, 0, 2, 1, 0, 2, 2}, {0, 1, ke, 0, 2, ke, 0, 1, kl, 0, 2, kl},
                          {0, je, 1, 0, je, 1, 0, jl, 2, 0, jl, 2}, {0, je, ke, 0, jl, ke, 0, je, kl, 0, jl, kl}};
#pragma unroll
    for (int m = 0; m < 12; m++) c[m] = (m % 3 == 0) ? x : row(T, m);
  }
#pragma unroll
  for (int m = 0; m < 4; m++) q[m] = ix(d, c[3 * m], c[3 * m + 1], c[3 * m + 2]);
}

template <typename T>
struct CornerArgs {
  Dims d;
  State<T> s;
  T gamma;
  Inf<T> inf;
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_bc_corner_body(const CornerArgs<T> &C, int bx, int, int) {
  const Dims &d = C.d;
  const State<T> s = C.s;
  const long E = d.E;
  T *W = s.w;
  const int il = d.il, jl = d.jl, kl = d.kl, ie = d.ie, je = d.je, ke = d.ke, ib = d.ib;
  const int n1 = 4 * (kl - 1), n2 = 4 * (jl - 1), n3 = 4 * (ib + 1);
  const int t = bx * blockDim.x + threadIdx.x;
  T v[5];
  long qt;
  if (t < n1 + n2 + n3) {
    int ph, u;
    if (t < n1) {
      ph = 1;
      u = t;
    } else if (t < n1 + n2) {
      ph = 2;
      u = t - n1;
    } else {
      ph = 3;
      u = t - n1 - n2;
    }
    const int e = u & 3, x = ph == 3 ? (u >> 2) : 2 + (u >> 2);
    if (ph == 3 && (x == 1 || x == ie)) return;  // corners, overwritten below
    // the i-edge line of :120 reads its own target; the two of them next to
    // a corner are that corner's input and are written by its thread
    if (ph == 3 && e == 2 && (x == 2 || x == il)) return;
    long q[4];
    edge_cells(d, ph, e, x, q);
    for (int n = 0; n < 5; n++) {
      v[n] = W[n * E + q[1]] + W[n * E + q[2]] - W[n * E + q[3]];
      W[n * E + q[0]] = v[n];
    }
    qt = q[0];
  } else if (t < n1 + n2 + n3 + 8) {
    const int c = t - (n1 + n2 + n3);
    const bool hi = c & 1, hj = (c & 2) != 0, hk = (c & 4) != 0;
    const int ci = hi ? ie : 1, cj = hj ? je : 1, ck = hk ? ke : 1;
    const int ai = hi ? il : 2, aj = hj ? jl : 2, ak = hk ? kl : 2;
    long qi[4], qj[4], qk[4];
    edge_cells(d, 3, (hk ? 1 : 0) + (hj ? 2 : 0), ai, qi);  // (ai, cj, ck)
    edge_cells(d, 2, (hi ? 1 : 0) + (hk ? 2 : 0), aj, qj);  // (ci, aj, ck)
    edge_cells(d, 1, (hi ? 1 : 0) + (hj ? 2 : 0), ak, qk);  // (ci, cj, ak)
    const T two = T(2), three = T(3);
    const long qa = ix(d, ai, aj, ak);
    const bool own = hj && !hk;  // the i-edge neighbour is a self-reading line of :120
    T ve[5];
    for (int n = 0; n < 5; n++) {
      T ei = W[n * E + qi[1]] + W[n * E + qi[2]] - W[n * E + qi[3]];
      T ej = W[n * E + qj[1]] + W[n * E + qj[2]] - W[n * E + qj[3]];
      T ek = W[n * E + qk[1]] + W[n * E + qk[2]] - W[n * E + qk[3]];
      v[n] = (ei + ej + ek) * two / three - W[n * E + qa];
      ve[n] = ei;
    }
    if (own) {
      for (int n = 0; n < 5; n++) W[n * E + qi[0]] = ve[n];
      s.p[qi[0]] = eos_st(C.gamma, C.inf, ve);
    }
    qt = ix(d, ci, cj, ck);
    for (int n = 0; n < 5; n++) W[n * E + qt] = v[n];
  } else {
    return;
  }
  s.p[qt] = eos_st(C.gamma, C.inf, v);
}
MGE_ENTRY_T(k_bc_corner, CornerArgs, 256, 1)

// ------------------------------------------------------------------ misc kernels

// ghost shell of src -> dst (w and p), so bcond sees the same ghost values
// that the reference's single in-place array holds at that point
template <typename T>
struct GhostArgs {
  Dims d;
  StateC<T> src;
  State<T> dst;
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_ghost_copy_body(const GhostArgs<T> &G, int bx, int, int) {
  const Dims &d = G.d;
  const StateC<T> &src = G.src;
  const State<T> &dst = G.dst;
  int i, j, k;
  if (!shell_cell(d, bx * (long)blockDim.x + threadIdx.x, i, j, k)) return;
  const long q = ix(d, i, j, k);
  for (int n = 0; n < 5; n++) dst.w[n * d.E + q] = src.w[n * d.E + q];
  dst.p[q] = src.p[q];
}
MGE_ENTRY_T(k_ghost_copy, GhostArgs, 256, 1)

template <typename T>
struct EosArgs {
  Dims d;
  State<T> s;
  T gamma;
  Inf<T> inf;
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_eos_all_body(const EosArgs<T> &A, int bx, int, int) {
  const Dims &d = A.d;
  long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  T v[5];
  for (int n = 0; n < 5; n++) v[n] = A.s.w[n * d.E + q];
  A.s.p[q] = eos_st(A.gamma, A.inf, v);
}
MGE_ENTRY_T(k_eos_all, EosArgs, 256, 1)

template <typename T>
struct FillArgs {
  Dims d;
  State<T> s;
  T v[5];
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_fill_body(const FillArgs<T> &A, int bx, int, int) {
  const Dims &d = A.d;
  long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  for (int n = 0; n < 5; n++) A.s.w[n * d.E + q] = A.v[n];
}
MGE_ENTRY_T(k_fill, FillArgs, 256, 1)

// n values src -> dst (the multigrid saved state ws = w on a first visit)
template <typename T>
struct CopyArgs {
  const T *src;
  T *dst;
  long n;
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_copy_body(const CopyArgs<T> &A, int bx, int, int) {
  long q = bx * (long)blockDim.x + threadIdx.x;
  if (q < A.n) A.dst[q] = A.src[q];
}
MGE_ENTRY_T(k_copy, CopyArgs, 256, 1)

// ------------------------------------------------------------------ transfers

template <typename T>
struct XferArgs {
  Dims dc, df;           // coarse, fine
  int rx, ry, rz;        // fine/coarse ratio per direction
  const T *wf, *dwf;
  T *wc;
  const T *wsc;
  T *wr;
  uint8_t *marker;
  T volc, volf;
  int fmg;               // mgl-1 == mlvl: interpolate the solution itself
  const uint8_t *flagsf;
  T *wfo;                // fine state (prolongation, in place)
  T *corr;               // coarse correction w - ws (5 components, coarse cells)
  // stretched tensor levels: per-cell volumes and the 1-D weight pairs of
  // the restriction (per coarse index) and prolongation (per fine index)
  const T *vc, *vf;
  const T *rw[3], *pw[3];
  // embedded blocks: the fine block's cell 2 lies in coarse cell 2 + off[a]
  // of the coarse block, whose cells 2 .. 1 + nrc[a] it covers (restriction
  // range); a full-coverage level pair has off = 0 and nrc = the coarse
  // interior extent
  int off[3], nrc[3];
  // eos: the launch also sets p from the EOS on every cell of the target
  // level (the coarse level of a restriction, the fine level of a
  // prolongation), as a k_eos_all pass after it would: its own cells from
  // the values it writes (or, untouched, reads), the ghost shell from blocks
  // past the cells' grid (gx x gy x gz0 blocks)
  int eos;
  T *peos;
  T gamma;
  Inf<T> inf;
  int gx, gy, gz0;
};

// the EOS of ghost-shell cell t of a level
template <typename T>
__device__ inline void eos_shell(const Dims &d, const T *w, T *p, T gamma, const Inf<T> &inf, long t) {
  int i, j, k;
  if (!shell_cell(d, t, i, j, k)) return;
  const long q = ix(d, i, j, k);
  T v[5];
#pragma unroll
  for (int n = 0; n < 5; n++) v[n] = w[n * d.E + q];
  p[q] = eos_st(gamma, inf, v);
}

// blocks past the cells' grid: the shell's EOS (true when this block was one)
template <typename T>
__device__ inline bool xfer_shell_block(const XferArgs<T> &X, const Dims &d, const T *w, int bx, int by, int bz) {
  if (!X.eos || bz < X.gz0) return false;
  const long t = (((long)(bz - X.gz0) * X.gy + by) * X.gx + bx) * (blockDim.x * blockDim.y) + threadIdx.x +
                 blockDim.x * threadIdx.y;
  eos_shell(d, w, X.peos, X.gamma, X.inf, t);
  return true;
}

// the coarse-level correction w - ws of every coarse cell, formed once so
// that the prolongation reads one value per coarse neighbour and component
// instead of two (addw.f90:156-173 forms the same difference per use)
template <typename T>
__device__ __attribute__((always_inline)) inline void k_corr_body(const XferArgs<T> &X, int bx, int, int) {
  const long q = bx * (long)blockDim.x + threadIdx.x;
  const long E = X.dc.E;
  if (q >= 5 * E) return;
  X.corr[q] = X.wc[q] - X.wsc[q];
}
MGE_ENTRY_T(k_corr, XferArgs, 256, 1)

template <typename T>
__device__ inline void rstencil(int I, int r, int *s, T *f) {
  if (r == 2) { *s = 2 * I - 2; f[0] = T(0.5); f[1] = T(0.5); }
  else { *s = I; f[0] = T(1); f[1] = T(0); }
}
template <typename T>
__device__ inline void pstencil(int i, int r, int *s, T *f) {
  if (r == 2) {
    if ((i & 1) == 0) { *s = i / 2; f[0] = T(0.25); f[1] = T(0.75); }
    else { *s = (i - 1) / 2 + 1; f[0] = T(0.75); f[1] = T(0.25); }
  } else { *s = i; f[0] = T(1); f[1] = T(0); }
}

// collc.f90:112-143 for aligned ratio-1/2 levels: w and the residual
// restricted with the cell-centre trilinear weights; marker = 1
template <typename T>
__device__ __attribute__((always_inline)) inline void k_restrict_body(const XferArgs<T> &X, int bx, int by, int bz) {
  const Dims &c = X.dc, &f = X.df;
  if (xfer_shell_block(X, c, X.wc, bx, by, bz)) return;
  // (ii, jj, kk): the coarse cell in the fine block's own coarse numbering
  const int ii = 2 + bx * blockDim.x + threadIdx.x;
  const int jj = 2 + by * blockDim.y + threadIdx.y;
  const int kk = 2 + bz;
  if (ii > 1 + X.nrc[0] || jj > 1 + X.nrc[1] || kk > 1 + X.nrc[2]) return;
  const int i = ii + X.off[0], j = jj + X.off[1], k = kk + X.off[2];
  int s[3];
  T fa[2], fb[2], fc[2], fac[8];
  rstencil(ii, X.rx, &s[0], fa);
  rstencil(jj, X.ry, &s[1], fb);
  rstencil(kk, X.rz, &s[2], fc);
  if (X.rw[0]) {
    fa[0] = X.rw[0][2 * i]; fa[1] = X.rw[0][2 * i + 1];
    fb[0] = X.rw[1][2 * j]; fb[1] = X.rw[1][2 * j + 1];
    fc[0] = X.rw[2][2 * k]; fc[1] = X.rw[2][2 * k + 1];
  }
  for (int c8 = 0; c8 < 8; c8++) fac[c8] = fa[c8 & 1] * fb[(c8 >> 1) & 1] * fc[(c8 >> 2) & 1];
  long qc = ix(c, i, j, k);
  long qf[8];
  for (int c8 = 0; c8 < 8; c8++) qf[c8] = ix(f, s[0] + (c8 & 1), s[1] + ((c8 >> 1) & 1), s[2] + ((c8 >> 2) & 1));
  for (int n = 0; n < 5; n++) {
    T acc = T(0);
    for (int c8 = 0; c8 < 8; c8++) {
      T val = X.dwf[n * f.E + qf[c8]] / (X.vf ? X.vf[qf[c8]] : X.volf);
      acc = acc + val * fac[c8];
    }
    X.wr[n * c.E + qc] = acc * (X.vc ? X.vc[qc] : X.volc);
  }
  T v[5];
  for (int n = 0; n < 5; n++) {
    T acc = T(0);
    for (int c8 = 0; c8 < 8; c8++) acc = acc + X.wf[n * f.E + qf[c8]] * fac[c8];
    X.wc[n * c.E + qc] = acc;
    v[n] = acc;
  }
  X.marker[qc] = 1;
  if (X.eos) X.peos[qc] = eos_st(X.gamma, X.inf, v);
}
MGE_ENTRY_T(k_restrict, XferArgs, 256, 1)

// addw.f90:156-173: trilinear prolongation of the correction w-ws (or of
// the solution on the full-multigrid step) to fluid cells
template <typename T>
__device__ __attribute__((always_inline)) inline void k_prolong_body(const XferArgs<T> &X, int bx, int by, int bz) {
  const Dims &c = X.dc, &f = X.df;
  if (xfer_shell_block(X, f, (const T *)X.wfo, bx, by, bz)) return;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int j = 2 + by * blockDim.y + threadIdx.y;
  int k = 2 + bz;
  if (i > f.il || j > f.jl || k > f.kl) return;
  int s[3];
  T fa[2], fb[2], fc[2], fac[8];
  pstencil(i, X.rx, &s[0], fa);
  pstencil(j, X.ry, &s[1], fb);
  pstencil(k, X.rz, &s[2], fc);
  if (X.pw[0]) {
    fa[0] = X.pw[0][2 * i]; fa[1] = X.pw[0][2 * i + 1];
    fb[0] = X.pw[1][2 * j]; fb[1] = X.pw[1][2 * j + 1];
    fc[0] = X.pw[2][2 * k]; fc[1] = X.pw[2][2 * k + 1];
  }
  for (int c8 = 0; c8 < 8; c8++) fac[c8] = fa[c8 & 1] * fb[(c8 >> 1) & 1] * fc[(c8 >> 2) & 1];
  long qf = ix(f, i, j, k);
  T v[5];
  if (!(X.flagsf[qf] & F_FLUID)) {
    if (X.eos) {
#pragma unroll
      for (int n = 0; n < 5; n++) v[n] = X.wfo[n * f.E + qf];
      X.peos[qf] = eos_st(X.gamma, X.inf, v);
    }
    return;
  }
  long qc[8];
  for (int c8 = 0; c8 < 8; c8++)
    qc[c8] = ix(c, X.off[0] + s[0] + (c8 & 1), X.off[1] + s[1] + ((c8 >> 1) & 1), X.off[2] + s[2] + ((c8 >> 2) & 1));
  const T *src = X.fmg ? X.wc : X.corr;
  for (int n = 0; n < 5; n++) {
    T acc = T(0);
    for (int c8 = 0; c8 < 8; c8++) acc = acc + src[n * c.E + qc[c8]] * fac[c8];
    v[n] = X.fmg ? acc : X.wfo[n * f.E + qf] + acc;
    X.wfo[n * f.E + qf] = v[n];
  }
  if (X.eos) X.peos[qf] = eos_st(X.gamma, X.inf, v);
}
MGE_ENTRY_T(k_prolong, XferArgs, 256, 1)

// k_prolong on a ratio-2 uniform level pair without block offsets (the
// single-case path): one thread per coarse cell c forms its eight fine cells
// (2c-2 and 2c-1 per direction) from the 3 x 3 x 3 coarse cells around it,
// each fine value with k_prolong's stencil, weights and summation order (so
// the same bits), and their p when eos.  A component's coarse values over
// the block's 64 x 4 cells, a one-cell halo and the planes ck-1..ck+1 are
// staged in LDS by coalesced row loads (the correction w - ws is formed
// there, so no k_corr pass runs first); the fine cells 2c-2, 2c-1 of a row
// move as one pair (rows, planes and components of a ratio-2 fine level have
// even lengths; the host checks the alignment).  blockDim = (64, 4).
constexpr int kPro2X = 66, kPro2Y = 6, kPro2N = 3 * kPro2Y * kPro2X;

template <typename T>
__device__ __attribute__((always_inline)) inline void k_prolong2_body(const XferArgs<T> &X, int bx, int by, int bz) {
  const Dims &c = X.dc, &f = X.df;
  if (xfer_shell_block(X, f, (const T *)X.wfo, bx, by, bz)) return;  // uniform per block
  __shared__ T cs[3][kPro2Y][kPro2X];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 64 * ty;
  const int ci0 = 2 + bx * 64, cj0 = 2 + by * 4, ck = 2 + bz;
  const int ci = ci0 + tx, cj = cj0 + ty;
  const bool act = ci <= c.il && cj <= c.jl;
  // the fine cell 2c-2 takes 1/4 of c-1 and 3/4 of c, the cell 2c-1 3/4 of c
  // and 1/4 of c+1 (pstencil)
  const T wlo[2] = {T(0.25), T(0.75)}, whi[2] = {T(0.75), T(0.25)};
  // fine rows r = (hj, hk) = (r & 1, r >> 1): the pair starts and flags
  long qr[4];
  unsigned fl[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    qr[r] = ix(f, 2 * ci - 2, 2 * cj - 2 + (r & 1), 2 * ck - 2 + (r >> 1));
    fl[r] = act ? *reinterpret_cast<const uint16_t *>(X.flagsf + qr[r]) : 0u;
  }
#pragma unroll 1
  for (int n = 0; n < 5; n++) {
    // the solution (full multigrid) or the correction w - ws, formed here as
    // k_corr forms it
    const T *wn = X.wc + n * c.E, *sn = X.wsc + n * c.E;
    for (int e = tid; e < kPro2N; e += 256) {
      const int xx = e % kPro2X, rr = e / kPro2X, yy = rr % kPro2Y, zz = rr / kPro2Y;
      const int gi = ci0 - 1 + xx < c.ib ? ci0 - 1 + xx : c.ib;  // past the block: never read
      const int gj = cj0 - 1 + yy < c.jb ? cj0 - 1 + yy : c.jb;
      const long qc = ix(c, gi, gj, ck - 1 + zz);
      (&cs[0][0][0])[e] = X.fmg ? wn[qc] : wn[qc] - sn[qc];
    }
    __syncthreads();
    if (act) {
      T cv[27];
#pragma unroll
      for (int g = 0; g < 27; g++) cv[g] = cs[g / 9][ty + (g / 3) % 3][tx + g % 3];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int hj = r & 1, hk = r >> 1;
        Pair<T> *dp = reinterpret_cast<Pair<T> *>(X.wfo + n * f.E + qr[r]);
        const Pair<T> old = *dp;
        T v[2] = {old.x, old.y};
#pragma unroll
        for (int hi = 0; hi < 2; hi++) {
          if (!((fl[r] >> (8 * hi)) & F_FLUID)) continue;  // blank: keeps its value
          const T *fa = hi ? whi : wlo, *fb = hj ? whi : wlo, *fc = hk ? whi : wlo;
          T acc = T(0);
#pragma unroll
          for (int c8 = 0; c8 < 8; c8++) {
            const T fac = fa[c8 & 1] * fb[(c8 >> 1) & 1] * fc[(c8 >> 2) & 1];
            acc = acc + cv[(hi + (c8 & 1)) + 3 * (hj + ((c8 >> 1) & 1)) + 9 * (hk + ((c8 >> 2) & 1))] * fac;
          }
          v[hi] = X.fmg ? acc : v[hi] + acc;
        }
        *dp = Pair<T>{v[0], v[1]};
      }
    }
    __syncthreads();  // the next component reuses the LDS
  }
  if (!X.eos || !act) return;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    T a[5], b[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      const Pair<T> v = *reinterpret_cast<const Pair<T> *>(X.wfo + n * f.E + qr[r]);
      a[n] = v.x;
      b[n] = v.y;
    }
    *reinterpret_cast<Pair<T> *>(X.peos + qr[r]) = Pair<T>{eos_st(X.gamma, X.inf, a), eos_st(X.gamma, X.inf, b)};
  }
}
MGE_ENTRY_T(k_prolong2, XferArgs, 256, 1)

// bc_embed.f90:50-143 for an aligned ratio-2 embedding: every ghost cell of
// the fine block (outside its cells 2..il, 2..jl, 2..kl) whose centre lies
// inside the coarse block's cells takes the cell-centre trilinear
// interpolation of the coarse state (the prolongation stencil, through
// off[]) where the fine cell is fluid, and p from the EOS of its state
// (depvars); ghost cells whose centre falls outside the coarse block's cells
// are left as they are (eval_x2cell's error return)
template <typename T>
struct EmbedArgs {
  Dims df, dc;
  State<T> sf;       // fine block state (ghosts written)
  StateC<T> sc;      // coarse block state
  const uint8_t *flagsf;
  int off[3];        // as XferArgs::off
  int nb[3];         // the other block's extent for k_overlap (cells 2 .. 1 + nb)
  T gamma;
  Inf<T> inf;
};
template <typename T>
__device__ __attribute__((always_inline)) inline void k_embed_body(const EmbedArgs<T> &X, int bx, int, int) {
  const Dims &f = X.df, &c = X.dc;
  const long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= f.E) return;
  const int i = (int)(q % f.sy), j = (int)((q / f.sy) % (f.jb + 1)), k = (int)(q / f.sz);
  if (i >= 2 && i <= f.il && j >= 2 && j <= f.jl && k >= 2 && k <= f.kl) return;   // a cell of the block
  const int id[3] = {i, j, k};
  const int cl[3] = {c.il, c.jl, c.kl};
  int s[3];
  T fw[3][2];
  for (int a = 0; a < 3; a++) {
    // the coarse cell holding the fine centre: fine cells 2m-2, 2m-1 of the
    // block's own coarse numbering lie in its cell m (fine 0, 1 in 1)
    const int mc = X.off[a] + 2 + (id[a] >= 2 ? (id[a] - 2) / 2 : -1);
    if (mc < 2 || mc > cl[a]) return;
    pstencil(id[a], 2, &s[a], fw[a]);
  }
  T fac[8];
  for (int c8 = 0; c8 < 8; c8++) fac[c8] = fw[0][c8 & 1] * fw[1][(c8 >> 1) & 1] * fw[2][(c8 >> 2) & 1];
  T v[5];
  if (X.flagsf[q] & F_FLUID) {
    for (int n = 0; n < 5; n++) {
      T acc = T(0);
      for (int c8 = 0; c8 < 8; c8++) {
        const long qc = ix(c, X.off[0] + s[0] + (c8 & 1), X.off[1] + s[1] + ((c8 >> 1) & 1),
                           X.off[2] + s[2] + ((c8 >> 2) & 1));
        acc = acc + X.sc.w[n * c.E + qc] * fac[c8];
      }
      v[n] = acc;
      X.sf.w[n * f.E + q] = acc;
    }
  } else {
    for (int n = 0; n < 5; n++) v[n] = X.sf.w[n * f.E + q];
  }
  X.sf.p[q] = eos_st(X.gamma, X.inf, v);
}
MGE_ENTRY_T(k_embed, EmbedArgs, 256, 1)

// bc_overlap.f90:32-90 for blocks of one level on one lattice: every ghost
// cell of the block (df, sf) whose cell lies among the cells 2 .. 1 + nb of
// the neighbour (dc, sc) -- its index there is its own plus off[] -- takes
// the neighbour's state (the trilinear weights of coinciding centres are 1
// and 0) and p from the EOS of it (depvars)
template <typename T>
__device__ __attribute__((always_inline)) inline void k_overlap_body(const EmbedArgs<T> &X, int bx, int, int) {
  const Dims &f = X.df, &c = X.dc;
  const long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= f.E) return;
  const int i = (int)(q % f.sy), j = (int)((q / f.sy) % (f.jb + 1)), k = (int)(q / f.sz);
  if (i >= 2 && i <= f.il && j >= 2 && j <= f.jl && k >= 2 && k <= f.kl) return;
  const int in = i + X.off[0], jn = j + X.off[1], kn = k + X.off[2];
  if (in < 2 || in > 1 + X.nb[0] || jn < 2 || jn > 1 + X.nb[1] || kn < 2 || kn > 1 + X.nb[2]) return;
  const long qn = ix(c, in, jn, kn);
  T v[5];
  for (int n = 0; n < 5; n++) {
    v[n] = X.sc.w[n * c.E + qn];
    X.sf.w[n * f.E + q] = v[n];
  }
  X.sf.p[q] = eos_st(X.gamma, X.inf, v);
}
MGE_ENTRY_T(k_overlap, EmbedArgs, 256, 1)

// wall force: momentum flux of the scheme's wall faces (efluxb.f90:66-68
// for JST; the full Roe wall flux, central + upwind, for the Roe schemes)
// in FP64; one thread per record writes its term, one thread adds them in
// record order
template <typename T>
__global__ void k_wall_terms(FluxCtx<T> fc, const DevRec *recs, const double *wsol, const double *psol, int nmir,
                             int roe, double gamma, double eps, double *terms, int *err) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmir) return;
  const Dims &d = fc.d;
  const StateC<T> &s = fc.s;
  const Geo<T> &g = fc.g;
  const DevRec &r = recs[m];
  if (roe) {
    double ws5[5], ve[5], vd[5];
    for (int n = 0; n < 5; n++) ws5[n] = wsol[n * nmir + m];
    roe_wall_rec(fc, r, ws5, psol[m], gamma, eps, 1.0, ve, vd, err);
    for (int dd = 0; dd < 3; dd++) terms[3 * m + dd] = ve[1 + dd] + vd[1 + dd];
    return;
  }
  int zz = r.zz;
  double sc[3], ss[3];
  axis_vec(sc, zz, (double)face_s_any(g, zz, r.c[0], r.c[1], r.c[2]));
  axis_vec(ss, zz, (double)face_s_any(g, zz, r.s[0], r.s[1], r.s[2]));
  long qc = ix(d, r.c[0], r.c[1], r.c[2]);
  double wc[5], ws[5], f1[5], f2[5];
  for (int n = 0; n < 5; n++) wc[n] = full(fc.inf, n, s.w[n * d.E + qc]);
  for (int n = 0; n < 5; n++) ws[n] = wsol[n * nmir + m];
  cflux3(wc, full(fc.inf, 5, s.p[qc]), sc, f1);
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

}  // namespace mge
