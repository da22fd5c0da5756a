// mge_smooth.h -- implicit residual smoothing (psmoo_old.f90) and the
// post-smoothing update, templated on the bulk type T.
//
// psmoo_old solves, for every grid line and each direction in turn, the
// constant-coefficient tridiagonal system (1+2 eps) x_s - eps (x_(s-1) +
// x_(s+1)) = dw_s with a Thomas recurrence whose pivots e1/e2 depend only
// on the line length (psmoo.f90:134-166); blank cells are skipped (their
// value is kept and they do not feed the recurrence).
#pragma once
#include "mge_kernels.h"

namespace mge {

template <typename T>
struct SmoothArgs {
  Dims d;
  T *dw;
  const uint8_t *flags;
  const T *e1, *e2;  // pivots of this direction (psmoo.f90:134-166)
  const T *fa;       // -bb/e1 (x lines: affine-map form of the forward step)
  const T *r1;       // 1/e1, rounded once from the FP64 pivots
  T bb;
};

// The forward step's division by the pivot e1_s.  FP64 multiplies by the
// reciprocal pivot instead (one rounding more per cell, a few ulp over a
// line, and no FP64 division on the recurrence's critical path); FP32
// divides, as the reference does.
template <typename T>
__device__ __attribute__((always_inline)) inline T piv_div(T x, const T *__restrict__ e1,
                                                           const T *__restrict__ r1, int s) {
  if constexpr (kDelta<T>) return x / e1[s];
  else return x * r1[s];
}

// a lane moves its CPL cells as pairs when CPL is even, the line's first
// cell (i = 2) sits on a pair boundary (so does every lane's first cell,
// i0 - 2 being a multiple of CPL) and all its cells are inside the line
template <int CPL, typename T>
__device__ __attribute__((always_inline)) inline bool xscan_pairs(const T *row, int i0, int il) {
  if constexpr (CPL % 2 != 0) return false;
  else return ((reinterpret_cast<uintptr_t>(row + 2) & (2 * sizeof(T) - 1)) == 0) && i0 + CPL - 1 <= il;
}

// psmoo.f90:171-188 (x lines): one wavefront per K consecutive (j,k)
// lines of a component.  Lane l owns CPL consecutive cells of each line, so
// loads and stores are coalesced.  Both recurrences are first composed as
// affine maps (forward y_i = A_i y_(i-1) + B_i, backward z_i = C_i z_(i+1)
// + y_i), scanned across the wave, and then every lane re-evaluates its own
// cells with the reference formula from the scanned boundary value.  Only
// the lane-boundary values differ from the sequential evaluation (by a few
// ulp); blank cells keep their value as in the reference.  The K lines are
// independent: their loads are in flight together and their shuffles
// interleave, and the pivots of the lane's cells are loaded once for all K.
template <int CPL, int K, typename T>
__device__ __attribute__((always_inline)) inline void smooth_x_scan(const SmoothArgs<T> &S, int bx) {
  const Dims &d = S.d;
  const int lane = threadIdx.x & 63;
  // the wave index is uniform: keep the line arithmetic scalar
  const int wave = __builtin_amdgcn_readfirstlane((int)((bx * (long)blockDim.x + threadIdx.x) >> 6));
  const int nl = d.nj * d.nk;
  const int nlines = nl * 5;
  const int l0 = wave * K;
  if (l0 >= nlines) return;  // uniform per wave
  const int il = d.il, ni = d.ni;
  const int i0 = 2 + lane * CPL;
  // pivots: e1 (FP32 divides by it) or 1/e1 (FP64 multiplies), -bb/e1, e2
  T pv[CPL], pa[CPL], pe2[CPL];
  bool ok[CPL];
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    const int i = i0 + c;
    ok[c] = i <= il;
    const int ic = ok[c] ? i : 2;
    pv[c] = kDelta<T> ? S.e1[ic] : S.r1[ic];
    pa[c] = S.fa[ic];
    pe2[c] = S.e2[ic];
  }
  auto pdiv = [](T x, T p) { if constexpr (kDelta<T>) return x / p; else return x * p; };
  T r[K][CPL];
  bool fl[K][CPL], lv[K];
  long base[K], fbase[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    lv[t] = l0 + t < nlines;
    const int lc = lv[t] ? l0 + t : l0;  // a missing line repeats the first; it is never stored
    const int n = lc / nl;
    const int l = lc - n * nl;
    const int j = 2 + l % d.nj, k = 2 + l / d.nj;
    fbase[t] = ix(d, 0, j, k);
    base[t] = (long)n * d.E + fbase[t];
    if (xscan_pairs<CPL>(S.dw + base[t], i0, il)) {
      const Pair<T> *pp = reinterpret_cast<const Pair<T> *>(S.dw + base[t] + i0);
#pragma unroll
      for (int h = 0; h < CPL / 2; h++) {
        const Pair<T> v = pp[h];
        r[t][2 * h] = v.x;
        r[t][2 * h + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; c++) r[t][c] = ok[c] ? S.dw[base[t] + i0 + c] : T(0);
    }
#pragma unroll
    for (int c = 0; c < CPL; c++) fl[t][c] = ok[c] && (S.flags[fbase[t] + i0 + c] & F_FLUID);
  }
  // forward: lane composite of y_i = A_i y_(i-1) + B_i
  T Al[K], Bl[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    Al[t] = T(1);
    Bl[t] = T(0);
#pragma unroll
    for (int c = 0; c < CPL; c++) {
      const int i = i0 + c;
      T A, B;
      if (!ok[c]) { A = T(1); B = T(0); }
      else if (!fl[t][c]) { A = T(0); B = r[t][c]; }
      else if (i == 2) { A = T(0); B = pdiv(r[t][c], pv[c]); }
      else { A = pa[c]; B = pdiv(r[t][c], pv[c]); }
      Bl[t] = A * Bl[t] + B;
      Al[t] = A * Al[t];
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int t = 0; t < K; t++) {
      T Ap = __shfl_up(Al[t], off, 64), Bp = __shfl_up(Bl[t], off, 64);
      if (lane >= off) {
        Bl[t] = Al[t] * Bp + Bl[t];
        Al[t] = Al[t] * Ap;
      }
    }
  }
  // re-evaluate the lane's cells from the incoming value (y kept in r)
#pragma unroll
  for (int t = 0; t < K; t++) {
    T yp = __shfl_up(Bl[t], 1, 64);
    if (lane == 0) yp = T(0);
#pragma unroll
    for (int c = 0; c < CPL; c++) {
      const int i = i0 + c;
      T v = r[t][c];
      if (ok[c] && fl[t][c]) v = (i == 2) ? pdiv(r[t][c], pv[c]) : pdiv(r[t][c] - S.bb * yp, pv[c]);
      r[t][c] = v;
      yp = v;
    }
  }
  // backward: z_i = C_i z_(i+1) + y_i, C_i = -e2_i on fluid cells 2..ni
  T Cl[K], Dl[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    Cl[t] = T(1);
    Dl[t] = T(0);
#pragma unroll
    for (int c = CPL - 1; c >= 0; c--) {
      const int i = i0 + c;
      T C, Dv;
      if (!ok[c]) { C = T(1); Dv = T(0); }
      else if (fl[t][c] && i <= ni) { C = -pe2[c]; Dv = r[t][c]; }
      else { C = T(0); Dv = r[t][c]; }
      Dl[t] = C * Dl[t] + Dv;
      Cl[t] = C * Cl[t];
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int t = 0; t < K; t++) {
      T Cn = __shfl_down(Cl[t], off, 64), Dn = __shfl_down(Dl[t], off, 64);
      if (lane + off < 64) {
        Dl[t] = Cl[t] * Dn + Dl[t];
        Cl[t] = Cl[t] * Cn;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    T zn = __shfl_down(Dl[t], 1, 64);
    if (lane == 63) zn = T(0);
#pragma unroll
    for (int c = CPL - 1; c >= 0; c--) {
      const int i = i0 + c;
      if (!ok[c]) continue;
      if (fl[t][c] && i <= ni) r[t][c] = r[t][c] - pe2[c] * zn;
      zn = r[t][c];
    }
    if (!lv[t]) continue;
    if (xscan_pairs<CPL>(S.dw + base[t], i0, il)) {
      // blank cells are written back with the value they were read with
      Pair<T> *pp = reinterpret_cast<Pair<T> *>(S.dw + base[t] + i0);
#pragma unroll
      for (int h = 0; h < CPL / 2; h++) pp[h] = Pair<T>{r[t][2 * h], r[t][2 * h + 1]};
    } else {
#pragma unroll
      for (int c = 0; c < CPL; c++)
        if (fl[t][c]) S.dw[base[t] + i0 + c] = r[t][c];
    }
  }
}

// lines per wave of the x scan: four lines of up to 256 cells share a
// wave, 512-cell lines take one (measured at 256^3 with single loads: one,
// two and four lines per wave give 527, 512 and 477 us in FP64, 409, 339
// and 324 us in FP32; with pairs, four lines 408 us FP64, 298 us FP32)
template <int CPL>
constexpr int xscan_lines() { return CPL >= 8 ? 1 : 4; }

// sequential x lines for blocks wider than 512 cells
template <typename T>
__device__ __attribute__((always_inline)) inline void smooth_x_seq(const SmoothArgs<T> &S, int bx) {
  const Dims &d = S.d;
  long t = bx * (long)blockDim.x + threadIdx.x;
  long nl = (long)d.nj * d.nk;
  if (t >= nl * 5) return;
  int n = (int)(t / nl);
  long l = t - (long)n * nl;
  int j = 2 + (int)(l % d.nj), k = 2 + (int)(l / d.nj);
  T *D = S.dw + n * d.E;
  long b0 = ix(d, 0, j, k);
  if (S.flags[b0 + 2] & F_FLUID) D[b0 + 2] = piv_div(D[b0 + 2], S.e1, S.r1, 2);
  for (int i = 3; i <= d.il; i++)
    if (S.flags[b0 + i] & F_FLUID) D[b0 + i] = piv_div(D[b0 + i] - S.bb * D[b0 + i - 1], S.e1, S.r1, i);
  for (int i = 2; i <= d.ni; i++) {
    int ir = d.il + 1 - i;
    if (S.flags[b0 + ir] & F_FLUID) D[b0 + ir] = D[b0 + ir] - S.e2[ir] * D[b0 + ir + 1];
  }
}

// x-line entry points: one per scan width (cells per lane, chosen from
// the line length) and the sequential form; a batch launches each width
// once over the cases that use it
template <int CPL, typename T>
__global__ void __launch_bounds__(256) k_smooth_x_scan(SmoothArgs<T> S) {
  smooth_x_scan<CPL, xscan_lines<CPL>()>(S, (int)blockIdx.x);
}
template <int CPL, typename T>
__global__ void __launch_bounds__(256) k_smooth_x_scan_b(const SmoothArgs<T> *As, const int4 *cg, int nc,
                                                         const int *act) {
  const int b = (int)blockIdx.x;
  const int c = batch_case(cg, nc, b);
  const int4 g = cg[c];
  if (!act[g.w]) return;
  smooth_x_scan<CPL, xscan_lines<CPL>()>(As[c], b - g.x);
}
template <typename T>
__device__ __attribute__((always_inline)) inline void k_smooth_x_seq_body(const SmoothArgs<T> &S, int bx, int, int) {
  smooth_x_seq(S, bx);
}
MGE_ENTRY_T(k_smooth_x_seq, SmoothArgs, 256, 1)

// One line of the exact sequential recurrence of psmoo_old
// (psmoo.f90:190-237): cells s = 2..nl at stride st.  Full chunks of CH
// cells are read into registers double-buffered with unconditional loads,
// so the compiler can keep the next chunk's loads in flight (counted waits)
// while the current chunk runs its dependent divisions.  The updates are
// branch-free: blank cells hold dw = +-0 (blanked before smoothing), so
// y = m*((r - bb*y)/e1) with m = 0 on blank cells gives the reference's
// values on fluid cells and keeps the zeros.  The backward sweep starts
// at s = nl with z(nl+1) = 0 and e2(nl) = 0, which returns y(nl) there.
template <int CH, typename T>
__device__ __attribute__((always_inline)) inline void line_sweep(T *__restrict__ D, const uint8_t *__restrict__ F,
                                                                 long b0, long st, int nl, const T *__restrict__ e1,
                                                                 const T *__restrict__ r1, const T *__restrict__ e2,
                                                                 T bb) {
  const int ncell = nl - 1;
  const int nfull = ncell / CH;
  T A[CH], B[CH], MA[CH], MB[CH];
  T y = T(0);
#define LD_F(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    long q_ = b0 + (long)(2 + (c) * CH + t) * st;                     \
    BUF[t] = D[q_];                                                   \
    MK[t] = (F[q_] & F_FLUID) ? T(1) : T(0);                          \
  }
#define PR_F(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    int s_ = 2 + (c) * CH + t;                                        \
    y = MK[t] * piv_div(BUF[t] - bb * y, e1, r1, s_);                 \
    D[b0 + (long)s_ * st] = y;                                        \
  }
#define LD_B(BUF, MK, c)                                              \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    long q_ = b0 + (long)(nl - (c) * CH - t) * st;                    \
    BUF[t] = D[q_];                                                   \
    MK[t] = (F[q_] & F_FLUID) ? T(1) : T(0);                          \
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
    T m_ = (F[q_] & F_FLUID) ? T(1) : T(0);
    y = m_ * piv_div(D[q_] - bb * y, e1, r1, s_);
    D[q_] = y;
  }
  // backward
  y = T(0);
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
    T m_ = (F[q_] & F_FLUID) ? T(1) : T(0);
    y = D[q_] - m_ * (e2[s_] * y);
    D[q_] = y;
  }
#undef LD_F
#undef PR_F
#undef LD_B
#undef PR_B
}

// The same recurrences with three passes over the line instead of four:
// the forward sweep keeps only the value y at the end of every chunk of CH
// cells (in LDS, one slot per thread and chunk) and writes nothing; the
// backward sweep, chunk by chunk from the end, re-reads the chunk's input,
// recomputes its forward values from the checkpoint (the same operations,
// so the same bits) and writes the final values.  The next chunk's loads
// are issued before the current chunk's recurrence in both sweeps.  Lines
// of up to CH * MAXC cells.
template <int CH, int MAXC, typename T, typename CK>
__device__ __attribute__((always_inline)) inline void line_sweep_ck(T *__restrict__ D, const uint8_t *__restrict__ F,
                                                                    long b0, long st, int nl, const T *__restrict__ e1,
                                                                    const T *__restrict__ r1, const T *__restrict__ e2,
                                                                    T bb, CK ck) {
  const int ncell = nl - 1;
  const int nfull = ncell / CH;
  T A[CH], B[CH], Y[CH];
  unsigned ma = 0, mb = 0;
  // full chunks: unconditional loads (the chunk index is clamped by the
  // caller), so the next chunk's loads stay in flight under counted waits
#define LDC(BUF, MK, c)                                               \
  {                                                                   \
    MK = 0;                                                           \
    _Pragma("unroll") for (int t = 0; t < CH; t++) {                  \
      long q_ = b0 + (long)(2 + (c) * CH + t) * st;                   \
      BUF[t] = D[q_];                                                 \
      MK |= (F[q_] & F_FLUID) ? (1u << t) : 0u;                       \
    }                                                                 \
  }
#define FWD(BUF, MK, c, OUT)                                          \
  _Pragma("unroll") for (int t = 0; t < CH; t++) {                    \
    T m_ = ((MK >> t) & 1u) ? T(1) : T(0);                            \
    y = m_ * piv_div(BUF[t] - bb * y, e1, r1, 2 + (c) * CH + t);      \
    OUT                                                               \
  }
#define BWD(MK, c)                                                    \
  _Pragma("unroll") for (int t = CH - 1; t >= 0; t--) {               \
    int s_ = 2 + (c) * CH + t;                                        \
    T m_ = ((MK >> t) & 1u) ? T(1) : T(0);                            \
    z = Y[t] - m_ * (e2[s_] * z);                                     \
    D[b0 + (long)s_ * st] = z;                                        \
  }
  // ---- forward: checkpoints after every full chunk, then the tail
  T y = T(0);
  if (nfull > 0) {
    LDC(A, ma, 0);
    for (int c = 0; c < nfull; c += 2) {
      int c1 = c + 1 < nfull ? c + 1 : c;
      LDC(B, mb, c1);
      FWD(A, ma, c, ;);
      ck[c] = y;
      if (c + 1 >= nfull) break;
      int c2 = c + 2 < nfull ? c + 2 : c + 1;
      LDC(A, ma, c2);
      FWD(B, mb, c + 1, ;);
      ck[c + 1] = y;
    }
  }
  // ---- backward: the tail (recomputed from the last checkpoint) ...
  T z = T(0);
  {
    T yt = nfull > 0 ? ck[nfull - 1] : T(0);
    const int s0 = 2 + nfull * CH;
    T Yt[CH];
    y = yt;
    for (int s_ = s0; s_ <= nl; s_++) {
      long q_ = b0 + (long)s_ * st;
      T m_ = (F[q_] & F_FLUID) ? T(1) : T(0);
      y = m_ * piv_div(D[q_] - bb * y, e1, r1, s_);
      Yt[s_ - s0] = y;
    }
    for (int s_ = nl; s_ >= s0; s_--) {
      long q_ = b0 + (long)s_ * st;
      T m_ = (F[q_] & F_FLUID) ? T(1) : T(0);
      z = Yt[s_ - s0] - m_ * (e2[s_] * z);
      D[q_] = z;
    }
  }
  // ... then the full chunks from the end, each recomputed from the
  // checkpoint before it
  if (nfull > 0) {
    LDC(A, ma, nfull - 1);
    for (int c = nfull - 1; c >= 0; c -= 2) {
      int c1 = c >= 1 ? c - 1 : 0;
      LDC(B, mb, c1);
      y = c > 0 ? ck[c - 1] : T(0);
      FWD(A, ma, c, Y[t] = y;);
      BWD(ma, c);
      if (c < 1) break;
      int c2 = c >= 2 ? c - 2 : 0;
      LDC(A, ma, c2);
      y = c - 1 > 0 ? ck[c - 2] : T(0);
      FWD(B, mb, c - 1, Y[t] = y;);
      BWD(mb, c - 1);
    }
  }
#undef LDC
#undef FWD
#undef BWD
}

// y/z line: the three-pass sweep for lines of up to 256 cells, else the
// four-pass one.  16 checkpoints of one wave take 8 KB of LDS in FP64 and
// the four-pass form runs 8-cell chunks, so neither limits the waves per
// CU below what the three-pass registers allow (measured at 256^3: 4% faster
// in FP64 and 17% in FP32 than 32 checkpoints and 16-cell fallback chunks).
template <typename T>
__device__ __attribute__((always_inline)) inline void smooth_line(T *D, const uint8_t *F, long b0, long st, int nl,
                                                                  const T *e1, const T *r1, const T *e2, T bb) {
  constexpr int CH = 16, MAXC = 16;
  __shared__ T ck[MAXC][64];
  if (nl - 1 <= CH * MAXC && blockDim.x * blockDim.y <= 64) {
    // ck[c][lane]: the lanes of one chunk index sit side by side
    T *mine = &ck[0][threadIdx.x + blockDim.x * threadIdx.y];
    struct Col {
      T *p;
      __device__ T &operator[](int c) const { return p[c * 64]; }
    } col{mine};
    line_sweep_ck<CH, MAXC>(D, F, b0, st, nl, e1, r1, e2, bb, col);
  } else {
    line_sweep<8>(D, F, b0, st, nl, e1, r1, e2, bb);
  }
}

// psmoo.f90:190-206 (y lines); thread per (i,k) and component
template <typename T>
__device__ __attribute__((always_inline)) inline void k_smooth_y_body(const SmoothArgs<T> &S, int bx, int by, int bz) {
  const Dims &d = S.d;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int k = 2 + by;
  int n = bz;
  if (i > d.il) return;
  smooth_line(S.dw + n * d.E, S.flags, ix(d, i, 0, k), d.sy, d.jl, S.e1, S.r1, S.e2, S.bb);
}
MGE_ENTRY_T_XCD(k_smooth_y, SmoothArgs, 64, 1)

// psmoo.f90:209-237 (z lines); thread per (i,j) and component
template <typename T>
__device__ __attribute__((always_inline)) inline void k_smooth_z_body(const SmoothArgs<T> &S, int bx, int by, int bz) {
  const Dims &d = S.d;
  int i = 2 + bx * blockDim.x + threadIdx.x;
  int j = 2 + by;
  int n = bz;
  if (i > d.il) return;
  smooth_line(S.dw + n * d.E, S.flags, ix(d, i, j, 0), d.sz, d.kl, S.e1, S.r1, S.e2, S.bb);
}
MGE_ENTRY_T_XCD(k_smooth_z, SmoothArgs, 64, 1)

// after smoothing: w = wn - dw and the EOS on interior cells
// (euler.f90:381-399), one thread per cell of the whole array so that the
// accesses stream through aligned rows; the ghost shell of the previous
// state is carried over where anything reads it before bcond rewrites it,
// as k_ghost_copy does on unsmoothed stages
template <typename T>
struct UpdArgs {
  Dims d;
  const T *wn, *dw;
  StateC<T> src;
  State<T> dst;
  T gamma;
  Inf<T> inf;
  // unsteady update on the finest active level: w = wn - dw/lam
  // (euler.f90:383-393)
  int ulam;
  const T *vdtim;
  const double *dtim;
  T fn, dtstep;
  int vt;
  int carry;   // carry the ghost shell over
};

template <typename T>
__device__ __attribute__((always_inline)) inline void k_update_ghost_body(const UpdArgs<T> &U, int bx, int, int) {
  const Dims &d = U.d;
  const T *wn = U.wn, *dw = U.dw;
  const StateC<T> &src = U.src;
  const State<T> &dst = U.dst;
  const T gamma = U.gamma;
  const long q = bx * (long)blockDim.x + threadIdx.x;
  if (q >= d.E) return;
  int i, j, k;
  if (d.E <= 0xffffffffL) {  // 32-bit index arithmetic
    const unsigned uq = (unsigned)q, sz = (unsigned)d.sz, sy = (unsigned)d.sy;
    const unsigned kk = uq / sz, r = uq - kk * sz, jj = r / sy;
    i = (int)(r - jj * sy);
    j = (int)jj;
    k = (int)kk;
  } else {
    k = (int)(q / d.sz);
    j = (int)((q - (long)k * d.sz) / d.sy);
    i = (int)(q - (long)k * d.sz - (long)j * d.sy);
  }
  if (!(i >= 2 && i <= d.il && j >= 2 && j <= d.jl && k >= 2 && k <= d.kl)) {
    if (U.carry) {
      for (int n = 0; n < 5; n++) dst.w[n * d.E + q] = src.w[n * d.E + q];
      dst.p[q] = src.p[q];
    }
    return;
  }
  T v[5];
  if (U.ulam) {
    const T lam = unsteady_lam(U.fn, U.vdtim, U.dtim, U.vt, U.dtstep, q);
    for (int n = 0; n < 5; n++) v[n] = wn[n * d.E + q] - dw[n * d.E + q] / lam;
  } else {
    for (int n = 0; n < 5; n++) v[n] = wn[n * d.E + q] - dw[n * d.E + q];
  }
  for (int n = 0; n < 5; n++) dst.w[n * d.E + q] = v[n];
  dst.p[q] = eos_st(gamma, U.inf, v);
}
MGE_ENTRY_T(k_update_ghost, UpdArgs, 256, 1)


// ---- psmoo (psmoo.f90:1-95) with progon3's sweep (progon3.f90:1-62) ------
// Rows (1+2e_r) x_r - e_r x_{r-1} - e_r x_{r+1} = d_r along each line of
// direction DIR with the per-cell coefficients e of timestep (the end
// couplings removed, as psmoo passes them, so progon3's sweep is the Thomas
// recurrence).  One thread per line carries all five components: the
// forward sweep leaves u_{r+1} in dw and v_{r+1} in vb, the backward sweep
// forms x_r = v_{r+1} x_{r+1} + u_{r+1} in place.  Lines are numbered with
// the fastest index of the other two first, so neighbouring threads touch
// neighbouring cells on y and z lines.
template <typename T>
struct PsmooArgs {
  Dims d;
  T *dw;
  const T *eps;  // this direction's coefficients
  T *vb;         // per-cell scratch of the forward sweep (E)
};

template <int DIR, typename T>
__device__ __attribute__((always_inline)) inline void k_psmoo_line_body(const PsmooArgs<T> &A, int bx, int, int) {
  const Dims &d = A.d;
  const int t = bx * blockDim.x + threadIdx.x;
  int n, na, nb;
  long st;
  if (DIR == 0) { n = d.ni; na = d.nj; nb = d.nk; st = 1; }
  else if (DIR == 1) { n = d.nj; na = d.ni; nb = d.nk; st = d.sy; }
  else { n = d.nk; na = d.ni; nb = d.nj; st = d.sz; }
  if (t >= na * nb) return;
  const int a = 2 + t % na, b = 2 + t / na;
  long q;
  if (DIR == 0) q = ix(d, 2, a, b);
  else if (DIR == 1) q = ix(d, a, 2, b);
  else q = ix(d, a, b, 2);
  const long E = d.E;
  T v = T(0), u[5] = {T(0), T(0), T(0), T(0), T(0)};
  for (int r = 1; r <= n; r++, q += st) {
    const T e = A.eps[q];
    const T aa = T(1) + T(2) * e;
    const T bb = r == n ? T(0) : -e;
    const T cc = r == 1 ? T(0) : -e;
    const T z = T(1) / (aa + cc * v);
    v = -bb * z;
    A.vb[q] = v;
#pragma unroll
    for (int m = 0; m < 5; m++) {
      u[m] = (-cc * u[m] + A.dw[m * E + q]) * z;
      A.dw[m * E + q] = u[m];
    }
  }
  q -= st;  // last cell: x_n = u_{n+1}, already in place
  for (int r = n - 1; r >= 1; r--) {
    const T vn = A.vb[q - st];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      u[m] = vn * u[m] + A.dw[m * E + q - st];
      A.dw[m * E + q - st] = u[m];
    }
    q -= st;
  }
}
template <typename T>
__device__ __attribute__((always_inline)) inline void k_psmoo_x_body(const PsmooArgs<T> &A, int bx, int by, int bz) {
  k_psmoo_line_body<0, T>(A, bx, by, bz);
}
template <typename T>
__device__ __attribute__((always_inline)) inline void k_psmoo_y_body(const PsmooArgs<T> &A, int bx, int by, int bz) {
  k_psmoo_line_body<1, T>(A, bx, by, bz);
}
template <typename T>
__device__ __attribute__((always_inline)) inline void k_psmoo_z_body(const PsmooArgs<T> &A, int bx, int by, int bz) {
  k_psmoo_line_body<2, T>(A, bx, by, bz);
}
MGE_ENTRY_T(k_psmoo_x, PsmooArgs, 64, 1)
MGE_ENTRY_T(k_psmoo_y, PsmooArgs, 64, 1)
MGE_ENTRY_T(k_psmoo_z, PsmooArgs, 64, 1)

}  // namespace mge
