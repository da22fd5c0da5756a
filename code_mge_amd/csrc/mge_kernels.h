// mge_kernels.h -- device-side data views and the per-cell arithmetic of
// the Jameson-RK stage, templated on the bulk type T (double or float).
//
// Each helper restates one piece of the reference with the same expression
// tree (the library is compiled with -ffp-contract=off); in FP64 the
// quotients of the convective flux and of the pressure switch are products
// with a reciprocal refined to within an ulp, so with T = double the GPU
// path agrees with the reference arithmetic to rounding:
//   eos           depvars.f90:37-52
//   cflux         ceflux.f90:1-25 with the face vector of one direction
//   fp_at/diw_at  dflux.f90:70-86 (pressure switch, enthalpy differences)
//   dface         dflux.f90:87-111 (one face of the JST dissipation)
//   dmatrix       dflux.f90:283-339 (matrix dissipation)
// Every literal is written as a T value so that T = float computes in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace mge {

// Block dimensions in the reference numbering: interior 2..il, ghosts
// 0,1 and ie,ib; arrays (0:ib,0:jb,0:kb) with i fastest.
struct Dims {
  int ni, nj, nk;
  int il, jl, kl, ie, je, ke, ib, jb, kb;
  long E;   // cells per component
  long sy;  // stride of j
  long sz;  // stride of k
};

__host__ __device__ inline long ix(const Dims &d, int i, int j, int k) {
  return (long)i + d.sy * (long)j + d.sz * (long)k;
}

// ghost cell t of the shell: the k planes {0,1,ke,kb} over all (i,j), then
// the j rows {0,1,je,jb} over all i and k = 2..kl, then the i columns
// {0,1,ie,ib} over the interior (j,k)
__device__ inline bool shell_cell(const Dims &d, long t, int &x, int &y, int &z) {
  const long ni4 = d.ib + 1, nj4 = d.jb + 1;
  const long nK = 4 * ni4 * nj4, nJ = 4 * ni4 * d.nk, nI = 4L * d.nj * d.nk;
  auto g = [](long m, int e) { return m < 2 ? (int)m : e + (int)m - 2; };
  if (t < nK) {
    x = (int)(t % ni4);
    const long r = t / ni4;
    y = (int)(r % nj4);
    z = g(r / nj4, d.ke);
    return true;
  }
  t -= nK;
  if (t < nJ) {
    x = (int)(t % ni4);
    const long r = t / ni4;
    z = 2 + (int)(r % d.nk);
    y = g(r / d.nk, d.je);
    return true;
  }
  t -= nJ;
  if (t < nI) {
    x = g(t % 4, d.ie);
    const long r = t / 4;
    y = 2 + (int)(r % d.nj);
    z = 2 + (int)(r / d.nj);
    return true;
  }
  return false;
}

// ghost cell t of the same shell, numbered by how many faces meet there:
// the faces' four layers over the interior of the face (i faces with the
// layer fastest, j and k faces with it slowest), then the twelve edges' 4 x 4 ghost pairs along their
// interior, then the 4 x 4 x 4 corner cells.  Consecutive threads then take
// the same faces, so a wave of k_bc_faces runs one chain shape.
__host__ __device__ inline bool shell_cell_by_depth(const Dims &d, long t, int &x, int &y, int &z) {
  const long ni = d.ni, nj = d.nj, nk = d.nk;
  auto g = [](long m, int e) { return m < 2 ? (int)m : e + (int)m - 2; };
  long n = 4 * nj * nk;
  if (t < n) {  // the i faces' four layers fastest: a row's ghost cells side by side
    const long r = t / 4;
    x = g(t % 4, d.ie);
    y = 2 + (int)(r % nj);
    z = 2 + (int)(r / nj);
    return true;
  }
  t -= n;
  n = 4 * ni * nk;
  if (t < n) {
    const long r = t % (ni * nk);
    y = g(t / (ni * nk), d.je);
    x = 2 + (int)(r % ni);
    z = 2 + (int)(r / ni);
    return true;
  }
  t -= n;
  n = 4 * ni * nj;
  if (t < n) {
    const long r = t % (ni * nj);
    z = g(t / (ni * nj), d.ke);
    x = 2 + (int)(r % ni);
    y = 2 + (int)(r / ni);
    return true;
  }
  t -= n;
  n = 16 * ni;
  if (t < n) {
    const long m = t / ni;
    x = 2 + (int)(t % ni);
    y = g(m & 3, d.je);
    z = g(m >> 2, d.ke);
    return true;
  }
  t -= n;
  n = 16 * nj;
  if (t < n) {
    const long m = t / nj;
    y = 2 + (int)(t % nj);
    x = g(m & 3, d.ie);
    z = g(m >> 2, d.ke);
    return true;
  }
  t -= n;
  n = 16 * nk;
  if (t < n) {
    const long m = t / nk;
    z = 2 + (int)(t % nk);
    x = g(m & 3, d.ie);
    y = g(m >> 2, d.je);
    return true;
  }
  t -= n;
  if (t < 64) {
    x = g(t & 3, d.ie);
    y = g((t >> 2) & 3, d.je);
    z = g(t >> 4, d.ke);
    return true;
  }
  return false;
}

__host__ __device__ inline long shell_cells(const Dims &d) {
  return 4L * (d.ib + 1) * (d.jb + 1) + 4L * (d.ib + 1) * d.nk + 4L * d.nj * d.nk;
}

// flags byte per cell
enum : uint8_t { F_FLUID = 1, F_TOUCHED = 2 };

// two-element vectors for paired loads and stores of neighbouring cells
template <typename T> struct PairOf;
template <> struct PairOf<double> { using type = double2; };
template <> struct PairOf<float> { using type = float2; };
template <typename T> using Pair = typename PairOf<T>::type;

// read-only view of one state buffer: w (5 components) followed by p
template <typename T>
struct StateC {
  const T *w;
  const T *p;
};
template <typename T>
struct State {
  T *w;
  T *p;
};

template <typename T>
struct Geo {
  const T *dx, *dy, *dz;  // vertex differences x(i+1)-x(i)
};

// The FP32 bulk path stores the state as its deviation from the freestream
// (w' = w - w_inf, p' = p - p_inf) and evaluates fluxes and the EOS in delta
// form, so that the freestream parts -- pressure and energy terms far larger
// than the changes that drive convergence -- cancel exactly instead of being
// rounded at the scale of the freestream.  FP64 stores the state itself and
// follows the reference expression trees.
template <typename T>
constexpr bool kDelta = std::is_same<T, float>::value;

template <typename T>
struct Inf {
  T w[5];  // freestream conservative state
  T p;     // freestream pressure
};

// full value of component n (5 = p) of a stored state value, in FP64
template <typename T>
__device__ inline double full(const Inf<T> &f, int n, T v) {
  if constexpr (kDelta<T>) return (double)v + (double)(n < 5 ? f.w[n] : f.p);
  else return v;
}
// stored value of a full FP64 state value
template <typename T>
__device__ inline T stored(const Inf<T> &f, int n, double v) {
  if constexpr (kDelta<T>) return (T)(v - (double)(n < 5 ? f.w[n] : f.p));
  else return (T)v;
}

// face vector of cell (i,j,k) in direction dir: the non-zero component of
// si/sj/sk (init_metric.f90:109-177) on an axis-aligned grid.
// si = (-dy*dz, 0, 0), sj = (0, -dx*dz, 0), sk = (0, 0, -dx*dy); the
// cross products of init_metric reduce to exactly these values.
template <int DIR, typename T>
__device__ inline T face_s(const Geo<T> &g, int i, int j, int k) {
  if (DIR == 0) return -(g.dy[j] * g.dz[k]);
  if (DIR == 1) return -(g.dx[i] * g.dz[k]);
  return -(g.dx[i] * g.dy[j]);
}

// ------------------------------------------------------------------ entry points
//
// Every kernel of the multigrid cycle is written as NAME_body(args, bx, by,
// bz) and gets two entry points:
//   NAME    one case; the block coordinates are blockIdx
//   NAME_b  a batch of cases in one launch: the 1-D grid is the
//           concatenation of the cases' own grids; cg[c] = {first block,
//           grid x, grid y, batch slot} of case c (sorted by first block)
//           and a block finds its case by binary search.  A case whose
//           slot is inactive (act[slot] == 0) returns at once, so a batch
//           stops each case exactly at its own convergence.
// Each case's blocks compute exactly what its single-case launch computes.
__device__ inline int batch_case(const int4 *cg, int nc, int b) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (cg[mid].x <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define MGE_ENTRY_T(NAME, ARGS, NT, MINB)                                                              \
  template <typename T>                                                                                \
  __global__ void __launch_bounds__(NT, MINB) NAME(ARGS<T> A) {                                        \
    NAME##_body(A, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);                                 \
  }                                                                                                    \
  template <typename T>                                                                                \
  __global__ void __launch_bounds__(NT, MINB) NAME##_b(const ARGS<T> *As, const int4 *cg, int nc,      \
                                                       const int *act) {                               \
    const int b = (int)blockIdx.x;                                                                     \
    const int c = batch_case(cg, nc, b);                                                               \
    const int4 g = cg[c];                                                                              \
    if (!act[g.w]) return;                                                                             \
    const int l = b - g.x;                                                                             \
    NAME##_body(As[c], l % g.y, (l / g.y) % g.z, l / (g.y * g.z));                                     \
  }

// Block coordinates of a single-case launch with each XCD given a
// contiguous range of blocks: workgroups are dealt to the 8 XCDs in turn
// (block w runs on XCD w % 8), so block w takes the logical index
// (w % 8) * (W / 8) + w / 8.  Neighbouring tiles (the segments of one row, a
// tile and its halo neighbours) then run on one XCD and share its L2.  Grids
// whose block count is not a multiple of 8 keep the plain order.
__device__ inline void xcd_block(int &bx, int &by, int &bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y, W = gx * gy * gridDim.z;
  const unsigned w = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned l = (W % 8u == 0u) ? (w % 8u) * (W / 8u) + w / 8u : w;
  bx = (int)(l % gx);
  by = (int)((l / gx) % gy);
  bz = (int)(l / (gx * gy));
}

// MGE_ENTRY_T with the XCD-contiguous block order in the single-case entry
#define MGE_ENTRY_T_XCD(NAME, ARGS, NT, MINB)                                                          \
  template <typename T>                                                                                \
  __global__ void __launch_bounds__(NT, MINB) NAME(ARGS<T> A) {                                        \
    int bx, by, bz;                                                                                    \
    xcd_block(bx, by, bz);                                                                             \
    NAME##_body(A, bx, by, bz);                                                                        \
  }                                                                                                    \
  template <typename T>                                                                                \
  __global__ void __launch_bounds__(NT, MINB) NAME##_b(const ARGS<T> *As, const int4 *cg, int nc,      \
                                                       const int *act) {                               \
    const int b = (int)blockIdx.x;                                                                     \
    const int c = batch_case(cg, nc, b);                                                               \
    const int4 g = cg[c];                                                                              \
    if (!act[g.w]) return;                                                                             \
    const int l = b - g.x;                                                                             \
    NAME##_body(As[c], l % g.y, (l / g.y) % g.z, l / (g.y * g.z));                                     \
  }

#define MGE_ENTRY(NAME, ARGS, NT, MINB)                                                                \
  __global__ void __launch_bounds__(NT, MINB) NAME(ARGS A) {                                           \
    NAME##_body(A, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);                                 \
  }                                                                                                    \
  __global__ void __launch_bounds__(NT, MINB) NAME##_b(const ARGS *As, const int4 *cg, int nc,         \
                                                       const int *act) {                               \
    const int b = (int)blockIdx.x;                                                                     \
    const int c = batch_case(cg, nc, b);                                                               \
    const int4 g = cg[c];                                                                              \
    if (!act[g.w]) return;                                                                             \
    const int l = b - g.x;                                                                             \
    NAME##_body(As[c], l % g.y, (l / g.y) % g.z, l / (g.y * g.z));                                     \
  }

template <typename T>
__device__ inline T dimd(T a, T b) { return a > b ? a - b : T(0); }
template <typename T>
__device__ inline T maxd(T a, T b) { return a > b ? a : b; }

// depvars_one: p = (gamma-1)*dim(rhoE, 0.5*|m|^2/rho)
template <typename T>
__device__ inline T eos(T gamma, const T *v) {
  T s = T(0);
  s = s + v[1] * v[1];
  s = s + v[2] * v[2];
  s = s + v[3] * v[3];
  T qq = T(0.5) * s / v[0];
  return (gamma - T(1)) * dimd(v[4], qq);
}

// EOS of a stored state: the reference form for FP64; for FP32 the delta
// form p' = (gamma-1)(E' - KE'), KE' = KE - KE_inf written from the
// deviations, KE' = (rho_inf (2 m_inf.m' + |m'|^2) - rho' |m_inf|^2) /
// (2 rho rho_inf), with the reference's clamp p >= 0
template <typename T>
__device__ inline T eos_st(T gamma, const Inf<T> &f, const T *v) {
  if constexpr (kDelta<T>) {
    const T r0 = f.w[0];
    T mm = f.w[1] * v[1] + f.w[2] * v[2] + f.w[3] * v[3];
    T m2 = v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
    T mi2 = f.w[1] * f.w[1] + f.w[2] * f.w[2] + f.w[3] * f.w[3];
    T rho = r0 + v[0];
    T ke = (r0 * (T(2) * mm + m2) - v[0] * mi2) / (T(2) * rho * r0);
    T pd = (gamma - T(1)) * (v[4] - ke);
    return f.p + pd < T(0) ? -f.p : pd;
  } else {
    return eos(gamma, v);
  }
}

template <typename T>
__device__ inline void load5(const T *w, long E, long q, T *v) {
#pragma unroll
  for (int n = 0; n < 5; n++) v[n] = w[n * E + q];
}

// 1/x for positive normal x: the hardware reciprocal estimate refined by
// Newton-Raphson steps r <- r + r (1 - x r), two in FP64 and one in FP32
// (within an ulp of the rounded quotient; the correctly rounded FP32
// division takes about three times the instructions)
template <typename T>
__device__ inline T rcp_pos(T x) {
  if constexpr (std::is_same<T, double>::value) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
  } else {
    float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
  }
}

// sqrt(x) for positive normal x: in FP64 the hardware reciprocal square
// root estimate refined by one Goldschmidt step and two Newton corrections
// (the iteration of the compiler's own expansion, without its range scaling
// and special-value selects, which positive normal arguments never need),
// in FP32 the hardware square root (within an ulp)
template <typename T>
__device__ inline T sqrt_pos(T x) {
  if constexpr (std::is_same<T, double>::value) {
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
  } else {
    return __builtin_amdgcn_sqrtf(x);
  }
}

// ceflux with ss having only component DIR non-zero (value S)
// (the quotients by rho as products with the refined reciprocal, which
// agree with the divisions to rounding)
// (h = v[4] + p, the total enthalpy per volume, formed by the caller)
template <int DIR, typename T>
__device__ inline void cflux_h(const T *v, T h, T p, T S, T *fs) {
  const T rr = rcp_pos(v[0]);
  T f1 = v[1 + DIR] * S;
  fs[0] = f1;
  fs[1] = f1 * v[1] * rr;
  fs[2] = f1 * v[2] * rr;
  fs[3] = f1 * v[3] * rr;
  fs[1 + DIR] = fs[1 + DIR] + p * S;
  fs[4] = f1 * h * rr;
}
template <int DIR, typename T>
__device__ inline void cflux(const T *v, T p, T S, T *fs) {
  cflux_h<DIR>(v, v[4] + p, p, S, fs);
}

// ceflux of a stored state: the reference form for FP64; for FP32 the flux
// deviation F - F_inf in delta form (f1 = f1_inf + f1', rho = rho_inf +
// rho'), e.g. F_c' = (rho_inf (m_inf,c f1' + m_c' f1_inf + m_c' f1') -
// rho' m_inf,c f1_inf) / (rho rho_inf) + p' S delta_c,DIR.  F_inf is the
// same for every cell of a grid line (S does not vary along it), so face
// differences of F' equal those of F.
// The _h form takes h = v[4] + p (stored form) from the caller; v[4] itself
// is not read.
template <int DIR, typename T>
__device__ inline void cflux_st_h(const Inf<T> &f, const T *v, T h, T p, T S, T *fs) {
  if constexpr (kDelta<T>) {
    const T r0 = f.w[0];
    T f1i = f.w[1 + DIR] * S, f1d = v[1 + DIR] * S;
    T den = (r0 + v[0]) * r0;
    fs[0] = f1d;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      T mi = f.w[1 + c], md = v[1 + c];
      fs[1 + c] = (r0 * (mi * f1d + md * f1i + md * f1d) - v[0] * mi * f1i) / den;
    }
    fs[1 + DIR] = fs[1 + DIR] + p * S;
    T hi = f.w[4] + f.p, hd = h;
    fs[4] = (r0 * (hi * f1d + hd * f1i + hd * f1d) - v[0] * hi * f1i) / den;
  } else {
    cflux_h<DIR>(v, h, p, S, fs);
  }
}
template <int DIR, typename T>
__device__ inline void cflux_st(const Inf<T> &f, const T *v, T p, T S, T *fs) {
  cflux_st_h<DIR>(f, v, v[4] + p, p, S, fs);
}

// general ceflux (full face vector), used by the wall records
template <typename T>
__device__ inline void cflux3(const T *v, T p, const T *ss, T *fs) {
  T f1 = T(0);
  f1 = f1 + v[1] * ss[0];
  f1 = f1 + v[2] * ss[1];
  f1 = f1 + v[3] * ss[2];
  fs[0] = f1;
  fs[1] = f1 * v[1] / v[0] + p * ss[0];
  fs[2] = f1 * v[2] / v[0] + p * ss[1];
  fs[3] = f1 * v[3] / v[0] + p * ss[2];
  fs[4] = f1 * (v[4] + p) / v[0];
}

// dflux_matrix (dflux.f90:283-339); u1/u2 carry the enthalpy in slot 4
template <typename T>
__device__ inline void dmatrix(T gamma, const T *u1, const T *u2, const T *dw, const T *ss, T *ff) {
  const T half = T(0.5);
  T gam1 = gamma - T(1), epsn = T(1), epsl = T(1);
  T rr = half * (u1[0] + u2[0]);
  T us[3];
  for (int d = 0; d < 3; d++) us[d] = half * (u1[1 + d] + u2[1 + d]) / rr;
  T ee = half * (u1[4] + u2[4]) / rr;
  T su = T(0);
  su = su + us[0] * us[0];
  su = su + us[1] * us[1];
  su = su + us[2] * us[2];
  T qq = half * su;
  T c2 = gamma * gam1 * (ee - qq);
  T hh = c2 / gam1 + qq;
  T sss = T(0);
  sss = sss + ss[0] * ss[0];
  sss = sss + ss[1] * ss[1];
  sss = sss + ss[2] * ss[2];
  T cs = sqrt(c2 * sss);
  T qn = T(0);
  qn = qn + ss[0] * us[0];
  qn = qn + ss[1] * us[1];
  qn = qn + ss[2] * us[2];
  T e0 = fabs(qn) + cs;
  T e1 = maxd(epsn * e0, T(fabs(qn + cs)));
  T e2 = maxd(epsn * e0, T(fabs(qn - cs)));
  T e3 = maxd(epsl * e0, T(fabs(qn)));
  T s1 = half * (e1 + e2), s2 = half * (e1 - e2), s3 = e3;
  T ak1 = (s1 - s3) / c2 * gam1;
  T ak2 = s2 / cs;
  T ak3 = c2 / gam1 / sss;
  T vd = qq * dw[0] - us[0] * dw[1] - us[1] * dw[2] - us[2] * dw[3] + dw[4];
  T qd = qn * dw[0] - ss[0] * dw[1] - ss[1] * dw[2] - ss[2] * dw[3];
  T f[5];
  f[0] = ak1 * vd - ak2 * qd + s3 * dw[0];
  f[1] = ak1 * (us[0] * vd - ss[0] * ak3 * qd) + ak2 * (gam1 * ss[0] * vd - us[0] * qd) + s3 * dw[1];
  f[2] = ak1 * (us[1] * vd - ss[1] * ak3 * qd) + ak2 * (gam1 * ss[1] * vd - us[1] * qd) + s3 * dw[2];
  f[3] = ak1 * (us[2] * vd - ss[2] * ak3 * qd) + ak2 * (gam1 * ss[2] * vd - us[2] * qd) + s3 * dw[3];
  f[4] = ak1 * (hh * vd - qn * ak3 * qd) + ak2 * (gam1 * qn * vd - hh * qd) + s3 * dw[4];
  for (int n = 0; n < 5; n++) ff[n] = f[n];
}

// everything a flux evaluation needs
template <typename T>
struct FluxCtx {
  Dims d;
  StateC<T> s;
  Geo<T> g;
  const T *radsum;  // radi+radj+radk at interior cells
  T gamma, vis2, vis4, cdis;
  int dissip;       // 0 scalar, 1 matrix
  Inf<T> inf;       // freestream (FP32: the state is stored as deviation)
};

template <int DIR>
__device__ inline long dstep(const Dims &d) {
  return DIR == 0 ? 1L : (DIR == 1 ? d.sy : d.sz);
}
template <int DIR>
__device__ inline int dne(const Dims &d) {
  return DIR == 0 ? d.ie : (DIR == 1 ? d.je : d.ke);
}
template <int DIR>
__device__ inline int dnl(const Dims &d) {
  return DIR == 0 ? d.il : (DIR == 1 ? d.jl : d.kl);
}

// pressure switch fp(s) of one line (dflux.f90:70-81); fp(0)=fp(ib)=0.
// Branch-free: the loads use an index clamped into [1, ie] (always inside
// the array) and the boundary convention is applied by a select, so the
// compiler can batch the loads of neighbouring calls.
// the switch value from the pressures at s-1, s, s+1 (the reference sums
// the j line in the opposite order)
template <int DIR, typename T>
__device__ inline T fp_val(const Inf<T> &inf, T pm, T pc, T pp) {
  T pb, num;
  if (DIR == 1) {
    pb = pp + T(2) * pc + pm;
    num = fabs(pp - T(2) * pc + pm);
  } else {
    pb = pm + T(2) * pc + pp;
    num = fabs(pm - T(2) * pc + pp);
  }
  if constexpr (kDelta<T>) pb = pb + T(4) * inf.p;
  pb = fabs(pb);
  // FP64: times the refined reciprocal (pb = 0, never for physical
  // pressures, is selected away); FP32 divides
  T v;
  if constexpr (kDelta<T>) v = num / pb;
  else v = num * rcp_pos(pb);
  return fabs(pb) > T(0) ? v : T(0);
}
template <int DIR, typename T>
__device__ inline T fp_at(const FluxCtx<T> &c, long base, int s) {
  int nb = dne<DIR>(c.d) + 1;
  int sc = s < 1 ? 1 : (s > nb - 1 ? nb - 1 : s);
  long st = dstep<DIR>(c.d);
  T pm = c.s.p[base + (sc - 1) * st], pc = c.s.p[base + sc * st], pp = c.s.p[base + (sc + 1) * st];
  T v = fp_val<DIR>(c.inf, pm, pc, pp);
  return (s <= 0 || s >= nb) ? T(0) : v;
}

// enthalpy-vector difference diw(s) = q(s)-q(s-1) with the end copies
// diw(1)=diw(2), diw(ib)=diw(ie) (dflux.f90:82-86)
template <int DIR, typename T>
__device__ inline void diw_at(const FluxCtx<T> &c, long base, int s, T *o) {
  int ne = dne<DIR>(c.d);
  s = s < 2 ? 2 : (s > ne ? ne : s);
  long st = dstep<DIR>(c.d);
  long qa = base + s * st, qb = qa - st;
  const long E = c.d.E;
#pragma unroll
  for (int n = 0; n < 4; n++) o[n] = c.s.w[n * E + qa] - c.s.w[n * E + qb];
  o[4] = (c.s.w[4 * E + qa] + c.s.p[qa]) - (c.s.w[4 * E + qb] + c.s.p[qb]);
}

template <int DIR, typename T>
__device__ inline T rad_at(const FluxCtx<T> &c, long base, int s) {
  int nl = dnl<DIR>(c.d);
  s = s < 2 ? 2 : (s > nl ? nl : s);
  return c.radsum[base + s * dstep<DIR>(c.d)];
}

// JST dissipation at face a (between a-1 and a) of the line `base`
// (dflux.f90:87-111); zero outside 2..ie like the unset dfsx entries.
template <int DIR, typename T>
__device__ inline void dface(const FluxCtx<T> &c, long base, int a, int li, int lj, int lk, T *out) {
  int ne = dne<DIR>(c.d);
  if (a < 2 || a > ne) {
#pragma unroll
    for (int n = 0; n < 5; n++) out[n] = T(0);
    return;
  }
  T f0 = fp_at<DIR>(c, base, a - 2), f1 = fp_at<DIR>(c, base, a - 1);
  T f2 = fp_at<DIR>(c, base, a), f3 = fp_at<DIR>(c, base, a + 1);
  T ep = maxd(maxd(maxd(f0, f1), f2), f3);
  T lp = rad_at<DIR>(c, base, a);
  T lm = rad_at<DIR>(c, base, a - 1);
  T lam = T(0.5) * (lm + lp);
  T dis2 = c.vis2 * ep;
  T dis4 = c.vis4;
  dis4 = dimd(dis4, dis2);
  dis2 = dis2 * c.cdis;
  dis4 = dis4 * c.cdis;
  T dm[5], da[5], dp[5];
  diw_at<DIR>(c, base, a - 1, dm);
  diw_at<DIR>(c, base, a, da);
  diw_at<DIR>(c, base, a + 1, dp);
  T d[5];
#pragma unroll
  for (int n = 0; n < 5; n++) {
    T t = dp[n] - T(2) * da[n] + dm[n];
    d[n] = dis2 * da[n] - dis4 * t;
  }
  if (c.dissip == 0) {
#pragma unroll
    for (int n = 0; n < 5; n++) out[n] = d[n] * lam;
  } else {
    long st = dstep<DIR>(c.d);
    long qm = base + (a - 1) * st, qa = base + a * st;
    T um[5], ua[5];
    load5(c.s.w, c.d.E, qm, um);
    load5(c.s.w, c.d.E, qa, ua);
    if constexpr (kDelta<T>) {
#pragma unroll
      for (int n = 0; n < 5; n++) {
        um[n] = um[n] + c.inf.w[n];
        ua[n] = ua[n] + c.inf.w[n];
      }
      um[4] = um[4] + (c.s.p[qm] + c.inf.p);
      ua[4] = ua[4] + (c.s.p[qa] + c.inf.p);
    } else {
      um[4] = um[4] + c.s.p[qm];
      ua[4] = ua[4] + c.s.p[qa];
    }
    int fi = DIR == 0 ? a : li, fj = DIR == 1 ? a : lj, fk = DIR == 2 ? a : lk;
    T ss[3] = {T(0), T(0), T(0)};
    ss[DIR] = face_s<DIR>(c.g, fi, fj, fk);
    dmatrix(c.gamma, um, ua, d, ss, out);
  }
}

// convective flux at face a: 0.5*(F(a)+F(a-1)) (eflux.f90:44-48)
template <int DIR, typename T>
__device__ inline void eface(const FluxCtx<T> &c, long base, int a, int li, int lj, int lk, T *out) {
  long st = dstep<DIR>(c.d);
  long qa = base + a * st, qm = qa - st;
  int ai = DIR == 0 ? a : li, aj = DIR == 1 ? a : lj, ak = DIR == 2 ? a : lk;
  int mi = DIR == 0 ? a - 1 : li, mj = DIR == 1 ? a - 1 : lj, mk = DIR == 2 ? a - 1 : lk;
  T va[5], vm[5], fa[5], fm[5];
  load5(c.s.w, c.d.E, qa, va);
  load5(c.s.w, c.d.E, qm, vm);
  cflux_st<DIR>(c.inf, va, c.s.p[qa], face_s<DIR>(c.g, ai, aj, ak), fa);
  cflux_st<DIR>(c.inf, vm, c.s.p[qm], face_s<DIR>(c.g, mi, mj, mk), fm);
#pragma unroll
  for (int n = 0; n < 5; n++) out[n] = T(0.5) * (fa[n] + fm[n]);
}

// the point-implicit factor of the physical time derivative in unsteady
// dual time stepping, lam = 1 + de*3/2/dtstep with de = fn*vdtim
// (euler.f90:322-326, 388-391)
template <typename T>
__device__ inline T unsteady_lam(T fn, const T *vdtim, const double *dtim, int vt, T dtstep, long q) {
  T de = fn * vdtim[q];
  if (vt == 1) de = fn * T(dtim[0]);
  return T(1) + de * T(3) / T(2) / dtstep;
}

}  // namespace mge
