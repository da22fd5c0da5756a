// mge_output.h -- the solution fields code_mge hands to its output and FSI
// layers (upload_solut, srcf90/upload_solut.f90:89-241).
//
// Per interior cell, in FP64 whatever the bulk precision: rho, u, v, w,
// cp = (p/p0 - 1) 2/(gamma M^2), Mach, total-pressure loss
// p (1 + 0.2 M^2)^3.5 / p_tot - 1, and for laminar flow also rho E / rho and
// three zero gradient slots (upload_solut.f90:213-234).  The output is
// field-major, i fastest over the ni*nj*nk interior cells, like the
// reference's solut(lowc:uppc, ..., nval).  One thread per cell: a
// streaming kernel (40 B read, 56 or 88 B written per cell).
#pragma once

#include "mge_kernels.h"

namespace mge {

template <typename T>
struct SolutArgs {
  Dims d;
  StateC<T> s;
  Inf<T> inf;
  double gamma, p0, rm, ptot;
  int ns;       // laminar Navier-Stokes: 11 fields, else 7
  long ncell;   // cells written by this launch: c0 .. c0+ncell-1
  double *out;  // nval*ostride: field f of cell c0+c at out[f*ostride + c]
  long c0 = 0;        // first cell (i fastest over ni*nj*nk)
  long ostride = 0;   // field stride of out (ncell when 0)
};

template <typename T>
__global__ void __launch_bounds__(256) k_upload_solut(SolutArgs<T> A) {
  const Dims &d = A.d;
  const long cl = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (cl >= A.ncell) return;
  const long c = A.c0 + cl;
  const long nij = (long)d.ni * d.nj;
  const int k = (int)(c / nij);
  const int j = (int)((c - (long)k * nij) / d.ni);
  const int i = (int)(c - (long)k * nij - (long)j * d.ni);
  const long q = ix(d, i + 2, j + 2, k + 2);
  double w[5];
#pragma unroll
  for (int n = 0; n < 5; n++) w[n] = full(A.inf, n, A.s.w[n * d.E + q]);
  const double rr = w[0];
  const double uu = w[1] / rr;
  const double vv = w[2] / rr;
  const double ww = w[3] / rr;
  const double re = w[4];
  const double qq = 0.5 * (uu * uu + vv * vv + ww * ww) * rr;
  double pp = (A.gamma - 1.0) * (re > qq ? re - qq : 0.0);
  const double mm = sqrt((uu * uu + vv * vv + ww * ww) / A.gamma / pp * rr);
  const double ptloss = pp * pow(1.0 + 0.2 * mm * mm, 3.5) / A.ptot - 1.0;
  pp = (pp / A.p0 - 1.0) * 2.0 / (A.gamma * A.rm * A.rm);
  double *o = A.out + cl;
  const long nc = A.ostride ? A.ostride : A.ncell;
  o[0] = rr;
  o[nc] = uu;
  o[2 * nc] = vv;
  o[3 * nc] = ww;
  o[4 * nc] = pp;
  o[5 * nc] = mm;
  o[6 * nc] = ptloss;
  if (A.ns) {
    o[7 * nc] = re / rr;
    o[8 * nc] = 0.0;
    o[9 * nc] = 0.0;
    o[10 * nc] = 0.0;
  }
}

}  // namespace mge
