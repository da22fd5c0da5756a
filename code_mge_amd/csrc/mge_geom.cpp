// mge_geom.cpp -- cut-cell geometry preprocessor for analytic bodies.
//
// The reference obtains cell types, wall records (txcut) and the
// reflected-node cloud weights from make_fsijob/solid_ipolcoef
// (srcf90/main.f90:44,57), whose library is not part of the reference
// tree.  This file defines that step for quadric bodies:
//  - a cell is solid when its centre is strictly inside the body
//    (blank = 0, the role of celltype==csolid in init_metric.f90:282-290);
//  - every (fluid, solid) face-neighbour pair of interior cells becomes one
//    record with the fields bc_cutcell.f90/efluxb.f90/dfluxb.f90 read:
//    faceid/zz (side and direction of the solid cell), ccid/scid,
//    psurf (where the segment between the two centres meets the surface),
//    the outward unit normal there, rr (distance from the solid centre to
//    the tangent plane) and radcur (principal curvatures, sphere/cylinder);
//  - the reflected node is the mirror image of the solid centre in the
//    tangent plane; its cloud is the 2x2x2 cell-centre stencil around it
//    with trilinear weights, solid cells dropped and the rest renormalised.
// Records are emitted in k, j, i order of the fluid cell and faceid order.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mge.h"

namespace {

struct Quadric {
  double M[3][3];
  double c[3];
  bool valid;
};

Quadric make_quadric(const mge_body *b) {
  Quadric q;
  std::memset(&q, 0, sizeof(q));
  q.valid = false;
  if (!b || b->kind == MGE_BODY_NONE) return q;
  for (int d = 0; d < 3; d++) q.c[d] = b->center[d];
  if (b->kind == MGE_BODY_SPHERE) {
    double r2 = b->semi[0] * b->semi[0];
    for (int d = 0; d < 3; d++) q.M[d][d] = 1.0 / r2;
  } else if (b->kind == MGE_BODY_CYLINDER) {
    double r2 = b->semi[0] * b->semi[0];
    for (int d = 0; d < 3; d++) q.M[d][d] = (d == b->axis) ? 0.0 : 1.0 / r2;
  } else if (b->kind == MGE_BODY_ELLIPSOID) {
    // x_b = R^T (x - c);  phi = sum (x_b,a / s_a)^2 - 1  =>  M = R D R^T
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) {
        double acc = 0.0;
        for (int a = 0; a < 3; a++)
          acc += b->rot[r * 3 + a] * b->rot[s * 3 + a] / (b->semi[a] * b->semi[a]);
        q.M[r][s] = acc;
      }
  } else {
    return q;
  }
  q.valid = true;
  return q;
}

double qform(const Quadric &q, const double *u, const double *v) {
  double acc = 0.0;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) acc += u[r] * q.M[r][s] * v[s];
  return acc;
}

double phi(const Quadric &q, const double *x) {
  double e[3] = {x[0] - q.c[0], x[1] - q.c[1], x[2] - q.c[2]};
  return qform(q, e, e) - 1.0;
}

// first crossing of the surface on the segment a -> b (phi(a) >= 0 > phi(b))
double crossing(const Quadric &q, const double *a, const double *b) {
  double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e[3] = {a[0] - q.c[0], a[1] - q.c[1], a[2] - q.c[2]};
  double A = qform(q, d, d), B = 2.0 * qform(q, e, d), C = qform(q, e, e) - 1.0;
  if (A <= 0.0) return 0.5;
  double disc = B * B - 4.0 * A * C;
  if (disc < 0.0) disc = 0.0;
  double sq = std::sqrt(disc);
  double qq = -0.5 * (B + (B >= 0.0 ? sq : -sq));
  double t1 = qq / A, t2 = (qq != 0.0) ? C / qq : t1;
  double best = -1.0;
  for (double t : {t1, t2})
    if (t >= 0.0 && t <= 1.0 && (best < 0.0 || t < best)) best = t;
  if (best < 0.0) best = 0.5;
  return best;
}

void normal_at(const Quadric &q, const double *x, double *n) {
  double e[3] = {x[0] - q.c[0], x[1] - q.c[1], x[2] - q.c[2]};
  double g[3];
  for (int r = 0; r < 3; r++) g[r] = q.M[r][0] * e[0] + q.M[r][1] * e[1] + q.M[r][2] * e[2];
  double nn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  if (nn <= 0.0) nn = 1.0;
  for (int r = 0; r < 3; r++) n[r] = g[r] / nn;
}

// index s in [lo, hi-1] with c[s] <= x < c[s+1] (clamped) and the weight t
void bracket(const std::vector<double> &cen, int lo, int hi, double x, int *s, double *t) {
  int a = lo;
  while (a + 1 < hi && cen[a + 1] <= x) a++;
  if (a > hi - 1) a = hi - 1;
  double tt = (x - cen[a]) / (cen[a + 1] - cen[a]);
  if (tt < 0.0) tt = 0.0;
  if (tt > 1.0) tt = 1.0;
  *s = a;
  *t = tt;
}

}  // namespace

extern "C" int mge_geom_level(const mge_body *body, int ni, int nj, int nk,
                              const double *xv, const double *yv, const double *zv,
                              uint8_t *blank, mge_cutrec *recs, int cap, int *nmir) {
  if (ni < 1 || nj < 1 || nk < 1 || !xv || !yv || !zv || !blank || !nmir) return MGE_ERR_ARG;
  const int il = ni + 1, jl = nj + 1, kl = nk + 1;
  Quadric q = make_quadric(body);
  std::vector<double> cx(ni + 3), cy(nj + 3), cz(nk + 3);
  for (int i = 1; i <= ni + 2; i++) cx[i] = 0.5 * (xv[i] + xv[i + 1]);
  for (int j = 1; j <= nj + 2; j++) cy[j] = 0.5 * (yv[j] + yv[j + 1]);
  for (int k = 1; k <= nk + 2; k++) cz[k] = 0.5 * (zv[k] + zv[k + 1]);
  auto bidx = [&](int i, int j, int k) -> long {
    return (long)(i - 2) + (long)ni * ((long)(j - 2) + (long)nj * (long)(k - 2));
  };
  for (int k = 2; k <= kl; k++)
    for (int j = 2; j <= jl; j++)
      for (int i = 2; i <= il; i++) {
        double x[3] = {cx[i], cy[j], cz[k]};
        blank[bidx(i, j, k)] = (q.valid && phi(q, x) < 0.0) ? 0 : 1;
      }
  int count = 0;
  if (!q.valid) {
    *nmir = 0;
    return MGE_OK;
  }
  const int off[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 2; k <= kl; k++)
    for (int j = 2; j <= jl; j++)
      for (int i = 2; i <= il; i++) {
        if (!blank[bidx(i, j, k)]) continue;
        for (int f = 0; f < 6; f++) {
          int is = i + off[f][0], js = j + off[f][1], ks = k + off[f][2];
          if (is < 2 || is > il || js < 2 || js > jl || ks < 2 || ks > kl) continue;
          if (blank[bidx(is, js, ks)]) continue;
          if (recs && count < cap) {
            mge_cutrec &r = recs[count];
            std::memset(&r, 0, sizeof(r));
            r.faceid = f + 1;
            r.zz = (f % 3) + 1;
            r.ccid[0] = i; r.ccid[1] = j; r.ccid[2] = k;
            r.scid[0] = is; r.scid[1] = js; r.scid[2] = ks;
            double xc[3] = {cx[i], cy[j], cz[k]};
            double xs[3] = {cx[is], cy[js], cz[ks]};
            double t = crossing(q, xc, xs);
            double ps[3], n[3];
            for (int d = 0; d < 3; d++) ps[d] = xc[d] + t * (xs[d] - xc[d]);
            normal_at(q, ps, n);
            double dist = (ps[0] - xs[0]) * n[0] + (ps[1] - xs[1]) * n[1] + (ps[2] - xs[2]) * n[2];
            if (dist < 0.0) dist = 0.0;
            for (int d = 0; d < 3; d++) {
              r.psurf[d] = ps[d];
              r.nvec[d] = n[d];
            }
            r.rr = dist;
            if (body->kind == MGE_BODY_SPHERE) {
              r.radcur[0] = r.radcur[1] = 1.0 / body->semi[0];
            } else if (body->kind == MGE_BODY_CYLINDER) {
              r.radcur[0] = 1.0 / body->semi[0];
              r.radcur[1] = 0.0;
            }
            double xr[3];
            for (int d = 0; d < 3; d++) xr[d] = xs[d] + 2.0 * dist * n[d];
            int s[3];
            double tw[3];
            bracket(cx, 2, il, xr[0], &s[0], &tw[0]);
            bracket(cy, 2, jl, xr[1], &s[1], &tw[1]);
            bracket(cz, 2, kl, xr[2], &s[2], &tw[2]);
            if (nk == 1) { s[2] = 2; tw[2] = 0.0; }
            double wsum = 0.0;
            int np = 0;
            for (int c8 = 0; c8 < 8; c8++) {
              int di = c8 & 1, dj = (c8 >> 1) & 1, dk = (c8 >> 2) & 1;
              int ci = s[0] + di, cj = s[1] + dj, ck = s[2] + dk;
              if (ci > il || cj > jl || ck > kl) continue;
              double wgt = (di ? tw[0] : 1.0 - tw[0]) * (dj ? tw[1] : 1.0 - tw[1]) *
                           (dk ? tw[2] : 1.0 - tw[2]);
              if (wgt <= 0.0) continue;
              if (!blank[bidx(ci, cj, ck)]) continue;
              r.ijk[np][0] = ci; r.ijk[np][1] = cj; r.ijk[np][2] = ck;
              r.coef[np] = wgt;
              wsum += wgt;
              np++;
            }
            if (np == 0 || wsum < 1e-12) {
              np = 1;
              r.ijk[0][0] = i; r.ijk[0][1] = j; r.ijk[0][2] = k;
              r.coef[0] = 1.0;
              wsum = 1.0;
            }
            for (int c = 0; c < np; c++) r.coef[c] = r.coef[c] / wsum;
            r.npoint = np;
          }
          count++;
        }
      }
  *nmir = count;
  return MGE_OK;
}
