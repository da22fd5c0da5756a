// mge_geom.cpp -- cut-cell geometry preprocessor for analytic bodies and
// triangulated surfaces.
//
// The reference obtains cell types, wall records (txcut) and the
// reflected-node cloud weights from make_fsijob/solid_ipolcoef
// (srcf90/main.f90:44,57), whose library is not part of the reference
// tree.  This file defines that step for quadric bodies and for closed
// triangulated surfaces (e.g. the patches of a PLOT3D surface file):
//  - a cell is solid when its centre is strictly inside the body (surfaces:
//    odd number of surface crossings of the grid line through the centre
//    before it, the line nudged off the grid planes by 1e-9 of a cell)
//    (blank = 0, the role of celltype==csolid in init_metric.f90:185-193);
//  - every (fluid, solid) face-neighbour pair of interior cells becomes one
//    record with the fields bc_cutcell.f90/efluxb.f90/dfluxb.f90 read:
//    faceid/zz (side and direction of the solid cell), ccid/scid,
//    psurf (where the segment between the two centres meets the surface),
//    the outward unit normal there (surfaces: the normal of the crossed
//    triangle), rr (distance from the solid centre to the tangent plane)
//    and radcur (principal curvatures, sphere/cylinder; 0 otherwise);
//  - the reflected node is the mirror image of the solid centre in the
//    tangent plane; its cloud is the 2x2x2 cell-centre stencil around it
//    with trilinear weights, solid cells dropped and the rest renormalised.
// Records are emitted in k, j, i order of the fluid cell and faceid order.
#include <algorithm>
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

// ---- closed triangulated surface
struct TriSurf {
  const double *t = nullptr;  // 9 values per triangle: three vertices
  int n = 0;
  // uniform bins over the triangles' bounding box (segment queries)
  double lo[3], bs[3];
  int nb[3];
  std::vector<std::vector<int>> bins;
  void build(const double *tri, int ntri) {
    t = tri;
    n = ntri;
    double hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = 1e300;
      hi[d] = -1e300;
    }
    for (int m = 0; m < n; m++)
      for (int v = 0; v < 3; v++)
        for (int d = 0; d < 3; d++) {
          lo[d] = std::min(lo[d], t[9 * m + 3 * v + d]);
          hi[d] = std::max(hi[d], t[9 * m + 3 * v + d]);
        }
    int per = std::max(1, (int)std::cbrt((double)n / 4.0));
    for (int d = 0; d < 3; d++) {
      nb[d] = per;
      double ext = std::max(hi[d] - lo[d], 1e-12);
      lo[d] -= 1e-9 * ext;
      bs[d] = ext * (1.0 + 2e-9) / nb[d];
    }
    bins.assign((size_t)nb[0] * nb[1] * nb[2], {});
    for (int m = 0; m < n; m++) {
      int a[3], b[3];
      for (int d = 0; d < 3; d++) {
        double mn = std::min(std::min(t[9 * m + d], t[9 * m + 3 + d]), t[9 * m + 6 + d]);
        double mx = std::max(std::max(t[9 * m + d], t[9 * m + 3 + d]), t[9 * m + 6 + d]);
        a[d] = bin_of(d, mn);
        b[d] = bin_of(d, mx);
      }
      for (int z = a[2]; z <= b[2]; z++)
        for (int y = a[1]; y <= b[1]; y++)
          for (int x = a[0]; x <= b[0]; x++) bins[((size_t)z * nb[1] + y) * nb[0] + x].push_back(m);
    }
  }
  int bin_of(int d, double v) const {
    int q = (int)std::floor((v - lo[d]) / bs[d]);
    return std::min(std::max(q, 0), nb[d] - 1);
  }
};

// crossing of the axis line {p + s e_ax} with triangle m: true and the
// parameter s when the line meets the triangle's interior (projected
// barycentric test with strict signs; the caller nudges the line off edges)
bool line_tri(const double *T, int ax, const double *p, double *s) {
  const int u = (ax + 1) % 3, v = (ax + 2) % 3;
  double au = T[u] - p[u], av = T[v] - p[v];
  double bu = T[3 + u] - p[u], bv = T[3 + v] - p[v];
  double cu = T[6 + u] - p[u], cv = T[6 + v] - p[v];
  double w0 = bu * cv - bv * cu, w1 = cu * av - cv * au, w2 = au * bv - av * bu;
  bool pos = w0 > 0.0 && w1 > 0.0 && w2 > 0.0, neg = w0 < 0.0 && w1 < 0.0 && w2 < 0.0;
  if (!pos && !neg) return false;
  double sw = w0 + w1 + w2;
  double x = (w0 * T[ax] + w1 * T[3 + ax] + w2 * T[6 + ax]) / sw;
  *s = x - p[ax];
  return true;
}

void tri_normal(const double *T, double *nv) {
  double e1[3], e2[3];
  for (int d = 0; d < 3; d++) {
    e1[d] = T[3 + d] - T[d];
    e2[d] = T[6 + d] - T[d];
  }
  nv[0] = e1[1] * e2[2] - e1[2] * e2[1];
  nv[1] = e1[2] * e2[0] - e1[0] * e2[2];
  nv[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  if (nn <= 0.0) nn = 1.0;
  for (int d = 0; d < 3; d++) nv[d] /= nn;
}

// first surface crossing on the axis-aligned segment xc -> xs (fluid to
// solid centre): parameter t in [0,1] and the crossed triangle's normal
// turned towards xc; false if none is found
bool tri_crossing(const TriSurf &S, const double *xc, const double *xs, double nudge, double *t, double *nv) {
  int ax = 0;
  for (int d = 0; d < 3; d++)
    if (xs[d] != xc[d]) ax = d;
  double p[3] = {xc[0], xc[1], xc[2]};
  p[(ax + 1) % 3] += 1.0e-9 * nudge;
  p[(ax + 2) % 3] += 1.3e-9 * nudge;
  double len = xs[ax] - xc[ax];
  int a[3], b[3];
  for (int d = 0; d < 3; d++) {
    a[d] = S.bin_of(d, std::min(xc[d], xs[d]));
    b[d] = S.bin_of(d, std::max(xc[d], xs[d]));
  }
  double best = 2.0;
  int bm = -1;
  for (int z = a[2]; z <= b[2]; z++)
    for (int y = a[1]; y <= b[1]; y++)
      for (int x = a[0]; x <= b[0]; x++)
        for (int m : S.bins[((size_t)z * S.nb[1] + y) * S.nb[0] + x]) {
          double sp;
          if (!line_tri(S.t + 9 * m, ax, p, &sp)) continue;
          double tt = sp / len;
          if (tt >= 0.0 && tt <= 1.0 && tt < best) {
            best = tt;
            bm = m;
          }
        }
  if (bm < 0) return false;
  *t = best;
  tri_normal(S.t + 9 * bm, nv);
  double dx[3] = {xc[0] - xs[0], xc[1] - xs[1], xc[2] - xs[2]};
  if (nv[0] * dx[0] + nv[1] * dx[1] + nv[2] * dx[2] < 0.0)
    for (int d = 0; d < 3; d++) nv[d] = -nv[d];
  return true;
}

// the common part of both entry points: records of every (fluid, solid)
// face pair, from the classification in blank and a crossing routine
template <class Hit>
int build_records_common(int ni, int nj, int nk, const std::vector<double> &cx, const std::vector<double> &cy,
                         const std::vector<double> &cz, const uint8_t *blank, mge_cutrec *recs, int cap, int *nmir,
                         Hit hit) {
  const int il = ni + 1, jl = nj + 1, kl = nk + 1;
  auto bidx = [&](int i, int j, int k) -> long {
    return (long)(i - 2) + (long)ni * ((long)(j - 2) + (long)nj * (long)(k - 2));
  };
  int count = 0;
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
            double ps[3], n[3];
            hit(xc, xs, ps, n, r.radcur);
            double dist = (ps[0] - xs[0]) * n[0] + (ps[1] - xs[1]) * n[1] + (ps[2] - xs[2]) * n[2];
            if (dist < 0.0) dist = 0.0;
            for (int d = 0; d < 3; d++) {
              r.psurf[d] = ps[d];
              r.nvec[d] = n[d];
            }
            r.rr = dist;
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

void centres(int ni, int nj, int nk, const double *xv, const double *yv, const double *zv, std::vector<double> &cx,
             std::vector<double> &cy, std::vector<double> &cz) {
  cx.assign(ni + 3, 0.0);
  cy.assign(nj + 3, 0.0);
  cz.assign(nk + 3, 0.0);
  for (int i = 1; i <= ni + 2; i++) cx[i] = 0.5 * (xv[i] + xv[i + 1]);
  for (int j = 1; j <= nj + 2; j++) cy[j] = 0.5 * (yv[j] + yv[j + 1]);
  for (int k = 1; k <= nk + 2; k++) cz[k] = 0.5 * (zv[k] + zv[k + 1]);
}

}  // namespace

extern "C" int mge_geom_level_tri(const double *tri, int ntri, int ni, int nj, int nk, const double *xv,
                                  const double *yv, const double *zv, uint8_t *blank, mge_cutrec *recs, int cap,
                                  int *nmir) {
  if (ni < 1 || nj < 1 || nk < 1 || !xv || !yv || !zv || !blank || !nmir || ntri < 0 || (ntri > 0 && !tri))
    return MGE_ERR_ARG;
  const int il = ni + 1, jl = nj + 1, kl = nk + 1;
  std::vector<double> cx, cy, cz;
  centres(ni, nj, nk, xv, yv, zv, cx, cy, cz);
  auto bidx = [&](int i, int j, int k) -> long {
    return (long)(i - 2) + (long)ni * ((long)(j - 2) + (long)nj * (long)(k - 2));
  };
  for (long q = 0; q < (long)ni * nj * nk; q++) blank[q] = 1;
  if (ntri == 0) {
    *nmir = 0;
    return MGE_OK;
  }
  TriSurf S;
  S.build(tri, ntri);
  double hmin = 1e300;
  for (int i = 2; i <= il; i++) hmin = std::min(hmin, xv[i + 1] - xv[i]);
  for (int j = 2; j <= jl; j++) hmin = std::min(hmin, yv[j + 1] - yv[j]);
  for (int k = 2; k <= kl; k++) hmin = std::min(hmin, zv[k + 1] - zv[k]);
  // classification along x lines: triangles bucketed by the (j,k) lines
  // their y-z bounding box covers
  std::vector<std::vector<int>> lines((size_t)nj * nk);
  auto lo_idx = [](const std::vector<double> &c, int a, int b, double v) {
    int q = a;
    while (q <= b && c[q] < v) q++;
    return q;
  };
  for (int m = 0; m < ntri; m++) {
    const double *T = tri + 9 * m;
    double y0 = std::min(std::min(T[1], T[4]), T[7]), y1 = std::max(std::max(T[1], T[4]), T[7]);
    double z0 = std::min(std::min(T[2], T[5]), T[8]), z1 = std::max(std::max(T[2], T[5]), T[8]);
    int ja = lo_idx(cy, 2, jl, y0 - 1e-6 * hmin), ka = lo_idx(cz, 2, kl, z0 - 1e-6 * hmin);
    for (int k = ka; k <= kl && cz[k] <= z1 + 1e-6 * hmin; k++)
      for (int j = ja; j <= jl && cy[j] <= y1 + 1e-6 * hmin; j++) lines[(size_t)(k - 2) * nj + (j - 2)].push_back(m);
  }
  std::vector<double> xs;
  for (int k = 2; k <= kl; k++)
    for (int j = 2; j <= jl; j++) {
      const auto &L = lines[(size_t)(k - 2) * nj + (j - 2)];
      if (L.empty()) continue;
      double p[3] = {0.0, cy[j] + 1.0e-9 * hmin, cz[k] + 1.3e-9 * hmin};
      xs.clear();
      for (int m : L) {
        double sp;
        if (line_tri(tri + 9 * m, 0, p, &sp)) xs.push_back(sp);
      }
      std::sort(xs.begin(), xs.end());
      size_t c = 0;
      for (int i = 2; i <= il; i++) {
        while (c < xs.size() && xs[c] < cx[i]) c++;
        blank[bidx(i, j, k)] = (c % 2 == 1) ? 0 : 1;
      }
    }
  return build_records_common(ni, nj, nk, cx, cy, cz, blank, recs, cap, nmir,
                              [&](const double *xc, const double *xsd, double *ps, double *n, double *rc) {
                                double t;
                                if (!tri_crossing(S, xc, xsd, hmin, &t, n)) {
                                  t = 0.5;
                                  for (int d = 0; d < 3; d++) n[d] = (xc[d] - xsd[d]);
                                  double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                                  for (int d = 0; d < 3; d++) n[d] /= nn;
                                }
                                for (int d = 0; d < 3; d++) ps[d] = xc[d] + t * (xsd[d] - xc[d]);
                                rc[0] = rc[1] = 0.0;
                              });
}

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
  if (!q.valid) {
    *nmir = 0;
    return MGE_OK;
  }
  return build_records_common(ni, nj, nk, cx, cy, cz, blank, recs, cap, nmir,
                              [&](const double *xc, const double *xs, double *ps, double *n, double *rc) {
                                double t = crossing(q, xc, xs);
                                for (int d = 0; d < 3; d++) ps[d] = xc[d] + t * (xs[d] - xc[d]);
                                normal_at(q, ps, n);
                                rc[0] = rc[1] = 0.0;
                                if (body->kind == MGE_BODY_SPHERE) {
                                  rc[0] = rc[1] = 1.0 / body->semi[0];
                                } else if (body->kind == MGE_BODY_CYLINDER) {
                                  rc[0] = 1.0 / body->semi[0];
                                }
                              });
}
