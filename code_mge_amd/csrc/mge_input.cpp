// mge_input.cpp -- the reference's input files for the solve step: the
// jobsheet, the .pin grid description and the PLOT3D surface (.p3d).
//
// The reference reads them in make_fsijob (srcf90/main.f90:44) through the
// ctrlib/jobpak libraries, which are not in the reference tree; the key ->
// variable mapping below is therefore this library's own:
//  - jobsheet: sections "NAME { key = value ... }", names and keys
//    case-insensitive, '!' starts a comment, list values separated by ':'.
//    CONTROL: job, model, iflow (1: start from <job>_restart.dat,
//    srcf90/main.f90:78), mcyc (one entry per multigrid level, coarsest
//    first: the cycles of the FMG ramp, srcf90/solver.f90:119-261),
//    runtype (execmode: 0 steady, 1 unsteady), period and step (the
//    physical time loop of solver.f90:110-112), dt (the physical time step
//    in seconds: the reference's fsi_dt, set by its absent FSI job library);
//    FLOW: mach, reynolds, gamma, prlmean, alphad, betad, pmean, tmean,
//    rgasmean (the freestream speed that makes dt nondimensional,
//    input_param.f90:51,92);
//    NUMERICS: cflf (negative: local time step), vis2, vis4, smoo, mgtype,
//    cstp, cdis, eqntype, distype, dissip, epsentr;
//    CFD: file (the .p3d surface); MESH: file (the .pin);
//    DOMAIN: xbound/ybound/zbound = (lo, hi) each followed by its
//    bc_type = (lo, hi) pair.
//  - .pin: BEGIN GRID DATA, then per grid GRID.NAME, POINTS (vertex counts),
//    ORIGIN, LENGTH; '#' and '!' comment lines.  One grid: the finest
//    level; coarser levels halve every direction that keeps >= 2 cells.
//    (Embedded hierarchies of several grids are not handled here.)
//  - .p3d: PLOT3D whole-format ASCII: block count, the dimensions of every
//    block, then X, Y and Z of each block, i fastest; every block is a
//    surface patch (one dimension 1) whose quads are split in two triangles.
//  - a block face lying on a DOMAIN bound takes that bound's bc_type; any
//    other face is far field (-20).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mge.h"

namespace {

std::string g_msg;

int fail(int code, const std::string &msg) {
  g_msg = msg;
  return code;
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool read_lines(const std::string &path, std::vector<std::string> &out) {
  std::ifstream f(path);
  if (!f) return false;
  std::string l;
  while (std::getline(f, l)) out.push_back(l);
  return true;
}

// strip a '!' comment that is not inside a quoted string
std::string strip_comment(const std::string &l) {
  bool q = false;
  for (size_t i = 0; i < l.size(); i++) {
    if (l[i] == '"') q = !q;
    if (l[i] == '!' && !q) return l.substr(0, i);
  }
  return l;
}

struct Section {
  std::string name;
  std::vector<std::pair<std::string, std::string>> kv;  // in file order
  const std::string *get(const std::string &k) const {
    for (auto &p : kv)
      if (p.first == k) return &p.second;
    return nullptr;
  }
};

// "key = value" items of one line: value is a quoted string, a
// parenthesised group or a bare token; items separated by blanks or commas
void parse_items(const std::string &l, std::vector<std::pair<std::string, std::string>> &kv) {
  size_t i = 0, n = l.size();
  while (i < n) {
    while (i < n && (std::isspace((unsigned char)l[i]) || l[i] == ',')) i++;
    size_t k0 = i;
    while (i < n && l[i] != '=' && !std::isspace((unsigned char)l[i]) && l[i] != ',') i++;
    std::string key = lower(l.substr(k0, i - k0));
    while (i < n && std::isspace((unsigned char)l[i])) i++;
    if (i >= n || l[i] != '=' || key.empty()) {
      while (i < n && l[i] != ',' && !std::isspace((unsigned char)l[i])) i++;
      continue;
    }
    i++;
    while (i < n && std::isspace((unsigned char)l[i])) i++;
    std::string val;
    if (i < n && l[i] == '"') {
      size_t e = l.find('"', i + 1);
      if (e == std::string::npos) e = n;
      val = l.substr(i + 1, e - i - 1);
      i = e + 1;
    } else if (i < n && l[i] == '(') {
      size_t e = l.find(')', i + 1);
      if (e == std::string::npos) e = n;
      val = l.substr(i + 1, e - i - 1);
      i = e + 1;
    } else if (i < n && l[i] == '\'') {
      size_t e = l.find('\'', i + 1);
      if (e == std::string::npos) e = n;
      val = l.substr(i + 1, e - i - 1);
      i = e + 1;
    } else {
      size_t v0 = i;
      while (i < n && l[i] != ',' && !std::isspace((unsigned char)l[i])) i++;
      val = l.substr(v0, i - v0);
    }
    kv.push_back({key, trim(val)});
  }
}

bool parse_jobsheet(const std::string &path, std::vector<Section> &secs) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return false;
  Section *cur = nullptr;
  std::string pending;
  for (auto &raw : lines) {
    std::string l = trim(strip_comment(raw));
    if (l.empty()) continue;
    if (!cur) {
      if (l == "{") {
        if (!pending.empty()) {
          secs.push_back(Section{lower(pending), {}});
          cur = &secs.back();
        }
        continue;
      }
      // a section name, possibly with the brace on the same line
      std::string name = l;
      bool brace = false;
      if (!name.empty() && name.back() == '{') {
        name = trim(name.substr(0, name.size() - 1));
        brace = true;
      }
      pending = name;
      if (brace) {
        secs.push_back(Section{lower(pending), {}});
        cur = &secs.back();
      }
      continue;
    }
    if (l == "}") {
      cur = nullptr;
      pending.clear();
      continue;
    }
    parse_items(l, cur->kv);
  }
  return true;
}

const Section *find_sec(const std::vector<Section> &secs, const std::string &name) {
  for (auto &s : secs)
    if (s.name == name) return &s;
  return nullptr;
}

std::vector<double> list_of(const std::string &v) {
  std::vector<double> out;
  std::string t = v;
  for (auto &c : t)
    if (c == ':' || c == ',') c = ' ';
  std::istringstream is(t);
  double x;
  while (is >> x) out.push_back(x);
  return out;
}

bool num(const Section *s, const char *key, double *out) {
  if (!s) return false;
  const std::string *v = s->get(key);
  if (!v) return false;
  char *end = nullptr;
  double x = std::strtod(v->c_str(), &end);
  if (end == v->c_str()) return false;
  *out = x;
  return true;
}

std::string dir_of(const std::string &path) {
  size_t p = path.find_last_of('/');
  return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

std::string join(const std::string &dir, const std::string &f) {
  if (!f.empty() && f[0] == '/') return f;
  return dir + "/" + f;
}

void copy_str(char *dst, size_t cap, const std::string &s) {
  std::strncpy(dst, s.c_str(), cap - 1);
  dst[cap - 1] = 0;
}

}  // namespace

extern "C" {

const char *mge_input_error(void) { return g_msg.c_str(); }

int mge_read_pin(const char *path, int cap, mge_pin_grid *grids, int *ngrid) {
  if (!path || !ngrid) return fail(MGE_ERR_ARG, "null argument");
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return fail(MGE_ERR_ARG, std::string("cannot read ") + path);
  std::vector<mge_pin_grid> gs;
  for (auto &raw : lines) {
    std::string l = trim(strip_comment(raw));
    if (l.empty() || l[0] == '#') continue;
    std::istringstream is(l);
    std::string w;
    is >> w;
    std::string W = lower(w);
    if (W == "grid.name") {
      mge_pin_grid g;
      std::memset(&g, 0, sizeof(g));
      std::string eq;
      is >> eq >> g.name;
      gs.push_back(g);
    } else if (!gs.empty() && W == "points") {
      is >> gs.back().points[0] >> gs.back().points[1] >> gs.back().points[2];
    } else if (!gs.empty() && W == "origin") {
      is >> gs.back().origin[0] >> gs.back().origin[1] >> gs.back().origin[2];
    } else if (!gs.empty() && W == "length") {
      is >> gs.back().length[0] >> gs.back().length[1] >> gs.back().length[2];
    }
  }
  *ngrid = (int)gs.size();
  if (grids)
    for (int g = 0; g < std::min(cap, (int)gs.size()); g++) grids[g] = gs[g];
  return MGE_OK;
}

int mge_read_p3d(const char *path, int cap, double *tri, int *ntri) {
  if (!path || !ntri) return fail(MGE_ERR_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(MGE_ERR_ARG, std::string("cannot read ") + path);
  int nb = 0;
  if (!(f >> nb) || nb < 1) return fail(MGE_ERR_ARG, "p3d: bad block count");
  std::vector<int> dims(3 * nb);
  for (int b = 0; b < nb; b++)
    if (!(f >> dims[3 * b] >> dims[3 * b + 1] >> dims[3 * b + 2])) return fail(MGE_ERR_ARG, "p3d: bad dimensions");
  int count = 0;
  for (int b = 0; b < nb; b++) {
    int a0 = dims[3 * b], a1 = dims[3 * b + 1], a2 = dims[3 * b + 2];
    long np = (long)a0 * a1 * a2;
    std::vector<double> X(np), Y(np), Z(np);
    for (long q = 0; q < np; q++)
      if (!(f >> X[q])) return fail(MGE_ERR_ARG, "p3d: short X");
    for (long q = 0; q < np; q++)
      if (!(f >> Y[q])) return fail(MGE_ERR_ARG, "p3d: short Y");
    for (long q = 0; q < np; q++)
      if (!(f >> Z[q])) return fail(MGE_ERR_ARG, "p3d: short Z");
    // the two non-unit dimensions span the patch
    int ext[3] = {a0, a1, a2};
    long st[3] = {1, (long)a0, (long)a0 * a1};
    int u = -1, v = -1;
    for (int d = 0; d < 3; d++)
      if (ext[d] > 1) {
        if (u < 0) u = d;
        else if (v < 0) v = d;
        else return fail(MGE_ERR_ARG, "p3d: a block is not a surface patch");
      }
    if (u < 0 || v < 0) continue;
    for (int jv = 0; jv + 1 < ext[v]; jv++)
      for (int iu = 0; iu + 1 < ext[u]; iu++) {
        long p00 = iu * st[u] + jv * st[v], p10 = p00 + st[u], p01 = p00 + st[v], p11 = p10 + st[v];
        const long tris[2][3] = {{p00, p10, p11}, {p00, p11, p01}};
        for (int t = 0; t < 2; t++) {
          if (tri && count < cap)
            for (int c = 0; c < 3; c++) {
              tri[9 * count + 3 * c] = X[tris[t][c]];
              tri[9 * count + 3 * c + 1] = Y[tris[t][c]];
              tri[9 * count + 3 * c + 2] = Z[tris[t][c]];
            }
          count++;
        }
      }
  }
  *ntri = count;
  return MGE_OK;
}

int mge_read_job(const char *path, mge_job *job) {
  if (!path || !job) return fail(MGE_ERR_ARG, "null argument");
  std::vector<Section> secs;
  if (!parse_jobsheet(path, secs)) return fail(MGE_ERR_ARG, std::string("cannot read ") + path);
  std::memset(job, 0, sizeof(*job));
  const Section *ctl = find_sec(secs, "control"), *flow = find_sec(secs, "flow"),
                *nm = find_sec(secs, "numerics"), *cfd = find_sec(secs, "cfd"), *mesh = find_sec(secs, "mesh"),
                *dom = find_sec(secs, "domain");
  if (!ctl || !flow || !nm || !mesh) return fail(MGE_ERR_ARG, "jobsheet: CONTROL, FLOW, NUMERICS and MESH are required");
  if (const std::string *v = ctl->get("job")) copy_str(job->job, sizeof(job->job), *v);
  if (const std::string *v = ctl->get("model")) copy_str(job->model, sizeof(job->model), *v);
  {
    double fl = 0.0, rt = 0.0, per = 1.0, stp = 1.0;
    num(ctl, "iflow", &fl);
    num(ctl, "runtype", &rt);
    num(ctl, "period", &per);
    num(ctl, "step", &stp);
    num(ctl, "dt", &job->dt);
    job->iflow = (int)fl;
    job->execmode = (int)rt;
    job->nperiod = (int)per;
    job->nstep = (int)stp;
  }
  const std::string *mc = ctl->get("mcyc");
  if (!mc) return fail(MGE_ERR_ARG, "jobsheet: mcyc missing");
  std::vector<double> m = list_of(*mc);
  if (m.empty() || m.size() > 16) return fail(MGE_ERR_ARG, "jobsheet: mcyc needs 1..16 levels");
  job->nlevel = (int)m.size();
  for (size_t l = 0; l < m.size(); l++) job->mcyc[l] = (int)m[l];
  // flow and numerics (defaults where a key is absent)
  job->mach = 0.2;
  job->reynolds = 500.0;
  job->gamma = 1.4;
  job->prlmean = 0.72;
  num(flow, "mach", &job->mach);
  num(flow, "reynolds", &job->reynolds);
  num(flow, "gamma", &job->gamma);
  num(flow, "prlmean", &job->prlmean);
  num(flow, "alphad", &job->alphad);
  num(flow, "betad", &job->betad);
  job->pmean = 101325.0;
  job->tmean = 288.15;
  job->rgasmean = 287.0;
  num(flow, "pmean", &job->pmean);
  num(flow, "tmean", &job->tmean);
  num(flow, "rgasmean", &job->rgasmean);
  job->cflf = -2.5;
  job->vis2 = 0.5;
  job->vis4 = 1.0;
  job->smoo = 0.0;
  job->epsentr = 0.05;
  double mg = 2.0;
  num(nm, "cflf", &job->cflf);
  num(nm, "vis2", &job->vis2);
  num(nm, "vis4", &job->vis4);
  num(nm, "smoo", &job->smoo);
  num(nm, "mgtype", &mg);
  num(nm, "epsentr", &job->epsentr);
  job->mgtype = (int)mg;
  std::vector<double> cs = {0.25, 0.16667, 0.375, 0.5, 1.0}, cd = {1.0, 0.0, 0.56, 0.0, 0.44};
  if (const std::string *v = nm->get("cstp")) cs = list_of(*v);
  if (const std::string *v = nm->get("cdis")) cd = list_of(*v);
  if (cs.empty() || cs.size() > 8 || cd.size() != cs.size())
    return fail(MGE_ERR_ARG, "jobsheet: cstp and cdis need the same 1..8 entries");
  job->mstage = (int)cs.size();
  for (size_t s = 0; s < cs.size(); s++) {
    job->cstp[s] = cs[s];
    job->cdis[s] = cd[s];
  }
  auto word = [&](const char *k, const char *def) {
    const std::string *v = nm->get(k);
    return lower(v ? *v : std::string(def));
  };
  std::string eq = word("eqntype", "e"), dt = word("distype", "c"), ds = word("dissip", "scalar");
  job->eqntype = (eq == "n") ? 1 : 0;
  job->distype = (dt == "r1") ? 1 : (dt == "r2" ? 2 : 0);
  job->dissip = (ds == "matrix") ? 1 : 0;
  if (eq != "e" && eq != "n") return fail(MGE_ERR_ARG, "jobsheet: eqntype must be e or n");
  if (dt != "c" && dt != "r1" && dt != "r2") return fail(MGE_ERR_ARG, "jobsheet: distype must be c, r1 or r2");
  // files
  std::string dir = dir_of(path);
  const std::string *pin = mesh->get("file");
  if (!pin) return fail(MGE_ERR_ARG, "jobsheet: MESH file missing");
  copy_str(job->mesh, sizeof(job->mesh), join(dir, *pin));
  if (cfd)
    if (const std::string *v = cfd->get("file")) copy_str(job->surface, sizeof(job->surface), join(dir, *v));
  // grid
  int ng = 0;
  mge_pin_grid g;
  {
    int rc = mge_read_pin(job->mesh, 1, &g, &ng);
    if (rc) return rc;
    if (ng != 1)
      return fail(MGE_ERR_ARG, "pin: " + std::to_string(ng) + " grids; only a single-block grid is handled");
  }
  for (int d = 0; d < 3; d++) {
    job->n[d] = g.points[d] - 1;
    job->lo[d] = g.origin[d];
    job->hi[d] = g.origin[d] + g.length[d];
    if (job->n[d] < 1) return fail(MGE_ERR_ARG, "pin: fewer than 2 points in a direction");
  }
  for (int f = 0; f < 6; f++) job->bctype[f] = -20;
  if (dom) {
    int axis = -1;
    const char *bn[3] = {"xbound", "ybound", "zbound"};
    for (auto &p : dom->kv) {
      for (int d = 0; d < 3; d++)
        if (p.first == bn[d]) {
          axis = d;
          std::vector<double> b = list_of(p.second);
          if (b.size() == 2) {
            job->dom_lo[d] = b[0];
            job->dom_hi[d] = b[1];
          }
        }
      if (p.first == "bc_type" && axis >= 0) {
        std::vector<double> b = list_of(p.second);
        if (b.size() == 2) {
          double tol = 1e-9 * (job->hi[axis] - job->lo[axis]);
          if (std::fabs(job->lo[axis] - job->dom_lo[axis]) <= tol) job->bctype[axis] = (int)b[0];
          if (std::fabs(job->hi[axis] - job->dom_hi[axis]) <= tol) job->bctype[3 + axis] = (int)b[1];
        }
        axis = -1;
      }
    }
  }
  return MGE_OK;
}

// parameters of the job in the order input_param.f90:13-115 derives them
// (rho0 = 1, p0 = 1/(gamma M^2), |U| = 1; angles from degrees)
int mge_job_params(const mge_job *job, int precision, mge_params *P) {
  if (!job || !P) return fail(MGE_ERR_ARG, "null argument");
  std::memset(P, 0, sizeof(*P));
  const double g = job->gamma, rm = job->mach;
  const double alpha = job->alphad * (M_PI / 180.0), beta = job->betad * (M_PI / 180.0);
  P->mstage = job->mstage;
  for (int s = 0; s < job->mstage; s++) {
    P->cstp[s] = job->cstp[s];
    P->cdis[s] = job->cdis[s];
  }
  P->vt = job->cflf < 0.0 ? 0 : 1;
  P->cfl = std::fabs(job->cflf);
  P->vis2 = job->vis2;
  P->vis4 = job->vis4 / 32.0;
  P->smoo = job->smoo;
  P->mgtype = job->mgtype;
  P->gamma = g;
  P->rho0 = 1.0;
  P->p0 = 1.0 / (g * (rm * rm));
  P->c0 = std::sqrt(g * P->p0 / P->rho0);
  P->ei0 = P->p0 / ((g - 1.0) * P->rho0);
  P->u0 = rm * P->c0 * std::cos(beta) * std::cos(alpha);
  P->v0 = rm * P->c0 * std::cos(beta) * std::sin(alpha);
  P->w0 = rm * P->c0 * std::sin(beta);
  P->h0 = g * P->ei0 + 0.5 * (P->u0 * P->u0 + P->v0 * P->v0 + P->w0 * P->w0);
  P->eqntype = job->eqntype;
  P->dissip = job->dissip;
  P->viscl = job->eqntype == 1 ? P->rho0 * P->u0 * 1.0 / job->reynolds : 0.0;
  P->prlmean = job->prlmean;
  P->gas_const = 1.0 / g;
  P->drop = 1.0e6;
  P->distype = job->distype;
  P->precision = precision;
  P->epsentr = job->epsentr;
  // physical time step made nondimensional by ref_leng/vel_fs (ref_leng = 1),
  // vel_fs = M sqrt(gamma pmean/rhomean), rhomean = pmean/(R tmean); laminar
  // unsteady flow takes u0/ref_leng/5 instead (input_param.f90:51,81-96)
  P->execmode = job->execmode;
  const double rhomean = job->pmean / (job->rgasmean * job->tmean);
  const double vel_fs = rm * std::sqrt(g * job->pmean / rhomean);
  P->dtstep = job->dt * vel_fs / 1.0;
  if (job->eqntype == 1 && job->execmode == 1) P->dtstep = P->u0 / 1.0 / 5.0;
  P->smoother = 0;  // euler.f90:376 calls psmoo_old
  P->ffreeze = 0.0;  // bc_far.f90's rule at every call
  P->lfreeze = 0.0;  // the limiter at every stage
  return MGE_OK;
}

// cells and coarsening ratio of level lvl (1 = coarsest) of the job's
// hierarchy: from the finest grid, halve each direction while it stays
// even and keeps >= 2 cells
int mge_job_level(const mge_job *job, int lvl, int n[3], int ratio[3]) {
  if (!job || lvl < 1 || lvl > job->nlevel) return fail(MGE_ERR_ARG, "bad level");
  std::vector<std::vector<int>> dims;
  int c[3] = {job->n[0], job->n[1], job->n[2]};
  for (int l = 0; l < job->nlevel; l++) {
    dims.push_back({c[0], c[1], c[2]});
    for (int d = 0; d < 3; d++) c[d] = (c[d] % 2 == 0 && c[d] / 2 >= 2) ? c[d] / 2 : c[d];
  }
  // dims[0] finest ... dims[nlevel-1] coarsest
  const std::vector<int> &me = dims[job->nlevel - lvl];
  for (int d = 0; d < 3; d++) n[d] = me[d];
  if (lvl > 1) {
    const std::vector<int> &co = dims[job->nlevel - lvl + 1];
    for (int d = 0; d < 3; d++) ratio[d] = me[d] / co[d];
  } else {
    for (int d = 0; d < 3; d++) ratio[d] = 2;
  }
  return MGE_OK;
}

}  // extern "C"
