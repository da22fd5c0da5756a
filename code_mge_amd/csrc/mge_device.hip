// mge_device.hip -- host side of the MI355X solve step and the C ABI of
// include/mge.h.  The kernels live in mge_stage.h (JST stage, wall records),
// mge_roe.h (Roe / MUSCL stage), mge_smooth.h (residual smoothing) and
// mge_bc.h (time step, norms, boundary conditions, transfers, forces); all
// are templated on the bulk type T and instantiated for double (FP64 case)
// and float (FP32 case).
//
// One Cartesian block per multigrid level lives in HBM for the whole solve:
//   S[3]     three state buffers (w: 5 components, then p), rotated across
//            RK stages so W^(0) (the reference's wn) is never copied
//   dw, fw   residual and blended dissipation (5 components each)
//   ws,wr,qf multigrid saved state, restricted residual, forcing
//   radsum   radi+radj+radk, vdtim: local time step (interior cells)
//   flags    bit0 fluid (blank != 0), bit1 touched by a wall record
// Bulk arrays hold T; the wall mirror states, residual-norm partials, the
// global time step and the force terms are FP64.  All fields use the
// reference numbering (0:ib,0:jb,0:kb), i fastest, component-major, so
// host<->device copies need no reshuffling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/mge.h"
#include "mge_bc.h"
#include "mge_output.h"
#include "mge_roe.h"
#include "mge_smooth.h"
#include "mge_stage.h"
#include "mge_visc.h"

using namespace mge;

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int set_err(int code, const char *msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      char b_[256];                                                        \
      snprintf(b_, sizeof b_, "%s: %s", #x, hipGetErrorString(e_));        \
      return set_err(MGE_ERR_HIP, b_);                                     \
    }                                                                      \
  } while (0)

#define CHKRC(x)            \
  do {                      \
    int r_ = (x);           \
    if (r_ < 0) return r_;  \
  } while (0)

// ================================================================== host state

namespace {

// Bulk arrays are untyped here: they hold T = double or float according to
// the case's precision and are cast by the typed views below.
struct Level {
  bool defined = false;
  Dims d{};
  double cellvolume = 0.0;
  int bctype[6] = {0, 0, 0, 0, 0, 0};
  int ratio[3] = {2, 2, 2};
  void *S[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  void *ws = nullptr, *wr = nullptr, *qf = nullptr, *dw = nullptr, *fw = nullptr;
  void *radsum = nullptr, *vdtim = nullptr;
  void *w1 = nullptr, *w2 = nullptr;  // unsteady: physical time levels n, n-1 (5E each)
  void *corr = nullptr;               // levels below the finest: correction w - ws for movfin (5E)
  void *eps = nullptr;                // psmoo (smoother 1): coefficients x, y, z, then the sweep scratch (4E)
  // stretched tensor level (cellvolume 0): per-cell volumes (E), the vertex
  // coordinates, and the 1-D weight pairs of the transfers from and to the
  // next coarser level (restriction per coarse index, prolongation per index)
  void *vol = nullptr;
  std::vector<double> hv[3];
  uint8_t *ffdec = nullptr;  // far-field branch per ghost point of the six faces (ffreeze > 0)
  unsigned long long *lim = nullptr;  // limiter words per direction and cell (Roe r2, lfreeze > 0)
  long ffoff[7] = {0, 0, 0, 0, 0, 0, 0};
  void *rw[3] = {nullptr, nullptr, nullptr}, *pw[3] = {nullptr, nullptr, nullptr};
  uint8_t *flags = nullptr, *marker = nullptr;
  void *dx = nullptr, *dy = nullptr, *dz = nullptr;
  void *piv = nullptr;  // e1x,e2x,e1y,e2y,e1z,e2z,1/e1x,1/e1y,1/e1z each npiv
  void *fax = nullptr;  // -bb/e1x
  int npiv = 0;
  int nmir = 0, nupd = 0, zero_at = -1;
  bool cut_interior = false;  // every cloud cell inside the ghost shell (bcond writes none of them)
  bool corner_stale = false;  // the current state's edge and corner ghost cells await bc_corner
  DevRec *recs = nullptr;
  double *wsol = nullptr, *psol = nullptr;
  std::vector<mge_cutrec> hrecs;
  int ntouch = 0, nops = 0;
  long *tcell = nullptr;
  int *beg = nullptr;
  int2 *ops = nullptr;
  int2 *chunks = nullptr;  // k_rec's workgroups (touched-cell ranges)
  int nchunk = 0;
  double *opval = nullptr;
  double *partial = nullptr;
  long npartial = 0;
  unsigned long long *dtim_bits = nullptr;
  double *dtim = nullptr;
  double *rtrms = nullptr;
  double *force = nullptr;
  int kc = 0;                   // autotuned k planes per stage workgroup (0: fill rule)
  int tj = 0;                   // autotuned tile rows of the stage kernels (0: the precision's default)
  void *vprim = nullptr;        // Navier-Stokes: velocity and temperature per cell (4E)
  int nvcell = 0;               // cut cells with wall records (viscous wall terms)
  long *vcell = nullptr;
  int *vbeg = nullptr, *vrec = nullptr;
  int mxcyc = 1;
  std::vector<void *> allocs;
  size_t bytes = 0;  // device bytes of the level
};

struct Timer {
  std::string name;
  hipEvent_t a, b;
  double bytes;
  int level;  // level of the launch (0: not level-specific)
};

// one recorded launch of a case (batch programs): both entry points of the
// kernel, the case's own grid and its argument struct
struct Rec {
  const char *name;
  const void *fs, *fb;
  dim3 grid, block;
  std::vector<char> args;
};

struct TStat {
  double ms = 0.0;
  long long count = 0;
  double bytes = 0.0;
};

}  // namespace

struct mge_case {
  mge_params prm{};
  int dev = 0;            // HIP device of the case's stream
  int *ffrozen = nullptr;     // device flags: [0] far-field branches frozen (ffreeze > 0),
                              // [1] limiter choices frozen (lfreeze > 0)
  double *ffini = nullptr;    // first finest-cycle residual for that test (device)
  double *scratch = nullptr;  // that device's address of the fixed staging array
  bool fp32 = false;
  size_t es = 8;  // bytes per bulk value
  int nlevel = 0;
  std::vector<Level> lv;
  hipStream_t stream = nullptr;
  int *err = nullptr;
  bool timing = false;
  std::vector<Timer> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, TStat> tstat;
  double next_bytes = 0.0;  // algorithmic bytes of the next timed launch
  int cur_level = 0;        // level the current operation works on (timer statistics)
  double smoo_bb = 0.0;
  std::vector<Rec> *rec = nullptr;  // when set, launches are recorded, not run
  hipError_t launch_err = hipSuccess;  // first failed launch, reported by check_err
  int itime = 0;            // physical time level of the current euler call (unsteady)
  bool lazy_corners = true; // stages may leave bc_corner until it is read (off inside a composite)
};

namespace {

// ---- host-synchronous copies and clears ---------------------------------
// On the calling thread's own default stream, after the stream whose work
// they must follow (if any): the legacy null stream would implicitly
// synchronise with every blocking stream and so join -- and break -- a
// batch's graph capture running on another thread.
hipError_t h2d_sync(void *dst, const void *src, size_t n, hipStream_t after = nullptr) {
  hipError_t e = after ? hipStreamSynchronize(after) : hipSuccess;
  if (e == hipSuccess) e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, hipStreamPerThread);
  return e == hipSuccess ? hipStreamSynchronize(hipStreamPerThread) : e;
}
hipError_t d2h_sync(void *dst, const void *src, size_t n, hipStream_t after = nullptr) {
  hipError_t e = after ? hipStreamSynchronize(after) : hipSuccess;
  if (e == hipSuccess) e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, hipStreamPerThread);
  return e == hipSuccess ? hipStreamSynchronize(hipStreamPerThread) : e;
}
hipError_t fill_sync(void *dst, int byte, size_t n) {
  hipError_t e = hipMemsetAsync(dst, byte, n, hipStreamPerThread);
  return e == hipSuccess ? hipStreamSynchronize(hipStreamPerThread) : e;
}
hipError_t clear_sync(void *dst, size_t n) { return fill_sync(dst, 0, n); }

int dalloc_bytes(Level &L, void **p, size_t bytes) {
  void *q = nullptr;
  if (hipMalloc(&q, bytes + 64) != hipSuccess) return set_err(MGE_ERR_NOMEM, "hipMalloc failed");
  if (clear_sync(q, bytes + 64) != hipSuccess) {
    hipFree(q);
    return set_err(MGE_ERR_HIP, "hipMemset failed");
  }
  L.allocs.push_back(q);
  L.bytes += bytes + 64;
  *p = q;
  return MGE_OK;
}

template <class U>
int dalloc(Level &L, U **p, size_t n) {
  void *q = nullptr;
  CHKRC(dalloc_bytes(L, &q, n * sizeof(U)));
  *p = (U *)q;
  return MGE_OK;
}

void level_free(Level &L) {
  for (void *p : L.allocs) hipFree(p);
  L.allocs.clear();
  L = Level();
}

hipEvent_t ev_get(mge_case *h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

void tic(mge_case *h, const char *name, Timer *t) {
  t->bytes = h->next_bytes;
  h->next_bytes = 0.0;
  if (!h->timing) return;
  t->name = name;
  t->level = h->cur_level;
  t->a = ev_get(h);
  t->b = ev_get(h);
  hipEventRecord(t->a, h->stream);
}
void toc(mge_case *h, Timer *t) {
  if (!h->timing) return;
  hipEventRecord(t->b, h->stream);
  h->pending.push_back(*t);
}
void resolve_timers(mge_case *h) {
  for (auto &t : h->pending) {
    hipEventSynchronize(t.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, t.a, t.b);
    // per kernel, and per kernel and level ("name@level")
    for (int v = 0; v < 2; v++) {
      if (v == 1 && t.level <= 0) break;
      auto &s = h->tstat[v == 0 ? t.name : t.name + "@" + std::to_string(t.level)];
      s.ms += ms;
      s.count += 1;
      s.bytes += t.bytes;
    }
    h->pool.push_back(t.a);
    h->pool.push_back(t.b);
  }
  h->pending.clear();
}

// Launch one kernel of the cycle for one case (fs: single-case entry point,
// fb: batch entry point), or record it when the case is being recorded
// into a batch program.  A grid with no blocks is a launch the case skips
// (no wall records, an inactive face); it is recorded so that every case
// of a batch records the same sequence.
template <class A>
void launch(mge_case *h, const char *name, const void *fs, const void *fb, dim3 g, dim3 b, const A &a) {
  if (h->rec) {
    Rec r;
    r.name = name;
    r.fs = fs;
    r.fb = fb;
    r.grid = g;
    r.block = b;
    r.args.assign((const char *)&a, (const char *)&a + sizeof(A));
    h->rec->push_back(std::move(r));
    h->next_bytes = 0.0;
    return;
  }
  if ((long)g.x * g.y * g.z == 0) {
    h->next_bytes = 0.0;
    return;
  }
  Timer t;
  tic(h, name, &t);
  void *p[1] = {(void *)&a};
  hipError_t e = hipLaunchKernel(fs, g, b, p, 0, h->stream);
  if (e != hipSuccess && h->launch_err == hipSuccess) h->launch_err = e;
  toc(h, &t);
}
#define LAUNCHT(h, name, K, T, g, b, a) launch(h, name, (const void *)&K<T>, (const void *)&K##_b<T>, g, b, a)
#define LAUNCHN(h, name, K, g, b, a) launch(h, name, (const void *)&K, (const void *)&K##_b, g, b, a)

// typed views of a level's bulk arrays
template <class T>
T *tp(void *p) {
  return static_cast<T *>(p);
}
template <class T>
Geo<T> geo(const Level &L) {
  return Geo<T>{tp<T>(L.dx), tp<T>(L.dy), tp<T>(L.dz)};
}
template <class T>
StateC<T> stc(const Level &L, int b) {
  return StateC<T>{tp<T>(L.S[b]), tp<T>(L.S[b]) + 5 * L.d.E};
}
template <class T>
State<T> stm(const Level &L, int b) {
  return State<T>{tp<T>(L.S[b]), tp<T>(L.S[b]) + 5 * L.d.E};
}

// freestream of the case in the bulk type (the FP32 state is stored as the
// deviation from it)
template <class T>
Inf<T> make_inf(const mge_params &P) {
  Inf<T> f;
  f.w[0] = (T)P.rho0;
  f.w[1] = (T)(P.rho0 * P.u0);
  f.w[2] = (T)(P.rho0 * P.v0);
  f.w[3] = (T)(P.rho0 * P.w0);
  f.w[4] = (T)(P.rho0 * P.h0 - P.p0);
  f.p = (T)P.p0;
  return f;
}

// run f with the case's bulk type: f(double{}) or f(float{})
template <class F>
int by_type(mge_case *h, F f) {
  return h->fp32 ? f(float{}) : f(double{});
}

dim3 cell_grid(const Dims &d) { return dim3((d.ni + 63) / 64, (d.nj + 3) / 4, d.nk); }
dim3 cell_block() { return dim3(64, 4, 1); }
dim3 flat_grid(long n) { return dim3((unsigned)((n + 255) / 256)); }

Level *get_level(mge_case *h, int lvl) {
  if (!h || lvl < 1 || lvl > h->nlevel) return nullptr;
  Level *L = &h->lv[lvl];
  return L->defined ? L : nullptr;
}

int check_err(mge_case *h) {
  if (h->launch_err != hipSuccess) {
    char b[256];
    snprintf(b, sizeof b, "kernel launch failed: %s", hipGetErrorString(h->launch_err));
    h->launch_err = hipSuccess;
    return set_err(MGE_ERR_HIP, b);
  }
  int e = 0;
  if (hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return set_err(MGE_ERR_HIP, "error-flag copy failed");
  if (hipStreamSynchronize(h->stream) != hipSuccess) return set_err(MGE_ERR_HIP, "stream sync failed");
  if (h->timing) resolve_timers(h);
  if (e < 0) {
    const char *m = "device error";
    if (e == MGE_ERR_NEG_C2) m = "timestep: negative or NaN speed of sound squared";
    else if (e == MGE_ERR_WALL_RHO) m = "efluxb: zero density at a wall face";
    else if (e == MGE_ERR_WSOL_NAN) m = "bc_cutcell: mirror state is NaN";
    else if (e == MGE_ERR_RES_NAN) m = "solver: residual is NaN";
    int z = 0;
    h2d_sync(h->err, &z, sizeof(int), h->stream);
    return set_err(e, m);
  }
  return MGE_OK;
}

// ---- boundary conditions ------------------------------------------------

// bcond.f90:16-39: the six faces in one launch (k_bc_faces), then bc_corner
// whether bcond rewrites every ghost cell of the level from interior cells
// and nothing between a stage's update and bcond reads the ghost shell (all
// six faces written, no one-cell direction, the cut-cell clouds interior):
// then a stage need not carry the old ghost shell into its new state
bool ghosts_rewritten(const Level &L) {
  const Dims &d = L.d;
  bool all = !(d.ni < 2 || d.nj < 2 || d.nk < 2) && (L.nmir == 0 || L.cut_interior);
  for (int f = 0; f < 6; f++) {
    const int t = L.bctype[f];
    all = all && (t == -20 || t == -50 || t == -60 || t == -80);
  }
  return all;
}

// Of the kernels a steady Euler solve runs between two bcond calls on a
// level, none reads a ghost cell with two or three ghost indices (the edge
// and corner cells bc_corner writes): the stage kernels, the time step and
// the smoothing read along one axis at a time, the faces' chains start from
// interior cells when all six faces are written, and the cut-cell clouds are
// interior.  There a stage's bcond leaves bc_corner for later and marks the
// level; the corners are formed (from the faces as they stand, the values
// bcond.f90:42 would have formed after the last stage) before anything that
// does read them: the prolongation from the level, the copy ws of the state,
// the end of a cycle, and the host reading the state.
bool corners_lazy(const mge_case *h, const Level &L) {
  return h->lazy_corners && h->prm.eqntype == 0 && ghosts_rewritten(L);
}

// bc_corner.f90 on buffer buf of the level (no blocks unless run)
template <class T>
void launch_corners(mge_case *h, Level &L, int buf, bool run) {
  const Dims &d = L.d;
  const mge_params &P = h->prm;
  const int items = 4 * (d.kl - 1) + 4 * (d.jl - 1) + 4 * (d.ib + 1) + 8;
  LAUNCHT(h, "k_bc_corner", k_bc_corner, T, run ? dim3((items + 255) / 256) : dim3(0), dim3(256),
          (CornerArgs<T>{L.d, stm<T>(L, buf), (T)P.gamma, make_inf<T>(P)}));
}

template <class T>
CutArgs<T> cut_args(mge_case *h, Level &L, int buf) {
  CutArgs<T> C{};
  C.d = L.d;
  C.s = stc<T>(L, buf);
  C.inf = make_inf<T>(h->prm);
  C.recs = L.recs;
  C.wsol = L.wsol;
  C.psol = L.psol;
  C.nmir = L.nmir;
  C.nupd = L.nupd;
  C.zero_at = L.zero_at;
  C.gamma = h->prm.gamma;
  C.eqntype = h->prm.eqntype;
  C.err = h->err;
  return C;
}

// bcond; with `cut`, bc_cutcell's records go into the faces' launch; with
// `lazy`, bc_corner is left for ensure_corners
template <class T>
int do_bcond(mge_case *h, Level &L, int buf, const CutArgs<T> *cut = nullptr, bool lazy = false) {
  const mge_params &P = h->prm;
  int warn = MGE_OK;
  FacesArgs<T> F{};
  F.d = L.d;
  F.s = stm<T>(L, buf);
  int active = 0;
  for (int f = 1; f <= 6; f++) {
    int t = L.bctype[f - 1];
    if (t == 0 || (t != -20 && t != -50 && t != -60 && t != -80)) {
      if (t != 0) warn = MGE_WARN_BCTYPE;
      continue;
    }
    if (t == -60 && f <= 3 && L.bctype[f + 2] != -60)
      return set_err(MGE_ERR_PERIODIC, "Periodic boundary must appear in pairs");
    F.kind[f - 1] = t;
    F.ffdec[f - 1] = (t == -20 && L.ffdec) ? L.ffdec + L.ffoff[f - 1] : nullptr;
    active++;
  }
  F.seq = (L.d.ni < 2 || L.d.nj < 2 || L.d.nk < 2) ? 1 : 0;
  F.gamma = P.gamma; F.rho0 = P.rho0; F.p0 = P.p0; F.c0 = P.c0;
  F.s0 = std::pow(P.rho0, P.gamma) / P.p0;
  F.u0 = P.u0; F.v0 = P.v0; F.w0 = P.w0;
  F.inf = make_inf<T>(P);
  F.frozen = h->ffrozen;
  {
    F.nfb = active == 0 ? 0 : (F.seq ? 1 : (shell_cells(L.d) + 255) / 256);
    long ncb = 0;
    if (cut) {
      F.cut = *cut;
      ncb = (cut->nmir + 255) / 256;
    }
    LAUNCHT(h, "k_bc_faces", k_bc_faces, T, dim3((unsigned)(F.nfb + ncb)), dim3(256), F);
  }
  launch_corners<T>(h, L, buf, !lazy);
  L.corner_stale = lazy;
  return warn;
}

// bc_corner on the level's current state if a lazy bcond left it (a launch
// with no blocks otherwise, so that every case of a batch records the same
// launches)
template <class T>
void ensure_corners(mge_case *h, Level &L) {
  launch_corners<T>(h, L, L.cur, L.corner_stale);
  L.corner_stale = false;
}

template <class T>
int do_bc_cutcell(mge_case *h, Level &L, int buf) {
  if (L.nmir == 0) {
    LAUNCHT(h, "k_bc_cut", k_bc_cut, T, dim3(0), dim3(64), CutArgs<T>{});
    return MGE_OK;
  }
  LAUNCHT(h, "k_bc_cut", k_bc_cut, T, dim3((L.nmir + 63) / 64), dim3(64), cut_args<T>(h, L, buf));
  return L.zero_at >= 0 ? MGE_WARN_ZERO_CLOUD : MGE_OK;
}

// bc_cutcell then bcond, as each RK stage ends (euler.f90): one launch for
// both when the level's clouds lie in the interior and its faces run in
// parallel (k_bc_cut keeps its place in the sequence with no blocks, so
// every case of a batch records the same launches)
template <class T>
int do_cut_bcond(mge_case *h, Level &L, int buf, int *warn) {
  const Dims &d = L.d;
  const bool seq = d.ni < 2 || d.nj < 2 || d.nk < 2;
  const bool lazy = corners_lazy(h, L);
  if (L.nmir == 0 || !L.cut_interior || seq) {
    int rc = do_bc_cutcell<T>(h, L, buf);
    if (rc > 0) *warn = rc;
    return do_bcond<T>(h, L, buf, nullptr, lazy);
  }
  LAUNCHT(h, "k_bc_cut", k_bc_cut, T, dim3(0), dim3(64), CutArgs<T>{});
  if (L.zero_at >= 0) *warn = MGE_WARN_ZERO_CLOUD;
  const CutArgs<T> C = cut_args<T>(h, L, buf);
  return do_bcond<T>(h, L, buf, &C, lazy);
}

// ---- timestep --------------------------------------------------------------

// timestep.f90; the global minimum of the time step only when need_dt (the
// global time step, vt = 1, and the host's reads): with the local step no
// kernel reads it, and its launches then have no blocks
template <class T>
int do_timestep(mge_case *h, Level &L, bool need_dt = true) {
  const mge_params &P = h->prm;
  const DtimArgs D{L.dtim_bits, L.dtim};
  LAUNCHN(h, "k_dtim_init", k_dtim_init, need_dt ? dim3(1) : dim3(0), dim3(64), D);
  TsArgs<T> A{};
  A.d = L.d;
  A.s = stc<T>(L, L.cur);
  A.g = geo<T>(L);
  A.flags = L.flags;
  A.radsum = tp<T>(L.radsum);
  A.vdtim = tp<T>(L.vdtim);
  A.dtim_bits = L.dtim_bits;
  A.gamma = (T)P.gamma;
  A.cfl = (T)P.cfl;
  A.cellvolume = (T)L.cellvolume;
  A.vol = L.vol ? tp<T>(L.vol) : nullptr;
  A.eqntype = P.eqntype;
  A.viscl = (T)P.viscl;
  A.prlmean = (T)P.prlmean;
  A.cfac = (T)(P.distype == 0 ? 4.0 : (P.distype == 2 ? 1.0 : 2.0));
  A.inf = make_inf<T>(P);
  A.err = h->err;
  A.execmode = P.execmode;
  A.dtime = (T)(2.0 * P.dtstep / (3.0 * P.cfl));
  for (int a = 0; a < 3; a++) A.eps[a] = L.eps ? tp<T>(L.eps) + a * L.d.E : nullptr;
  A.smoo = (T)P.smoo;
  A.cflrat = (T)std::sqrt(1.0 + 4.0 * P.smoo);
  A.want_dt = need_dt ? 1 : 0;
  LAUNCHT(h, "k_timestep", k_timestep, T, cell_grid(L.d), cell_block(), A);
  LAUNCHN(h, "k_dtim_decode", k_dtim_decode, need_dt ? dim3(1) : dim3(0), dim3(64), D);
  return MGE_OK;
}

// ---- euler -------------------------------------------------------------------

// arguments of the viscous kernels on state buffer buf
template <class T>
ViscArgs<T> visc_args(mge_case *h, Level &L, int buf, T cdis, T blend, int m) {
  const mge_params &P = h->prm;
  ViscArgs<T> V{};
  V.d = L.d;
  V.s = stc<T>(L, buf);
  V.flags = L.flags;
  V.g = geo<T>(L);
  V.prim = tp<T>(L.vprim);
  V.fw = tp<T>(L.fw);
  V.inf = make_inf<T>(P);
  V.uinf[0] = (T)P.u0;
  V.uinf[1] = (T)P.v0;
  V.uinf[2] = (T)P.w0;
  V.tinf = (T)(P.p0 / (P.rho0 * P.gas_const));
  V.rgas = (T)P.gas_const;
  V.mu = (T)P.viscl;
  V.kap = (T)(P.viscl * P.gamma * P.gas_const / ((P.gamma - 1.0) * P.prlmean));
  V.cdis = cdis;
  V.blend = blend;
  V.m = m;
  return V;
}

// the stage kernel's arguments for stage m reading state buffer src and
// writing dst (w0: the stage-0 buffer)
template <class T>
StageArgs<T> stage_args(mge_case *h, Level &L, int m, int src, int dst, int w0, int mgl, int mlvl, int lres,
                        int lfir) {
  const mge_params &P = h->prm;
  const Dims &d = L.d;
  const long E = d.E;
  T *dwp = tp<T>(L.dw);
  StageArgs<T> A{};
  A.fx.d = d;
  A.fx.s = stc<T>(L, src);
  A.fx.g = geo<T>(L);
  A.fx.radsum = tp<T>(L.radsum);
  A.fx.gamma = (T)P.gamma;
  A.fx.vis2 = (T)P.vis2;
  A.fx.vis4 = (T)P.vis4;
  A.fx.cdis = (T)P.cdis[m - 1];
  A.fx.dissip = P.dissip;
  A.fx.inf = make_inf<T>(P);
  A.wdst = tp<T>(L.S[dst]);
  A.pdst = tp<T>(L.S[dst]) + 5 * E;
  A.wn = tp<T>(L.S[w0]);
  A.dw = dwp;
  A.fw = tp<T>(L.fw);
  A.qf = tp<T>(L.qf);
  A.wr = tp<T>(L.wr);
  A.marker = L.marker;
  A.flags = L.flags;
  A.vdtim = tp<T>(L.vdtim);
  A.dtim = L.dtim;
  A.cellvolume = (T)L.cellvolume;
  A.vol = L.vol ? tp<T>(L.vol) : nullptr;
  A.blend = (T)(1.0 - P.cdis[m - 1]);
  A.fn = (T)(P.cstp[m - 1] * P.cfl);
  A.m = m;
  A.has_diss = P.cdis[m - 1] > 0.0;
  A.coarse = mgl < mlvl;
  A.lfir1 = lfir && m == 1;
  A.lres = lres;
  A.smoothed = (m % 2 != 0) && P.smoo > 0.0;
  A.store_fw = (m < P.mstage) && (A.has_diss || m == 1 || A.blend != T(1));
  A.vt = P.vt;
  A.order = P.distype == 2 ? 2 : 1;
  A.eps = (T)P.epsentr;
  A.lim = L.lim;
  A.lfrozen = h->ffrozen + 1;
  A.g1 = A.fx.gamma / (A.fx.gamma - T(1));
  if (P.execmode == 1) {
    A.tsrc = h->itime > 1 ? 2 : (h->itime == 1 ? 1 : 0);
    A.ulam = A.coarse ? 1 : 2;
    A.w1 = tp<T>(L.w1);
    A.w2 = tp<T>(L.w2);
    A.tcoef = (T)(0.5 * L.cellvolume / P.dtstep);
    A.dtstep = (T)P.dtstep;
  }
  return A;
}

// tile rows of the LDS-tiled stage kernels: the level's autotuned value
// when set, else the precision's default
template <class T>
int stage_tj(const Level &L, int distype) {
  if (L.tj > 0) return L.tj;
  return distype != 0 ? RoeTile<T>::TJ : Stage2Tile<T>::TJ;
}

// k planes per workgroup of the LDS-tiled stage kernels: the level's
// autotuned value when set, else halve from nk while the launch has fewer
// than 2048 workgroups (enough to fill 256 CUs)
int stage_kc(const Level &L) {
  const Dims &d = L.d;
  if (L.kc > 0) return std::min(L.kc, d.nk);
  int kc = d.nk;
  long tiles = (long)((d.ni + TI - 1) / TI) * ((d.nj + TJ - 1) / TJ);
  while (kc > 8 && tiles * ((d.nk + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
  return kc;
}

// single-case and batch entry points of the LDS-tiled stage kernel of a
// scheme (distype 0: JST, else Roe, with the limiter-freezing paths when
// lim) with tj tile rows (4 or 8), in finish_cell's general form when gen
template <class T, int TJ_, bool GEN>
void stage_fns_t(int distype, bool lim, const void **fs, const void **fb) {
  if (distype != 0 && lim) {
    *fs = (const void *)&k_stage_roe<T, TJ_, true, GEN>;
    *fb = (const void *)&k_stage_roe_b<T, TJ_, true, GEN>;
  } else if (distype != 0) {
    *fs = (const void *)&k_stage_roe<T, TJ_, false, GEN>;
    *fb = (const void *)&k_stage_roe_b<T, TJ_, false, GEN>;
  } else {
    *fs = (const void *)&k_stage2<T, TJ_, GEN>;
    *fb = (const void *)&k_stage2_b<T, TJ_, GEN>;
  }
}
template <class T>
void stage_fns(int distype, int tj, bool lim, bool gen, const void **fs, const void **fb) {
  if (tj == 4) {
    if (gen) stage_fns_t<T, 4, true>(distype, lim, fs, fb);
    else stage_fns_t<T, 4, false>(distype, lim, fs, fb);
  } else {
    if (gen) stage_fns_t<T, 8, true>(distype, lim, fs, fb);
    else stage_fns_t<T, 8, false>(distype, lim, fs, fb);
  }
}
// whether a stage launch needs finish_cell's general form: a coarse level,
// unsteady terms, per-cell volumes or the global time step
template <class T>
bool stage_gen(const StageArgs<T> &A) {
  return A.coarse || A.tsrc || A.ulam != 0 || A.vol != nullptr || A.vt == 1;
}



// euler.f90; the residual norm only when want_res (a cycle reads it on the
// finest active level only; its launches then have no blocks elsewhere)
template <class T>
int do_euler(mge_case *h, Level &L, int ncyc, int mgl, int mlvl, int lres, int lfir, bool want_res = true) {
  const mge_params &P = h->prm;
  const Dims &d = L.d;
  struct LevelTag {  // timer statistics of this call carry the level number
    mge_case *h;
    ~LevelTag() { h->cur_level = 0; }
  } tag{h};
  h->cur_level = (int)(&L - h->lv.data());  // lv[1..nlevel]
  const long E = d.E;
  if (P.smoo > 0.0 && (d.ni < 2 || d.nj < 2 || d.nk < 2))
    return set_err(MGE_ERR_SMOO_DIM, "residual smoothing needs at least 2 cells per direction");
  if (lfir) {
    ensure_corners<T>(h, L);
    LAUNCHT(h, "k_copy", k_copy, T, flat_grid(5 * E), dim3(256),
            (CopyArgs<T>{(const T *)tp<T>(L.S[L.cur]), tp<T>(L.ws), 5 * E}));
  }
  T *dwp = tp<T>(L.dw);
  const int w0 = L.cur;
  // the last stage always lands in `target` (0 <-> 1 from a start in 0 or
  // 1), whatever the stage count, so the state buffer toggles with period
  // two per call; earlier stages alternate between target and the third
  // buffer, never writing the buffer they read
  const int target = w0 == 0 ? 1 : 0;
  const int other = 3 - w0 - target;
  int src = w0;
  int warn = MGE_OK;
  for (int m = 1; m <= P.mstage; m++) {
    const int dst = ((P.mstage - m) % 2 == 0) ? target : other;
    StageArgs<T> A = stage_args<T>(h, L, m, src, dst, w0, mgl, mlvl, lres, lfir);
    // laminar viscous fluxes (euler.f90:137-181): fw = blend fw + cdis V
    // before the stage kernel adds its dissipation
    A.fw_in = P.eqntype == 1 && A.has_diss;
    if (A.fw_in) {
      ViscArgs<T> V = visc_args<T>(h, L, src, A.fx.cdis, A.blend, m);
      LAUNCHT(h, "k_visc_prim", k_visc_prim, T, flat_grid(E), dim3(256), V);
      LAUNCHT(h, "k_visc", k_visc, T, cell_grid(d), cell_block(), V);
      LAUNCHT(h, "k_visc_wall", k_visc_wall, T, dim3((L.nvcell + 63) / 64), dim3(64),
              (ViscWallArgs<T>{V, L.recs, L.wsol, L.psol, L.nmir, L.vcell, L.vbeg, L.vrec, L.nvcell}));
    }
    {
      // algorithmic bytes per cell: read W^(m-1); write W^(m) and read W^(0)
      // when this kernel updates, else write dw; forcing on coarse levels;
      // one flag byte
      double s8 = (double)sizeof(T), per = 5 * s8 + 1;
      bool upd = !lres && !A.smoothed;
      per += upd ? 10 * s8 : 5 * s8;
      if (A.coarse) per += A.lfir1 ? 10 * s8 : 5 * s8;
      h->next_bytes = per * (double)d.ni * d.nj * d.nk;
    }
    {
      const int kc = stage_kc(L);
      const int tj = stage_tj<T>(L, P.distype);  // tile rows of the LDS-tiled kernel
      dim3 g((d.ni + TI - 1) / TI, (d.nj + tj - 1) / tj, (d.nk + kc - 1) / kc);
      A.kc = kc;
      if (P.distype != 0 || P.dissip == 0) {
        const void *fs, *fb;
        stage_fns<T>(P.distype, tj, A.lim != nullptr, stage_gen(A), &fs, &fb);
        launch(h, "k_stage", fs, fb, g, dim3(TI, tj), A);
      } else {
        LAUNCHT(h, "k_stage", k_stage, T, cell_grid(d), cell_block(), A);  // matrix dissipation
      }
    }
    {
      RecArgs<T> R;
      R.st = A;
      R.st.store_fw = 1;
      R.recs = L.recs;
      R.wsol = L.wsol;
      R.psol = L.psol;
      R.nmir = L.nmir;
      R.tcell = L.tcell;
      R.beg = L.beg;
      R.ops = L.ops;
      R.chunks = L.chunks;
      R.ntouch = L.ntouch;
      R.nops = L.nops;
      R.err = h->err;
      LAUNCHT(h, "k_rec", k_rec, T, dim3((unsigned)L.nchunk), dim3(kRecOps), R);
    }
    if (lres) return MGE_OK;
    if (A.smoothed && P.smoother == 1) {
      // psmoo with progon3's sweep: x, y, then z lines (psmoo.f90:40-87)
      PsmooArgs<T> S{d, dwp, tp<T>(L.eps), tp<T>(L.eps) + 3 * E};
      auto lines = [](long nl) { return dim3((unsigned)((nl + 63) / 64)); };
      LAUNCHT(h, "k_psmoo_x", k_psmoo_x, T, lines((long)d.nj * d.nk), dim3(64), S);
      S.eps = tp<T>(L.eps) + E;
      LAUNCHT(h, "k_psmoo_y", k_psmoo_y, T, lines((long)d.ni * d.nk), dim3(64), S);
      S.eps = tp<T>(L.eps) + 2 * E;
      LAUNCHT(h, "k_psmoo_z", k_psmoo_z, T, lines((long)d.ni * d.nj), dim3(64), S);
      UpdArgs<T> U{d, (const T *)tp<T>(L.S[w0]), (const T *)dwp, stc<T>(L, src), stm<T>(L, dst),
                   (T)P.gamma, make_inf<T>(P)};
      U.ulam = A.ulam == 2;
      U.vdtim = A.vdtim;
      U.dtim = A.dtim;
      U.fn = A.fn;
      U.dtstep = A.dtstep;
      U.vt = A.vt;
      U.carry = ghosts_rewritten(L) ? 0 : 1;
      LAUNCHT(h, "k_update_ghost", k_update_ghost, T, flat_grid(E), dim3(256), U);
    } else if (A.smoothed) {
      SmoothArgs<T> S;
      S.d = d;
      S.dw = dwp;
      S.flags = L.flags;
      S.bb = (T)h->smoo_bb;
      T *piv = tp<T>(L.piv);
      S.e1 = piv;
      S.e2 = piv + L.npiv;
      S.fa = tp<T>(L.fax);
      S.r1 = piv + 6 * L.npiv;
      long nwave = (long)d.nj * d.nk * 5;
      auto gxk = [&](int K) { return dim3((unsigned)(((nwave + K - 1) / K * 64 + 255) / 256)); };
#define XSCAN(C)                                                                                            \
  launch(h, "k_smooth_x", (const void *)&k_smooth_x_scan<C, T>, (const void *)&k_smooth_x_scan_b<C, T>,     \
         gxk(xscan_lines<C>()), dim3(256), S)
      if (d.ni <= 64) XSCAN(1);
      else if (d.ni <= 128) XSCAN(2);
      else if (d.ni <= 256) XSCAN(4);
      else if (d.ni <= 512) XSCAN(8);
      else LAUNCHT(h, "k_smooth_x", k_smooth_x_seq, T, flat_grid(nwave), dim3(256), S);
#undef XSCAN
      S.e1 = piv + 2 * L.npiv;
      S.e2 = piv + 3 * L.npiv;
      S.r1 = piv + 7 * L.npiv;
      LAUNCHT(h, "k_smooth_y", k_smooth_y, T, dim3((d.ni + 63) / 64, d.nk, 5), dim3(64), S);
      S.e1 = piv + 4 * L.npiv;
      S.e2 = piv + 5 * L.npiv;
      S.r1 = piv + 8 * L.npiv;
      LAUNCHT(h, "k_smooth_z", k_smooth_z, T, dim3((d.ni + 63) / 64, d.nj, 5), dim3(64), S);
      UpdArgs<T> U{d, (const T *)tp<T>(L.S[w0]), (const T *)dwp, stc<T>(L, src), stm<T>(L, dst),
                   (T)P.gamma, make_inf<T>(P)};
      U.ulam = A.ulam == 2;
      U.vdtim = A.vdtim;
      U.dtim = A.dtim;
      U.fn = A.fn;
      U.dtstep = A.dtstep;
      U.vt = A.vt;
      U.carry = ghosts_rewritten(L) ? 0 : 1;
      LAUNCHT(h, "k_update_ghost", k_update_ghost, T, flat_grid(E), dim3(256), U);
    } else {
      // the ghost shell of the new state is the old one until bcond; when
      // bcond rewrites every ghost cell nothing reads the copy and it is
      // skipped
      LAUNCHT(h, "k_ghost_copy", k_ghost_copy, T, ghosts_rewritten(L) ? dim3(0) : flat_grid(shell_cells(d)), dim3(256),
              (GhostArgs<T>{d, stc<T>(L, src), stm<T>(L, dst)}));
    }
    int rc = do_cut_bcond<T>(h, L, dst, &warn);
    if (rc < 0) return rc;
    if (rc > 0) warn = rc;
    src = dst;
  }
  L.cur = src;
  RtArgs<T> R;
  R.d = d;
  R.dw = dwp;
  R.vdtim = tp<T>(L.vdtim);
  R.dtim = L.dtim;
  R.flags = L.flags;
  R.cfl = P.cfl;
  R.vt = P.vt;
  R.partial = L.partial;
  dim3 g = cell_grid(d);
  g.z = (g.z + kRtrmsPlanes - 1) / kRtrmsPlanes;
  R.gx = (int)g.x;
  R.gy = (int)g.y;
  LAUNCHT(h, "k_rtrms", k_rtrms, T, want_res ? g : dim3(0), cell_block(), R);
  long nb = (long)g.x * g.y * g.z;
  int c = ncyc >= 1 && ncyc <= L.mxcyc ? ncyc : 0;
  LAUNCHN(h, "k_rtrms_final", k_rtrms_final, want_res ? dim3(1) : dim3(0), dim3(256),
          (RfArgs{(const double *)L.partial, nb, (double)((long)d.ni * d.nj * d.nk), L.rtrms + c, h->err}));
  return warn;
}

// ---- transfers ------------------------------------------------------------------

// fine level = ratio x coarse level in every direction (aligned levels)
bool levels_nest(const Level &c, const Level &f) {
  int nc[3] = {c.d.ni, c.d.nj, c.d.nk}, nf[3] = {f.d.ni, f.d.nj, f.d.nk};
  for (int a = 0; a < 3; a++)
    if (nf[a] != f.ratio[a] * nc[a]) return false;
  return true;
}

// the grid of a transfer onto level L that also forms p on every cell of L
// (XferArgs::eos): the cells' grid and then enough z-slices of blocks for
// the ghost shell
template <class T>
dim3 xfer_eos_grid(XferArgs<T> &X, Level &L, const mge_params &P, dim3 g) {
  const long per = (long)g.x * g.y * 256;
  X.eos = 1;
  X.peos = tp<T>(L.S[L.cur]) + 5 * L.d.E;
  X.gamma = (T)P.gamma;
  X.inf = make_inf<T>(P);
  X.gx = (int)g.x;
  X.gy = (int)g.y;
  X.gz0 = (int)g.z;
  return dim3(g.x, g.y, g.z + (unsigned)((shell_cells(L.d) + per - 1) / per));
}

template <class T>
int do_movco(mge_case *h, int mgl) {
  Level &c = h->lv[mgl], &f = h->lv[mgl + 1];
  if (!levels_nest(c, f)) return set_err(MGE_ERR_ARG, "levels do not nest with the given ratio");
  XferArgs<T> X{};
  X.dc = c.d;
  X.df = f.d;
  X.rx = f.ratio[0]; X.ry = f.ratio[1]; X.rz = f.ratio[2];
  X.wf = tp<T>(f.S[f.cur]);
  X.dwf = tp<T>(f.dw);
  X.wc = tp<T>(c.S[c.cur]);
  X.wsc = tp<T>(c.ws);
  X.wr = tp<T>(c.wr);
  X.marker = c.marker;
  X.volc = (T)c.cellvolume;
  X.volf = (T)f.cellvolume;
  X.vc = c.vol ? tp<T>(c.vol) : nullptr;
  X.vf = f.vol ? tp<T>(f.vol) : nullptr;
  for (int a = 0; a < 3; a++) X.rw[a] = f.rw[a] ? tp<T>(f.rw[a]) : nullptr;
  if ((X.vc != nullptr) != (X.vf != nullptr) || (X.vf && !X.rw[0]))
    return set_err(MGE_ERR_ARG, "a stretched level and a uniform level do not nest");
  X.fmg = 0;
  X.flagsf = f.flags;
  X.wfo = nullptr;
  X.nrc[0] = c.d.ni; X.nrc[1] = c.d.nj; X.nrc[2] = c.d.nk;
  // the restriction covers every coarse cell: it sets their p, and blocks
  // past its grid the ghost shell's (the EOS pass over the level)
  const dim3 gr = xfer_eos_grid(X, c, h->prm, cell_grid(c.d));
  LAUNCHT(h, "k_restrict", k_restrict, T, gr, cell_block(), X);
  return do_bcond<T>(h, c, c.cur);
}

template <class T>
int do_movfin(mge_case *h, int mgl, int mlvl) {
  Level &f = h->lv[mgl], &c = h->lv[mgl - 1];
  if (!levels_nest(c, f)) return set_err(MGE_ERR_ARG, "levels do not nest with the given ratio");
  XferArgs<T> X{};
  X.dc = c.d;
  X.df = f.d;
  X.rx = f.ratio[0]; X.ry = f.ratio[1]; X.rz = f.ratio[2];
  X.wf = nullptr;
  X.dwf = nullptr;
  X.wc = tp<T>(c.S[c.cur]);
  X.wsc = tp<T>(c.ws);
  X.wr = nullptr;
  X.marker = nullptr;
  X.volc = (T)c.cellvolume;
  X.volf = (T)f.cellvolume;
  for (int a = 0; a < 3; a++) X.pw[a] = f.pw[a] ? tp<T>(f.pw[a]) : nullptr;
  if ((c.vol != nullptr) != (f.vol != nullptr) || (f.vol && !X.pw[0]))
    return set_err(MGE_ERR_ARG, "a stretched level and a uniform level do not nest");
  X.fmg = (mgl - 1 == mlvl);
  X.flagsf = f.flags;
  X.wfo = tp<T>(f.S[f.cur]);
  X.corr = tp<T>(c.corr);
  ensure_corners<T>(h, c);  // the prolongation stencils reach the coarse level's edge cells
  auto pair_aligned = [&](const void *p) { return (reinterpret_cast<uintptr_t>(p) & (2 * sizeof(T) - 1)) == 0; };
  const bool two = f.ratio[0] == 2 && f.ratio[1] == 2 && f.ratio[2] == 2 && !X.pw[0] && f.d.sy % 2 == 0 &&
                   f.d.sz % 2 == 0 && f.d.E % 2 == 0 && pair_aligned(X.wfo) &&
                   (reinterpret_cast<uintptr_t>(f.flags) & 1) == 0;
  // k_prolong2 forms the correction itself: its k_corr launch has no blocks
  // (kept, so that the cases of a batch record the same launches)
  if (!X.fmg) LAUNCHT(h, "k_corr", k_corr, T, two ? dim3(0) : flat_grid(5 * c.d.E), dim3(256), X);
  if (two) {
    // a thread per coarse cell and its eight fine cells (pairs of fine cells
    // in a row move together: even strides and aligned arrays)
    const dim3 gp = xfer_eos_grid(X, f, h->prm, cell_grid(c.d));
    LAUNCHT(h, "k_prolong", k_prolong2, T, gp, cell_block(), X);
  } else {
    const dim3 gp = xfer_eos_grid(X, f, h->prm, cell_grid(f.d));
    LAUNCHT(h, "k_prolong", k_prolong, T, gp, cell_block(), X);
  }
  return do_bcond<T>(h, f, f.cur);
}

// unsteady: both earlier physical time levels start from the current state
// (init_flow.f90:79-86)
int init_time_levels(mge_case *h, Level &L) {
  const size_t n = h->es * 5 * (size_t)L.d.E;
  HIPCHK(hipMemcpyAsync(L.w1, L.S[L.cur], n, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(L.w2, L.S[L.cur], n, hipMemcpyDeviceToDevice, h->stream));
  return MGE_OK;
}

// solver.f90:263-276: w2 = w1, w1 = w on every level (whole arrays)
int shift_time_levels(mge_case *h) {
  for (int l = 1; l <= h->nlevel; l++) {
    Level &L = h->lv[l];
    const size_t n = h->es * 5 * (size_t)L.d.E;
    HIPCHK(hipMemcpyAsync(L.w2, L.w1, n, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(L.w1, L.S[L.cur], n, hipMemcpyDeviceToDevice, h->stream));
  }
  return MGE_OK;
}

template <class T>
int do_init_flow(mge_case *h) {
  const mge_params &P = h->prm;
  int warn = MGE_OK;
  for (int l = 1; l <= h->nlevel; l++) {
    Level &L = h->lv[l];
    if (!L.defined) return set_err(MGE_ERR_STATE, "level not defined");
    // the freestream state itself (FP64) or a zero deviation from it (FP32)
    const Inf<T> fi = make_inf<T>(P);
    FillArgs<T> F{L.d, stm<T>(L, L.cur), {fi.w[0], fi.w[1], fi.w[2], fi.w[3], fi.w[4]}};
    if constexpr (kDelta<T>)
      for (int n = 0; n < 5; n++) F.v[n] = T(0);
    LAUNCHT(h, "k_fill", k_fill, T, flat_grid(L.d.E), dim3(256), F);
  }
  for (int l = 1; l <= h->nlevel; l++) {
    Level &L = h->lv[l];
    LAUNCHT(h, "k_eos_all", k_eos_all, T, flat_grid(L.d.E), dim3(256), (EosArgs<T>{L.d, stm<T>(L, L.cur), (T)P.gamma, make_inf<T>(P)}));
    if (P.execmode == 1) CHKRC(init_time_levels(h, L));
    int rc = do_bc_cutcell<T>(h, L, L.cur);
    if (rc > 0) warn = rc;
    rc = do_bcond<T>(h, L, L.cur);
    if (rc < 0) return rc;
    if (rc > 0) warn = rc;
  }
  return warn;
}

// one multigrid cycle with finest active level mlvl: the body of lopmg in
// solver.f90:127-234.  The residual of the mlvl relax is left in
// L.rtrms[ncyc] (read by the caller after the cycle: the finest level is
// relaxed once per cycle, so the value is final when the cycle ends).
template <class T>
int do_cycle_levels(mge_case *h, int mlvl, int ncyc, long long *cs);

// one W/V-cycle, the levels' corner cells formed at its end
template <class T>
int do_cycle(mge_case *h, int mlvl, int ncyc, long long *cs) {
  const int rc = do_cycle_levels<T>(h, mlvl, ncyc, cs);
  if (rc < 0) return rc;
  for (int l = 1; l <= h->nlevel; l++)
    if (h->lv[l].defined) ensure_corners<T>(h, h->lv[l]);
  return rc;
}

template <class T>
int do_cycle_levels(mge_case *h, int mlvl, int ncyc, long long *cs) {
  const int nlevel = h->nlevel;
  const mge_params &P = h->prm;
  std::vector<int> icyc(nlevel + 1, 0);
  int mgl = mlvl;
  for (;;) {
    for (;;) {
      icyc[mgl] += 1;
      int lfir = (mgl == mlvl) ? 0 : 1;
      Level &B = h->lv[mgl];
      CHKRC(do_timestep<T>(h, B, P.vt == 1));
      CHKRC(do_euler<T>(h, B, ncyc, mgl, mlvl, 0, lfir, mgl == mlvl));
      *cs += (long long)P.mstage * B.d.ni * B.d.nj * B.d.nk;
      if (mlvl == 1) return MGE_OK;
      CHKRC(do_euler<T>(h, B, ncyc, mgl, mlvl, 1, 0));
      *cs += (long long)B.d.ni * B.d.nj * B.d.nk;
      mgl -= 1;
      CHKRC(do_movco<T>(h, mgl));
      if (mgl == 1) break;
    }
    int lfir = 1;
    for (;;) {
      Level &B = h->lv[mgl];
      CHKRC(do_timestep<T>(h, B, P.vt == 1));
      CHKRC(do_euler<T>(h, B, ncyc, mgl, mlvl, 0, lfir, mgl == mlvl));
      *cs += (long long)P.mstage * B.d.ni * B.d.nj * B.d.nk;
      lfir = 0;
      mgl += 1;
      CHKRC(do_movfin<T>(h, mgl, mlvl));
      if (mgl == mlvl) return MGE_OK;
      if (icyc[mgl] != P.mgtype) break;
      icyc[mgl] = 0;
    }
  }
}

// the far-field freeze test of one case after a finest-level cycle, on the
// device so that it needs no host round trip (k_converge makes the same
// test for batched cases)
__global__ void k_ffreeze(const double *rtrms_slot, double *ini, int ncyc, double ffreeze, double lfreeze,
                          int *frozen) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r = *rtrms_slot;
  if (ncyc == 1) *ini = r;
  if (ffreeze > 0.0 && *ini / r > ffreeze) frozen[0] = 1;
  if (lfreeze > 0.0 && *ini / r > lfreeze) frozen[1] = 1;
}

// after finest-level cycle ncyc of a single case: the freeze tests (far
// field and limiter)
int ffreeze_test(mge_case *h, int mlvl, int ncyc) {
  if (!(h->prm.ffreeze > 0.0 || h->prm.lfreeze > 0.0) || mlvl != h->nlevel || h->rec) return MGE_OK;
  Level &B = h->lv[mlvl];
  const int c = ncyc >= 1 && ncyc <= B.mxcyc ? ncyc : 0;
  hipLaunchKernelGGL(k_ffreeze, dim3(1), dim3(64), 0, h->stream, (const double *)(B.rtrms + c), h->ffini, ncyc,
                     h->prm.ffreeze, h->prm.lfreeze, h->ffrozen);
  HIPCHK(hipGetLastError());
  return MGE_OK;
}

int read_rtrms(mge_case *h, int mlvl, int ncyc, double *r) {
  Level &B = h->lv[mlvl];
  int c = ncyc >= 1 && ncyc <= B.mxcyc ? ncyc : 0;
  HIPCHK(hipMemcpyAsync(r, B.rtrms + c, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return check_err(h);
}

// solver.f90:119-261, one physical time step at level itime (steady flow:
// nperiod = nstep = 1, itime = 0): FMG ramp over the levels, cycles with the
// residual log and the drop test
template <class T>
int do_solve(mge_case *h, const int *mcyc, int itime, double *reslog, int cap, int *nlog, long long *cs_out) {
  const int nlevel = h->nlevel;
  struct TimeTag {  // euler calls of this solve run at time level itime
    mge_case *h;
    ~TimeTag() { h->itime = 0; }
  } ttag{h};
  h->itime = itime;
  const mge_params &P = h->prm;
  int nl = 0;
  double ini_res = 0.0, max_res = 0.0;
  long long cs = 0;
  for (int mlvl = 1; mlvl <= nlevel; mlvl++) {
    for (int ncyc = 1; ncyc <= mcyc[mlvl - 1]; ncyc++) {
      CHKRC(do_cycle<T>(h, mlvl, ncyc, &cs));
      CHKRC(ffreeze_test(h, mlvl, ncyc));
      double r = 0.0;
      CHKRC(read_rtrms(h, mlvl, ncyc, &r));
      max_res = 1000.0;
      if (ncyc == 1) ini_res = r;
      if (r < max_res) max_res = r;
      if (r != r) return set_err(MGE_ERR_RES_NAN, "solver: residual is NaN");
      if (reslog && nl < cap) {
        reslog[3 * nl] = mlvl;
        reslog[3 * nl + 1] = ncyc;
        reslog[3 * nl + 2] = max_res;
      }
      nl++;
      if (mlvl > 1 && ini_res / max_res > P.drop) break;
    }
    if (mlvl < nlevel) CHKRC(do_movfin<T>(h, mlvl + 1, mlvl));
  }
  CHKRC(check_err(h));
  if (nlog) *nlog = nl;
  if (cs_out) *cs_out = cs;
  return MGE_OK;
}

// ---- autotuner for the LDS-tiled stage kernels ---------------------------
//
// The launch shape of k_stage2/k_stage_roe changes neither the arithmetic
// nor the results: every face value has the same expression whatever tile
// evaluates it.  Two parameters are tuned together: the tile rows tj (4 or
// 8: 256 or 512 threads, the LDS tile and its halo share) and kc, the number
// of k planes a workgroup marches through (the column carry restarts at
// every chunk): few planes give many workgroups, many planes amortise the
// chunk start.  The tuner times stage 1 of the level for tj = 4, 8 and
// kc = 1, 2, 4, ... nk (one warm-up launch, then the best of three) on the
// level's own arrays, writing only a state buffer that is not current, dw,
// fw and the time-step arrays (all rewritten by the next timestep/euler
// call).  Results are cached per (shape, precision, scheme) for the process.
struct TuneKey {
  int ni, nj, nk, fp32, distype, smoothed;
  bool operator<(const TuneKey &o) const {
    return std::tie(ni, nj, nk, fp32, distype, smoothed) < std::tie(o.ni, o.nj, o.nk, o.fp32, o.distype, o.smoothed);
  }
};
struct TuneVal {
  int tj, kc;
  double us;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

template <class T>
int do_autotune(mge_case *h, Level &L, int *kc_out, double *us_out) {
  const mge_params &P = h->prm;
  const Dims &d = L.d;
  if (h->rec) return set_err(MGE_ERR_STATE, "autotune while a batch is being recorded");
  if (P.distype == 0 && P.dissip != 0) {  // matrix dissipation: one thread per cell, nothing to tune
    L.kc = 0;
    L.tj = 0;
    if (kc_out) *kc_out = 0;
    if (us_out) *us_out = 0.0;
    return MGE_OK;
  }
  TuneKey key{d.ni, d.nj, d.nk, h->fp32 ? 1 : 0, P.distype, P.smoo > 0.0 ? 1 : 0};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      L.kc = it->second.kc;
      L.tj = it->second.tj;
      if (kc_out) *kc_out = L.kc;
      if (us_out) *us_out = it->second.us;
      return MGE_OK;
    }
  }
  CHKRC(do_timestep<T>(h, L));
  const int w0 = L.cur;
  int dst = 0;
  while (dst == w0) dst++;
  StageArgs<T> A = stage_args<T>(h, L, 1, w0, dst, w0, 1, 1, 0, 0);
  A.fw_in = 0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // the best k-chunk per tile row count; the row count other than the
  // precision's default is taken only when it is faster by more than the
  // timing's spread between runs (isolated repeats of a launch rank the two
  // within a few percent either way)
  constexpr double kOtherMargin = 0.98;
  L.tj = 0;
  const int tj_def = stage_tj<T>(L, P.distype);
  int best = 0, best_tj = 0, best_o = 0;
  double best_us = 1e300, best_o_us = 1e300;
  hipError_t err = hipSuccess;
  for (int tj : {4, 8}) {
    const void *fn, *fb;
    stage_fns<T>(P.distype, tj, A.lim != nullptr, stage_gen(A), &fn, &fb);
    for (int kc = 1; err == hipSuccess; kc = std::min(2 * kc, d.nk)) {
      A.kc = kc;
      dim3 g((d.ni + TI - 1) / TI, (d.nj + tj - 1) / tj, (d.nk + kc - 1) / kc);
      void *args[1] = {(void *)&A};
      double t = 1e300;
      for (int rep = 0; rep < 4 && err == hipSuccess; rep++) {
        err = hipEventRecord(e0, h->stream);
        if (err == hipSuccess) err = hipLaunchKernel(fn, g, dim3(TI, tj), args, 0, h->stream);
        if (err == hipSuccess) err = hipEventRecord(e1, h->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && err == hipSuccess) t = std::min(t, 1e3 * (double)ms);
      }
      if (err != hipSuccess) break;
      if (tj == tj_def && t < best_us) {
        best_us = t;
        best = kc;
        best_tj = tj;
      } else if (tj != tj_def && t < best_o_us) {
        best_o_us = t;
        best_o = kc;
      }
      if (kc == d.nk) break;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (err != hipSuccess) return set_err(MGE_ERR_HIP, hipGetErrorString(err));
  if (best_o_us < kOtherMargin * best_us) {
    best_us = best_o_us;
    best = best_o;
    best_tj = tj_def == 4 ? 8 : 4;
  }
  CHKRC(check_err(h));
  L.kc = best;
  L.tj = best_tj;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {best_tj, best, best_us};
  }
  if (kc_out) *kc_out = best;
  if (us_out) *us_out = best_us;
  return MGE_OK;
}

template <class T>
int do_wall_force(mge_case *h, Level &L) {
  FluxCtx<T> fc{};
  fc.d = L.d;
  fc.s = stc<T>(L, L.cur);
  fc.g = geo<T>(L);
  fc.inf = make_inf<T>(h->prm);
  hipLaunchKernelGGL(k_wall_terms<T>, dim3((L.nmir + 63) / 64), dim3(64), 0, h->stream, fc, (const DevRec *)L.recs,
                     (const double *)L.wsol, (const double *)L.psol, L.nmir, h->prm.distype != 0 ? 1 : 0,
                     h->prm.gamma, h->prm.epsentr, L.opval, h->err);
  if (h->prm.eqntype == 1 && L.nmir > 0) {
    ViscArgs<T> V = visc_args<T>(h, L, L.cur, T(1), T(0), 1);
    hipLaunchKernelGGL(k_wall_visc_terms<T>, dim3((L.nmir + 63) / 64), dim3(64), 0, h->stream, V,
                       (const DevRec *)L.recs, (const double *)L.wsol, (const double *)L.psol, L.nmir, L.opval);
  }
  hipLaunchKernelGGL(k_wall_sum, dim3(1), dim3(64), 0, h->stream, (const double *)L.opval, L.nmir, L.force);
  return MGE_OK;
}

// host-side record bookkeeping: touched cells and their ordered op lists.
// JST: the dfluxb/efluxb undo/add terms on the cut cell and its inner
// neighbour; Roe: one wall-face term on the cut cell.
int build_records(mge_case *h, Level &L, int nmir, const mge_cutrec *recs) {
  const bool roe = h->prm.distype != 0;
  const Dims &d = L.d;
  L.nmir = nmir;
  L.hrecs.assign(recs, recs + nmir);
  L.zero_at = -1;
  L.nupd = nmir;
  for (int m = 0; m < nmir; m++)
    if (recs[m].npoint == 0) {
      L.zero_at = m;
      L.nupd = m;
      break;
    }
  if (nmir == 0) return MGE_OK;
  bool interior = true;
  std::vector<DevRec> dr(nmir);
  for (int m = 0; m < nmir; m++) {
    const mge_cutrec &r = recs[m];
    DevRec &o = dr[m];
    std::memset(&o, 0, sizeof(o));
    if (r.faceid < 1 || r.faceid > 6 || r.zz != (r.faceid - 1) % 3 + 1 || r.npoint < 0 || r.npoint > 8)
      return set_err(MGE_ERR_ARG, "bad wall record");
    o.faceid = r.faceid;
    o.zz = r.zz;
    for (int a = 0; a < 3; a++) {
      o.c[a] = r.ccid[a];
      o.s[a] = r.scid[a];
      o.nvec[a] = r.nvec[a];
    }
    int cmax[3] = {d.il, d.jl, d.kl};
    for (int a = 0; a < 3; a++) {
      // the cut cell and its stencil along zz must stay inside the arrays
      if (r.ccid[a] < 2 || r.ccid[a] > cmax[a]) return set_err(MGE_ERR_ARG, "record cell outside the block");
      if (r.scid[a] < 0 || r.scid[a] > cmax[a] + 2) return set_err(MGE_ERR_ARG, "record solid cell outside");
    }
    int s0 = r.ccid[r.zz - 1];
    int nl = cmax[r.zz - 1];
    if (s0 < 3 || s0 > nl - 1) return set_err(MGE_ERR_ARG, "wall record too close to the block edge");
    o.npoint = r.npoint;
    for (int c = 0; c < r.npoint; c++) {
      for (int a = 0; a < 3; a++) {
        if (r.ijk[c][a] < 0 || r.ijk[c][a] > cmax[a] + 2) return set_err(MGE_ERR_ARG, "cloud point outside");
        if (r.ijk[c][a] < 2 || r.ijk[c][a] > cmax[a]) interior = false;
      }
      o.cloud[c] = ix(d, r.ijk[c][0], r.ijk[c][1], r.ijk[c][2]);
      o.coef[c] = r.coef[c];
    }
  }
  L.cut_interior = interior;
  // ordered op streams per touched cell
  std::map<long, std::pair<std::vector<int2>, std::vector<int2>>> per;
  for (int m = 0; m < nmir; m++) {
    const mge_cutrec &r = recs[m];
    long st = r.zz == 1 ? 1L : (r.zz == 2 ? d.sy : d.sz);
    long qc = ix(d, r.ccid[0], r.ccid[1], r.ccid[2]);
    bool hi = r.faceid >= 4;
    long qn = hi ? qc - st : qc + st;
    auto &pc = per[qc];
    if (roe) {
      pc.first.push_back(make_int2(m, OP_ROE_WALL));
      continue;
    }
    auto &pn = per[qn];
    if (hi) {
      pc.first.push_back(make_int2(m, OP_HI_C_UNDO));
      pn.first.push_back(make_int2(m, OP_HI_N_UNDO));
      pc.first.push_back(make_int2(m, OP_HI_C_ADD));
      pn.first.push_back(make_int2(m, OP_HI_N_ADD));
    } else {
      pc.first.push_back(make_int2(m, OP_LO_C_UNDO));
      pn.first.push_back(make_int2(m, OP_LO_N_UNDO));
      pc.first.push_back(make_int2(m, OP_LO_C_ADD));
      pn.first.push_back(make_int2(m, OP_LO_N_ADD));
    }
    pc.second.push_back(make_int2(m, OP_E_UNDO));
    pc.second.push_back(make_int2(m, OP_E_ADD));
  }
  std::vector<long> tcell;
  std::vector<int> begv(1, 0);
  std::vector<int2> opv;
  for (auto &kv : per) {
    tcell.push_back(kv.first);
    for (auto &o : kv.second.first) opv.push_back(o);
    for (auto &o : kv.second.second) opv.push_back(o);
    begv.push_back((int)opv.size());
  }
  L.ntouch = (int)tcell.size();
  L.nops = (int)opv.size();
  // workgroups of k_rec: consecutive touched cells with at most kRecOps ops
  std::vector<int2> chunks;
  for (int t = 0, t0 = 0; t < L.ntouch; t++) {
    if (begv[t + 1] - begv[t] > kRecOps) return set_err(MGE_ERR_ARG, "too many wall ops on one cell");
    if (begv[t + 1] - begv[t0] > kRecOps) {
      chunks.push_back(make_int2(t0, t));
      t0 = t;
    }
    if (t == L.ntouch - 1) chunks.push_back(make_int2(t0, L.ntouch));
  }
  L.nchunk = (int)chunks.size();
  CHKRC(dalloc(L, &L.recs, nmir));
  CHKRC(dalloc(L, &L.wsol, (size_t)5 * nmir));
  CHKRC(dalloc(L, &L.psol, nmir));
  CHKRC(dalloc(L, &L.tcell, tcell.size()));
  CHKRC(dalloc(L, &L.beg, begv.size()));
  CHKRC(dalloc(L, &L.ops, opv.size() + 1));
  CHKRC(dalloc(L, &L.chunks, chunks.size() + 1));
  CHKRC(dalloc(L, &L.opval, (opv.size() + 1) * 10));  // the wall force's per-record terms
  HIPCHK(h2d_sync(L.recs, dr.data(), sizeof(DevRec) * nmir, h->stream));
  {
    // cut cells with wall records, in order of their first record, and
    // their records in record order (viscous wall terms)
    std::map<long, int> slot;
    std::vector<long> vc;
    std::vector<std::vector<int>> lists;
    for (int m = 0; m < nmir; m++) {
      long qc = ix(d, recs[m].ccid[0], recs[m].ccid[1], recs[m].ccid[2]);
      auto it = slot.find(qc);
      int sl;
      if (it == slot.end()) {
        sl = (int)vc.size();
        slot[qc] = sl;
        vc.push_back(qc);
        lists.emplace_back();
      } else {
        sl = it->second;
      }
      lists[sl].push_back(m);
    }
    std::vector<int> vb(1, 0), vr;
    for (auto &l : lists) {
      vr.insert(vr.end(), l.begin(), l.end());
      vb.push_back((int)vr.size());
    }
    L.nvcell = (int)vc.size();
    CHKRC(dalloc(L, &L.vcell, vc.size()));
    CHKRC(dalloc(L, &L.vbeg, vb.size()));
    CHKRC(dalloc(L, &L.vrec, vr.size()));
    HIPCHK(h2d_sync(L.vcell, vc.data(), sizeof(long) * vc.size(), h->stream));
    HIPCHK(h2d_sync(L.vbeg, vb.data(), sizeof(int) * vb.size()));
    HIPCHK(h2d_sync(L.vrec, vr.data(), sizeof(int) * vr.size()));
  }
  HIPCHK(h2d_sync(L.tcell, tcell.data(), sizeof(long) * tcell.size(), h->stream));
  HIPCHK(h2d_sync(L.beg, begv.data(), sizeof(int) * begv.size()));
  if (!opv.empty()) HIPCHK(h2d_sync(L.ops, opv.data(), sizeof(int2) * opv.size()));
  if (!chunks.empty()) HIPCHK(h2d_sync(L.chunks, chunks.data(), sizeof(int2) * chunks.size()));
  std::vector<uint8_t> fl(d.E);
  HIPCHK(d2h_sync(fl.data(), L.flags, d.E, h->stream));
  for (long q : tcell) fl[q] |= F_TOUCHED;
  HIPCHK(h2d_sync(L.flags, fl.data(), d.E));
  std::vector<double> w5(5 * (size_t)nmir), p1(nmir);
  for (int m = 0; m < nmir; m++) {
    for (int n = 0; n < 5; n++) w5[(size_t)n * nmir + m] = recs[m].wsol[n];
    p1[m] = recs[m].psol;
  }
  HIPCHK(h2d_sync(L.wsol, w5.data(), sizeof(double) * w5.size(), h->stream));
  HIPCHK(h2d_sync(L.psol, p1.data(), sizeof(double) * p1.size()));
  return MGE_OK;
}

// host FP64 array -> device bulk array of the case's type
int put_bulk(mge_case *h, void *dev, const double *src, size_t n) {
  if (!h->fp32) {
    HIPCHK(h2d_sync(dev, src, sizeof(double) * n, h->stream));
    return MGE_OK;
  }
  std::vector<float> t(n);
  for (size_t q = 0; q < n; q++) t[q] = (float)src[q];
  HIPCHK(h2d_sync(dev, t.data(), sizeof(float) * n, h->stream));
  return MGE_OK;
}

// Transfer weights between a stretched level lvl and the coarser level
// lvl-1, from the cells' real centres (the mean of their two vertices):
// restriction -- the coarse centre between its children a, a+1 takes
// h_a/(h_a+h_{a+1}) and h_{a+1}/(h_a+h_{a+1}); prolongation -- the fine
// centre between coarse centres x0 < x1 takes (x1-x)/(x1-x0) and
// (x-x0)/(x1-x0); a ratio-1 direction takes (1, 0).  On a uniform level these
// are the exact 1/2 and 1/4, 3/4 of the integer stencils.
int stretched_weights(mge_case *h, int lvl) {
  if (lvl < 2 || lvl > h->nlevel) return MGE_OK;
  Level &f = h->lv[lvl], &c = h->lv[lvl - 1];
  if (!f.defined || !c.defined || (!f.vol && !c.vol)) return MGE_OK;
  if (!f.vol || !c.vol) return set_err(MGE_ERR_ARG, "a stretched level and a uniform level do not nest");
  const int nf[3] = {f.d.ni, f.d.nj, f.d.nk}, nc[3] = {c.d.ni, c.d.nj, c.d.nk};
  for (int a = 0; a < 3; a++) {
    const std::vector<double> &vf = f.hv[a], &vc = c.hv[a];
    std::vector<double> r((size_t)2 * (nc[a] + 4), 0.0), p((size_t)2 * (nf[a] + 4), 0.0);
    for (int I = 2; I <= nc[a] + 1; I++) {
      if (f.ratio[a] == 2) {
        const int s = 2 * I - 2;
        const double ha = vf[s + 1] - vf[s], hb = vf[s + 2] - vf[s + 1];
        r[2 * I] = ha / (ha + hb);
        r[2 * I + 1] = hb / (ha + hb);
      } else {
        r[2 * I] = 1.0;
      }
    }
    for (int i = 2; i <= nf[a] + 1; i++) {
      if (f.ratio[a] == 2) {
        const int m = (i % 2 == 0) ? i / 2 : (i - 1) / 2 + 1;
        const double x = 0.5 * (vf[i] + vf[i + 1]);
        const double x0 = 0.5 * (vc[m] + vc[m + 1]), x1 = 0.5 * (vc[m + 1] + vc[m + 2]);
        p[2 * i] = (x1 - x) / (x1 - x0);
        p[2 * i + 1] = (x - x0) / (x1 - x0);
      } else {
        p[2 * i] = 1.0;
      }
    }
    if (!f.rw[a]) CHKRC(dalloc_bytes(f, &f.rw[a], h->es * r.size()));
    if (!f.pw[a]) CHKRC(dalloc_bytes(f, &f.pw[a], h->es * p.size()));
    CHKRC(put_bulk(h, f.rw[a], r.data(), r.size()));
    CHKRC(put_bulk(h, f.pw[a], p.data(), p.size()));
  }
  return MGE_OK;
}


// ================================================================== batches
//
// A batch advances many cases together.  Every case of a batch follows the
// same cycle (same number of levels, scheme, stage count, smoothing and
// precision), so the launches its cycle makes line up one to one across
// the cases.  The batch records two cycles of each active case (two cycles
// bring every level's state buffers back to the same rotation), merges the
// recorded launches position by position into one launch of the kernel's
// batch entry point over all cases, uploads the argument structs once, and
// replays the merged program -- as a hipGraph or as a launch loop.  After
// each cycle k_converge updates every case's cycle count and residual and
// clears the case's active flag when its residual has dropped by `drop`
// (the drop test of solver.f90:238), when it fails, or at its cycle limit;
// the remaining launches of the replay skip an inactive case.  Each case
// computes exactly what its own single-case run computes.

struct BCtl {
  const double *rtrms;  // residual slots of the case's finest level
  const int *err;       // the case's device error flag
  double *log;          // residual history (cap entries)
  double ini, last, drop;
  int ncyc, maxcyc, status, cap;
  int *frozen;          // the case's freeze flags (far field, limiter)
  double ffreeze;       // its far-field freeze drop (0: never)
  double lfreeze;       // its limiter freeze drop (0: never)
};

__global__ void k_converge(BCtl *ctl, int *act, int n, int slot) {
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    if (!act[c]) continue;
    BCtl &C = ctl[c];
    double r = C.rtrms[slot];
    C.ncyc += 1;
    if (C.ncyc == 1) C.ini = r;
    C.last = r;
    if (C.ffreeze > 0.0 && C.ini / r > C.ffreeze) C.frozen[0] = 1;
    if (C.lfreeze > 0.0 && C.ini / r > C.lfreeze) C.frozen[1] = 1;
    if (C.ncyc - 1 < C.cap) C.log[C.ncyc - 1] = r;
    int e = *C.err;
    if (r != r || e < 0) {
      act[c] = 0;
      C.status = e < 0 ? e : MGE_ERR_RES_NAN;
    } else if (C.ini / r > C.drop) {
      act[c] = 0;
      C.status = 1;
    } else if (C.ncyc >= C.maxcyc) {
      act[c] = 0;
      C.status = 2;
    }
  }
}

}  // namespace

struct mge_batch {
  struct Step {
    const void *fb;      // batch entry point; null: k_converge of cycle `slot`
    dim3 block;
    unsigned nblocks;
    const void *args;    // device: nc argument structs
    const int4 *cg;      // device: nc {first block, gx, gy, slot}
    int nc, slot;
  };
  std::vector<mge_case *> cs;          // slot -> case (null after removal)
  std::vector<BCtl> hctl;
  std::vector<int> hact;
  BCtl *d_ctl = nullptr;
  int *d_act = nullptr;
  double *d_log = nullptr;
  int nslot_cap = 0, cap = 0;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  bool dirty = true;
  std::vector<Step> prog;
  std::vector<int> prog_slots;         // slots the program was built for
  std::vector<std::vector<int>> cur1;  // per program case: level buffers after cycle 1
  char *d_arena = nullptr;
  size_t arena_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nlevel = 0, fp32 = -1, distype = -1, dissip = -1, mstage = -1, mgtype = -1, smoothed = -1;
};

namespace {

int batch_grow(mge_batch *b, int need) {
  if (need <= b->nslot_cap) return MGE_OK;
  int cap = std::max(need, 2 * b->nslot_cap + 8);
  BCtl *nc = nullptr;
  int *na = nullptr;
  double *nl = nullptr;
  if (hipMalloc(&nc, sizeof(BCtl) * cap) != hipSuccess || hipMalloc(&na, sizeof(int) * cap) != hipSuccess ||
      hipMalloc(&nl, sizeof(double) * (size_t)cap * b->cap) != hipSuccess)
    return set_err(MGE_ERR_NOMEM, "batch control allocation failed");
  HIPCHK(hipStreamSynchronize(b->stream));
  b->hctl.resize(cap);
  b->hact.resize(cap, 0);
  // move the residual logs of existing slots to the new log array
  std::vector<double> lg((size_t)cap * b->cap, 0.0);
  if (b->nslot_cap) HIPCHK(d2h_sync(lg.data(), b->d_log, sizeof(double) * (size_t)b->nslot_cap * b->cap));
  HIPCHK(h2d_sync(nl, lg.data(), sizeof(double) * lg.size()));
  for (int c = 0; c < cap; c++) b->hctl[c].log = nl + (size_t)c * b->cap;
  if (b->d_ctl) hipFree(b->d_ctl);
  if (b->d_act) hipFree(b->d_act);
  if (b->d_log) hipFree(b->d_log);
  b->d_ctl = nc;
  b->d_act = na;
  b->d_log = nl;
  b->nslot_cap = cap;
  HIPCHK(h2d_sync(b->d_ctl, b->hctl.data(), sizeof(BCtl) * cap));
  HIPCHK(h2d_sync(b->d_act, b->hact.data(), sizeof(int) * cap));
  return MGE_OK;
}

void batch_drop_graph(mge_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int batch_enqueue(mge_batch *b) {
  for (const auto &st : b->prog) {
    if (!st.fb) {
      hipLaunchKernelGGL(k_converge, dim3(1), dim3(256), 0, b->stream, b->d_ctl, b->d_act, (int)b->cs.size(),
                         st.slot);
      continue;
    }
    const void *a = st.args;
    const int4 *cg = st.cg;
    int nc = st.nc;
    const int *act = b->d_act;
    void *p[4] = {(void *)&a, (void *)&cg, (void *)&nc, (void *)&act};
    HIPCHK(hipLaunchKernel(st.fb, dim3(st.nblocks), st.block, p, 0, b->stream));
  }
  return MGE_OK;
}

// record two cycles of every active case and merge them into the program
int batch_build(mge_batch *b) {
  HIPCHK(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  b->prog.clear();
  b->prog_slots.clear();
  b->cur1.clear();
  for (int c = 0; c < (int)b->cs.size(); c++)
    if (b->cs[c] && b->hact[c]) b->prog_slots.push_back(c);
  b->dirty = false;
  const int nc = (int)b->prog_slots.size();
  if (nc == 0) return MGE_OK;
  std::vector<std::vector<Rec>> R(nc);
  b->cur1.resize(nc);
  for (int i = 0; i < nc; i++) {
    mge_case *h = b->cs[b->prog_slots[i]];
    h->rec = &R[i];
    long long cs = 0;
    int rc = MGE_OK;
    // the program replays these two cycles from the buffers they start in
    std::vector<int> cur0;
    for (int l = 1; l <= h->nlevel; l++) cur0.push_back(h->lv[l].cur);
    for (int cyc = 1; cyc <= 2 && rc >= 0; cyc++) {
      rc = by_type(h, [&](auto z) { return do_cycle<decltype(z)>(h, b->nlevel, cyc, &cs); });
      Rec m;
      m.name = "k_converge";
      m.fs = m.fb = nullptr;
      m.grid = dim3(cyc);
      R[i].push_back(std::move(m));
      if (cyc == 1) {
        for (int l = 1; l <= h->nlevel; l++) b->cur1[i].push_back(h->lv[l].cur);
      }
    }
    h->rec = nullptr;
    if (rc < 0) return rc;
    for (int l = 1; l <= h->nlevel; l++)
      if (h->lv[l].cur != cur0[l - 1])
        return set_err(MGE_ERR_STATE, "two recorded cycles do not return a level's state to its buffer");
  }
  const size_t np = R[0].size();
  for (int i = 1; i < nc; i++)
    if (R[i].size() != np) return set_err(MGE_ERR_STATE, "cases of a batch must follow the same cycle");
  // per position, the cases grouped by entry point (cases whose line
  // lengths pick different scan widths launch different kernels there)
  struct Group {
    size_t p;
    std::vector<int> cases;
  };
  std::vector<Group> groups;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t total = 0;
  for (size_t p = 0; p < np; p++) {
    for (int i = 1; i < nc; i++)
      if (std::strcmp(R[i][p].name, R[0][p].name) != 0 || (R[i][p].fb == nullptr) != (R[0][p].fb == nullptr))
        return set_err(MGE_ERR_STATE, "cases of a batch must follow the same cycle");
    if (!R[0][p].fb) {
      groups.push_back(Group{p, {}});
      continue;
    }
    size_t first_group = groups.size();
    for (int i = 0; i < nc; i++) {
      const Rec &r = R[i][p];
      size_t gsel = groups.size();
      for (size_t g = first_group; g < groups.size(); g++) {
        const Rec &q = R[groups[g].cases[0]][p];
        if (q.fb == r.fb && q.block.x == r.block.x && q.block.y == r.block.y && q.block.z == r.block.z &&
            q.args.size() == r.args.size()) {
          gsel = g;
          break;
        }
      }
      if (gsel == groups.size()) groups.push_back(Group{p, {}});
      groups[gsel].cases.push_back(i);
    }
    for (size_t g = first_group; g < groups.size(); g++)
      total += al(R[groups[g].cases[0]][p].args.size() * groups[g].cases.size()) +
               al(sizeof(int4) * groups[g].cases.size());
  }
  if (total > b->arena_cap) {
    if (b->d_arena) hipFree(b->d_arena);
    b->d_arena = nullptr;
    b->arena_cap = 0;
    if (hipMalloc(&b->d_arena, total) != hipSuccess) return set_err(MGE_ERR_NOMEM, "batch arena allocation failed");
    b->arena_cap = total;
  }
  std::vector<char> host(total + 1, 0);
  size_t off = 0;
  for (const Group &G : groups) {
    const size_t p = G.p;
    if (G.cases.empty()) {
      mge_batch::Step st{};
      st.fb = nullptr;
      st.slot = (int)R[0][p].grid.x;
      b->prog.push_back(st);
      continue;
    }
    const int ng = (int)G.cases.size();
    const Rec &r0 = R[G.cases[0]][p];
    const size_t as = r0.args.size();
    const size_t aoff = off, coff = off + al(as * ng);
    off = coff + al(sizeof(int4) * ng);
    long first = 0;
    for (int k = 0; k < ng; k++) {
      const Rec &r = R[G.cases[k]][p];
      std::memcpy(&host[aoff + as * k], r.args.data(), as);
      long n = (long)r.grid.x * r.grid.y * r.grid.z;
      int4 g = make_int4((int)first, (int)r.grid.x, (int)r.grid.y, b->prog_slots[G.cases[k]]);
      std::memcpy(&host[coff + sizeof(int4) * k], &g, sizeof(int4));
      first += n;
    }
    if (first == 0) continue;  // no case of the group launches here
    if (first > 0x7fffffffL) return set_err(MGE_ERR_ARG, "batch launch too large");
    mge_batch::Step st{};
    st.fb = r0.fb;
    st.block = r0.block;
    st.nblocks = (unsigned)first;
    st.args = b->d_arena + aoff;
    st.cg = (const int4 *)(b->d_arena + coff);
    st.nc = ng;
    b->prog.push_back(st);
  }
  if (total) HIPCHK(h2d_sync(b->d_arena, host.data(), total, b->stream));
  if (b->use_graph) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc < 0) return rc;
    if (e != hipSuccess) return set_err(MGE_ERR_HIP, "hipStreamEndCapture failed");
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  return MGE_OK;
}

int batch_pull(mge_batch *b) {
  const int n = (int)b->cs.size();
  HIPCHK(hipMemcpyAsync(b->hctl.data(), b->d_ctl, sizeof(BCtl) * n, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(b->hact.data(), b->d_act, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return MGE_OK;
}

int batch_push(mge_batch *b) {
  const int n = (int)b->cs.size();
  HIPCHK(hipMemcpyAsync(b->d_ctl, b->hctl.data(), sizeof(BCtl) * n, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->d_act, b->hact.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return MGE_OK;
}

}  // namespace

// ---- fixed device scratch -----------------------------------------------
// Staging array of mge_upload_solut_noalloc: every device the library runs
// on holds its own copy for the life of the process.  In host code the name
// `scratch` is only a stand-in for the symbol; its device address comes from
// hipGetSymbolAddress, once per device (at the first mge_case_create there),
// and only that address is ever passed to kernels or copies.
static __device__ __align__(256) double scratch[MGE_SCRATCH_BYTES / sizeof(double)];

namespace {
std::mutex g_scr_mu;
std::vector<double *> g_scr_ptr;  // per device, nullptr until resolved

int scratch_on_device(int dev, double **out) {
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (dev < 0) return set_err(MGE_ERR_HIP, "bad device");
  if ((int)g_scr_ptr.size() <= dev) g_scr_ptr.resize(dev + 1, nullptr);
  if (!g_scr_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device is dev
    if (e != hipSuccess || !p) {
      char b[160];
      snprintf(b, sizeof b, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return set_err(MGE_ERR_HIP, b);
    }
    g_scr_ptr[dev] = (double *)p;
  }
  *out = g_scr_ptr[dev];
  return MGE_OK;
}
}  // namespace

// ================================================================== C ABI

extern "C" {

const char *mge_last_error(void) { return g_err.c_str(); }

int mge_case_create(const mge_params *prm, int nlevel, mge_case **out) {
  if (!prm || !out || nlevel < 1 || nlevel > 16) return set_err(MGE_ERR_ARG, "bad case arguments");
  if (prm->mstage < 1 || prm->mstage > 8) return set_err(MGE_ERR_ARG, "mstage must be 1..8");
  if (prm->distype < 0 || prm->distype > 2) return set_err(MGE_ERR_DISTYPE, "distype must be c (0), r1 (1) or r2 (2)");
  if (prm->dissip != 0 && prm->dissip != 1) return set_err(MGE_ERR_DISTYPE, "dissip must be scalar or matrix");
  if (prm->precision != 64 && prm->precision != 32 && prm->precision != 0)
    return set_err(MGE_ERR_ARG, "precision must be 64 or 32");
  if (prm->execmode != 0 && prm->execmode != 1) return set_err(MGE_ERR_ARG, "execmode must be 0 or 1");
  if (prm->execmode == 1 && !(prm->dtstep > 0.0)) return set_err(MGE_ERR_ARG, "unsteady flow needs dtstep > 0");
  if (prm->smoother != 0 && prm->smoother != 1) return set_err(MGE_ERR_ARG, "smoother must be 0 (psmoo_old) or 1 (psmoo)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(MGE_ERR_HIP, "no HIP device");
  int cur_dev = 0;
  if (hipGetDevice(&cur_dev) != hipSuccess) return set_err(MGE_ERR_HIP, "hipGetDevice failed");
  double *scr = nullptr;
  CHKRC(scratch_on_device(cur_dev, &scr));
  mge_case *h = new mge_case();
  h->prm = *prm;
  h->dev = cur_dev;
  h->scratch = scr;
  h->fp32 = prm->precision == 32;
  h->es = h->fp32 ? sizeof(float) : sizeof(double);
  h->nlevel = nlevel;
  h->lv.resize(nlevel + 1);
  h->smoo_bb = -prm->smoo;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return set_err(MGE_ERR_HIP, "hipStreamCreate failed");
  }
  if (hipMalloc(&h->err, sizeof(int)) != hipSuccess || hipMalloc(&h->ffrozen, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc(&h->ffini, sizeof(double)) != hipSuccess) {
    if (h->err) hipFree(h->err);
    if (h->ffrozen) hipFree(h->ffrozen);
    hipStreamDestroy(h->stream);
    delete h;
    return set_err(MGE_ERR_NOMEM, "hipMalloc failed");
  }
  clear_sync(h->err, sizeof(int));
  clear_sync(h->ffrozen, 2 * sizeof(int));
  clear_sync(h->ffini, sizeof(double));
  *out = h;
  return MGE_OK;
}

int mge_case_destroy(mge_case *h) {
  if (!h) return MGE_OK;
  hipStreamSynchronize(h->stream);
  for (auto &L : h->lv) level_free(L);
  for (auto &t : h->pending) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  for (auto e : h->pool) hipEventDestroy(e);
  hipFree(h->err);
  hipFree(h->ffrozen);
  hipFree(h->ffini);
  hipStreamDestroy(h->stream);
  delete h;
  return MGE_OK;
}

int mge_level_define(mge_case *h, int lvl, int ni, int nj, int nk, const double *xv, const double *yv,
                     const double *zv, double cellvolume, const int bctype[6], const int ratio[3],
                     const uint8_t *blank, int nmir, const mge_cutrec *recs, int mxcyc) {
  if (!h || lvl < 1 || lvl > h->nlevel) return set_err(MGE_ERR_ARG, "bad level");
  if (ni < 1 || nj < 1 || nk < 1 || !xv || !yv || !zv || !bctype || !blank || nmir < 0 ||
      (nmir > 0 && !recs))
    return set_err(MGE_ERR_ARG, "bad level arguments");
  Level &L = h->lv[lvl];
  level_free(L);
  Dims &d = L.d;
  d.ni = ni; d.nj = nj; d.nk = nk;
  d.il = ni + 1; d.jl = nj + 1; d.kl = nk + 1;
  d.ie = ni + 2; d.je = nj + 2; d.ke = nk + 2;
  d.ib = ni + 3; d.jb = nj + 3; d.kb = nk + 3;
  d.sy = d.ib + 1;
  d.sz = (long)(d.ib + 1) * (d.jb + 1);
  d.E = d.sz * (d.kb + 1);
  const long E = d.E;
  const size_t es = h->es;
  L.cellvolume = cellvolume;
  for (int f = 0; f < 6; f++) L.bctype[f] = bctype[f];
  for (int a = 0; a < 3; a++) L.ratio[a] = ratio ? ratio[a] : 2;
  for (int a = 0; a < 3; a++)
    if (L.ratio[a] != 1 && L.ratio[a] != 2) return set_err(MGE_ERR_ARG, "ratio must be 1 or 2");
  L.mxcyc = mxcyc > 0 ? mxcyc : 1;
  for (int b = 0; b < 3; b++) CHKRC(dalloc_bytes(L, &L.S[b], es * 6 * E));
  CHKRC(dalloc_bytes(L, &L.ws, es * 5 * E));
  CHKRC(dalloc_bytes(L, &L.wr, es * 5 * E));
  CHKRC(dalloc_bytes(L, &L.qf, es * 5 * E));
  CHKRC(dalloc_bytes(L, &L.dw, es * 5 * E));
  CHKRC(dalloc_bytes(L, &L.fw, es * 5 * E));
  CHKRC(dalloc_bytes(L, &L.radsum, es * E));
  CHKRC(dalloc_bytes(L, &L.vdtim, es * E));
  if (h->prm.smoother == 1 && h->prm.smoo > 0.0) CHKRC(dalloc_bytes(L, &L.eps, es * 4 * E));
  CHKRC(dalloc(L, &L.flags, E));
  CHKRC(dalloc(L, &L.marker, E));
  int nmax = std::max(d.ib, std::max(d.jb, d.kb)) + 2;
  CHKRC(dalloc_bytes(L, &L.dx, es * nmax));
  CHKRC(dalloc_bytes(L, &L.dy, es * nmax));
  CHKRC(dalloc_bytes(L, &L.dz, es * nmax));
  L.npiv = nmax;
  CHKRC(dalloc_bytes(L, &L.piv, es * 9 * nmax));
  CHKRC(dalloc_bytes(L, &L.fax, es * nmax));
  dim3 g = cell_grid(d);
  L.npartial = (long)g.x * g.y * ((g.z + kRtrmsPlanes - 1) / kRtrmsPlanes);
  CHKRC(dalloc(L, &L.partial, L.npartial));
  CHKRC(dalloc(L, &L.dtim_bits, 1));
  CHKRC(dalloc(L, &L.dtim, 1));
  CHKRC(dalloc(L, &L.rtrms, L.mxcyc + 1));
  CHKRC(dalloc(L, &L.force, 3));
  if (h->prm.ffreeze > 0.0) {
    for (int f = 1; f <= 6; f++) {
      const int dir = (f - 1) % 3;
      const long amax = dir == 0 ? d.jb : d.ib, cmax = dir == 2 ? d.jb : d.kb;
      L.ffoff[f] = L.ffoff[f - 1] + (amax + 1) * (cmax + 1);
    }
    CHKRC(dalloc(L, &L.ffdec, L.ffoff[6]));
  }
  if (h->prm.lfreeze > 0.0 && h->prm.distype == 2) {
    // every byte starts "not recorded" (255): the ghost cells, whose values
    // no stage records, keep the limiter itself
    CHKRC(dalloc(L, &L.lim, 3 * E));
    HIPCHK(fill_sync(L.lim, 0xff, (size_t)E * 3 * sizeof(unsigned long long)));
  }
  if (h->prm.eqntype == 1) CHKRC(dalloc_bytes(L, &L.vprim, es * 4 * E));
  if (h->prm.execmode == 1) {
    CHKRC(dalloc_bytes(L, &L.w1, es * 5 * E));
    CHKRC(dalloc_bytes(L, &L.w2, es * 5 * E));
  }
  if (lvl < h->nlevel) CHKRC(dalloc_bytes(L, &L.corr, es * 5 * E));
  // vertex differences (ghost entries of index 0 stay unused)
  std::vector<double> hx(nmax, 0.0), hy(nmax, 0.0), hz(nmax, 0.0);
  for (int i = 1; i < d.ib; i++) hx[i] = xv[i + 1] - xv[i];
  for (int j = 1; j < d.jb; j++) hy[j] = yv[j + 1] - yv[j];
  for (int k = 1; k < d.kb; k++) hz[k] = zv[k + 1] - zv[k];
  CHKRC(put_bulk(h, L.dx, hx.data(), nmax));
  CHKRC(put_bulk(h, L.dy, hy.data(), nmax));
  CHKRC(put_bulk(h, L.dz, hz.data(), nmax));
  // a stretched tensor level (cellvolume <= 0; the reference block has one
  // cellvolume): each interior cell's volume is the product of its vertex
  // differences, ghost cells 1
  if (!(cellvolume > 0.0)) {
    L.cellvolume = 0.0;
    std::vector<double> v((size_t)E, 1.0);
    for (int k = 2; k <= d.kl; k++)
      for (int j = 2; j <= d.jl; j++)
        for (int i = 2; i <= d.il; i++)
          v[(size_t)ix(d, i, j, k)] = ((xv[i + 1] - xv[i]) * (yv[j + 1] - yv[j])) * (zv[k + 1] - zv[k]);
    CHKRC(dalloc_bytes(L, &L.vol, es * E));
    CHKRC(put_bulk(h, L.vol, v.data(), (size_t)E));
    L.hv[0].assign(xv, xv + d.ib + 1);
    L.hv[1].assign(yv, yv + d.jb + 1);
    L.hv[2].assign(zv, zv + d.kb + 1);
  }
  // smoothing pivots (psmoo.f90:134-166), computed once on the host in FP64
  std::vector<double> piv((size_t)9 * nmax, 0.0);
  if (h->prm.smoo > 0.0) {
    double aa = 1.0 + 2.0 * h->prm.smoo, bb = -h->prm.smoo;
    int nn[3] = {ni, nj, nk};
    for (int a = 0; a < 3; a++) {
      double *e1 = &piv[(size_t)(2 * a) * nmax], *e2 = &piv[(size_t)(2 * a + 1) * nmax];
      int lm = nn[a], l = nn[a] + 1;
      if (lm < 2) continue;
      e1[2] = aa;
      e2[2] = bb / aa;
      for (int s = 3; s <= lm; s++) {
        e1[s] = aa - bb * e2[s - 1];
        e2[s] = bb / e1[s];
      }
      e1[l] = aa - bb * e2[lm];
      double *r1 = &piv[(size_t)(6 + a) * nmax];
      for (int s = 2; s <= l; s++) r1[s] = 1.0 / e1[s];
    }
  }
  CHKRC(put_bulk(h, L.piv, piv.data(), piv.size()));
  std::vector<double> fa(nmax, 0.0);
  if (h->prm.smoo > 0.0)
    for (int i = 2; i <= d.il && i < nmax; i++) fa[i] = h->prm.smoo / piv[i];  // -bb/e1, bb = -smoo
  CHKRC(put_bulk(h, L.fax, fa.data(), nmax));
  // flags: fluid bit with the ghost copies of init_metric.f90:207-218
  std::vector<uint8_t> bl(E, 1);
  auto X = [&](int i, int j, int k) { return ix(d, i, j, k); };
  for (int k = 2; k <= d.kl; k++)
    for (int j = 2; j <= d.jl; j++)
      for (int i = 2; i <= d.il; i++)
        bl[X(i, j, k)] = blank[(long)(i - 2) + (long)ni * ((long)(j - 2) + (long)nj * (k - 2))] ? 1 : 0;
  for (int k = 0; k <= d.kb; k++)
    for (int j = 0; j <= d.jb; j++) {
      bl[X(0, j, k)] = bl[X(2, j, k)];
      bl[X(1, j, k)] = bl[X(2, j, k)];
      bl[X(d.ie, j, k)] = bl[X(d.il, j, k)];
      bl[X(d.ib, j, k)] = bl[X(d.il, j, k)];
    }
  for (int k = 0; k <= d.kb; k++)
    for (int i = 0; i <= d.ib; i++) {
      bl[X(i, 0, k)] = bl[X(i, 2, k)];
      bl[X(i, 1, k)] = bl[X(i, 2, k)];
      bl[X(i, d.je, k)] = bl[X(i, d.jl, k)];
      bl[X(i, d.jb, k)] = bl[X(i, d.jl, k)];
    }
  for (int j = 0; j <= d.jb; j++)
    for (int i = 0; i <= d.ib; i++) {
      bl[X(i, j, 0)] = bl[X(i, j, 2)];
      bl[X(i, j, 1)] = bl[X(i, j, 2)];
      bl[X(i, j, d.ke)] = bl[X(i, j, d.kl)];
      bl[X(i, j, d.kb)] = bl[X(i, j, d.kl)];
    }
  HIPCHK(h2d_sync(L.flags, bl.data(), E, h->stream));
  L.cur = 0;
  L.defined = true;
  int rc = build_records(h, L, nmir, recs);
  if (rc >= 0) rc = stretched_weights(h, lvl);
  if (rc >= 0) rc = stretched_weights(h, lvl + 1);
  if (rc < 0) {
    level_free(L);
    return rc;
  }
  return MGE_OK;
}

int mge_init_flow(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  HIPCHK(hipMemsetAsync(h->ffrozen, 0, 2 * sizeof(int), h->stream));  // a new solve chooses afresh
  int rc = by_type(h, [&](auto z) { return do_init_flow<decltype(z)>(h); });
  if (rc < 0) return rc;
  int e = check_err(h);
  return e < 0 ? e : rc;
}

int mge_timestep(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return by_type(h, [&](auto z) { return do_timestep<decltype(z)>(h, *L); });
}

int mge_euler(mge_case *h, int lvl, int blk, int itime, int ncyc, int mgl, int mlvl, int lres, int lfir,
              double *rtrms) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || itime < 0) return set_err(MGE_ERR_ARG, "bad level/block");
  h->itime = itime;
  int rc = by_type(h, [&](auto z) { return do_euler<decltype(z)>(h, *L, ncyc, mgl, mlvl, lres, lfir); });
  h->itime = 0;
  if (rc < 0) return rc;
  if (rtrms && !lres) {
    int c = ncyc >= 1 && ncyc <= L->mxcyc ? ncyc : 0;
    HIPCHK(hipMemcpyAsync(rtrms, L->rtrms + c, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    int e = check_err(h);
    if (e < 0) return e;
  }
  return rc;
}

int mge_bcond(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return by_type(h, [&](auto z) { return do_bcond<decltype(z)>(h, *L, L->cur); });
}

int mge_bc_cutcell(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  return by_type(h, [&](auto z) { return do_bc_cutcell<decltype(z)>(h, *L, L->cur); });
}

int mge_bc_overlap(mge_case *h, int lvl) {
  // one block per level: no same-level overlaps (bc_overlap.f90:32-90 loops over none)
  if (!get_level(h, lvl)) return set_err(MGE_ERR_ARG, "bad level");
  return MGE_OK;
}

int mge_movco(mge_case *h, int mgl) {
  if (!get_level(h, mgl) || !get_level(h, mgl + 1)) return set_err(MGE_ERR_ARG, "bad level");
  return by_type(h, [&](auto z) { return do_movco<decltype(z)>(h, mgl); });
}

int mge_movfin(mge_case *h, int mgl, int mlvl) {
  if (!get_level(h, mgl) || !get_level(h, mgl - 1)) return set_err(MGE_ERR_ARG, "bad level");
  return by_type(h, [&](auto z) { return do_movfin<decltype(z)>(h, mgl, mlvl); });
}

int mge_solve(mge_case *h, const int *mcyc, double *reslog, int cap, int *nlog, long long *cell_stages) {
  if (!h || !mcyc) return set_err(MGE_ERR_ARG, "bad solve arguments");
  for (int l = 1; l <= h->nlevel; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  return by_type(h, [&](auto z) { return do_solve<decltype(z)>(h, mcyc, 0, reslog, cap, nlog, cell_stages); });
}

int mge_solve_step(mge_case *h, const int *mcyc, int itime, double *reslog, int cap, int *nlog,
                   long long *cell_stages) {
  if (!h || !mcyc || itime < 0) return set_err(MGE_ERR_ARG, "bad solve arguments");
  for (int l = 1; l <= h->nlevel; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  CHKRC(by_type(h, [&](auto z) { return do_solve<decltype(z)>(h, mcyc, itime, reslog, cap, nlog, cell_stages); }));
  if (h->prm.execmode == 1) {
    CHKRC(shift_time_levels(h));
    CHKRC(check_err(h));
  }
  return MGE_OK;
}

int mge_shift_time_levels(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  if (h->prm.execmode != 1) return set_err(MGE_ERR_STATE, "time levels exist for unsteady cases only");
  CHKRC(shift_time_levels(h));
  return check_err(h);
}

int mge_cycles(mge_case *h, int mlvl, int ncyc0, int count, double *res, long long *cell_stages) {
  if (!h || mlvl < 1 || mlvl > h->nlevel || count < 0) return set_err(MGE_ERR_ARG, "bad cycle arguments");
  for (int l = 1; l <= mlvl; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  long long cs = 0;
  for (int c = 0; c < count; c++) {
    CHKRC(by_type(h, [&](auto z) { return do_cycle<decltype(z)>(h, mlvl, ncyc0 + c, &cs); }));
    CHKRC(ffreeze_test(h, mlvl, ncyc0 + c));
  }
  if (res && count > 0) CHKRC(read_rtrms(h, mlvl, ncyc0 + count - 1, res));
  else CHKRC(check_err(h));
  if (cell_stages) *cell_stages = cs;
  return MGE_OK;
}

// fields holding a state (FP32 stores them as the deviation from the
// freestream): "w" and "ws" (5 components), "p" (1)
static int state_field(const char *name) {
  if (!strcmp(name, "w") || !strcmp(name, "ws") || !strcmp(name, "w1") || !strcmp(name, "w2")) return 5;
  if (!strcmp(name, "p")) return 1;
  return 0;
}

static void *field_ptr(Level &L, size_t es, const char *name, long *n) {
  const long E = L.d.E;
  *n = 5 * E;
  if (!strcmp(name, "w")) return L.S[L.cur];
  if (!strcmp(name, "dw")) return L.dw;
  if (!strcmp(name, "fw")) return L.fw;
  if (!strcmp(name, "ws")) return L.ws;
  if (!strcmp(name, "wr")) return L.wr;
  if (!strcmp(name, "qf")) return L.qf;
  if (!strcmp(name, "w1")) return L.w1;
  if (!strcmp(name, "w2")) return L.w2;
  *n = E;
  if (!strcmp(name, "p")) return (char *)L.S[L.cur] + es * 5 * E;
  if (!strcmp(name, "vdtim")) return L.vdtim;
  if (!strcmp(name, "radsum")) return L.radsum;
  return nullptr;
}

// the level's corner cells formed before the host reads or writes its state
int sync_corners(mge_case *h, Level &L) {
  if (!L.corner_stale) return MGE_OK;
  by_type(h, [&](auto z) {
    ensure_corners<decltype(z)>(h, L);
    return MGE_OK;
  });
  return MGE_OK;
}

int mge_download(mge_case *h, int lvl, int blk, const char *field, double *out) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !field || !out) return set_err(MGE_ERR_ARG, "bad download arguments");
  sync_corners(h, *L);
  if (!strcmp(field, "lim")) {  // limiter bytes (limiter freezing): out[(5 dir + n) E + q]
    if (!L->lim) return set_err(MGE_ERR_ARG, "the level has no limiter words (lfreeze = 0 or not Roe r2)");
    const long E = L->d.E;
    std::vector<unsigned long long> t(3 * E);
    HIPCHK(hipMemcpyAsync(t.data(), L->lim, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost,
                          h->stream));
    int e = check_err(h);
    for (int d = 0; d < 3; d++)
      for (int n = 0; n < 5; n++)
        for (long q = 0; q < E; q++) out[(long)(5 * d + n) * E + q] = (double)((t[d * E + q] >> (8 * n)) & 255u);
    return e;
  }
  long n;
  void *p = field_ptr(*L, h->es, field, &n);
  if (!p) return set_err(MGE_ERR_ARG, "unknown field");
  if (!h->fp32) {
    HIPCHK(hipMemcpyAsync(out, p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    return check_err(h);
  }
  std::vector<float> t(n);
  HIPCHK(hipMemcpyAsync(t.data(), p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  int e = check_err(h);
  const int nc = state_field(field);
  const Inf<float> f = make_inf<float>(h->prm);
  const long E = L->d.E;
  for (long q = 0; q < n; q++) {
    double off = nc == 5 ? (double)f.w[q / E] : (nc == 1 ? (double)f.p : 0.0);
    out[q] = (double)t[q] + off;
  }
  return e;
}

int mge_upload(mge_case *h, int lvl, int blk, const char *field, const double *in) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !field || !in) return set_err(MGE_ERR_ARG, "bad upload arguments");
  sync_corners(h, *L);
  long n;
  void *p = field_ptr(*L, h->es, field, &n);
  if (!p) return set_err(MGE_ERR_ARG, "unknown field");
  HIPCHK(hipStreamSynchronize(h->stream));
  const int nc = state_field(field);
  if (!h->fp32 || nc == 0) return put_bulk(h, p, in, n);
  const Inf<float> f = make_inf<float>(h->prm);
  const long E = L->d.E;
  std::vector<float> t(n);
  for (long q = 0; q < n; q++) {
    double off = nc == 5 ? (double)f.w[q / E] : (double)f.p;
    t[q] = (float)(in[q] - off);
  }
  HIPCHK(h2d_sync(p, t.data(), sizeof(float) * n, h->stream));
  return MGE_OK;
}

// ---- restart files (srcf90/wsolut.f90:47-93, srcf90/rsolut.f90:1-41) -------
//
// Text (format 0), as write_solut/read_solut: per block, levels 1..nlevel
// with the block code 100*lvl + 1 (the .pin's GRID.NAME convention LNN),
// a "(5I5)" header code, low(4), upp(4) with low = (0,0,0,1) and
// upp = (ib,jb,kb,5), then w(0:ib,0:jb,0:kb,1:5) i fastest in "(5f16.7)".
// That file carries w only: reading it recomputes p from w on every cell
// and the wall mirror states from the restored state (bc_cutcell).
// Binary (format 1), the full-precision variant: "MGESOL2\0", nlevel, then
// per block code, ib, jb, kb, nmir (int32), w (5E doubles, same order), p
// (E doubles) and the mirror states of the wall records (5 nmir + nmir
// doubles): a run restarted from it continues exactly as the run that wrote
// it.  Reading checks code and dimensions (the reference's "INCONSISTENT
// FILE" stop).
}  // extern "C"
namespace {
template <class T>
int restore_level(mge_case *h, Level &L, bool mirror) {
  LAUNCHT(h, "k_eos_all", k_eos_all, T, flat_grid(L.d.E), dim3(256),
          (EosArgs<T>{L.d, stm<T>(L, L.cur), (T)h->prm.gamma, make_inf<T>(h->prm)}));
  if (mirror) {
    int rc = do_bc_cutcell<T>(h, L, L.cur);
    if (rc < 0) return rc;
  }
  return check_err(h);
}
}  // namespace
extern "C" {

int mge_write_solut(mge_case *h, const char *path, int format) {
  if (!h || !path || (format != 0 && format != 1)) return set_err(MGE_ERR_ARG, "bad write_solut arguments");
  for (int l = 1; l <= h->nlevel; l++)
    if (h->lv[l].defined) sync_corners(h, h->lv[l]);
  FILE *f = std::fopen(path, format ? "wb" : "w");
  if (!f) return set_err(MGE_ERR_ARG, "cannot open the restart file for writing");
  int rc = MGE_OK;
  if (format == 1) {
    const char magic[8] = {'M', 'G', 'E', 'S', 'O', 'L', '2', 0};
    std::fwrite(magic, 1, 8, f);
    int32_t nl = h->nlevel;
    std::fwrite(&nl, 4, 1, f);
  }
  for (int lvl = 1; lvl <= h->nlevel && rc == MGE_OK; lvl++) {
    Level &L = h->lv[lvl];
    const Dims &d = L.d;
    std::vector<double> w((size_t)5 * d.E);
    rc = mge_download(h, lvl, 1, "w", w.data());
    if (rc < 0) break;
    const int code = 100 * lvl + 1;
    if (format == 1) {
      std::vector<double> p(d.E), ws((size_t)5 * L.nmir), ps(L.nmir);
      rc = mge_download(h, lvl, 1, "p", p.data());
      if (rc < 0) break;
      if (L.nmir > 0) {
        HIPCHK(hipMemcpyAsync(ws.data(), L.wsol, sizeof(double) * ws.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(ps.data(), L.psol, sizeof(double) * ps.size(), hipMemcpyDeviceToHost, h->stream));
        rc = check_err(h);
        if (rc < 0) break;
      }
      int32_t hd[5] = {code, d.ib, d.jb, d.kb, L.nmir};
      std::fwrite(hd, 4, 5, f);
      std::fwrite(w.data(), sizeof(double), w.size(), f);
      std::fwrite(p.data(), sizeof(double), p.size(), f);
      std::fwrite(ws.data(), sizeof(double), ws.size(), f);
      std::fwrite(ps.data(), sizeof(double), ps.size(), f);
    } else {
      const int hdr[9] = {code, 0, 0, 0, 1, d.ib, d.jb, d.kb, 5};
      for (int m = 0; m < 9; m++) {
        std::fprintf(f, "%5d", hdr[m]);
        if (m == 4 || m == 8) std::fputc('\n', f);
      }
      for (size_t q = 0; q < w.size(); q++) {
        std::fprintf(f, "%16.7f", w[q]);
        if (q % 5 == 4 || q + 1 == w.size()) std::fputc('\n', f);
      }
    }
  }
  if (std::fclose(f) != 0 && rc == MGE_OK) rc = set_err(MGE_ERR_ARG, "writing the restart file failed");
  return rc;
}

int mge_read_solut(mge_case *h, const char *path, int format) {
  if (!h || !path || (format != 0 && format != 1)) return set_err(MGE_ERR_ARG, "bad read_solut arguments");
  FILE *f = std::fopen(path, format ? "rb" : "r");
  if (!f) return set_err(MGE_ERR_ARG, "cannot open the restart file");
  // init_flow(lrflow = .true.) (srcf90/init_flow.f90:25-31): everything the
  // freestream initialisation sets up, then the state from the file
  int rc = mge_init_flow(h);
  if (rc < 0) {
    std::fclose(f);
    return rc;
  }
  if (format == 1) {
    char magic[8];
    int32_t nl = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "MGESOL2", 8) != 0 || std::fread(&nl, 4, 1, f) != 1 ||
        nl != h->nlevel)
      rc = set_err(MGE_ERR_RESTART, "restart file: not a binary solution of this case");
  }
  for (int lvl = 1; lvl <= h->nlevel && rc >= 0; lvl++) {
    Level &L = h->lv[lvl];
    const Dims &d = L.d;
    std::vector<double> w((size_t)5 * d.E);
    int hdr[9] = {0};
    int nmir = L.nmir;
    bool ok;
    if (format == 1) {
      int32_t hd[5];
      ok = std::fread(hd, 4, 5, f) == 5;
      hdr[0] = hd[0];
      hdr[5] = hd[1];
      hdr[6] = hd[2];
      hdr[7] = hd[3];
      hdr[8] = 5;
      nmir = hd[4];
    } else {
      ok = true;
      for (int m = 0; m < 9 && ok; m++) ok = std::fscanf(f, "%d", &hdr[m]) == 1;
    }
    // rsolut.f90:28-32
    if (!ok || hdr[0] != 100 * lvl + 1 || hdr[5] != d.ib || hdr[6] != d.jb || hdr[7] != d.kb || hdr[8] != 5 ||
        nmir != L.nmir) {
      rc = set_err(MGE_ERR_RESTART, "INCONSISTENT FILE DOF & GRID DIMENSION");
      break;
    }
    if (format == 1) {
      ok = std::fread(w.data(), sizeof(double), w.size(), f) == w.size();
    } else {
      for (size_t q = 0; q < w.size() && ok; q++) ok = std::fscanf(f, "%lf", &w[q]) == 1;
    }
    if (!ok) {
      rc = set_err(MGE_ERR_RESTART, "restart file: short data");
      break;
    }
    rc = mge_upload(h, lvl, 1, "w", w.data());
    if (rc < 0) break;
    if (format == 1) {
      std::vector<double> p(d.E), ws((size_t)5 * nmir), ps(nmir);
      ok = std::fread(p.data(), sizeof(double), p.size(), f) == p.size() &&
           std::fread(ws.data(), sizeof(double), ws.size(), f) == ws.size() &&
           std::fread(ps.data(), sizeof(double), ps.size(), f) == ps.size();
      if (!ok) {
        rc = set_err(MGE_ERR_RESTART, "restart file: short data");
        break;
      }
      rc = mge_upload(h, lvl, 1, "p", p.data());
      if (rc < 0) break;
      if (nmir > 0) {
        HIPCHK(h2d_sync(L.wsol, ws.data(), sizeof(double) * ws.size(), h->stream));
        HIPCHK(h2d_sync(L.psol, ps.data(), sizeof(double) * ps.size()));
      }
    } else {
      rc = by_type(h, [&](auto z) { return restore_level<decltype(z)>(h, L, true); });
      if (rc < 0) break;
    }
    // unsteady: the earlier time levels start from the restart state (init_flow.f90:79-86)
    if (h->prm.execmode == 1) rc = init_time_levels(h, L);
  }
  std::fclose(f);
  return rc < 0 ? rc : MGE_OK;
}

int mge_get_cuts(mge_case *h, int lvl, int blk, mge_cutrec *out) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1 || !out) return set_err(MGE_ERR_ARG, "bad arguments");
  int n = L->nmir;
  if (n == 0) return 0;
  std::vector<double> w5((size_t)5 * n), p1(n);
  HIPCHK(hipMemcpyAsync(w5.data(), L->wsol, sizeof(double) * w5.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(p1.data(), L->psol, sizeof(double) * p1.size(), hipMemcpyDeviceToHost, h->stream));
  int e = check_err(h);
  if (e < 0) return e;
  for (int m = 0; m < n; m++) {
    out[m] = L->hrecs[m];
    for (int c = 0; c < 5; c++) out[m].wsol[c] = w5[(size_t)c * n + m];
    out[m].psol = p1[m];
  }
  return n;
}

double mge_get_dtim(mge_case *h, int lvl, int blk) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return NAN;
  double v = NAN;
  hipMemcpyAsync(&v, L->dtim, sizeof(double), hipMemcpyDeviceToHost, h->stream);
  hipStreamSynchronize(h->stream);
  return v;
}

int mge_wall_force(mge_case *h, int lvl, double out[3]) {
  Level *L = get_level(h, lvl);
  if (!L || !out) return set_err(MGE_ERR_ARG, "bad arguments");
  if (L->nmir == 0) {
    out[0] = out[1] = out[2] = 0.0;
    return MGE_OK;
  }
  by_type(h, [&](auto z) { return do_wall_force<decltype(z)>(h, *L); });
  HIPCHK(hipMemcpyAsync(out, L->force, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
  return check_err(h);
}

// ---- solution fields (upload_solut.f90:89-241) -------------------------------
}  // extern "C"
namespace {
// freestream Mach number from the nondimensional state (|U| / c0)
double freestream_mach(const mge_params &P) {
  return std::sqrt(P.u0 * P.u0 + P.v0 * P.v0 + P.w0 * P.w0) / P.c0;
}

template <class T>
int do_upload_solut(mge_case *h, Level &L, double ptot, double *host, int nval) {
  const long nc = (long)L.d.ni * L.d.nj * L.d.nk;
  double *dev = nullptr;
  HIPCHK(hipMallocAsync((void **)&dev, sizeof(double) * nval * nc, h->stream));
  SolutArgs<T> A{L.d, stc<T>(L, L.cur), make_inf<T>(h->prm), h->prm.gamma, h->prm.p0,
                 freestream_mach(h->prm), ptot, nval == 11, nc, dev};
  hipLaunchKernelGGL(k_upload_solut<T>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, h->stream, A);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, dev, sizeof(double) * nval * nc, hipMemcpyDeviceToHost, h->stream);
  hipError_t f = hipFreeAsync(dev, h->stream);
  if (e == hipSuccess) e = f;
  if (e != hipSuccess) return set_err(MGE_ERR_HIP, hipGetErrorString(e));
  return check_err(h);
}

// the same fields through the fixed scratch array: chunks of whole k-planes,
// each written by one launch and copied out field by field; no allocation
template <class T>
int do_upload_solut_noalloc(mge_case *h, Level &L, double ptot, double *host, int nval) {
  const long plane = (long)L.d.ni * L.d.nj;
  const long nc = plane * L.d.nk;
  const long cap = (long)(MGE_SCRATCH_BYTES / sizeof(double));
  const long pc = std::min<long>(cap / (plane * nval), L.d.nk);  // k-planes per chunk (>= 1, checked)
  for (long k0 = 0; k0 < L.d.nk; k0 += pc) {
    const long c0 = k0 * plane;
    const long ncc = std::min<long>(pc, L.d.nk - k0) * plane;
    SolutArgs<T> A{L.d, stc<T>(L, L.cur), make_inf<T>(h->prm), h->prm.gamma, h->prm.p0,
                   freestream_mach(h->prm), ptot, nval == 11, ncc, h->scratch};
    A.c0 = c0;
    A.ostride = ncc;
    hipLaunchKernelGGL(k_upload_solut<T>, dim3((unsigned)((ncc + 255) / 256)), dim3(256), 0, h->stream, A);
    hipError_t e = hipGetLastError();
    for (int f = 0; f < nval && e == hipSuccess; f++)
      e = hipMemcpyAsync(host + (long)f * nc + c0, h->scratch + (long)f * ncc, sizeof(double) * ncc,
                         hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return set_err(MGE_ERR_HIP, hipGetErrorString(e));
  }
  return check_err(h);
}
}  // namespace
extern "C" {

int mge_upload_solut_noalloc(mge_case *h, int lvl, int blk, double p_tot, double *solut, int *nval) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  sync_corners(h, *L);
  const int nv = h->prm.eqntype == 1 ? 11 : 7;
  if (nval) *nval = nv;
  if (!solut) return MGE_OK;
  const double plane_bytes = (double)L->d.ni * L->d.nj * nv * sizeof(double);
  if (plane_bytes > (double)MGE_SCRATCH_BYTES) {
    char b[200];
    snprintf(b, sizeof b,
             "mge_upload_solut_noalloc: one k-plane of output fields (%.0f bytes) exceeds the %llu-byte scratch",
             plane_bytes, (unsigned long long)MGE_SCRATCH_BYTES);
    return set_err(MGE_ERR_ARG, b);
  }
  if (!h->scratch) return set_err(MGE_ERR_STATE, "device scratch not resolved");
  const mge_params &P = h->prm;
  if (!(p_tot > 0.0)) {
    const double m = freestream_mach(P);
    p_tot = P.p0 * std::pow(1.0 + 0.2 * m * m, 3.5);
  }
  return by_type(h, [&](auto z) { return do_upload_solut_noalloc<decltype(z)>(h, *L, p_tot, solut, nv); });
}

int mge_upload_solut(mge_case *h, int lvl, int blk, double p_tot, double *solut, int *nval) {
  Level *L = get_level(h, lvl);
  if (!L || blk != 1) return set_err(MGE_ERR_ARG, "bad level/block");
  sync_corners(h, *L);
  const int nv = h->prm.eqntype == 1 ? 11 : 7;
  if (nval) *nval = nv;
  if (!solut) return MGE_OK;
  const mge_params &P = h->prm;
  if (!(p_tot > 0.0)) {
    const double m = freestream_mach(P);
    p_tot = P.p0 * std::pow(1.0 + 0.2 * m * m, 3.5);
  }
  return by_type(h, [&](auto z) { return do_upload_solut<decltype(z)>(h, *L, p_tot, solut, nv); });
}

// ---- batches ---------------------------------------------------------------

int mge_batch_create(int use_graph, int log_cap, mge_batch **out) {
  if (!out || log_cap < 0) return set_err(MGE_ERR_ARG, "bad batch arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(MGE_ERR_HIP, "no HIP device");
  mge_batch *b = new mge_batch();
  b->use_graph = use_graph != 0;
  b->cap = log_cap > 0 ? log_cap : 1;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return set_err(MGE_ERR_HIP, "hipStreamCreate failed");
  }
  *out = b;
  return MGE_OK;
}

int mge_batch_destroy(mge_batch *b) {
  if (!b) return MGE_OK;
  hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_arena) hipFree(b->d_arena);
  if (b->d_ctl) hipFree(b->d_ctl);
  if (b->d_act) hipFree(b->d_act);
  if (b->d_log) hipFree(b->d_log);
  hipStreamDestroy(b->stream);
  delete b;
  return MGE_OK;
}

int mge_batch_add(mge_batch *b, mge_case *h, double drop, int maxcyc) {
  if (!b || !h || maxcyc < 1) return set_err(MGE_ERR_ARG, "bad batch arguments");
  for (int l = 1; l <= h->nlevel; l++)
    if (!h->lv[l].defined) return set_err(MGE_ERR_STATE, "level not defined");
  if (h->lv[h->nlevel].mxcyc < 2) return set_err(MGE_ERR_ARG, "a batched case needs mxcyc >= 2");
  if (h->prm.execmode != 0) return set_err(MGE_ERR_ARG, "unsteady cases are solved one at a time, not batched");
  const int sm = h->prm.smoo > 0.0 ? 1 + h->prm.smoother : 0;  // smoothing off, psmoo_old, psmoo
  if (b->fp32 < 0) {
    b->nlevel = h->nlevel;
    b->fp32 = h->fp32;
    b->distype = h->prm.distype;
    b->dissip = h->prm.dissip;
    b->mstage = h->prm.mstage;
    b->mgtype = h->prm.mgtype;
    b->smoothed = sm;
  } else if (b->nlevel != h->nlevel || b->fp32 != (int)h->fp32 || b->distype != h->prm.distype ||
             b->dissip != h->prm.dissip || b->mstage != h->prm.mstage || b->mgtype != h->prm.mgtype ||
             b->smoothed != sm) {
    return set_err(MGE_ERR_ARG, "cases of a batch need the same levels, scheme, stages, cycle and precision");
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  int slot = (int)b->cs.size();
  CHKRC(batch_grow(b, slot + 1));
  b->cs.push_back(h);
  BCtl &C = b->hctl[slot];
  double *log = C.log;
  C = BCtl{};
  C.log = log;
  C.rtrms = h->lv[h->nlevel].rtrms;
  C.err = h->err;
  C.drop = drop;
  C.maxcyc = maxcyc;
  C.cap = b->cap;
  C.frozen = h->ffrozen;
  C.ffreeze = h->prm.ffreeze;
  C.lfreeze = h->prm.lfreeze;
  b->hact[slot] = 1;
  CHKRC(batch_push(b));
  b->dirty = true;
  return slot;
}

int mge_batch_remove(mge_batch *b, int slot) {
  if (!b || slot < 0 || slot >= (int)b->cs.size()) return set_err(MGE_ERR_ARG, "bad batch slot");
  CHKRC(batch_pull(b));
  b->hact[slot] = 0;
  b->cs[slot] = nullptr;
  CHKRC(batch_push(b));
  b->dirty = true;
  return MGE_OK;
}

int mge_batch_run(mge_batch *b, int periods, int *nactive) {
  if (!b || periods < 0) return set_err(MGE_ERR_ARG, "bad batch arguments");
  for (int p = 0; p < periods; p++) {
    // rebuild when members changed or a quarter of the program has stopped
    int stopped = 0;
    for (int s : b->prog_slots)
      if (!b->hact[s]) stopped++;
    if (b->dirty || (stopped > 0 && 4 * stopped >= (int)b->prog_slots.size())) CHKRC(batch_build(b));
    if (b->prog.empty()) break;
    const int np = (int)b->prog_slots.size();
    std::vector<int> n0(np), a0(np);
    for (int i = 0; i < np; i++) {
      n0[i] = b->hctl[b->prog_slots[i]].ncyc;
      a0[i] = b->hact[b->prog_slots[i]];
    }
    if (b->use_graph) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    else CHKRC(batch_enqueue(b));
    CHKRC(batch_pull(b));
    // every case active at the start of a period runs at least one cycle
    for (int i = 0; i < np; i++)
      if (a0[i] && b->hctl[b->prog_slots[i]].ncyc == n0[i])
        return set_err(MGE_ERR_STATE, "batch period made no progress");
    int alive = 0;
    for (int i = 0; i < np; i++) {
      int s = b->prog_slots[i];
      if (a0[i] && !b->hact[s] && b->hctl[s].ncyc - n0[i] == 1) {
        // stopped after the first cycle of the period: its buffers are
        // in the rotation recorded after cycle 1
        mge_case *h = b->cs[s];
        for (int l = 1; l <= h->nlevel; l++) h->lv[l].cur = b->cur1[i][l - 1];
      }
      alive += b->hact[s];
    }
    if (!alive) break;
  }
  if (nactive) {
    int a = 0;
    for (int c = 0; c < (int)b->cs.size(); c++) a += b->cs[c] && b->hact[c];
    *nactive = a;
  }
  return MGE_OK;
}

int mge_batch_status(mge_batch *b, int slot, int *status, int *ncyc, double *ini, double *last) {
  if (!b || slot < 0 || slot >= (int)b->cs.size()) return set_err(MGE_ERR_ARG, "bad batch slot");
  const BCtl &C = b->hctl[slot];
  if (status) *status = b->hact[slot] ? 0 : (C.status ? C.status : 3);
  if (ncyc) *ncyc = C.ncyc;
  if (ini) *ini = C.ini;
  if (last) *last = C.last;
  return MGE_OK;
}

int mge_batch_log(mge_batch *b, int slot, double *out, int cap) {
  if (!b || slot < 0 || slot >= (int)b->cs.size() || !out) return set_err(MGE_ERR_ARG, "bad batch slot");
  int n = std::min(std::min(b->hctl[slot].ncyc, b->cap), cap);
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(out, b->hctl[slot].log, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  return n;
}

long long mge_case_bytes(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  long long b = 0;
  for (auto &L : h->lv) b += (long long)L.bytes;
  return b;
}

int mge_synchronize(mge_case *h) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  return check_err(h);
}

int mge_autotune(mge_case *h, int lvl, int *kc, double *us) {
  Level *L = get_level(h, lvl);
  if (!L) return set_err(MGE_ERR_ARG, "bad level");
  return by_type(h, [&](auto z) { return do_autotune<decltype(z)>(h, *L, kc, us); });
}

int mge_set_kc(mge_case *h, int lvl, int kc) {
  Level *L = get_level(h, lvl);
  if (!L || kc < 0) return set_err(MGE_ERR_ARG, "bad level or kc");
  L->kc = kc;
  return MGE_OK;
}

int mge_set_freeze(mge_case *h, int far, int lim) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  const int v[2] = {far ? 1 : 0, lim ? 1 : 0};
  HIPCHK(h2d_sync(h->ffrozen, v, sizeof(v), h->stream));
  return MGE_OK;
}

int mge_stage_tile(mge_case *h, int lvl, int *tj, int *kc) {
  Level *L = get_level(h, lvl);
  if (!L || !tj || !kc) return set_err(MGE_ERR_ARG, "bad level or output pointer");
  const int distype = h->prm.distype;
  *tj = h->fp32 ? stage_tj<float>(*L, distype) : stage_tj<double>(*L, distype);
  *kc = stage_kc(*L);
  return MGE_OK;
}

int mge_set_tile(mge_case *h, int lvl, int tj, int kc) {
  Level *L = get_level(h, lvl);
  if (!L || !(tj == 0 || tj == 4 || tj == 8) || kc < 0) return set_err(MGE_ERR_ARG, "bad level, tile rows or kc");
  L->tj = tj;
  L->kc = kc;
  return MGE_OK;
}

int mge_kernel_timing(mge_case *h, int enable, const char *name, double *ms, long long *count,
                      double *bytes) {
  if (!h) return set_err(MGE_ERR_ARG, "null handle");
  if (h->timing) {
    hipStreamSynchronize(h->stream);
    resolve_timers(h);
  }
  h->timing = enable != 0;
  if (!name) {
    h->tstat.clear();
    return MGE_OK;
  }
  auto it = h->tstat.find(name);
  if (it == h->tstat.end()) {
    if (ms) *ms = 0.0;
    if (count) *count = 0;
    if (bytes) *bytes = 0.0;
    return MGE_OK;
  }
  const TStat &t = it->second;
  if (ms) *ms = t.count ? t.ms / t.count : 0.0;
  if (count) *count = t.count;
  if (bytes) *bytes = t.count ? t.bytes / t.count : 0.0;
  return MGE_OK;
}

}  // extern "C"

// embedded multi-block hierarchies (composite grids)
#include "mge_comp.h"
