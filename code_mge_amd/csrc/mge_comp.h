// mge_comp.h -- embedded multi-block hierarchies (composite grids), part of
// the translation unit of mge_device.hip (it uses its Level, mge_case and
// do_* routines).
//
// A composite is a set of blocks, each a one-level mge_case, arranged in
// multigrid levels: every block of level L > 1 lies inside one block of
// level L-1 (its parent) on the lattice refined by 2 in every direction, and
// the blocks of one level may overlap.  The solve follows solver.f90:119-261
// with its block loops: each block's timestep/euler is followed by
// bc_overlap of the level; movco restricts every child into the part of its
// parent it covers (collc.f90:58-146) and clears the parent's marker
// elsewhere, so the uncovered coarse cells carry no forcing; movfin
// prolongs the parent's correction into every child (addw.f90:84-188), then
// bc_embed fills the children's ghost cells from their parents
// (bc_embed.f90:50-143) and bc_overlap from their same-level neighbours
// (bc_overlap.f90:32-90); movco.f90:21-49 and movfin.f90:20-40 give the
// order, with bcond of every block last.  The reference locates cells with
// eval_x2cell/shft_cell/trilinear_factor (absent from its tree); on the
// aligned lattices here the interpolation weights are the integer-index
// stencils of the single-block transfers (1/2 for restriction, 1/4 and 3/4
// for prolongation and embedding, 1 for coinciding centres), and a stencil
// next to the parent's boundary reads the parent's ghost cell as the
// single-block prolongation does.
//
// All member cases run on the composite's stream while they belong to it.

struct mge_comp {
  struct Block {
    mge_case *h;
    int level;           // 1 .. nlevel
    int parent;          // index of the block of level - 1 holding it (-1 on level 1)
    int g0[3];           // index of the block's cell 2 on its level's lattice
    hipStream_t own;     // the case's own stream, restored when the composite ends
    void *corr = nullptr;  // correction w - ws of the block as a parent (5 components)
  };
  std::vector<Block> b;
  std::vector<std::vector<int>> lvl;   // lvl[L]: the blocks of level L
  int nlevel = 0;
  hipStream_t stream = nullptr;
  int fp32 = 0;
};

namespace {

Level &blev(mge_comp *C, int i) { return C->b[i].h->lv[1]; }

// bc_overlap of level L: for each block, every same-level neighbour whose
// cells its ghost layers reach
template <class T>
int comp_overlap(mge_comp *C, int L) {
  const std::vector<int> &bl = C->lvl[L];
  if (bl.size() < 2) return MGE_OK;
  for (int a : bl) {
    mge_case *h = C->b[a].h;
    Level &A = blev(C, a);
    for (int n : bl) {
      if (n == a) continue;
      Level &N = blev(C, n);
      EmbedArgs<T> X{};
      X.df = A.d;
      X.dc = N.d;
      X.sf = stm<T>(A, A.cur);
      X.sc = stc<T>(N, N.cur);
      X.nb[0] = N.d.ni; X.nb[1] = N.d.nj; X.nb[2] = N.d.nk;
      bool reach = true;
      const int na[3] = {A.d.ni, A.d.nj, A.d.nk};
      for (int d = 0; d < 3; d++) {
        X.off[d] = C->b[a].g0[d] - C->b[n].g0[d];
        // A's cells -2 .. na+1 (ghosts included) against N's 0 .. nb-1
        const int lo = X.off[d] - 2, hi = X.off[d] + na[d] + 1;
        if (hi < 0 || lo > X.nb[d] - 1) reach = false;
      }
      if (!reach) continue;
      X.gamma = (T)h->prm.gamma;
      X.inf = make_inf<T>(h->prm);
      LAUNCHT(h, "k_overlap", k_overlap, T, flat_grid(A.d.E), dim3(256), X);
    }
  }
  return MGE_OK;
}

// bc_embed of level L: every block's ghost cells from its parent
template <class T>
int comp_embed(mge_comp *C, int L) {
  for (int f : C->lvl[L]) {
    const int p = C->b[f].parent;
    if (p < 0) continue;
    mge_case *h = C->b[f].h;
    Level &F = blev(C, f), &P = blev(C, p);
    EmbedArgs<T> X{};
    X.df = F.d;
    X.dc = P.d;
    X.sf = stm<T>(F, F.cur);
    X.sc = stc<T>(P, P.cur);
    X.flagsf = F.flags;
    for (int d = 0; d < 3; d++) X.off[d] = C->b[f].g0[d] / 2 - C->b[p].g0[d];
    X.gamma = (T)h->prm.gamma;
    X.inf = make_inf<T>(h->prm);
    LAUNCHT(h, "k_embed", k_embed, T, flat_grid(F.d.E), dim3(256), X);
  }
  return MGE_OK;
}

template <class T>
int comp_bcond(mge_comp *C, int L) {
  for (int i : C->lvl[L]) {
    Level &B = blev(C, i);
    int rc = do_bcond<T>(C->b[i].h, B, B.cur);
    if (rc < 0) return rc;
  }
  return MGE_OK;
}

// movco.f90:21-49 at coarse level mgl
template <class T>
int comp_movco(mge_comp *C, int mgl) {
  for (int c : C->lvl[mgl]) {
    mge_case *h = C->b[c].h;
    Level &Cb = blev(C, c);
    HIPCHK(hipMemsetAsync(Cb.marker, 0, (size_t)Cb.d.E, h->stream));   // collc.f90:52
    for (int f : C->lvl[mgl + 1]) {
      if (C->b[f].parent != c) continue;
      Level &F = blev(C, f);
      XferArgs<T> X{};
      X.dc = Cb.d;
      X.df = F.d;
      X.rx = X.ry = X.rz = 2;
      X.wf = tp<T>(F.S[F.cur]);
      X.dwf = tp<T>(F.dw);
      X.wc = tp<T>(Cb.S[Cb.cur]);
      X.wsc = tp<T>(Cb.ws);
      X.wr = tp<T>(Cb.wr);
      X.marker = Cb.marker;
      X.volc = (T)Cb.cellvolume;
      X.volf = (T)F.cellvolume;
      X.flagsf = F.flags;
      for (int d = 0; d < 3; d++) X.off[d] = C->b[f].g0[d] / 2 - C->b[c].g0[d];
      X.nrc[0] = F.d.ni / 2; X.nrc[1] = F.d.nj / 2; X.nrc[2] = F.d.nk / 2;
      Dims cov = Cb.d;
      cov.ni = X.nrc[0]; cov.nj = X.nrc[1]; cov.nk = X.nrc[2];
      LAUNCHT(h, "k_restrict", k_restrict, T, cell_grid(cov), cell_block(), X);
    }
    LAUNCHT(h, "k_eos_all", k_eos_all, T, flat_grid(Cb.d.E), dim3(256),
            (EosArgs<T>{Cb.d, stm<T>(Cb, Cb.cur), (T)h->prm.gamma, make_inf<T>(h->prm)}));
  }
  CHKRC(comp_overlap<T>(C, mgl));
  return comp_bcond<T>(C, mgl);
}

// movfin.f90:20-40 at fine level mgl
template <class T>
int comp_movfin(mge_comp *C, int mgl, int mlvl) {
  const bool fmg = (mgl - 1 == mlvl);
  if (!fmg)
    for (int p : C->lvl[mgl - 1]) {
      mge_case *h = C->b[p].h;
      Level &P = blev(C, p);
      XferArgs<T> X{};
      X.dc = P.d;
      X.wc = tp<T>(P.S[P.cur]);
      X.wsc = tp<T>(P.ws);
      X.corr = tp<T>(C->b[p].corr);
      LAUNCHT(h, "k_corr", k_corr, T, flat_grid(5 * P.d.E), dim3(256), X);
    }
  for (int f : C->lvl[mgl]) {
    const int p = C->b[f].parent;
    mge_case *h = C->b[f].h;
    Level &F = blev(C, f), &P = blev(C, p);
    XferArgs<T> X{};
    X.dc = P.d;
    X.df = F.d;
    X.rx = X.ry = X.rz = 2;
    X.wc = tp<T>(P.S[P.cur]);
    X.wsc = tp<T>(P.ws);
    X.volc = (T)P.cellvolume;
    X.volf = (T)F.cellvolume;
    X.fmg = fmg;
    X.flagsf = F.flags;
    X.wfo = tp<T>(F.S[F.cur]);
    X.corr = tp<T>(C->b[p].corr);
    for (int d = 0; d < 3; d++) X.off[d] = C->b[f].g0[d] / 2 - C->b[p].g0[d];
    LAUNCHT(h, "k_prolong", k_prolong, T, cell_grid(F.d), cell_block(), X);
    LAUNCHT(h, "k_eos_all", k_eos_all, T, flat_grid(F.d.E), dim3(256),
            (EosArgs<T>{F.d, stm<T>(F, F.cur), (T)h->prm.gamma, make_inf<T>(h->prm)}));
  }
  CHKRC(comp_embed<T>(C, mgl));
  CHKRC(comp_overlap<T>(C, mgl));
  return comp_bcond<T>(C, mgl);
}

// timestep + euler of every block of level mgl, each followed by bc_overlap
// of the level (solver.f90:151-157, 196-201, 214-221)
template <class T>
int comp_relax(mge_comp *C, int mgl, int ncyc, int mlvl, int lres, int lfir, long long *cs) {
  for (int i : C->lvl[mgl]) {
    mge_case *h = C->b[i].h;
    Level &B = blev(C, i);
    if (!lres) CHKRC(do_timestep<T>(h, B));
    CHKRC(do_euler<T>(h, B, ncyc, mgl, mlvl, lres, lfir));
    *cs += (long long)(lres ? 1 : h->prm.mstage) * B.d.ni * B.d.nj * B.d.nk;
    CHKRC(comp_overlap<T>(C, mgl));
  }
  return MGE_OK;
}

// solver.f90:119-261 over the composite (steady flow)
template <class T>
int comp_solve(mge_comp *C, const int *mcyc, double *reslog, int cap, int *nlog, long long *cs_out) {
  const int nlevel = C->nlevel;
  const mge_params &P = C->b[0].h->prm;
  int nl = 0;
  long long cs = 0;
  double ini_res = 0.0, max_res = 0.0;
  for (int mlvl = 1; mlvl <= nlevel; mlvl++) {
    int mgl = mlvl;
    for (int ncyc = 1; ncyc <= mcyc[mlvl - 1]; ncyc++) {
      std::vector<int> icyc(nlevel + 1, 0);
      bool next_cycle = false;
      for (;;) {       // lopmg
        for (;;) {     // lopr
          icyc[mgl] += 1;
          const int lfir = (mgl == mlvl) ? 0 : 1;
          CHKRC(comp_relax<T>(C, mgl, ncyc, mlvl, 0, lfir, &cs));
          if (mgl == mlvl) {
            max_res = 1000.0;
            for (int i : C->lvl[mgl]) {
              double r = 0.0;
              CHKRC(read_rtrms(C->b[i].h, 1, ncyc, &r));
              if (ncyc == 1) ini_res = r;
              if (r < max_res) max_res = r;
              if (r != r) return set_err(MGE_ERR_RES_NAN, "solver: residual is NaN");
            }
            if (reslog && nl < cap) {
              reslog[3 * nl] = mgl;
              reslog[3 * nl + 1] = ncyc;
              reslog[3 * nl + 2] = max_res;
            }
            nl++;
          }
          if (mlvl == 1) {
            next_cycle = true;
            break;
          }
          CHKRC(comp_relax<T>(C, mgl, ncyc, mlvl, 1, 0, &cs));
          mgl -= 1;
          CHKRC(comp_movco<T>(C, mgl));
          if (mgl == 1) break;
        }
        if (next_cycle) break;
        int lfir = 1;
        bool exit_mg = false;
        for (;;) {     // lopp
          CHKRC(comp_relax<T>(C, mgl, ncyc, mlvl, 0, lfir, &cs));
          lfir = 0;
          mgl += 1;
          CHKRC(comp_movfin<T>(C, mgl, mlvl));
          if (mgl == mlvl) {
            exit_mg = true;
            break;
          }
          if (icyc[mgl] != P.mgtype) break;
          icyc[mgl] = 0;
        }
        if (exit_mg) break;
      }
      if (next_cycle) continue;
      if (mgl == mlvl && ini_res / max_res > P.drop) break;
    }
    if (mlvl < nlevel) CHKRC(comp_movfin<T>(C, mlvl + 1, mlvl));
  }
  if (nlog) *nlog = nl;
  if (cs_out) *cs_out = cs;
  for (int i = 0; i < (int)C->b.size(); i++) CHKRC(check_err(C->b[i].h));
  return MGE_OK;
}

}  // namespace

extern "C" {

int mge_comp_create(int nblock, mge_case *const *blocks, const int *level, const int *parent, const int *g0,
                    mge_comp **out) {
  if (!out || nblock < 1 || !blocks || !level || !parent || !g0) return set_err(MGE_ERR_ARG, "bad composite arguments");
  *out = nullptr;
  mge_comp *C = new mge_comp();
  auto fail = [&](const char *m) {
    delete C;
    return set_err(MGE_ERR_ARG, m);
  };
  for (int i = 0; i < nblock; i++) {
    mge_case *h = blocks[i];
    if (!h || h->nlevel != 1 || !h->lv[1].defined) return fail("a composite block must be a one-level case, defined");
    if (h->rec) return fail("a composite block is recording a batch");
    if (h->fp32 != blocks[0]->fp32 || std::memcmp(&h->prm, &blocks[0]->prm, sizeof(mge_params)) != 0)
      return fail("the blocks of a composite must share their parameters");
    if (h->prm.ffreeze > 0.0 || h->prm.lfreeze > 0.0 || h->prm.execmode != 0)
      return fail("composite solves are steady and without freezes");
    mge_comp::Block B;
    B.h = h;
    B.level = level[i];
    B.parent = parent[i];
    for (int d = 0; d < 3; d++) B.g0[d] = g0[3 * i + d];
    B.own = h->stream;
    C->b.push_back(B);
    C->nlevel = std::max(C->nlevel, level[i]);
  }
  C->lvl.assign(C->nlevel + 1, {});
  for (int i = 0; i < nblock; i++) {
    const int L = C->b[i].level;
    if (L < 1) return fail("block levels start at 1");
    C->lvl[L].push_back(i);
  }
  for (int L = 1; L <= C->nlevel; L++)
    if (C->lvl[L].empty()) return fail("every level of a composite needs a block");
  for (int i = 0; i < nblock; i++) {
    const mge_comp::Block &B = C->b[i];
    const Dims &d = B.h->lv[1].d;
    const int n[3] = {d.ni, d.nj, d.nk};
    if (B.level == 1) {
      if (B.parent != -1) return fail("a block of level 1 has no parent");
      continue;
    }
    if (B.parent < 0 || B.parent >= nblock || C->b[B.parent].level != B.level - 1)
      return fail("a block's parent must be a block of the next coarser level");
    const mge_comp::Block &Pb = C->b[B.parent];
    const Dims &pd = Pb.h->lv[1].d;
    const int np[3] = {pd.ni, pd.nj, pd.nk};
    for (int a = 0; a < 3; a++) {
      if (B.g0[a] % 2 != 0 || n[a] % 2 != 0) return fail("an embedded block starts and ends on its parent's cell faces");
      if (B.g0[a] < 2 * Pb.g0[a] || B.g0[a] + n[a] > 2 * (Pb.g0[a] + np[a]))
        return fail("an embedded block lies inside its parent");
    }
    for (int f = 0; f < 6; f++) {
      // a face inside the parent's cells is embedded (bctype 0); a face on
      // the parent's boundary keeps the parent's condition there
      const int a = f % 3;
      const bool inner = f < 3 ? B.g0[a] > 2 * Pb.g0[a] : B.g0[a] + n[a] < 2 * (Pb.g0[a] + np[a]);
      if (inner && B.h->lv[1].bctype[f] != 0) return fail("an embedded face must have bctype 0");
    }
  }
  for (int i = 0; i < nblock; i++) {
    bool is_parent = false;
    for (int j = 0; j < nblock; j++) is_parent = is_parent || C->b[j].parent == i;
    if (!is_parent) continue;
    const Level &L = C->b[i].h->lv[1];
    void *p = nullptr;
    if (hipMalloc(&p, C->b[i].h->es * 5 * (size_t)L.d.E) != hipSuccess) {
      for (auto &B : C->b) if (B.corr) hipFree(B.corr);
      return fail("composite correction array allocation failed");
    }
    C->b[i].corr = p;
  }
  C->fp32 = blocks[0]->fp32;
  if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess) {
    for (auto &B : C->b) if (B.corr) hipFree(B.corr);
    delete C;
    return set_err(MGE_ERR_HIP, "composite stream creation failed");
  }
  for (auto &B : C->b) {
    // the composite's transfers read the blocks' ghost cells at any time:
    // every bcond forms the corner cells at once
    sync_corners(B.h, B.h->lv[1]);
    B.h->lazy_corners = false;
    hipStreamSynchronize(B.own);
    B.h->stream = C->stream;
  }
  *out = C;
  return MGE_OK;
}

int mge_comp_solve(mge_comp *C, const int *mcyc, double *reslog, int cap, int *nlog, long long *cell_stages) {
  if (!C || !mcyc) return set_err(MGE_ERR_ARG, "bad composite arguments");
  for (auto &B : C->b)
    for (int L = 1; L <= C->nlevel; L++)
      if (mcyc[L - 1] > B.h->lv[1].mxcyc) return set_err(MGE_ERR_ARG, "mcyc beyond a block's residual slots");
  const int rc = C->fp32 ? comp_solve<float>(C, mcyc, reslog, cap, nlog, cell_stages)
                         : comp_solve<double>(C, mcyc, reslog, cap, nlog, cell_stages);
  HIPCHK(hipStreamSynchronize(C->stream));
  return rc;
}

int mge_comp_destroy(mge_comp *C) {
  if (!C) return MGE_OK;
  hipStreamSynchronize(C->stream);
  for (auto &B : C->b) {
    B.h->stream = B.own;
    B.h->lazy_corners = true;
    if (B.corr) hipFree(B.corr);
  }
  hipStreamDestroy(C->stream);
  delete C;
  return MGE_OK;
}

}  // extern "C"
