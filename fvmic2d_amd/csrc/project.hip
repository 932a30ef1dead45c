// Deterministic marker -> node projection (markerProjection.c:71-250 and the
// scatter parts of markers.c:1507-1594, subgridStressChanges.c:128-233):
// single-instance launches and the public projection entry points.
//
// Pass 1 (k_partials): a group of G lanes per lattice cell walks the (at
//   most four) buckets that make up the cell, in bucket order and ascending
//   marker index within a bucket (lane j takes entries j, j+16, ...), and a
//   fixed shuffle tree sums the lanes' corner partials.  No atomics: each
//   partial is owned by one lane group.
// Pass 2 (k_nodes): one thread per node adds the corner partials of the (at
//   most four) lattice cells around it in a fixed order and normalises.
// Both passes are pure functions of the bucket order, so the result is
// bitwise reproducible; it differs from the reference's sequential sum only
// by the association of the additions.
#include <utility>
#include <vector>

#include "proj_common.h"

// Lanes per lattice cell (G) of the single-instance launches: 8 by default,
// 4, 16 or 32 through fvm_ctx_set_lanes.

template <int METHOD, int NF, int G>
__global__ void __launch_bounds__(256) k_partials(DGrid g, ProjGroup pg, const int32_t *__restrict__ start,
                                                  const double *__restrict__ xs, const double *__restrict__ ys,
                                                  const int32_t *__restrict__ ms, double *__restrict__ part, int nLx,
                                                  int nLy, int lox, int loy) {
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t t = (tile * blockDim.x + threadIdx.x) / G;
  const int j = threadIdx.x % G;
  const int64_t nL = (int64_t)nLx * nLy;
  const bool live = t < nL;
  const int lx = live ? (int)(t % nLx) + lox : 0;
  const int ly = live ? (int)(t / nLx) + loy : 0;
  __shared__ double smat[NF * FVM_MAXMAT];
  const double *mtab = stage_mat_tables<NF>(pg, smat);
  partials_cell_sorted<METHOD, NF, G>(g, pg, start, xs, ys, ms, mtab, part, nL, t, lx, ly, j, live);
}

// Several groups of one kind (method, field count) in one launch: the launch
// covers lattice-cell coordinates (cx, cy) in [-1, ncx] x [-1, NY-1] and each
// lane group computes, family after family, the lattice cell at (cx, cy) of
// every group whose lattice contains it.  Neighbouring families' cells at one
// coordinate cover nearly the same markers, so the later families read them
// from cache.
#define MAX_MULTI 4
struct MultiGroups {
  int ng;
  ProjGroup pg[MAX_MULTI];
  double *part[MAX_MULTI];
};

template <int METHOD, int NF, int G>
__global__ void __launch_bounds__(256) k_partials_multi(DGrid g, MultiGroups mg, const int32_t *__restrict__ start,
                                                        const double *__restrict__ xs, const double *__restrict__ ys,
                                                        const int32_t *__restrict__ ms) {
  // consecutive blocks (one XCD, after xcd_tile) take the same range of cells
  // for each group in turn, so every group after the first reads the markers
  // from L2
  const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
  const int q = (int)(b % mg.ng);
  const int64_t tile = b / mg.ng;
  const int64_t t = (tile * blockDim.x + threadIdx.x) / G;
  const int j = threadIdx.x % G;
  const int ntx = g.ncx() + 2;
  const int64_t nT = (int64_t)ntx * (g.NY + 1);
  const bool inside = t < nT;
  const int cx = inside ? (int)(t % ntx) - 1 : -2;
  const int cy = inside ? (int)(t / ntx) - 1 : -2;
  const ProjGroup &pg = mg.pg[q];
  const Lat Lx = lattice(pg.stagx, g.ncx() + 1), Ly = lattice(pg.stagy, g.NY);
  const bool live = cx >= Lx.lo && cx < Lx.lo + Lx.n && cy >= Ly.lo && cy < Ly.lo + Ly.n;
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  const int64_t tl = live ? (int64_t)(cy - Ly.lo) * Lx.n + (cx - Lx.lo) : 0;
  __shared__ double smat[NF * FVM_MAXMAT];
  const double *mtab = stage_mat_tables<NF>(pg, smat);
  partials_cell_sorted<METHOD, NF, G>(g, pg, start, xs, ys, ms, mtab, mg.part[q], nL, tl, live ? cx : 0,
                                      live ? cy : 0, j, live);
}

template <int G>
__global__ void __launch_bounds__(256) k_partials_subt(DGrid g, ProjGroup pg, SubT sb,
                                                       const int32_t *__restrict__ start,
                                                       const int32_t *__restrict__ perm, const double *__restrict__ xs,
                                                       const double *__restrict__ ys, double *__restrict__ part,
                                                       int nLx, int nLy) {
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t t = (tile * blockDim.x + threadIdx.x) / G;
  const int j = threadIdx.x % G;
  const int64_t nL = (int64_t)nLx * nLy;
  const bool live = t < nL;
  const int lx = live ? (int)(t % nLx) : 0;
  const int ly = live ? (int)(t / nLx) : 0;
  partials_cell_subt_sorted<G>(g, pg, start, perm, xs, ys, part, nL, t, lx, ly, j, live, sb);
}

__global__ void __launch_bounds__(256) k_gather_payload(int32_t n, const int32_t *__restrict__ total,
                                                        const int32_t *__restrict__ perm, const double *__restrict__ X,
                                                        const double *__restrict__ Y, const int32_t *__restrict__ Mat,
                                                        double *__restrict__ xs, double *__restrict__ ys,
                                                        int32_t *__restrict__ ms, PayloadCols pc, int with_xy) {
  const int64_t p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n || p >= *total) return;
  gather_entry(pc, with_xy, X, Y, Mat, xs, ys, ms, p, perm[p]);
}

// one thread per node; launch row blockIdx.y = group (all its fields)
__global__ void __launch_bounds__(256) k_nodes(int NX, int NY, int per, NodeGroups G) {
  const int64_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  if (per) node_group_periodic(G.q[blockIdx.y], ix, jy, i, NX);
  else node_group(G.q[blockIdx.y], ix, jy, i);
}

struct Payload {
  const double *xs, *ys;
  const int32_t *ms;
};

template <int METHOD, int NF>
static int launch_partials(fvm_ctx *ctx, const fvm_grid *g, const Payload &pl, const ProjGroup &pg,
                           double *part) {
  const Lat Lx = lattice(pg.stagx, fvm_ncx(g) + 1), Ly = lattice(pg.stagy, g->NY);
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  FVM_PHASE(ctx, FVM_PHASE_PARTIALS);
  const int G = ctx->lanes;
#define LAUNCH_P(GG)                                                                                     \
  k_partials<METHOD, NF, GG><<<fvm_blocks(nL * GG, 256), 256, 0, ctx->stream>>>(                          \
      dgrid(g), pg, (const int32_t *)ctx->start.ptr, pl.xs, pl.ys, pl.ms, part, Lx.n, Ly.n, Lx.lo, Ly.lo)
  if (G == 4) LAUNCH_P(4);
  else if (G == 8) LAUNCH_P(8);
  else if (G == 32) LAUNCH_P(32);
  else LAUNCH_P(16);
#undef LAUNCH_P
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

template <int METHOD, int NF>
static int launch_multi(fvm_ctx *ctx, const fvm_grid *g, const Payload &pl, const MultiGroups &mg) {
  const int64_t nT = (int64_t)(fvm_ncx(g) + 2) * (g->NY + 1);
  FVM_PHASE(ctx, FVM_PHASE_PARTIALS);
  const int G = ctx->lanes;
#define LAUNCH_M(GG)                                                                                     \
  k_partials_multi<METHOD, NF, GG><<<(unsigned)fvm_blocks(nT * GG, 256) * (unsigned)mg.ng, 256, 0,              \
                                     ctx->stream>>>(                                                       \
      dgrid(g), mg, (const int32_t *)ctx->start.ptr, pl.xs, pl.ys, pl.ms)
  if (G == 4) LAUNCH_M(4);
  else if (G == 8) LAUNCH_M(8);
  else if (G == 32) LAUNCH_M(32);
  else LAUNCH_M(16);
#undef LAUNCH_M
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

// multi-group launches exist for the common kinds; others go one group per launch
static int dispatch_multi(fvm_ctx *ctx, const fvm_grid *g, const Payload &pl, const MultiGroups &mg, bool *done) {
  *done = true;
  const int m = mg.pg[0].method, nf = mg.pg[0].nf;
  if (m == PM_ARITH) {
    switch (nf) {
      case 1: return launch_multi<PM_ARITH, 1>(ctx, g, pl, mg);
      case 2: return launch_multi<PM_ARITH, 2>(ctx, g, pl, mg);
      case 3: return launch_multi<PM_ARITH, 3>(ctx, g, pl, mg);
      case 4: return launch_multi<PM_ARITH, 4>(ctx, g, pl, mg);
      case 5: return launch_multi<PM_ARITH, 5>(ctx, g, pl, mg);
    }
  } else if (m == PM_LOCAL) {
    switch (nf) {
      case 1: return launch_multi<PM_LOCAL, 1>(ctx, g, pl, mg);
      case 2: return launch_multi<PM_LOCAL, 2>(ctx, g, pl, mg);
    }
  }
  *done = false;
  return FVM_OK;
}

template <int METHOD>
static int dispatch_nf(fvm_ctx *ctx, const fvm_grid *g, const Payload &mk, const ProjGroup &pg, double *part) {
  switch (pg.nf) {
    case 1: return launch_partials<METHOD, 1>(ctx, g, mk, pg, part);
    case 2: return launch_partials<METHOD, 2>(ctx, g, mk, pg, part);
    case 3: return launch_partials<METHOD, 3>(ctx, g, mk, pg, part);
    case 4: return launch_partials<METHOD, 4>(ctx, g, mk, pg, part);
    case 5: return launch_partials<METHOD, 5>(ctx, g, mk, pg, part);
    case 6: return launch_partials<METHOD, 6>(ctx, g, mk, pg, part);
    case 7: return launch_partials<METHOD, 7>(ctx, g, mk, pg, part);
    case 8: return launch_partials<METHOD, 8>(ctx, g, mk, pg, part);
    default: return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "projection group with %d fields", pg.nf);
  }
}

static int dispatch_method(fvm_ctx *ctx, const fvm_grid *g, const Payload &mk, const ProjGroup &pg,
                           double *part, const SubT *sb) {
  if (sb) {
    if (pg.stagx != 0 || pg.stagy != 0 || pg.method != PM_ARITH || pg.nf != 1)
      return fvm_set_err(ctx, FVM_ERR_ARG, "fused subgrid pass needs one ARITHMETIC field on basic nodes");
    const Lat Lx = lattice(0, fvm_ncx(g) + 1), Ly = lattice(0, g->NY);
    const int64_t nL = (int64_t)Lx.n * Ly.n;
    FVM_PHASE(ctx, FVM_PHASE_SUBGRID_T1);
    const int G = ctx->lanes;
#define LAUNCH_S(GG)                                                                                     \
  k_partials_subt<GG><<<fvm_blocks(nL * GG, 256), 256, 0, ctx->stream>>>(                                 \
      dgrid(g), pg, *sb, (const int32_t *)ctx->start.ptr, (const int32_t *)ctx->perm.ptr, mk.xs, mk.ys, part, \
      Lx.n, Ly.n)
    // 4 lanes per cell split the per-marker dTsub writes into more partial
    // lines: the pass takes 8 then
    if (G <= 8) LAUNCH_S(8);
    else if (G == 32) LAUNCH_S(32);
    else LAUNCH_S(16);
#undef LAUNCH_S
    FVM_LAUNCH_CHECK(ctx);
    return FVM_OK;
  }
  switch (pg.method) {
    case PM_ARITH: return dispatch_nf<PM_ARITH>(ctx, g, mk, pg, part);
    case PM_LOCAL: return dispatch_nf<PM_LOCAL>(ctx, g, mk, pg, part);
    case PM_HARM: return dispatch_nf<PM_HARM>(ctx, g, mk, pg, part);
    case PM_TIES: return dispatch_nf<PM_TIES>(ctx, g, mk, pg, part);
    case PM_CENTRE: return dispatch_nf<PM_CENTRE>(ctx, g, mk, pg, part);
    default: return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "projection method %d not implemented", pg.method);
  }
}

// Project several groups over one bucket order: one partials launch per group,
// then one node launch for (up to MAX_NODE_GROUPS of) them.  With `sub`, the
// single group is subgridTemperatureChanges' dTsub, computed in the partials pass.
int fvm_project_groups(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, int ng, const ProjGroup *pgs,
                       const SubT *sub) {
  if (sub && ng != 1) return fvm_set_err(ctx, FVM_ERR_ARG, "fused subgrid pass takes one group");
  size_t total = 0;
  size_t off[64];
  if (ng > 64) return fvm_set_err(ctx, FVM_ERR_ARG, "too many projection groups");
  for (int i = 0; i < ng; i++) {
    const ProjGroup &pg = pgs[i];
    if (pg.stagx < -1 || pg.stagx > 1 || pg.stagy < -1 || pg.stagy > 1)
      return fvm_set_err(ctx, FVM_ERR_BAD_STAG, "invalid stagger (%d,%d)", pg.stagx, pg.stagy);
    if (pg.nf < 1 || pg.nf > FVM_MAXF)
      return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "projection group with %d fields", pg.nf);
  }
  // ARITHMETIC groups one field short of the largest (<= 5) get a zero field,
  // so that all of them share one launch (blocks of one cell range take every
  // group in turn and read the markers once from HBM); the padded group's
  // partials carry the extra field, which no node pass reads
  std::vector<ProjGroup> sg(pgs, pgs + ng);
  if (!sub) {
    int nmax = 0;
    for (int i = 0; i < ng; i++)
      if (sg[i].method == PM_ARITH && sg[i].nf <= 5) nmax = sg[i].nf > nmax ? sg[i].nf : nmax;
    for (int i = 0; i < ng; i++)
      if (sg[i].method == PM_ARITH && sg[i].nf == nmax - 1) {
        ProjSrc &z = sg[i].src[sg[i].nf];
        memset(&z, 0, sizeof(z));
        z.use_mat = 1;  // material table of zeros: no load, adds 0
        sg[i].out[sg[i].nf] = nullptr;
        sg[i].nf++;
      }
  }
  for (int i = 0; i < ng; i++) {
    const Lat Lx = lattice(sg[i].stagx, fvm_ncx(g) + 1), Ly = lattice(sg[i].stagy, g->NY);
    off[i] = total;
    total += (size_t)Lx.n * Ly.n * 4 * (sg[i].nf + 1);
  }
  FVM_TRY(fvm_ws(ctx, ctx->partials, total * sizeof(double)));
  double *part = (double *)ctx->partials.ptr;
  // a rebuilt bucket order invalidates every cached payload column
  const bool resort = !fvm_sort_valid(ctx, g, mk);
  // payload: X, Y, Mat and every distinct source column, in bucket order
  PayloadCols pc;
  memset(&pc, 0, sizeof(pc));
  auto col_slot = [&](const double *c) -> int {
    for (int k = 0; k < pc.nc; k++)
      if (pc.src[k] == c) return k;
    if (pc.nc == MAX_PAYLOAD_COLS) return -1;
    pc.src[pc.nc] = c;
    return pc.nc++;
  };
  std::vector<std::pair<int, int>> uses;  // (group * FVM_MAXF + field, slot)
  for (int i = 0; i < ng; i++)
    for (int k = 0; k < sg[i].nf; k++)
      if (!sg[i].src[k].use_mat && !sub) {
        const int sl = col_slot(sg[i].src[k].col);
        if (sl < 0) return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "more than %d distinct marker columns", MAX_PAYLOAD_COLS);
        uses.push_back({i * FVM_MAXF + k, sl});
      }
  const int32_t n = mk->nMark;
  const size_t np = ((size_t)(n > 0 ? n : 1) + 31) / 32 * 32;
  SubT ssub;
  const double *cached_mt = nullptr;
  if (sub) {
    ssub = *sub;
    if (!resort && ctx->reuse_payload && ctx->payload.ptr && ctx->payload_serial == ctx->sort_serial &&
        ctx->payload_stride == np)
      for (int k = 0; k < ctx->payload_nc; k++)
        if (ctx->payload_src[k] == sub->MT) cached_mt = (const double *)ctx->payload.ptr + np * k;
    if (!cached_mt) col_slot(sub->MT);
  }
  // X, Y, Mat in bucket order are kept for the life of the sort
  const bool have_xy = !resort && ctx->xy_serial == ctx->sort_serial && ctx->payload_xy.bytes >= np * 20;
  if (!have_xy) FVM_TRY(fvm_ws(ctx, ctx->payload_xy, np * (2 * sizeof(double) + sizeof(int32_t))));
  double *xs = (double *)ctx->payload_xy.ptr, *ys = xs + np;
  int32_t *ms = (int32_t *)(ys + np);
  if (pc.nc > 0) FVM_TRY(fvm_ws(ctx, ctx->payload, np * sizeof(double) * pc.nc));
  for (int k = 0; k < pc.nc; k++) pc.dst[k] = (double *)ctx->payload.ptr + np * k;
  for (auto &u : uses) sg[u.first / FVM_MAXF].src[u.first % FVM_MAXF].col = pc.dst[u.second];
  if (sub) ssub.MT = cached_mt ? cached_mt : pc.dst[0];
  // a rebuilt order gathers the payload in its final pass
  SortGather sgat;
  memset(&sgat, 0, sizeof(sgat));
  sgat.with_xy = have_xy ? 0 : 1;
  sgat.X = mk->X;
  sgat.Y = mk->Y;
  sgat.Mat = mk->Mat;
  sgat.xs = xs;
  sgat.ys = ys;
  sgat.ms = ms;
  sgat.pc = pc;
  bool gathered = false;
  FVM_TRY(fvm_sort_markers(ctx, g, mk, &sgat, &gathered));
  if (pc.nc > 0) {  // record what the payload now holds
    ctx->payload_nc = pc.nc;
    for (int k = 0; k < pc.nc; k++) ctx->payload_src[k] = pc.src[k];
    ctx->payload_stride = np;
    ctx->payload_serial = ctx->sort_serial;
  }
  if (!gathered && n > 0 && (!have_xy || pc.nc > 0)) {
    FVM_PHASE(ctx, FVM_PHASE_SORT);
    const int64_t nk = 4LL * fvm_ncx(g) * (g->NY - 1);
    k_gather_payload<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, (const int32_t *)ctx->start.ptr + nk,
                                                                  (const int32_t *)ctx->perm.ptr, mk->X, mk->Y,
                                                                  mk->Mat, xs, ys, ms, pc, have_xy ? 0 : 1);
    FVM_LAUNCH_CHECK(ctx);
  }
  ctx->xy_serial = ctx->sort_serial;
  const Payload pl = {xs, ys, ms};
  // groups of one kind share a launch (up to MAX_MULTI per launch)
  std::vector<bool> launched(ng, false);
  if (!sub) {
    for (int i = 0; i < ng; i++) {
      if (launched[i]) continue;
      MultiGroups mg;
      memset(&mg, 0, sizeof(mg));
      std::vector<int> members;
      for (int k = i; k < ng && mg.ng < MAX_MULTI; k++) {
        if (launched[k] || sg[k].method != sg[i].method || sg[k].nf != sg[i].nf) continue;
        mg.pg[mg.ng] = sg[k];
        mg.part[mg.ng] = part + off[k];
        mg.ng++;
        members.push_back(k);
      }
      if (mg.ng < 2) continue;
      bool done = false;
      FVM_TRY(dispatch_multi(ctx, g, pl, mg, &done));
      if (done)
        for (int k : members) launched[k] = true;
    }
  }
  for (int i = 0; i < ng; i++)
    if (!launched[i]) FVM_TRY(dispatch_method(ctx, g, pl, sg[i], part + off[i], sub ? &ssub : nullptr));
  const int64_t nn = (int64_t)g->NX * g->NY;
  for (int i0 = 0; i0 < ng; i0 += MAX_NODE_GROUPS) {
    NodeGroups G;
    memset(&G, 0, sizeof(G));
    for (int i = i0; i < ng && G.ng < MAX_NODE_GROUPS; i++) {
      const ProjGroup &pg = pgs[i];
      const Lat Lx = lattice(pg.stagx, fvm_ncx(g) + 1), Ly = lattice(pg.stagy, g->NY);
      NodeGroup &q = G.q[G.ng++];
      q.part = part + off[i];
      q.nf = pg.nf;
      q.pst = sg[i].nf + 1;
      q.harm = pg.method == PM_HARM;
      q.nLx = Lx.n;
      q.nLy = Ly.n;
      q.lox = Lx.lo;
      q.loy = Ly.lo;
      q.epilogue = pg.epilogue;
      for (int k = 0; k < pg.nf; k++) {
        q.out[k] = pg.out[k];
        G.pg[G.npair] = (unsigned char)(G.ng - 1);
        G.pk[G.npair] = (unsigned char)k;
        G.npair++;
      }
      q.ea = pg.ea;
      q.eb = pg.eb;
    }
    FVM_PHASE(ctx, FVM_PHASE_NODES);
    k_nodes<<<dim3(fvm_blocks(nn, 256), G.ng), 256, 0, ctx->stream>>>(g->NX, g->NY, g->xperiodic ? 1 : 0, G);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

int fvm_project_group(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg) {
  return fvm_project_groups(ctx, g, mk, 1, &pg, nullptr);
}

// ---- public projection entry points ------------------------------------------
extern "C" int fvm_project_marker_fields(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, int32_t n,
                                         const fvm_projection *list) {
  if (!ctx || !mk || (n > 0 && !list)) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  // group entries by (stagx, stagy, method) keeping first-appearance order
  bool done[256] = {false};
  if (n > 256) return fvm_set_err(ctx, FVM_ERR_ARG, "at most 256 projections per call");
  for (int i = 0; i < n; i++) {
    const fvm_projection &e = list[i];
    if (e.method != ARITHMETIC && e.method != ARITHMETIC_LOCAL && e.method != HARMONIC)
      return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "projection method %d not implemented", e.method);
    if (!e.nodalField || (!e.markerField && !e.materialProperty))
      return fvm_set_err(ctx, FVM_ERR_ARG, "projection %d: missing field or output", i);
  }
  ProjGroup groups[64];
  int ng = 0;
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    const fvm_projection &a = list[i];
    ProjGroup &pg = groups[ng++];
    memset(&pg, 0, sizeof(pg));
    pg.stagx = a.stagx;
    pg.stagy = a.stagy;
    pg.method = a.method == ARITHMETIC ? PM_ARITH : (a.method == ARITHMETIC_LOCAL ? PM_LOCAL : PM_HARM);
    pg.epilogue = EPI_DIVIDE;
    for (int j = i; j < n; j++) {
      const fvm_projection &b = list[j];
      if (done[j] || b.stagx != a.stagx || b.stagy != a.stagy || b.method != a.method) continue;
      if (pg.nf == FVM_MAXF) break;
      ProjSrc &s = pg.src[pg.nf];
      if (b.materialProperty) {
        s.use_mat = 1;
        for (int k = 0; k < FVM_MAXMAT; k++) s.mat[k] = b.materialProperty[k];
      } else {
        s.col = b.markerField;
      }
      pg.out[pg.nf] = b.nodalField;
      pg.nf++;
      done[j] = true;
    }
    if (ng == 64) break;
  }
  return fvm_project_groups(ctx, g, mk, ng, groups);
}

extern "C" int fvm_project_markers_nodes_from_scalar(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                                     const double *markerField, double *nodalField) {
  if (!ctx || !mk || !markerField || !nodalField) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  ProjGroup pg;
  memset(&pg, 0, sizeof(pg));
  pg.stagx = pg.stagy = 0;
  pg.method = PM_ARITH;
  pg.nf = 1;
  pg.src[0].col = markerField;
  pg.out[0] = nodalField;
  pg.epilogue = EPI_DIVIDE;
  return fvm_project_group(ctx, g, mk, pg);
}

__global__ void k_fill(double *__restrict__ a, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

int fvm_fill(fvm_ctx *ctx, double *a, int64_t n, double v) {
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_fill<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(a, n, v);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

extern "C" int fvm_project_markers_nodes_all2(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                              const NodalArrays *nf, const Materials *mat) {
  if (!ctx || !mk || !nf || !mat) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int64_t nn = (int64_t)g->NX * g->NY;
  // markerProjection.c:19-58, grouped by node family and method
  fvm_projection list[9] = {
      {mk->rho, nullptr, nf->rho, 0, 0, ARITHMETIC},
      {mk->eta, nullptr, nf->etaN, -1, -1, ARITHMETIC},
      {mk->eta, nullptr, nf->etaS, 0, 0, ARITHMETIC},
      {nullptr, mat->materialkThermal, nf->kThermal, 0, 0, ARITHMETIC},
      {mk->T, nullptr, nf->lastT, 0, 0, ARITHMETIC},
      {nullptr, mat->materialCp, nf->Cp, 0, 0, ARITHMETIC},
      {mk->sxx, nullptr, nf->soxx, -1, -1, ARITHMETIC_LOCAL},
      {mk->syy, nullptr, nf->soyy, -1, -1, ARITHMETIC_LOCAL},
      {mk->sxy, nullptr, nf->soxy, 0, 0, ARITHMETIC_LOCAL},
  };
  FVM_TRY(fvm_project_marker_fields(ctx, g, mk, 9, list));
  FVM_HIP(ctx, hipMemsetAsync(nf->rhodot, 0, (size_t)nn * sizeof(double), ctx->stream));  // :21
  FVM_TRY(fvm_fill(ctx, nf->muS, nn, 1e99));                                               // :36
  FVM_TRY(fvm_fill(ctx, nf->muN, nn, 1e99));                                               // :37
  return FVM_OK;
}
