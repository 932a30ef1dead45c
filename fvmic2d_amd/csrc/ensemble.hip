// Batched ensemble of independent marker problems (uneven grids and marker
// counts): every phase of the thermal marker step runs as ONE launch over all
// instances, and the whole step can be captured as a hipGraph.
//
// There is no reference counterpart: the reference runs parameter sweeps as
// separate processes (the Monte-Carlo loop of main_subduction.c:167-172).
// Each instance follows the single-instance step exactly (same device bodies,
// same bucket order), so with the same lane-group width an instance stepped
// here is bitwise identical to the same instance stepped alone.
//
// Work decomposition: for each kind of work (markers, lattice cells of one
// projection group, nodes) the instances' ranges are padded to whole blocks,
// so every block belongs to exactly one instance; a block finds its instance
// by a uniform binary search over the block offsets.
#include <vector>

#include "marker_common.h"
#include "proj_common.h"

int fvm_fill(fvm_ctx *ctx, double *a, int64_t n, double v);

struct DensTab {
  double rho[FVM_MAXMAT], alpha[FVM_MAXMAT];
};

struct InstDev {
  DGrid g;
  int32_t nMark;
  int64_t mark_off, bucket_off, node_off;
  double *X, *Y, *VX, *VY, *T, *Tdot, *sxx, *syy, *sxy, *exx, *eyy, *exy, *Eii, *wxy, *rho, *rhodot;
  int32_t *Mat, *cellX, *cellY, *cpu;
  const double *vx, *vy, *wxyn, *thisT, *lastT, *rhon, *Cpn, *kn;
  DensTab dens;
  double rho0;  // materialRho[0] (subgridTemperatureChanges.c:62)
  BoundaryValues bv;
  double dt, dsub;
  int32_t jaumann;
};

struct fvm_ensemble {
  int n = 0, ng = 0;
  fvm_tuning tune = {256, 8};
  std::vector<fvm_instance> inst;
  std::vector<InstDev> hinst;
  std::vector<ProjGroup> hgroups;   // [group][instance]; group ng is the dTsub projection
  std::vector<int> gstag_x, gstag_y, gmethod, gnf;
  std::vector<int> gsnf;  // fields stored in a group's partials (gnf, or gnf + 1 when padded)
  int64_t nmark = 0, nbucket = 0, nnode = 0, ncell = 0;
  // device memory
  InstDev *dinst = nullptr;
  ProjGroup *dgroups = nullptr;
  NodeGroup *dnodes = nullptr;        // [group][instance]
  int64_t *dblk = nullptr;            // block offsets: markers, nodes, then one row per group
  std::vector<int64_t> hblk;
  // instance of every block of every row (row r starts at bi_off[r]), and of
  // every 8 x 8 cell tile of the ordering pass
  int32_t *dbi = nullptr, *dtile_inst = nullptr;
  int64_t *dbi_off = nullptr;
  std::vector<int64_t> hbi_off;
  size_t bi_cap = 0;
  int64_t *dpartoff = nullptr;        // [group][instance] offsets into part
  int32_t *key = nullptr, *slot = nullptr, *perm = nullptr, *count = nullptr, *start = nullptr,
          *sums = nullptr, *counters = nullptr;
  double *part = nullptr, *dTsub = nullptr, *deltaT = nullptr;
  // bucket-ordered payload at global positions: X, Y, Mat and the distinct
  // marker columns the projections and the dTsub pass read ("roles"), filled
  // by the ordering pass; the partials passes read it as contiguous runs
  int ncol = 0, mt_role = -1;
  double *xs = nullptr, *ys = nullptr, *pcols = nullptr;
  int32_t *ms = nullptr;
  const double **dcolsrc = nullptr;     // [role][instance] source column
  int64_t *dcell_off = nullptr;         // [instance + 1] first global basic cell, then [instance + 1] first tile
  int64_t ntile = 0;                    // 8 x 8 cell tiles over all instances
  ProjGroup *dgroups_sorted = nullptr;  // [group][instance], columns -> payload roles
  size_t lds_locate = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// ---- block -> instance ----------------------------------------------------------------
// Every kind of work pads each instance's range to whole blocks; a table maps
// a block (row table `bi`) or an ordering tile to its instance, so a block
// finds its instance with one load.

// ---- kernels ------------------------------------------------------------------------------
__global__ void __launch_bounds__(512) kb_locate(const InstDev *__restrict__ insts, int n,
                                                 const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                 int32_t *__restrict__ key,
                                                 int32_t *__restrict__ slot, int32_t *__restrict__ count,
                                                 int32_t *__restrict__ counters) {
  extern __shared__ double sgrid[];
  const int ii = bi[blockIdx.x];
  const InstDev &I = insts[ii];
  const DGrid &g = I.g;
  for (int i = threadIdx.x; i < g.NX; i += blockDim.x) sgrid[i] = g.x[i];
  for (int i = threadIdx.x; i < g.NY; i += blockDim.x) sgrid[g.NX + i] = g.y[i];
  for (int i = threadIdx.x; i <= g.NX; i += blockDim.x) sgrid[g.NX + g.NY + i] = g.xc[i];
  for (int i = threadIdx.x; i <= g.NY; i += blockDim.x) sgrid[2 * g.NX + g.NY + 1 + i] = g.yc[i];
  __syncthreads();
  const double *gx = sgrid, *gy = sgrid + g.NX, *gxc = sgrid + g.NX + g.NY, *gyc = sgrid + 2 * g.NX + g.NY + 1;
  const int64_t m = (blockIdx.x - blk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  int out = 0, bad = 0, k = -1;
  if (m < I.nMark) {
    const double xm = I.X[m], ym = I.Y[m];
    int cx = -1, cy = -1;
    if (xm >= 0.0 && ym >= 0.0 && xm <= g.LX && ym <= g.LY) {
      cx = guess_grid(gx, g.NX, xm, (double)(g.NX - 1) / (gx[g.NX - 1] - gx[0]));
      cy = guess_grid(gy, g.NY, ym, (double)(g.NY - 1) / (gy[g.NY - 1] - gy[0]));
      if (cx < 0 || cx > g.NX - 2 || cy < 0 || cy > g.NY - 2) {
        bad = 1;
      } else {
        const int bx = xm >= gxc[cx + 1] ? 1 : 0;
        const int by = ym >= gyc[cy + 1] ? 1 : 0;
        k = (int)I.bucket_off + (((cy * (g.NX - 1) + cx) << 2) | bx | (by << 1));
      }
    } else {
      out = 1;
    }
    I.cellX[m] = cx;
    I.cellY[m] = cy;
    I.cpu[m] = 0;
    if (cx != -1) {  // main_subduction.c:255-264
      const int mm = I.Mat[m] & (FVM_MAXMAT - 1);
      I.rho[m] = -I.dens.rho[mm] * (I.dens.alpha[mm] * I.T[m]);
      I.rhodot[m] = 0.0;
    }
    key[I.mark_off + m] = k;
  }
  const int sl = wave_count_slot(k, count);
  if (k >= 0) slot[I.mark_off + m] = sl;
  const int lane = threadIdx.x & 63;
  const unsigned long long bo = __ballot(out), bb = __ballot(bad);
  if (lane == 0) {
    if (bo) atomicAdd(&counters[0], __popcll(bo));
    if (bb) atomicAdd(&counters[1], __popcll(bb));
  }
}

__global__ void __launch_bounds__(512) kb_scatter(const InstDev *__restrict__ insts, int n,
                                                  const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                  const int32_t *__restrict__ key, const int32_t *__restrict__ slot,
                                                  const int32_t *__restrict__ start, int32_t *__restrict__ perm) {
  const int ii = bi[blockIdx.x];
  const InstDev &I = insts[ii];
  const int64_t m = (blockIdx.x - blk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  if (m >= I.nMark) return;
  const int k = key[I.mark_off + m];
  if (k < 0) return;
  perm[start[k] + slot[I.mark_off + m]] = (int32_t)m;
}

// Bucket order of every instance's cells (one wavefront per cell) with the
// payload gathered from the owning instance's columns.  Cells are visited in
// 8 x 8 tiles per instance (y fastest inside a tile; tile_off = first tile of
// each instance), tiles spread over the XCDs in contiguous runs, as in the
// single-instance ordering pass.
__global__ void __launch_bounds__(256) kb_order_gather(const InstDev *__restrict__ insts, int n,
                                                       const int64_t *__restrict__ cell_off,
                                                       const int64_t *__restrict__ tile_off,
                                                       const int32_t *__restrict__ tile_inst,
                                                       const int32_t *__restrict__ start, int32_t *__restrict__ perm,
                                                       double *__restrict__ xs, double *__restrict__ ys,
                                                       int32_t *__restrict__ ms, int ncol,
                                                       const double *const *__restrict__ colsrc,
                                                       double *__restrict__ pcols, int64_t nmark) {
  // the wavefront index is uniform: say so, so that the instance record and
  // its column pointers are read with scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t w = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wave;
  const int64_t tile = w >> 6;
  if (tile >= tile_off[n]) return;  // whole wavefront exits together
  const int ii = tile_inst[tile];
  const InstDev &I = insts[ii];
  const int ncx = I.g.NX - 1, ncy = I.g.NY - 1;
  const uint32_t ntx = (uint32_t)(ncx + 7) >> 3;
  const uint32_t tl = (uint32_t)(tile - tile_off[ii]);
  const int l = (int)(w & 63);
  const int cx = (int)(tl % ntx) * 8 + (l >> 3), cy = (int)(tl / ntx) * 8 + (l & 7);
  if (cx >= ncx || cy >= ncy) return;
  const int64_t c = cell_off[ii] + (int64_t)cy * ncx + cx;
  const double *__restrict__ X = I.X, *__restrict__ Y = I.Y;
  const int32_t *__restrict__ Mat = I.Mat;
  bucket_order_cell(c, start, perm, [&](int p, int v) {
    xs[p] = X[v];
    ys[p] = Y[v];
    ms[p] = Mat[v];
    for (int k = 0; k < ncol; k++) pcols[(int64_t)k * nmark + p] = colsrc[(int64_t)k * n + ii][v];
  });
}

// Groups of one (method, field count) share a launch: launch row blockIdx.y
// is group gi[blockIdx.y]; a row's blocks beyond that group's count exit.
#define KB_MAX_MULTI 8
struct KbGroups {
  int gi[KB_MAX_MULTI], sx[KB_MAX_MULTI], sy[KB_MAX_MULTI];
};

// Every group of one (method, stored field count) in one launch over the
// universal lattice ((NX+1) x (NY+1) cell coordinates per instance, block row
// `urow`): consecutive blocks take the same cell range for each group in turn
// (group fastest, one XCD after xcd_tile), so the groups after the first read
// the payload from L2.
template <int METHOD, int NF, int G>
__global__ void __launch_bounds__(256) kb_partials(const InstDev *__restrict__ insts, int n,
                                                   const int64_t *__restrict__ ublk, const int32_t *__restrict__ ubi,
                                                   const ProjGroup *__restrict__ all_groups,
                                                   const int64_t *__restrict__ all_partoff,
                                                   const int32_t *__restrict__ start, const double *__restrict__ xs,
                                                   const double *__restrict__ ys, const int32_t *__restrict__ ms,
                                                   double *__restrict__ part, KbGroups kg, int nrow) {
  const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
  const int q = (int)(b % nrow);
  const int64_t ub = b / nrow;
  if (ub >= ublk[n]) return;
  const int gi = kg.gi[q], sx = kg.sx[q], sy = kg.sy[q];
  const int ii = ubi[ub];
  const InstDev &I = insts[ii];
  const int64_t lt = (ub - ublk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  const int64_t t = lt / G;
  const int j = (int)(lt % G);
  const int64_t nT = (int64_t)(I.g.NX + 1) * (I.g.NY + 1);
  const bool inside = t < nT;
  const int cx = inside ? (int)(t % (I.g.NX + 1)) - 1 : -2;
  const int cy = inside ? (int)(t / (I.g.NX + 1)) - 1 : -2;
  const Lat Lx = lattice(sx, I.g.NX), Ly = lattice(sy, I.g.NY);
  const bool live = cx >= Lx.lo && cx < Lx.lo + Lx.n && cy >= Ly.lo && cy < Ly.lo + Ly.n;
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  const int64_t tl = live ? (int64_t)(cy - Ly.lo) * Lx.n + (cx - Lx.lo) : 0;
  partials_cell_sorted<METHOD, NF, G>(I.g, all_groups[(int64_t)gi * n + ii], start + I.bucket_off, xs, ys, ms,
                                      part + all_partoff[(int64_t)gi * n + ii], nL, tl, live ? cx : 0,
                                      live ? cy : 0, j, live);
}

// node pass: launch row blockIdx.y = group g0 + blockIdx.y, every field of it
__global__ void __launch_bounds__(256) kb_nodes(const InstDev *__restrict__ insts, int n,
                                                const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                const NodeGroup *__restrict__ nodes, int g0) {
  const int ii = bi[blockIdx.x];
  const InstDev &I = insts[ii];
  const int64_t i = (blockIdx.x - blk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)I.g.NX * I.g.NY) return;
  const int ix = (int)(i % I.g.NX), jy = (int)(i / I.g.NX);
  node_group(nodes[(int64_t)(g0 + blockIdx.y) * n + ii], ix, jy, i);
}

template <int G>
__global__ void __launch_bounds__(256) kb_partials_subt(const InstDev *__restrict__ insts, int n,
                                                        const int64_t *__restrict__ blk,
                                                        const int32_t *__restrict__ bi,
                                                        const ProjGroup *__restrict__ groups,
                                                        const int64_t *__restrict__ partoff,
                                                        const int32_t *__restrict__ start,
                                                        const int32_t *__restrict__ perm,
                                                        const double *__restrict__ xs, const double *__restrict__ ys,
                                                        const double *__restrict__ mt, double *__restrict__ part,
                                                        double *__restrict__ dTsub) {
  const int ii = bi[blockIdx.x];
  const InstDev &I = insts[ii];
  const int nLx = I.g.NX - 1, nLy = I.g.NY - 1;
  const int64_t nL = (int64_t)nLx * nLy;
  const int64_t lt = (blockIdx.x - blk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  const int64_t t = lt / G;
  const int j = (int)(lt % G);
  const bool live = t < nL;
  const int lx = live ? (int)(t % nLx) : 0;
  const int ly = live ? (int)(t / nLx) : 0;
  SubT sb;
  sb.T = I.lastT;
  sb.rho = I.rhon;
  sb.Cp = I.Cpn;
  sb.k = I.kn;
  sb.MT = mt;                     // payload column of marker T (global positions)
  sb.dTsub = dTsub + I.mark_off;  // written at perm[p], the instance's marker index
  sb.rho0 = I.rho0;
  sb.dt = I.dt;
  sb.dsub = I.dsub;
  partials_cell_subt_sorted<G>(I.g, groups[ii], start + I.bucket_off, perm, xs, ys, part + partoff[ii], nL, t, lx,
                               ly, j, live, sb);
}

__global__ void __launch_bounds__(512) kb_advect(const InstDev *__restrict__ insts, int n,
                                                 const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                 const double *__restrict__ deltaT, const double *__restrict__ dTsub) {
  // the instance's grid coordinates are staged in LDS (as in kb_locate)
  extern __shared__ double sgrid[];
  const int ii = bi[blockIdx.x];
  const InstDev &I = insts[ii];
  const DGrid &g = I.g;
  for (int i = threadIdx.x; i < g.NX; i += blockDim.x) sgrid[i] = g.x[i];
  for (int i = threadIdx.x; i < g.NY; i += blockDim.x) sgrid[g.NX + i] = g.y[i];
  for (int i = threadIdx.x; i <= g.NX; i += blockDim.x) sgrid[g.NX + g.NY + i] = g.xc[i];
  for (int i = threadIdx.x; i <= g.NY; i += blockDim.x) sgrid[2 * g.NX + g.NY + 1 + i] = g.yc[i];
  __syncthreads();
  DGrid gl = g;
  gl.x = sgrid;
  gl.y = sgrid + g.NX;
  gl.xc = sgrid + g.NX + g.NY;
  gl.yc = sgrid + 2 * g.NX + g.NY + 1;
  const int64_t m = (blockIdx.x - blk[ii]) * (int64_t)blockDim.x + threadIdx.x;
  if (m >= I.nMark) return;
  advect_marker<false>(m, gl, I.X, I.Y, I.VX, I.VY, I.wxy, I.cellX, I.cellY, I.vx, I.vy, I.wxyn, I.bv, I.dt, I.T, I.Tdot,
                deltaT + I.node_off, dTsub + I.mark_off, I.dt);
  // the rotation fused behind the RK4 stages: the batched launch keeps one kernel per phase
  rotate_eii_marker(m, g.ncx(), g.NY, I.cellX, I.cellY, I.wxy, I.sxx, I.syy, I.sxy, I.exx, I.eyy, I.exy, I.Eii, I.dt,
                    I.jaumann);
}

// ---- host side ----------------------------------------------------------------------------
// block offsets rows: 0 markers (tune.block), 1 nodes (256), 2 + g lattice cells of
// group g, then the universal lattice ((NX+1) x (NY+1) cells) of the shared
// partials launches
static int64_t *blk_row(fvm_ensemble *E, int r) { return E->dblk + (int64_t)r * (E->n + 1); }
static const int32_t *bi_row(fvm_ensemble *E, int r) { return E->dbi + E->hbi_off[r]; }

static int urow(const fvm_ensemble *E) { return 2 + E->ng + 1; }

static int build_blocks(fvm_ctx *ctx, fvm_ensemble *E) {
  const int n = E->n, rows = urow(E) + 1;
  E->hblk.assign((size_t)rows * (n + 1), 0);
  const int B = E->tune.block, G = E->tune.lanes;
  for (int i = 0; i < n; i++) {
    const InstDev &I = E->hinst[i];
    const int64_t nm = I.nMark, nn = (int64_t)I.g.NX * I.g.NY;
    E->hblk[0 * (n + 1) + i + 1] = E->hblk[0 * (n + 1) + i] + (nm + B - 1) / B;
    E->hblk[1 * (n + 1) + i + 1] = E->hblk[1 * (n + 1) + i] + (nn + 255) / 256;
    for (int gi = 0; gi <= E->ng; gi++) {
      const Lat Lx = lattice(E->gstag_x[gi], I.g.NX), Ly = lattice(E->gstag_y[gi], I.g.NY);
      const int64_t th = (int64_t)Lx.n * Ly.n * G;
      const int r = 2 + gi;
      E->hblk[r * (n + 1) + i + 1] = E->hblk[r * (n + 1) + i] + (th + 255) / 256;
    }
    const int64_t tu = (int64_t)(I.g.NX + 1) * (I.g.NY + 1) * G, u = urow(E);
    E->hblk[u * (n + 1) + i + 1] = E->hblk[u * (n + 1) + i] + (tu + 255) / 256;
  }
  E->hbi_off.assign(rows + 1, 0);
  for (int r = 0; r < rows; r++) E->hbi_off[r + 1] = E->hbi_off[r] + E->hblk[(size_t)r * (n + 1) + n];
  std::vector<int32_t> hbi((size_t)E->hbi_off[rows]);
  for (int r = 0; r < rows; r++)
    for (int i = 0; i < n; i++)
      for (int64_t b = E->hblk[(size_t)r * (n + 1) + i]; b < E->hblk[(size_t)r * (n + 1) + i + 1]; b++)
        hbi[(size_t)(E->hbi_off[r] + b)] = i;
  const size_t need = hbi.size() * sizeof(int32_t) + 8;
  if (need > E->bi_cap) {
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (E->dbi) (void)hipFree(E->dbi);
    E->dbi = nullptr;
    E->bi_cap = 0;
    FVM_HIP(ctx, hipMalloc((void **)&E->dbi, need));
    E->bi_cap = need;
  }
  if (!E->dbi_off) FVM_HIP(ctx, hipMalloc((void **)&E->dbi_off, sizeof(int64_t) * (rows + 1)));
  FVM_HIP(ctx, hipMemcpyAsync(E->dblk, E->hblk.data(), E->hblk.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dbi, hbi.data(), hbi.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dbi_off, E->hbi_off.data(), sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

template <int METHOD, int NF>
static int launch_kb_partials(fvm_ctx *ctx, fvm_ensemble *E, const KbGroups &kg, int nrow) {
  const int n = E->n, u = urow(E);
  const int64_t nblk = E->hblk[(size_t)u * (n + 1) + n];
  if (nblk == 0) return FVM_OK;
  FVM_PHASE(ctx, FVM_PHASE_PARTIALS);
  const unsigned grid = (unsigned)(nblk * nrow);
  const int64_t *ub = blk_row(E, u);
  const int32_t *ubi = bi_row(E, u);
  switch (E->tune.lanes) {
    case 8:
      kb_partials<METHOD, NF, 8><<<grid, 256, 0, ctx->stream>>>(E->dinst, n, ub, ubi, E->dgroups_sorted, E->dpartoff,
                                                                E->start, E->xs, E->ys, E->ms, E->part, kg, nrow);
      break;
    case 32:
      kb_partials<METHOD, NF, 32><<<grid, 256, 0, ctx->stream>>>(E->dinst, n, ub, ubi, E->dgroups_sorted, E->dpartoff,
                                                                 E->start, E->xs, E->ys, E->ms, E->part, kg, nrow);
      break;
    default:
      kb_partials<METHOD, NF, 16><<<grid, 256, 0, ctx->stream>>>(E->dinst, n, ub, ubi, E->dgroups_sorted, E->dpartoff,
                                                                 E->start, E->xs, E->ys, E->ms, E->part, kg, nrow);
  }
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

template <int METHOD>
static int kb_partials_nf(fvm_ctx *ctx, fvm_ensemble *E, int nf, const KbGroups &kg, int nrow) {
  switch (nf) {
    case 1: return launch_kb_partials<METHOD, 1>(ctx, E, kg, nrow);
    case 2: return launch_kb_partials<METHOD, 2>(ctx, E, kg, nrow);
    case 3: return launch_kb_partials<METHOD, 3>(ctx, E, kg, nrow);
    case 4: return launch_kb_partials<METHOD, 4>(ctx, E, kg, nrow);
    case 5: return launch_kb_partials<METHOD, 5>(ctx, E, kg, nrow);
    case 6: return launch_kb_partials<METHOD, 6>(ctx, E, kg, nrow);
    case 7: return launch_kb_partials<METHOD, 7>(ctx, E, kg, nrow);
    case 8: return launch_kb_partials<METHOD, 8>(ctx, E, kg, nrow);
    default: return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "ensemble group with %d fields", nf);
  }
}

// every projection group, groups of one (method, field count) in one launch
static int kb_partials_all(fvm_ctx *ctx, fvm_ensemble *E) {
  std::vector<bool> done(E->ng, false);
  for (int g0 = 0; g0 < E->ng; g0++) {
    if (done[g0]) continue;
    KbGroups kg;
    memset(&kg, 0, sizeof(kg));
    int nrow = 0;
    for (int g = g0; g < E->ng && nrow < KB_MAX_MULTI; g++) {
      if (done[g] || E->gmethod[g] != E->gmethod[g0] || E->gsnf[g] != E->gsnf[g0]) continue;
      kg.gi[nrow] = g;
      kg.sx[nrow] = E->gstag_x[g];
      kg.sy[nrow] = E->gstag_y[g];
      nrow++;
      done[g] = true;
    }
    if (E->gmethod[g0] == PM_ARITH) FVM_TRY(kb_partials_nf<PM_ARITH>(ctx, E, E->gsnf[g0], kg, nrow));
    else if (E->gmethod[g0] == PM_LOCAL) FVM_TRY(kb_partials_nf<PM_LOCAL>(ctx, E, E->gsnf[g0], kg, nrow));
    else return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "ensemble projections support ARITHMETIC and ARITHMETIC_LOCAL");
  }
  return FVM_OK;
}

static int launch_kb_partials_subt(fvm_ctx *ctx, fvm_ensemble *E) {
  const int n = E->n, gi = E->ng;
  const int64_t nblk = E->hblk[(2 + gi) * (n + 1) + n];
  if (nblk == 0) return FVM_OK;
  const ProjGroup *groups = E->dgroups + (int64_t)gi * n;
  const int64_t *po = E->dpartoff + (int64_t)gi * n;
  const int64_t *blk = blk_row(E, 2 + gi);
  const double *mt = E->pcols + (int64_t)E->mt_role * E->nmark;
  const int32_t *bi = bi_row(E, 2 + gi);
  FVM_PHASE(ctx, FVM_PHASE_SUBGRID_T1);
  switch (E->tune.lanes) {
    case 8:
      kb_partials_subt<8><<<(unsigned)nblk, 256, 0, ctx->stream>>>(E->dinst, n, blk, bi, groups, po, E->start, E->perm,
                                                                   E->xs, E->ys, mt, E->part, E->dTsub);
      break;
    case 32:
      kb_partials_subt<32><<<(unsigned)nblk, 256, 0, ctx->stream>>>(E->dinst, n, blk, bi, groups, po, E->start, E->perm,
                                                                    E->xs, E->ys, mt, E->part, E->dTsub);
      break;
    default:
      kb_partials_subt<16><<<(unsigned)nblk, 256, 0, ctx->stream>>>(E->dinst, n, blk, bi, groups, po, E->start, E->perm,
                                                                    E->xs, E->ys, mt, E->part, E->dTsub);
  }
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

static int launch_kb_nodes(fvm_ctx *ctx, fvm_ensemble *E, int g0, int g1) {
  const int n = E->n;
  const int64_t nblk = E->hblk[1 * (n + 1) + n];
  if (nblk == 0 || g1 <= g0) return FVM_OK;
  FVM_PHASE(ctx, FVM_PHASE_NODES);
  kb_nodes<<<dim3((unsigned)nblk, (unsigned)(g1 - g0)), 256, 0, ctx->stream>>>(E->dinst, n, blk_row(E, 1),
                                                                               bi_row(E, 1), E->dnodes, g0);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

static int ensemble_step(fvm_ctx *ctx, fvm_ensemble *E) {
  const int n = E->n;
  const int64_t mblk = E->hblk[0 * (n + 1) + n];
  const int B = E->tune.block;
  // findMarkerCells + density loop + bucket keys (main_subduction.c:243, 255-264)
  const int64_t ntiles = fvm_scan_tiles(E->nbucket);
  unsigned long long *scan_state = (unsigned long long *)E->sums;
  FVM_TRY(fvm_launch_zero_sort(ctx, E->count, E->nbucket + 1, E->counters, scan_state, ntiles + 2));
  if (mblk) {
    FVM_PHASE(ctx, FVM_PHASE_LOCATE);
    kb_locate<<<(unsigned)mblk, B, E->lds_locate, ctx->stream>>>(E->dinst, n, blk_row(E, 0), bi_row(E, 0), E->key,
                                                                  E->slot, E->count, E->counters);
    FVM_LAUNCH_CHECK(ctx);
  }
  FVM_TRY(fvm_launch_scan_onepass(ctx, E->count, E->start, scan_state, E->nbucket));
  if (mblk) {
    FVM_PHASE(ctx, FVM_PHASE_SORT);
    kb_scatter<<<(unsigned)mblk, B, 0, ctx->stream>>>(E->dinst, n, blk_row(E, 0), bi_row(E, 0), E->key, E->slot,
                                                      E->start, E->perm);
    FVM_LAUNCH_CHECK(ctx);
  }
  {
    FVM_PHASE(ctx, FVM_PHASE_SORT);
    kb_order_gather<<<fvm_blocks(E->ntile * 64 * 64, 256), 256, 0, ctx->stream>>>(
        E->dinst, n, E->dcell_off, E->dcell_off + n + 1, E->dtile_inst, E->start, E->perm, E->xs, E->ys, E->ms, E->ncol,
        E->dcolsrc, E->pcols, E->nmark);
    FVM_LAUNCH_CHECK(ctx);
  }
  // projections (:268)
  FVM_TRY(kb_partials_all(ctx, E));
  FVM_TRY(launch_kb_nodes(ctx, E, 0, E->ng));
  // subgridTemperatureChanges (:415): pass 1 inside the dTsub projection, whose
  // node pass applies the VecAXPBYPCZ epilogue
  FVM_TRY(launch_kb_partials_subt(ctx, E));
  FVM_TRY(launch_kb_nodes(ctx, E, E->ng, E->ng + 1));
  // pass 2 fused into advectMarkersRK (:435)
  if (mblk) {
    FVM_PHASE(ctx, FVM_PHASE_ADVECT);
    kb_advect<<<(unsigned)mblk, B, E->lds_locate, ctx->stream>>>(E->dinst, n, blk_row(E, 0), bi_row(E, 0), E->deltaT,
                                                                  E->dTsub);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

static void free_dev(void *p) {
  if (p) (void)hipFree(p);
}

extern "C" int fvm_ensemble_destroy(fvm_ensemble *E) {
  if (!E) return FVM_OK;
  if (E->exec) (void)hipGraphExecDestroy(E->exec);
  if (E->graph) (void)hipGraphDestroy(E->graph);
  void *bufs[] = {E->dinst, E->dgroups, E->dnodes, E->dblk,    E->dpartoff, E->key,       E->slot,
                  E->perm,  E->count,   E->start,  E->sums,    E->counters, E->part,      E->dTsub,
                  E->deltaT, E->xs,     E->ys,     E->pcols,   E->ms,       (void *)E->dcolsrc, E->dcell_off,
                  E->dgroups_sorted, E->dbi, E->dtile_inst, E->dbi_off};
  for (void *b : bufs) free_dev(b);
  delete E;
  return FVM_OK;
}

#define ENS_ALLOC(ctx, ptr, bytes)                                                      \
  do {                                                                                   \
    hipError_t e_ = hipMalloc((void **)&(ptr), (bytes) > 0 ? (bytes) : 8);             \
    if (e_ != hipSuccess) {                                                              \
      fvm_ensemble_destroy(E);                                                           \
      return fvm_set_err(ctx, FVM_ERR_NOMEM, "ensemble hipMalloc: %s", hipGetErrorString(e_)); \
    }                                                                                    \
  } while (0)

extern "C" int fvm_ensemble_create(fvm_ctx *ctx, int32_t n, const fvm_instance *inst, fvm_ensemble **out) {
  if (!ctx || !inst || !out || n < 1) return FVM_ERR_ARG;
  // projection groups from instance 0's list; every instance must have the same pattern
  const fvm_instance &I0 = inst[0];
  if (I0.nproj < 1 || !I0.proj || I0.nproj > 64) return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: projection list");
  for (int i = 0; i < n; i++) {
    FVM_TRY(fvm_grid_check_nonperiodic(ctx, &inst[i].grid, "ensemble instance"));
    if (inst[i].jaumann != 0 && inst[i].jaumann != 1)
      return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: instance %d: jaumann must be 0 or 1", i);
    if (inst[i].nproj != I0.nproj || !inst[i].proj)
      return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: instance %d has a different projection list", i);
    for (int p = 0; p < I0.nproj; p++) {
      const fvm_projection &a = I0.proj[p], &b = inst[i].proj[p];
      if (a.stagx != b.stagx || a.stagy != b.stagy || a.method != b.method ||
          (a.materialProperty == nullptr) != (b.materialProperty == nullptr))
        return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: instance %d projection %d differs in kind", i, p);
    }
    const NodalArrays &nf = inst[i].nodal;
    if (!nf.vx || !nf.vy || !nf.wxy || !nf.thisT || !nf.lastT || !nf.rho || !nf.Cp || !nf.kThermal)
      return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: instance %d lacks a nodal field", i);
  }
  fvm_ensemble *E = new fvm_ensemble();
  E->n = n;
  E->inst.assign(inst, inst + n);
  // group assignment: same rule as fvm_project_marker_fields (first appearance order)
  std::vector<int> gof(I0.nproj, -1), slot(I0.nproj, 0);
  for (int p = 0; p < I0.nproj; p++) {
    if (gof[p] >= 0) continue;
    const int gi = E->ng++;
    E->gstag_x.push_back(I0.proj[p].stagx);
    E->gstag_y.push_back(I0.proj[p].stagy);
    const int m = I0.proj[p].method;
    E->gmethod.push_back(m == ARITHMETIC ? PM_ARITH : (m == ARITHMETIC_LOCAL ? PM_LOCAL : PM_HARM));
    int nf = 0;
    for (int q = p; q < I0.nproj; q++) {
      const fvm_projection &b = I0.proj[q];
      if (gof[q] >= 0 || b.stagx != I0.proj[p].stagx || b.stagy != I0.proj[p].stagy || b.method != m) continue;
      if (nf == FVM_MAXF) break;
      gof[q] = gi;
      slot[q] = nf++;
    }
    E->gnf.push_back(nf);
  }
  // the dTsub projection group (subgridTemperatureChanges.c:74)
  E->gstag_x.push_back(0);
  E->gstag_y.push_back(0);
  E->gmethod.push_back(PM_ARITH);
  E->gnf.push_back(1);
  const int NG = E->ng + 1;
  // ARITHMETIC groups one field short of the largest (<= 5) store a zero field
  // so that they share the largest group's launch
  E->gsnf = E->gnf;
  {
    int nmax = 0;
    for (int gi = 0; gi < E->ng; gi++)
      if (E->gmethod[gi] == PM_ARITH && E->gnf[gi] <= 5) nmax = std::max(nmax, E->gnf[gi]);
    for (int gi = 0; gi < E->ng; gi++)
      if (E->gmethod[gi] == PM_ARITH && E->gnf[gi] == nmax - 1) E->gsnf[gi] = nmax;
  }

  // per-instance device views and offsets
  E->hinst.resize(n);
  size_t max_lds = 0;
  std::vector<int64_t> partoff((size_t)NG * n);
  int64_t parttot = 0;
  for (int i = 0; i < n; i++) {
    const fvm_instance &S = inst[i];
    InstDev &D = E->hinst[i];
    memset(&D, 0, sizeof(D));
    D.g = dgrid(&S.grid);
    D.nMark = S.markers.nMark;
    D.mark_off = E->nmark;
    D.bucket_off = E->nbucket;
    D.node_off = E->nnode;
    const fvm_markers &M = S.markers;
    D.X = M.X, D.Y = M.Y, D.VX = M.VX, D.VY = M.VY, D.T = M.T, D.Tdot = M.Tdot;
    D.sxx = M.sxx, D.syy = M.syy, D.sxy = M.sxy, D.exx = M.exx, D.eyy = M.eyy, D.exy = M.exy;
    D.Eii = M.Eii, D.wxy = M.wxy, D.rho = M.rho, D.rhodot = M.rhodot;
    D.Mat = M.Mat, D.cellX = M.cellX, D.cellY = M.cellY, D.cpu = M.cpu;
    D.vx = S.nodal.vx, D.vy = S.nodal.vy, D.wxyn = S.nodal.wxy, D.thisT = S.nodal.thisT;
    D.lastT = S.nodal.lastT, D.rhon = S.nodal.rho, D.Cpn = S.nodal.Cp, D.kn = S.nodal.kThermal;
    for (int k = 0; k < FVM_MAXMAT; k++) {
      D.dens.rho[k] = S.materials.materialRho[k];
      D.dens.alpha[k] = S.materials.materialAlpha[k];
    }
    D.rho0 = S.materials.materialRho[0];
    D.bv = S.bv;
    D.dt = S.dt;
    D.dsub = S.dsubgrid;
    D.jaumann = S.jaumann;
    const int64_t ncell = (int64_t)(S.grid.NX - 1) * (S.grid.NY - 1);
    E->nmark += S.markers.nMark;
    E->nbucket += 4 * ncell;
    E->ncell += ncell;
    E->nnode += (int64_t)S.grid.NX * S.grid.NY;
    max_lds = std::max(max_lds, (size_t)(2 * S.grid.NX + 2 * S.grid.NY + 2) * sizeof(double));
    for (int gi = 0; gi < NG; gi++) {
      const Lat Lx = lattice(E->gstag_x[gi], S.grid.NX), Ly = lattice(E->gstag_y[gi], S.grid.NY);
      partoff[(size_t)gi * n + i] = parttot;
      parttot += (int64_t)Lx.n * Ly.n * 4 * (E->gsnf[gi] + 1);
    }
  }
  if (max_lds > 64 * 1024) {
    fvm_ensemble_destroy(E);
    return fvm_set_err(ctx, FVM_ERR_ARG, "ensemble: grid too large for the batched location kernel");
  }
  E->lds_locate = max_lds;
  const int64_t ntiles = fvm_scan_tiles(E->nbucket);
  ENS_ALLOC(ctx, E->dinst, sizeof(InstDev) * n);
  ENS_ALLOC(ctx, E->dgroups, sizeof(ProjGroup) * NG * n);
  ENS_ALLOC(ctx, E->dnodes, sizeof(NodeGroup) * NG * n);
  ENS_ALLOC(ctx, E->dblk, sizeof(int64_t) * (3 + NG) * (n + 1));
  ENS_ALLOC(ctx, E->dpartoff, sizeof(int64_t) * NG * n);
  ENS_ALLOC(ctx, E->key, sizeof(int32_t) * E->nmark);
  ENS_ALLOC(ctx, E->slot, sizeof(int32_t) * E->nmark);
  ENS_ALLOC(ctx, E->perm, sizeof(int32_t) * E->nmark);
  ENS_ALLOC(ctx, E->count, sizeof(int32_t) * (E->nbucket + 1));
  ENS_ALLOC(ctx, E->start, sizeof(int32_t) * (E->nbucket + 1));
  ENS_ALLOC(ctx, E->sums, sizeof(unsigned long long) * (ntiles + 2));  // one-pass scan state
  ENS_ALLOC(ctx, E->counters, sizeof(int32_t) * 4);
  ENS_ALLOC(ctx, E->part, sizeof(double) * parttot);
  ENS_ALLOC(ctx, E->dTsub, sizeof(double) * E->nmark);
  ENS_ALLOC(ctx, E->deltaT, sizeof(double) * E->nnode);

  // projection and node groups per (group, instance)
  E->hgroups.assign((size_t)NG * n, ProjGroup());
  std::vector<NodeGroup> hnodes((size_t)NG * n);
  for (int i = 0; i < n; i++) {
    const fvm_instance &S = inst[i];
    for (int gi = 0; gi < NG; gi++) {
      ProjGroup &pg = E->hgroups[(size_t)gi * n + i];
      memset(&pg, 0, sizeof(pg));
      pg.stagx = E->gstag_x[gi];
      pg.stagy = E->gstag_y[gi];
      pg.method = E->gmethod[gi];
      pg.nf = E->gsnf[gi];
      for (int k = E->gnf[gi]; k < E->gsnf[gi]; k++) pg.src[k].use_mat = 1;  // zero field (padding)
      pg.epilogue = EPI_DIVIDE;
    }
    for (int p = 0; p < I0.nproj; p++) {
      ProjGroup &pg = E->hgroups[(size_t)gof[p] * n + i];
      ProjSrc &s = pg.src[slot[p]];
      const fvm_projection &b = S.proj[p];
      if (b.materialProperty) {
        s.use_mat = 1;
        for (int k = 0; k < FVM_MAXMAT; k++) s.mat[k] = b.materialProperty[k];
      } else {
        s.col = b.markerField;
      }
      pg.out[slot[p]] = b.nodalField;
    }
    ProjGroup &dg = E->hgroups[(size_t)E->ng * n + i];
    dg.src[0].col = E->dTsub + E->hinst[i].mark_off;
    dg.out[0] = E->deltaT + E->hinst[i].node_off;
    dg.epilogue = EPI_SUBGRID_T;
    dg.ea = S.nodal.thisT;
    dg.eb = S.nodal.lastT;
    for (int gi = 0; gi < NG; gi++) {
      const ProjGroup &pg = E->hgroups[(size_t)gi * n + i];
      NodeGroup &q = hnodes[(size_t)gi * n + i];
      memset(&q, 0, sizeof(q));
      const Lat Lx = lattice(pg.stagx, S.grid.NX), Ly = lattice(pg.stagy, S.grid.NY);
      q.part = E->part + partoff[(size_t)gi * n + i];
      q.nf = E->gnf[gi];
      q.pst = E->gsnf[gi] + 1;
      q.harm = pg.method == PM_HARM;
      q.nLx = Lx.n;
      q.nLy = Ly.n;
      q.lox = Lx.lo;
      q.loy = Ly.lo;
      q.epilogue = pg.epilogue;
      for (int k = 0; k < q.nf; k++) q.out[k] = pg.out[k];
      q.ea = pg.ea;
      q.eb = pg.eb;
    }
  }
  // payload roles: fields whose column pointers agree in every instance share one
  std::vector<std::vector<const double *>> roles;
  auto role_of = [&](const std::vector<const double *> &v) -> int {
    for (size_t r = 0; r < roles.size(); r++)
      if (roles[r] == v) return (int)r;
    roles.push_back(v);
    return (int)roles.size() - 1;
  };
  std::vector<int> role_gk((size_t)E->ng * FVM_MAXF, -1);  // role of (group, field), -1 = material table
  for (int gi = 0; gi < E->ng; gi++)
    for (int k = 0; k < E->gnf[gi]; k++) {
      if (E->hgroups[(size_t)gi * n].src[k].use_mat) continue;
      std::vector<const double *> v(n);
      for (int i = 0; i < n; i++) v[i] = E->hgroups[(size_t)gi * n + i].src[k].col;
      role_gk[(size_t)gi * FVM_MAXF + k] = role_of(v);
    }
  {
    std::vector<const double *> v(n);
    for (int i = 0; i < n; i++) v[i] = E->hinst[i].T;
    E->mt_role = role_of(v);
  }
  E->ncol = (int)roles.size();
  ENS_ALLOC(ctx, E->xs, sizeof(double) * E->nmark);
  ENS_ALLOC(ctx, E->ys, sizeof(double) * E->nmark);
  ENS_ALLOC(ctx, E->ms, sizeof(int32_t) * E->nmark);
  ENS_ALLOC(ctx, E->pcols, sizeof(double) * E->nmark * E->ncol);
  ENS_ALLOC(ctx, E->dcolsrc, sizeof(const double *) * E->ncol * n);
  ENS_ALLOC(ctx, E->dcell_off, sizeof(int64_t) * 2 * (n + 1));
  ENS_ALLOC(ctx, E->dgroups_sorted, sizeof(ProjGroup) * NG * n);
  std::vector<ProjGroup> hsorted = E->hgroups;  // the same groups reading the payload columns
  for (int gi = 0; gi < E->ng; gi++)
    for (int k = 0; k < E->gnf[gi]; k++) {
      const int r = role_gk[(size_t)gi * FVM_MAXF + k];
      if (r < 0) continue;
      for (int i = 0; i < n; i++) hsorted[(size_t)gi * n + i].src[k].col = E->pcols + (int64_t)r * E->nmark;
    }
  std::vector<const double *> hcolsrc((size_t)E->ncol * n);
  for (int r = 0; r < E->ncol; r++)
    for (int i = 0; i < n; i++) hcolsrc[(size_t)r * n + i] = roles[r][i];
  std::vector<int64_t> hcell(2 * (n + 1), 0);
  for (int i = 0; i < n; i++) {
    const int64_t ncx = inst[i].grid.NX - 1, ncy = inst[i].grid.NY - 1;
    hcell[i + 1] = hcell[i] + ncx * ncy;
    hcell[n + 1 + i + 1] = hcell[n + 1 + i] + ((ncx + 7) / 8) * ((ncy + 7) / 8);
  }
  E->ntile = hcell[2 * n + 1];
  std::vector<int32_t> htile((size_t)E->ntile);
  for (int i = 0; i < n; i++)
    for (int64_t t = hcell[n + 1 + i]; t < hcell[n + 1 + i + 1]; t++) htile[(size_t)t] = i;
  ENS_ALLOC(ctx, E->dtile_inst, sizeof(int32_t) * E->ntile);
  FVM_HIP(ctx, hipMemcpyAsync(E->dtile_inst, htile.data(), sizeof(int32_t) * E->ntile, hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dcolsrc, hcolsrc.data(), sizeof(const double *) * E->ncol * n,
                              hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dcell_off, hcell.data(), sizeof(int64_t) * 2 * (n + 1), hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dgroups_sorted, hsorted.data(), sizeof(ProjGroup) * NG * n, hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dinst, E->hinst.data(), sizeof(InstDev) * n, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dgroups, E->hgroups.data(), sizeof(ProjGroup) * NG * n, hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dnodes, hnodes.data(), sizeof(NodeGroup) * NG * n, hipMemcpyHostToDevice,
                              ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->dpartoff, partoff.data(), sizeof(int64_t) * NG * n, hipMemcpyHostToDevice,
                              ctx->stream));
  int rc = build_blocks(ctx, E);
  if (rc) {
    fvm_ensemble_destroy(E);
    return rc;
  }
  *out = E;
  return FVM_OK;
}

extern "C" int fvm_ensemble_step(fvm_ctx *ctx, fvm_ensemble *E) {
  if (!ctx || !E) return FVM_ERR_ARG;
  return ensemble_step(ctx, E);
}

extern "C" int fvm_ensemble_set_tuning(fvm_ctx *ctx, fvm_ensemble *E, const fvm_tuning *t) {
  if (!ctx || !E || !t) return FVM_ERR_ARG;
  if ((t->block != 128 && t->block != 256 && t->block != 512) || (t->lanes != 8 && t->lanes != 16 && t->lanes != 32))
    return fvm_set_err(ctx, FVM_ERR_ARG, "tuning: block in {128,256,512}, lanes in {8,16,32}");
  if (E->tune.block == t->block && E->tune.lanes == t->lanes) return FVM_OK;
  E->tune = *t;
  if (E->exec) {  // a captured graph holds the old launch shapes
    (void)hipGraphExecDestroy(E->exec);
    E->exec = nullptr;
  }
  if (E->graph) {
    (void)hipGraphDestroy(E->graph);
    E->graph = nullptr;
  }
  return build_blocks(ctx, E);
}

extern "C" int fvm_ensemble_get_tuning(const fvm_ensemble *E, fvm_tuning *t) {
  if (!E || !t) return FVM_ERR_ARG;
  *t = E->tune;
  return FVM_OK;
}

extern "C" int64_t fvm_ensemble_markers(const fvm_ensemble *E) { return E ? E->nmark : 0; }

extern "C" int fvm_ensemble_capture(fvm_ctx *ctx, fvm_ensemble *E) {
  if (!ctx || !E) return FVM_ERR_ARG;
  if (E->exec) {
    (void)hipGraphExecDestroy(E->exec);
    E->exec = nullptr;
  }
  if (E->graph) {
    (void)hipGraphDestroy(E->graph);
    E->graph = nullptr;
  }
  const int prof = ctx->profiling;
  ctx->profiling = 0;  // no event records inside the graph
  FVM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = ensemble_step(ctx, E);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->profiling = prof;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fvm_set_err(ctx, FVM_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  E->graph = g;
  FVM_HIP(ctx, hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
  return FVM_OK;
}

extern "C" int fvm_ensemble_launch(fvm_ctx *ctx, fvm_ensemble *E) {
  if (!ctx || !E) return FVM_ERR_ARG;
  if (!E->exec) FVM_TRY(fvm_ensemble_capture(ctx, E));
  FVM_HIP(ctx, hipGraphLaunch(E->exec, ctx->stream));
  return FVM_OK;
}

// Autotuner: time every (block, lanes) pair on `reps` graph launches of the
// real step and keep the fastest.  Each trial advances the ensemble's state.
extern "C" int fvm_ensemble_autotune(fvm_ctx *ctx, fvm_ensemble *E, int32_t reps, fvm_tuning *best,
                                     float *ms_per_step) {
  if (!ctx || !E || reps < 1) return FVM_ERR_ARG;
  const int blocks[3] = {128, 256, 512}, lanes[3] = {8, 16, 32};
  fvm_tuning bt = E->tune;
  float bms = 1e30f;
  for (int bi = 0; bi < 3; bi++)
    for (int li = 0; li < 3; li++) {
      fvm_tuning t = {blocks[bi], lanes[li]};
      FVM_TRY(fvm_ensemble_set_tuning(ctx, E, &t));
      FVM_TRY(fvm_ensemble_capture(ctx, E));
      FVM_TRY(fvm_ensemble_launch(ctx, E));  // warm
      FVM_TRY(fvm_timer_start(ctx));
      for (int r = 0; r < reps; r++) FVM_TRY(fvm_ensemble_launch(ctx, E));
      float ms = 0.f;
      FVM_TRY(fvm_timer_stop(ctx, &ms));
      ms /= reps;
      if (ms < bms) {
        bms = ms;
        bt = t;
      }
    }
  FVM_TRY(fvm_ensemble_set_tuning(ctx, E, &bt));
  FVM_TRY(fvm_ensemble_capture(ctx, E));
  if (best) *best = bt;
  if (ms_per_step) *ms_per_step = bms;
  return FVM_OK;
}

extern "C" int fvm_ensemble_counters(fvm_ctx *ctx, fvm_ensemble *E, int32_t *nOut, int32_t *nBad) {
  if (!ctx || !E) return FVM_ERR_ARG;
  int32_t h[2] = {0, 0};
  FVM_HIP(ctx, hipMemcpyAsync(h, E->counters, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nOut) *nOut = h[0];
  if (nBad) *nBad = h[1];
  return FVM_OK;
}
